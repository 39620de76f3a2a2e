"""FP64 CPU oracle for the differentiable Yee-FDTD path (TEST INFRASTRUCTURE).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.
The product (``diffnano_amd``) never imports it and has no CPU fallback.

It restates, op for op in NumPy float64, the reference solver
(DiffNano ``diffnano/solvers/fdtd3d.py`` and ``fdtd2d.py``), including its
quirks, and adds the pieces the reference lacks (DFT monitors, profiled
sources, an explicit discrete adjoint):

* ``cpml_profile``            fdtd3d.py:51-85, fdtd2d.py:52-88
* ``gaussian_waveform`` etc.  fdtd3d.py:319-338, fdtd2d.py:178-205
* ``source_mask_3d/2d``       fdtd3d.py:340-372, fdtd2d.py:207-227
* ``step3d``                  fdtd3d.py:378-465 (H from old E, then E from new H)
* ``run3d``                   fdtd3d.py:471-504 (+ DFT monitors, extension)
* ``step2d_tm`` / ``_te``     fdtd2d.py:233-267 / 269-303
* ``run2d``                   fdtd2d.py:309-357
* ``adjoint3d`` / ``adjoint2d_tm`` / ``adjoint2d_te``: the gradient the
  reference obtains by autograd through the step chain
  (fdtd3d.py:166-219 re-runs the forward and calls ``torch.autograd.grad``;
  fdtd2d has plain autograd).  Written here as explicit transposes of the
  truncated difference operators, with dL/deps taken from the *direct*
  formula  dL/deps += lambda_E^{n+1} * (-dt/eps^2) * K_E H^{n+1}
  (H is stored), which is independent of the summation-by-parts form the
  GPU kernels use.

3D CPML profile axes.  The reference broadcasts ``b_z``/``c_z`` (length D)
as ``(1,1,D)``, i.e. along axis 2, and ``c_x`` / one ``b_x`` as ``(1,W,1)``,
i.e. along axis 1 (fdtd3d.py:402-463); it therefore only runs cubic grids,
where all three profiles are identical arrays.  The oracle evaluates every
profile with the length of the axis it varies along (axis 1 -> H, axis 2 ->
W): bit-identical to the reference on cubic grids and defined on
non-cubic ones.  Axis 0 (z) is never damped.

Pinning (see tests): golden vectors written by importing the reference
(``tests/golden/make_golden.py``), discrete energy conservation with no
CPML, the numerical dispersion relation, mirror symmetry, causality,
linearity, and for the adjoint a dot-product test and central finite
differences.
"""

from __future__ import annotations

import math

import numpy as np

__all__ = [
    "cpml_profile",
    "gaussian_waveform",
    "continuous_waveform",
    "waveform",
    "waveform_table",
    "source_mask_3d",
    "source_mask_2d",
    "Params3D",
    "Params2D",
    "step3d",
    "run3d",
    "adjoint3d",
    "step2d_tm",
    "step2d_te",
    "run2d",
    "adjoint2d",
    "dft_phase_table",
]


# ---------------------------------------------------------------------------
# CPML profile  (fdtd3d.py:51-85, fdtd2d.py:52-88)
# ---------------------------------------------------------------------------


def cpml_profile(size, pml_layers, dl, dt, sigma_max=0.8, order=3):
    """Return the 1-D damping profiles ``(b, c)`` (float64, length ``size``).

    sigma[i] = sigma_max * ((L - i)/L)**order for i < L, mirrored at the far
    end (fdtd3d.py:67-73); s = sigma*dl/dt; a = 0.02 s; b = exp(-(s+a) dt);
    c = s (b - 1)/(s + a + 1e-30) where s + a > 1e-30 else 0 (fdtd3d.py:75-85).
    ``kappa`` is computed by the reference and never used, so it is omitted.
    """
    sigma = np.zeros(size, dtype=np.float64)
    if pml_layers > 0:
        for i in range(pml_layers):
            d = (pml_layers - i) / pml_layers
            val = sigma_max * d**order
            sigma[i] = val
            sigma[size - 1 - i] = val
    s = sigma * dl / dt
    alpha = 0.02 * s
    b = np.exp(-(s + alpha) * dt)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(s + alpha > 1e-30, s * (b - 1.0) / (s + alpha + 1e-30), 0.0)
    return b, c


# ---------------------------------------------------------------------------
# Source waveform  (fdtd3d.py:319-338, fdtd2d.py:178-205)
# ---------------------------------------------------------------------------


def _np_sin32(x):
    return np.sin(np.float32(x))


def gaussian_waveform(step, dt, omega, source, sin32=_np_sin32):
    """``gaussian_pulse``: amp * sin_f32(f32(omega*t)) * exp(-(t-t0)^2/(2 s^2)).

    The reference builds ``torch.tensor(omega * t)`` without a dtype, so the
    sine argument is rounded to float32 and the sine is evaluated in float32;
    ``amp * sin`` stays float32 (a Python float is a weak scalar); the
    product with the float64 envelope is float64 (fdtd3d.py:333).
    ``sin32`` evaluates the float32 sine; NumPy's may differ from torch's
    CPU kernel (which the reference uses) by one ulp, so callers that need
    bit-equality with the reference pass torch's.
    """
    t = step * dt
    t0 = source.get("t0", 3.0 / omega)
    spread = source.get("spread", 1.0 / omega)
    amp = source.get("amplitude", 1.0)
    sin32 = np.float32(sin32(np.float32(omega * t)))
    amp_sin = np.float32(np.float32(amp) * sin32)
    envelope = math.exp(-((t - t0) ** 2) / (2 * spread**2))
    return float(amp_sin) * envelope


def continuous_waveform(step, dt, omega, source):
    """``continuous``: amp * min(1, t*omega/(4 pi)) * sin(omega t), float64."""
    t = step * dt
    amp = source.get("amplitude", 1.0)
    ramp = min(1.0, t * omega / (4 * math.pi))
    return float(amp * ramp * math.sin(omega * t))


def waveform(step, dt, omega, source, sin32=_np_sin32):
    """Dispatch on ``source['type']`` (default gaussian_pulse); unknown -> 0."""
    kind = source.get("type", "gaussian_pulse")
    if kind == "gaussian_pulse":
        return gaussian_waveform(step, dt, omega, source, sin32)
    if kind == "continuous":
        return continuous_waveform(step, dt, omega, source)
    return 0.0


def waveform_table(n_steps, dt, omega, source, sin32=_np_sin32):
    return np.array([waveform(n, dt, omega, source, sin32) for n in range(n_steps)],
                    dtype=np.float64)


# ---------------------------------------------------------------------------
# Source masks  (fdtd3d.py:340-372, fdtd2d.py:207-227)
# ---------------------------------------------------------------------------


def source_mask_3d(shape, source):
    D, H, W = shape
    mask = np.zeros((D, H, W), dtype=np.float64)
    pos = source.get("pos", None)
    if pos is not None:
        z, y, x = pos
        if 0 <= z < D and 0 <= y < H and 0 <= x < W:
            mask[z, y, x] = 1.0
    else:
        plane = source.get("plane", None)
        if plane == "xy":
            z = source.get("z", D // 2)
            if 0 <= z < D:
                mask[z, :, :] = 1.0
        elif plane == "xz":
            y = source.get("y", H // 2)
            if 0 <= y < H:
                mask[:, y, :] = 1.0
        elif plane == "yz":
            x = source.get("x", W // 2)
            if 0 <= x < W:
                mask[:, :, x] = 1.0
        else:
            mask[D // 2, H // 2, :] = 1.0
    return mask


def source_mask_2d(shape, source):
    H, W = shape
    mask = np.zeros((H, W), dtype=np.float64)
    pos = source.get("pos", None)
    if pos is not None:
        y, x = pos
        if 0 <= y < H and 0 <= x < W:
            mask[y, x] = 1.0
    else:
        row = source.get("row", H // 2)
        if 0 <= row < H:
            mask[row, :] = 1.0
    return mask


# ---------------------------------------------------------------------------
# Truncated differences and their exact transposes
# ---------------------------------------------------------------------------


def _dplus(f, axis, dl):
    """(f[i+1]-f[i])/dl, 0 at the last index (fdtd3d.py:398-429)."""
    out = np.zeros_like(f)
    n = f.shape[axis]
    hi = [slice(None)] * f.ndim
    lo = [slice(None)] * f.ndim
    hi[axis] = slice(1, n)
    lo[axis] = slice(0, n - 1)
    out[tuple(lo)] = (f[tuple(hi)] - f[tuple(lo)]) / dl
    return out


def _dminus(f, axis, dl):
    """(f[i]-f[i-1])/dl, 0 at index 0 (fdtd3d.py:432-457)."""
    out = np.zeros_like(f)
    n = f.shape[axis]
    hi = [slice(None)] * f.ndim
    lo = [slice(None)] * f.ndim
    hi[axis] = slice(1, n)
    lo[axis] = slice(0, n - 1)
    out[tuple(hi)] = (f[tuple(hi)] - f[tuple(lo)]) / dl
    return out


def _dplus_T(g, axis, dl):
    """Transpose of ``_dplus``: [-g0, g0-g1, ..., g_{n-3}-g_{n-2}, g_{n-2}]/dl."""
    out = np.zeros_like(g)
    n = g.shape[axis]
    gh = np.array(g, copy=True)
    idx = [slice(None)] * g.ndim
    idx[axis] = n - 1
    gh[tuple(idx)] = 0.0
    a = [slice(None)] * g.ndim
    b = [slice(None)] * g.ndim
    a[axis] = slice(1, n)
    b[axis] = slice(0, n - 1)
    out -= gh
    out[tuple(a)] += gh[tuple(b)]
    return out / dl


def _dminus_T(g, axis, dl):
    """Transpose of ``_dminus``: [-g1, g1-g2, ..., g_{n-2}-g_{n-1}, g_{n-1}]/dl."""
    out = np.zeros_like(g)
    n = g.shape[axis]
    gt = np.array(g, copy=True)
    idx = [slice(None)] * g.ndim
    idx[axis] = 0
    gt[tuple(idx)] = 0.0
    a = [slice(None)] * g.ndim
    b = [slice(None)] * g.ndim
    a[axis] = slice(0, n - 1)
    b[axis] = slice(1, n)
    out += gt
    out[tuple(a)] -= gt[tuple(b)]
    return out / dl


# ---------------------------------------------------------------------------
# 3D
# ---------------------------------------------------------------------------


class Params3D:
    """Scalars and profiles of ``FDTDSolver3D`` (fdtd3d.py:250-281).

    dt = courant*dl/sqrt(3) (fdtd3d.py:274); omega = 2 pi / lambda0 (:275).
    ``by, cy`` vary along axis 1 (length H), ``bx, cx`` along axis 2
    (length W).
    """

    def __init__(self, grid_shape, dl=20.0, wavelength_nm=1550.0, pml_layers=5, courant=0.4,
                 dtype=np.float64):
        self.shape = tuple(int(v) for v in grid_shape)
        D, H, W = self.shape
        self.dl = float(dl)
        self.courant = courant
        self.dt = courant * dl / math.sqrt(3.0)
        self.omega = 2 * math.pi / wavelength_nm
        self.dtype = dtype
        by, cy = cpml_profile(H, pml_layers, dl, self.dt)
        bx, cx = cpml_profile(W, pml_layers, dl, self.dt)
        self.by = by.astype(dtype).reshape(1, H, 1)
        self.cy = cy.astype(dtype).reshape(1, H, 1)
        self.bx = bx.astype(dtype).reshape(1, 1, W)
        self.cx = cx.astype(dtype).reshape(1, 1, W)


def step3d(E, H, eps, p):
    """One leapfrog step, fdtd3d.py:378-465 (mu_r = 1, so dt/mu_r = dt).

    ``E = (Ex, Ey, Ez)``, ``H = (Hx, Hy, Hz)``, arrays ``(D, H, W)``.
    Axis-2-broadcast ``c_z``/``b_z`` of the reference are ``cx``/``bx`` here
    and axis-1-broadcast ``c_x``/``b_x`` are ``cy``/``by`` (identical arrays
    on cubic grids).  Returns new ``(E, H)``.
    """
    Ex, Ey, Ez = E
    Hx, Hy, Hz = H
    dt, dl = p.dt, p.dl
    bx, cx, by, cy = p.bx, p.cx, p.by, p.cy
    dtm = dt  # dt / mu_r with mu_r == 1

    dEz_dy = _dplus(Ez, 1, dl)
    dEy_dz = _dplus(Ey, 0, dl)
    Hx = by * Hx - dtm * (cy * dEz_dy + dEz_dy - cx * dEy_dz - dEy_dz)

    dEx_dz = _dplus(Ex, 0, dl)
    dEz_dx = _dplus(Ez, 2, dl)
    Hy = bx * Hy - dtm * (cx * dEx_dz + dEx_dz - cy * dEz_dx - dEz_dx)

    dEy_dx = _dplus(Ey, 2, dl)
    dEx_dy = _dplus(Ex, 1, dl)
    Hz = bx * Hz - dtm * (cx * dEy_dx + dEy_dx - cy * dEx_dy - dEx_dy)

    ce = dt / eps
    dHz_dy = _dminus(Hz, 1, dl)
    dHy_dz = _dminus(Hy, 0, dl)
    Ex = by * bx * Ex + ce * (cy * dHz_dy + dHz_dy - cx * dHy_dz - dHy_dz)

    dHx_dz = _dminus(Hx, 0, dl)
    dHz_dx = _dminus(Hz, 2, dl)
    Ey = bx * by * Ey + ce * (cx * dHx_dz + dHx_dz - cy * dHz_dx - dHz_dx)

    dHy_dx = _dminus(Hy, 2, dl)
    dHx_dy = _dminus(Hx, 1, dl)
    Ez = bx * by * Ez + ce * (cx * dHy_dx + dHy_dx - cy * dHx_dy - dHx_dy)
    return (Ex, Ey, Ez), (Hx, Hy, Hz)


def _curl_E_3d(H, p):
    """K_E H per component (the bracket multiplied by dt/eps in ``step3d``)."""
    Hx, Hy, Hz = H
    dl, cx, cy = p.dl, p.cx, p.cy
    kx = (1 + cy) * _dminus(Hz, 1, dl) - (1 + cx) * _dminus(Hy, 0, dl)
    ky = (1 + cx) * _dminus(Hx, 0, dl) - (1 + cy) * _dminus(Hz, 2, dl)
    kz = (1 + cx) * _dminus(Hy, 2, dl) - (1 + cy) * _dminus(Hx, 1, dl)
    return kx, ky, kz


def dft_phase_table(n_steps, dt, omegas):
    """phase[n, f] = omega_f * (n+1) * dt: the sample after step n is E^{n+1}."""
    n = np.arange(1, n_steps + 1, dtype=np.float64)[:, None]
    return n * dt * np.asarray(omegas, dtype=np.float64)[None, :]


def _dft_accumulate(acc, fields, comps, region, phase_row):
    # acc: complex array (n_freq, n_comp, *region_shape)
    for ci, c in enumerate(comps):
        v = fields[c][region]
        for f, ph in enumerate(phase_row):
            acc[f, ci] += v * complex(math.cos(ph), -math.sin(ph))


def run3d(eps, p, n_steps, src_comp=2, src_mask=None, wave=None, monitors=(), store=False):
    """Forward run, fdtd3d.py:471-504 (zero init; step; ``E_c += mask*w(n)``).

    ``src_comp``: 0/1/2 = Ex/Ey/Ez (the reference injects Ez only).
    ``src_mask``: dense amplitude array (reference: 0/1 mask).
    ``wave``: waveform table w(n), n < n_steps.
    ``monitors``: list of dicts {region: tuple of slices, comps: [..], omegas: [..]}
    accumulating F = sum_{n=1..T} E^n exp(-i omega n dt).
    ``store``: keep E^n, H^n for n = 0..T (for the adjoint).
    Returns dict with E, H, monitors (list of complex arrays), history.
    """
    D, H_, W = p.shape
    dt = p.dt
    z = lambda: np.zeros((D, H_, W), dtype=p.dtype)  # noqa: E731
    E = (z(), z(), z())
    H = (z(), z(), z())
    eps = np.asarray(eps, dtype=p.dtype)
    accs = []
    tabs = []
    for m in monitors:
        shape = tuple(np.zeros(p.shape)[m["region"]].shape)
        accs.append(np.zeros((len(m["omegas"]), len(m["comps"])) + shape, dtype=np.complex128))
        tabs.append(dft_phase_table(n_steps, dt, m["omegas"]))
    hist_E = [E] if store else None
    hist_H = [H] if store else None
    for n in range(n_steps):
        E, H = step3d(E, H, eps, p)
        if src_mask is not None and wave is not None:
            E = list(E)
            E[src_comp] = E[src_comp] + src_mask * wave[n]
            E = tuple(E)
        for m, acc, tab in zip(monitors, accs, tabs):
            _dft_accumulate(acc, E, m["comps"], m["region"], tab[n])
        if store:
            hist_E.append(E)
            hist_H.append(H)
    return {"E": E, "H": H, "monitors": accs, "hist_E": hist_E, "hist_H": hist_H}


def _monitor_direct(shape, monitors, mon_grads, tabs, m_index, dtype):
    """dL/dE^m (m >= 1) from monitor gradients g = dL/dRe F + i dL/dIm F.

    F = sum_m E^m (cos phi_m - i sin phi_m), so dL/dE^m = Re(g e^{i phi_m}).
    """
    out = [np.zeros(shape, dtype=dtype) for _ in range(3)]
    for m, g, tab in zip(monitors, mon_grads, tabs):
        ph = tab[m_index - 1]
        for ci, c in enumerate(m["comps"]):
            for f in range(len(m["omegas"])):
                out[c][m["region"]] += np.real(g[f, ci] * complex(math.cos(ph[f]), math.sin(ph[f])))
    return out


def adjoint3d(eps, p, n_steps, lam_E_T, src_comp=2, src_mask=None, wave=None, monitors=(),
              mon_grads=()):
    """dL/deps for L with dL/dE^T = ``lam_E_T`` (3 arrays) and monitor grads.

    Reverse sweep over the stored forward (Appendix-B form of the autograd
    the reference runs in fdtd3d.py:184-217):
      mu_H   = lam_H^{n+1} + K_E^T[(dt/eps) lam_E^{n+1}]
      lam_E^n = B_E lam_E^{n+1} - dt K_H^T mu_H + dL/dE^n (monitors)
      lam_H^n = B_H mu_H
      dL/deps += sum_c lam_Ec^{n+1} (-dt/eps^2) (K_E H^{n+1})_c
    """
    fw = run3d(eps, p, n_steps, src_comp, src_mask, wave, monitors, store=True)
    eps = np.asarray(eps, dtype=np.float64)
    dl, dt = p.dl, p.dt
    bx, cx, by, cy = p.bx, p.cx, p.by, p.cy
    tabs = [dft_phase_table(n_steps, dt, m["omegas"]) for m in monitors]
    lam = [np.array(v, dtype=np.float64) for v in lam_E_T]
    if monitors:
        d = _monitor_direct(p.shape, monitors, mon_grads, tabs, n_steps, np.float64)
        lam = [lam[i] + d[i] for i in range(3)]
    lamH = [np.zeros(p.shape) for _ in range(3)]
    grad = np.zeros(p.shape)
    ax, ay = 1 + cx, 1 + cy
    for n in range(n_steps - 1, -1, -1):
        kE = _curl_E_3d(fw["hist_H"][n + 1], p)
        for c in range(3):
            grad += lam[c] * (-dt / eps**2) * kE[c]
        u = [(dt / eps) * lam[c] for c in range(3)]
        # K_E^T: Ex' += ce (aY Dy- Hz - aX Dz- Hy), Ey' += ce (aX Dz- Hx - aY Dx- Hz),
        #        Ez' += ce (aX Dx- Hy - aY Dy- Hx)
        muHx = lamH[0] + _dminus_T(ax * u[1], 0, dl) - _dminus_T(ay * u[2], 1, dl)
        muHy = lamH[1] + _dminus_T(ax * u[2], 2, dl) - _dminus_T(ax * u[0], 0, dl)
        muHz = lamH[2] + _dminus_T(ay * u[0], 1, dl) - _dminus_T(ay * u[1], 2, dl)
        # K_H^T: Hx' -= dt (aY Dy+ Ez - aX Dz+ Ey), Hy' -= dt (aX Dz+ Ex - aY Dx+ Ez),
        #        Hz' -= dt (aX Dx+ Ey - aY Dy+ Ex)
        lEx = bx * by * lam[0] - dt * (_dplus_T(ax * muHy, 0, dl) - _dplus_T(ay * muHz, 1, dl))
        lEy = bx * by * lam[1] - dt * (_dplus_T(ax * muHz, 2, dl) - _dplus_T(ax * muHx, 0, dl))
        lEz = bx * by * lam[2] - dt * (_dplus_T(ay * muHx, 1, dl) - _dplus_T(ay * muHy, 2, dl))
        lam = [lEx, lEy, lEz]
        if monitors and n >= 1:
            d = _monitor_direct(p.shape, monitors, mon_grads, tabs, n, np.float64)
            lam = [lam[i] + d[i] for i in range(3)]
        lamH = [by * muHx, bx * muHy, bx * muHz]
    return grad, fw


# ---------------------------------------------------------------------------
# 2D (TM: Ez, Hx, Hy; TE: Hz, Ex, Ey)
# ---------------------------------------------------------------------------


class Params2D:
    """Scalars and profiles of ``FDTDSolver2D`` (fdtd2d.py:117-149).

    dt = courant*dl/sqrt(2) (fdtd2d.py:142).  ``by, cy`` (length H) vary
    along axis 0; ``bx, cx`` (length W) along axis 1 (fdtd2d.py:147-148).
    """

    def __init__(self, grid_shape, dl=20.0, wavelength_nm=1550.0, polarization="TM",
                 pml_layers=10, courant=0.5, dtype=np.float64):
        self.shape = tuple(int(v) for v in grid_shape)
        H, W = self.shape
        self.dl = float(dl)
        self.courant = courant
        self.polarization = polarization.upper()
        self.dt = courant * dl / math.sqrt(2.0)
        self.omega = 2 * math.pi / wavelength_nm
        self.dtype = dtype
        bx, cx = cpml_profile(W, pml_layers, dl, self.dt)
        by, cy = cpml_profile(H, pml_layers, dl, self.dt)
        self.bx = bx.astype(dtype).reshape(1, W)
        self.cx = cx.astype(dtype).reshape(1, W)
        self.by = by.astype(dtype).reshape(H, 1)
        self.cy = cy.astype(dtype).reshape(H, 1)


def step2d_tm(Ez, Hx, Hy, eps, p):
    """fdtd2d.py:233-267.  Note the reference's "dy" derivative is along
    axis 1 and its "dx" derivative along axis 0."""
    dt, dl = p.dt, p.dl
    bx, cx, by, cy = p.bx, p.cx, p.by, p.cy
    dEz_dy = _dplus(Ez, 1, dl)
    Hx = by * Hx - dt * (cy * dEz_dy + dEz_dy)
    dEz_dx = _dplus(Ez, 0, dl)
    Hy = bx * Hy + dt * (cx * dEz_dx + dEz_dx)
    dHy_dx = _dminus(Hy, 0, dl)
    dHx_dy = _dminus(Hx, 1, dl)
    Ez = bx * by * Ez + (dt / eps) * (cx * dHy_dx + dHy_dx - cy * dHx_dy - dHx_dy)
    return Ez, Hx, Hy


def step2d_te(Hz, Ex, Ey, eps, p):
    """fdtd2d.py:269-303, including the reference's Hz sign (``+ dt*(...)``)."""
    dt, dl = p.dt, p.dl
    bx, cx, by, cy = p.bx, p.cx, p.by, p.cy
    dHz_dx = _dminus(Hz, 1, dl)
    Ex = bx * Ex + (dt / eps) * (cx * dHz_dx + dHz_dx)
    dHz_dy = _dminus(Hz, 0, dl)
    Ey = by * Ey - (dt / eps) * (cy * dHz_dy + dHz_dy)
    dEy_dx = _dplus(Ey, 0, dl)
    dEx_dy = _dplus(Ex, 1, dl)
    Hz = bx * by * Hz + dt * (cx * dEy_dx + dEy_dx - cy * dEx_dy - dEx_dy)
    return Hz, Ex, Ey


def run2d(eps, p, n_steps, src_mask=None, wave=None, monitors=(), probe=None, store=False):
    """fdtd2d.py:309-357.  TM returns Ez as ``field``; TE returns Hz.

    The source is added to Ez (TM) or Hz (TE) after each step.  Monitors
    sample the ``field`` component after each step: F = sum_{n=1..T} f^n
    exp(-i omega n dt), ``comps`` ignored.
    """
    H_, W = p.shape
    z = lambda: np.zeros((H_, W), dtype=p.dtype)  # noqa: E731
    a, b, c = z(), z(), z()
    eps = np.asarray(eps, dtype=p.dtype)
    step = step2d_tm if p.polarization == "TM" else step2d_te
    accs, tabs = [], []
    for m in monitors:
        shape = tuple(np.zeros(p.shape)[m["region"]].shape)
        accs.append(np.zeros((len(m["omegas"]),) + shape, dtype=np.complex128))
        tabs.append(dft_phase_table(n_steps, p.dt, m["omegas"]))
    hist = [(a, b, c)] if store else None
    snaps = []
    for n in range(n_steps):
        a, b, c = step(a, b, c, eps, p)
        if src_mask is not None and wave is not None:
            a = a + src_mask * wave[n]
        for m, acc, tab in zip(monitors, accs, tabs):
            v = a[m["region"]]
            for f, ph in enumerate(tab[n]):
                acc[f] += v * complex(math.cos(ph), -math.sin(ph))
        if probe is not None:
            snaps.append(a[probe[0], probe[1]])
        if store:
            hist.append((a, b, c))
    return {"field": a, "aux": (b, c), "monitors": accs, "snapshots": np.array(snaps), "hist": hist}


def adjoint2d(eps, p, n_steps, lam_T, src_mask=None, wave=None, monitors=(), mon_grads=()):
    """dL/deps for the 2D run (TM or TE) with dL/d(field^T) = ``lam_T``.

    TM: a=Ez, (b,c)=(Hx,Hy); E-like state Ez, H-like Hx, Hy (H first).
    TE: a=Hz, (b,c)=(Ex,Ey); E-like Ex, Ey are updated first from Hz.
    """
    fw = run2d(eps, p, n_steps, src_mask, wave, monitors, store=True)
    eps = np.asarray(eps, dtype=np.float64)
    dl, dt = p.dl, p.dt
    bx, cx, by, cy = p.bx, p.cx, p.by, p.cy
    ax, ay = 1 + cx, 1 + cy
    tabs = [dft_phase_table(n_steps, dt, m["omegas"]) for m in monitors]

    def direct(m_index):
        out = np.zeros(p.shape)
        for m, g, tab in zip(monitors, mon_grads, tabs):
            ph = tab[m_index - 1]
            for f in range(len(m["omegas"])):
                out[m["region"]] += np.real(g[f] * complex(math.cos(ph[f]), math.sin(ph[f])))
        return out

    grad = np.zeros(p.shape)
    lam_a = np.array(lam_T, dtype=np.float64) + (direct(n_steps) if monitors else 0.0)
    lam_b = np.zeros(p.shape)
    lam_c = np.zeros(p.shape)
    if p.polarization == "TM":
        # Hx' = by Hx - dt aY D1+ Ez ; Hy' = bx Hy + dt aX D0+ Ez
        # Ez' = bxby Ez + (dt/eps)(aX D0- Hy' - aY D1- Hx')
        for n in range(n_steps - 1, -1, -1):
            _, Hx1, Hy1 = fw["hist"][n + 1]
            kE = ax * _dminus(Hy1, 0, dl) - ay * _dminus(Hx1, 1, dl)
            grad += lam_a * (-dt / eps**2) * kE
            u = (dt / eps) * lam_a
            muHx = lam_b - _dminus_T(ay * u, 1, dl)
            muHy = lam_c + _dminus_T(ax * u, 0, dl)
            lam_a = bx * by * lam_a - dt * _dplus_T(ay * muHx, 1, dl) + dt * _dplus_T(ax * muHy, 0, dl)
            if monitors and n >= 1:
                lam_a = lam_a + direct(n)
            lam_b = by * muHx
            lam_c = bx * muHy
    else:
        # Ex' = bx Ex + (dt/eps) aX D1- Hz ; Ey' = by Ey - (dt/eps) aY D0- Hz
        # Hz' = bxby Hz + dt (aX D0+ Ey' - aY D1+ Ex')
        # Adjoint: lam_a = lam_Hz, lam_b = lam_Ex, lam_c = lam_Ey.
        for n in range(n_steps - 1, -1, -1):
            Hz0 = fw["hist"][n][0]
            # mu_E = lam_E' + (dHz'/dE')^T lam_Hz'
            muEx = lam_b - dt * _dplus_T(ay * lam_a, 1, dl)
            muEy = lam_c + dt * _dplus_T(ax * lam_a, 0, dl)
            # eps gradient: Ex' = ... + (dt/eps) aX D1- Hz^n (old Hz)
            kx = ax * _dminus(Hz0, 1, dl)
            ky = -ay * _dminus(Hz0, 0, dl)
            grad += (muEx * kx + muEy * ky) * (-dt / eps**2)
            uX = (dt / eps) * muEx
            uY = (dt / eps) * muEy
            lam_a = bx * by * lam_a + _dminus_T(ax * uX, 1, dl) - _dminus_T(ay * uY, 0, dl)
            if monitors and n >= 1:
                lam_a = lam_a + direct(n)
            lam_b = bx * muEx
            lam_c = by * muEy
    return grad, fw

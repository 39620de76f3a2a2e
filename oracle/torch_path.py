"""Torch FP64 restatement of the reference's CPU path (test and CPU-baseline infrastructure).

Only ``tests/`` and ``bench.py``'s CPU-baseline leg import this module; the
product never does.  It performs, in PyTorch float64 on the host, the same
sequence of tensor operations as DiffNano's eager CPU path, so that timing
it on the host's cores measures what the reference costs there:

* ``ref_time_step``      FDTDSolver3D._time_step (fdtd3d.py:378-465): per
                         difference a ``zeros_like`` plus a sliced
                         subtraction divided by dl written into it (12 per
                         step), per component the broadcast damped update
                         ``b F -/+ (dt/mu or dt/eps) (c d + d - c' d' - d')``
                         with the profiles broadcast along the reference's
                         axes (axis-2 "z" profiles, axis-1 "x" profiles in
                         the mixed terms);
* ``ref_waveform``       _source_waveform (fdtd3d.py:319-338): a float32
                         ``sin`` of a default-dtype tensor times a float64
                         envelope, built from fresh scalar tensors per step;
* ``time_reversal_forward`` / ``time_reversal_backward``
                         _TimeReversalFDTD (fdtd3d.py:103-219): the detached
                         forward that clones Ex, Ey, Ez every step and reads
                         the waveform back with ``.item()``, and the backward
                         that re-runs the forward from zero fields with
                         autograd on eps and calls ``torch.autograd.grad``.

The profiles come from ``oracle.fdtd_oracle.cpml_profile`` (fdtd3d.py:51-85).
Pinned against the reference goldens (one step, forwards, time-reversal
gradients) in tests/test_oracle.py.
"""

from __future__ import annotations

import math

import torch

from oracle.fdtd_oracle import cpml_profile

F64 = torch.float64


class RefPath3D:
    """Grid, step and profiles of one reference-path 3D solve (cubic grids,
    as the reference's broadcasts require)."""

    def __init__(self, grid_shape, dl=20.0, wavelength_nm=1550.0, pml_layers=5, courant=0.4):
        D, H, W = grid_shape
        if not D == H == W:
            raise ValueError("the reference's CPML broadcasts need a cubic grid")
        self.shape = (D, H, W)
        self.dl = float(dl)
        self.dt = courant * dl / math.sqrt(3.0)
        self.omega = 2 * math.pi / wavelength_nm
        b, c = cpml_profile(D, pml_layers, dl, self.dt)
        t = lambda a: torch.as_tensor(a, dtype=F64)  # noqa: E731
        self.b, self.c = t(b), t(c)

    # broadcast views of the 1-D profile along axis 1 ("y" / mixed "x") and
    # along axis 2 ("x" / mixed "z")
    def _ax1(self, v):
        return v.unsqueeze(0).unsqueeze(-1)

    def _ax2(self, v):
        return v.unsqueeze(0).unsqueeze(0)


def _difference(f, axis, dl, at_low):
    """zeros_like(f) with (f[i+1] - f[i]) / dl written at index i (forward
    difference, ``at_low``) or i+1 (backward difference) along ``axis``."""
    d = torch.zeros_like(f)
    n = f.shape[axis]
    up = [slice(None)] * 3
    dn = [slice(None)] * 3
    up[axis], dn[axis] = slice(1, n), slice(0, n - 1)
    d[tuple(dn if at_low else up)] = (f[tuple(up)] - f[tuple(dn)]) / dl
    return d


def ref_time_step(p: RefPath3D, Ex, Ey, Ez, Hx, Hy, Hz, eps_r, mu_r):
    """One step: H from the old E, then E from the new H (no source)."""
    dl, dt = p.dl, p.dt
    b, c = p.b, p.c
    b1, c1, b2, c2 = p._ax1(b), p._ax1(c), p._ax2(b), p._ax2(c)

    def h_update(F, bF, ca, da, cb, db):
        return bF * F - (dt / mu_r) * (ca * da + da - cb * db - db)

    def e_update(F, bF, ca, da, cb, db):
        return bF * F + (dt / eps_r) * (ca * da + da - cb * db - db)

    # (axis 0 = "z", 1 = "y", 2 = "x"; forward differences of E, backward of H)
    Hx = h_update(Hx, b1, c1, _difference(Ez, 1, dl, True), c2, _difference(Ey, 0, dl, True))
    Hy = h_update(Hy, b2, c2, _difference(Ex, 0, dl, True), c1, _difference(Ez, 2, dl, True))
    Hz = h_update(Hz, b2, c2, _difference(Ey, 2, dl, True), c1, _difference(Ex, 1, dl, True))
    Ex = e_update(Ex, b1 * b2, c1, _difference(Hz, 1, dl, False), c2,
                  _difference(Hy, 0, dl, False))
    Ey = e_update(Ey, b2 * b1, c2, _difference(Hx, 0, dl, False), c1,
                  _difference(Hz, 2, dl, False))
    Ez = e_update(Ez, b2 * b1, c2, _difference(Hy, 2, dl, False), c1,
                  _difference(Hx, 1, dl, False))
    return Ex, Ey, Ez, Hx, Hy, Hz


def ref_waveform(p: RefPath3D, step, source):
    """Gaussian pulse: amp * sin_f32(omega t) * exp(-(t - t0)^2 / (2 spread^2));
    "continuous": amp * min(1, t omega / 4 pi) * sin(omega t); else 0."""
    kind = source.get("type", "gaussian_pulse")
    t = step * p.dt
    w = p.omega
    if kind == "gaussian_pulse":
        amp = source.get("amplitude", 1.0)
        tt = torch.tensor(t, dtype=F64)
        t0 = torch.tensor(source.get("t0", 3.0 / w), dtype=F64)
        sp = torch.tensor(source.get("spread", 1.0 / w), dtype=F64)
        env = torch.exp(-((tt - t0) ** 2) / (2 * sp ** 2))
        return amp * torch.sin(torch.tensor(w * t)) * env
    if kind == "continuous":
        amp = source.get("amplitude", 1.0)
        return torch.tensor(amp * min(1.0, t * w / (4 * math.pi)) * math.sin(w * t), dtype=F64)
    return torch.tensor(0.0, dtype=F64)


def _zeros(p):
    return [torch.zeros(p.shape, dtype=F64) for _ in range(6)]


def time_reversal_forward(p: RefPath3D, eps_r, source, mask, n_steps):
    """Detached forward with per-step E snapshots; returns the field
    (3, D, H, W) = [Ez, Ex, Ey] and the snapshot lists."""
    Ex, Ey, Ez, Hx, Hy, Hz = _zeros(p)
    mu = torch.ones_like(eps_r)
    e = eps_r.detach()
    snaps = ([], [], [], [])
    for n in range(n_steps):
        Ex, Ey, Ez, Hx, Hy, Hz = ref_time_step(p, Ex, Ey, Ez, Hx, Hy, Hz, e, mu)
        wv = ref_waveform(p, n, source)
        Ez = Ez + mask * wv
        snaps[0].append(Ex.clone())
        snaps[1].append(Ey.clone())
        snaps[2].append(Ez.clone())
        snaps[3].append(wv.item())
    return torch.stack([Ez, Ex, Ey], dim=0), snaps


def time_reversal_backward(p: RefPath3D, eps_r, source, mask, n_steps, grad_output):
    """dL/deps: the forward re-run from zero fields with autograd on eps,
    then ``torch.autograd.grad`` of the final field against grad_output."""
    with torch.enable_grad():
        e = eps_r.detach().clone().requires_grad_(True)
        mu = torch.ones_like(e)
        Ex, Ey, Ez, Hx, Hy, Hz = _zeros(p)
        for n in range(n_steps):
            Ex, Ey, Ez, Hx, Hy, Hz = ref_time_step(p, Ex, Ey, Ez, Hx, Hy, Hz, e, mu)
            Ez = Ez + mask * ref_waveform(p, n, source)
        final = torch.stack([Ez, Ex, Ey], dim=0)
        return torch.autograd.grad(final, e, grad_output)[0]


def source_mask(shape, source):
    """Dense 0/1 mask of a point / plane / default-line source (fdtd3d.py:340-372)."""
    from oracle.fdtd_oracle import source_mask_3d
    return torch.as_tensor(source_mask_3d(shape, source), dtype=F64)


# ---------------------------------------------------------------------------
# 2D TM (fdtd2d.py:233-267, run loop :309-337; gradient by plain autograd)
# ---------------------------------------------------------------------------


class RefPath2D:
    """Grid, step and profiles of one reference-path 2D TM solve."""

    def __init__(self, grid_shape, dl=20.0, wavelength_nm=1550.0, pml_layers=10, courant=0.5):
        H, W = grid_shape
        self.shape = (H, W)
        self.dl = float(dl)
        self.dt = courant * dl / math.sqrt(2.0)
        self.omega = 2 * math.pi / wavelength_nm
        t = lambda a: torch.as_tensor(a, dtype=F64)  # noqa: E731
        by, cy = cpml_profile(H, pml_layers, dl, self.dt)
        bx, cx = cpml_profile(W, pml_layers, dl, self.dt)
        # "y" profiles vary along axis 0 (rows), "x" profiles along axis 1
        self.by, self.cy = t(by).unsqueeze(1), t(cy).unsqueeze(1)
        self.bx, self.cx = t(bx).unsqueeze(0), t(cx).unsqueeze(0)


def _difference2(f, axis, dl, at_low):
    d = torch.zeros_like(f)
    n = f.shape[axis]
    up = [slice(None)] * 2
    dn = [slice(None)] * 2
    up[axis], dn[axis] = slice(1, n), slice(0, n - 1)
    d[tuple(dn if at_low else up)] = (f[tuple(up)] - f[tuple(dn)]) / dl
    return d


def ref_time_step_tm(p: RefPath2D, Ez, Hx, Hy, eps_r, mu_r):
    """One TM step: Hx (difference of Ez along axis 1, damped by the axis-0
    profile), Hy (along axis 0, axis-1 profile), then Ez."""
    dl, dt = p.dl, p.dt
    d_ez_1 = _difference2(Ez, 1, dl, True)
    Hx = p.by * Hx - (dt / mu_r) * (p.cy * d_ez_1 + d_ez_1)
    d_ez_0 = _difference2(Ez, 0, dl, True)
    Hy = p.bx * Hy + (dt / mu_r) * (p.cx * d_ez_0 + d_ez_0)
    d_hy_0 = _difference2(Hy, 0, dl, False)
    d_hx_1 = _difference2(Hx, 1, dl, False)
    Ez = p.bx * p.by * Ez + (dt / eps_r) * (p.cx * d_hy_0 + d_hy_0 - p.cy * d_hx_1 - d_hx_1)
    return Ez, Hx, Hy


def tm_run(p: RefPath2D, eps_r, source, mask, n_steps, line=None, omega=None):
    """Forward from zero fields (autograd records when eps requires grad).
    ``line`` = (row, col0, col1): also the running DFT of Ez on that line at
    ``omega``, F = sum_n Ez^{n+1} exp(-i omega (n+1) dt) as (re, im)."""
    H, W = p.shape
    Ez, Hx, Hy = (torch.zeros(H, W, dtype=F64) for _ in range(3))
    mu = torch.ones_like(eps_r)
    fr = fi = None
    if line is not None:
        fr = torch.zeros(line[2] - line[1], dtype=F64)
        fi = torch.zeros_like(fr)
    for n in range(n_steps):
        Ez, Hx, Hy = ref_time_step_tm(p, Ez, Hx, Hy, eps_r, mu)
        Ez = Ez + mask * ref_waveform(p, n, source)
        if line is not None:
            ph = omega * (n + 1) * p.dt
            seg = Ez[line[0], line[1]:line[2]]
            fr = fr + seg * math.cos(ph)
            fi = fi - seg * math.sin(ph)
    return Ez, (fr, fi)

"""Pin the FP64 oracle: golden vectors from the reference plus invariants.

The golden fixtures were written by ``tests/golden/make_golden.py``, which
runs the reference solver (fdtd2d.py / fdtd3d.py).  The mathematical checks
(energy conservation, causality, mirror symmetry, linearity, adjoint
dot-product, central finite differences) pin what the goldens leave open and
catch a dropped term, a wrong sign or a transposed index independently.
"""

import json
import math

import numpy as np
import torch
import pytest

from oracle import fdtd_oracle as O
from tests._helpers import golden, rel_l2, torch_sin32

# ---------------------------------------------------------------------------
# Tables
# ---------------------------------------------------------------------------


def test_cpml_tables_match_reference():
    t = golden("tables.npz")
    keys = [k[:-2] for k in t.files if k.startswith("cpml") and k.endswith("_b")]
    assert len(keys) >= 8
    for key in keys:
        _, dim, n, L, dl, cour = key.split("_")
        dim = int(dim[0])
        dt = float(cour) * float(dl) / math.sqrt(float(dim))
        b, c = O.cpml_profile(int(n), int(L), float(dl), dt)
        np.testing.assert_array_equal(b, t[key + "_b"], err_msg=key)
        np.testing.assert_allclose(c, t[key + "_c"], rtol=1e-15, atol=0, err_msg=key)


def test_cpml_interior_is_plain_yee():
    b, c = O.cpml_profile(64, 8, 25.0, 0.4 * 25 / math.sqrt(3))
    assert np.all(b[8:56] == 1.0) and np.all(c[8:56] == 0.0)
    assert np.all(b[:8] < 1.0) and np.all(c[:8] < 0.0)
    np.testing.assert_array_equal(b, b[::-1])
    b0, c0 = O.cpml_profile(16, 0, 20.0, 5.0)
    assert np.all(b0 == 1.0) and np.all(c0 == 0.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_waveform_tables_match_reference(dim):
    t = golden("tables.npz")
    srcs = json.loads(str(t["wave_meta"]))
    dl, T = (25.0, 4000) if dim == 3 else (20.0, 2000)
    dt = (0.4 * dl / math.sqrt(3)) if dim == 3 else (0.5 * dl / math.sqrt(2))
    om = 2 * math.pi / 1550.0
    for name, src in srcs.items():
        ref = t[f"wave{dim}d_{name}"]
        exact = O.waveform_table(T, dt, om, src, sin32=torch_sin32)
        # float64 exp of the envelope may differ by one ulp from torch's.
        np.testing.assert_allclose(exact, ref, rtol=1e-14, atol=0, err_msg=name)
        # NumPy's float32 sine: within one float32 ulp of the sine.
        approx = O.waveform_table(T, dt, om, src)
        assert np.max(np.abs(approx - ref)) <= 1.2e-7 * max(1.0, np.abs(ref).max())


def test_source_masks():
    m = O.source_mask_3d((6, 5, 4), {})
    assert m.sum() == 4 and np.all(m[3, 2, :] == 1)
    m = O.source_mask_3d((6, 5, 4), {"pos": [6, 0, 0]})
    assert m.sum() == 0
    m = O.source_mask_3d((6, 5, 4), {"plane": "xz"})
    assert m.sum() == 6 * 4 and np.all(m[:, 2, :] == 1)
    m = O.source_mask_2d((6, 5), {"row": 9})
    assert m.sum() == 0
    m = O.source_mask_2d((6, 5), {"pos": [1, 4]})
    assert m[1, 4] == 1 and m.sum() == 1


# ---------------------------------------------------------------------------
# 3D step and forward vs reference goldens
# ---------------------------------------------------------------------------


def test_step3d_matches_reference():
    g = golden("step3d.npz")
    N = g["eps"].shape[0]
    p = O.Params3D((N, N, N), dl=20.0, pml_layers=3, courant=0.4)
    f = g["fin"]
    E, H = O.step3d((f[0], f[1], f[2]), (f[3], f[4], f[5]), g["eps"], p)
    out = np.stack(E + H)
    assert rel_l2(out, g["fout"]) < 1e-14
    np.testing.assert_allclose(out, g["fout"], rtol=1e-12, atol=1e-13)


def _fwd3d_case(src, N, T=60, L=3, dl=20.0, cour=0.4):
    g = golden("fwd3d.npz")
    p = O.Params3D((N, N, N), dl=dl, pml_layers=L, courant=cour)
    src = src or {"type": "gaussian_pulse"}
    wave = O.waveform_table(T, p.dt, p.omega, src, sin32=torch_sin32)
    mask = O.source_mask_3d(p.shape, src)
    r = O.run3d(g["eps"], p, T, 2, mask, wave)
    Ex, Ey, Ez = r["E"]
    return np.stack([Ez, Ex, Ey])


@pytest.mark.parametrize("name,src", [
    ("line", None),
    ("point", {"type": "gaussian_pulse", "pos": [6, 5, 4], "amplitude": 2.0}),
    ("xy", {"type": "gaussian_pulse", "plane": "xy", "z": 3}),
    ("xz", {"type": "gaussian_pulse", "plane": "xz"}),
    ("yz", {"type": "continuous", "plane": "yz", "x": 2}),
    ("oob", {"type": "gaussian_pulse", "pos": [12, 0, 0]}),
])
def test_forward3d_matches_reference(name, src):
    g = golden("fwd3d.npz")
    out = _fwd3d_case(src, 12)
    ref = g[f"field_{name}"]
    if name == "oob":
        assert np.all(ref == 0) and np.all(out == 0)
        return
    assert rel_l2(out, ref) < 1e-13, name


def test_time_series3d_matches_reference():
    g = golden("fwd3d.npz")
    p = O.Params3D((12, 12, 12), dl=20.0, pml_layers=3, courant=0.4)
    src = {"type": "gaussian_pulse"}
    wave = O.waveform_table(40, p.dt, p.omega, src, sin32=torch_sin32)
    mask = O.source_mask_3d(p.shape, src)
    r = O.run3d(g["eps"], p, 40, 2, mask, wave, store=True)
    ts = np.array([r["hist_E"][n + 1][2][6, 6, 9] for n in range(40)])
    np.testing.assert_allclose(ts, g["ts"], rtol=1e-12, atol=1e-16)


# ---------------------------------------------------------------------------
# 3D gradient (time_reversal mode == autograd) vs reference goldens
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("tag,N,L,T,cour", [("small", 8, 0, 6, 0.35), ("pml", 10, 2, 24, 0.4)])
def test_adjoint3d_matches_reference(tag, N, L, T, cour):
    g = golden("grad3d.npz")
    eps = g[f"{tag}_eps"]
    p = O.Params3D((N, N, N), dl=20.0, pml_layers=L, courant=cour)
    src = {"type": "gaussian_pulse"}
    wave = O.waveform_table(T, p.dt, p.omega, src, sin32=torch_sin32)
    mask = O.source_mask_3d(p.shape, src)
    ones = np.ones(p.shape)
    grad, fw = O.adjoint3d(eps, p, T, [ones, ones, ones], 2, mask, wave)
    Ex, Ey, Ez = fw["E"]
    assert rel_l2(np.stack([Ez, Ex, Ey]), g[f"{tag}_field"]) < 1e-13
    assert rel_l2(grad, g[f"{tag}_grad_sum"]) < 1e-11
    grad, _ = O.adjoint3d(eps, p, T, [2 * Ex, 2 * Ey, 2 * Ez], 2, mask, wave)
    assert rel_l2(grad, g[f"{tag}_grad_energy"]) < 1e-11
    # Point source, loss Ez[c]^2 (test_time_reversal.py:174-221 fixture).
    c = N // 2
    psrc = {"type": "gaussian_pulse", "pos": [c, c, c]}
    pm = O.source_mask_3d(p.shape, psrc)
    fw = O.run3d(eps, p, T, 2, pm, wave)
    lam = [np.zeros(p.shape) for _ in range(3)]
    lam[2][c, c, c] = 2 * fw["E"][2][c, c, c]
    grad, _ = O.adjoint3d(eps, p, T, lam, 2, pm, wave)
    assert rel_l2(grad, g[f"{tag}_grad_point"]) < 1e-10


def _dft3d_setup():
    g = golden("dft3d.npz")
    meta = json.loads(str(g["meta"]))
    D, H, W = meta["shape"]
    p = O.Params3D((D, H, W), dl=meta["dl"], pml_layers=meta["L"], courant=meta["courant"])
    mask = np.zeros((D, H, W))
    mask[:, :, meta["src_x"]] = g["profile"]
    wave = O.waveform_table(meta["T"], p.dt, p.omega, {"type": "gaussian_pulse"}, sin32=torch_sin32)
    omegas = [2 * math.pi / lam for lam in meta["lams"]]
    mon = [{"region": (slice(None), slice(None), meta["mon_x"]), "comps": [0, 1, 2], "omegas": omegas}]
    return g, meta, p, mask, wave, mon


def test_dft_monitor3d_and_gradient_match_reference_harness():
    g, meta, p, mask, wave, mon = _dft3d_setup()
    r = O.run3d(g["eps"], p, meta["T"], 1, mask, wave, monitors=mon)
    F = r["monitors"][0]  # (freq, comp, D, H)
    assert rel_l2(F.real, g["Fr"]) < 1e-12 and rel_l2(F.imag, g["Fi"]) < 1e-12
    # loss = sum |F[:, Ey]|^2  ->  g = 2 F on the Ey slot.
    G = np.zeros_like(F)
    G[:, 1] = 2 * F[:, 1]
    zero = np.zeros(p.shape)
    grad, _ = O.adjoint3d(g["eps"], p, meta["T"], [zero, zero, zero], 1, mask, wave,
                          monitors=mon, mon_grads=[G])
    assert rel_l2(grad, g["grad"]) < 1e-10


# ---------------------------------------------------------------------------
# 2D vs reference goldens
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("pol", ["TM", "TE"])
@pytest.mark.parametrize("name,src", [
    ("point", {"type": "gaussian_pulse", "pos": [9, 7]}),
    ("row", {"type": "gaussian_pulse", "row": 15, "amplitude": 0.5}),
    ("cw", {"type": "continuous"}),
])
def test_fdtd2d_matches_reference(pol, name, src):
    g = golden("fdtd2d.npz")
    eps = g["eps"]
    p = O.Params2D(eps.shape, dl=20.0, polarization=pol, pml_layers=4)
    wave = O.waveform_table(80, p.dt, p.omega, src, sin32=torch_sin32)
    mask = O.source_mask_2d(p.shape, src)
    r = O.run2d(eps, p, 80, mask, wave)
    assert rel_l2(r["field"][None], g[f"{pol}_{name}_field"]) < 1e-13
    grad, _ = O.adjoint2d(eps, p, 80, 2 * r["field"], mask, wave)
    assert rel_l2(grad, g[f"{pol}_{name}_grad"]) < 1e-10


def test_time_series2d_matches_reference():
    g = golden("fdtd2d.npz")
    p = O.Params2D(g["eps"].shape, dl=20.0, pml_layers=4)
    src = {"type": "gaussian_pulse"}
    wave = O.waveform_table(50, p.dt, p.omega, src, sin32=torch_sin32)
    r = O.run2d(g["eps"], p, 50, O.source_mask_2d(p.shape, src), wave, probe=(12, 10))
    np.testing.assert_allclose(r["snapshots"], g["ts"], rtol=1e-12, atol=1e-16)


def test_c0_full_case_matches_reference():
    """200x200 TM, 2000 steps, 10-cell CPML: final Ez, 1x60 DFT line, gradient."""
    g = golden("c0.npz")
    meta = json.loads(str(g["meta"]))
    src = meta["src"]
    T = meta["T"]
    p = O.Params2D((200, 200), dl=20.0, pml_layers=10, courant=0.5)
    wave = O.waveform_table(T, p.dt, p.omega, src, sin32=torch_sin32)
    mask = O.source_mask_2d(p.shape, src)
    mon = [{"region": (160, slice(70, 130)), "omegas": [p.omega]}]
    r = O.run2d(g["eps"], p, T, mask, wave, monitors=mon)
    assert rel_l2(r["field"], g["Ez"]) < 1e-12
    assert rel_l2(r["field"][None], g["field"]) < 1e-12
    F = r["monitors"][0][0]
    assert rel_l2(F.real, g["Fr"]) < 1e-12 and rel_l2(F.imag, g["Fi"]) < 1e-12
    assert abs(np.sum(np.abs(F) ** 2) - float(g["loss"])) < 1e-10 * float(g["loss"])
    grad, _ = O.adjoint2d(g["eps"], p, T, np.zeros(p.shape), mask, wave, monitors=mon,
                          mon_grads=[2 * r["monitors"][0]])
    assert rel_l2(grad, g["grad"]) < 1e-9


# ---------------------------------------------------------------------------
# Invariants that pin the oracle independently of the goldens
# ---------------------------------------------------------------------------


def _compact_init(shape, rng, margin):
    D, H, W = shape
    f = np.zeros(shape)
    f[margin:D - margin, margin:H - margin, margin:W - margin] = rng.standard_normal(
        (D - 2 * margin, H - 2 * margin, W - 2 * margin))
    return f


def test_energy_conserved_without_cpml_3d():
    """Leapfrog invariant Q = <eps E^n, E^n> + <H^n, H^{n+1}> is exact while the
    light cone stays inside the truncated-difference boundaries."""
    rng = np.random.default_rng(0)
    shape = (20, 20, 20)
    p = O.Params3D(shape, dl=20.0, pml_layers=0, courant=0.5)
    eps = 1.5 + rng.random(shape)
    E = tuple(_compact_init(shape, rng, 8) for _ in range(3))
    H = tuple(_compact_init(shape, rng, 8) for _ in range(3))
    qs = []
    for _ in range(5):
        E1, H1 = O.step3d(E, H, eps, p)
        qs.append(sum(np.sum(eps * e * e) for e in E) + sum(np.sum(a * b) for a, b in zip(H, H1)))
        E, H = E1, H1
    qs = np.array(qs)
    assert np.max(np.abs(qs - qs[0])) < 1e-11 * abs(qs[0])


def test_energy_conserved_without_cpml_2d_tm():
    rng = np.random.default_rng(1)
    shape = (30, 26)
    p = O.Params2D(shape, dl=20.0, pml_layers=0, courant=0.6)
    eps = 1.5 + rng.random(shape)
    f = lambda: np.pad(rng.standard_normal((10, 8)), ((10, 10), (9, 9)))  # noqa: E731
    Ez, Hx, Hy = f(), f(), f()
    qs = []
    for _ in range(6):
        Ez1, Hx1, Hy1 = O.step2d_tm(Ez, Hx, Hy, eps, p)
        qs.append(np.sum(eps * Ez * Ez) + np.sum(Hx * Hx1) + np.sum(Hy * Hy1))
        Ez, Hx, Hy = Ez1, Hx1, Hy1
    qs = np.array(qs)
    assert np.max(np.abs(qs - qs[0])) < 1e-11 * abs(qs[0])


def test_causality_and_mirror_symmetry_3d():
    shape = (13, 13, 13)
    p = O.Params3D(shape, dl=20.0, pml_layers=0, courant=0.5)
    src = {"type": "gaussian_pulse", "pos": [6, 6, 6]}
    T = 4
    wave = O.waveform_table(T, p.dt, p.omega, src)
    r = O.run3d(np.full(shape, 2.25), p, T, 2, O.source_mask_3d(shape, src), wave)
    Ex, Ey, Ez = r["E"]
    # Each step moves information by at most one cell per axis (two half
    # updates, each a one-cell stencil) -> support within the L1 ball of
    # radius 2*T around the source.
    zz, yy, xx = np.meshgrid(*[np.arange(13)] * 3, indexing="ij")
    far = (np.abs(zz - 6) + np.abs(yy - 6) + np.abs(xx - 6)) > 2 * T
    assert all(np.all(e[far] == 0) for e in r["E"])
    # A centred Ez point source: Ez is symmetric under z -> -z mirror, since
    # Ez lives on integer z planes.
    np.testing.assert_allclose(Ez, Ez[::-1, :, :], rtol=0, atol=1e-15 * np.abs(Ez).max())


def test_linearity_in_amplitude():
    shape = (10, 10, 10)
    p = O.Params3D(shape, dl=20.0, pml_layers=2, courant=0.4)
    eps = 2.0 + np.random.default_rng(2).random(shape)
    m = O.source_mask_3d(shape, {})
    w1 = O.waveform_table(30, p.dt, p.omega, {"type": "gaussian_pulse"})
    w2 = O.waveform_table(30, p.dt, p.omega, {"type": "continuous"})
    r1 = O.run3d(eps, p, 30, 2, m, w1)["E"]
    r2 = O.run3d(eps, p, 30, 2, m, w2)["E"]
    r3 = O.run3d(eps, p, 30, 2, m, 0.7 * w1 - 1.3 * w2)["E"]
    for a, b, c in zip(r1, r2, r3):
        np.testing.assert_allclose(c, 0.7 * a - 1.3 * b, rtol=1e-10, atol=1e-14)


def test_adjoint3d_dot_product_and_central_fd():
    """<dL/deps, v> equals the central finite difference of L along v."""
    rng = np.random.default_rng(5)
    shape = (8, 9, 10)  # non-cubic: the oracle's defined generalisation
    p = O.Params3D(shape, dl=25.0, pml_layers=2, courant=0.4)
    eps = 2.0 + 3.0 * rng.random(shape)
    src = {"type": "gaussian_pulse", "plane": "yz", "x": 3}
    T = 40
    wave = O.waveform_table(T, p.dt, p.omega, src)
    mask = O.source_mask_3d(shape, src)
    wts = [rng.standard_normal(shape) for _ in range(3)]
    omegas = [2 * math.pi / 1550.0]
    mon = [{"region": (slice(None), slice(None), 7), "comps": [0, 1, 2], "omegas": omegas}]

    def loss(e):
        r = O.run3d(e, p, T, 1, mask, wave, monitors=mon)
        F = r["monitors"][0]
        return sum(np.sum(w * f) for w, f in zip(wts, r["E"])) + np.sum(np.abs(F[0, 1]) ** 2), r

    L0, r = loss(eps)
    G = np.zeros_like(r["monitors"][0])
    G[0, 1] = 2 * r["monitors"][0][0, 1]
    grad, _ = O.adjoint3d(eps, p, T, wts, 1, mask, wave, monitors=mon, mon_grads=[G])
    v = rng.standard_normal(shape)
    h = 1e-5
    fd = (loss(eps + h * v)[0] - loss(eps - h * v)[0]) / (2 * h)
    assert abs(np.sum(grad * v) - fd) < 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_adjoint2d_central_fd(pol):
    rng = np.random.default_rng(6)
    shape = (14, 11)
    p = O.Params2D(shape, dl=20.0, polarization=pol, pml_layers=3)
    eps = 1.5 + 2.0 * rng.random(shape)
    src = {"type": "gaussian_pulse", "pos": [5, 4]}
    wave = O.waveform_table(50, p.dt, p.omega, src)
    mask = O.source_mask_2d(shape, src)
    w = rng.standard_normal(shape)

    def loss(e):
        return np.sum(w * O.run2d(e, p, 50, mask, wave)["field"])

    grad, _ = O.adjoint2d(eps, p, 50, w, mask, wave)
    for _ in range(3):
        v = rng.standard_normal(shape)
        h = 1e-5
        fd = (loss(eps + h * v) - loss(eps - h * v)) / (2 * h)
        assert abs(np.sum(grad * v) - fd) < 1e-6 * max(1.0, abs(fd))


# ---------------------------------------------------------------------------
# FNO oracle: the spectral layer against a brute-force DFT
# ---------------------------------------------------------------------------


def _spectral_brute(h, sr, si, m1, m2):
    """Direct sums, no FFT: forward DFT X(fy, fx) = sum_{y,x} h e^{-2 pi i (fy y/H + fx x/W)}
    at the kept frequencies fy in {0..m1-1} U {-m1..-1} (row index kyi, fy = kyi for
    kyi < m1, else kyi - 2 m1) and fx in {0..m2-1}; channel mixing per mode; real
    inverse over the kept half-spectrum: y = (1/HW) sum c(fx) Re[Y e^{+2 pi i (...)}],
    c(0) = 1, c(fx) = 2 for 0 < fx < W/2 (the Hermitian partner of the column)."""
    B, I, H, W = h.shape
    O = sr.shape[2]
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    out = np.zeros((B, O, H, W))
    for kyi in range(2 * m1):
        fy = kyi if kyi < m1 else kyi - 2 * m1
        for fx in range(m2):
            ph = np.exp(-2j * np.pi * (fy * yy / H + fx * xx / W))
            X = (h * ph[None, None]).sum(axis=(2, 3))             # [B, I]
            Wm = sr[kyi * m2 + fx] + 1j * si[kyi * m2 + fx]       # [I, O]
            Y = X @ Wm                                            # [B, O]
            c = 1.0 if fx == 0 else 2.0
            out += c * np.real(Y[:, :, None, None] * np.conj(ph)[None, None]) / (H * W)
    return out


@pytest.mark.parametrize("H,W,m1,m2", [(8, 8, 2, 3), (6, 10, 3, 2), (9, 7, 1, 3)])
def test_fno_oracle_spectral_layer_matches_brute_force_dft(H, W, m1, m2):
    from oracle import fno_oracle as FO
    rng = np.random.default_rng(H * 100 + W * 10 + m1)
    B, I, Oc = 2, 3, 2
    h = rng.standard_normal((B, I, H, W))
    sr = rng.standard_normal((2 * m1 * m2, I, Oc))
    si = rng.standard_normal((2 * m1 * m2, I, Oc))
    got = FO.spectral_conv(torch.from_numpy(h), torch.from_numpy(sr), torch.from_numpy(si),
                           m1, m2).numpy()
    ref = _spectral_brute(h, sr, si, m1, m2)
    assert np.abs(got - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
    # a transposed mode map (negative rows taken as H - kyi) must not pass
    bad = _spectral_brute(h, sr[::-1].copy(), si[::-1].copy(), m1, m2)
    assert np.abs(got - bad).max() > 1e-3


# ---------------------------------------------------------------------------
# The torch restatement of the reference's CPU path (the CPU baseline)
# ---------------------------------------------------------------------------


def test_torch_path_step_matches_reference_step_golden():
    from oracle import torch_path as TP
    g = golden("step3d.npz")
    p = TP.RefPath3D((10, 10, 10), dl=20.0, pml_layers=3, courant=0.4)
    f = [torch.from_numpy(a) for a in g["fin"]]
    eps = torch.from_numpy(g["eps"])
    out = TP.ref_time_step(p, *f, eps, torch.ones_like(eps))
    assert rel_l2(np.stack([o.numpy() for o in out]), g["fout"]) < 1e-15


@pytest.mark.parametrize("name,src", [("point", {"type": "gaussian_pulse", "pos": [6, 5, 4],
                                                 "amplitude": 2.0}),
                                      ("xy", {"type": "gaussian_pulse", "plane": "xy", "z": 3}),
                                      ("yz", {"type": "continuous", "plane": "yz", "x": 2}),
                                      ("line", {"type": "gaussian_pulse"})])
def test_torch_path_forward_matches_reference_golden(name, src):
    from oracle import torch_path as TP
    g = golden("fwd3d.npz")
    p = TP.RefPath3D((12, 12, 12), dl=20.0, pml_layers=3, courant=0.4)
    field, snaps = TP.time_reversal_forward(p, torch.from_numpy(g["eps"]), src,
                                            TP.source_mask((12, 12, 12), src), 60)
    assert rel_l2(field.numpy(), g[f"field_{name}"]) < 1e-14
    assert len(snaps[0]) == 60 and len(snaps[3]) == 60
    assert torch.equal(snaps[2][-1], field[0])


@pytest.mark.parametrize("tag,N,L,T,cour", [("small", 8, 0, 6, 0.35), ("pml", 10, 2, 24, 0.4)])
def test_torch_path_time_reversal_gradient_matches_reference_golden(tag, N, L, T, cour):
    from oracle import torch_path as TP
    g = golden("grad3d.npz")
    p = TP.RefPath3D((N, N, N), dl=20.0, pml_layers=L, courant=cour)
    src = {"type": "gaussian_pulse"}
    eps = torch.from_numpy(g[f"{tag}_eps"])
    mask = TP.source_mask((N, N, N), src)
    field, _ = TP.time_reversal_forward(p, eps, src, mask, T)
    assert rel_l2(field.numpy(), g[f"{tag}_field"]) < 1e-14
    grad = TP.time_reversal_backward(p, eps, src, mask, T, torch.ones_like(field))
    assert rel_l2(grad.numpy(), g[f"{tag}_grad_sum"]) < 1e-13
    grad2 = TP.time_reversal_backward(p, eps, src, mask, T, 2 * field)
    assert rel_l2(grad2.numpy(), g[f"{tag}_grad_energy"]) < 1e-13


def test_torch_path_tm_matches_reference_goldens():
    from oracle import fdtd_oracle as O
    from oracle import torch_path as TP
    g = golden("fdtd2d.npz")
    p = TP.RefPath2D((24, 20), dl=20.0, pml_layers=4, courant=0.5)
    for name, src in (("point", {"type": "gaussian_pulse", "pos": [9, 7]}),
                      ("cw", {"type": "continuous"})):
        mask = torch.as_tensor(O.source_mask_2d((24, 20), src))
        eps = torch.from_numpy(g["eps"]).clone().requires_grad_(True)
        Ez, _ = TP.tm_run(p, eps, src, mask, 80)
        assert rel_l2(Ez.detach().numpy(), g[f"TM_{name}_field"][0]) < 1e-14
        (Ez ** 2).sum().backward()
        assert rel_l2(eps.grad.numpy(), g[f"TM_{name}_grad"]) < 1e-13
    # the full 200 x 200 / 2000-step case: final Ez and the 1550 nm DFT line
    c = golden("c0.npz")
    meta = json.loads(str(c["meta"]))
    p0 = TP.RefPath2D((200, 200), dl=20.0, pml_layers=10, courant=0.5)
    mask = torch.as_tensor(O.source_mask_2d((200, 200), meta["src"]))
    nt = torch.get_num_threads()
    torch.set_num_threads(1)       # 200 x 200 arrays: threads only add overhead
    try:
        Ez, (fr, fi) = TP.tm_run(p0, torch.from_numpy(c["eps"]), meta["src"], mask, 2000,
                                 line=(160, 70, 130), omega=2 * math.pi / 1550.0)
    finally:
        torch.set_num_threads(nt)
    assert rel_l2(Ez.numpy(), c["Ez"]) < 1e-13
    assert rel_l2(fr.numpy(), c["Fr"]) < 1e-13 and rel_l2(fi.numpy(), c["Fi"]) < 1e-13

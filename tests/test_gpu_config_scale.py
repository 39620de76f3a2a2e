"""GPU parity at the benchmark workloads' full grid sizes.

The toy-grid parity tests pin every kernel variant; these run the kernels
in the configurations the throughput numbers are quoted on:

* 256^3 (16-cell CPML, Ey Gaussian sheet, 5-wavelength DFT plane): all 256
  lanes of every row live, 8 z-chunks, two-step forward / replay / adjoint
  with checkpoint seams and the forward-pair monitor recompute.
  - a short run (T = 12, K = 8) against the FP64 NumPy oracle: fields and
    monitors to 1e-5, dL/deps to 1e-4 in float32; 1e-10 in float64;
  - the full run (T = 8000, K = 64) in float32 against the float64 kernels,
    two float64 directional central differences (h = 1e-3), and the
    reference's own 8000-step forward (tests/golden/c2.npz).
* 128^3 x 4000 steps (configuration 1): float32 and float64 forward +
  adjoint against the reference's forward and autograd gradient
  (tests/golden/c1.npz) and 16 float64 central finite differences
  (h = 1e-3) at seeded design cells.
* One design of every batched grid class at full shape: batched against the
  single-design solver (3000 steps), and float64 against the oracle at
  T = 12 (a monitor plane within the pulse's reach, seeded weights on the
  monitor values and the final field).

Tolerances (relative L2 unless stated): fields / monitors 1e-5 and
gradients 1e-4 in float32, 1e-10 (oracle) or 1e-9 (reference goldens,
which sum in another order over thousands of steps) in float64.
"""

import hashlib
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import fdtd_oracle as O
from tests._helpers import GOLDEN, golden, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need an MI355X")
    import diffnano_amd._lib as L
    L.lib()
    yield


def _say(msg):
    print(msg, flush=True)


def _solve(eps, kw, src, mons, dtype, n_steps=None, ckpt=None, field_w=None, mon_w=None,
           grad=True):
    """One drop-in solve on the GPU: returns (monitor F, field, loss, dL/deps)
    on the host; loss = sum |F|^2, or sum Re(conj(mon_w) F) when mon_w is
    given, plus sum field_w * field."""
    from diffnano_amd import FDTDSolver3D
    kw = dict(kw)
    if n_steps is not None:
        kw["n_steps"] = n_steps
    if ckpt is not None:
        kw["checkpoint_every"] = ckpt
    s = FDTDSolver3D(**kw, device="cuda", dtype=dtype)
    e = torch.tensor(eps, device="cuda", requires_grad=grad)
    with torch.set_grad_enabled(grad):
        r = s.forward(e, source=src, monitors=mons)
        F = r.metadata["monitors"][0]
        if mon_w is None:
            loss = (F.abs() ** 2).sum()
        else:
            loss = (F.real * mon_w.real + F.imag * mon_w.imag).sum()
        if field_w is not None:
            loss = loss + (r.field * field_w).sum()
    g = None
    if grad:
        loss.backward()
        g = e.grad.cpu().numpy()
    out = (F.detach().cpu().numpy(), r.field.detach().cpu().numpy(), float(loss.detach()), g)
    del s, e, r, F, loss
    torch.cuda.empty_cache()
    return out


def _oracle_problem(eps, kw, src, mons):
    p = O.Params3D(kw["grid_shape"], dl=kw["dl"], wavelength_nm=kw["wavelength_nm"],
                   pml_layers=kw["pml_layers"], courant=kw["courant"])
    mask = np.zeros(kw["grid_shape"])
    mask[:, :, src["x"]] = src["profile"].numpy()[:, :, 0]
    z0, z1, y0, y1, x0, x1 = mons[0]["box"]
    mon = {"region": (slice(z0, z1), slice(y0, y1), slice(x0, x1)), "comps": [1],
           "omegas": [2 * math.pi / lam for lam in mons[0]["wavelengths"]]}
    return p, mask, mon


def _eps_sha(eps):
    return hashlib.sha256(np.ascontiguousarray(eps).tobytes()).hexdigest()


# ---------------------------------------------------------------------------
# 256^3: short run against the oracle
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c2_short_oracle():
    """The oracle's forward and adjoint of a linear loss (seeded weights on
    the final field and on the monitor values), 12 steps of the full grid."""
    from diffnano_amd import workloads
    T = 12
    eps, kw, src, mons = workloads.config_c2(n_steps=T)
    # in 12 steps the pulse travels at most 12 cells: the monitor plane sits
    # 4 cells past the source sheet (the output plane would record zeros)
    D, H, _ = kw["grid_shape"]
    xm = src["x"] + 4
    mons = [dict(mons[0], box=(0, D, 0, H, xm, xm + 1))]
    rng = np.random.default_rng(5)
    w = rng.standard_normal((3,) + kw["grid_shape"])          # [Ez, Ex, Ey] field weights
    nf = len(mons[0]["wavelengths"])
    c = rng.standard_normal((nf, 1) + kw["grid_shape"][:2] + (1,)) * (1 + 1j)
    p, mask, mon = _oracle_problem(eps, kw, src, mons)
    wave = O.waveform_table(T, p.dt, p.omega, {"type": "gaussian_pulse"})
    _say(f"c2 short: oracle forward + adjoint, {T} steps of {kw['grid_shape']}")
    lam_T = [w[1], w[2], w[0]]
    grad, fw = O.adjoint3d(eps, p, T, lam_T, 1, mask, wave, [mon], [c])
    Ex, Ey, Ez = fw["E"]
    return dict(eps=eps, kw=kw, src=src, mons=mons, w=w, c=c, F=fw["monitors"][0],
                field=np.stack([Ez, Ex, Ey]), grad=grad, T=T)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("dtype,tf,tg", [(torch.float32, 1e-5, 1e-4),
                                         (torch.float64, 1e-10, 1e-10)])
def test_c2_full_grid_short_run_vs_oracle(c2_short_oracle, dtype, tf, tg):
    o = c2_short_oracle
    wt = torch.from_numpy(o["w"]).cuda()
    ct = torch.from_numpy(o["c"]).cuda()
    _say(f"c2 short: GPU {dtype}")
    F, field, _, g = _solve(o["eps"], o["kw"], o["src"], o["mons"], dtype, ckpt=8, field_w=wt,
                            mon_w=ct)
    assert np.abs(o["F"]).max() > 0 and np.abs(o["field"]).max() > 0
    ef = rel_l2(field, o["field"])
    em = rel_l2(F[:, 0, :, :, 0], o["F"][:, 0, :, :, 0])
    eg = rel_l2(g, o["grad"])
    _say(f"c2 short {dtype}: field {ef:.3e} monitors {em:.3e} grad {eg:.3e} "
         f"(|F| max {np.abs(o['F']).max():.3e})")
    assert ef < tf and em < tf and eg < tg


# ---------------------------------------------------------------------------
# 256^3 x 8000 steps: float32 vs float64 kernels, central differences,
# the reference's own forward
# ---------------------------------------------------------------------------


def _slab_directions(eps_shape):
    D, H, W = eps_shape
    dz, dy, dx = 16, 128, 128
    cz, cy, cx = (D - dz) // 2, (H - dy) // 2, (W - dx) // 2
    rng = np.random.default_rng(17)
    out = []
    for kind in ("sign", "gauss"):
        v = np.zeros(eps_shape)
        r = rng.choice([-1.0, 1.0], (dz, dy, dx)) if kind == "sign" else \
            rng.standard_normal((dz, dy, dx))
        v[cz:cz + dz, cy:cy + dy, cx:cx + dx] = r
        out.append(v)
    return out


@pytest.fixture(scope="module")
def c2_full():
    from diffnano_amd import workloads
    eps, kw, src, mons = workloads.config_c2()
    res = {"eps": eps, "kw": kw, "src": src, "mons": mons}
    for dt in (torch.float32, torch.float64):
        _say(f"c2 full: {kw['n_steps']} steps fwd+adj {dt}")
        res[dt] = _solve(eps, kw, src, mons, dt)
    return res


@pytest.mark.timeout(1200)
def test_c2_full_run_float32_against_float64_kernels(c2_full):
    F32, f32, L32, g32 = c2_full[torch.float32]
    F64, f64, L64, g64 = c2_full[torch.float64]
    em = rel_l2(F32, F64)
    ef = rel_l2(f32, f64)
    eg = rel_l2(g32, g64)
    el = abs(L32 - L64) / abs(L64)
    _say(f"c2 full T=8000: f32 vs f64 monitors {em:.3e} loss {el:.3e} final field {ef:.3e} "
         f"grad {eg:.3e}")
    assert em < 1e-5 and el < 1e-5
    assert eg < 1e-4
    # the final field is the pulse's absorbed tail (~1e-6 of its peak): the
    # float32 rounding of the late steps dominates it, so it is bounded loosely
    assert ef < 1e-2


@pytest.mark.timeout(1200)
def test_c2_full_run_gradient_vs_central_differences(c2_full):
    from diffnano_amd import FDTDSolver3D
    eps, kw, src, mons = c2_full["eps"], c2_full["kw"], c2_full["src"], c2_full["mons"]
    g64 = c2_full[torch.float64][3]
    g32 = c2_full[torch.float32][3]
    h = 1e-3
    s = FDTDSolver3D(**kw, device="cuda", dtype=torch.float64)

    def loss(e):
        with torch.no_grad():
            F = s.forward(torch.tensor(e, device="cuda"), source=src,
                          monitors=mons).metadata["monitors"][0]
        return float((F.abs() ** 2).sum())

    for v in _slab_directions(eps.shape):
        fd = (loss(eps + h * v) - loss(eps - h * v)) / (2 * h)
        a64 = float((g64 * v).sum())
        a32 = float((g32 * v).sum())
        _say(f"c2 directional FD {fd:.9e}  adjoint f64 {a64:.9e} ({abs(a64 - fd) / abs(fd):.2e})"
             f"  f32 {a32:.9e} ({abs(a32 - fd) / abs(fd):.2e})")
        assert abs(a64 - fd) < 1e-5 * abs(fd)
        assert abs(a32 - fd) < 1e-4 * abs(fd)


@pytest.mark.timeout(1200)
def test_c2_full_run_vs_reference_forward(c2_full):
    g = golden("c2.npz")
    meta = json.loads(str(g["meta"]))
    assert _eps_sha(c2_full["eps"]) == meta["eps_sha256"]
    a, b = meta["window"]
    ref = g["Fr_win"] + 1j * g["Fi_win"]
    for dt, tol in ((torch.float64, 1e-9), (torch.float32, 1e-5)):
        F = c2_full[dt][0][:, 0, :, :, 0]
        ew = rel_l2(F[:, a:b, a:b], ref)
        pe = (np.abs(F) ** 2).reshape(F.shape[0], -1).sum(1)
        ep = np.abs(pe - g["plane_energy"]).max() / np.abs(g["plane_energy"]).max()
        el = abs(c2_full[dt][2] - float(g["loss"])) / float(g["loss"])
        _say(f"c2 vs reference ({dt}): window {ew:.3e} plane energy {ep:.3e} loss {el:.3e}")
        assert ew < tol and ep < tol and el < tol


# ---------------------------------------------------------------------------
# 128^3 x 4000 steps (configuration 1)
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c1_full():
    from diffnano_amd import workloads
    eps, kw, src, mons = workloads.config_c1()
    res = {"eps": eps, "kw": kw, "src": src, "mons": mons}
    for dt in (torch.float32, torch.float64):
        _say(f"c1: {kw['n_steps']} steps fwd+adj {dt}")
        res[dt] = _solve(eps, kw, src, mons, dt)
    return res


@pytest.mark.timeout(900)
def test_c1_vs_reference_forward_and_gradient(c1_full):
    g = golden("c1.npz")
    meta = json.loads(str(g["meta"]))
    assert _eps_sha(c1_full["eps"]) == meta["eps_sha256"]
    cz, cy, cx = meta["design_origin"]
    dz, dy, dx = meta["design_shape"]
    w = np.random.default_rng(meta["wsum_seed"]).standard_normal(c1_full["eps"].shape)
    ref_F = g["Fr"] + 1j * g["Fi"]
    for dt, tf, tg in ((torch.float64, 1e-9, 1e-9), (torch.float32, 1e-5, 1e-4)):
        F, field, loss, grad = c1_full[dt]
        em = rel_l2(F[:, 0, :, :, 0], ref_F)
        el = abs(loss - float(g["loss"])) / float(g["loss"])
        eg = rel_l2(grad[cz:cz + dz, cy:cy + dy, cx:cx + dx], g["grad_design"])
        en = abs(np.linalg.norm(grad) - float(g["grad_l2"])) / float(g["grad_l2"])
        ew = abs((grad * w).sum() - float(g["grad_wsum"])) / abs(float(g["grad_wsum"]))
        ey = rel_l2(field[2][grad.shape[0] // 2], g["Ey_mid_z"])
        _say(f"c1 vs reference ({dt}): monitors {em:.3e} loss {el:.3e} grad(design) {eg:.3e} "
             f"|grad| {en:.3e} grad.w {ew:.3e} Ey(z mid) {ey:.3e}")
        assert em < tf and el < tf
        assert eg < tg and en < tg and ew < tg
        assert ey < (1e-9 if dt == torch.float64 else 1e-2)


@pytest.mark.timeout(900)
def test_c1_sixteen_central_differences(c1_full):
    from diffnano_amd import FDTDSolver3D
    eps, kw, src, mons = c1_full["eps"], c1_full["kw"], c1_full["src"], c1_full["mons"]
    D, H, W = eps.shape
    dz, dy, dx = 8, 64, 64
    cz, cy, cx = (D - dz) // 2, (H - dy) // 2, (W - dx) // 2
    g64, g32 = c1_full[torch.float64][3], c1_full[torch.float32][3]
    # 16 seeded probe cells among the design cells with the larger gradients
    # (a cell whose derivative is ~0 checks nothing)
    blk = np.abs(g64[cz:cz + dz, cy:cy + dy, cx:cx + dx])
    cand = np.argwhere(blk > 0.1 * blk.max())
    rng = np.random.default_rng(2024)
    picks = cand[rng.choice(len(cand), 16, replace=False)] + np.array([cz, cy, cx])
    s = FDTDSolver3D(**kw, device="cuda", dtype=torch.float64)
    h = 1e-3

    def loss(e):
        with torch.no_grad():
            F = s.forward(e, source=src, monitors=mons).metadata["monitors"][0]
        return float((F.abs() ** 2).sum())

    base = torch.tensor(eps, device="cuda")
    fd = []
    for (z, y, x) in picks:
        e = base.clone()
        e[z, y, x] += h
        lp = loss(e)
        e[z, y, x] -= 2 * h
        lm = loss(e)
        fd.append((lp - lm) / (2 * h))
    fd = np.array(fd)
    a64 = np.array([g64[tuple(c)] for c in picks])
    a32 = np.array([g32[tuple(c)] for c in picks])
    for c, f, b, d in zip(picks, fd, a64, a32):
        _say(f"c1 FD {tuple(int(v) for v in c)}: {f:.9e}  f64 {b:.9e}  f32 {d:.9e}")
    e64, e32 = rel_l2(a64, fd), rel_l2(a32, fd)
    _say(f"c1 16-probe FD: f64 adjoint {e64:.3e}  f32 adjoint {e32:.3e}")
    assert e64 < 1e-5
    assert e32 < 1e-4
    assert np.all(np.abs(a32 - fd) < 1e-4 * np.abs(fd).max())


# ---------------------------------------------------------------------------
# Batched grid classes at full shape
# ---------------------------------------------------------------------------

CLASSES = [(96, 96, 96), (128, 128, 128), (160, 160, 96), (192, 128, 128)]


@pytest.mark.timeout(900)
def test_batched_grid_classes_full_shape_match_single_design():
    """One design of every grid class at full shape and step count (3000),
    one batched launch sequence against the single-design solver: monitors
    and the FoM gradient."""
    from diffnano_amd import FDTDSolver3D
    from diffnano_amd.batched import BatchedFDTD3D
    from diffnano_amd.inverse_design import DesignProblem
    T, K = 3000, 64
    probs = [DesignProblem(s, 25.0, 12, seed=31 + i) for i, s in enumerate(CLASSES)]
    b = BatchedFDTD3D([p.shape for p in probs], [p.source for p in probs],
                      [p.monitor for p in probs], dl=25.0, pml_layers=12, n_steps=T,
                      courant=0.4, checkpoint_every=K, device="cuda")
    es = [torch.tensor(p.eps_bg, device="cuda", requires_grad=True) for p in probs]
    _say(f"batched classes: {CLASSES}, pairs {b.uses_pairs}, graph {b.uses_graph}")
    outs = b.forward(es)
    sum((F.abs() ** 2).sum() for F in outs).backward()
    for p, Fb, e in zip(probs, outs, es):
        s = FDTDSolver3D(grid_shape=p.shape, dl=25.0, pml_layers=12, n_steps=T, courant=0.4,
                         device="cuda", backward="time_reversal", checkpoint_every=K)
        e1 = torch.tensor(p.eps_bg, device="cuda", requires_grad=True)
        F1 = s.forward(e1, source=p.source, monitors=[p.monitor]).metadata["monitors"][0]
        (F1.abs() ** 2).sum().backward()
        F1n = F1.detach().cpu().numpy()
        assert np.abs(F1n).max() > 0 and float(e1.grad.abs().max()) > 0
        em = rel_l2(Fb.detach().cpu().numpy(), F1n)
        eg = rel_l2(e.grad.cpu().numpy(), e1.grad.cpu().numpy())
        _say(f"batched vs single {p.shape}: monitors {em:.3e} grad {eg:.3e} "
             f"(|F| max {np.abs(F1n).max():.3e})")
        assert em < 1e-6 and eg < 1e-5


@pytest.mark.timeout(900)
@pytest.mark.parametrize("shape", CLASSES)
def test_grid_class_full_shape_float64_vs_oracle(shape):
    """Every grid class at full shape, float64 kernels against the oracle
    (non-cubic classes use each profile along the axis it varies on)."""
    from diffnano_amd.inverse_design import DesignProblem
    T = 12
    pr = DesignProblem(shape, 25.0, 12, seed=7)
    kw = dict(grid_shape=shape, dl=25.0, wavelength_nm=1550.0, pml_layers=12, n_steps=T,
              courant=0.4, backward="time_reversal")
    D, H, W = shape
    xm = pr.source["x"] + 3   # within reach of the pulse in T steps
    mons = [dict(pr.monitor, box=(0, D, 0, H, xm, xm + 1), wavelengths=[1500.0, 1550.0])]
    rng = np.random.default_rng(sum(shape))
    w = rng.standard_normal((3,) + shape)    # [Ez, Ex, Ey] weights on the final field
    c = rng.standard_normal((2, 1, D, H, 1)) * (1 - 0.5j)
    p, mask, mon = _oracle_problem(pr.eps_bg, kw, pr.source, mons)
    wave = O.waveform_table(T, p.dt, p.omega, {"type": "gaussian_pulse"})
    go, fw = O.adjoint3d(pr.eps_bg, p, T, [w[1], w[2], w[0]], 1, mask, wave, [mon], [c])
    Fo = fw["monitors"][0]
    Ex, Ey, Ez = fw["E"]
    assert np.abs(Fo).max() > 0 and np.abs(go).max() > 0
    F, field, _, g = _solve(pr.eps_bg, kw, pr.source, mons, torch.float64, ckpt=4,
                            field_w=torch.from_numpy(w).cuda(), mon_w=torch.from_numpy(c).cuda())
    em = rel_l2(F[:, 0, :, :, 0], Fo[:, 0, :, :, 0])
    ef = rel_l2(field, np.stack([Ez, Ex, Ey]))
    eg = rel_l2(g, go)
    _say(f"class {shape} f64 vs oracle: monitors {em:.3e} field {ef:.3e} grad {eg:.3e}")
    assert em < 1e-10 and ef < 1e-10 and eg < 1e-10


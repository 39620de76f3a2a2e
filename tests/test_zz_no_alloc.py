"""GPU: the allocation-free 3D adjoint entry point
(``dnx_fdtd3d_adjoint_run_noalloc``, fixed device scratch) against
``dnx_fdtd3d_adjoint_run`` with a torch-allocated workspace, through the C
ABI: bitwise equal adjoint states and gradient accumulators on two calls in
a row, the second on a smaller grid and monitor (the scratch keeps the
first call's contents), plus the size limit."""

import ctypes as C
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(shape, mon_box, comp_mask, seed, m_hi=9, m_lo=2):
    from diffnano_amd import FDTDSolver3D
    from diffnano_amd._engine import MonitorSpec, SourceSpec
    dev = torch.device("cuda", 0)
    s = FDTDSolver3D(grid_shape=shape, dl=25.0, pml_layers=3, n_steps=m_hi + 2, courant=0.4,
                     device=dev)
    eng = s._engine(dev)
    rng = np.random.default_rng(seed)
    eng.set_eps(torch.tensor(1.5 + 3.0 * rng.random(shape), device=dev))
    D, H, W = shape
    N = D * H * W
    src = SourceSpec(2, (0, D, 0, H, W // 3, W // 3 + 1),
                     torch.from_numpy(rng.random((D, H, 1))), rng.standard_normal(m_hi + 2))
    spec = MonitorSpec(mon_box, comp_mask, [2 * math.pi / 1550.0, 2 * math.pi / 1400.0])
    grad = torch.from_numpy(rng.standard_normal(spec.out_shape(3))).to(dev)
    srcs = eng.sources([src])
    mons = eng.monitors([spec], m_hi + 2, grads=[grad])
    f32 = dict(device=dev, dtype=torch.float32)
    t = lambda n: torch.from_numpy(rng.standard_normal(n).astype(np.float32)).to(dev)  # noqa
    return {"eng": eng, "srcs": srcs, "mons": mons, "lam": t(6 * N), "S": t(N),
            "e_lo": t(3 * N), "estore": t((m_hi - m_lo - 1) * 3 * N), "m": (m_hi, m_lo),
            "work": torch.empty(6 * N, **f32)}


def _run(entry, case):
    """One adjoint sweep on copies of the case's buffers; returns the final
    adjoint state and S."""
    from diffnano_amd import _lib as L
    eng = case["eng"]
    lam, S, work = case["lam"].clone(), case["S"].clone(), case["work"].clone()
    m_hi, m_lo = case["m"]
    fin = C.c_int32(-1)
    g = eng.grid()
    srcs, mons = case["srcs"], case["mons"]
    common = (g, L.ptr(lam), L.ptr(work), m_hi, m_lo, L.ptr(case["e_lo"]), L.ptr(case["estore"]),
              srcs[0], srcs[1], mons[0], mons[1], L.ptr(S))
    if entry == "workspace":
        n = int(L.lib().dnx_fdtd3d_adjoint_workspace_bytes(g, mons[0], mons[1]))
        ws = torch.empty(max(n, 1), dtype=torch.uint8, device=lam.device)
        rc = L.lib().dnx_fdtd3d_adjoint_run(*common, L.ptr(ws), n, eng.stream(), C.byref(fin))
    else:
        rc = L.lib().dnx_fdtd3d_adjoint_run_noalloc(*common, eng.stream(), C.byref(fin))
    L.check(rc)
    torch.cuda.synchronize()
    assert fin.value in (0, 1)
    return (work if fin.value else lam), S


def test_noalloc_adjoint_matches_workspace_entry_point_twice():
    from diffnano_amd import _lib as L
    cases = [_inputs((20, 24, 96), (0, 20, 0, 24, 70, 71), 7, seed=1),    # 3 components
             _inputs((10, 12, 40), (2, 8, 3, 9, 30, 31), 2, seed=2)]      # smaller, Ey only
    for case in cases:
        eng = case["eng"]
        need = L.lib().dnx_fdtd3d_adjoint_workspace_bytes(eng.grid(), *case["mons"][:2])
        assert need > 0, "the two-step adjoint (which uses the scratch) must run"
        ref_lam, ref_S = _run("workspace", case)
        got_lam, got_S = _run("noalloc", case)
        assert torch.equal(got_lam, ref_lam) and torch.equal(got_S, ref_S)
        assert float(ref_S.abs().max()) > 0 and float(ref_lam.abs().max()) > 0
    assert L.lib().dnx_fdtd3d_adjoint_static_scratch_bytes() == 4 << 20


def test_noalloc_adjoint_rejects_monitors_beyond_its_scratch():
    from diffnano_amd import _lib as L
    shape = (16, 256, 256)                     # 3 components x 1 M cells = 12 MB > 4 MiB
    case = _inputs(shape, (0, 16, 0, 256, 0, 256), 7, seed=3, m_hi=4, m_lo=1)
    eng = case["eng"]
    assert L.lib().dnx_fdtd3d_adjoint_workspace_bytes(eng.grid(), *case["mons"][:2]) > 4 << 20
    fin = C.c_int32(-1)
    rc = L.lib().dnx_fdtd3d_adjoint_run_noalloc(
        eng.grid(), L.ptr(case["lam"]), L.ptr(case["work"]), 4, 1, L.ptr(case["e_lo"]),
        L.ptr(case["estore"]), case["srcs"][0], case["srcs"][1], case["mons"][0],
        case["mons"][1], L.ptr(case["S"]), eng.stream(), C.byref(fin))
    assert rc == -3 and b"beyond the library's fixed" in L.lib().dnx_last_error()

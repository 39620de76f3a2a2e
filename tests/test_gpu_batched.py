"""GPU: the batched multi-design path equals the single-design solver on
each design (forward monitors and gradients), with and without the captured
graph, on uneven grids."""

import numpy as np
import pytest
import torch

from tests._helpers import rel_l2

pytestmark = pytest.mark.gpu


def _cases():
    shapes = [(20, 18, 40), (16, 24, 64), (24, 16, 33)]
    rng = np.random.default_rng(3)
    eps = [2.085 + 10.0 * rng.random(s) for s in shapes]
    srcs = [{"type": "gaussian_pulse", "plane": "yz", "x": 4, "component": "Ey"},
            {"type": "gaussian_pulse", "pos": [8, 12, 20]},
            {"type": "continuous", "plane": "xz", "y": 5, "component": "Ex", "amplitude": 0.5}]
    mons = [{"box": (0, s[0], 0, s[1], s[2] - 5, s[2] - 4), "components": "y",
             "wavelengths": [1500.0, 1550.0]} for s in shapes]
    mons[2] = {"box": (3, 9, 2, 12, 10, 20), "components": "xyz", "wavelengths": [1550.0]}
    return shapes, eps, srcs, mons


def _plain(shape, eps, src, mon, T, K):
    from diffnano_amd import FDTDSolver3D
    s = FDTDSolver3D(grid_shape=shape, dl=25.0, pml_layers=3, n_steps=T, courant=0.4,
                     backward="time_reversal", checkpoint_every=K, device="cuda")
    e = torch.tensor(eps, device="cuda", requires_grad=True)
    F = s.forward(e, source=dict(src), monitors=[mon]).metadata["monitors"][0]
    loss = (F.abs() ** 2).sum()
    loss.backward()
    return F.detach(), e.grad.detach()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batched_equals_single_design(use_graph):
    from diffnano_amd.batched import BatchedFDTD3D
    shapes, eps, srcs, mons = _cases()
    T, K = 50, 8  # 50 steps: six full 8-step segments and a 2-step tail
    b = BatchedFDTD3D(shapes, [dict(s) for s in srcs], mons, dl=25.0, pml_layers=3, n_steps=T,
                      courant=0.4, checkpoint_every=K, device="cuda", use_graph=use_graph)
    assert b.uses_graph == use_graph
    es = [torch.tensor(e, device="cuda", requires_grad=True) for e in eps]
    outs = b.forward(es)
    loss = sum((F.abs() ** 2).sum() for F in outs)
    loss.backward()
    for d in range(len(shapes)):
        F1, g1 = _plain(shapes[d], eps[d], srcs[d], mons[d], T, K)
        assert rel_l2(outs[d].detach().cpu().numpy(), F1.cpu().numpy()) < 1e-6, d
        assert rel_l2(es[d].grad.cpu().numpy(), g1.cpu().numpy()) < 1e-5, d
    # a second pass reuses the plan (and its graphs) with new eps
    es2 = [torch.tensor(e * 1.01, device="cuda", requires_grad=True) for e in eps]
    outs2 = b.forward(es2)
    sum((F.abs() ** 2).sum() for F in outs2).backward()
    F1, g1 = _plain(shapes[1], eps[1] * 1.01, srcs[1], mons[1], T, K)
    assert rel_l2(outs2[1].detach().cpu().numpy(), F1.cpu().numpy()) < 1e-6
    assert rel_l2(es2[1].grad.cpu().numpy(), g1.cpu().numpy()) < 1e-5


def test_autotuner_measures_caches_and_applies(tmp_path, monkeypatch):
    from diffnano_amd import autotune
    monkeypatch.setenv("DNX_AUTOTUNE_CACHE", str(tmp_path / "tiles.json"))
    saved = [autotune.get_tiles(torch.float32, k) for k in (0, 1, 2, 3)]
    try:
        out = autotune.tune((24, 40, 64), steps=4)
        assert out["forward"][0] in (2, 4, 8) and out["adjoint"][0] in (2, 4, 8)
        assert out["forward2"][0] in (0, 2, 4) and out["adjoint2"][0] in (0, 1, 2, 4)
        for k, name in enumerate(("forward", "adjoint", "forward2", "adjoint2")):
            assert autotune.get_tiles(torch.float32, k) == out[name]
        (tmp_path / "tiles.json").write_text("garbage")  # corrupt cache -> re-tune
        out2 = autotune.tune((24, 40, 64), steps=4)
        assert set(out2) == {"forward", "adjoint", "forward2", "adjoint2"}
        assert autotune._load_cache(str(tmp_path / "tiles.json"))
        # a switched-off two-step kernel: the solver runs one-step kernels only
        autotune.set_tiles(torch.float32, 2, 0, 1)
        autotune.set_tiles(torch.float32, 3, 0, 1)
        F0, g0 = _plain((12, 14, 40), 2.0 + np.zeros((12, 14, 40)), {"type": "gaussian_pulse",
                        "pos": [6, 7, 20]}, {"box": (0, 12, 0, 14, 30, 31)}, 10, 4)
        for k, (ty, zc) in enumerate(saved):
            autotune.set_tiles(torch.float32, k, ty, zc)
        F1, g1 = _plain((12, 14, 40), 2.0 + np.zeros((12, 14, 40)), {"type": "gaussian_pulse",
                        "pos": [6, 7, 20]}, {"box": (0, 12, 0, 14, 30, 31)}, 10, 4)
        assert torch.equal(F0, F1) and torch.equal(g0, g1)
    finally:
        for k, (ty, zc) in enumerate(saved):
            autotune.set_tiles(torch.float32, k, ty, zc)


def _batched_run(T, K, opts, e_store_gb=None, dtype=torch.float32, tails=None):
    from diffnano_amd import _lib as L
    from diffnano_amd.batched import BatchedFDTD3D
    shapes, eps, srcs, mons = _cases()
    with L.options(**opts):  # the plan picks its kernels when it is built
        b = BatchedFDTD3D(shapes, [dict(s) for s in srcs], mons, dl=25.0, pml_layers=3,
                          n_steps=T, courant=0.4, checkpoint_every=K, device="cuda",
                          use_graph=True, e_store_gb=e_store_gb, dtype=dtype)
    if tails is not None:
        tails.append(b.tail_segments)
    es = [torch.tensor(e, device="cuda", requires_grad=True) for e in eps]
    outs = b.forward(es)
    fields = [f.clone() for f in b.final_fields()]
    sum((F.abs() ** 2).sum() for F in outs).backward()
    return b.uses_pairs, [o.detach().clone() for o in outs], fields, [e.grad.clone() for e in es]


@pytest.mark.parametrize("T,K", [(50, 8), (37, 6), (9, 2)])
@pytest.mark.parametrize("opts,dtype", [({}, torch.float32), ({"one_step": 1}, torch.float32),
                                        ({}, torch.float64)])
def test_batched_e_store_of_final_segments_changes_nothing(T, K, opts, dtype):
    """A recording forward that keeps E of the final segments (so the
    backward skips their replay) gives bitwise the monitors, final fields
    and gradients of the plan that replays every segment: no store, a
    two-segment store, and the default (every segment but the first)."""
    shapes = _cases()[0]
    tot = sum(int(np.prod(s)) for s in shapes)
    seg_gb = K * 3 * tot * (4 if dtype == torch.float32 else 8) / 1e9
    tails = []
    runs = [_batched_run(T, K, opts, e_store_gb=g, dtype=dtype, tails=tails)
            for g in (0.0, 2.5 * seg_gb, None)]
    nseg = -(-T // K)
    assert tails == [0, min(2, nseg - 1), nseg - 1]
    ref = runs[0]
    for run in runs[1:]:
        assert run[0] == ref[0]
        for a, b in zip(run[1] + run[2] + run[3], ref[1] + ref[2] + ref[3]):
            assert torch.equal(a, b)


@pytest.mark.parametrize("T,K", [(50, 8), (37, 6), (21, 4), (9, 2)])
def test_batched_two_step_matches_one_step_bitwise(T, K):
    """The batched two-step kernels (pairs of steps per persistent launch,
    their own buffer order, single steps at odd tails) reproduce the
    one-step batched path bitwise: monitors, final fields and gradients, for
    step counts and checkpoint intervals that leave odd segments."""
    p2, F2, E2, G2 = _batched_run(T, K, {})
    pm, Fm, Em, Gm = _batched_run(T, K, {"adj2": 0})
    p1, F1, E1, G1 = _batched_run(T, K, {"one_step": 1})
    assert p2 == (True, True) and pm == (True, False) and p1 == (False, False)
    for run in ((F2, E2, G2), (Fm, Em, Gm)):
        for a, b in zip(run[0], F1):
            assert torch.equal(a, b)
        for a, b in zip(run[1], E1):
            assert torch.equal(a, b)
        for a, b in zip(run[2], G1):
            assert torch.equal(a, b), (a - b).abs().max()


def test_batched_two_step_switched_off_by_tiles():
    """Tile height 0 for the two-step kernels (the autotuner's choice when
    the one-step kernels are faster) gives a plan on the one-step path."""
    from diffnano_amd import autotune
    saved = [autotune.get_tiles(torch.float32, k) for k in (2, 3)]
    try:
        autotune.set_tiles(torch.float32, 2, 0, 1)
        autotune.set_tiles(torch.float32, 3, 0, 1)
        p0, F0, E0, G0 = _batched_run(21, 4, {})
    finally:
        for k, (ty, zc) in zip((2, 3), saved):
            autotune.set_tiles(torch.float32, k, ty, zc)
    p1, F1, E1, G1 = _batched_run(21, 4, {"one_step": 1})
    assert p0 == (False, False)
    for a, b in zip(F0 + E0 + G0, F1 + E1 + G1):
        assert torch.equal(a, b)


def test_plans_sharing_stores_match_one_plan():
    """Two plans that share one set of checkpoint and E stores (sized for
    the larger, run one after the other) give bitwise the monitors and
    gradients of one plan over all designs; a backward after the other
    plan's forward raises."""
    from diffnano_amd.batched import BatchedFDTD3D, SharedStores
    shapes, eps, srcs, mons = _cases()
    T, K = 37, 6
    kw = dict(dl=25.0, pml_layers=3, n_steps=T, courant=0.4, checkpoint_every=K, device="cuda")

    def run(plan, idx):
        es = [torch.tensor(eps[i], device="cuda", requires_grad=True) for i in idx]
        outs = plan.forward(es)
        sum((F.abs() ** 2).sum() for F in outs).backward()
        return [o.detach().clone() for o in outs], [e.grad.clone() for e in es]

    whole = BatchedFDTD3D(shapes, [dict(s) for s in srcs], mons, **kw)
    F0, G0 = run(whole, [0, 1, 2])
    groups = [[0, 2], [1]]
    cells = max(sum(int(np.prod(shapes[i])) for i in g) for g in groups)
    for gb in (0.0, None):   # replay everything; E of all segments but the first
        st = SharedStores(cells, T, checkpoint_every=K, device="cuda", e_store_gb=gb)
        assert st.tail_segments == (0 if gb == 0.0 else -(-T // K) - 1)
        plans = [BatchedFDTD3D([shapes[i] for i in g], [dict(srcs[i]) for i in g],
                               [mons[i] for i in g], stores=st, **kw) for g in groups]
        F, G = [None] * 3, [None] * 3
        for g, p in zip(groups, plans):
            f, gr = run(p, g)
            for j, i in enumerate(g):
                F[i], G[i] = f[j], gr[j]
        for a, b in zip(F + G, F0 + G0):
            assert torch.equal(a, b)
        # plan 0's backward after plan 1's forward: its checkpoints are gone
        e0 = [torch.tensor(eps[i], device="cuda", requires_grad=True) for i in groups[0]]
        out0 = plans[0].forward(e0)
        run(plans[1], groups[1])
        with pytest.raises(RuntimeError, match="another forward ran"):
            sum((F.abs() ** 2).sum() for F in out0).backward()
    with pytest.raises(ValueError, match="shared stores"):
        BatchedFDTD3D(shapes, [dict(s) for s in srcs], mons, stores=st, **kw)

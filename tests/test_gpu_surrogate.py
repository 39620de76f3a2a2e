"""GPU: the batched 2D TM forward against the single-design solver, and the
surrogate pieces (dataset, training, prescreen loop bookkeeping)."""

import numpy as np
import pytest
import torch

from tests._helpers import rel_l2

pytestmark = pytest.mark.gpu


def _cfg(n_steps):
    from diffnano_amd.workloads import SurrogateConfig
    return SurrogateConfig(n_steps=n_steps)


def test_batched_tm_matches_single_design_solver():
    from diffnano_amd import FDTDSolver2D
    from diffnano_amd.batched2d import BatchedTM2D
    from diffnano_amd.workloads import blob_noise
    cfg = _cfg(500)
    noise = torch.stack([blob_noise(s) for s in (3, 4, 5)])
    eps = cfg.eps(noise)
    mon = {"box": (0, cfg.H, 0, cfg.W), "wavelengths": list(cfg.wavelengths)}
    bt = BatchedTM2D(**cfg.solver, source=cfg.source, monitor=mon, device="cuda")
    F, Ez = bt.forward(eps, return_field=True)
    solver = FDTDSolver2D(**cfg.solver, device="cuda")
    for b in range(3):
        r = solver.forward(eps[b], source=cfg.source, monitors=[mon])
        ref = r.metadata["monitors"][0].cpu().numpy()          # [nf, H, W]
        got = F[b].cpu().numpy()
        assert rel_l2(got, ref) < 1e-5, b
        assert rel_l2(Ez[b].cpu().numpy(), r.field[0].cpu().numpy()) < 1e-6, b


def test_batched_tm_line_monitor_and_chunking():
    from diffnano_amd.batched2d import BatchedTM2D
    from diffnano_amd.workloads import blob_noise
    cfg = _cfg(300)
    eps = cfg.eps(torch.stack([blob_noise(s) for s in range(5)]))
    y0, y1, x0, x1 = cfg.line
    full = BatchedTM2D(**cfg.solver, source=cfg.source, device="cuda",
                       monitor={"box": (0, cfg.H, 0, cfg.W), "wavelengths": [1550.0]})
    line = BatchedTM2D(**cfg.solver, source=cfg.source, device="cuda", max_batch=2,
                       monitor={"box": (y0, y1, x0, x1), "wavelengths": [1550.0]})
    a = full.forward(eps)[:, :, y0:y1, x0:x1]
    b = line.forward(eps)
    assert torch.equal(a, b)
    with pytest.raises(RuntimeError):
        line.forward(torch.ones(2, 64, 64))


def test_surrogate_fit_and_prescreen_bookkeeping():
    from diffnano_amd.surrogate import FieldSurrogate, PrescreenDesigner, generate_dataset
    cfg = _cfg(400)
    noise, eps, fields = generate_dataset(cfg, range(96), device="cuda", batch=64)
    assert fields.shape == (96, 3, cfg.H, cfg.W) and fields.dtype == torch.complex64
    sur = FieldSurrogate(cfg, device="cuda", layers=2)
    hist = sur.fit(eps, fields, epochs=4, batch=16, lr=3e-3)
    assert hist[-1] < hist[0]
    err = sur.rel_error(eps[:8], fields[:8])
    assert err.shape == (8,) and torch.isfinite(err).all()
    d = PrescreenDesigner(cfg, sur, n_candidates=32, top_k=4, max_rel_err=0.0, seed=1)
    f0 = d.parent_fom
    # anchored estimate of the parent itself: its own FDTD figure of merit
    assert float(d._predict_fom(cfg.eps(d.parent))[0]) == f0
    d.run(2)
    # max_rel_err 0: every step falls back to evaluating all candidates
    assert d.stats["fallbacks"] == 2 and d.stats["fdtd_solves"] == 1 + 2 * (4 + 32)
    assert d.parent_fom >= f0 and len(d.history) == 3
    # the parent's line field follows the accepted design
    assert float((d.parent_line.abs() ** 2).sum()) == pytest.approx(d.parent_fom, rel=1e-6)
    r = PrescreenDesigner(cfg, None, mode="random", n_candidates=32, top_k=4, seed=1)
    r.run(1)
    assert r.stats["fdtd_solves"] == 1 + 4


def test_ablation_network_ranking_beats_random_selection():
    """What the trained network gains over the method without it: for the
    candidates of a design step (256 mutations of a parent), the 8 the FNO
    ranks highest have a higher true (FDTD) figure of merit than 8 picked at
    random, whose expectation is the mean over all 256."""
    from diffnano_amd.surrogate import FieldSurrogate, PrescreenDesigner, generate_dataset
    cfg = _cfg(2000)
    _, eps, fields = generate_dataset(cfg, range(4096), device="cuda")
    sur = FieldSurrogate(cfg, device="cuda")
    sur.fit(eps, fields, epochs=8, batch=64, lr=3e-3)
    del eps, fields
    lifts, best_pct, lifts_a, err, err_a = [], [], [], [], []
    for seed in (101, 202, 303):
        d = PrescreenDesigner(cfg, sur, mode="surrogate", seed=seed, device="cuda")
        cand = cfg.eps(d._candidates())
        true_fom, _ = d._evaluate(cand)
        pred_fom = cfg.fom(sur.predict(cand)).double()
        top = torch.argsort(pred_fom, descending=True)[:8]
        lifts.append(float(true_fom[top].mean() / true_fom.mean()))
        best_pct.append(float((true_fom < true_fom[top].max()).double().mean()))
        err.append(float(((pred_fom - true_fom).abs() / true_fom).mean()))
        # anchored on the parent's FDTD field (the prescreen loop's default)
        pa = d._predict_fom(cand)
        ta = torch.argsort(pa, descending=True)[:8]
        lifts_a.append(float(true_fom[ta].mean() / true_fom.mean()))
        err_a.append(float(((pa - true_fom).abs() / true_fom).mean()))
    print("top-8 mean FoM / all-candidate mean:", lifts, "best-of-8 percentile:", best_pct)
    print("anchored: top-8 lift", lifts_a, "FoM rel. error", err_a, "unanchored", err)
    assert np.mean(lifts) > 1.0
    assert np.mean(best_pct) > 0.5
    assert np.mean(lifts_a) > 1.0


@pytest.mark.parametrize("shape,src,mon", [
    ((128, 128), {"type": "gaussian_pulse"}, None),
    ((100, 70), {"type": "gaussian_pulse", "pos": [37, 11]}, {"box": (10, 90, 60, 61)}),
    ((33, 128), {"type": "continuous", "pos": [16, 100]}, {"box": (0, 33, 0, 128),
                                                          "wavelengths": [1500.0, 1600.0]}),
    ((128, 5), {"type": "gaussian_pulse", "pos": [64, 2]}, {"box": (120, 128, 0, 5)}),
    ((17, 19), {"type": "gaussian_pulse", "pos": [16, 18]}, {"box": (0, 17, 18, 19)}),
])
def test_resident_batched_tm_matches_per_step_launches(shape, src, mon, lib_opts):
    """The resident kernel (one work-group per design for the whole call,
    grids up to 128 x 128) gives bitwise the final fields and FP32 DFT of
    the two-launches-per-step kernels, and so does the library's pick
    between them: full-grid and line monitors, sources at the grid's last
    row and column, ragged rows and columns."""
    from diffnano_amd.batched2d import BatchedTM2D
    g = torch.Generator().manual_seed(shape[0] * 1000 + shape[1])
    eps = 1.5 + 10.0 * torch.rand(6, *shape, generator=g, dtype=torch.float64)
    outs = []
    for variant in (1, 2, 0):   # per-step launches, resident, the library's pick
        lib_opts.set("fdtd2d_per_step", variant)
        bt = BatchedTM2D(grid_shape=shape, dl=20.0, pml_layers=4, n_steps=300, courant=0.5,
                         source=src, monitor=mon, device="cuda")
        outs.append(bt.forward(eps, return_field=True))
        torch.cuda.synchronize()
    (F0, E0) = outs[0]
    assert float(E0.abs().max()) > 0.0
    for F1, E1 in outs[1:]:
        assert torch.equal(F0, F1) and torch.equal(E0, E1)

"""Benchmark: 3D Yee FDTD forward + time-reversal adjoint on one MI355X.

Workload (default): 256^3 cells, dl 25 nm, 16-cell CPML, 8000 steps,
checkpoint every 64 steps; 128x128x16 random-rho design slab between input
and output Si waveguides; Ey Gaussian sheet source with a broadband pulse
at 1550 nm; 5-wavelength DFT of Ey on the output plane; loss = sum |DFT|^2.
One bench "step" = forward + loss.backward() (dL/d eps over the full grid).

Prints one JSON line.  value = Mcell-updates/s of fwd+adj = N_cells * T /
(t_fwd + t_adj).  roofline: the dominant phase's algorithmic HBM bytes per
cell-update (forward 52 B, replay 64 B, adjoint 72 B, FP32) times the cell
updates it performs, over its HIP-event time, against 8 TB/s.
cpu_baseline: the FP64 NumPy oracle (a port of the reference's CPU path),
fwd+adj over a few steps of the same 256^3 grid on the host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES = {"forward": 52, "replay": 64, "adjoint": 72}
METRIC = "Mcell-updates/s fwd+adjoint, %HBM roofline; designs/s batched vs plain @1e-4"


def _cpu_baseline(eps, solver_kw, steps):
    """FP64 oracle fwd+adj over `steps` steps of the full grid (host cores)."""
    from oracle import fdtd_oracle as O

    shape = solver_kw["grid_shape"]
    p = O.Params3D(shape, dl=solver_kw["dl"], pml_layers=solver_kw["pml_layers"],
                   courant=solver_kw["courant"])
    src = {"type": "gaussian_pulse", "plane": "yz", "x": solver_kw["pml_layers"] + 4}
    wave = O.waveform_table(steps, p.dt, p.omega, src)
    mask = O.source_mask_3d(shape, src)
    lam = [np.ones(shape) for _ in range(3)]
    t0 = time.perf_counter()
    O.adjoint3d(eps, p, steps, lam, 1, mask, wave)
    t = time.perf_counter() - t0
    n = int(np.prod(shape))
    return {"value": round(n * steps / t / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"FP64 NumPy oracle fwd+adj, {steps} steps of the full {shape} grid, "
                      f"{t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--n-steps", type=int, default=None, help="time steps (default: config's)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding of the FDTD "
                                   "domain is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from diffnano_amd import FDTDSolver3D, workloads
    from diffnano_amd._engine import PhaseTimer, set_phase_timer

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    cfg = workloads.config_c2 if args.config == "c2" else workloads.config_c1
    kw = {} if args.n_steps is None else {"n_steps": args.n_steps}
    eps_np, solver_kw, src, mons = cfg(seed=0, **kw)
    solver = FDTDSolver3D(**solver_kw, device=dev)
    eps0 = torch.tensor(eps_np, dtype=torch.float64, device=dev)
    D, H, W = solver_kw["grid_shape"]
    N = D * H * W
    T = solver_kw["n_steps"]

    def one_step():
        eps = eps0.clone().requires_grad_(True)
        res = solver.forward(eps, source=src, monitors=mons)
        loss = workloads.monitor_loss(res.metadata["monitors"])
        loss.backward()
        return loss, eps.grad

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    timer = PhaseTimer()
    set_phase_timer(timer)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, grad = one_step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    set_phase_timer(None)
    phases = timer.totals_ms()
    ms_per_step = wall / args.steps * 1e3
    value = N * T / (wall / args.steps) / 1e6

    # Per-phase algorithmic bandwidth; the dominant phase is the roofline row.
    updates = {"forward": N * T * args.steps, "replay": N * (T - (T + 63) // 64) * args.steps,
               "adjoint": N * (T - 1) * args.steps}
    gbs = {k: BYTES[k] * updates[k] / (phases[k] / 1e3) / 1e9 for k in phases if phases[k] > 0}
    dom = max(phases, key=lambda k: phases[k])
    total_bytes = sum(BYTES[k] * updates[k] for k in BYTES)
    overall_gbs = total_bytes / (sum(phases.values()) / 1e3) / 1e9

    cpu = _cpu_baseline(eps_np, solver_kw, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-rho design slab, waveguides)",
        "config": {"workload": f"3D Yee FDTD {D}x{H}x{W}, {T} steps, "
                               f"{solver_kw['pml_layers']}-cell CPML, fwd+time-reversal adjoint, "
                               f"checkpoint every {solver_kw['checkpoint_every']}",
                   "grid": [D, H, W], "n_steps": T, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": f"{dom} phase", "achieved": round(gbs[dom], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs[dom] / HBM_PEAK_GBS, 4), "traffic": None},
        "phases_ms_per_step": {k: round(v / args.steps, 2) for k, v in phases.items()},
        "phase_gbs": {k: round(v, 1) for k, v in gbs.items()},
        "fwd_adj_roofline_frac": round(overall_gbs / HBM_PEAK_GBS, 4),
        "fwd_mcells": round(N * T / (phases["forward"] / args.steps / 1e3) / 1e6, 2),
        "loss": float(loss.item()),
        "grad_l2": float(grad.norm().item()),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

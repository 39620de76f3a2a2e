"""Adam inverse-design loop over many 3D designs, batched or one at a time.

The driver follows the reference's FDTD-capable design loop
(``WaveguideDesigner.optimize_bend``, waveguide.py:269-378: density ->
permittivity -> solver.forward -> loss -> backward -> Adam step -> clamp),
with the batched solver in place of per-design solves.

Problem per design (the 3D batched configuration): a straight Si waveguide
along x in SiO2, a centred (8, H/2, W/2) design block with
eps = 2.085 + 10.015 * rho, rho in [0, 1] (seeded uniform start), an Ey
Gaussian sheet source at x = L + 4 and a DFT monitor of Ey on the output
plane x = W - L - 4; the figure of merit is sum |DFT Ey(lambda0)|^2
(maximised; the loss is its negative).
"""

from __future__ import annotations

import time

import numpy as np
import torch

from . import workloads
from ._common import spare_device_bytes
from .batched import BatchedFDTD3D, SharedStores
from .fdtd3d import FDTDSolver3D

GRID_CLASSES = [(96, 96, 96), (128, 128, 128), (160, 160, 96), (192, 128, 128)]


class DesignProblem:
    def __init__(self, shape, dl, pml_layers, seed):
        self.shape = tuple(shape)
        D, H, W = self.shape
        dz, dy, dx = 8, H // 2, W // 2
        eps = workloads.waveguide_3d(self.shape, dl, design=(dz, dy, dx), seed=seed)
        cz, cy, cx = (D - dz) // 2, (H - dy) // 2, (W - dx) // 2
        self.box = (slice(cz, cz + dz), slice(cy, cy + dy), slice(cx, cx + dx))
        self.rho0 = (eps[self.box] - workloads.EPS_SIO2) / (workloads.EPS_SI - workloads.EPS_SIO2)
        self.eps_bg = eps
        self.source = {"type": "gaussian_pulse", "plane": "yz", "x": pml_layers + 4,
                       "component": "Ey",
                       "profile": torch.from_numpy(workloads.gaussian_sheet(self.shape, dl))}
        xm = W - pml_layers - 4
        self.monitor = {"box": (0, D, 0, H, xm, xm + 1), "components": ["Ey"],
                        "wavelengths": [1550.0]}

    def eps(self, rho, eps_bg_dev):
        e = eps_bg_dev.clone()
        e[self.box] = workloads.EPS_SIO2 + (workloads.EPS_SI - workloads.EPS_SIO2) * rho.clamp(0, 1)
        return e


def batched_problems(n_designs=48, seed=0, dl=25.0, pml_layers=12, classes=GRID_CLASSES):
    """Designs with grid classes drawn by seed."""
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(classes), n_designs)
    return [DesignProblem(classes[int(c)], dl, pml_layers, seed * 100003 + d)
            for d, c in enumerate(picks)]


def _group_plans(problems, group, dl, pml_layers, n_steps, checkpoint_every, dev, use_graph):
    """Batched plans over consecutive groups of designs (one plan for all
    when ``group`` is None).  Grouped plans run one forward + backward at a
    time and share one set of checkpoint and E stores (``SharedStores``),
    sized for the largest group, with the device memory left after every
    plan's own buffers given to E of the final segments."""
    n = len(problems)
    if not group:
        return [(list(range(n)), BatchedFDTD3D(
            [p.shape for p in problems], [p.source for p in problems],
            [p.monitor for p in problems], dl=dl, pml_layers=pml_layers, n_steps=n_steps,
            courant=0.4, checkpoint_every=checkpoint_every, device=dev, use_graph=use_graph))]
    cells = [int(np.prod(p.shape)) for p in problems]
    K = max(2, int(checkpoint_every) + (int(checkpoint_every) & 1))
    nseg = -(-n_steps // K)
    own = 4 * 28 * sum(cells)          # every plan's states, adjoints, S, ce, eps (bytes)
    shared_per_cell = 4 * (nseg * 6 + max(K - 1, 1) * 3)
    spare = max(0, spare_device_bytes(dev) - (16 << 30) - own)
    g = max(1, min(int(group), n))
    chunks = [list(range(i, min(n, i + g))) for i in range(0, n, g)]
    max_cells = max(sum(cells[i] for i in c) for c in chunks)
    budget = max(0, spare - shared_per_cell * max_cells) * 0.95
    stores = SharedStores(max_cells, n_steps, checkpoint_every=checkpoint_every, device=dev,
                          e_store_gb=budget / 1e9)
    plans = []
    for c in chunks:
        ps = [problems[i] for i in c]
        plans.append((c, BatchedFDTD3D([p.shape for p in ps], [p.source for p in ps],
                                       [p.monitor for p in ps], dl=dl, pml_layers=pml_layers,
                                       n_steps=n_steps, courant=0.4,
                                       checkpoint_every=checkpoint_every, device=dev,
                                       use_graph=use_graph, stores=stores)))
    return plans


def adam_loop(problems, iterations=20, lr=0.05, n_steps=3000, dl=25.0, pml_layers=12,
              checkpoint_every=64, mode="batched", device="cuda", use_graph=True, log=None,
              tiles=None, group=None):
    """Run `iterations` Adam steps on every design.  Returns dict with the
    FoM history (iterations x designs), final rho, wall time (s) and
    design-iterations per second.  ``tiles``: batched mode, one autotuner
    result ({"forward": (ty, zc), "adjoint": ..., "forward2": ...,
    "adjoint2": ...}) applied before the plan is built; plain mode, a dict
    shape -> such a result applied before each design's solve.  ``group``:
    batched mode, designs per batched plan (None: one plan for all); the
    plans share their checkpoint and E stores and run in turn, so the
    checkpoints of a batch too large for the device fit, and smaller groups
    leave room for E of more final segments, whose replay the backward then
    skips."""
    from . import autotune

    dev = torch.device(device)
    if mode == "batched" and tiles:
        autotune.apply_tiles(torch.float32, tiles)
    rhos = [torch.tensor(p.rho0, dtype=torch.float64, device=dev, requires_grad=True)
            for p in problems]
    bgs = [torch.tensor(p.eps_bg, dtype=torch.float64, device=dev) for p in problems]
    opt = torch.optim.Adam(rhos, lr=lr)
    hist = []
    if mode == "batched":
        plans = _group_plans(problems, group, dl, pml_layers, n_steps, checkpoint_every, dev,
                             use_graph)
    else:
        solvers = {}
        for p in problems:
            if p.shape not in solvers:
                solvers[p.shape] = FDTDSolver3D(grid_shape=p.shape, dl=dl, pml_layers=pml_layers,
                                                n_steps=n_steps, courant=0.4, device=dev,
                                                backward="time_reversal",
                                                checkpoint_every=checkpoint_every)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for it in range(iterations):
        opt.zero_grad()
        if mode == "batched":
            foms = [None] * len(problems)
            for idx, solver in plans:
                outs = solver.forward([problems[i].eps(rhos[i], bgs[i]) for i in idx])
                fg = [(F.abs() ** 2).sum() for F in outs]
                (-sum(fg)).backward()
                for i, f in zip(idx, fg):
                    foms[i] = f
        else:
            foms = []
            for p, r, bg in zip(problems, rhos, bgs):
                if tiles and p.shape in tiles:
                    autotune.apply_tiles(torch.float32, tiles[p.shape])
                F = solvers[p.shape].forward(p.eps(r, bg), source=p.source,
                                             monitors=[p.monitor]).metadata["monitors"][0]
                fom = (F.abs() ** 2).sum()
                (-fom).backward()
                foms.append(fom)
        opt.step()
        with torch.no_grad():
            for r in rhos:
                r.clamp_(0.0, 1.0)
        hist.append([float(f.detach()) for f in foms])
        if log:
            log(f"{mode} iter {it}: mean FoM {np.mean(hist[-1]):.6e}")
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    return {"fom": np.array(hist), "rho": [r.detach() for r in rhos], "wall_s": wall,
            "design_iters_per_s": len(problems) * iterations / wall}


def final_fom(problems, rhos, n_steps=3000, dl=25.0, pml_layers=12, device="cuda"):
    """FoM of each design at the given rho, single-design solver, no gradient."""
    dev = torch.device(device)
    out = []
    for p, r in zip(problems, rhos):
        s = FDTDSolver3D(grid_shape=p.shape, dl=dl, pml_layers=pml_layers, n_steps=n_steps,
                         courant=0.4, device=dev)
        with torch.no_grad():
            F = s.forward(p.eps(r, torch.tensor(p.eps_bg, device=dev)), source=p.source,
                          monitors=[p.monitor]).metadata["monitors"][0]
        out.append(float((F.abs() ** 2).sum()))
    return np.array(out)

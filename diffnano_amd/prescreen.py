"""Two-stage proxy prescreen, drop-in for DiffNano's ``TwoStageOptimizer``
(``diffnano/solvers/proxy_prescreen.py:147-274``), re-targeted from the
reference's RCWA proxy and RCWA solver to the FNO surrogate and the batched
2D TM FDTD.

Behaviour kept from the reference:

* the objective maps a ``SimResult`` to a scalar tensor and is minimised;
* ``prescreen`` scores every candidate with the proxy, sorts ascending
  (a stable sort: equal scores keep candidate order) and keeps the first
  ``self.top_k`` (proxy_prescreen.py:180-194);
* ``refine`` solves each kept candidate exactly and returns the first one
  with the strictly smallest score, or ``(None, inf)`` for an empty list
  (:196-211);
* ``optimize`` draws ``max(k + 1, n_candidates)`` candidates with
  ``k = min(top_k, self.top_k)``, prescreens, refines, then solves every
  candidate exactly to report ``true_best_objective``, ``exact_hit`` (the
  exact best is among the proxy's picks) and ``hit_rate`` (the share of the
  exact top k among the proxy's picks), with the same result keys and
  timings (:213-274).

What differs: a proxy or solver whose ``batched`` attribute is true takes
the whole stack ``[B, ...]`` of candidates in one call and returns one
``SimResult`` whose field has a leading batch axis; the optimizer splits it
into per-candidate results before calling the objective, so an objective
written for single results runs unchanged.  ``FNOProxy`` (the FNO of
``surrogate.FieldSurrogate`` on the BF16 MFMA kernels) and ``FDTDEvaluator``
(``batched2d.BatchedTM2D``) are such models.  Candidates are identified by
position, not by object identity.
"""

from __future__ import annotations

import time
from typing import Callable

import torch

from ._result import SimResult

__all__ = ["TwoStageOptimizer", "FNOProxy", "FDTDEvaluator"]


def _as_batch(geometry):
    g = torch.as_tensor(geometry)
    single = g.dim() == 2
    return (g.unsqueeze(0) if single else g), single


class FNOProxy:
    """Proxy with the solver's interface: permittivity ``[H, W]`` or
    ``[B, H, W]`` -> ``SimResult`` with the predicted complex DFT of Ez
    (``[n_freq, H, W]`` or ``[B, n_freq, H, W]``) and ``metadata["proxy"]``
    true, as the reference's proxy result (proxy_prescreen.py:140-144)."""

    batched = True

    def __init__(self, surrogate):
        self.surrogate = surrogate
        self.wavelengths = torch.tensor(surrogate.cfg.wavelengths, dtype=torch.float64)

    def __call__(self, geometry):
        return self.forward(geometry)

    def forward(self, geometry):
        eps, single = _as_batch(geometry)
        f = self.surrogate.predict(eps)
        return SimResult(f[0] if single else f, self.wavelengths, {"proxy": True})


class FDTDEvaluator:
    """Exact solver for the prescreen: ``BatchedTM2D.forward`` over a stack of
    permittivity maps in one launch sequence; field = its DFT monitor
    (``[B, n_freq, my, mx]``, or without the batch axis for one map)."""

    batched = True

    def __init__(self, solver):
        self.solver = solver
        self.wavelengths = torch.tensor(solver.wavelengths, dtype=torch.float64)

    def __call__(self, geometry):
        return self.forward(geometry)

    def forward(self, geometry):
        eps, single = _as_batch(geometry)
        f = self.solver.forward(eps)
        meta = {"n_steps": self.solver.n_steps, "dt": self.solver.dt,
                "courant": self.solver.courant, "grid_shape": self.solver.grid_shape}
        return SimResult(f[0] if single else f, self.wavelengths, meta)


def _scores(call, batched, candidates, objective) -> list[float]:
    """Objective value of every candidate under one model."""
    if not candidates:
        return []
    if batched:
        with torch.no_grad():
            res = call(torch.stack([torch.as_tensor(c) for c in candidates]))
        if res.field.shape[0] != len(candidates):
            raise RuntimeError(f"batched model returned {res.field.shape[0]} results for "
                               f"{len(candidates)} candidates")
        return [float(objective(SimResult(res.field[i], res.wavelengths,
                                          dict(res.metadata))).item())
                for i in range(len(candidates))]
    out = []
    for geo in candidates:
        with torch.no_grad():
            res = call(geo)
        out.append(float(objective(res).item()))
    return out


class TwoStageOptimizer:
    """Stage 1: the proxy ranks many candidates; stage 2: the exact solver
    evaluates the proxy's top k (proxy_prescreen.py:147-274)."""

    def __init__(self, proxy, solver, top_k: int = 10, n_candidates: int = 100):
        self.proxy = proxy
        self.solver = solver
        self.top_k = top_k
        self.n_candidates = n_candidates

    # -- the reference's steps --------------------------------------------
    def generate_candidates(self, n: int, geometry_shape: tuple[int, ...] = (5, 20),
                            eps_range: tuple[float, float] = (1.0, 12.0), *,
                            generator: torch.Generator | None = None) -> list[torch.Tensor]:
        """``n`` float64 maps uniform in ``eps_range`` (proxy_prescreen.py:166-178)."""
        lo, hi = eps_range
        return [lo + torch.rand(*geometry_shape, dtype=torch.float64, generator=generator)
                * (hi - lo) for _ in range(n)]

    def _proxy_scores(self, candidates, objective):
        return _scores(self.proxy, getattr(self.proxy, "batched", False), candidates, objective)

    def _exact_scores(self, candidates, objective):
        return _scores(self.solver.forward, getattr(self.solver, "batched", False),
                       candidates, objective)

    def _rank(self, candidates, objective):
        s = self._proxy_scores(candidates, objective)
        return sorted(range(len(candidates)), key=lambda i: s[i])[: self.top_k]

    def prescreen(self, candidates: list[torch.Tensor],
                  objective: Callable[[SimResult], torch.Tensor]) -> list[torch.Tensor]:
        """The ``top_k`` candidates with the smallest proxy objective."""
        return [candidates[i] for i in self._rank(candidates, objective)]

    def refine(self, candidates: list[torch.Tensor],
               objective: Callable[[SimResult], torch.Tensor]) -> tuple[torch.Tensor, float]:
        """Exact solve of each candidate; the first strict minimum wins."""
        best_geo, best = None, float("inf")
        for geo, s in zip(candidates, self._exact_scores(candidates, objective)):
            if s < best:
                best, best_geo = s, geo
        return best_geo, best

    def optimize(self, objective: Callable[[SimResult], torch.Tensor], n_candidates: int = 100,
                 top_k: int = 10, geometry_shape: tuple[int, ...] = (5, 20), *,
                 eps_range: tuple[float, float] = (1.0, 12.0),
                 candidates: list[torch.Tensor] | None = None,
                 generator: torch.Generator | None = None) -> dict:
        """Both stages plus the exhaustive exact check
        (proxy_prescreen.py:213-274).  ``candidates`` replaces the random
        draw (its length then sets the candidate count)."""
        k = min(top_k, self.top_k)
        t0 = time.perf_counter()
        if candidates is None:
            n = max(k + 1, n_candidates)
            candidates = self.generate_candidates(n, geometry_shape, eps_range,
                                                  generator=generator)
        n = len(candidates)
        picked = self._rank(candidates, objective)
        t_proxy = time.perf_counter() - t0

        t1 = time.perf_counter()
        best_geo, best = self.refine([candidates[i] for i in picked], objective)
        t_refine = time.perf_counter() - t1

        t2 = time.perf_counter()
        exact = self._exact_scores(candidates, objective)
        order = sorted(range(n), key=lambda i: exact[i])
        t_full = time.perf_counter() - t2

        picked_set = set(picked)
        return {
            "best_geometry": best_geo,
            "best_objective": best,
            "true_best_objective": exact[order[0]],
            "hit_rate": len(set(order[:k]) & picked_set) / k,
            "exact_hit": order[0] in picked_set,
            "n_candidates": n,
            "top_k": k,
            "proxy_time_s": t_proxy,
            "refine_time_s": t_refine,
            "full_rcwa_time_s": t_full,
            "speedup": t_full / max(t_proxy, 1e-10),
        }

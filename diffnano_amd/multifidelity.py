"""Cost-aware multi-fidelity design: FNO screening, FDTD verification.

Drop-in for the interface of DiffNano's ``diffnano/design/multifidelity.py``
(``FoundryConstraints`` :80-253, ``FidelityOracle`` :256-296,
``MultiFidelityDesigner`` :299-489), re-targeted from RCWA screening plus
FDTD verification to the FNO surrogate (``surrogate.FieldSurrogate``,
BF16 MFMA) as the low fidelity and the batched 2D FDTD
(``batched2d.BatchedTM2D``) as the high fidelity.

Kept from the reference: class and method names, arguments, the returned
keys, the foundry rules (run lengths of the thresholded design along rows
and columns; opening with the minimum feature, then closing with the
minimum space, first along axis 0 and then along axis 1, with windows of
2 * (k // 2) + 1 pixels clipped at the edges and k <= 1 leaving a line
unchanged), the screening score (negative MSE to the target response in
``design``, the response's sum in ``screen_candidates``), the verification
FoM (negative MSE) and when budget is consumed (after the evaluations, only
if affordable).

Differs, for the GPU: the morphology and the rule check are vectorised
over all lines and candidates (torch pooling and NumPy run lengths instead
of per-pixel Python loops), and an evaluator that declares ``batched =
True`` receives the whole candidate stack in one call (one FNO inference,
one batched FDTD launch sequence) instead of one call per candidate.  The
reference's cost model comes from ``diff_surrogate`` (a dependency absent
here); ``CostModel`` below provides the two methods the designer calls
(``can_afford``, ``consume``).
"""

from __future__ import annotations

import math
from collections.abc import Callable
from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F

from . import workloads

__all__ = ["FoundryConstraints", "FidelityOracle", "MultiFidelityDesigner", "CostModel",
           "FNOScreen", "FDTDVerify"]


# ---------------------------------------------------------------------------
# Foundry rules
# ---------------------------------------------------------------------------


def _pool_lines(x: torch.Tensor, k: int, mode: str) -> torch.Tensor:
    """Binary erosion ("min") or dilation ("max") of every row of ``x``
    ([..., n]) with a window of 2 * (k // 2) + 1 pixels clipped at the
    ends; k <= 1 returns ``x`` unchanged."""
    if k <= 1 or x.shape[-1] == 0:
        return x.clone()
    half = k // 2
    b = (x > 0.5).to(torch.float32)
    lead = b.shape[:-1]
    b = b.reshape(-1, 1, b.shape[-1])
    if mode == "max":
        y = F.max_pool1d(b, 2 * half + 1, stride=1, padding=half)
    else:
        y = -F.max_pool1d(-b, 2 * half + 1, stride=1, padding=half)
    return y.reshape(*lead, -1).to(x.dtype)


def _open_close(lines: torch.Tensor, k_feat: int, k_space: int) -> torch.Tensor:
    opened = _pool_lines(_pool_lines(lines, k_feat, "min"), k_feat, "max")
    return _pool_lines(_pool_lines(opened, k_space, "max"), k_space, "min")


def _run_stats(binary: np.ndarray, k_feat: int, k_space: int) -> tuple[int, int]:
    """(violations, runs) over the rows of a 2D boolean array: every maximal
    run is one check; a run of 1s shorter than k_feat or of 0s shorter than
    k_space is a violation."""
    n, m = binary.shape
    if n == 0 or m == 0:
        return 0, 0
    start = np.ones((n, m), dtype=bool)
    start[:, 1:] = binary[:, 1:] != binary[:, :-1]
    idx = np.flatnonzero(start.ravel())
    lengths = np.diff(np.append(idx, n * m))     # every row begins a run
    vals = binary.ravel()[idx]
    bad = np.where(vals, lengths < k_feat, lengths < k_space)
    return int(bad.sum()), int(idx.size)


@dataclass
class FoundryConstraints:
    """Minimum feature and spacing rules on a pixel grid."""

    min_feature_nm: float = 40.0
    min_space_nm: float = 40.0
    pixel_size_nm: float = 1.0

    @property
    def min_feature_px(self) -> int:
        return max(1, math.ceil(self.min_feature_nm / self.pixel_size_nm))

    @property
    def min_space_px(self) -> int:
        return max(1, math.ceil(self.min_space_nm / self.pixel_size_nm))

    def check(self, design: torch.Tensor) -> dict:
        """Run-length check of ``design`` ((H, W) or (1, H, W)) thresholded at
        0.5, along every row and then every column."""
        d = design.detach()
        if d.dim() == 3:
            d = d.squeeze(0)
        b = (d > 0.5).cpu().numpy()
        v0, n0 = _run_stats(b, self.min_feature_px, self.min_space_px)
        v1, n1 = _run_stats(b.T, self.min_feature_px, self.min_space_px)
        viol, total = v0 + v1, n0 + n1
        return {"passed": viol == 0, "violations": viol,
                "violation_rate": viol / max(total, 1)}

    def project_batch(self, designs: torch.Tensor) -> torch.Tensor:
        """``project`` of every design of a stack [n, H, W]."""
        d = designs.detach().clone()
        kf, ks = self.min_feature_px, self.min_space_px
        # lines along axis 0 (columns), then along axis 1 (rows)
        d = _open_close(d.transpose(-1, -2), kf, ks).transpose(-1, -2)
        return _open_close(d, kf, ks).contiguous()

    def project(self, design: torch.Tensor) -> torch.Tensor:
        """Opening with the minimum feature and closing with the minimum
        space, along axis 0 and then axis 1; (H, W) or (1, H, W)."""
        d = design.detach()
        squeeze = d.dim() == 3
        out = self.project_batch(d if squeeze else d.unsqueeze(0))
        return out if squeeze else out[0]


# ---------------------------------------------------------------------------
# Fidelities and cost
# ---------------------------------------------------------------------------


def _evaluate_stack(fn: Callable, designs: torch.Tensor) -> torch.Tensor:
    if getattr(fn, "batched", False):
        return fn(designs)
    return torch.stack([fn(d) for d in designs])


class FidelityOracle:
    """The low (screening) and high (verification) fidelity evaluators.
    The argument names are the reference's; here ``rcwa_fn`` is the low
    fidelity (the FNO surrogate) and ``fdtd_fn`` the high one."""

    def __init__(self, rcwa_fn: Callable[[torch.Tensor], torch.Tensor],
                 fdtd_fn: Callable[[torch.Tensor], torch.Tensor]) -> None:
        self.rcwa_fn = rcwa_fn
        self.fdtd_fn = fdtd_fn

    def evaluate_low(self, design: torch.Tensor) -> torch.Tensor:
        return self.rcwa_fn(design[None])[0] if getattr(self.rcwa_fn, "batched", False) \
            else self.rcwa_fn(design)

    def evaluate_high(self, design: torch.Tensor) -> torch.Tensor:
        return self.fdtd_fn(design[None])[0] if getattr(self.fdtd_fn, "batched", False) \
            else self.fdtd_fn(design)

    def evaluate(self, design: torch.Tensor, fidelity: str = "low") -> torch.Tensor:
        if fidelity == "low":
            return self.evaluate_low(design)
        if fidelity == "high":
            return self.evaluate_high(design)
        raise ValueError(f"fidelity must be 'low' or 'high', got {fidelity!r}")

    def evaluate_batch(self, designs: torch.Tensor, fidelity: str = "low") -> torch.Tensor:
        """Responses of a stack [n, H, W] (one call for batched evaluators)."""
        if fidelity not in ("low", "high"):
            raise ValueError(f"fidelity must be 'low' or 'high', got {fidelity!r}")
        return _evaluate_stack(self.rcwa_fn if fidelity == "low" else self.fdtd_fn, designs)


class CostModel:
    """Budget bookkeeping: a cost per evaluation of each fidelity and a
    total budget.  ``can_afford(f, n)``: spending n more evaluations of f
    stays within the budget; ``consume(f, n)`` records them (and raises
    when they do not fit)."""

    def __init__(self, cost_low: float = 1.0, cost_high: float = 100.0,
                 budget: float = float("inf")) -> None:
        self.costs = {"low": float(cost_low), "high": float(cost_high)}
        self.budget = float(budget)
        self.spent = 0.0
        self.counts = {"low": 0, "high": 0}

    @property
    def remaining(self) -> float:
        return self.budget - self.spent

    def can_afford(self, fidelity: str, n: int = 1) -> bool:
        return self.spent + self.costs[fidelity] * n <= self.budget

    def consume(self, fidelity: str, n: int = 1) -> None:
        if not self.can_afford(fidelity, n):
            raise ValueError(f"budget exceeded: {n} x {fidelity} costs "
                             f"{self.costs[fidelity] * n}, {self.remaining} left")
        self.spent += self.costs[fidelity] * n
        self.counts[fidelity] += n


# ---------------------------------------------------------------------------
# Designer
# ---------------------------------------------------------------------------


class MultiFidelityDesigner:
    """Generate candidates, enforce the foundry rules (and an optional
    quantizer), screen them with the low fidelity, verify the top ones
    with the high fidelity, keep the best verified design."""

    def __init__(self, oracle: FidelityOracle, cost_model: object, foundry: FoundryConstraints,
                 quantizer: torch.nn.Module | None = None) -> None:
        self.oracle = oracle
        self.cost_model = cost_model
        self.foundry = foundry
        self.quantizer = quantizer

    def _apply_quantize(self, design: torch.Tensor) -> torch.Tensor:
        return self.quantizer(design) if self.quantizer is not None else design

    def _enforce_foundry(self, candidates: torch.Tensor) -> tuple[torch.Tensor, list[dict]]:
        projected = self.foundry.project_batch(candidates)
        if self.quantizer is not None:
            projected = torch.stack([self._apply_quantize(d) for d in projected])
        return projected, [self.foundry.check(d) for d in projected]

    def _score_candidates(self, candidates: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
        """Negative MSE of each low-fidelity response to ``target``."""
        resp = self.oracle.evaluate_batch(candidates, "low")
        t = torch.as_tensor(target, dtype=resp.dtype, device=resp.device)
        return -(resp - t).pow(2).flatten(1).mean(1)

    def _spend(self, fidelity: str, n: int) -> None:
        cm = self.cost_model
        if hasattr(cm, "can_afford") and cm.can_afford(fidelity, n):
            cm.consume(fidelity, n)

    @torch.no_grad()
    def screen_candidates(self, candidates: torch.Tensor,
                          n_top: int) -> tuple[torch.Tensor, torch.Tensor]:
        """Foundry-projected candidates ranked by the sum of their
        low-fidelity response; the top ``n_top`` and their scores."""
        projected, _ = self._enforce_foundry(candidates)
        scores = self.oracle.evaluate_batch(projected, "low").flatten(1).sum(1)
        top = torch.topk(scores, min(n_top, projected.shape[0])).indices
        return projected[top.to(projected.device)], scores[top]

    @torch.no_grad()
    def verify_candidates(self, candidates: torch.Tensor,
                          target: torch.Tensor | None = None) -> tuple[torch.Tensor, dict]:
        """High-fidelity responses of every candidate and, with a target,
        their FoMs (negative MSE); the high-fidelity budget is consumed."""
        resp = self.oracle.evaluate_batch(candidates, "high")
        foms = []
        if target is not None:
            t = torch.as_tensor(target, dtype=resp.dtype, device=resp.device)
            foms = [float(v) for v in -(resp - t).pow(2).flatten(1).mean(1)]
        self._spend("high", candidates.shape[0])
        return resp, {"hf_foms": foms}

    @torch.no_grad()
    def design(self, target_response: torch.Tensor, n_initial: int = 50, n_top: int = 5,
               grid_size: int = 16) -> dict:
        """Random candidates (uniform, float64, torch's global generator) ->
        foundry projection -> low-fidelity screen (negative MSE to the
        target) -> high-fidelity verification of the top ``n_top`` -> the
        best verified design."""
        candidates = torch.rand(n_initial, grid_size, grid_size, dtype=torch.float64)
        projected, reports = self._enforce_foundry(candidates)
        scores = self._score_candidates(projected, target_response)
        top = torch.topk(scores, min(n_top, projected.shape[0])).indices
        screened = projected[top.to(projected.device)]
        self._spend("low", n_initial)
        hf_resp, info = self.verify_candidates(screened, target_response)
        if info["hf_foms"]:
            best = int(np.argmax(info["hf_foms"]))
            best_score = torch.tensor(info["hf_foms"][best])
        else:
            best, best_score = 0, scores[top[0]]
        return {"best": screened[best], "best_score": best_score, "candidates": candidates,
                "screened": screened, "screen_scores": scores[top], "hf_responses": hf_resp,
                "foundry_reports": reports}


# ---------------------------------------------------------------------------
# The two fidelities on the surrogate workload
# ---------------------------------------------------------------------------


def design_to_eps(designs: torch.Tensor, cfg) -> torch.Tensor:
    """Density stack [n, S, S] in [0, 1] -> permittivity [n, H, W]: each
    design pixel covers (window / S)^2 cells of the configuration's design
    window (nearest), eps = SiO2 + (Si - SiO2) * density; SiO2 elsewhere."""
    n, S, _ = designs.shape
    size = cfg.size
    if size % S:
        raise ValueError(f"design grid {S} must divide the {size}-cell design window")
    up = designs.to(torch.float64).repeat_interleave(size // S, 1).repeat_interleave(size // S, 2)
    eps = torch.full((n, cfg.H, cfg.W), workloads.EPS_SIO2, dtype=torch.float64,
                     device=designs.device)
    eps[:, cfg.lo:cfg.lo + size, cfg.lo:cfg.lo + size] = \
        workloads.EPS_SIO2 + (workloads.EPS_SI - workloads.EPS_SIO2) * up.clamp(0, 1)
    return eps


def _line_response(fields: torch.Tensor, cfg) -> torch.Tensor:
    """|DFT Ez|^2 on the figure-of-merit line at every wavelength: [n, nf, cells]."""
    y0, y1, x0, x1 = cfg.line
    f = fields[:, :, y0:y1, x0:x1]
    return f.abs().pow(2).flatten(2).to(torch.float64)


class FNOScreen:
    """Low fidelity: the FNO surrogate's predicted line response."""

    batched = True

    def __init__(self, surrogate, cfg):
        self.surrogate, self.cfg = surrogate, cfg

    def __call__(self, designs: torch.Tensor) -> torch.Tensor:
        eps = design_to_eps(designs.to(self.surrogate.dev), self.cfg)
        return _line_response(self.surrogate.predict(eps), self.cfg).cpu()


class FDTDVerify:
    """High fidelity: the batched 2D TM FDTD's line response (a DFT monitor
    on the figure-of-merit line at the configuration's wavelengths)."""

    batched = True

    def __init__(self, cfg, device=None):
        from .batched2d import BatchedTM2D
        y0, y1, x0, x1 = cfg.line
        self.cfg = cfg
        self.fdtd = BatchedTM2D(**cfg.solver, source=cfg.source, device=device,
                                monitor={"box": (y0, y1, x0, x1),
                                         "wavelengths": list(cfg.wavelengths)})

    def __call__(self, designs: torch.Tensor) -> torch.Tensor:
        eps = design_to_eps(designs.to(self.fdtd.dev), self.cfg)
        F_ = self.fdtd.forward(eps)                       # [n, nf, 1, cells]
        return F_.abs().pow(2).flatten(2).to(torch.float64).cpu()

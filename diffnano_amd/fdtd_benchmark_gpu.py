"""GPU benchmark matrix of ``FDTDSolver3D`` over gradient memory strategies,
drop-in for ``FDTDGPUBenchmark`` and its schemas in DiffNano's
``diffnano/solvers/fdtd_benchmark_gpu.py`` (lines 59-398 and 644-870).

* ``GPUMemoryStrategy`` (:59-78) and the keywords each one sets: time
  reversal -> ``backward="time_reversal"``, checkpoint ->
  ``use_checkpoint=True, checkpoint_segments=4``, full autodiff -> the
  default (:242-247).
* ``FDTDBenchmarkConfig`` (:86-113), ``StabilityReport`` (:121-144).
* ``_estimate_peak_mb`` (:166-186): the allocator's peak on the GPU, else
  the reference's storage model of each strategy.
* ``FDTDGPUBenchmark.run_single`` (:214-296): a CPML-free solver (courant
  0.35, dl 20, 1550 nm) on the seeded ε grid, synchronised forward and
  ``field.sum()`` backward times, peak memory; the record keeps the
  gradient until ``run_all`` (:298-338) has turned it into the cosine
  against the full-autodiff run (which runs first per grid).
* ``benchmark_memory_scaling`` (:340-356), ``cpu_fallback_benchmark``
  (:358-398): all strategies on the first grid with at most 20 steps and
  results on the host; valid when every cosine is at least 0.99.
* ``FDTDXCrossValidator`` (:644-870): an injected or self-generated
  reference field and gradient, compared per seed into a ``StabilityReport``.

Here every strategy runs the same hand-written adjoint (the strategies only
choose its checkpoint interval), so the cosines are 1, and "cpu" places
the results on the host while the kernels still run on the GPU (there is
no CPU solver to fall back to).
"""

from __future__ import annotations

import enum
import gc
import time
from dataclasses import dataclass, field as dc_field
from typing import Any

import torch

from .fdtd_benchmark import _cosine_sim, _make_eps_grid

__all__ = ["GPUMemoryStrategy", "FDTDBenchmarkConfig", "StabilityReport", "FDTDGPUBenchmark",
           "FDTDXCrossValidator"]


class GPUMemoryStrategy(enum.Enum):
    TIME_REVERSAL = "time_reversal"
    CHECKPOINT = "checkpoint"
    FULL_AUTODIFF = "full_autodiff"


@dataclass
class FDTDBenchmarkConfig:
    grid_sizes: list[tuple[int, int, int]] = dc_field(
        default_factory=lambda: [(32, 32, 32), (64, 64, 64)])
    memory_strategies: list[GPUMemoryStrategy] = dc_field(
        default_factory=lambda: [GPUMemoryStrategy.TIME_REVERSAL, GPUMemoryStrategy.CHECKPOINT,
                                 GPUMemoryStrategy.FULL_AUTODIFF])
    n_time_steps: int = 100
    device: str = "cpu"


@dataclass
class StabilityReport:
    forward_errors: list[float] = dc_field(default_factory=list)
    gradient_cosines: list[float] = dc_field(default_factory=list)
    memory_used_mb: dict[str, float] = dc_field(default_factory=dict)
    is_valid: bool = False


def _estimate_peak_mb(grid_size: tuple[int, int, int], n_steps: int,
                      strategy: GPUMemoryStrategy, device: str,
                      actual_cuda_allocated: float | None = None) -> float:
    """Measured peak on the GPU; otherwise FP64 storage of the strategy:
    10 x 6 fields x T (full autodiff), 3 x T (time reversal), 4 segments x 6."""
    if device == "cuda" and actual_cuda_allocated is not None:
        return actual_cuda_allocated / 1e6
    D, H, W = grid_size
    cells = D * H * W
    if strategy == GPUMemoryStrategy.FULL_AUTODIFF:
        return 10 * 6 * n_steps * cells * 8 / 1e6
    if strategy == GPUMemoryStrategy.TIME_REVERSAL:
        return 3 * n_steps * cells * 8 / 1e6
    return 4 * 6 * cells * 8 / 1e6


def _sync():
    if torch.cuda.is_available():
        torch.cuda.synchronize()


class FDTDGPUBenchmark:
    """FDTDSolver3D across grid sizes and gradient memory strategies."""

    def __init__(self, config: FDTDBenchmarkConfig) -> None:
        self.config = config
        self._results: list[dict[str, Any]] = []

    @staticmethod
    def detect_gpu() -> bool:
        return torch.cuda.is_available()

    def run_single(self, grid_size: tuple[int, int, int], strategy: GPUMemoryStrategy,
                   device: str) -> dict[str, Any]:
        from .fdtd3d import FDTDSolver3D
        effective = "cpu" if device == "cuda" and not torch.cuda.is_available() else device
        gc.collect()
        if effective == "cuda":
            torch.cuda.empty_cache()
            torch.cuda.reset_peak_memory_stats()
            mem_before = torch.cuda.memory_allocated()
        else:
            mem_before = 0
        kw: dict[str, Any] = {}
        if strategy == GPUMemoryStrategy.TIME_REVERSAL:
            kw["backward"] = "time_reversal"
        elif strategy == GPUMemoryStrategy.CHECKPOINT:
            kw.update(use_checkpoint=True, checkpoint_segments=4)
        solver = FDTDSolver3D(grid_shape=grid_size, dl=20.0, wavelength_nm=1550.0, pml_layers=0,
                              n_steps=self.config.n_time_steps, device=effective, courant=0.35,
                              **kw)
        eps = _make_eps_grid(grid_size, effective).detach().requires_grad_(True)

        _sync()
        t0 = time.perf_counter()
        result = solver.forward(eps)
        _sync()
        fwd_ms = (time.perf_counter() - t0) * 1e3
        loss = result.field.sum()
        t0 = time.perf_counter()
        loss.backward()
        _sync()
        bwd_ms = (time.perf_counter() - t0) * 1e3

        grad = eps.grad.detach().clone() if eps.grad is not None else torch.zeros_like(eps)
        alloc = torch.cuda.max_memory_allocated() - mem_before if effective == "cuda" else None
        peak_mb = _estimate_peak_mb(grid_size, self.config.n_time_steps, strategy, effective,
                                    alloc)
        record = {"grid_size": grid_size, "strategy": strategy.value, "device": effective,
                  "forward_time_ms": round(fwd_ms, 3), "backward_time_ms": round(bwd_ms, 3),
                  "peak_memory_mb": round(peak_mb, 3), "gradient": grad}
        self._results.append(record)
        return record

    def run_all(self) -> list[dict[str, Any]]:
        """Every grid x strategy, full autodiff first per grid (the cosine's
        reference); the records' gradients are replaced by the cosine."""
        self._results = []
        device = self.config.device
        if device == "cuda" and not torch.cuda.is_available():
            device = "cpu"
        for grid_size in self.config.grid_sizes:
            ref_grad = None
            ordered = sorted(self.config.memory_strategies,
                             key=lambda s: 0 if s == GPUMemoryStrategy.FULL_AUTODIFF else 1)
            for strategy in ordered:
                rec = self.run_single(grid_size, strategy, device)
                if strategy == GPUMemoryStrategy.FULL_AUTODIFF:
                    ref_grad = rec["gradient"]
                    rec["gradient_cosine"] = None
                elif ref_grad is not None:
                    rec["gradient_cosine"] = round(_cosine_sim(ref_grad, rec["gradient"]), 6)
                else:
                    rec["gradient_cosine"] = None
                del rec["gradient"]
        return self._results

    def benchmark_memory_scaling(self) -> list[dict[str, Any]]:
        """Peak memory of every grid x strategy."""
        out = []
        for grid_size in self.config.grid_sizes:
            for strategy in self.config.memory_strategies:
                rec = self.run_single(grid_size, strategy, self.config.device)
                out.append({"grid_size": grid_size, "strategy": strategy.value,
                            "peak_memory_mb": rec["peak_memory_mb"]})
        return out

    def cpu_fallback_benchmark(self) -> StabilityReport:
        """All strategies on the first grid (at most 20 steps, results on the
        host); valid when every cosine against full autodiff is >= 0.99."""
        cfg = FDTDBenchmarkConfig(
            grid_sizes=self.config.grid_sizes[:1],
            memory_strategies=[GPUMemoryStrategy.FULL_AUTODIFF, GPUMemoryStrategy.TIME_REVERSAL,
                               GPUMemoryStrategy.CHECKPOINT],
            n_time_steps=min(self.config.n_time_steps, 20), device="cpu")
        results = FDTDGPUBenchmark(cfg).run_all()
        cosines = [r["gradient_cosine"] for r in results if r["gradient_cosine"] is not None]
        memory = {r["strategy"]: r["peak_memory_mb"] for r in results}
        return StabilityReport(forward_errors=[0.0], gradient_cosines=cosines,
                               memory_used_mb=memory,
                               is_valid=bool(cosines) and all(c >= 0.99 for c in cosines))


class FDTDXCrossValidator:
    """Cross-validation against an external solver's field and gradient
    (fdtd_benchmark_gpu.py:644-870).  ``reference_fn(grid_size, n_steps)``
    returns ``{"field": (3, D, H, W), "gradient": (D, H, W)}``; without it
    the reference is this solver itself (a self-consistency run of the
    pipeline).  The ε grid is the seed-42 one for every seed, as in the
    reference."""

    def __init__(self, reference_fn=None) -> None:
        self.reference_fn = reference_fn

    @staticmethod
    def _solve(grid_size, n_steps):
        from .fdtd3d import FDTDSolver3D
        solver = FDTDSolver3D(grid_shape=grid_size, dl=20.0, wavelength_nm=1550.0, pml_layers=0,
                              n_steps=n_steps, device="cpu", courant=0.35)
        eps = _make_eps_grid(grid_size, "cpu").detach().requires_grad_(True)
        result = solver.forward(eps)
        result.field.sum().backward()
        grad = eps.grad.detach().clone() if eps.grad is not None else torch.zeros_like(eps)
        return result.field.detach().clone(), grad

    def generate_synthetic_reference(self, grid_size: tuple[int, int, int],
                                     n_steps: int = 50) -> dict[str, torch.Tensor]:
        field, grad = self._solve(grid_size, n_steps)
        return {"field": field, "gradient": grad}

    def validate_forward(self, result: torch.Tensor, reference: torch.Tensor,
                         rtol: float = 1e-3) -> dict[str, Any]:
        diff = result.detach().to(torch.float64) - reference.detach().to(torch.float64)
        rel = diff.norm().item() / max(reference.detach().to(torch.float64).norm().item(), 1e-30)
        return {"relative_error": rel, "max_absolute_error": diff.abs().max().item(),
                "passed": rel < rtol}

    def validate_gradient(self, grad: torch.Tensor, reference_grad: torch.Tensor,
                          cosine_threshold: float = 0.99) -> dict[str, Any]:
        g = grad.detach().to(torch.float64)
        r = reference_grad.detach().to(torch.float64)
        cos = _cosine_sim(g, r)
        rel = (g - r).norm().item() / max(r.norm().item(), 1e-30)
        return {"cosine_similarity": cos, "relative_error": rel,
                "passed": cos >= cosine_threshold}

    def run_cross_validation(self, n_seeds: int = 3, grid_size: tuple[int, int, int] = (16, 16, 16),
                             n_steps: int = 30) -> StabilityReport:
        """Per seed: reference, this solver's field and gradient of sum(field),
        both compared; valid when every forward error is below 1e-3 and
        every cosine at least 0.99."""
        fwd, cos = [], []
        for seed in range(n_seeds):
            torch.manual_seed(seed * 100 + 42)
            ref = (self.reference_fn(grid_size, n_steps) if self.reference_fn is not None
                   else self.generate_synthetic_reference(grid_size, n_steps))
            torch.manual_seed(seed * 100 + 42)
            field, grad = self._solve(grid_size, n_steps)
            fwd.append(self.validate_forward(field, ref["field"])["relative_error"])
            cos.append(self.validate_gradient(grad, ref["gradient"])["cosine_similarity"])
        memory = {"full_autodiff": _estimate_peak_mb(grid_size, n_steps,
                                                     GPUMemoryStrategy.FULL_AUTODIFF, "cpu")}
        return StabilityReport(forward_errors=fwd, gradient_cosines=cos, memory_used_mb=memory,
                               is_valid=bool(fwd) and all(e < 1e-3 for e in fwd)
                               and all(c >= 0.99 for c in cos))

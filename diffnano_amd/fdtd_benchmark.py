"""FDTD benchmark suite and external cross-validation, drop-in for DiffNano's
``diffnano/solvers/fdtd_benchmark.py``: the same result schemas, suite loop,
summary table and JSON exchange format, driving ``FDTDSolver3D`` on the
hand-written kernels.

* ``BenchmarkConfig`` / ``BenchmarkResult`` (fdtd_benchmark.py:36-56).
* ``FDTDBenchmarkSuite.run`` (:86-181): per grid size one seeded ε (1.5 + U[0, 1)
  with a centred block of 4.0, seed 42, :66-73); per backward mode a solver
  without CPML (courant 0.35, dl 20, 1550 nm), forward timed, ``field.sum()``
  backpropagated and timed, peak memory, and the gradient's cosine against
  the "autograd" mode's (None for that mode itself and for modes run before
  it).  Mode "time_reversal" sets ``backward="time_reversal"``,
  "checkpoint" sets ``use_checkpoint=True, checkpoint_segments=2``, any
  other name runs the default.
* ``summary_table`` (:183-209), ``_estimate_cpu_peak_mb`` (:212-221).
* ``ExternalCrossValidator`` (:224-310): forward relative L2 error against
  an external field, gradient cosine, the seeded test-case description and
  the JSON loader (``{"field": {"shape", "data"}, "gradient": {...}, ...}``).

Differences: the timed regions end with ``torch.cuda.synchronize()`` when a
GPU is present (the kernels run asynchronously, so an unsynchronised clock
would time the launch, not the solve); peak memory on the GPU counts
torch's allocator (the solver's checkpoints and states are torch tensors).
Every backward mode gives the same gradient here (the modes only choose
the adjoint's checkpoint interval), so the cosines are 1 up to rounding.
"""

from __future__ import annotations

import gc
import json
import time
from dataclasses import dataclass, field as dc_field
from pathlib import Path
from typing import Any

import torch

__all__ = ["BenchmarkConfig", "BenchmarkResult", "FDTDBenchmarkSuite", "ExternalCrossValidator"]


@dataclass
class BenchmarkConfig:
    grid_sizes: list[tuple[int, int, int]] = dc_field(
        default_factory=lambda: [(16, 16, 16), (32, 32, 32)])
    n_time_steps: int = 100
    backward_modes: list[str] = dc_field(
        default_factory=lambda: ["time_reversal", "checkpoint", "autograd"])
    device: str = "cpu"


@dataclass
class BenchmarkResult:
    grid_size: tuple[int, int, int]
    backward_mode: str
    forward_time_ms: float
    backward_time_ms: float
    peak_memory_mb: float
    gradient_cosine_vs_autograd: float | None
    device: str


def _resolve_device(requested: str) -> tuple[str, str]:
    if requested == "cuda":
        if torch.cuda.is_available():
            return "cuda", "GPU (ROCm)"
        return "cpu", "CPU (GPU requested but not available)"
    return "cpu", "CPU"


def _make_eps_grid(grid_size: tuple[int, int, int], device: str) -> torch.Tensor:
    """1.5 + U[0, 1) with a centred block of 4.0 (half of each extent), seed 42."""
    D, H, W = grid_size
    torch.manual_seed(42)
    eps = 1.5 + torch.rand(D, H, W, dtype=torch.float64, device=device)
    d, h, w = D // 4, H // 4, W // 4
    if d > 0 and h > 0 and w > 0:
        eps[D // 2 - d:D // 2 + d, H // 2 - h:H // 2 + h, W // 2 - w:W // 2 + w] = 4.0
    return eps


def _cosine_sim(a: torch.Tensor, b: torch.Tensor) -> float:
    fa = a.flatten().to(torch.float64)
    fb = b.flatten().to(torch.float64)
    denom = (fa.norm() * fb.norm()).item()
    if denom < 1e-30:
        return 0.0
    return torch.dot(fa, fb).item() / denom


def _sync():
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def _estimate_cpu_peak_mb(eps: torch.Tensor, n_steps: int,
                          grid_size: tuple[int, int, int]) -> float:
    """Six FP64 field arrays, times 4 when ε requires grad (:212-221)."""
    D, H, W = grid_size
    return 6 * D * H * W * 8 * (4 if eps.requires_grad else 1) / 1e6


class FDTDBenchmarkSuite:
    """Forward/backward time, peak memory and gradient agreement of a 3D
    FDTD class across grid sizes and backward modes."""

    def __init__(self) -> None:
        self._results: list[BenchmarkResult] = []

    def run(self, fdtd3d_class: type, config: BenchmarkConfig) -> list[BenchmarkResult]:
        device, _ = _resolve_device(config.device)
        self._results = []
        for grid_size in config.grid_sizes:
            eps_template = _make_eps_grid(grid_size, device)
            reference_grad: torch.Tensor | None = None
            for mode in config.backward_modes:
                gc.collect()
                if device == "cuda":
                    torch.cuda.empty_cache()
                    torch.cuda.reset_peak_memory_stats()
                    mem_before = torch.cuda.memory_allocated()
                else:
                    mem_before = 0
                kw: dict[str, Any] = {}
                if mode == "time_reversal":
                    kw["backward"] = "time_reversal"
                elif mode == "checkpoint":
                    kw.update(use_checkpoint=True, checkpoint_segments=2)
                solver = fdtd3d_class(grid_shape=grid_size, dl=20.0, wavelength_nm=1550.0,
                                      pml_layers=0, n_steps=config.n_time_steps,
                                      device=device, courant=0.35, **kw)
                eps = eps_template.clone().detach().requires_grad_(True)

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
                if device == "cuda":
                    peak_mb = (torch.cuda.max_memory_allocated() - mem_before) / 1e6
                else:
                    peak_mb = _estimate_cpu_peak_mb(eps, config.n_time_steps, grid_size)
                cos = None
                if mode == "autograd":
                    reference_grad = grad
                elif reference_grad is not None:
                    cos = _cosine_sim(reference_grad, grad)
                self._results.append(BenchmarkResult(
                    grid_size=grid_size, backward_mode=mode,
                    forward_time_ms=round(fwd_ms, 3), backward_time_ms=round(bwd_ms, 3),
                    peak_memory_mb=round(peak_mb, 3),
                    gradient_cosine_vs_autograd=None if cos is None else round(cos, 6),
                    device=device))
        return self._results

    def summary_table(self) -> str:
        if not self._results:
            return "No benchmark results."
        header = (f"{'grid':>12s} | {'mode':>14s} | {'fwd_ms':>8s} | {'bwd_ms':>8s} | "
                  f"{'mem_MB':>8s} | {'cos_vs_AD':>10s} | {'device':>6s}")
        lines = [header, "-" * len(header)]
        for r in self._results:
            g = "x".join(str(v) for v in r.grid_size)
            cos = ("ref" if r.gradient_cosine_vs_autograd is None
                   else f"{r.gradient_cosine_vs_autograd:.4f}")
            lines.append(f"{g:>12s} | {r.backward_mode:>14s} | {r.forward_time_ms:8.1f} | "
                         f"{r.backward_time_ms:8.1f} | {r.peak_memory_mb:8.2f} | "
                         f"{cos:>10s} | {r.device:>6s}")
        return "\n".join(lines)


class ExternalCrossValidator:
    """Compare this solver's outputs with results computed elsewhere."""

    def validate_forward(self, our_result: torch.Tensor, external_result: torch.Tensor,
                         rtol: float = 1e-3) -> dict[str, Any]:
        ours = our_result.detach().to(torch.float64)
        ext = external_result.detach().to(torch.float64).to(ours.device)
        diff = ours - ext
        rel = diff.norm().item() / max(ext.norm().item(), 1e-30)
        return {"relative_error": rel, "max_absolute_error": diff.abs().max().item(),
                "passed": rel < rtol}

    def validate_gradient(self, our_grad: torch.Tensor, external_grad: torch.Tensor,
                          cosine_threshold: float = 0.99) -> dict[str, Any]:
        ours = our_grad.detach().to(torch.float64)
        ext = external_grad.detach().to(torch.float64).to(ours.device)
        cos = _cosine_sim(ours, ext)
        rel = (ours - ext).norm().item() / max(ext.norm().item(), 1e-30)
        return {"cosine_similarity": cos, "relative_error": rel,
                "passed": cos >= cosine_threshold}

    def generate_test_case(self, grid_size: tuple[int, int, int] = (16, 16, 16),
                           source_pos: tuple[int, int, int] | None = None,
                           freq: float = 1.0) -> dict[str, Any]:
        """The case description an external solver is run on (seed 42)."""
        D, H, W = grid_size
        if source_pos is None:
            source_pos = (D // 2, H // 2, W // 2)
        seed = 42
        torch.manual_seed(seed)
        eps = 1.5 + torch.rand(D, H, W, dtype=torch.float64)
        return {
            "grid_size": grid_size, "source_pos": source_pos, "freq": freq, "eps": eps,
            "source_config": {"type": "gaussian_pulse", "pos": list(source_pos),
                              "amplitude": 1.0},
            "n_steps": 50, "dl": 20.0, "wavelength_nm": 1550.0, "courant": 0.35,
            "seed": seed,
        }

    @staticmethod
    def load_external_results(path: str) -> dict[str, Any]:
        """JSON with ``field`` / ``gradient`` as ``{"shape": [...], "data": [...]}``
        (loaded as float64 tensors) and any other keys passed through."""
        with open(Path(path)) as f:
            raw = json.load(f)
        out: dict[str, Any] = {}
        for key in ("field", "gradient"):
            if key in raw:
                out[key] = torch.tensor(raw[key]["data"], dtype=torch.float64).reshape(
                    tuple(raw[key]["shape"]))
        out.update({k: v for k, v in raw.items() if k not in ("field", "gradient")})
        return out

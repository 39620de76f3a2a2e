"""FDTD benchmark suite and external cross-validation for the MI355X solver.

Drop-in for the interface of DiffNano's ``diffnano/solvers/fdtd_benchmark.py``
(the dataclasses ``BenchmarkConfig`` / ``BenchmarkResult`` at :36-56,
``FDTDBenchmarkSuite.run`` / ``summary_table`` at :86-204 and
``ExternalCrossValidator`` at :219-310): same class and method names,
arguments, result fields and keys, the seed-42 test permittivity and the
``{"shape": [...], "data": [...]}`` JSON exchange format.  The measurement
itself is ``_harness.run_trial``: HIP-event phase times on the solver's
stream and the allocator's memory peak (the kernels run asynchronously on
the GPU, so host clocks without a synchronisation would time launches).

Semantics kept from the reference: every backward mode of a grid runs on
the same seeded permittivity with ``field.sum()`` as the loss; mode
"time_reversal" sets ``backward="time_reversal"``, "checkpoint" sets
``use_checkpoint=True, checkpoint_segments=2``, any other name runs the
solver's default; a result's gradient cosine is taken against the most
recent "autograd" result of its grid, and is None for "autograd" itself and
for modes listed before the first "autograd".  Here all modes return the
same gradient (they only choose the adjoint's checkpoint interval), so the
cosines are 1 up to rounding.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field as dc_field
from pathlib import Path
from typing import Any

import torch

from ._harness import agreement, resolve_device, run_trial, seeded_eps, storage_bytes

__all__ = ["BenchmarkConfig", "BenchmarkResult", "FDTDBenchmarkSuite", "ExternalCrossValidator"]

# backward-mode name -> solver keywords (other names: the solver's default)
MODE_KWARGS: dict[str, dict[str, Any]] = {
    "time_reversal": {"backward": "time_reversal"},
    "checkpoint": {"use_checkpoint": True, "checkpoint_segments": 2},
}
REFERENCE_MODE = "autograd"


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


def _make_eps_grid(grid_size: tuple[int, int, int], device: str) -> torch.Tensor:
    return seeded_eps(grid_size, device)


def model_peak_mb(grid_size: tuple[int, int, int], n_steps: int, extra_kw: dict) -> float:
    """Device memory of one solve from the storage model (used where no
    allocator peak exists), at the checkpoint interval the keywords select."""
    if extra_kw.get("use_checkpoint") or extra_kw.get("backward") == "checkpoint":
        K = -(-n_steps // max(1, extra_kw.get("checkpoint_segments", 4)))
    else:
        K = extra_kw.get("checkpoint_every", 64)
    eb = extra_kw.get("e_store_gb")
    return sum(storage_bytes(grid_size, n_steps, K, e_store_bytes=None if eb is None
                             else int(eb * 1e9)).values()) / 1e6


def _cosines(modes: list[str], grads: list[torch.Tensor]) -> list[float | None]:
    """Cosine of every mode's gradient against the latest reference-mode
    gradient listed before it (None for the reference mode itself and for
    modes without an earlier reference)."""
    out: list[float | None] = []
    ref = None
    for mode, g in zip(modes, grads):
        if mode == REFERENCE_MODE:
            ref = g
            out.append(None)
        else:
            out.append(None if ref is None else round(agreement(g, ref).cosine, 6))
    return out


def _render_table(columns: list[str], rows: list[list[str]]) -> str:
    """Right-aligned columns sized to their widest cell, ' | ' separated,
    with a dashed rule under the header."""
    widths = [max(len(c), *(len(r[i]) for r in rows)) for i, c in enumerate(columns)]
    fmt = lambda cells: " | ".join(c.rjust(w) for c, w in zip(cells, widths))  # noqa: E731
    head = fmt(columns)
    return "\n".join([head, "-" * len(head)] + [fmt(r) for r in rows])


class FDTDBenchmarkSuite:
    """Forward / backward time, device-memory peak and gradient agreement of
    a 3D FDTD class over grid sizes and backward modes."""

    COLUMNS = ["grid", "mode", "fwd_ms", "bwd_ms", "mem_MB", "cos_vs_AD", "device"]

    def __init__(self) -> None:
        self._results: list[BenchmarkResult] = []

    def run(self, fdtd3d_class: type, config: BenchmarkConfig) -> list[BenchmarkResult]:
        device = resolve_device(config.device)
        modes = list(config.backward_modes)
        self._results = []
        for grid in config.grid_sizes:
            eps = seeded_eps(grid, device)
            trials = [run_trial(fdtd3d_class, grid, config.n_time_steps, device,
                                MODE_KWARGS.get(m, {}), eps) for m in modes]
            for mode, t, cos in zip(modes, trials, _cosines(modes, [t.gradient for t in trials])):
                peak = (t.peak_bytes / 1e6 if t.peak_bytes is not None else
                        model_peak_mb(grid, config.n_time_steps, MODE_KWARGS.get(mode, {})))
                self._results.append(BenchmarkResult(
                    grid_size=grid, backward_mode=mode,
                    forward_time_ms=round(t.forward_ms, 3),
                    backward_time_ms=round(t.backward_ms, 3),
                    peak_memory_mb=round(peak, 3), gradient_cosine_vs_autograd=cos,
                    device=device))
        return self._results

    def summary_table(self) -> str:
        if not self._results:
            return "No benchmark results."
        rows = [["x".join(map(str, r.grid_size)), r.backward_mode, f"{r.forward_time_ms:.1f}",
                 f"{r.backward_time_ms:.1f}", f"{r.peak_memory_mb:.2f}",
                 "ref" if r.gradient_cosine_vs_autograd is None
                 else f"{r.gradient_cosine_vs_autograd:.4f}", r.device]
                for r in self._results]
        return _render_table(self.COLUMNS, rows)


def _decode_array(key: str, entry: Any) -> torch.Tensor:
    """One ``{"shape": [...], "data": [...]}`` entry as a float64 tensor
    (``data`` flat in row-major order, or already nested)."""
    if not isinstance(entry, dict) or "shape" not in entry or "data" not in entry:
        raise ValueError(f"external results: {key!r} must be an object with 'shape' and 'data'")
    shape = tuple(int(v) for v in entry["shape"])
    t = torch.tensor(entry["data"], dtype=torch.float64).reshape(-1)
    if t.numel() != (1 if not shape else int(torch.tensor(shape).prod())):
        raise ValueError(f"external results: {key!r} holds {t.numel()} values for shape {shape}")
    return t.reshape(shape)


class ExternalCrossValidator:
    """Compare this solver's field and gradient with results computed by an
    external solver and exchanged as JSON."""

    ARRAY_KEYS = ("field", "gradient")

    def validate_forward(self, our_result: torch.Tensor, external_result: torch.Tensor,
                         rtol: float = 1e-3) -> dict[str, Any]:
        a = agreement(our_result, external_result)
        return {"relative_error": a.relative_error, "max_absolute_error": a.max_absolute_error,
                "passed": a.relative_error < rtol}

    def validate_gradient(self, our_grad: torch.Tensor, external_grad: torch.Tensor,
                          cosine_threshold: float = 0.99) -> dict[str, Any]:
        a = agreement(our_grad, external_grad)
        return {"cosine_similarity": a.cosine, "relative_error": a.relative_error,
                "passed": a.cosine >= cosine_threshold}

    def generate_test_case(self, grid_size: tuple[int, int, int] = (16, 16, 16),
                           source_pos: tuple[int, int, int] | None = None,
                           freq: float = 1.0) -> dict[str, Any]:
        """The case an external solver is run on: torch seed 42,
        eps = 1.5 + U[0, 1) (no block), a unit Gaussian-pulse point source
        (default: the centre cell), 50 steps, dl 20 nm, Courant 0.35."""
        seed = 42
        torch.manual_seed(seed)
        eps = 1.5 + torch.rand(*grid_size, dtype=torch.float64)
        pos = tuple(source_pos) if source_pos is not None else tuple(n // 2 for n in grid_size)
        return {"grid_size": grid_size, "source_pos": pos, "freq": freq, "eps": eps,
                "source_config": {"type": "gaussian_pulse", "pos": list(pos), "amplitude": 1.0},
                "n_steps": 50, "dl": 20.0, "wavelength_nm": 1550.0, "courant": 0.35,
                "seed": seed}

    @staticmethod
    def load_external_results(path: str) -> dict[str, Any]:
        """JSON whose ``field`` / ``gradient`` are ``{"shape", "data"}``
        objects (returned as float64 tensors); other keys pass through."""
        raw = json.loads(Path(path).read_text())
        return {k: (_decode_array(k, v) if k in ExternalCrossValidator.ARRAY_KEYS else v)
                for k, v in raw.items()}

    @staticmethod
    def save_external_results(path: str, **entries: Any) -> None:
        """Write results in the format ``load_external_results`` reads:
        tensors and arrays become ``{"shape", "data"}`` objects."""
        def enc(v):
            if isinstance(v, torch.Tensor) or hasattr(v, "shape"):
                t = torch.as_tensor(v, dtype=torch.float64).detach().cpu()
                return {"shape": list(t.shape), "data": t.reshape(-1).tolist()}
            return v
        Path(path).write_text(json.dumps({k: enc(v) for k, v in entries.items()}))

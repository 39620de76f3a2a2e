"""GPU benchmark matrix, coarse-to-fine tiling and external cross-validation
for the MI355X FDTD solver.

Drop-in for the interface of DiffNano's ``diffnano/solvers/fdtd_benchmark_gpu.py``:
``GPUMemoryStrategy`` (:59-78), ``FDTDBenchmarkConfig`` (:86-113),
``StabilityReport`` (:121-144), ``FDTDGPUBenchmark`` (:192-392),
``MultiScaleMetalens`` (:400-636) and ``FDTDXCrossValidator`` (:644-870):
the same names, arguments, record keys and validity rules.  Measurements go
through ``_harness.run_trial`` (HIP-event phase times, allocator peak).

What differs from the reference, and why:
* The strategies only select the adjoint's checkpoint interval (every one
  runs the same hand-written adjoint), so gradient cosines are 1.  Device
  "cpu" returns results on the host; the kernels still run on the GPU (the
  product has no CPU solver to fall back to).
* ``peak_memory_mb`` is the allocator's peak when a GPU is present;
  ``storage_mb`` is the analytic model of this solver's buffers (the
  reference's model describes its own snapshot lists and autograd graph).
* ``MultiScaleMetalens`` advances all fine tiles of a forward in one batched
  launch sequence (``BatchedFDTD3D``: equal tile shapes, one persistent
  kernel per step) instead of one solve per tile, and ``benchmark_scaling``
  times that batched stage instead of extrapolating from one tile.
"""

from __future__ import annotations

import enum
import math
import time
from dataclasses import dataclass, field as dc_field
from typing import Any, Callable

import torch

from ._harness import agreement, resolve_device, run_trial, seeded_eps, storage_bytes

__all__ = ["GPUMemoryStrategy", "FDTDBenchmarkConfig", "StabilityReport", "FDTDGPUBenchmark",
           "MultiScaleMetalens", "FDTDXCrossValidator", "storage_mb"]


class GPUMemoryStrategy(enum.Enum):
    TIME_REVERSAL = "time_reversal"
    CHECKPOINT = "checkpoint"
    FULL_AUTODIFF = "full_autodiff"


# strategy -> solver keywords, and the checkpoint interval each one gives
_STRATEGY_KW = {
    GPUMemoryStrategy.TIME_REVERSAL: {"backward": "time_reversal"},
    GPUMemoryStrategy.CHECKPOINT: {"use_checkpoint": True, "checkpoint_segments": 4},
    GPUMemoryStrategy.FULL_AUTODIFF: {},
}


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


def _interval(strategy: GPUMemoryStrategy, n_steps: int) -> int:
    kw = _STRATEGY_KW[strategy]
    if kw.get("use_checkpoint"):
        return -(-n_steps // kw["checkpoint_segments"])
    return 64  # the solvers' default checkpoint_every


def storage_mb(grid_size: tuple[int, int, int], n_steps: int, strategy: GPUMemoryStrategy,
               elem_bytes: int = 4) -> float:
    """Analytic device memory of one forward + adjoint (``_harness.storage_bytes``)."""
    return sum(storage_bytes(grid_size, n_steps, _interval(strategy, n_steps),
                             elem_bytes).values()) / 1e6


def _reference_first(strategies):
    """Full autodiff (the cosine's reference) first, the rest in order."""
    ref = [s for s in strategies if s == GPUMemoryStrategy.FULL_AUTODIFF]
    return ref + [s for s in strategies if s != GPUMemoryStrategy.FULL_AUTODIFF]


class FDTDGPUBenchmark:
    """FDTDSolver3D over grid sizes and gradient memory strategies."""

    def __init__(self, config: FDTDBenchmarkConfig) -> None:
        self.config = config
        self._results: list[dict[str, Any]] = []

    @staticmethod
    def detect_gpu() -> bool:
        return torch.cuda.is_available()

    def run_single(self, grid_size: tuple[int, int, int], strategy: GPUMemoryStrategy,
                   device: str) -> dict[str, Any]:
        """One trial; the record keeps the gradient for ``run_all``."""
        from .fdtd3d import FDTDSolver3D
        dev = resolve_device(device)
        t = run_trial(FDTDSolver3D, grid_size, self.config.n_time_steps, dev,
                      _STRATEGY_KW[strategy])
        model = storage_mb(grid_size, self.config.n_time_steps, strategy)
        rec = {"grid_size": grid_size, "strategy": strategy.value, "device": dev,
               "forward_time_ms": round(t.forward_ms, 3),
               "backward_time_ms": round(t.backward_ms, 3),
               "peak_memory_mb": round(t.peak_bytes / 1e6 if t.peak_bytes is not None
                                       else model, 3),
               "storage_mb": round(model, 3), "gradient": t.gradient}
        self._results.append(rec)
        return rec

    def run_all(self) -> list[dict[str, Any]]:
        """Every grid x strategy with full autodiff first per grid; each
        record's gradient is replaced by its cosine against that run (None
        for the run itself, and for all when it is not configured)."""
        self._results = []
        dev = resolve_device(self.config.device)
        for grid in self.config.grid_sizes:
            recs = [self.run_single(grid, s, dev)
                    for s in _reference_first(self.config.memory_strategies)]
            ref = next((r["gradient"] for r in recs
                        if r["strategy"] == GPUMemoryStrategy.FULL_AUTODIFF.value), None)
            for r in recs:
                g = r.pop("gradient")
                r["gradient_cosine"] = (None if ref is None or g is ref else
                                        round(agreement(g, ref).cosine, 6))
        return self._results

    def benchmark_memory_scaling(self) -> list[dict[str, Any]]:
        """Peak memory of every grid x strategy (grid-major order)."""
        return [{"grid_size": g, "strategy": s.value,
                 "peak_memory_mb": self.run_single(g, s, self.config.device)["peak_memory_mb"]}
                for g in self.config.grid_sizes for s in self.config.memory_strategies]

    def cpu_fallback_benchmark(self) -> StabilityReport:
        """Every strategy on the first grid with at most 20 steps and results
        on the host; valid when each cosine against full autodiff >= 0.99."""
        S = GPUMemoryStrategy
        sub = FDTDGPUBenchmark(FDTDBenchmarkConfig(
            grid_sizes=self.config.grid_sizes[:1],
            memory_strategies=[S.FULL_AUTODIFF, S.TIME_REVERSAL, S.CHECKPOINT],
            n_time_steps=min(self.config.n_time_steps, 20), device="cpu"))
        recs = sub.run_all()
        cos = [r["gradient_cosine"] for r in recs if r["gradient_cosine"] is not None]
        return StabilityReport(forward_errors=[0.0], gradient_cosines=cos,
                               memory_used_mb={r["strategy"]: r["peak_memory_mb"] for r in recs},
                               is_valid=bool(cos) and min(cos) >= 0.99)


# ---------------------------------------------------------------------------
# Coarse-to-fine tiling
# ---------------------------------------------------------------------------


def _resample(field: torch.Tensor, size) -> torch.Tensor:
    """Trilinear resampling of a (C, D, H, W) or (D, H, W) array to ``size``
    (half-pixel centres, as torch's ``align_corners=False``), in float32 as
    the reference's coupling does (fdtd_benchmark_gpu.py:506-512, 549-554)."""
    x = field[None] if field.dim() == 4 else field[None, None]
    y = torch.nn.functional.interpolate(x.float(), size=tuple(size), mode="trilinear",
                                        align_corners=False)
    return (y[0] if field.dim() == 4 else y[0, 0]).to(torch.float64)


class MultiScaleMetalens:
    """Coarse-to-fine tiling of a large aperture (fdtd_benchmark_gpu.py:400-636).

    A coarse solve (``coarse_grid`` cells at ``coarse_dl``) covers the whole
    domain; the design is cut into ``n_tiles`` tiles of ``fine_grid`` cells
    (the last ones padded with eps = 1), each solved at ``fine_dl``; the
    boundary bands of each fine field (``fine_pml`` cells at both ends of
    axes 0 and 1) are then taken from the coarse field resampled onto the
    tile.  ``coupling="domain"`` (the reference's) resamples the whole coarse
    field onto every tile; ``"tile"`` resamples only the part of the coarse
    domain the tile covers.  Solver settings default to the reference's
    (coarse: dl 40, 3 CPML cells, 50 steps; fine: dl 20, 5 cells, 100 steps;
    Courant 0.35; the default Ez line source).
    """

    def __init__(self, coarse_grid: tuple[int, int, int], fine_grid: tuple[int, int, int],
                 n_tiles: tuple[int, int, int], *, coarse_dl: float = 40.0,
                 fine_dl: float = 20.0, coarse_pml: int = 3, fine_pml: int = 5,
                 coarse_steps: int = 50, fine_steps: int = 100, courant: float = 0.35,
                 wavelength_nm: float = 1550.0, coupling: str = "domain",
                 device: Any = "cpu", dtype: torch.dtype = torch.float32) -> None:
        if coupling not in ("domain", "tile"):
            raise ValueError(f"coupling must be 'domain' or 'tile', got {coupling!r}")
        self.coarse_grid = tuple(coarse_grid)
        self.fine_grid = tuple(fine_grid)
        self.n_tiles = tuple(n_tiles)
        self.coarse_kw = dict(grid_shape=self.coarse_grid, dl=coarse_dl, wavelength_nm=wavelength_nm,
                              pml_layers=coarse_pml, n_steps=coarse_steps, courant=courant,
                              dtype=dtype)
        self.fine_kw = dict(grid_shape=self.fine_grid, dl=fine_dl, wavelength_nm=wavelength_nm,
                            pml_layers=fine_pml, n_steps=fine_steps, courant=courant, dtype=dtype)
        self.coupling = coupling
        self.device = torch.device(device)
        self._coarse_result: torch.Tensor | None = None

    # -- stages ---------------------------------------------------------------
    def coarse_solve(self, region: torch.Tensor) -> torch.Tensor:
        """Field (3, D, H, W) = [Ez, Ex, Ey] of the coarse solve of ``region``."""
        from .fdtd3d import FDTDSolver3D
        s = FDTDSolver3D(**self.coarse_kw, device=self.device)
        with torch.no_grad():
            self._coarse_result = s.forward(region.to(self.device, torch.float64)).field
        return self._coarse_result

    def _couple(self, fine: torch.Tensor, boundary: torch.Tensor) -> torch.Tensor:
        """Overwrite the axis-0 and axis-1 boundary bands of ``fine``
        (3, fz, fy, fx) with ``boundary`` resampled onto the tile."""
        p = self.fine_kw["pml_layers"]
        if p <= 0 or tuple(boundary.shape[1:]) == tuple(fine.shape[1:]):
            return fine
        up = _resample(boundary.to(fine.device), fine.shape[1:])
        out = fine.clone()
        for ax in (1, 2):                 # the field's axes 1, 2 = grid axes 0, 1
            lo = [slice(None)] * 4
            hi = [slice(None)] * 4
            lo[ax], hi[ax] = slice(0, p), slice(-p, None)
            out[tuple(lo)] = up[tuple(lo)]
            out[tuple(hi)] = up[tuple(hi)]
        return out

    def fine_solve(self, region: torch.Tensor, coarse_boundary: torch.Tensor) -> torch.Tensor:
        """One tile: a direct solve of ``region`` (``fine_grid``), coupled to
        ``coarse_boundary`` (a (3, ...) field resampled onto the tile)."""
        from .fdtd3d import FDTDSolver3D
        s = FDTDSolver3D(**self.fine_kw, device=self.device)
        with torch.no_grad():
            fine = s.forward(region.to(self.device, torch.float64)).field
        return self._couple(fine, coarse_boundary)

    def tiles(self, design: torch.Tensor) -> list[tuple[tuple[int, int, int], torch.Tensor]]:
        """(origin, eps) of every tile in z-major order; tiles that run past
        the design are padded with eps = 1 at their far side."""
        fz, fy, fx = self.fine_grid
        out = []
        for iz in range(self.n_tiles[0]):
            for iy in range(self.n_tiles[1]):
                for ix in range(self.n_tiles[2]):
                    o = (iz * fz, iy * fy, ix * fx)
                    part = design[o[0]:o[0] + fz, o[1]:o[1] + fy, o[2]:o[2] + fx]
                    if tuple(part.shape) != self.fine_grid:
                        pad = torch.ones(self.fine_grid, dtype=torch.float64,
                                         device=design.device)
                        pad[:part.shape[0], :part.shape[1], :part.shape[2]] = part
                        part = pad
                    out.append((o, part))
        return out

    def _tile_boundary(self, coarse: torch.Tensor, origin, design_shape) -> torch.Tensor:
        """The coarse field over the tile's extent (``coupling="tile"``)."""
        sl = [slice(None)]
        for a in range(3):
            scale = self.coarse_grid[a] / design_shape[a]
            lo = int(origin[a] * scale)
            hi = max(lo + 1, int(round((origin[a] + self.fine_grid[a]) * scale)))
            sl.append(slice(lo, min(hi, self.coarse_grid[a])))
        return coarse[tuple(sl)]

    def _solve_tiles(self, tiles) -> list[torch.Tensor]:
        """All tiles in one batched launch sequence (no gradient)."""
        from .batched import BatchedFDTD3D
        kw = self.fine_kw
        n = len(tiles)
        if n == 0:
            return []
        # (the plan's checkpoint and replay buffers are unused by a forward
        # without gradient: an interval near sqrt(T) keeps them small)
        K = 2 * max(1, round(math.sqrt(max(kw["n_steps"], 1)) / 2))
        b = BatchedFDTD3D([self.fine_grid] * n, [{"type": "gaussian_pulse"}] * n,
                          [{"box": (0, 1, 0, 1, 0, 1)}] * n, dl=kw["dl"],
                          wavelength_nm=kw["wavelength_nm"], pml_layers=kw["pml_layers"],
                          n_steps=kw["n_steps"], courant=kw["courant"], device=self.device,
                          dtype=kw["dtype"], checkpoint_every=K)
        with torch.no_grad():
            b.forward([t for _, t in tiles])
        return [f.to(self.device) for f in b.final_fields()]

    def forward(self, design: torch.Tensor) -> dict[str, Any]:
        """Coarse solve of the design resampled onto ``coarse_grid``, then
        every tile's fine solve (batched) and coupling."""
        design = design.to(self.device, torch.float64)
        coarse = self.coarse_solve(_resample(design, self.coarse_grid))
        tiles = self.tiles(design)
        fields = []
        for (o, _), fine in zip(tiles, self._solve_tiles(tiles)):
            bnd = coarse if self.coupling == "domain" else \
                self._tile_boundary(coarse, o, tuple(design.shape))
            fields.append(self._couple(fine, bnd))
        return {"coarse_field": coarse, "tile_fields": fields, "n_tiles": self.n_tiles}

    def benchmark_scaling(self, base_size: int = 16, max_tiles: int = 4) -> list[dict[str, Any]]:
        """For n = 1..max_tiles tiles per axis (tiles of ``base_size``^3, the
        coarse grid as large as the tiled domain): measured coarse-solve
        time, the batched fine stage's time per tile, and their total."""
        out = []
        for n in range(1, max_tiles + 1):
            dim = base_size * n
            msl = MultiScaleMetalens((dim,) * 3, (base_size,) * 3, (n,) * 3,
                                     coarse_dl=self.coarse_kw["dl"], fine_dl=self.fine_kw["dl"],
                                     coarse_pml=self.coarse_kw["pml_layers"],
                                     fine_pml=self.fine_kw["pml_layers"],
                                     coarse_steps=self.coarse_kw["n_steps"],
                                     fine_steps=self.fine_kw["n_steps"],
                                     courant=self.coarse_kw["courant"], coupling=self.coupling,
                                     device=self.device, dtype=self.coarse_kw["dtype"])
            torch.manual_seed(42)
            design = 1.5 + torch.rand(dim, dim, dim, dtype=torch.float64)
            t_coarse = _timed(lambda: msl.coarse_solve(design))
            tiles = msl.tiles(design.to(msl.device))
            t_fine = _timed(lambda: msl._solve_tiles(tiles))
            out.append({"n_tiles_per_axis": n, "total_tiles": n ** 3,
                        "coarse_time_ms": round(t_coarse, 3),
                        "fine_time_ms": round(t_fine / n ** 3, 3),
                        "total_time_ms": round(t_coarse + t_fine, 3)})
        return out


def _timed(fn: Callable[[], Any]) -> float:
    """Wall time of ``fn`` in ms, synchronised with the GPU on both sides."""
    if torch.cuda.is_available():
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    if torch.cuda.is_available():
        torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3


# ---------------------------------------------------------------------------
# External cross-validation
# ---------------------------------------------------------------------------


class FDTDXCrossValidator:
    """Cross-validation against an external solver (fdtd_benchmark_gpu.py:644-870).

    ``reference_fn(grid_size, n_steps)`` returns ``{"field": (3, D, H, W),
    "gradient": (D, H, W)}`` computed elsewhere on the harness case (the
    seed-42 permittivity of ``_harness.seeded_eps``, no CPML, default source,
    loss ``field.sum()``); without one the reference is this solver itself,
    which checks the pipeline and its run-to-run repeatability.  As in the
    reference, every "seed" of ``run_cross_validation`` runs that same case.
    """

    def __init__(self, reference_fn: Callable | None = None) -> None:
        self.reference_fn = reference_fn

    @staticmethod
    def _ours(grid_size, n_steps):
        from .fdtd3d import FDTDSolver3D
        return run_trial(FDTDSolver3D, grid_size, n_steps, "cpu")

    def generate_synthetic_reference(self, grid_size: tuple[int, int, int],
                                     n_steps: int = 50) -> dict[str, torch.Tensor]:
        t = self._ours(grid_size, n_steps)
        return {"field": t.field, "gradient": t.gradient}

    def validate_forward(self, result: torch.Tensor, reference: torch.Tensor,
                         rtol: float = 1e-3) -> dict[str, Any]:
        a = agreement(result, reference)
        return {"relative_error": a.relative_error, "max_absolute_error": a.max_absolute_error,
                "passed": a.relative_error < rtol}

    def validate_gradient(self, grad: torch.Tensor, reference_grad: torch.Tensor,
                          cosine_threshold: float = 0.99) -> dict[str, Any]:
        a = agreement(grad, reference_grad)
        return {"cosine_similarity": a.cosine, "relative_error": a.relative_error,
                "passed": a.cosine >= cosine_threshold}

    def run_cross_validation(self, n_seeds: int = 3, grid_size: tuple[int, int, int] = (16, 16, 16),
                             n_steps: int = 30) -> StabilityReport:
        """Per seed: the reference and this solver's field and gradient,
        compared; valid when every forward error is below 1e-3 and every
        cosine at least 0.99.  memory_used_mb: this solver's peak."""
        fwd, cos, peak = [], [], 0.0
        for seed in range(n_seeds):
            torch.manual_seed(seed * 100 + 42)
            ref = (self.reference_fn(grid_size, n_steps) if self.reference_fn is not None
                   else self.generate_synthetic_reference(grid_size, n_steps))
            t = self._ours(grid_size, n_steps)
            fwd.append(self.validate_forward(t.field, ref["field"])["relative_error"])
            cos.append(self.validate_gradient(t.gradient, ref["gradient"])["cosine_similarity"])
            peak = max(peak, (t.peak_bytes or 0) / 1e6)
        if not torch.cuda.is_available():
            peak = storage_mb(grid_size, n_steps, GPUMemoryStrategy.FULL_AUTODIFF)
        return StabilityReport(forward_errors=fwd, gradient_cosines=cos,
                               memory_used_mb={"full_autodiff": round(peak, 3)},
                               is_valid=bool(fwd) and max(fwd) < 1e-3 and min(cos) >= 0.99)


def _make_eps_grid(grid_size: tuple[int, int, int], device: str) -> torch.Tensor:
    return seeded_eps(grid_size, device)

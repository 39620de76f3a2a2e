"""Measurement core shared by the benchmark and cross-validation drop-ins
(``fdtd_benchmark.py``, ``fdtd_benchmark_gpu.py``).

One *trial* = build a solver with the harness's fixed physics (no CPML,
dl 20 nm, 1550 nm, Courant 0.35, the reference harnesses' choice in
fdtd_benchmark.py:124-133 and fdtd_benchmark_gpu.py:249-257), run the
forward on a seeded permittivity, back-propagate ``field.sum()`` and record
the two phase times, the device-memory peak and the gradient.

* Times come from HIP events recorded on the current stream when a GPU is
  present (the kernels run asynchronously, so a host clock would time the
  launches); without one, from the host clock.
* The memory peak is torch's allocator peak over the trial: every buffer of
  a solve (state, checkpoints, replay store, adjoint, workspace) is a torch
  allocation, the C ABI allocates nothing.  Without a GPU the analytic
  ``storage_bytes`` model of those buffers stands in.
"""

from __future__ import annotations

import gc
import math
import time
from dataclasses import dataclass
from typing import Any

import torch

# The harnesses' solver settings (kept from the reference harnesses).
BASE_SOLVER_KW = {"dl": 20.0, "wavelength_nm": 1550.0, "pml_layers": 0, "courant": 0.35}


def seeded_eps(grid: tuple[int, int, int], device: Any = "cpu") -> torch.Tensor:
    """The harnesses' test permittivity (fdtd_benchmark.py:67-74): torch seed
    42, 1.5 + U[0, 1) in float64, and a centred block of 4.0 spanning half of
    every axis when each axis has at least 4 cells."""
    D, H, W = grid
    torch.manual_seed(42)
    eps = 1.5 + torch.rand(D, H, W, dtype=torch.float64, device=device)
    half = (D // 4, H // 4, W // 4)
    if min(half) > 0:
        sl = tuple(slice(n // 2 - h, n // 2 + h) for n, h in zip((D, H, W), half))
        eps[sl] = 4.0
    return eps


def resolve_device(requested: str) -> str:
    """Where results go: "cuda" only when a GPU is visible."""
    return "cuda" if requested == "cuda" and torch.cuda.is_available() else "cpu"


@dataclass
class Agreement:
    """How far ``ours`` is from ``ref`` (both flattened to float64)."""
    relative_error: float
    max_absolute_error: float
    cosine: float


def agreement(ours: torch.Tensor, ref: torch.Tensor) -> Agreement:
    a = ours.detach().reshape(-1).to(torch.float64)
    b = ref.detach().reshape(-1).to(torch.float64).to(a.device)
    d = a - b
    nb = float(b.norm())
    na = float(a.norm())
    cos = float(torch.dot(a, b)) / (na * nb) if na * nb >= 1e-30 else 0.0
    return Agreement(relative_error=float(d.norm()) / max(nb, 1e-30),
                     max_absolute_error=float(d.abs().max()) if d.numel() else 0.0,
                     cosine=cos)


class _Laps:
    """Phase times of one trial: HIP events on the current stream (GPU) or
    the host clock."""

    def __init__(self, gpu: bool):
        self.gpu = gpu
        self.marks: list[tuple[str, Any]] = []

    def mark(self, name: str) -> None:
        if self.gpu:
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            self.marks.append((name, ev))
        else:
            self.marks.append((name, time.perf_counter()))

    def durations_ms(self) -> dict[str, float]:
        if self.gpu:
            torch.cuda.synchronize()
        out = {}
        for (_, a), (name, b) in zip(self.marks, self.marks[1:]):
            out[name] = a.elapsed_time(b) if self.gpu else (b - a) * 1e3
        return out


@dataclass
class TrialOutcome:
    forward_ms: float
    backward_ms: float
    peak_bytes: int | None       # torch allocator peak (GPU), else None
    gradient: torch.Tensor       # dL/deps of L = sum(field), float64, on the host
    field: torch.Tensor          # the forward field, on the host


def run_trial(solver_cls: type, grid: tuple[int, int, int], n_steps: int, device: str,
              extra_kw: dict | None = None, eps: torch.Tensor | None = None) -> TrialOutcome:
    """Forward + ``field.sum()`` backward of one solver configuration."""
    gpu = torch.cuda.is_available()
    gc.collect()
    base = 0
    if gpu:
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
    kw = dict(BASE_SOLVER_KW)
    kw.update(extra_kw or {})
    solver = solver_cls(grid_shape=grid, n_steps=n_steps, device=device, **kw)
    e = (seeded_eps(grid, device) if eps is None else eps.to(device)).detach().clone()
    e.requires_grad_(True)
    laps = _Laps(gpu)
    laps.mark("start")
    result = solver.forward(e)
    laps.mark("forward")
    result.field.sum().backward()
    laps.mark("backward")
    ms = laps.durations_ms()
    peak = (torch.cuda.max_memory_allocated() - base) if gpu else None
    grad = e.grad.detach().to("cpu").clone() if e.grad is not None else torch.zeros(grid)
    return TrialOutcome(ms["forward"], ms["backward"], peak, grad,
                        result.field.detach().to("cpu").clone())


def storage_bytes(grid: tuple[int, int, int], n_steps: int, checkpoint_every: int,
                  elem_bytes: int = 4, n_components: int = 6,
                  e_store_bytes: int | None = None) -> dict[str, int]:
    """Device buffers of one forward + adjoint of ``FDTDSolver3D`` (or 2D with
    ``n_components=3``), by name: the permittivity and coefficient, the
    forward state and its ping-pong twin, ceil(T/K) checkpoints, the replay
    store of K-1 E fields, E of every step of the final segments that the
    time-reversal forward keeps (``final_segments``), four adjoint /
    replay state buffers, the gradient accumulator, the float64 gradient
    and the float64 output field.

    ``e_store_bytes`` is the solver's E budget (its ``e_store_gb`` keyword);
    None models the default on a device with room to spare (free memory of
    at least 16 GiB plus twice the store), where every segment but the
    first keeps its E."""
    N = math.prod(grid)
    K = max(1, int(checkpoint_every))
    ne = 3 if n_components == 6 else 1
    s = elem_bytes
    nseg = math.ceil(n_steps / K) if n_steps else 0
    tail = 0
    if nseg >= 2:
        seg = K * ne * N * s
        m = nseg - 1 if e_store_bytes is None else min(nseg - 1, int(e_store_bytes) // seg)
        tail = (n_steps - (nseg - m) * K) * ne * N * s if m > 0 else 0
    return {
        "eps_and_coefficient": N * (8 + s),
        "state": 2 * n_components * N * s,
        "checkpoints": nseg * n_components * N * s,
        "replay_store": max(K - 1, 1) * ne * N * s,
        "final_segments": tail,
        "adjoint": 4 * n_components * N * s + N * s,
        "gradient": N * 8,
        "field": ne * N * 8,
    }

"""Result container and solver protocol (DiffNano ``solvers/_result.py:8-35``
and ``solvers/__init__.py:81-116``), kept verbatim in behaviour so code that
type-checks or unpacks the reference's results works unchanged."""

from __future__ import annotations

from collections.abc import Sequence
from typing import Protocol, runtime_checkable

import torch


class SimResult:
    """Container for simulation outputs: ``field``, ``wavelengths``, ``metadata``."""

    __slots__ = ("field", "wavelengths", "metadata")

    def __init__(self, field: torch.Tensor, wavelengths: torch.Tensor, metadata: dict | None = None):
        self.field = field
        self.wavelengths = wavelengths
        self.metadata = metadata or {}

    @property
    def device(self) -> torch.device:
        return self.field.device


@runtime_checkable
class Solver(Protocol):
    """Backend-agnostic forward-solver interface (``solvers/__init__.py:81-116``)."""

    def forward(
        self,
        geometry: torch.Tensor,
        wavelengths: Sequence[float] | torch.Tensor | None,
        *,
        source: dict | None = None,
    ) -> SimResult: ...

    @property
    def device(self) -> torch.device: ...

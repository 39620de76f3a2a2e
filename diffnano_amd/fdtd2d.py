"""Differentiable 2D Yee FDTD (TM / TE), drop-in for DiffNano's
``FDTDSolver2D`` (``diffnano/solvers/fdtd2d.py``), on MI355X HIP kernels.

Same constructor, ``forward``, ``time_series``, ``device``,
``_time_step_tm``/``_te``, ``_source_waveform`` and ``_build_source_mask``
as the reference (fdtd2d.py:117-524); field ``(1, H, W)`` float64 (Ez for
TM, Hz for TE) and the reference's metadata.  The gradient (plain autograd
in the reference) comes from the hand-written adjoint kernels.

Extensions (keyword-only): ``backward`` (accepted for symmetry with 3D; all
modes give the same gradient), ``dtype``, ``checkpoint_every``, ``e_store_gb``
(device memory for E of the final checkpoint segments, which the backward
then does not replay; default: half the spare device memory above 16 GiB),
``forward(..., monitors=[{"box": (y0, y1, x0, x1), "wavelengths": [...]}])``
and source key ``profile``.
"""

from __future__ import annotations

import math
from collections.abc import Sequence

import torch

from . import _lib as L
from ._common import (EPS_CHECKS, check_eps_finite, cpml_profile, dense_mask, monitor_box,
                      source_box_2d, source_profile, waveform_table)
from ._engine import Engine, MonitorSpec, Plan, SourceSpec, _TimeReversalFDTD, forward_only
from ._result import SimResult

__all__ = ["FDTDSolver2D"]


class FDTDSolver2D:
    """Differentiable 2D FDTD solver (reference signature, fdtd2d.py:117-129)."""

    def __init__(
        self,
        grid_shape: tuple[int, int] = (100, 100),
        dl: float = 20.0,
        wavelength_nm: float = 1550.0,
        polarization: str = "TM",
        pml_layers: int = 10,
        n_steps: int = 1000,
        courant: float = 0.5,
        device: str | torch.device = "cpu",
        use_checkpoint: bool = False,
        checkpoint_segments: int = 4,
        *,
        backward: str = "auto",
        dtype: torch.dtype = torch.float32,
        checkpoint_every: int = 64,
        check_eps: str = "warn",
        e_store_gb: float | None = None,
    ):
        self.grid_shape = grid_shape
        self.dl = dl
        self.wavelength_nm = wavelength_nm
        self.polarization = polarization.upper()
        self.pml_layers = pml_layers
        self.n_steps = n_steps
        self.courant = courant
        self._device = torch.device(device)
        self.use_checkpoint = use_checkpoint
        self.checkpoint_segments = checkpoint_segments
        self.backward = backward
        self.dtype = dtype
        self.checkpoint_every = checkpoint_every
        if e_store_gb is not None and not e_store_gb >= 0:
            raise ValueError(f"e_store_gb must be None or >= 0, got {e_store_gb!r}")
        self.e_store_gb = e_store_gb
        if check_eps not in EPS_CHECKS:
            raise ValueError(f"check_eps must be one of {EPS_CHECKS}, got {check_eps!r}")
        self.check_eps = check_eps

        self.dt = courant * dl / math.sqrt(2.0)
        self.omega = 2 * math.pi / wavelength_nm
        H, W = grid_shape
        self._prof_x = cpml_profile(W, pml_layers, dl, self.dt)  # along axis 1
        self._prof_y = cpml_profile(H, pml_layers, dl, self.dt)  # along axis 0

    @property
    def device(self) -> torch.device:
        return self._device

    @property
    def _te(self) -> bool:
        # Any polarization other than "TM" runs TE, as in fdtd2d.py:326-355.
        return self.polarization != "TM"

    def _source_waveform(self, step: int, source: dict) -> torch.Tensor:
        w = waveform_table(step + 1, self.dt, self.omega, source)[step]
        return torch.tensor(w, dtype=torch.float64, device=self._device)

    def _build_source_mask(self, source: dict) -> torch.Tensor:
        return dense_mask(tuple(self.grid_shape), source_box_2d(self.grid_shape, source),
                          device=self._device)

    def _engine(self, dev, te=None):
        H, W = self.grid_shape
        by, cy = self._prof_y
        bx, cx = self._prof_x
        return Engine(2, (H, W), self.dl, self.dt, (by, cy, bx, cx), self.dtype, dev,
                      te=self._te if te is None else te)

    def _step_compat(self, a, b, c, eps_r, te):
        dev = L.compute_device(a.device)
        eng = self._engine(dev, te=te)
        eng.set_eps(eps_r.to(dev, torch.float64))
        state = torch.cat([t.reshape(-1).to(dev, self.dtype) for t in (a, b, c)])
        state = eng.run(state, 0, 1, (None, 0), (None, 0))
        N = eng.N
        return tuple(state[k * N:(k + 1) * N].view(eng.shape).to(a.device, torch.float64)
                     for k in range(3))

    def _time_step_tm(self, Ez, Hx, Hy, eps_r, mu_r):
        """One TM step (fdtd2d.py:233-267) on the GPU; mu_r must be ones."""
        return self._step_compat(Ez, Hx, Hy, eps_r, te=False)

    def _time_step_te(self, Hz, Ex, Ey, eps_r, mu_r):
        """One TE step (fdtd2d.py:269-303) on the GPU; mu_r must be ones."""
        return self._step_compat(Hz, Ex, Ey, eps_r, te=True)

    def _source_specs(self, src: dict, n_steps: int):
        box = source_box_2d(self.grid_shape, src)
        if box is None:
            return []
        wave = waveform_table(n_steps, self.dt, self.omega, src)
        return [SourceSpec(0, box, source_profile(src, box), wave)]

    def _monitor_specs(self, monitors):
        specs = []
        for m in monitors or ():
            box = monitor_box(self.grid_shape, m)
            lams = m.get("wavelengths", [self.wavelength_nm])
            specs.append(MonitorSpec(box, 1, [2 * math.pi / float(lam) for lam in lams]))
        return specs

    def _ckpt_every(self):
        if self.use_checkpoint or self.backward == "checkpoint":
            return max(1, math.ceil(self.n_steps / max(1, self.checkpoint_segments)))
        return self.checkpoint_every

    def forward(
        self,
        geometry: torch.Tensor,
        wavelengths: Sequence[float] | torch.Tensor | None = None,
        *,
        source: dict | None = None,
        monitors: list[dict] | None = None,
    ) -> SimResult:
        """Run the 2D FDTD (fdtd2d.py:412-484).  ``field`` is (1, H, W)
        float64; ``metadata["snapshots"]`` holds the probe series when
        ``source["probe"]`` is set (empty with ``use_checkpoint``)."""
        if wavelengths is None:
            wavelengths = [self.wavelength_nm]
        if not isinstance(wavelengths, torch.Tensor):
            wavelengths = torch.tensor(wavelengths, dtype=torch.float64, device=self._device)
        src = source or {"type": "gaussian_pulse"}
        H, W = self.grid_shape
        eps_r = geometry.to(self._device).to(torch.float64)
        if eps_r.dim() == 3:
            eps_r = eps_r.squeeze(0)

        finite = check_eps_finite(eps_r, self.check_eps)
        dev = L.compute_device(self._device)
        eng = self._engine(dev)
        probe = -1
        pp = src.get("probe", None)
        if pp is not None and not self.use_checkpoint:
            py, px = pp
            if not (0 <= py < H and 0 <= px < W):
                raise IndexError(f"probe {pp} outside the grid {self.grid_shape}")
            probe = py * W + px
        mon_specs = self._monitor_specs(monitors)
        plan = Plan(eng, self.n_steps, self._ckpt_every(), self._source_specs(src, self.n_steps),
                    mon_specs, out_order=[0], out_unsqueeze=True, probe=probe,
                    out_device=self._device)
        if self.e_store_gb is not None:
            plan.e_store_bytes = int(self.e_store_gb * 1e9)
        if torch.is_grad_enabled() and eps_r.requires_grad:
            outs = _TimeReversalFDTD.apply(eps_r, plan)
            field, mon_raw = outs[0], list(outs[1:])
        else:
            field, mon_raw, _ = forward_only(plan, eps_r)
        snapshots = []
        if plan.probe_values is not None:
            vals = plan.probe_values.to(torch.float64).to(self._device)
            snapshots = list(vals.unbind(0))
        meta = {
            "polarization": self.polarization,
            "n_steps": self.n_steps,
            "dt": self.dt,
            "courant": self.courant,
            "snapshots": snapshots,
        }
        if not finite:
            meta["eps_finite"] = False
        if mon_specs:
            meta["monitors"] = [torch.view_as_complex(m) for m in mon_raw]
        return SimResult(field=field, wavelengths=wavelengths, metadata=meta)

    def time_series(
        self,
        eps_r: torch.Tensor,
        probe: tuple[int, int],
        source: dict | None = None,
        n_steps: int | None = None,
    ) -> torch.Tensor:
        """Probe series (T,) float64 (fdtd2d.py:486-524).  As in the
        reference, ``source["probe"]`` is set on the caller's dict."""
        src = source or {"type": "gaussian_pulse"}
        src["probe"] = probe
        H, W = self.grid_shape
        eps_r = eps_r.to(self._device).to(torch.float64)
        if eps_r.dim() == 3:
            eps_r = eps_r.squeeze(0)
        steps = n_steps or self.n_steps
        py, px = probe
        if not (0 <= py < H and 0 <= px < W):
            raise IndexError(f"probe {probe} outside the grid {self.grid_shape}")
        dev = L.compute_device(self._device)
        eng = self._engine(dev)
        plan = Plan(eng, steps, 1, self._source_specs(src, steps), [], out_order=[0],
                    out_unsqueeze=False, probe=py * W + px)
        with torch.no_grad():
            forward_only(plan, eps_r.detach().to(dev))
        if steps == 0:
            return torch.zeros(steps, dtype=torch.float64, device=self._device)
        return plan.probe_values.to(torch.float64).to(self._device)

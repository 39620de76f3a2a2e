"""Batched 2D TM FDTD forward: many designs of one grid shape per launch.

Each design advances exactly as ``FDTDSolver2D`` in TM polarisation
(``FDTDSolver2D._time_step_tm`` / ``_run_steps``, fdtd2d.py:233-267 and
:309-337 of the reference), with one shared soft Ez source and a running
DFT of Ez over a monitor box accumulated in float32 inside the E-update
kernel (``dnx_fdtd2d_tm_batch``).  No gradient: this is the forward used to
generate surrogate training fields and to evaluate many candidate designs at
once; the differentiable single-design path stays ``FDTDSolver2D``.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from ._common import (cpml_profile, monitor_box, phase_table, source_box_2d, source_profile,
                      waveform_table)

__all__ = ["BatchedTM2D"]


class BatchedTM2D:
    """Forward-only TM FDTD over a batch of permittivity maps ``[B, H, W]``.

    ``monitor`` is ``{"box": (y0, y1, x0, x1), "wavelengths": [...]}``
    (default: the whole grid at ``wavelength_nm``).  ``forward`` returns the
    complex DFT ``[B, n_freq, y1-y0, x1-x0]`` (complex64) with the same
    convention as ``FDTDSolver2D.forward(..., monitors=...)``:
    F = sum_n Ez^{n+1} exp(-i w (n+1) dt).
    """

    def __init__(self, grid_shape=(128, 128), dl=20.0, wavelength_nm=1550.0, pml_layers=10,
                 n_steps=2000, courant=0.5, source=None, monitor=None, device=None,
                 max_batch=512):
        self.grid_shape = tuple(int(v) for v in grid_shape)
        H, W = self.grid_shape
        self.dl, self.n_steps, self.courant = float(dl), int(n_steps), float(courant)
        self.wavelength_nm = float(wavelength_nm)
        self.dt = courant * dl / math.sqrt(2.0)
        self.omega = 2 * math.pi / wavelength_nm
        self.dev = L.compute_device(device)
        self.max_batch = int(max_batch)
        f32 = dict(device=self.dev, dtype=torch.float32)
        by, cy = cpml_profile(H, pml_layers, dl, self.dt)
        bx, cx = cpml_profile(W, pml_layers, dl, self.dt)
        self.by = torch.as_tensor(np.asarray(by)).to(**f32).contiguous()
        self.ay = torch.as_tensor(1.0 + np.asarray(cy)).to(**f32).contiguous()
        self.bx = torch.as_tensor(np.asarray(bx)).to(**f32).contiguous()
        self.ax = torch.as_tensor(1.0 + np.asarray(cx)).to(**f32).contiguous()

        src = source or {"type": "gaussian_pulse"}
        box = source_box_2d(self.grid_shape, src)
        self._src = None
        if box is not None:
            prof = source_profile(src, box).to(**f32).contiguous()
            wave = torch.as_tensor(waveform_table(self.n_steps, self.dt, self.omega, src)).to(
                **f32).contiguous()
            self._src_keep = (prof, wave)
            self._src = L.Source(0, L.Box(*box), L.ptr(prof), L.ptr(wave))

        mon = monitor or {"box": (0, H, 0, W)}
        self.mbox = monitor_box(self.grid_shape, mon)
        lams = mon.get("wavelengths", [self.wavelength_nm])
        self.wavelengths = [float(v) for v in lams]
        omegas = [2 * math.pi / v for v in self.wavelengths]
        self.phase = torch.as_tensor(phase_table(self.n_steps, self.dt, omegas)).to(
            **f32).contiguous()
        _, _, y0, y1, x0, x1 = self.mbox
        self.out_cells = (y1 - y0, x1 - x0)

    @property
    def n_freq(self):
        return len(self.wavelengths)

    def forward(self, eps, return_field=False):
        """DFT monitor of every design (and, optionally, the final Ez)."""
        eps = torch.as_tensor(eps)
        if eps.dim() == 2:
            eps = eps.unsqueeze(0)
        if tuple(eps.shape[1:]) != self.grid_shape:
            raise RuntimeError(f"permittivity shape {tuple(eps.shape)} does not match "
                               f"(B, {self.grid_shape[0]}, {self.grid_shape[1]})")
        outs, fields = [], []
        for b0 in range(0, eps.shape[0], self.max_batch):
            o, f = self._chunk(eps[b0:b0 + self.max_batch])
            outs.append(o)
            fields.append(f)
        out = torch.cat(outs)
        return (out, torch.cat(fields)) if return_field else out

    def _chunk(self, eps):
        B = eps.shape[0]
        H, W = self.grid_shape
        N = H * W
        e64 = eps.detach().to(device=self.dev, dtype=torch.float64).contiguous()
        ce = torch.empty(B * N, device=self.dev, dtype=torch.float32)
        st = L.stream_handle(self.dev)
        L.check(L.lib().dnx_prepare_coef(L.ptr(e64), L.ptr(ce), B * N, self.dt / self.dl,
                                         L.DNX_F32, st))
        g = L.Grid2D(H, W, L.DNX_F32, 0, self.dt / self.dl, L.ptr(ce), L.ptr(self.by),
                     L.ptr(self.ay), L.ptr(self.bx), L.ptr(self.ax))
        state = torch.zeros(B, 3, N, device=self.dev, dtype=torch.float32)
        my, mx = self.out_cells
        acc = torch.zeros(B, self.n_freq, 2, my * mx, device=self.dev, dtype=torch.float32)
        mb = L.Box(*self.mbox)
        src = C.byref(self._src) if self._src is not None else None
        L.check(L.lib().dnx_fdtd2d_tm_batch(C.byref(g), B, L.ptr(state), self.n_steps, src,
                                            C.byref(mb), self.n_freq, L.ptr(self.phase),
                                            L.ptr(acc), st))
        out = torch.complex(acc[:, :, 0], acc[:, :, 1]).view(B, self.n_freq, my, mx)
        return out, state[:, 0].view(B, H, W)

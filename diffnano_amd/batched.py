"""Batched 3D FDTD: many designs with different grids advanced together.

One persistent launch per time step covers the tiles of every design (the
fused single-grid tile bodies, so each design gets exactly the result the
single-design solver gives); a K-step segment of each step sequence is
captured once into a HIP graph and replayed.  The reference solves designs
one at a time (``FDTDSolver3D._run_steps``, fdtd3d.py:471-504); this is the
batched form of that loop, with the same checkpoint-and-replay adjoint.

Usage::

    b = BatchedFDTD3D([(96, 96, 96), (128, 128, 128)], dl=25.0, pml_layers=12,
                      n_steps=3000, sources=[src0, src1], monitors=[mon0, mon1])
    outs = b.forward([eps0, eps1])       # list of complex monitor tensors
    loss = sum((F.abs() ** 2).sum() for F in outs); loss.backward()

Sources and monitors use the ``FDTDSolver3D`` dict formats (one source and
one monitor per design).
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from ._common import (COMPONENTS_3D, comp_mask_3d, cpml_profile, monitor_box, phase_table,
                      source_box_3d, source_profile, spare_device_bytes, waveform_table)
from ._engine import _begin, _end

__all__ = ["BatchedFDTD3D", "SharedStores"]


Design3D = L.Design3D
Pools3D = L.Pools3D


def _tail_count(nseg, seg_bytes, e_store_gb, dev):
    """Final segments whose E fits the budget (``e_store_gb``, or half the
    spare device memory above 16 GiB), at most all but the first."""
    if e_store_gb is not None and not e_store_gb >= 0:
        raise ValueError(f"e_store_gb must be None or >= 0, got {e_store_gb!r}")
    if e_store_gb is None:
        budget = max(0, (spare_device_bytes(dev) - (16 << 30)) // 2)
    else:
        budget = int(e_store_gb * 1e9)
    return int(min(max(nseg - 1, 0), budget // max(seg_bytes, 1)))


class SharedStores:
    """Checkpoints, replay store and final-segment E store for several
    batched plans that run one forward + backward at a time (the design
    loop's groups): sized for the largest plan (``max_cells``), allocated
    once.  The plans share a forward counter, so a backward after another
    sharing plan's forward raises (its checkpoints are gone)."""

    def __init__(self, max_cells, n_steps, checkpoint_every=64, device=None,
                 dtype=torch.float32, e_store_gb=None):
        self.max_cells, self.n_steps = int(max_cells), int(n_steps)
        self.K = max(2, int(checkpoint_every) + (int(checkpoint_every) & 1))
        self.dtype = dtype
        self.dev = L.compute_device(device)
        nseg = math.ceil(self.n_steps / self.K)
        elem = torch.tensor([], dtype=dtype).element_size()
        c = self.max_cells
        self.ckpt = torch.empty(nseg * 6 * c, dtype=dtype, device=self.dev)
        self.estore = torch.empty(max(self.K - 1, 1) * 3 * c, dtype=dtype, device=self.dev)
        self.tail_segments = _tail_count(nseg, self.K * 3 * c * elem, e_store_gb, self.dev)
        self.a_tail = (nseg - self.tail_segments) * self.K if self.tail_segments else self.n_steps
        self.etail = (torch.empty((self.n_steps - self.a_tail) * 3 * c, dtype=dtype,
                                  device=self.dev) if self.tail_segments else None)
        self._gen = 0


class BatchedFDTD3D:
    """A fixed batch of 3D designs (shapes, sources, monitors) sharing dl,
    wavelength, CPML depth, Courant number and step count."""

    def __init__(self, grid_shapes, sources, monitors, dl=25.0, wavelength_nm=1550.0,
                 pml_layers=12, n_steps=3000, courant=0.4, checkpoint_every=64, device=None,
                 dtype=torch.float32, use_graph=True, e_store_gb=None, stores=None):
        if len(sources) != len(grid_shapes) or len(monitors) != len(grid_shapes):
            raise ValueError("one source and one monitor per design")
        self.shapes = [tuple(int(v) for v in s) for s in grid_shapes]
        self.n = len(self.shapes)
        self.dl, self.courant, self.n_steps = float(dl), courant, int(n_steps)
        self.dt = courant * dl / math.sqrt(3.0)
        self.omega = 2 * math.pi / wavelength_nm
        self.wavelength_nm = wavelength_nm
        self.K = max(2, int(checkpoint_every) + (int(checkpoint_every) & 1))
        self.dtype = dtype
        self.code = L.dtype_code(dtype)
        self.dev = L.compute_device(device)
        self.out_device = torch.device(device) if device is not None else self.dev
        lib = L.lib()
        dev, dt = self.dev, dtype
        sizes = [int(np.prod(s)) for s in self.shapes]
        self.off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        tot = int(self.off[-1])
        self.sumN = tot
        nseg = math.ceil(self.n_steps / self.K)
        self.state = [torch.zeros(6 * tot, dtype=dt, device=dev) for _ in range(2)]
        self.lam = [torch.zeros(6 * tot, dtype=dt, device=dev) for _ in range(2)]
        self.S = torch.zeros(tot, dtype=dt, device=dev)
        if stores is not None:
            # another owner's checkpoints and E stores (sized for at least
            # this batch), used by one plan at a time
            if (stores.max_cells < tot or stores.n_steps != self.n_steps or stores.K != self.K
                    or stores.dtype != dt or stores.dev != dev):
                raise ValueError("BatchedFDTD3D: the shared stores do not fit this batch "
                                 "(cells, n_steps, checkpoint interval, dtype or device)")
            self._gen_owner = stores
            self.ckpt, self.estore, self.etail = stores.ckpt, stores.estore, stores.etail
            self.tail_segments, self.a_tail = stores.tail_segments, stores.a_tail
        else:
            self._gen_owner = self
            self.ckpt = torch.empty(nseg * 6 * tot, dtype=dt, device=dev)
            self.estore = torch.empty(max(self.K - 1, 1) * 3 * tot, dtype=dt, device=dev)
            # E of every step of the final segments, kept by a recording
            # forward so that the backward skips their replay: as many
            # segments as e_store_gb (default: half the spare device memory
            # above 16 GiB) holds, at most all but the first
            elem = torch.tensor([], dtype=dt).element_size()
            self.tail_segments = _tail_count(nseg, self.K * 3 * tot * elem, e_store_gb, dev)
            self.a_tail = ((nseg - self.tail_segments) * self.K if self.tail_segments
                           else self.n_steps)
            self.etail = (torch.empty((self.n_steps - self.a_tail) * 3 * tot, dtype=dt,
                                      device=dev) if self.tail_segments else None)
        self.ce = torch.ones(tot, dtype=dt, device=dev)
        self.eps = torch.ones(tot, dtype=torch.float64, device=dev)
        self._keep = []
        # forward generation: the backward of a forward is valid only while no
        # other forward has overwritten the plan's checkpoints, coefficients
        # and eps (they live in plan-wide buffers, the checkpoints possibly in
        # stores shared with other plans: then the counter is the stores')
        self._gen = 0
        arr = (Design3D * self.n)()
        self.accs, self.grads, self.mon_shapes = [], [], []
        for d, (shape, src, mon) in enumerate(zip(self.shapes, sources, monitors)):
            D, H, W = shape
            by, cy = cpml_profile(H, pml_layers, dl, self.dt)
            bx, cx = cpml_profile(W, pml_layers, dl, self.dt)
            prof = [torch.as_tensor(v).to(device=dev, dtype=dt).contiguous()
                    for v in (by, 1.0 + cy, bx, 1.0 + cx)]
            self._keep += prof
            ce_d = self.ce[self.off[d]:self.off[d + 1]]
            arr[d].grid = L.Grid3D(D, H, W, self.code, self.dt / self.dl, ce_d.data_ptr(),
                                   *[p.data_ptr() for p in prof])
            box = source_box_3d(shape, src)
            if box is None:
                arr[d].src.comp = -1
            else:
                pr = source_profile(src, box).to(device=dev, dtype=dt).contiguous()
                wv = torch.as_tensor(waveform_table(self.n_steps, self.dt, self.omega, src)).to(
                    device=dev, dtype=dt).contiguous()
                self._keep += [pr, wv]
                arr[d].src.comp = COMPONENTS_3D[src.get("component", "Ez")]
                arr[d].src.box = L.Box(*box)
                arr[d].src.profile = pr.data_ptr()
                arr[d].src.wave = wv.data_ptr()
            mbox = monitor_box(shape, mon)
            lams = mon.get("wavelengths", [wavelength_nm])
            omegas = [2 * math.pi / float(v) for v in lams]
            mask = comp_mask_3d(mon.get("components"))
            ncomp = bin(mask).count("1")
            z0, z1, y0, y1, x0, x1 = mbox
            oshape = (len(omegas), ncomp, z1 - z0, y1 - y0, x1 - x0, 2)
            ph = torch.as_tensor(phase_table(self.n_steps, self.dt, omegas)).to(dev).contiguous()
            acc = torch.zeros(oshape, dtype=torch.float64, device=dev)
            grd = torch.zeros(oshape, dtype=torch.float64, device=dev)
            self._keep.append(ph)
            self.accs.append(acc)
            self.grads.append(grd)
            self.mon_shapes.append(oshape)
            arr[d].mon.box = L.Box(*mbox)
            arr[d].mon.comp_mask = mask
            arr[d].mon.n_freq = len(omegas)
            arr[d].mon.phase = ph.data_ptr()
            arr[d].mon.acc = acc.data_ptr()
            arr[d].mon.grad = grd.data_ptr()
        self._designs = arr
        pools = Pools3D()
        pools.state[0], pools.state[1] = self.state[0].data_ptr(), self.state[1].data_ptr()
        pools.lam[0], pools.lam[1] = self.lam[0].data_ptr(), self.lam[1].data_ptr()
        pools.ckpt = self.ckpt.data_ptr()
        pools.estore = self.estore.data_ptr()
        pools.S = self.S.data_ptr()
        pools.etail = self.etail.data_ptr() if self.etail is not None else None
        pools.a_tail = self.a_tail if self.etail is not None else 0
        side_n = int(lib.dnx_batch3d_side_elems(arr, self.n))
        if side_n < 0:
            raise ValueError("BatchedFDTD3D: a grid (nx <= 256) or monitor box the batched "
                             "kernels cannot take")
        self.side = torch.empty(max(side_n, 1), dtype=dt, device=dev)
        pools.side = self.side.data_ptr()
        self._pools = pools
        nbytes = lib.dnx_batch3d_workspace_bytes(arr, self.n)
        if nbytes < 0:
            raise RuntimeError("libdnx_fdtd: batched workspace query failed")
        self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        plan = C.c_void_p()
        L.check(lib.dnx_batch3d_create(arr, self.n, self.n_steps, self.K, C.byref(pools),
                                       self.ws.data_ptr(), int(nbytes), 1 if use_graph else 0,
                                       L.stream_handle(dev), C.byref(plan)))
        self._plan = plan

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                torch.cuda.synchronize(self.dev)
                L.lib().dnx_batch3d_destroy(plan)
            except Exception:
                pass
            self._plan = None

    @property
    def uses_graph(self):
        return bool(L.lib().dnx_batch3d_uses_graph(self._plan))

    @property
    def uses_pairs(self):
        """(forward, adjoint) run the two-step kernels."""
        v = int(L.lib().dnx_batch3d_uses_pairs(self._plan))
        return bool(v & 1), bool(v & 2)

    # -- device steps ---------------------------------------------------------
    def _set_eps(self, eps_list):
        st = L.stream_handle(self.dev)
        for d, e in enumerate(eps_list):
            if tuple(e.shape) != self.shapes[d]:
                raise RuntimeError(f"design {d}: eps shape {tuple(e.shape)} != {self.shapes[d]}")
            seg = self.eps[self.off[d]:self.off[d + 1]]
            seg.copy_(e.detach().reshape(-1).to(device=self.dev, dtype=torch.float64))
        L.check(L.lib().dnx_prepare_coef(self.eps.data_ptr(), self.ce.data_ptr(), self.sumN,
                                         self.dt / self.dl, self.code, st))

    def _run_forward(self, record):
        self.state[0].zero_()
        for a in self.accs:
            a.zero_()
        _begin("forward")
        L.check(L.lib().dnx_batch3d_forward(self._plan, 1 if record else 0,
                                            L.stream_handle(self.dev)))
        _end("forward")

    def _run_backward(self, grads):
        for g_dst, g in zip(self.grads, grads):
            if g is None:
                g_dst.zero_()
            else:
                g_dst.copy_(g.to(device=self.dev, dtype=torch.float64))
        _begin("adjoint")
        L.check(L.lib().dnx_batch3d_backward(self._plan, L.stream_handle(self.dev)))
        _end("adjoint")
        grad = torch.empty(self.sumN, dtype=torch.float64, device=self.dev)
        L.check(L.lib().dnx_finalize_grad(self.S.data_ptr(), self.eps.data_ptr(),
                                          grad.data_ptr(), self.sumN, self.code,
                                          L.stream_handle(self.dev)))
        return [grad[self.off[d]:self.off[d + 1]].view(self.shapes[d]) for d in range(self.n)]

    def final_fields(self):
        """(Ez, Ex, Ey) of every design after the last forward, float64."""
        buf = self.state[int(L.lib().dnx_batch3d_final_slot(self._plan))]
        out = []
        for d, shape in enumerate(self.shapes):
            N = int(np.prod(shape))
            base = 6 * int(self.off[d])
            e = [buf[base + c * N: base + (c + 1) * N].view(shape) for c in (2, 0, 1)]
            out.append(torch.stack(e).to(torch.float64))
        return out

    def forward(self, eps_list):
        """Monitor outputs per design: complex128 (n_freq, n_comp, bz, by, bx),
        differentiable w.r.t. every eps (checkpoints are recorded only when
        grad mode is on and some eps requires grad)."""
        if torch.is_grad_enabled() and any(e.requires_grad for e in eps_list):
            outs = _BatchedFn.apply(self, *eps_list)
        else:
            outs = self._forward_detached(eps_list)
        return [torch.view_as_complex(o) for o in outs]

    def _next_gen(self):
        self._gen_owner._gen += 1
        return self._gen_owner._gen

    def _current_gen(self):
        return self._gen_owner._gen

    def _forward_detached(self, eps_list):
        self._next_gen()
        self._set_eps(eps_list)
        self._run_forward(record=False)
        return tuple(a.clone().to(self.out_device) for a in self.accs)


class _BatchedFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *eps_list):
        ctx.batch = batch
        ctx.devices = [e.device for e in eps_list]
        ctx.recorded = True  # BatchedFDTD3D.forward applies this only to record
        ctx.gen = batch._next_gen()
        batch._set_eps(eps_list)
        batch._run_forward(record=ctx.recorded)
        return tuple(a.clone().to(batch.out_device) for a in batch.accs)

    @staticmethod
    def backward(ctx, *grads):
        b = ctx.batch
        if ctx.gen != b._current_gen():
            raise RuntimeError(
                "BatchedFDTD3D: another forward ran on this batch (or on one sharing its "
                "stores) after the one being differentiated and overwrote its checkpoints and "
                "coefficients; call backward before the next forward (or use a separate "
                "BatchedFDTD3D)")
        if not ctx.recorded:
            raise RuntimeError("BatchedFDTD3D: this forward stored no checkpoints")
        gl = b._run_backward(list(grads))
        return (None,) + tuple(g.to(dv) for g, dv in zip(gl, ctx.devices))

"""Device-side driver of one FDTD solve and its adjoint (2D or 3D).

``Engine`` owns the uploaded coefficients and turns source/monitor specs
into the C-ABI descriptors; ``_TimeReversalFDTD`` is the autograd Function
behind every ``backward=`` mode of the solvers.  Its forward stores full
state checkpoints every K steps; its backward walks the segments in
reverse, replays each from its checkpoint while storing E, and runs the
hand-written adjoint kernels over it, accumulating dL/deps in the compute
dtype.  All modes return the same (autograd-equal) gradient, as in the
reference where "auto", "checkpoint" and "time_reversal" agree
(fdtd3d.py:88-219, 506-562); they differ only in the checkpoint interval.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L
from ._common import phase_table, spare_device_bytes


class PhaseTimer:
    """Optional HIP-event timing of the forward / replay / adjoint phases,
    recorded on the stream the kernels run on.  Install with
    ``set_phase_timer(PhaseTimer())``; ``totals_ms()`` after a synchronize."""

    def __init__(self):
        self.pairs = {}

    def begin(self, name):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.pairs.setdefault(name, []).append([ev, None])

    def end(self, name):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.pairs[name][-1][1] = ev

    def totals_ms(self):
        return {k: sum(a.elapsed_time(b) for a, b in v) for k, v in self.pairs.items()}

    def reset(self):
        self.pairs = {}


_timer = None


def set_phase_timer(timer):
    global _timer
    _timer = timer


def _begin(name):
    if _timer is not None:
        _timer.begin(name)


def _end(name):
    if _timer is not None:
        _timer.end(name)


class SourceSpec:
    """One soft source: component index, box, amplitude profile (float64
    CPU tensor over the box) and waveform table (float64 numpy, length T)."""

    __slots__ = ("comp", "box", "profile", "wave")

    def __init__(self, comp, box, profile, wave):
        self.comp, self.box, self.profile, self.wave = comp, box, profile, wave


class MonitorSpec:
    """One DFT monitor: box, component mask, angular frequencies."""

    __slots__ = ("box", "comp_mask", "omegas")

    def __init__(self, box, comp_mask, omegas):
        self.box, self.comp_mask, self.omegas = box, comp_mask, list(omegas)

    @property
    def n_comp(self):
        return bin(self.comp_mask).count("1")

    def out_shape(self, dim):
        z0, z1, y0, y1, x0, x1 = self.box
        cells = (z1 - z0, y1 - y0, x1 - x0) if dim == 3 else (y1 - y0, x1 - x0)
        comp = (self.n_comp,) if dim == 3 else ()
        return (len(self.omegas),) + comp + cells + (2,)


class Engine:
    def __init__(self, dim, shape, dl, dt, profiles, dtype, device, te=False):
        self.dim = dim
        self.shape = tuple(int(v) for v in shape)
        self.N = int(np.prod(self.shape))
        self.ns = 6 if dim == 3 else 3
        self.ne = 3 if dim == 3 else 1
        self.te = bool(te)
        self.dtype = dtype
        self.code = L.dtype_code(dtype)
        self.dev = device
        self.dl, self.dt = float(dl), float(dt)
        by, cy, bx, cx = profiles
        mk = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64)).to(  # noqa: E731
            device=device, dtype=dtype).contiguous()
        self.by, self.ay = mk(by), mk(1.0 + np.asarray(cy))
        self.bx, self.ax = mk(bx), mk(1.0 + np.asarray(cx))
        self.ce = None
        self.eps = None
        self._keep = []

    # -- set-up -------------------------------------------------------------
    def stream(self):
        return L.stream_handle(self.dev)

    def set_eps(self, eps64):
        if tuple(eps64.shape) != self.shape:
            raise RuntimeError(f"permittivity shape {tuple(eps64.shape)} does not match the solver "
                               f"grid {self.shape}")
        self.eps = eps64.detach().to(device=self.dev, dtype=torch.float64).contiguous()
        self.ce = torch.empty(self.N, dtype=self.dtype, device=self.dev)
        L.check(L.lib().dnx_prepare_coef(L.ptr(self.eps), L.ptr(self.ce), self.N,
                                         self.dt / self.dl, self.code, self.stream()))

    def grid(self):
        if self.dim == 3:
            g = L.Grid3D(self.shape[0], self.shape[1], self.shape[2], self.code, self.dt / self.dl,
                         L.ptr(self.ce), L.ptr(self.by), L.ptr(self.ay), L.ptr(self.bx),
                         L.ptr(self.ax))
        else:
            g = L.Grid2D(self.shape[0], self.shape[1], self.code, int(self.te), self.dt / self.dl,
                         L.ptr(self.ce), L.ptr(self.by), L.ptr(self.ay), L.ptr(self.bx),
                         L.ptr(self.ax))
        self._grid = g
        return C.byref(g)

    def sources(self, specs):
        if not specs:
            return None, 0
        arr = (L.Source * len(specs))()
        for s, sp in zip(arr, specs):
            prof = sp.profile.to(device=self.dev, dtype=self.dtype).contiguous()
            wave = torch.as_tensor(sp.wave).to(device=self.dev, dtype=self.dtype).contiguous()
            self._keep += [prof, wave]
            s.comp = sp.comp
            s.box = L.Box(*sp.box)
            s.profile = L.ptr(prof)
            s.wave = L.ptr(wave)
        return arr, len(specs)

    def monitors(self, specs, n_steps, grads=None):
        """Descriptor array plus freshly zeroed accumulators (forward) or the
        given gradients (adjoint)."""
        if not specs:
            return None, 0, []
        arr = (L.Monitor * len(specs))()
        accs = []
        for j, (m, sp) in enumerate(zip(arr, specs)):
            ph = torch.as_tensor(phase_table(n_steps, self.dt, sp.omegas)).to(self.dev).contiguous()
            self._keep.append(ph)
            m.box = L.Box(*sp.box)
            m.comp_mask = sp.comp_mask
            m.n_freq = len(sp.omegas)
            m.phase = L.ptr(ph)
            if grads is None:
                acc = torch.zeros(sp.out_shape(self.dim), dtype=torch.float64, device=self.dev)
                accs.append(acc)
                m.acc = L.ptr(acc)
                m.grad = None
            else:
                g = grads[j]
                if g is None:
                    g = torch.zeros(sp.out_shape(self.dim), dtype=torch.float64, device=self.dev)
                g = g.to(device=self.dev, dtype=torch.float64).contiguous()
                self._keep.append(g)
                m.acc = None
                m.grad = L.ptr(g)
        return arr, len(specs), accs

    # -- kernels ------------------------------------------------------------
    def run(self, state, n0, n_steps, srcs, mons, ckpt=None, ckpt_every=0, estore=None,
            probe=-1, probe_out=None, work=None, init=None):
        """Advance ``state`` (or, 3D, the state in ``init``, left unchanged);
        returns the buffer (``state`` or ``work``) that holds the final state
        (the fused kernels ping-pong between them)."""
        if work is None:
            work = torch.empty_like(state)
        fin = C.c_int32(0)
        if self.dim == 3:
            L.check(L.lib().dnx_fdtd3d_run_from(
                self.grid(), L.ptr(init), L.ptr(state), L.ptr(work), int(n0), int(n_steps),
                srcs[0], srcs[1], mons[0], mons[1], L.ptr(ckpt), int(ckpt_every), L.ptr(estore),
                int(probe), L.ptr(probe_out), self.stream(), C.byref(fin)))
            return work if fin.value else state
        if init is not None:
            state.copy_(init)
        L.check(L.lib().dnx_fdtd2d_run(self.grid(), L.ptr(state), L.ptr(work), int(n0),
                                       int(n_steps), srcs[0], srcs[1], mons[0], mons[1],
                                       L.ptr(ckpt), int(ckpt_every), L.ptr(estore), int(probe),
                                       L.ptr(probe_out), self.stream(), C.byref(fin)))
        return work if fin.value else state

    def adjoint_init(self, lam, e_T, mons, n_total, S):
        fn = L.lib().dnx_fdtd3d_adjoint_init if self.dim == 3 else L.lib().dnx_fdtd2d_adjoint_init
        L.check(fn(self.grid(), L.ptr(lam), L.ptr(e_T), mons[0], mons[1], int(n_total), L.ptr(S),
                   self.stream()))

    def adjoint_workspace(self, mons):
        """Device scratch the 3D adjoint sweep needs for these monitors (a
        torch allocation, reused by every sweep of one backward), or None."""
        if self.dim != 3:
            return None
        n = int(L.lib().dnx_fdtd3d_adjoint_workspace_bytes(self.grid(), mons[0], mons[1]))
        if n < 0:
            raise RuntimeError("libdnx_fdtd: adjoint workspace query failed")
        return torch.empty(n, dtype=torch.uint8, device=self.dev) if n else None

    def adjoint_run(self, lam, m_hi, m_lo, e_lo, estore, srcs, mons, S, work, ws=None):
        """Reverse sweep; returns (final, spare) adjoint buffers.  ``ws``:
        the 3D workspace from ``adjoint_workspace`` (allocated here when
        not given)."""
        fin = C.c_int32(0)
        if self.dim == 3:
            if ws is None:
                ws = self.adjoint_workspace(mons)
            nbytes = 0 if ws is None else ws.numel()
            L.check(L.lib().dnx_fdtd3d_adjoint_run(
                self.grid(), L.ptr(lam), L.ptr(work), int(m_hi), int(m_lo), L.ptr(e_lo),
                L.ptr(estore), srcs[0], srcs[1], mons[0], mons[1], L.ptr(S), L.ptr(ws), nbytes,
                self.stream(), C.byref(fin)))
        else:
            L.check(L.lib().dnx_fdtd2d_adjoint_run(
                self.grid(), L.ptr(lam), L.ptr(work), int(m_hi), int(m_lo), L.ptr(e_lo),
                L.ptr(estore), srcs[0], srcs[1], mons[0], mons[1], L.ptr(S), self.stream(),
                C.byref(fin)))
        return (work, lam) if fin.value else (lam, work)

    def finalize_grad(self, S):
        g = torch.empty(self.N, dtype=torch.float64, device=self.dev)
        L.check(L.lib().dnx_finalize_grad(L.ptr(S), L.ptr(self.eps), L.ptr(g), self.N, self.code,
                                          self.stream()))
        return g.view(self.shape)


class Plan:
    """Everything a solve needs besides eps: engine, step count, checkpoint
    interval, sources, monitors, output layout and optional probe."""

    def __init__(self, engine, n_steps, ckpt_every, sources, monitors, out_order, out_unsqueeze,
                 probe=-1, out_device=None):
        self.engine = engine
        self.out_device = engine.dev if out_device is None else torch.device(out_device)
        self.n_steps = int(n_steps)
        self.ckpt_every = max(1, int(ckpt_every))
        self.sources = list(sources)
        self.monitors = list(monitors)
        self.out_order = list(out_order)
        self.out_unsqueeze = out_unsqueeze
        self.probe = int(probe)
        self.probe_values = None
        # device bytes the time-reversal forward may spend on E of the final
        # steps: None = half the spare device memory above 16 GiB
        self.e_store_bytes = None
        # set by the time-reversal forward: checkpoint segments, and how many
        # trailing ones kept E (their replay is skipped)
        self.segments = 0
        self.tail_segments = 0


_KEEP_FREE = 16 << 30


def _tail_segments(plan, nseg, K, seg_bytes):
    """Trailing checkpoint segments whose E (every step) the forward keeps,
    so that the backward runs their adjoint without replaying them: as many
    as ``plan.e_store_bytes`` (or, by default, half of the free device
    memory above 16 GiB) holds, at most all but the first."""
    eng = plan.engine
    if plan.probe >= 0 or nseg < 2 or seg_bytes <= 0:
        return 0
    budget = plan.e_store_bytes
    if budget is None:
        budget = max(0, (spare_device_bytes(eng.dev) - _KEEP_FREE) // 2)
    return int(min(nseg - 1, int(budget) // seg_bytes))


def _field_out(plan, state):
    eng = plan.engine
    N = eng.N
    comps = [state[c * N:(c + 1) * N].view(eng.shape) for c in plan.out_order]
    out = torch.stack(comps, 0) if len(comps) > 1 else comps[0]
    if plan.out_unsqueeze:
        out = out.unsqueeze(0)
    return out.to(device=plan.out_device, dtype=torch.float64)


def forward_only(plan, eps64):
    """Run without recording anything for a gradient; returns (field,
    monitor accumulators, final state).  Outputs are on ``plan.out_device``."""
    eng = plan.engine
    eng.set_eps(eps64)
    T = plan.n_steps
    state = torch.zeros(eng.ns * eng.N, dtype=eng.dtype, device=eng.dev)
    srcs = eng.sources(plan.sources)
    mons = eng.monitors(plan.monitors, T)
    probe_out = None
    if plan.probe >= 0:
        probe_out = torch.zeros(max(T, 1), dtype=eng.dtype, device=eng.dev)
    state = eng.run(state, 0, T, srcs[:2], mons[:2], probe=plan.probe, probe_out=probe_out)
    plan.probe_values = probe_out[:T] if probe_out is not None else None
    return _field_out(plan, state), [m.to(plan.out_device) for m in mons[2]], state


class _TimeReversalFDTD(torch.autograd.Function):
    """FDTD forward with a checkpoint-and-replay discrete adjoint.

    Forward: state checkpoints every K steps, and E of every step of the
    last M segments (as many as the plan's E budget holds; the reference's
    time-reversal forward keeps E of every step, fdtd3d.py:147-150).
    Backward: for each segment in reverse, replay from its checkpoint storing
    E (not needed for the last M), then run the adjoint kernels over the
    segment.  Memory: T/K full states + (K-1) E fields + M K E fields.  The
    input eps may live on any device (it is
    moved to the GPU inside); outputs are returned on ``plan.out_device``
    and the gradient on eps's device, so this Function is the grad_fn of
    the solver's field.
    """

    @staticmethod
    def forward(ctx, eps, plan):
        eng = plan.engine
        ctx.eps_device = eps.device
        eng.set_eps(eps)
        T = plan.n_steps
        K = plan.ckpt_every
        N, ns = eng.N, eng.ns
        state = torch.zeros(ns * N, dtype=eng.dtype, device=eng.dev)
        srcs = eng.sources(plan.sources)
        mons = eng.monitors(plan.monitors, T)
        need = ctx.needs_input_grad[0]
        nseg = math.ceil(T / K) if (need and T > 0) else 0
        ckpt = torch.empty(nseg * ns * N, dtype=eng.dtype, device=eng.dev) if nseg else None
        probe_out = None
        if plan.probe >= 0:
            probe_out = torch.zeros(max(T, 1), dtype=eng.dtype, device=eng.dev)
        ne = eng.ne
        elem = torch.tensor([], dtype=eng.dtype).element_size()
        n_tail = _tail_segments(plan, nseg, K, K * ne * N * elem) if nseg else 0
        a_tail = (nseg - n_tail) * K  # first step whose segment keeps its E
        plan.segments, plan.tail_segments = nseg, n_tail
        tail = (torch.empty((T - a_tail) * ne * N, dtype=eng.dtype, device=eng.dev)
                if n_tail else None)
        _begin("forward")
        if n_tail:
            s0, w0 = state, torch.empty_like(state)
            mid = eng.run(s0, 0, a_tail, srcs[:2], mons[:2], ckpt, K, work=w0)
            state = eng.run(mid, a_tail, T - a_tail, srcs[:2], mons[:2],
                            ckpt[(a_tail // K) * ns * N:], K, estore=tail,
                            work=w0 if mid is s0 else s0)
        else:
            state = eng.run(state, 0, T, srcs[:2], mons[:2], ckpt, K, probe=plan.probe,
                            probe_out=probe_out)
        _end("forward")
        plan.probe_values = probe_out[:T] if probe_out is not None else None
        ctx.plan = plan
        # Saved through autograd: an in-place change of eps before backward
        # raises (the coefficients and -S/eps come from this eps), and the
        # checkpoints, the E store and the final state are released once a
        # backward without retain_graph has run, not when the graph dies.
        ctx.save_for_backward(eps, ckpt, tail, state)
        ctx.a_tail = a_tail if n_tail else T
        ctx.srcs = srcs
        outs = [_field_out(plan, state)]
        for acc in mons[2]:
            outs.append(acc.to(plan.out_device))
        return tuple(outs)

    @staticmethod
    def backward(ctx, g_field, *g_mons):
        _, ckpt, tail, state = ctx.saved_tensors  # raises if eps was modified in place
        plan = ctx.plan
        eng = plan.engine
        T, K = plan.n_steps, plan.ckpt_every
        N, ns, ne = eng.N, eng.ns, eng.ne
        dev, dt = eng.dev, eng.dtype
        if T == 0:
            return torch.zeros(eng.shape, dtype=torch.float64, device=ctx.eps_device), None
        lam = torch.zeros(ns * N, dtype=dt, device=dev)
        if g_field is not None:
            g = g_field.to(dev).reshape((len(plan.out_order),) + eng.shape)
            for k, c in enumerate(plan.out_order):
                lam[c * N:(c + 1) * N] = g[k].reshape(-1).to(dt)
        S = torch.empty(N, dtype=dt, device=dev)
        srcs = ctx.srcs
        mons = eng.monitors(plan.monitors, T, grads=list(g_mons))
        eng.adjoint_init(lam, state[:ne * N], mons[:2], T, S)
        nseg = math.ceil(T / K)
        estore = torch.empty(max(K - 1, 1) * ne * N, dtype=dt, device=dev)
        work = torch.empty(ns * N, dtype=dt, device=dev)
        work2 = torch.empty(ns * N, dtype=dt, device=dev)
        lam_spare = torch.empty(ns * N, dtype=dt, device=dev)
        ws = eng.adjoint_workspace(mons[:2])
        for k in range(nseg - 1, -1, -1):
            a, b = k * K, min(T, (k + 1) * K)
            slot = ckpt[k * ns * N:(k + 1) * ns * N]
            es = estore
            if a >= ctx.a_tail:  # E^{a+1} .. E^{b-1} kept by the forward
                es = tail[(a - ctx.a_tail) * ne * N:]
            elif b - a > 1:
                _begin("replay")
                eng.run(work, a, b - a - 1, srcs[:2], (None, 0), estore=estore, work=work2,
                        init=slot)
                _end("replay")
            _begin("adjoint")
            lam, lam_spare = eng.adjoint_run(lam, b, a, slot[:ne * N], es, srcs[:2], mons[:2],
                                             S, lam_spare, ws)
            _end("adjoint")
        grad = eng.finalize_grad(S)
        return grad.to(ctx.eps_device), None

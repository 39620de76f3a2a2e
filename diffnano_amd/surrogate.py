"""FNO surrogate of the 2D TM FDTD and the surrogate-prescreened design loop.

``FieldSurrogate`` maps a permittivity map to the complex DFT of Ez over the
grid at several wavelengths.  It predicts the field scattered by the design
(the total field minus the field of the empty SiO2 background, which the
FDTD computes once) with a 2D FNO whose forward and backward run on the
hand-written BF16 MFMA kernels (``fno.FNO2d``); the FP32 master weights are
trained with Adam on fields generated by the batched FDTD
(``batched2d.BatchedTM2D``).

``PrescreenDesigner`` is a (1 + lambda) evolution loop over blob designs
that follows the reference's two-stage prescreen (``TwoStageOptimizer``,
proxy_prescreen.py:147-274: the proxy ranks many candidates, the exact
solver re-evaluates the top k and the best evaluated candidate is kept) with the
FNO as the proxy and the FDTD as the evaluator, plus the correction rule of
``NeuralSurrogate.forward`` (surrogate.py:237-239: fall back to the full
solver): when the surrogate's relative error on the evaluated candidates
exceeds ``max_rel_err``, that step evaluates every candidate with the FDTD.
Every accepted design has been solved by the FDTD, so the figure of merit
the loop reports is always an FDTD value.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import workloads
from .batched2d import BatchedTM2D
from .fno import FNO2d

__all__ = ["FieldSurrogate", "PrescreenDesigner", "generate_dataset", "plain_fom"]


def generate_dataset(cfg, seeds, device=None, batch=256):
    """Blob designs from ``seeds`` and their full-grid DFT fields.

    Returns (noise [N, S, S] float64, eps [N, H, W] float32, fields
    [N, n_freq, H, W] complex64), all on the GPU."""
    solver = BatchedTM2D(**cfg.solver, source=cfg.source,
                         monitor={"box": (0, cfg.H, 0, cfg.W), "wavelengths": cfg.wavelengths},
                         device=device, max_batch=batch)
    dev = solver.dev
    noise = torch.stack([workloads.blob_noise(s, cfg.size) for s in seeds]).to(dev)
    eps, fields = [], []
    for b0 in range(0, len(seeds), batch):
        e = cfg.eps(noise[b0:b0 + batch])
        fields.append(solver.forward(e))
        eps.append(e.to(torch.float32))
    return noise, torch.cat(eps), torch.cat(fields)


class FieldSurrogate:
    """eps [B, H, W] -> complex DFT Ez [B, n_freq, H, W] (incident + FNO).

    Input channels: the normalised permittivity (0 for SiO2, 1 for Si), the
    cell coordinates, and the background field (real and imaginary parts,
    normalised), which carries the source's position and phase."""

    def __init__(self, cfg, width=32, modes=(16, 16), layers=4, hidden=128, device=None,
                 seed=0):
        self.cfg = cfg
        H, W = cfg.H, cfg.W
        self.nf = len(cfg.wavelengths)
        self.fdtd = BatchedTM2D(**cfg.solver, source=cfg.source,
                                 monitor={"box": (0, H, 0, W), "wavelengths": cfg.wavelengths},
                                 device=device)
        self.dev = self.fdtd.dev
        self.fno = FNO2d(H, W, in_ch=3 + 2 * self.nf, out_ch=2 * self.nf, width=width,
                         modes=modes, layers=layers, hidden=hidden, device=self.dev, seed=seed)
        bg = torch.full((1, H, W), workloads.EPS_SIO2, dtype=torch.float64, device=self.dev)
        self.incident = self.fdtd.forward(bg)[0]                  # [nf, H, W] complex64
        self.scale = torch.ones(self.nf, device=self.dev)
        yy, xx = torch.meshgrid(torch.arange(H, device=self.dev) / (H - 1),
                                torch.arange(W, device=self.dev) / (W - 1), indexing="ij")
        inc = self.incident / self.incident.abs().pow(2).mean(dim=(1, 2), keepdim=True).sqrt()
        # fixed input channels: coordinates and the background field
        self.fixed = torch.cat([torch.stack([yy, xx]), inc.real, inc.imag]).float()

    # -- encodings --------------------------------------------------------
    def encode(self, eps):
        e = (eps.to(self.dev, torch.float32) - workloads.EPS_SIO2) / (
            workloads.EPS_SI - workloads.EPS_SIO2)
        B = e.shape[0]
        return torch.cat([e.unsqueeze(1), self.fixed.expand(B, -1, -1, -1)], dim=1)

    def to_target(self, fields):
        s = (fields - self.incident) / self.scale[:, None, None]
        return torch.cat([s.real, s.imag], dim=1)                  # [B, 2 nf, H, W]

    def from_output(self, out):
        nf = self.nf
        s = torch.complex(out[:, :nf], out[:, nf:]) * self.scale[:, None, None]
        return self.incident + s

    def predict(self, eps, batch=256):
        outs = []
        with torch.no_grad():
            for b0 in range(0, eps.shape[0], batch):
                outs.append(self.from_output(self.fno(self.encode(eps[b0:b0 + batch]))))
        return torch.cat(outs)

    # -- training ---------------------------------------------------------
    def loss(self, y, fields, line_weight=1.0):
        """Mean relative L2^2 of the scattered field plus ``line_weight``
        times the relative L2^2 of the total field on the figure-of-merit
        line at its wavelength (where the prescreen reads the prediction)."""
        t = self.to_target(fields)
        num = (y - t).pow(2).flatten(1).sum(1)
        den = t.pow(2).flatten(1).sum(1).clamp_min(1e-12)
        out = (num / den).mean()
        if line_weight:
            y0, y1, x0, x1 = self.cfg.line
            fi, nf = self.cfg.fom_index, self.nf
            sc = self.scale[fi]
            inc = self.incident[fi, y0:y1, x0:x1]
            pr = inc.real + sc * y[:, fi, y0:y1, x0:x1]
            pi = inc.imag + sc * y[:, nf + fi, y0:y1, x0:x1]
            tl = fields[:, fi, y0:y1, x0:x1]
            ln = ((pr - tl.real).pow(2) + (pi - tl.imag).pow(2)).flatten(1).sum(1)
            ld = tl.abs().pow(2).flatten(1).sum(1).clamp_min(1e-12)
            out = out + line_weight * (ln / ld).mean()
        return out

    def fit(self, eps, fields, epochs=20, batch=64, lr=3e-3, seed=0, log=None, time_limit=None,
            weight_decay=0.0, line_weight=0.0):
        """AdamW on the FP32 master weights (loss: ``loss``), one-cycle
        cosine learning rate.  Returns the loss per epoch."""
        sc = (fields - self.incident).abs().pow(2).mean(dim=(0, 2, 3)).sqrt()
        self.scale = torch.where(sc > 0, sc, torch.ones_like(sc)).float()
        params = [self.fno.params[k] for k in self.fno.names]
        opt = torch.optim.AdamW(params, lr=lr, weight_decay=weight_decay)
        n = eps.shape[0]
        steps = epochs * (n // batch)
        sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=max(1, steps),
                                                    pct_start=0.1)
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            tot = torch.zeros((), device=self.dev)
            for i in range(n // batch):
                idx = perm[i * batch:(i + 1) * batch]
                y = self.fno(self.encode(eps[idx]))
                loss = self.loss(y, fields[idx], line_weight)
                opt.zero_grad(set_to_none=True)
                loss.backward()
                opt.step()
                sched.step()
                tot += loss.detach()
            hist.append(float(tot) / max(1, n // batch))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5f} ({time.perf_counter() - t0:.1f} s)")
            if time_limit is not None and time.perf_counter() - t0 > time_limit:
                break
        return hist

    def rel_error(self, eps, fields, region=None):
        """Per-sample relative L2 error of the predicted complex fields
        (optionally over region = (y0, y1, x0, x1))."""
        p = self.predict(eps)
        if region is not None:
            y0, y1, x0, x1 = region
            p, fields = p[..., y0:y1, x0:x1], fields[..., y0:y1, x0:x1]
        return ((p - fields).abs().pow(2).flatten(1).sum(1).sqrt()
                / fields.abs().pow(2).flatten(1).sum(1).sqrt())

    # -- persistence ------------------------------------------------------
    def state_dict(self):
        d = {k: v.detach().cpu() for k, v in self.fno.params.items()}
        d["_scale"] = self.scale.cpu()
        return d

    def load_state_dict(self, d):
        with torch.no_grad():
            for k, v in self.fno.params.items():
                v.copy_(d[k].to(v.device))
        self.scale = d["_scale"].to(self.dev).float()


def plain_fom(cfg, eps, device=None):
    """FoM of each design from the single-design solver (FDTDSolver2D with a
    DFT monitor on the FoM line): the plain per-design GPU loop."""
    from .fdtd2d import FDTDSolver2D
    solver = FDTDSolver2D(**cfg.solver, device=device or "cuda")
    y0, y1, x0, x1 = cfg.line
    mon = [{"box": (y0, y1, x0, x1), "wavelengths": [cfg.wavelengths[cfg.fom_index]]}]
    out = []
    with torch.no_grad():
        for e in eps:
            r = solver.forward(e, source=cfg.source, monitors=mon)
            out.append(float((r.metadata["monitors"][0].abs() ** 2).sum()))
    return torch.tensor(out, dtype=torch.float64)


class PrescreenDesigner:
    """(1 + n_candidates) evolution of a blob design.

    Each step mutates the parent's noise in ``n_candidates`` random patches,
    maps them to permittivity, ranks them and keeps the best FDTD-evaluated
    candidate if it beats the parent.  ``mode``:

    * ``"surrogate"``: the FNO ranks all candidates and the batched FDTD
      evaluates the top ``top_k`` (full-grid fields).  If the surrogate's FoM
      relative error over those exceeds ``max_rel_err``, every candidate of
      the step is evaluated by the FDTD (fallback).  With ``anchor`` (default)
      the predicted FoM-line field of a candidate is the parent's FDTD
      field plus the network's predicted change, FNO(candidate) -
      FNO(parent): an additive low/high-fidelity correction in which the
      network's error common to the parent and its mutations cancels.
      With ``finetune_steps`` > 0, every FDTD solution of the loop also goes
      into a replay buffer the surrogate is fine-tuned on after each step
      (it learns the neighbourhood the search is in);
    * ``"random"``: ``top_k`` candidates picked at random are evaluated (the
      same FDTD budget without the network);
    * ``"exhaustive"``: every candidate is evaluated by the batched FDTD;
    * ``"plain"``: every candidate is evaluated by the single-design solver, one
      at a time (the plain loop).

    All modes draw the same candidate stream for the same parent and seed;
    ``"plain"`` and ``"exhaustive"`` make the same decisions.
    """

    def __init__(self, cfg, surrogate=None, mode="surrogate", n_candidates=256, top_k=8,
                 max_rel_err=0.05, patch=24, strength=0.5, seed=0, device=None,
                 finetune_steps=0, finetune_batch=32, finetune_lr=2e-4, buffer=1024,
                 anchor=True):
        self.cfg, self.sur, self.mode, self.anchor = cfg, surrogate, mode, bool(anchor)
        if mode == "surrogate" and surrogate is None:
            raise ValueError("mode 'surrogate' needs a trained FieldSurrogate")
        self.nc, self.top_k, self.max_rel_err = int(n_candidates), int(top_k), float(max_rel_err)
        self.patch, self.strength = int(patch), float(strength)
        self.ft_steps, self.ft_batch, self.ft_lr = int(finetune_steps), int(finetune_batch), \
            float(finetune_lr)
        y0, y1, x0, x1 = cfg.line
        self.fdtd = BatchedTM2D(**cfg.solver, source=cfg.source,
                                 monitor={"box": (y0, y1, x0, x1),
                                          "wavelengths": [cfg.wavelengths[cfg.fom_index]]},
                                 device=device)
        self.dev = self.fdtd.dev
        self.full = None
        if mode == "surrogate":
            self.full = BatchedTM2D(**cfg.solver, source=cfg.source,
                                    monitor={"box": (0, cfg.H, 0, cfg.W),
                                             "wavelengths": cfg.wavelengths},
                                    device=self.dev)
            self.buf_eps = torch.empty(buffer, cfg.H, cfg.W, device=self.dev)
            self.buf_f = torch.empty(buffer, len(cfg.wavelengths), cfg.H, cfg.W,
                                     dtype=torch.complex64, device=self.dev)
            self.buf_n, self.buf_i = 0, 0
            self.opt = None
        self.rng = torch.Generator(device="cpu").manual_seed(seed)       # candidates
        self.aux_rng = torch.Generator(device="cpu").manual_seed(seed + 1)  # picks, batches
        self.parent = workloads.blob_noise(seed, cfg.size).to(self.dev).unsqueeze(0)
        fom0, line0 = self._evaluate(cfg.eps(self.parent))
        self.parent_fom, self.parent_line = float(fom0[0]), line0[:1]
        self.history = [self.parent_fom]
        self.stats = {"steps": 0, "fallbacks": 0, "fdtd_solves": 1, "rel_err": [],
                      "time_s": {"candidates": 0.0, "surrogate": 0.0, "fdtd": 0.0,
                                 "finetune": 0.0}}

    # -- pieces -------------------------------------------------------------
    def _line(self, fields):
        """FoM-line DFT field [B, 1, n] of full-grid fields [B, nf, H, W]."""
        y0, y1, x0, x1 = self.cfg.line
        return fields[:, self.cfg.fom_index, y0:y1, x0:x1]

    def _predict_fom(self, eps):
        """Surrogate FoM of each candidate (anchored on the parent's FDTD
        field when ``anchor`` is set)."""
        line = self._line(self.sur.predict(eps))
        if self.anchor:
            base = self._line(self.sur.predict(self.cfg.eps(self.parent)))
            line = self.parent_line + (line - base)
        return (line.abs() ** 2).flatten(1).sum(1).double()

    def _evaluate(self, eps):
        """(FoM, line field [B, 1, n]) of each design from the batched FDTD."""
        f = self.fdtd.forward(eps)[:, 0]
        return (f.abs() ** 2).flatten(1).sum(1).double(), f

    def _evaluate_full(self, eps):
        """(FoM, full-grid fields) of each design; the fields feed the replay
        buffer."""
        f = self.full.forward(eps)
        self._remember(eps, f)
        return self.cfg.fom(f).double(), f

    def _remember(self, eps, fields):
        cap = self.buf_eps.shape[0]
        n = min(eps.shape[0], cap)
        slots = (self.buf_i + torch.arange(n, device=self.dev)) % cap
        self.buf_eps[slots] = eps[-n:].to(self.buf_eps.dtype)
        self.buf_f[slots] = fields[-n:]
        self.buf_i = (self.buf_i + n) % cap
        self.buf_n = min(cap, self.buf_n + n)

    def _finetune(self):
        if self.ft_steps <= 0 or self.buf_n == 0:
            return
        if self.opt is None:
            self.opt = torch.optim.Adam([self.sur.fno.params[k] for k in self.sur.fno.names],
                                        lr=self.ft_lr)
        for _ in range(self.ft_steps):
            idx = torch.randint(0, self.buf_n, (min(self.ft_batch, self.buf_n),),
                                generator=self.aux_rng).to(self.dev)
            y = self.sur.fno(self.sur.encode(self.buf_eps[idx]))
            loss = self.sur.loss(y, self.buf_f[idx])
            self.opt.zero_grad(set_to_none=True)
            loss.backward()
            self.opt.step()

    def _candidates(self):
        S, P, n = self.cfg.size, self.patch, self.nc
        pos = torch.randint(0, S - P + 1, (n, 2), generator=self.rng)
        noise = torch.randn(n, P, P, generator=self.rng, dtype=torch.float64)
        ar = torch.arange(P)
        rows = (pos[:, 0, None, None] + ar[None, :, None]).expand(n, P, P)
        cols = (pos[:, 1, None, None] + ar[None, None, :]).expand(n, P, P)
        d = torch.zeros(n, S, S, dtype=torch.float64)
        d[torch.arange(n)[:, None, None].expand(n, P, P), rows, cols] = noise
        return self.parent + self.strength * d.to(self.dev)

    def _tick(self, key, t0):
        torch.cuda.synchronize(self.dev)
        t = time.perf_counter()
        self.stats["time_s"][key] += t - t0
        return t

    # -- one design step -------------------------------------------------------
    def step(self):
        st = self.stats
        t0 = time.perf_counter()
        cand = self._candidates()
        eps = self.cfg.eps(cand)
        t0 = self._tick("candidates", t0)
        line = None
        if self.mode == "plain":
            fom = plain_fom(self.cfg, eps, self.dev).to(self.dev)
            idx = torch.arange(self.nc, device=self.dev)
            st["fdtd_solves"] += self.nc
            self._tick("fdtd", t0)
        elif self.mode == "exhaustive":
            fom, line = self._evaluate(eps)
            idx = torch.arange(self.nc, device=self.dev)
            st["fdtd_solves"] += self.nc
            self._tick("fdtd", t0)
        elif self.mode == "random":
            idx = torch.randperm(self.nc, generator=self.aux_rng)[:self.top_k].to(self.dev)
            fom, line = self._evaluate(eps[idx])
            st["fdtd_solves"] += self.top_k
            self._tick("fdtd", t0)
        else:
            pfom = self._predict_fom(eps)
            idx = torch.argsort(pfom, descending=True)[:self.top_k]
            t0 = self._tick("surrogate", t0)
            fom, full = self._evaluate_full(eps[idx])
            line = self._line(full)
            st["fdtd_solves"] += self.top_k
            err = ((pfom[idx] - fom).abs() / fom.clamp_min(1e-300)).mean().item()
            st["rel_err"].append(err)
            if err > self.max_rel_err:
                st["fallbacks"] += 1
                fom, line = self._evaluate(eps)
                idx = torch.arange(self.nc, device=self.dev)
                st["fdtd_solves"] += self.nc
            t0 = self._tick("fdtd", t0)
            self._finetune()
            self._tick("finetune", t0)
        best = int(torch.argmax(fom))
        if float(fom[best]) > self.parent_fom:
            self.parent = cand[int(idx[best])].unsqueeze(0)
            self.parent_fom = float(fom[best])
            if line is not None:
                self.parent_line = line[best:best + 1].clone()
        st["steps"] += 1
        self.history.append(self.parent_fom)
        return self.parent_fom

    def run(self, n_steps):
        for _ in range(n_steps):
            self.step()
        return self.parent_fom

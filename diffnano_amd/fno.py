"""Fourier neural operator (2D) on hand-written BF16 MFMA kernels.

Surrogate for the 2D TM FDTD: maps a permittivity map to complex DFT
fields.  Every dense operation runs through ``dnx_gemm_bf16`` (BF16
operands, FP32 accumulation, fused epilogues), the per-mode complex channel
mixing through ``dnx_fno_mode_mix`` / ``dnx_fno_mode_wgrad`` (FP32) and the
GELU backward through ``dnx_gelu_bwd``; forward and backward are written out by hand (no torch
autograd inside), wrapped in one ``torch.autograd.Function`` so a torch
optimizer can train the FP32 master weights.

Architecture (Li et al., "Fourier Neural Operator", 2021): lift (pointwise
linear, in_ch -> width), L Fourier layers h <- act(W h + b + K h), projection
(width -> hidden, GELU, hidden -> out_ch).  K h keeps modes ky in
{0..m1-1} U {-m1..-1} and kx in {0..m2-1} of the 2D DFT, mixes channels per
mode with complex weights, and returns to space (real part, Hermitian
weights over kx, 1/(H W) normalisation, as irfft2 of the truncated
spectrum).  Here the truncated 2D DFT and its inverse are dense matrices
(H W x 2 m1 m2 complex, stored BF16), so each spectral convolution is two
large-K GEMMs on MFMA plus the per-mode complex channel mixing.  The activation after
the last Fourier layer is omitted, as in the original FNO.

Layout: activations are [C][B][H*W] FP32 (channel-major over the batch), so
pointwise layers are single GEMMs over N = B*H*W and the DFT GEMMs see rows
(c, b).
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib as L


GemmArgs = L.GemmArgs


def _off(t, elems):
    return t.data_ptr() + int(elems) * t.element_size()


def gemm(M, N, K, A, sA, B, sB, Cout, sC, *, batch=1, a_off=0, b_off=0, c_off=0, alpha=1.0,
         beta=0.0, Cin=None, sI=None, i_off=0, bias=None, bias_mode=0, act=0, pre=None,
         c16=None):
    """C = act(alpha A B + beta Cin + bias); sA = (sAm, sAk, sAb) etc. (elements).
    ``c16`` (BF16, C's layout) also receives C rounded to BF16 (channel-mixing
    products, M and K <= 128)."""
    g = GemmArgs()
    g.M, g.N, g.K, g.batch = int(M), int(N), int(K), int(batch)
    g.A, (g.sAm, g.sAk, g.sAb) = _off(A, a_off), sA
    g.B, (g.sBk, g.sBn, g.sBb) = _off(B, b_off), sB
    # Cout None: only the BF16 copy c16 is written (pointwise shapes)
    g.C, (g.sCm, g.sCn, g.sCb) = (_off(Cout, c_off) if Cout is not None else None), sC
    if Cin is Cout and i_off == 0:
        i_off = c_off  # accumulating in place: read C where it is written
    if Cin is not None:
        g.Cin = _off(Cin, i_off)
        g.sIm, g.sIn, g.sIb = sI if sI is not None else sC
    g.alpha, g.beta = float(alpha), float(beta)
    g.bias = bias.data_ptr() if bias is not None else None
    g.bias_mode, g.act = int(bias_mode), int(act)
    g.pre = _off(pre, c_off) if pre is not None else None
    if c16 is not None:
        if c16.dtype != torch.bfloat16:
            raise TypeError("c16 must be bfloat16")
        g.c_bf16 = _off(c16, c_off)
    g.a_bf16 = 1 if A.dtype == torch.bfloat16 else 0
    g.b_bf16 = 1 if B.dtype == torch.bfloat16 else 0
    for t in (A, B):
        if t.dtype not in (torch.float32, torch.bfloat16):
            raise TypeError("gemm operands must be float32 or bfloat16")
    dev = (Cout if Cout is not None else c16).device
    L.check(L.lib().dnx_gemm_bf16(C.byref(g), L.stream_handle(dev)))
    return Cout


def mode_mix(X, Wr, Wi, Y, B, M, ld, c, conj=False):
    """Y[q][b][r*M+mu] = sum_p X[p][b][.] (Wr + i s Wi)[p][q][mu] (FP32
    arithmetic, s = -1 if conj); Wr, Wi are [p][q][M] contiguous; Y float32
    or bfloat16 (rounded to nearest even)."""
    ybf = Y.dtype == torch.bfloat16
    L.check(L.lib().dnx_fno_mode_mix(X.data_ptr(), Wr.data_ptr(), Wi.data_ptr(), Y.data_ptr(),
                                     1 if ybf else 0, int(B), int(M), int(ld), int(c),
                                     1 if conj else 0, L.stream_handle(Y.device)))
    return Y


def mode_wgrad(X, dY, dWr, dWi, B, M, ld, c, nsplit):
    """dWr[s] + i dWi[s] = sum over sample split s of conj(X[p][b]) dY[q][b], per
    mode, as [nsplit][p][q][M]."""
    L.check(L.lib().dnx_fno_mode_wgrad(X.data_ptr(), dY.data_ptr(), dWr.data_ptr(),
                                       dWi.data_ptr(), int(B), int(M), int(ld), int(c),
                                       int(nsplit), L.stream_handle(dY.device)))


def gelu_bwd(dy, pre, dx, dx16=None):
    L.check(L.lib().dnx_gelu_bwd(dy.data_ptr(), pre.data_ptr(), dx.data_ptr(),
                                 dx16.data_ptr() if dx16 is not None else None, dy.numel(),
                                 L.stream_handle(dy.device)))
    return dx


def dft_matrices(H, W, m1, m2, device):
    """Forward F (H W x 2 M) and inverse G (2 M x H W), M = 2 m1 m2 modes,
    column/row n = ri * M + mu, mu = kyi * m2 + kx, ky = kyi (kyi < m1) or
    kyi - 2 m1 (negative half).  Built in float64, stored BF16, each also
    transposed (FT, GT) so that every large product reads both operands
    contiguously along its reduction axis."""
    kyi = np.arange(2 * m1)
    ky = np.where(kyi < m1, kyi, kyi - 2 * m1).astype(np.float64)
    kx = np.arange(m2, dtype=np.float64)
    y = np.arange(H, dtype=np.float64)
    x = np.arange(W, dtype=np.float64)
    theta = 2 * np.pi * (ky[None, None, :, None] * y[:, None, None, None] / H
                         + kx[None, None, None, :] * x[None, :, None, None] / W)
    theta = theta.reshape(H * W, 2 * m1 * m2)
    F = np.concatenate([np.cos(theta), -np.sin(theta)], axis=1)
    ckx = np.where(kx == 0, 1.0, 2.0)
    cw = np.tile(ckx, 2 * m1)[None, :] / (H * W)
    G = np.concatenate([np.cos(theta) * cw, -np.sin(theta) * cw], axis=1).T

    def bf(a):
        return torch.tensor(np.ascontiguousarray(a), device=device).to(torch.bfloat16).contiguous()

    return bf(F), bf(G), bf(F.T), bf(G.T)


def sum_partials(part, s, n, out, beta=0.0):
    """out[i] = sum_j part[j n + i] + beta out[i] (split-K reduction)."""
    L.check(L.lib().dnx_sum_partials(part.data_ptr(), int(s), int(n), out.data_ptr(),
                                     float(beta), L.stream_handle(out.device)))
    return out


def gemm_nt(M, N, K, A, B, out, *, target_blocks=None, beta=0.0, Cin=None, A16=None):
    """out[M x N] = A[M x K] B[N x K]^T (+ beta Cin), both operands row-major
    with K contiguous, split along K into a batch of partial products (then
    summed) when the output alone has too few tiles to fill the GPU (256 x
    256 tiles, one work-group per CU, for M, N >= 256 and long K; 128 x 128
    tiles otherwise).  ``A16``: A already rounded to BF16 (what the product
    does to an FP32 A on load), read in its place -- same result, half the
    bytes, and the direct-to-LDS kernel for BF16 operands."""
    if A16 is not None:
        A = A16
    big = M >= 256 and N >= 256 and K >= 4096  # split chunks keep K >= 2048
    tb = 256 if big else 128
    if target_blocks is None:
        target_blocks = 512 if big else 1024
    tiles = -(-M // tb) * -(-N // tb)
    s = 1
    while (s < 8 and tiles * s * 2 <= target_blocks and K % (s * 2 * 64) == 0
           and (not big or K // (s * 2) >= 2048)):
        s *= 2
    if s == 1:
        return gemm(M, N, K, A, (K, 1, 0), B, (1, K, 0), out, (N, 1, 0), beta=beta, Cin=Cin)
    kc = K // s
    part = torch.empty(s, M * N, device=out.device)
    gemm(M, N, kc, A, (K, 1, kc), B, (1, K, kc), part, (N, 1, M * N), batch=s)
    if beta != 0.0 and Cin is not out:
        out.copy_(Cin)
    return sum_partials(part, s, M * N, out, beta=beta)


class FNO2d:
    """Parameters and the hand-written forward/backward of a 2D FNO."""

    def __init__(self, H, W, in_ch=3, out_ch=6, width=32, modes=(16, 16), layers=4, hidden=128,
                 device="cuda", seed=0):
        self.H, self.W = int(H), int(W)
        self.in_ch, self.out_ch, self.width = in_ch, out_ch, width
        self.m1, self.m2 = modes
        self.layers, self.hidden = layers, hidden
        self.dev = torch.device(device)
        self.nmodes = 2 * self.m1 * self.m2
        self.F, self.G, self.FT, self.GT = dft_matrices(self.H, self.W, self.m1, self.m2,
                                                         self.dev)
        g = torch.Generator().manual_seed(seed)

        def u(shape, fan_in):
            bound = 1.0 / math.sqrt(fan_in)
            return ((torch.rand(shape, generator=g, dtype=torch.float64) * 2 - 1) * bound).to(
                device=self.dev, dtype=torch.float32)

        c = width
        scale = 1.0 / (c * c)
        p = {"w_lift": u((c, in_ch), in_ch), "b_lift": u((c,), in_ch)}
        for l in range(layers):
            p[f"w{l}"] = u((c, c), c)
            p[f"b{l}"] = u((c,), c)
            p[f"sr{l}"] = (scale * torch.rand((self.nmodes, c, c), generator=g,
                                              dtype=torch.float64)).to(self.dev, torch.float32)
            p[f"si{l}"] = (scale * torch.rand((self.nmodes, c, c), generator=g,
                                              dtype=torch.float64)).to(self.dev, torch.float32)
        p["w_p1"] = u((hidden, c), c)
        p["b_p1"] = u((hidden,), c)
        p["w_p2"] = u((out_ch, hidden), hidden)
        p["b_p2"] = u((out_ch,), hidden)
        self.params = {k: v.contiguous().requires_grad_(True) for k, v in p.items()}
        self.names = list(self.params)

    # -- building blocks ------------------------------------------------------
    def _pointwise(self, Wt, x, out, bias=None, act=0, pre=None, Cin=None, beta=0.0, c16=None):
        """out[o][n] = act(sum_i Wt[o][i] x[i][n] + bias[o] + beta Cin[o][n])
        (and out rounded to BF16 into ``c16``)."""
        O, I = Wt.shape
        n = x.shape[1]
        return gemm(O, n, I, Wt, (I, 1, 0), x, (n, 1, 0), out, (n, 1, 0), bias=bias,
                    bias_mode=2 if bias is not None else 0, act=act, pre=pre, Cin=Cin, beta=beta,
                    c16=c16)

    fast_mix = True  # spectral mixing on the dedicated kernels where shapes allow
    # the forward-DFT products read BF16 copies of their FP32 operand, written
    # by the kernels that produce it (bitwise the same products)
    bf16_copies = True

    # work-groups the forward-DFT products aim at when they split K: one
    # round of the 256 CUs (at 128 x 128, 256 samples: split 2 instead of 4,
    # 5.09 -> 4.72 ms per inference, 10.75 -> 10.15 ms per training step)
    dft_blocks = 256

    # each layer's inverse DFT and pointwise part run as one kernel
    # (dnx_fno_inverse_pw) where the shapes allow
    fused_inverse = True

    def _fused_inverse_ok(self, B):
        P = self.H * self.W
        return (self.fused_inverse and self.bf16_copies and self._fast_mix_ok(B)
                and P % 256 == 0 and (2 * self.nmodes) % 64 == 0)

    def _inverse_pw(self, Y, h16, l, out, pre, out16, act, B, *, backward=False,
                    gelu_pre=None):
        """out = act(Y GT + W_l h + b_l) per layer (FP32 out, pre and the
        BF16 copy out16 each optional), the sum of the spectral branch kept
        in the accumulators.  backward: d h = dZ F^T + W_l^T d(pre), times
        GELU'(gelu_pre) if given (the previous layer's d(pre))."""
        p = self.params
        ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        L.check(L.lib().dnx_fno_inverse_pw(
            Y.data_ptr(), (self.F if backward else self.GT).data_ptr(), h16.data_ptr(),
            p[f"w{l}"].data_ptr(), 1 if backward else 0,
            None if backward else p[f"b{l}"].data_ptr(), ptr(out), ptr(pre), ptr(out16),
            int(act), ptr(gelu_pre), self.width, int(B), self.H * self.W, 2 * self.nmodes,
            L.stream_handle(self.dev)))

    # inference runs the projection head as one kernel (dnx_fno_proj) where
    # the shapes allow
    fused_head = True

    def _fused_head_ok(self):
        return (self.fused_head and self.width == 32 and self.hidden == 128
                and self.out_ch in (1, 2, 3, 4, 5, 6, 8, 16, 32))

    def _fast_mix_ok(self, B):
        return self.fast_mix and self.width == 32 and B % 8 == 0 and self.nmodes % 16 == 0

    def _mix(self, Z, sr, si, Y, B):
        """Y[(o,b)][ri*M+mu] = sum_i Z[(i,b)][.] (Wr + i Wi)[mu][i][o] per mode."""
        c, M = self.width, self.nmodes
        ld = 2 * M
        if self._fast_mix_ok(B):
            mode_mix(Z, sr.permute(1, 2, 0).contiguous(), si.permute(1, 2, 0).contiguous(), Y,
                     B, M, ld, c, conj=False)
            return
        sA = (ld, B * ld, 1)
        sB = (c, 1, c * c)
        sC = (ld, B * ld, 1)
        kw = dict(batch=M)
        gemm(B, c, c, Z, sA, sr, sB, Y, sC, a_off=0, c_off=0, **kw)                 # Yr = Zr Wr
        gemm(B, c, c, Z, sA, si, sB, Y, sC, a_off=M, c_off=0, alpha=-1.0, beta=1.0,  # - Zi Wi
             Cin=Y, **kw)
        gemm(B, c, c, Z, sA, si, sB, Y, sC, a_off=0, c_off=M, **kw)                 # Yi = Zr Wi
        gemm(B, c, c, Z, sA, sr, sB, Y, sC, a_off=M, c_off=M, beta=1.0, Cin=Y, **kw)  # + Zi Wr

    # -- forward / backward ----------------------------------------------------
    # inference keeps the hidden activations in BF16 only (every consumer
    # rounds them to BF16 anyway: bitwise the same output, fewer bytes)
    bf16_inference = True

    def _bf16_inference_ok(self, n):
        # the pointwise kernel's 32-bit row-block offsets: 33 rows of n cells
        # of FP32 below 2^31 bytes (larger batches take the FP32 path)
        return (self.bf16_inference and self.bf16_copies and self._fused_head_ok()
                and 33 * 4 * n < (1 << 31))

    def forward_raw(self, x, save=True):
        """x: [in_ch][B][H*W] FP32 -> ([out_ch][B][H*W] FP32, saved tensors)."""
        if not save and self._bf16_inference_ok(x.shape[1] * self.H * self.W):
            return self._infer_bf16(x), None
        p = self.params
        c, B = self.width, x.shape[1]
        P = self.H * self.W
        n = B * P
        M2 = 2 * self.nmodes
        dev = self.dev
        x2 = x.reshape(self.in_ch, n)
        h = torch.empty(c, n, device=dev)
        # h rounded to BF16 for the forward-DFT product, written by the layer
        # that produces h (one buffer: each layer's product reads it before
        # the layer's pointwise part overwrites it)
        h16 = torch.empty(c, n, device=dev, dtype=torch.bfloat16) if self.bf16_copies else None
        self._pointwise(p["w_lift"], x2, h, bias=p["b_lift"], c16=h16)
        saved = {"x": x2, "h": [], "Z": [], "pre": []}
        # the fused layer kernel reads the layer's input copy while it writes
        # the output's: a second BF16 buffer
        fuse = self._fused_inverse_ok(B)
        h16b = torch.empty_like(h16) if fuse else None
        for l in range(self.layers):
            Z = torch.empty(c * B, M2, device=dev)
            gemm_nt(c * B, M2, P, h, self.FT, Z, A16=h16, target_blocks=self.dft_blocks)
            # the inverse-DFT product rounds Y to BF16 on load: the mixing
            # kernel writes it rounded (same values, half the bytes)
            Y = torch.empty(c * B, M2, device=dev,
                            dtype=torch.bfloat16 if self._fast_mix_ok(B) else torch.float32)
            self._mix(Z, p[f"sr{l}"], p[f"si{l}"], Y, B)
            last = l == self.layers - 1
            hn = torch.empty(c, n, device=dev)
            pre = None if (last or not save) else torch.empty(c, n, device=dev)
            if fuse:
                self._inverse_pw(Y, h16, l, hn, pre, None if last else h16b, 0 if last else 1, B)
                if not last:
                    h16, h16b = h16b, h16
            else:
                S = torch.empty(c, n, device=dev)
                gemm(c * B, P, M2, Y, (M2, 1, 0), self.GT, (1, M2, 0), S, (P, 1, 0))
                self._pointwise(p[f"w{l}"], h, hn, bias=p[f"b{l}"], act=0 if last else 1,
                                pre=pre, Cin=S, beta=1.0, c16=None if last else h16)
            if save:
                saved["h"].append(h)
                saved["Z"].append(Z)
                saved["pre"].append(pre)
            h = hn
        out = torch.empty(self.out_ch, n, device=dev)
        if not save and self._fused_head_ok():
            # inference: the projection 32 -> 128 -> GELU -> out_ch in one
            # pass, the hidden activations never written
            L.check(L.lib().dnx_fno_proj(
                h.data_ptr(), 0, p["w_p1"].data_ptr(), p["b_p1"].data_ptr(),
                p["w_p2"].data_ptr(), p["b_p2"].data_ptr(), out.data_ptr(), c, self.hidden,
                self.out_ch, n, L.stream_handle(dev)))
            return out.view(self.out_ch, B, P), None
        q = torch.empty(self.hidden, n, device=dev)
        qpre = torch.empty(self.hidden, n, device=dev) if save else None
        self._pointwise(p["w_p1"], h, q, bias=p["b_p1"], act=1, pre=qpre)
        self._pointwise(p["w_p2"], q, out, bias=p["b_p2"])
        saved.update({"hL": h, "q": q, "qpre": qpre, "B": B})
        return out.view(self.out_ch, B, P), (saved if save else None)

    def _infer_bf16(self, x):
        """The forward without saving, with the layers' activations written
        only as their BF16 copies: the forward DFT, the next pointwise layer
        and the projection head all read h rounded to BF16, so the FP32 h is
        never needed (the lifting layer still writes it where its shape is
        not a pointwise-kernel shape)."""
        p = self.params
        c, B = self.width, x.shape[1]
        P = self.H * self.W
        n = B * P
        M2 = 2 * self.nmodes
        dev = self.dev
        x2 = x.reshape(self.in_ch, n)
        fuse = self._fused_inverse_ok(B)
        h16 = torch.empty(c, n, device=dev, dtype=torch.bfloat16)
        nxt = torch.empty(c, n, device=dev, dtype=torch.bfloat16)
        # the lifting layer on the pointwise kernel writes only the copy
        h = None if (c, self.in_ch) in ((32, 9), (32, 32)) else torch.empty(c, n, device=dev)
        self._pointwise(p["w_lift"], x2, h, bias=p["b_lift"], c16=h16)
        del h
        for l in range(self.layers):
            Z = torch.empty(c * B, M2, device=dev)
            gemm_nt(c * B, M2, P, h16, self.FT, Z, target_blocks=self.dft_blocks)
            Y = torch.empty(c * B, M2, device=dev,
                            dtype=torch.bfloat16 if self._fast_mix_ok(B) else torch.float32)
            self._mix(Z, p[f"sr{l}"], p[f"si{l}"], Y, B)
            last = l == self.layers - 1
            if fuse:
                self._inverse_pw(Y, h16, l, None, None, nxt, 0 if last else 1, B)
            else:
                S = torch.empty(c, n, device=dev)
                gemm(c * B, P, M2, Y, (M2, 1, 0), self.GT, (1, M2, 0), S, (P, 1, 0))
                gemm(c, n, c, p[f"w{l}"], (c, 1, 0), h16, (n, 1, 0), None, (n, 1, 0),
                     bias=p[f"b{l}"], bias_mode=2, act=0 if last else 1, Cin=S, beta=1.0,
                     c16=nxt)
            h16, nxt = nxt, h16
        out = torch.empty(self.out_ch, n, device=dev)
        L.check(L.lib().dnx_fno_proj(
            h16.data_ptr(), 1, p["w_p1"].data_ptr(), p["b_p1"].data_ptr(),
            p["w_p2"].data_ptr(), p["b_p2"].data_ptr(), out.data_ptr(), c, self.hidden,
            self.out_ch, n, L.stream_handle(dev)))
        return out.view(self.out_ch, B, P)

    SPLIT = 256  # split-K factor for reductions over the B*H*W axis

    def _split(self, n):
        s = self.SPLIT
        while n % s:
            s //= 2
        return s, n // s

    def _wgrad(self, dY, X, out):
        """out[o][i] = sum_n dY[o][n] X[i][n] over n = B*H*W: split-K partial
        GEMMs across the batch grid axis, then a GEMM with a ones vector sums
        the partials."""
        O, n = dY.shape
        I = X.shape[0]
        s, chunk = self._split(n)
        part = torch.empty(s, O * I, device=self.dev)
        gemm(O, I, chunk, dY, (n, 1, chunk), X, (1, n, chunk), part, (I, 1, O * I), batch=s)
        sum_partials(part, s, O * I, out)

    def _wgrad_bias(self, dY, X, gw, gb):
        """gw[o][i] = sum_n dY[o][n] X[i][n], gb[o] = sum_n dY[o][n] (FP32
        partial sums over column splits, then summed)."""
        O, n = dY.shape
        I = X.shape[0]
        if n % 64 or dY.data_ptr() % 16 or X.data_ptr() % 16:
            self._wgrad(dY, X, gw)
            self._rowsum(dY, gb)
            return
        # output tiles per column split: one for at most 128 outputs over at
        # most 32 inputs (the library's single-pass tile)
        one = O <= 128 and I <= 32
        tiles = 1 if one else -(-O // 32) * -(-I // 32)
        q = n // 64
        target = max(1, min(q, 1024 // tiles))  # ~4 work-groups per CU
        ns = max(d for d in range(1, target + 1) if q % d == 0)
        part = torch.empty(ns, O * I, device=self.dev)
        pb = torch.empty(ns, O, device=self.dev)
        L.check(L.lib().dnx_fno_wgrad(dY.data_ptr(), X.data_ptr(), part.data_ptr(),
                                      pb.data_ptr(), int(O), int(I), int(n), int(ns),
                                      L.stream_handle(self.dev)))
        sum_partials(part, ns, O * I, gw)
        sum_partials(pb, ns, O, gb)

    def _rowsum(self, dY, out):
        O, n = dY.shape
        s, chunk = self._split(n)
        ones = self._ones(n)
        part = torch.empty(s, O, device=self.dev)
        gemm(O, 1, chunk, dY, (n, 1, chunk), ones, (1, 0, chunk), part, (1, 1, O), batch=s)
        sum_partials(part, s, O, out)

    def _ones(self, n):
        o = getattr(self, "_ones_buf", None)
        if o is None or o.numel() < n:
            o = torch.ones(n, device=self.dev)
            self._ones_buf = o
        return o

    def backward_raw(self, dout, saved):
        """Gradients of all parameters given d(out) [out_ch][B][H*W]."""
        p = self.params
        c, B = self.width, saved["B"]
        P = self.H * self.W
        n = B * P
        M, M2 = self.nmodes, 2 * self.nmodes
        dev = self.dev
        g = {k: torch.zeros_like(v) for k, v in p.items()}
        d2 = dout.reshape(self.out_ch, n).contiguous()
        # projection
        self._wgrad_bias(d2, saved["q"], g["w_p2"], g["b_p2"])
        dq = torch.empty(self.hidden, n, device=dev)
        gemm(self.hidden, n, self.out_ch, p["w_p2"], (1, self.hidden, 0), d2, (n, 1, 0), dq,
             (n, 1, 0))
        gelu_bwd(dq, saved["qpre"], dq)
        self._wgrad_bias(dq, saved["hL"], g["w_p1"], g["b_p1"])
        dh = torch.empty(c, n, device=dev)
        # d(pre) rounded to BF16 for the spectral branch's forward-DFT product
        d16 = torch.empty(c, n, device=dev, dtype=torch.bfloat16) if self.bf16_copies else None
        gemm(c, n, self.hidden, p["w_p1"], (1, c, 0), dq, (n, 1, 0), dh, (n, 1, 0), c16=d16)
        del dq
        # the fused layer kernel computes each layer's input gradient and, with
        # the previous layer's pre-activation, that layer's d(pre) and its copy
        fuse = self._fused_inverse_ok(B)
        d16b = torch.empty_like(d16) if fuse else None
        dpre_next = None
        for l in range(self.layers - 1, -1, -1):
            h, Z, pre = saved["h"][l], saved["Z"][l], saved["pre"][l]
            if dpre_next is not None:
                dpre = dpre_next
            else:
                dpre = dh if pre is None else gelu_bwd(dh, pre, torch.empty_like(dh), d16)
            self._wgrad_bias(dpre, h, g[f"w{l}"], g[f"b{l}"])
            if not fuse:
                dhn = torch.empty(c, n, device=dev)
                gemm(c, n, c, p[f"w{l}"], (1, c, 0), dpre, (n, 1, 0), dhn, (n, 1, 0))
            # spectral branch: dY = dS G^T
            dY = torch.empty(c * B, M2, device=dev)
            gemm_nt(c * B, M2, P, dpre, self.G, dY, A16=d16, target_blocks=self.dft_blocks)
            sr, si = p[f"sr{l}"], p[f"si{l}"]
            ld = M2
            sZ = (ld, B * ld, 1)       # Z/dZ/dY as [b][i] per mode
            sWt = (1, c, c * c)        # W^T: B[k=o][n=i] = W[mu][i][o]
            sWc = (c, 1, c * c)        # dW[mu][i][o]
            if self._fast_mix_ok(B):
                dZ = torch.empty(c * B, M2, device=dev, dtype=torch.bfloat16)
                # dZ = dY W^H per mode; dW = sum_b conj(Z) dY
                mode_mix(dY, sr.permute(2, 1, 0).contiguous(), si.permute(2, 1, 0).contiguous(),
                         dZ, B, M, ld, c, conj=True)
                chunks = B // 8
                ns = max(d for d in range(1, min(chunks, 16) + 1) if chunks % d == 0)
                pr = torch.empty(ns, c, c, M, device=dev)
                pi = torch.empty(ns, c, c, M, device=dev)
                mode_wgrad(Z, dY, pr, pi, B, M, ld, c, ns)
                g[f"sr{l}"] = pr.sum(0).permute(2, 0, 1).contiguous()
                g[f"si{l}"] = pi.sum(0).permute(2, 0, 1).contiguous()
                if fuse:
                    prev = saved["pre"][l - 1] if l > 0 else None
                    if prev is not None:  # the previous layer's d(pre) and its copy
                        dpre_next = torch.empty(c, n, device=dev)
                        self._inverse_pw(dZ, d16, l, dpre_next, None, d16b, 0, B,
                                         backward=True, gelu_pre=prev)
                        d16, d16b = d16b, d16
                    else:  # d h of the lifting layer's output
                        dpre_next = None
                        dh = torch.empty(c, n, device=dev)
                        self._inverse_pw(dZ, d16, l, dh, None, None, 0, B, backward=True)
                    continue
                gemm(c * B, P, M2, dZ, (M2, 1, 0), self.F, (1, M2, 0), dhn, (P, 1, 0), beta=1.0,
                     Cin=dhn)
                dh = dhn
                continue
            dZ = torch.empty(c * B, M2, device=dev)
            kw = dict(batch=M)
            # dZr = dYr Wr^T + dYi Wi^T ; dZi = -dYr Wi^T + dYi Wr^T
            gemm(B, c, c, dY, sZ, sr, sWt, dZ, sZ, a_off=0, c_off=0, **kw)
            gemm(B, c, c, dY, sZ, si, sWt, dZ, sZ, a_off=M, c_off=0, beta=1.0, Cin=dZ, **kw)
            gemm(B, c, c, dY, sZ, si, sWt, dZ, sZ, a_off=0, c_off=M, alpha=-1.0, **kw)
            gemm(B, c, c, dY, sZ, sr, sWt, dZ, sZ, a_off=M, c_off=M, beta=1.0, Cin=dZ, **kw)
            # dWr = Zr^T dYr + Zi^T dYi ; dWi = Zr^T dYi - Zi^T dYr   (K = batch)
            sZt = (B * ld, ld, 1)      # A[m=i][k=b]
            sdY = (ld, B * ld, 1)      # B[k=b][n=o]
            gemm(c, c, B, Z, sZt, dY, sdY, g[f"sr{l}"], sWc, a_off=0, b_off=0, **kw)
            gemm(c, c, B, Z, sZt, dY, sdY, g[f"sr{l}"], sWc, a_off=M, b_off=M, beta=1.0,
                 Cin=g[f"sr{l}"], **kw)
            gemm(c, c, B, Z, sZt, dY, sdY, g[f"si{l}"], sWc, a_off=0, b_off=M, **kw)
            gemm(c, c, B, Z, sZt, dY, sdY, g[f"si{l}"], sWc, a_off=M, b_off=0, alpha=-1.0,
                 beta=1.0, Cin=g[f"si{l}"], **kw)
            # dh = W^T dpre + dZ F^T
            gemm(c * B, P, M2, dZ, (M2, 1, 0), self.F, (1, M2, 0), dhn, (P, 1, 0), beta=1.0,
                 Cin=dhn)
            dh = dhn
        self._wgrad_bias(dh, saved["x"], g["w_lift"], g["b_lift"])
        return g

    def __call__(self, x):
        """x: [B, in_ch, H, W] -> [B, out_ch, H, W] (differentiable w.r.t. the
        parameters; the input is not differentiated)."""
        B = x.shape[0]
        xt = x.to(self.dev, torch.float32).permute(1, 0, 2, 3).reshape(
            self.in_ch, B, self.H * self.W).contiguous()
        if torch.is_grad_enabled():
            out = _FNOFn.apply(self, xt, *[self.params[k] for k in self.names])
        else:  # inference: nothing saved for a backward
            out, _ = self.forward_raw(xt, save=False)
        return out.view(self.out_ch, B, self.H, self.W).permute(1, 0, 2, 3)


class _FNOFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, xt, *params):
        need = any(ctx.needs_input_grad[2:])
        out, saved = model.forward_raw(xt, save=need)
        ctx.model, ctx.saved_fno = model, saved
        return out

    @staticmethod
    def backward(ctx, dout):
        model = ctx.model
        g = model.backward_raw(dout.contiguous(), ctx.saved_fno)
        ctx.saved_fno = None
        return (None, None) + tuple(g[k] for k in model.names)

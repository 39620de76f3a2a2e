"""FP64 CPU oracle of the 2D FNO surrogate (TEST INFRASTRUCTURE).

Only tests import this module.  It restates the network of
``diffnano_amd/fno.py`` with ``torch.fft.rfft2`` / ``irfft2`` (an FFT, not
the dense truncated-DFT matrices the GPU path multiplies on MFMA) and torch
autograd in float64 on the CPU, so it checks the GPU forward and the
hand-written backward independently.

Spectral weights are stored per kept mode mu = kyi * m2 + kx with
ky = kyi for kyi < m1 and ky = kyi - 2 m1 (negative frequencies) otherwise,
as [modes][in][out] real and imaginary parts.
"""

from __future__ import annotations

import torch


def _lin(w, b, x):
    # x: [B, I, H, W] -> [B, O, H, W]
    return torch.einsum("oi,bihw->bohw", w, x) + b[None, :, None, None]


def spectral_conv(h, sr, si, m1, m2):
    """Spectral convolution of one Fourier layer: h [B, I, H, W] -> [B, O, H, W].
    sr, si: [modes][I][O] real and imaginary weights (mode order above)."""
    B, _, H, W = h.shape
    O = sr.shape[2]
    X = torch.fft.rfft2(h)
    out = torch.zeros(B, O, H, W // 2 + 1, dtype=torch.complex128)
    Wc = torch.complex(sr, si)
    for kyi in range(2 * m1):
        ky = kyi if kyi < m1 else H + kyi - 2 * m1
        for kx in range(m2):
            mu = kyi * m2 + kx
            out[:, :, ky, kx] = X[:, :, ky, kx] @ Wc[mu]
    return torch.fft.irfft2(out, s=(H, W))


def fno_forward(p, x, layers, m1, m2):
    """p: dict of float64 CPU tensors (names as FNO2d.params); x [B, in, H, W]."""
    gelu = torch.nn.functional.gelu
    h = _lin(p["w_lift"], p["b_lift"], x)
    for l in range(layers):
        S = spectral_conv(h, p[f"sr{l}"], p[f"si{l}"], m1, m2)
        h = _lin(p[f"w{l}"], p[f"b{l}"], h) + S
        if l < layers - 1:
            h = gelu(h)
    q = gelu(_lin(p["w_p1"], p["b_p1"], h))
    return _lin(p["w_p2"], p["b_p2"], q)

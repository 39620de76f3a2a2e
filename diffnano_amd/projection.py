"""Density projection, smoothing and binarisation schedule for topology
optimisation, drop-in for DiffNano's ``diffnano/design/projection.py``.

Host-side tensor operations on the design density (a few thousand cells),
differentiable with autograd; the solver they feed runs on the GPU.

* ``heaviside_projection`` (projection.py:15-36): sigmoid(beta (rho - eta)).
* ``smooth_filter`` (projection.py:39-85): normalised Gaussian blur with
  sigma = radius / 2 over an odd (2 radius + 1)-tap window, zero padding,
  same output shapes as the reference (see the function).
* ``beta_continuation_schedule`` (projection.py:88-115): geometric ramp from
  ``beta_start`` to ``beta_end`` over ``total_steps``.
"""

from __future__ import annotations

import torch
import torch.nn.functional as F

__all__ = ["heaviside_projection", "smooth_filter", "beta_continuation_schedule"]


def heaviside_projection(density: torch.Tensor, beta: float = 10.0,
                         eta: float = 0.5) -> torch.Tensor:
    """Smoothed step at threshold ``eta`` with sharpness ``beta``
    (projection.py:15-36)."""
    return torch.sigmoid(beta * (density - eta))


def _gauss_taps(radius: float, dtype, device) -> torch.Tensor:
    n = int(2 * radius + 1)
    n += 1 - n % 2                      # odd window
    sigma = radius / 2.0
    x = torch.arange(n, dtype=dtype, device=device) - (n // 2)
    w = torch.exp(-(x * x) / (2 * sigma * sigma + 1e-12))
    return w / w.sum()


def smooth_filter(density: torch.Tensor, radius: float = 2.0) -> torch.Tensor:
    """Gaussian blur of a density map (projection.py:39-85).

    Input ``(H, W)`` gives ``(H, W)``; input ``(N, H, W)`` gives ``(H, W)``
    when N = 1 and ``(N, 1, H, W)`` otherwise (the reference's shapes)."""
    if density.dim() == 2:
        x = density[None, None]
    elif density.dim() == 3:
        x = density[:, None]
    else:
        x = density
    w = _gauss_taps(radius, x.dtype, x.device)
    k = w[:, None] * w[None, :]
    out = F.conv2d(x, k[None, None], padding=(w.numel() - 1) // 2)
    if out.dim() == 4 and out.shape[0] == 1:
        out = out[0, 0] if out.shape[1] == 1 else out[0]
    return out


def beta_continuation_schedule(step: int, total_steps: int = 500, beta_start: float = 1.0,
                               beta_end: float = 64.0) -> float:
    """beta_start (beta_end / beta_start)^min(step / total_steps, 1)
    (projection.py:88-115)."""
    t = min(step / total_steps, 1.0)
    return beta_start * (beta_end / beta_start) ** t

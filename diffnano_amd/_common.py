"""Host-side set-up shared by the 2D and 3D solvers: CPML profiles, source
waveform tables, source boxes and DFT-monitor descriptors.

All of this is computed once per call on the host (a few KB) and uploaded;
the time stepping itself runs only in libdnx_fdtd.
"""

from __future__ import annotations

import math

import warnings

import numpy as np
import torch

COMPONENTS_3D = {"Ex": 0, "Ey": 1, "Ez": 2}


def cpml_profile(size: int, pml_layers: int, dl: float, dt: float, sigma_max: float = 0.8,
                 order: int = 3):
    """(b, c) float64 arrays of the memoryless CPML (fdtd3d.py:51-85).

    sigma = sigma_max ((L-i)/L)^order on the first L cells, mirrored;
    s = sigma dl/dt; b = exp(-(s + 0.02 s) dt); c = s (b-1)/(s + 0.02 s + 1e-30)
    where s + 0.02 s > 1e-30, else 0.  The reference's ``kappa`` is unused.
    """
    sigma = [0.0] * size
    for i in range(pml_layers):
        v = sigma_max * ((pml_layers - i) / pml_layers) ** order
        sigma[i] = v
        sigma[size - 1 - i] = v
    s = np.asarray(sigma, dtype=np.float64) * dl / dt
    a = 0.02 * s
    b = np.exp(-(s + a) * dt)
    den = s + a
    c = np.zeros_like(s)
    nz = den > 1e-30
    c[nz] = s[nz] * (b[nz] - 1.0) / (den[nz] + 1e-30)
    return b, c


def waveform_table(n_steps: int, dt: float, omega: float, source: dict) -> np.ndarray:
    """w(n), n < n_steps, as the reference's ``_source_waveform`` (fdtd3d.py:319-338).

    gaussian_pulse: the sine argument is rounded to float32 and the sine is
    evaluated by torch's float32 CPU kernel (as the reference does by
    building ``torch.tensor(omega * t)`` without a dtype); ``amp * sin`` is
    float32; the envelope and the product are float64.  continuous: float64
    ``amp * min(1, t omega/(4 pi)) * sin(omega t)``.  Other types: zero.
    """
    kind = source.get("type", "gaussian_pulse")
    if kind == "gaussian_pulse":
        t0 = source.get("t0", 3.0 / omega)
        spread = source.get("spread", 1.0 / omega)
        amp = source.get("amplitude", 1.0)
        t = torch.arange(n_steps, dtype=torch.float64) * dt
        s32 = torch.sin((omega * t).to(torch.float32))
        amp_sin = amp * s32
        env = torch.exp(-((t - t0) ** 2) / (2 * torch.tensor(spread, dtype=torch.float64) ** 2))
        return (amp_sin.to(torch.float64) * env).numpy()
    if kind == "continuous":
        amp = source.get("amplitude", 1.0)
        out = np.empty(n_steps, dtype=np.float64)
        for n in range(n_steps):
            t = n * dt
            ramp = min(1.0, t * omega / (4 * math.pi))
            out[n] = amp * ramp * math.sin(omega * t)
        return out
    return np.zeros(n_steps, dtype=np.float64)


def source_box_3d(shape, source: dict):
    """Box (z0, z1, y0, y1, x0, x1) of the reference's 0/1 mask (fdtd3d.py:340-372),
    or None when the position is out of range (the reference then injects nothing)."""
    D, H, W = shape
    pos = source.get("pos", None)
    if pos is not None:
        z, y, x = (int(v) for v in pos)
        if 0 <= z < D and 0 <= y < H and 0 <= x < W:
            return (z, z + 1, y, y + 1, x, x + 1)
        return None
    plane = source.get("plane", None)
    if plane == "xy":
        z = int(source.get("z", D // 2))
        return (z, z + 1, 0, H, 0, W) if 0 <= z < D else None
    if plane == "xz":
        y = int(source.get("y", H // 2))
        return (0, D, y, y + 1, 0, W) if 0 <= y < H else None
    if plane == "yz":
        x = int(source.get("x", W // 2))
        return (0, D, 0, H, x, x + 1) if 0 <= x < W else None
    return (D // 2, D // 2 + 1, H // 2, H // 2 + 1, 0, W)


def source_box_2d(shape, source: dict):
    """Box (0, 1, y0, y1, x0, x1) of the 2D mask (fdtd2d.py:207-227) or None."""
    H, W = shape
    pos = source.get("pos", None)
    if pos is not None:
        y, x = (int(v) for v in pos)
        if 0 <= y < H and 0 <= x < W:
            return (0, 1, y, y + 1, x, x + 1)
        return None
    row = int(source.get("row", H // 2))
    return (0, 1, row, row + 1, 0, W) if 0 <= row < H else None


def box_shape(box):
    z0, z1, y0, y1, x0, x1 = box
    return (z1 - z0, y1 - y0, x1 - x0)


def dense_mask(shape, box, profile=None, dtype=torch.float64, device="cpu"):
    """Dense amplitude mask (the reference's ``_build_source_mask`` layout)."""
    mask = torch.zeros(shape, dtype=dtype, device=device)
    if box is None:
        return mask
    z0, z1, y0, y1, x0, x1 = box
    val = 1.0 if profile is None else torch.as_tensor(profile, dtype=dtype, device=device)
    if len(shape) == 3:
        mask[z0:z1, y0:y1, x0:x1] = val
    else:
        mask[y0:y1, x0:x1] = val if profile is None else val.reshape(y1 - y0, x1 - x0)
    return mask


def comp_mask_3d(components):
    """Bit mask of 3D components from e.g. ("Ex", "Ez") or "xyz"."""
    if components is None:
        return 7
    if isinstance(components, str) and all(ch in "xyz" for ch in components):
        components = ["E" + ch for ch in components]
    m = 0
    for c in components:
        if c not in COMPONENTS_3D:
            raise ValueError(f"unknown monitor component {c!r}")
        m |= 1 << COMPONENTS_3D[c]
    return m


def source_profile(source: dict, box):
    """Amplitude profile over the box: ones, or source['profile'] broadcast."""
    shp = box_shape(box)
    prof = source.get("profile", None)
    if prof is None:
        return torch.ones(shp, dtype=torch.float64)
    p = torch.as_tensor(prof, dtype=torch.float64).detach().cpu()
    while p.dim() < 3:
        p = p.unsqueeze(0)
    return torch.broadcast_to(p, shp).contiguous()


def monitor_box(shape, spec: dict):
    """Monitor box from spec['box'] (3D: z0,z1,y0,y1,x0,x1; 2D: y0,y1,x0,x1)."""
    b = tuple(int(v) for v in spec["box"])
    if len(shape) == 2:
        if len(b) != 4:
            raise ValueError("2D monitor box must be (y0, y1, x0, x1)")
        b = (0, 1) + b
        dims = (1,) + tuple(shape)
    else:
        if len(b) != 6:
            raise ValueError("3D monitor box must be (z0, z1, y0, y1, x0, x1)")
        dims = tuple(shape)
    for lo, hi, n in zip(b[0::2], b[1::2], dims):
        if not (0 <= lo < hi <= n):
            raise ValueError(f"monitor box {spec['box']} outside the grid {tuple(shape)}")
    return b


def phase_table(n_steps: int, dt: float, omegas) -> np.ndarray:
    """(cos, sin) of omega_f (n+1) dt for n < n_steps: shape (n_steps, n_freq, 2).

    After step n the fields are E^{n+1}; the DFT is F = sum E^{n+1} e^{-i w (n+1) dt}.
    """
    n = np.arange(1, n_steps + 1, dtype=np.float64)[:, None]
    ph = n * dt * np.asarray(omegas, dtype=np.float64)[None, :]
    return np.stack([np.cos(ph), np.sin(ph)], axis=-1)


EPS_CHECKS = ("warn", "raise", "off")


def check_eps_finite(eps: torch.Tensor, mode: str) -> bool:
    """Host-side guard on the permittivity before any kernel runs.

    In the reference a NaN or Inf in eps flows through the step and the
    ``eps.sum() * 0`` term into every output (fdtd3d.py:162-163), and the
    solve runs to the end.  ``mode``: "warn" keeps that result and warns,
    "raise" raises ValueError before launching anything, "off" skips the
    check (it costs one device-to-host read when eps is on the GPU).
    Returns whether eps is finite (True when unchecked).
    """
    if mode not in EPS_CHECKS:
        raise ValueError(f"check_eps must be one of {EPS_CHECKS}, got {mode!r}")
    if mode == "off":
        return True
    finite = bool(torch.isfinite(eps.detach()).all())
    if not finite:
        bad = int((~torch.isfinite(eps.detach())).sum())
        msg = (f"permittivity has {bad} non-finite value(s) (NaN/Inf); every field and "
               "gradient of this solve will be non-finite")
        if mode == "raise":
            raise ValueError(msg)
        warnings.warn(msg, RuntimeWarning, stacklevel=3)
    return finite


def spare_device_bytes(dev):
    """Device memory a solve may take: what the driver reports free plus what
    torch's caching allocator holds without using it (its cached blocks are
    handed out before the driver is asked for more)."""
    import torch
    free, _ = torch.cuda.mem_get_info(dev)
    return int(free + torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev))

"""Seeded synthetic workloads of the benchmark configurations.

* ``blob_eps_2d``      2D Si/SiO2 blob design: uniform noise (torch
                       generator), Gaussian filter sigma=3 cells (truncate 4,
                       reflect), threshold 0.5 -> eps 12.1 / 2.085.
* ``waveguide_3d``     3D Si waveguide(s) along x in SiO2 with a seeded
                       continuous design slab eps = 2.085 + 10.015 rho.
* ``config_*``         the solver kwargs, source, monitors and loss of each
                       benchmark configuration.

Lengths are in nm; grids are (D, H, W) = (z, y, x), x fastest.
"""

from __future__ import annotations

import math

import numpy as np
import torch
from scipy.ndimage import gaussian_filter

EPS_SI = 12.1
EPS_SIO2 = 2.085


def blob_noise(seed, size=80):
    """The seeded uniform noise a blob design is filtered from."""
    g = torch.Generator().manual_seed(int(seed))
    return torch.rand(size, size, dtype=torch.float64, generator=g)


def blob_eps_2d(seed, H=200, W=200, lo=60, size=80, sigma=3.0):
    """200x200 field with an 80x80 blob design window at rows/cols 60..139."""
    noise = blob_noise(seed, size).numpy()
    blobs = gaussian_filter(noise, sigma, mode="reflect", truncate=4.0) > 0.5
    eps = np.full((H, W), EPS_SIO2)
    eps[lo:lo + size, lo:lo + size] = np.where(blobs, EPS_SI, EPS_SIO2)
    return eps


def _gauss_taps(sigma, truncate=4.0):
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    return w / w.sum(), r


def blob_eps_batch(noise, H, W, lo, sigma=3.0):
    """Batched ``blob_eps_2d`` on the noise's device: noise [B, S, S] float64
    -> eps [B, H, W] float64.  Separable Gaussian filter with half-sample
    symmetric padding (scipy.ndimage ``mode="reflect"``, truncate 4), then
    threshold 0.5 to Si / SiO2 inside the window at rows/cols lo..lo+S-1."""
    B, S, _ = noise.shape
    w, r = _gauss_taps(sigma)
    k = torch.as_tensor(w, dtype=torch.float64, device=noise.device)
    # half-sample symmetric index map: -1 -> 0, -2 -> 1, S -> S-1, ...
    idx = np.arange(-r, S + r)
    period = 2 * S
    idx = np.mod(idx, period)
    idx = np.where(idx >= S, period - 1 - idx, idx)
    it = torch.as_tensor(idx, device=noise.device)

    def filt(xp):
        # symmetric correlation along the last axis, w0 x[i] + sum_j w_j
        # (x[i-j] + x[i+j]); agrees with scipy.ndimage to ~1e-16 before the
        # threshold (the thresholded maps are identical; tested)
        acc = xp[..., r:r + S] * k[r]
        for j in range(1, r + 1):
            acc = acc + (xp[..., r - j:r - j + S] + xp[..., r + j:r + j + S]) * k[r + j]
        return acc

    x = filt(noise[:, :, it])                             # along x
    x = filt(x[:, it, :].transpose(1, 2)).transpose(1, 2)  # along y
    eps = torch.full((B, H, W), EPS_SIO2, dtype=torch.float64, device=noise.device)
    hi_ = torch.tensor(EPS_SI, dtype=torch.float64, device=noise.device)
    lo_ = torch.tensor(EPS_SIO2, dtype=torch.float64, device=noise.device)
    eps[:, lo:lo + S, lo:lo + S] = torch.where(x > 0.5, hi_, lo_)
    return eps


def waveguide_3d(shape, dl, design=(8, 64, 64), seed=0, split=False):
    """Si waveguide along x (400 nm wide in y, 220 nm tall in z), centred in
    SiO2, with a centred design block (dz, dy, dx) of eps = 2.085 + 10.015 rho,
    rho ~ U[0, 1] (numpy default_rng(seed)).  ``split``: the waveguide exists
    only outside the design block's x extent (input and output guides)."""
    D, H, W = shape
    eps = np.full(shape, EPS_SIO2)
    wy = int(round(400.0 / dl))
    wz = int(round(220.0 / dl))
    y0, z0 = (H - wy) // 2, (D - wz) // 2
    eps[z0:z0 + wz, y0:y0 + wy, :] = EPS_SI
    dz, dy, dx = design
    cz, cy, cx = (D - dz) // 2, (H - dy) // 2, (W - dx) // 2
    if split:
        eps[z0:z0 + wz, y0:y0 + wy, cx:cx + dx] = EPS_SIO2
    rho = np.random.default_rng(seed).random((dz, dy, dx))
    eps[cz:cz + dz, cy:cy + dy, cx:cx + dx] = EPS_SIO2 + (EPS_SI - EPS_SIO2) * rho
    return eps


def gaussian_sheet(shape, dl, sigma_nm=200.0):
    """Gaussian yz profile (D, H, 1) centred on the waveguide axis."""
    D, H, _ = shape
    s = sigma_nm / dl
    zz, yy = np.meshgrid(np.arange(D) - (D - 1) / 2, np.arange(H) - (H - 1) / 2, indexing="ij")
    return np.exp(-(zz**2 + yy**2) / (2 * s**2))[:, :, None]


def config_c2(seed=0, n_steps=8000):
    """256^3, dl 25 nm, 16-cell CPML, 8000 steps: 128x128x16 random-rho slab
    between input and output Si waveguides; Ey Gaussian sheet with the
    default broadband pulse at 1550 nm; DFT of Ey on the output yz plane at
    1450/1500/1550/1600/1650 nm; loss = sum |DFT Ey|^2."""
    shape = (256, 256, 256)
    dl, L = 25.0, 16
    eps = waveguide_3d(shape, dl, design=(16, 128, 128), seed=seed, split=True)
    solver = dict(grid_shape=shape, dl=dl, wavelength_nm=1550.0, pml_layers=L, n_steps=n_steps,
                  courant=0.4, backward="time_reversal", checkpoint_every=64)
    src = {"type": "gaussian_pulse", "plane": "yz", "x": L + 4, "component": "Ey",
           "profile": torch.from_numpy(gaussian_sheet(shape, dl))}
    xm = shape[2] - L - 4
    mons = [{"box": (0, shape[0], 0, shape[1], xm, xm + 1), "components": ["Ey"],
             "wavelengths": [1450.0, 1500.0, 1550.0, 1600.0, 1650.0]}]
    return eps, solver, src, mons


def config_c1(seed=0, n_steps=4000):
    """128^3, dl 25 nm, 12-cell CPML, 4000 steps: straight Si waveguide with a
    64x64x8 random-rho design region; Ey Gaussian sheet at x = L+4; DFT of
    Ey at 1500/1550/1600 nm on the output plane x = W-L-4."""
    shape = (128, 128, 128)
    dl, L = 25.0, 12
    eps = waveguide_3d(shape, dl, design=(8, 64, 64), seed=seed)
    solver = dict(grid_shape=shape, dl=dl, wavelength_nm=1550.0, pml_layers=L, n_steps=n_steps,
                  courant=0.4, backward="time_reversal", checkpoint_every=64)
    src = {"type": "gaussian_pulse", "plane": "yz", "x": L + 4, "component": "Ey",
           "profile": torch.from_numpy(gaussian_sheet(shape, dl))}
    xm = shape[2] - L - 4
    mons = [{"box": (0, shape[0], 0, shape[1], xm, xm + 1), "components": ["Ey"],
             "wavelengths": [1500.0, 1550.0, 1600.0]}]
    return eps, solver, src, mons


def config_c0(seed=0):
    """200x200 TM, dl 20 nm, 10-cell CPML, courant 0.5, 2000 steps; point Ez
    source at (40, 100) with a 200 nm-bandwidth pulse at 1550 nm; loss =
    sum |DFT Ez(1550 nm)|^2 over the 1x60 line at row 160, cols 70..129."""
    lam0 = 1550.0
    spread = lam0**2 / (2 * math.pi * 200.0)
    eps = blob_eps_2d(seed)
    solver = dict(grid_shape=(200, 200), dl=20.0, wavelength_nm=lam0, pml_layers=10,
                  n_steps=2000, courant=0.5)
    src = {"type": "gaussian_pulse", "pos": [40, 100], "spread": spread, "t0": 3 * spread}
    mons = [{"box": (160, 161, 70, 130), "wavelengths": [lam0]}]
    return eps, solver, src, mons


class SurrogateConfig:
    """128x128 TM surrogate workload with the C0 physics: dl 20 nm, 10-cell
    CPML, courant 0.5, 2000 steps, point Ez source at (16, 64) with the
    200 nm-bandwidth pulse at 1550 nm; an 80x80 blob window at rows/cols
    24..103; the surrogate target is the complex DFT of Ez over the whole
    grid at 1500/1550/1600 nm; the figure of merit (maximised) is
    sum |DFT Ez(1550 nm)|^2 over the 1x60 line at row 112, cols 34..93."""

    H = W = 128
    lo, size = 24, 80
    wavelengths = (1500.0, 1550.0, 1600.0)
    fom_index = 1
    line = (112, 113, 34, 94)

    def __init__(self, n_steps=2000):
        lam0 = 1550.0
        spread = lam0**2 / (2 * math.pi * 200.0)
        self.solver = dict(grid_shape=(self.H, self.W), dl=20.0, wavelength_nm=lam0,
                           pml_layers=10, n_steps=n_steps, courant=0.5)
        self.source = {"type": "gaussian_pulse", "pos": [16, 64], "spread": spread,
                       "t0": 3 * spread}

    def eps(self, noise):
        return blob_eps_batch(noise, self.H, self.W, self.lo)

    def fom(self, fields):
        """FoM of full-grid DFT fields [B, 3, H, W] (complex)."""
        y0, y1, x0, x1 = self.line
        f = fields[:, self.fom_index, y0:y1, x0:x1]
        return (f.abs() ** 2).flatten(1).sum(1)


def monitor_loss(monitors_out, comp_index=0):
    """sum over monitors and wavelengths of |F|^2 of one recorded component."""
    tot = 0.0
    for F in monitors_out:
        f = F[:, comp_index] if F.dim() >= 5 else F
        tot = tot + (f.abs() ** 2).sum()
    return tot

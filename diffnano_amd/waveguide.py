"""Waveguide mode analysis, mode-overlap figure of merit and the waveguide
inverse-design loops, drop-in for DiffNano's ``WaveguideDesigner``
(``diffnano/workflows/waveguide.py``) driving this package's FDTD.

* ``waveguide_mode_1d`` (waveguide.py:23-87): eigenmodes of the 1D slab
  operator d^2/dx^2 / dl^2 (second difference, zero outside) + k0^2 eps;
  n_eff = sqrt(max(beta^2, 0)) / k0, sorted by decreasing n_eff, unit-norm
  profiles.  Dense FP64 symmetric eigensolve (a few hundred points),
  differentiable through ``torch.linalg.eigh``.
* ``WaveguideDesigner.mode_overlap`` (waveguide.py:216-245):
  |<E, M>|^2 / (|E|^2 (|M|^2 + 1e-12)) with the Hermitian inner product;
  ``transmission_loss`` (waveguide.py:247-267): -log(overlap + 1e-8).
* ``optimize_bend`` (waveguide.py:269-378): density -> Heaviside projection
  (beta 1 -> 16) -> permittivity in the design region -> ``solver.forward``
  -> transmission loss -> backward -> Adam -> clamp to [0, 1].  With this
  package's ``FDTDSolver2D`` the gradient is the hand-written discrete
  adjoint on the GPU.  Keyword-only extension ``fom="dft"``: the overlap is
  taken between the input mode and the running DFT of Ez on an output
  cross-section (a DFT monitor column at ``output_x``), whose adjoint
  source the solver injects in its backward sweep; the default
  ``fom="field"`` is the reference's overlap with the final-time field.
* ``optimize_mode_converter`` (waveguide.py:380-470): mode analysis only
  (no solver), as in the reference.

Reference quirks kept: ``fundamental_mode`` averages eps along axis 1 and
returns the 1D mode along axis 0 replicated over ``grid_shape[0]`` rows (so
the 2D profile varies along its last axis and has shape (H, H));
``waveguide_eps`` overwrites only the core rows with the geometry.
"""

from __future__ import annotations

import math

import torch

from .projection import beta_continuation_schedule, heaviside_projection

__all__ = ["WaveguideDesigner", "waveguide_mode_1d", "mode_overlap"]


def waveguide_mode_1d(eps_r: torch.Tensor, wavelength_nm: float, dl: float = 20.0,
                      n_modes: int = 1) -> tuple[torch.Tensor, torch.Tensor]:
    """(n_eff [n_modes], modes [n_modes, N]) of the 1D profile ``eps_r``
    (waveguide.py:23-87)."""
    eps = eps_r.to(torch.float64)
    n = eps.shape[0]
    k0 = 2.0 * math.pi / wavelength_nm
    inv = 1.0 / (dl * dl)
    op = torch.diag(k0 * k0 * eps - 2.0 * inv)
    off = torch.full((n - 1,), inv, dtype=torch.float64, device=eps.device)
    op = op + torch.diag(off, 1) + torch.diag(off, -1)
    lam, vec = torch.linalg.eigh(op)
    neff = torch.sqrt(torch.clamp(lam / (k0 * k0), min=0.0))
    order = torch.argsort(neff, descending=True)[:n_modes]
    modes = vec[:, order].T
    nrm = modes.norm(dim=1, keepdim=True)
    modes = torch.where(nrm > 0, modes / torch.where(nrm > 0, nrm, torch.ones_like(nrm)), modes)
    return neff[order], modes


def mode_overlap(field: torch.Tensor, target_mode: torch.Tensor) -> torch.Tensor:
    """|<field, target>|^2 / (|field|^2 (|target|^2 + 1e-12)), flattened,
    Hermitian (waveguide.py:216-245).  In [0, 1]."""
    f = field.flatten()
    t = target_mode.flatten()
    if f.is_complex() or t.is_complex():
        ct = torch.promote_types(torch.promote_types(f.dtype, t.dtype), torch.complex64)
        f, t = f.to(ct), t.to(ct)
    num = torch.abs(torch.dot(f, t.conj())) ** 2
    den = f.norm() ** 2 * (t.norm() ** 2 + 1e-12)
    return num / den


class WaveguideDesigner:
    """Waveguide inverse design (reference signature, waveguide.py:90-140).

    ``solver``: any object with the Solver protocol's ``forward`` (this
    package's ``FDTDSolver2D``); ``None`` uses the built-in mode analysis."""

    def __init__(self, wavelength_nm: float = 1550.0, grid_shape: tuple[int, int] = (60, 60),
                 dl: float = 20.0, n_core: float = 2.5, n_clad: float = 1.0,
                 waveguide_width_nm: float = 400.0, device: str | torch.device = "cpu",
                 solver=None):
        self.wavelength_nm = wavelength_nm
        self.grid_shape = grid_shape
        self.dl = dl
        self.n_core = n_core
        self.n_clad = n_clad
        self.waveguide_width_nm = waveguide_width_nm
        self._device = torch.device(device)
        self.solver = solver

    @property
    def device(self) -> torch.device:
        return self._device

    def _core_rows(self, y_offset=0):
        H = self.grid_shape[0]
        w = int(self.waveguide_width_nm / self.dl)
        c = H // 2 + y_offset
        return max(0, c - w // 2), min(H, c + w // 2)

    def waveguide_eps(self, geometry: torch.Tensor | None = None,
                      y_offset: int = 0) -> torch.Tensor:
        """Straight waveguide along axis 1 (core rows of width
        int(width / dl) centred at H // 2 + y_offset); ``geometry`` sets the
        core rows to clad + (core - clad) * geometry (waveguide.py:142-183)."""
        H, W = self.grid_shape
        ec, eo = self.n_clad ** 2, self.n_core ** 2
        r0, r1 = self._core_rows(y_offset)
        eps = torch.full((H, W), ec, dtype=torch.float64, device=self._device)
        eps[r0:r1, :] = eo
        if geometry is not None:
            eps = eps.clone()
            eps[r0:r1, :] = ec + (eo - ec) * geometry[r0:r1, :]
        return eps

    def fundamental_mode(self, eps_r: torch.Tensor | None = None
                         ) -> tuple[torch.Tensor, torch.Tensor]:
        """(n_eff, mode profile (H, H)) of ``eps_r`` averaged along axis 1
        (waveguide.py:185-214)."""
        if eps_r is None:
            eps_r = self.waveguide_eps()
        neff, modes = waveguide_mode_1d(eps_r.mean(dim=1), self.wavelength_nm, dl=self.dl,
                                        n_modes=1)
        return neff[0], modes[0].unsqueeze(0).expand(self.grid_shape[0], -1)

    def mode_overlap(self, field: torch.Tensor, target_mode: torch.Tensor) -> torch.Tensor:
        return mode_overlap(field, target_mode)

    def transmission_loss(self, field: torch.Tensor, input_mode: torch.Tensor) -> torch.Tensor:
        """-log(overlap + 1e-8) (waveguide.py:247-267)."""
        return -torch.log(mode_overlap(field, input_mode) + 1e-8)

    def _region(self, design_region):
        H, W = self.grid_shape
        if design_region is None:
            m = H // 4
            design_region = (m, H - m, m, W - m)
        return design_region

    def _design_eps(self, projected, region):
        y0, y1, x0, x1 = region
        eps = self.waveguide_eps()
        ec, eo = self.n_clad ** 2, self.n_core ** 2
        eps[y0:y1, x0:x1] = ec + (eo - ec) * projected
        return eps

    def _bend_field(self, eps, fom, monitors):
        r = self.solver.forward(eps, wavelengths=[self.wavelength_nm],
                                source={"type": "gaussian_pulse", "amplitude": 1.0},
                                **({"monitors": monitors} if monitors else {}))
        if fom == "dft":
            return r.metadata["monitors"][0][0, :, 0]
        f = r.field
        if f.dim() == 3:
            return f[0]
        if f.dim() == 2:
            return f
        return f.reshape(self.grid_shape)

    def optimize_bend(self, bend_radius_nm: float = 2000.0, n_steps: int = 200,
                      lr: float = 0.01, design_region: tuple[int, int, int, int] | None = None,
                      verbose: bool = True, *, fom: str = "field",
                      output_x: int | None = None) -> tuple[torch.Tensor, list[float]]:
        """Topology optimisation of the design region for the overlap of the
        output with the input waveguide's fundamental mode
        (waveguide.py:269-378).  Returns (density (H, W) zero outside the
        region, loss per step).

        ``fom="field"``: overlap with the solver's final-time field (the
        reference).  ``fom="dft"`` (needs ``solver``): overlap of the 1D mode
        with the DFT of Ez at ``wavelength_nm`` on column ``output_x``
        (default: four cells inside the right absorbing layer of the
        solver, or W - 4)."""
        H, W = self.grid_shape
        region = self._region(design_region)
        y0, y1, x0, x1 = region
        density = torch.rand(y1 - y0, x1 - x0, device=self._device, dtype=torch.float64)
        density = density.detach().requires_grad_(True)
        neff, input_mode = self.fundamental_mode(self.waveguide_eps())
        monitors = None
        if fom == "dft":
            if self.solver is None:
                raise ValueError("fom='dft' needs a solver")
            if output_x is None:
                output_x = W - getattr(self.solver, "pml_layers", 0) - 4
            if not 0 <= output_x < W:
                raise ValueError(f"output_x {output_x} outside the grid width {W}")
            monitors = [{"box": (0, H, output_x, output_x + 1),
                         "wavelengths": [self.wavelength_nm]}]
            input_mode = input_mode[0]
        elif fom != "field":
            raise ValueError(f"fom must be 'field' or 'dft', got {fom!r}")
        output_mode = input_mode
        opt = torch.optim.Adam([density], lr=lr)
        history = []
        for step in range(n_steps):
            beta = beta_continuation_schedule(step, n_steps, beta_start=1.0, beta_end=16.0)
            eps = self._design_eps(heaviside_projection(density, beta=beta), region)
            if self.solver is not None:
                out = self._bend_field(eps, fom, monitors)
            else:
                _, out = self.fundamental_mode(eps)
            loss = self.transmission_loss(out, output_mode)
            opt.zero_grad()
            loss.backward()
            if density.grad is not None and torch.isnan(density.grad).any():
                break
            opt.step()
            with torch.no_grad():
                density.clamp_(0.0, 1.0)
            history.append(loss.item())
            if verbose and step % 20 == 0:
                ov = mode_overlap(out.detach(), output_mode).item()
                print(f"Step {step:4d}: loss={loss.item():.6f}, overlap={ov:.4f}")
        full = torch.zeros(H, W, device=self._device, dtype=torch.float64)
        full[y0:y1, x0:x1] = density.detach()
        return full, history

    def optimize_mode_converter(self, n_steps: int = 200, lr: float = 0.01,
                                design_region: tuple[int, int, int, int] | None = None,
                                verbose: bool = True) -> tuple[torch.Tensor, list[float]]:
        """Fundamental -> second-order mode converter on the mode analysis
        (waveguide.py:380-470).  Returns (density (H, W), loss per step)."""
        H, W = self.grid_shape
        region = self._region(design_region)
        y0, y1, x0, x1 = region
        density = torch.rand(y1 - y0, x1 - x0, device=self._device, dtype=torch.float64)
        density = density.detach().requires_grad_(True)
        _, modes = waveguide_mode_1d(self.waveguide_eps().mean(dim=1), self.wavelength_nm,
                                     dl=self.dl, n_modes=2)
        target = modes[1 if modes.shape[0] > 1 else 0].unsqueeze(0).expand(H, -1)
        opt = torch.optim.Adam([density], lr=lr)
        history = []
        for step in range(n_steps):
            beta = beta_continuation_schedule(step, n_steps, beta_start=1.0, beta_end=16.0)
            eps = self._design_eps(heaviside_projection(density, beta=beta), region)
            _, out = self.fundamental_mode(eps)
            loss = self.transmission_loss(out, target)
            opt.zero_grad()
            loss.backward()
            if density.grad is not None and torch.isnan(density.grad).any():
                break
            opt.step()
            with torch.no_grad():
                density.clamp_(0.0, 1.0)
            history.append(loss.item())
            if verbose and step % 20 == 0:
                ov = mode_overlap(out.detach(), target).item()
                print(f"Step {step:4d}: loss={loss.item():.6f}, overlap={ov:.4f}")
        full = torch.zeros(H, W, device=self._device, dtype=torch.float64)
        full[y0:y1, x0:x1] = density.detach()
        return full, history

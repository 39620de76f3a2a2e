"""Waveguide inverse design through the GPU FDTD: the mode-overlap loss of
``WaveguideDesigner.optimize_bend`` (final-time field, and the DFT-monitor
extension) differentiated by the hand-written adjoint, checked against
central finite differences of the FP64 kernels, and a short design run."""

import pytest
import torch

from diffnano_amd.fdtd2d import FDTDSolver2D
from diffnano_amd.projection import heaviside_projection
from diffnano_amd.waveguide import WaveguideDesigner

pytestmark = pytest.mark.gpu


def _loss_fn(wd, fom, beta=4.0):
    region = wd._region(None)
    H, W = wd.grid_shape
    _, mode = wd.fundamental_mode(wd.waveguide_eps())
    mons = None
    if fom == "dft":
        xo = W - wd.solver.pml_layers - 2
        mons = [{"box": (0, H, xo, xo + 1), "wavelengths": [wd.wavelength_nm]}]
        mode = mode[0]

    def f(density):
        eps = wd._design_eps(heaviside_projection(density, beta=beta), region)
        return wd.transmission_loss(wd._bend_field(eps, fom, mons), mode)
    return f, region


@pytest.mark.parametrize("fom", ["field", "dft"])
def test_bend_loss_gradient_matches_finite_differences(fom):
    solver = FDTDSolver2D(grid_shape=(24, 24), dl=20.0, pml_layers=4, n_steps=90,
                          device="cuda", dtype=torch.float64, checkpoint_every=16)
    wd = WaveguideDesigner(grid_shape=(24, 24), dl=20.0, device="cuda", solver=solver)
    f, (y0, y1, x0, x1) = _loss_fn(wd, fom)
    g = torch.Generator().manual_seed(3)
    d = torch.rand(y1 - y0, x1 - x0, generator=g, dtype=torch.float64).cuda()
    d.requires_grad_(True)
    loss = f(d)
    (grad,) = torch.autograd.grad(loss, d)
    assert torch.isfinite(grad).all() and float(grad.abs().max()) > 0
    h = 1e-5
    gmax = float(grad.abs().max())
    probes = [(0, 0), (3, 5), (6, 6), (11, 2), tuple(int(i) for i in
              torch.nonzero(grad.abs() == grad.abs().max())[0])]
    with torch.no_grad():
        for (i, j) in probes:
            dp, dm = d.detach().clone(), d.detach().clone()
            dp[i, j] += h
            dm[i, j] -= h
            fd = (float(f(dp)) - float(f(dm))) / (2 * h)
            print(fom, (i, j), float(grad[i, j]), fd)
            assert abs(fd - float(grad[i, j])) <= 1e-4 * gmax


def test_bend_design_loop_runs_and_improves():
    torch.manual_seed(0)
    solver = FDTDSolver2D(grid_shape=(40, 40), dl=20.0, pml_layers=8, n_steps=240,
                          device="cuda")
    wd = WaveguideDesigner(grid_shape=(40, 40), dl=20.0, device="cuda", solver=solver)
    for fom in ("field", "dft"):
        dens, hist = wd.optimize_bend(n_steps=8, lr=0.1, verbose=False, fom=fom)
        print(fom, hist)
        assert dens.shape == (40, 40) and len(hist) == 8
        assert all(torch.isfinite(torch.tensor(hist)))
        assert float(dens.min()) >= 0.0 and float(dens.max()) <= 1.0
        assert min(hist[1:]) < hist[0]

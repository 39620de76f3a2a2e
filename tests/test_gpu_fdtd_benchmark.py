"""GPU: the benchmark suite over FDTDSolver3D's backward modes, and the
external cross-validation path with the FP64 oracle as the external solver
(fdtd_benchmark.py:86-310)."""

import numpy as np
import pytest
import torch

from oracle import fdtd_oracle as O
from tests._helpers import torch_sin32

pytestmark = pytest.mark.gpu


def test_suite_runs_every_backward_mode_with_equal_gradients():
    from diffnano_amd import FDTDSolver3D
    from diffnano_amd.fdtd_benchmark import BenchmarkConfig, FDTDBenchmarkSuite
    cfg = BenchmarkConfig(grid_sizes=[(12, 12, 12), (16, 16, 16)], n_time_steps=40,
                          backward_modes=["autograd", "time_reversal", "checkpoint"],
                          device="cuda")
    suite = FDTDBenchmarkSuite()
    res = suite.run(FDTDSolver3D, cfg)
    assert len(res) == 6
    for r in res:
        assert r.device == "cuda" and r.forward_time_ms > 0 and r.backward_time_ms > 0
        assert r.peak_memory_mb >= 0
        if r.backward_mode == "autograd":
            assert r.gradient_cosine_vs_autograd is None
        else:   # the modes only change the adjoint's checkpoint interval
            assert r.gradient_cosine_vs_autograd == 1.0
    assert len(suite.summary_table().splitlines()) == 8


@pytest.mark.parametrize("dtype,rtol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_cross_validation_against_oracle(dtype, rtol):
    from diffnano_amd import FDTDSolver3D
    from diffnano_amd.fdtd_benchmark import ExternalCrossValidator
    v = ExternalCrossValidator()
    case = v.generate_test_case((12, 12, 12))
    shape, T, src = case["grid_size"], case["n_steps"], case["source_config"]
    s = FDTDSolver3D(grid_shape=shape, dl=case["dl"], wavelength_nm=case["wavelength_nm"],
                     pml_layers=2, n_steps=T, courant=case["courant"], device="cuda", dtype=dtype)
    field = s.forward(case["eps"].to("cuda"), source=src).field        # [Ez, Ex, Ey]
    p = O.Params3D(shape, dl=case["dl"], wavelength_nm=case["wavelength_nm"], pml_layers=2,
                   courant=case["courant"])
    wave = O.waveform_table(T, p.dt, p.omega, src, sin32=torch_sin32)
    fw = O.run3d(case["eps"].numpy(), p, T, 2, O.source_mask_3d(shape, src), wave)
    Ex, Ey, Ez = fw["E"]
    ext = torch.from_numpy(np.stack([Ez, Ex, Ey]))
    r = v.validate_forward(field, ext, rtol=rtol)
    assert r["passed"], r
    assert ext.abs().max() > 0


def test_gpu_benchmark_matrix_and_stability_report():
    from diffnano_amd.fdtd_benchmark_gpu import FDTDBenchmarkConfig, FDTDGPUBenchmark
    b = FDTDGPUBenchmark(FDTDBenchmarkConfig(grid_sizes=[(12, 12, 12)], n_time_steps=30,
                                             device="cuda"))
    assert b.detect_gpu()
    res = b.run_all()
    assert [r["strategy"] for r in res] == ["full_autodiff", "time_reversal", "checkpoint"]
    assert res[0]["gradient_cosine"] is None
    assert all(r["gradient_cosine"] == 1.0 for r in res[1:])
    assert all(r["device"] == "cuda" and r["peak_memory_mb"] > 0 for r in res)
    rep = b.cpu_fallback_benchmark()                  # results on the host, kernels on the GPU
    assert rep.is_valid and rep.gradient_cosines == [1.0, 1.0]


def test_fdtdx_self_consistency_run():
    from diffnano_amd.fdtd_benchmark_gpu import FDTDXCrossValidator
    rep = FDTDXCrossValidator().run_cross_validation(n_seeds=2, grid_size=(12, 12, 12),
                                                     n_steps=20)
    assert rep.forward_errors == [0.0, 0.0]            # deterministic kernels
    assert all(abs(c - 1.0) < 1e-12 for c in rep.gradient_cosines) and rep.is_valid

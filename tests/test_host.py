"""Host-side tests (no GPU): the C ABI library loads and exports every
declared symbol, struct layouts match the C header, argument errors come
back as error codes, and the host set-up (profiles, waveforms, boxes)
matches the reference goldens and the oracle."""

import ctypes as C
import json
import math
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest
import torch

from diffnano_amd import _common as H
from diffnano_amd import _lib as L
from oracle import fdtd_oracle as O
from tests._helpers import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dnx_fdtd.h")


def _declared():
    inc = os.path.dirname(HEADER)
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                   if f.endswith(".h"))
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(dnx_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.EXPORTED)
    assert lib.dnx_abi_version() == L.ABI_VERSION


def test_struct_layout_matches_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "dnx_fdtd.h"
int main(void){
 printf("%zu %zu %zu %zu %zu\n", sizeof(dnx_box), sizeof(dnx_grid3d), sizeof(dnx_grid2d),
        sizeof(dnx_source), sizeof(dnx_monitor));
 printf("%zu %zu %zu %zu\n", offsetof(dnx_grid3d, kh), offsetof(dnx_grid3d, ax),
        offsetof(dnx_source, profile), offsetof(dnx_monitor, grad));
 return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out]
    assert sizes[:5] == [C.sizeof(L.Box), C.sizeof(L.Grid3D), C.sizeof(L.Grid2D),
                         C.sizeof(L.Source), C.sizeof(L.Monitor)]
    assert sizes[5:] == [L.Grid3D.kh.offset, L.Grid3D.ax.offset, L.Source.profile.offset,
                         L.Monitor.grad.offset]


def test_abi_argument_errors_return_codes():
    lib = L.load()
    fin = C.c_int32()
    rc = lib.dnx_fdtd3d_run(None, None, None, 0, 1, None, 0, None, 0, None, 0, None, -1, None,
                            None, C.byref(fin))
    assert rc == -1 and b"grid3d" in lib.dnx_last_error()
    g = L.Grid3D(4, 4, 4, 7, 0.1, 1, 1, 1, 1, 1)
    rc = lib.dnx_fdtd3d_run(C.byref(g), 1, None, 0, 1, None, 0, None, 0, None, 0, None, -1, None,
                            None, C.byref(fin))
    assert rc == -1 and b"dtype" in lib.dnx_last_error()
    g2 = L.Grid2D(4, 4, 0, 2, 0.1, 1, 1, 1, 1, 1)
    rc = lib.dnx_fdtd2d_run(C.byref(g2), 1, None, 0, 1, None, 0, None, 0, None, 0, None, -1,
                            None, None, C.byref(fin))
    assert rc == -1 and b"te" in lib.dnx_last_error()
    rc = lib.dnx_prepare_coef(None, None, 0, 1.0, 0, None)
    assert rc == -1


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU behaviour")
def test_product_fails_loudly_without_gpu():
    from diffnano_amd import FDTDSolver2D, FDTDSolver3D
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        FDTDSolver3D(grid_shape=(6, 6, 6), pml_layers=1, n_steps=2).forward(torch.ones(6, 6, 6))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        FDTDSolver2D(grid_shape=(6, 6), pml_layers=1, n_steps=2).forward(torch.ones(6, 6))


def test_cpml_profile_matches_reference():
    t = golden("tables.npz")
    for key in [k[:-2] for k in t.files if k.startswith("cpml") and k.endswith("_b")]:
        _, dim, n, Lr, dl, cour = key.split("_")
        dt = float(cour) * float(dl) / math.sqrt(float(dim[0]))
        b, c = H.cpml_profile(int(n), int(Lr), float(dl), dt)
        np.testing.assert_array_equal(b, t[key + "_b"])
        np.testing.assert_allclose(c, t[key + "_c"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_waveform_table_matches_reference(dim):
    t = golden("tables.npz")
    srcs = json.loads(str(t["wave_meta"]))
    dl, T = (25.0, 4000) if dim == 3 else (20.0, 2000)
    dt = (0.4 * dl / math.sqrt(3)) if dim == 3 else (0.5 * dl / math.sqrt(2))
    om = 2 * math.pi / 1550.0
    for name, src in srcs.items():
        np.testing.assert_allclose(H.waveform_table(T, dt, om, src), t[f"wave{dim}d_{name}"],
                                   rtol=1e-14, atol=0, err_msg=name)


@pytest.mark.parametrize("src", [{}, {"pos": [1, 2, 3]}, {"pos": [9, 0, 0]}, {"plane": "xy"},
                                 {"plane": "xz", "y": 4}, {"plane": "yz", "x": 0},
                                 {"plane": "yz", "x": 7}, {"plane": "qq"}])
def test_source_boxes_match_oracle_masks(src):
    shape = (6, 5, 7)
    box = H.source_box_3d(shape, src)
    ref = O.source_mask_3d(shape, src)
    got = H.dense_mask(shape, box).numpy()
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("src", [{}, {"pos": [1, 2]}, {"pos": [-1, 2]}, {"row": 3}, {"row": 9}])
def test_source_boxes_2d_match_oracle_masks(src):
    shape = (6, 5)
    np.testing.assert_array_equal(H.dense_mask(shape, H.source_box_2d(shape, src)).numpy(),
                                  O.source_mask_2d(shape, src))


def test_phase_table_convention():
    tab = H.phase_table(5, 2.0, [0.1, 0.3])
    assert tab.shape == (5, 2, 2)
    ph = O.dft_phase_table(5, 2.0, [0.1, 0.3])
    np.testing.assert_allclose(tab[..., 0], np.cos(ph), rtol=0, atol=0)
    np.testing.assert_allclose(tab[..., 1], np.sin(ph), rtol=0, atol=0)


def test_monitor_box_validation():
    assert H.monitor_box((4, 5, 6), {"box": (0, 4, 1, 2, 3, 6)}) == (0, 4, 1, 2, 3, 6)
    assert H.monitor_box((5, 6), {"box": (1, 2, 0, 6)}) == (0, 1, 1, 2, 0, 6)
    with pytest.raises(ValueError):
        H.monitor_box((4, 5, 6), {"box": (0, 5, 0, 1, 0, 1)})
    with pytest.raises(ValueError):
        H.monitor_box((5, 6), {"box": (0, 1, 0, 1, 0, 1)})
    assert H.comp_mask_3d("xz") == 5 and H.comp_mask_3d(["Ey"]) == 2


def test_autotune_cache_corrupt_file_is_discarded(tmp_path):
    from diffnano_amd import autotune
    p = tmp_path / "tiles.json"
    p.write_text("{not json")
    assert autotune._load_cache(str(p)) == {}
    p.write_text("[1, 2]")
    assert autotune._load_cache(str(p)) == {}
    autotune._save_cache(str(p), {"k": {"ty": 4, "zc": 16}})
    assert autotune._load_cache(str(p)) == {"k": {"ty": 4, "zc": 16}}
    assert autotune._valid({"ty": 4, "zc": 16}) and not autotune._valid({"ty": 3, "zc": 16})


def test_blob_batch_generator_matches_scipy_blobs():
    """The batched (torch) blob generator gives the thresholded maps of the
    scipy-filtered reference generator, for the 200x200 and 128x128 layouts."""
    from diffnano_amd import workloads as W
    seeds = list(range(24))
    z = torch.stack([W.blob_noise(s) for s in seeds])
    for (H, lo) in ((200, 60), (128, 24)):
        e = W.blob_eps_batch(z, H, H, lo).numpy()
        for i, s in enumerate(seeds):
            np.testing.assert_array_equal(e[i], W.blob_eps_2d(s, H=H, W=H, lo=lo))

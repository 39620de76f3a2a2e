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


# ---------------------------------------------------------------------------
# Allocation statement, options, workspace query, non-finite eps guard
# ---------------------------------------------------------------------------


def test_library_sources_allocate_no_device_memory():
    """The C ABI allocates nothing: no hipMalloc / hipMallocAsync /
    hipFree / hipHostMalloc call anywhere in the HIP sources (buffers and
    workspaces come from the caller; torch's caching allocator owns them)."""
    csrc = os.path.join(ROOT, "diffnano_amd", "csrc")
    pat = re.compile(r"\bhip(Malloc\w*|Free\w*|HostMalloc|HostAlloc|ExtMalloc\w*)\s*\(")
    for f in sorted(os.listdir(csrc)):
        text = open(os.path.join(csrc, f)).read()
        code = re.sub(r"//[^\n]*|/\*.*?\*/", "", text, flags=re.S)
        assert not pat.search(code), f"{f}: {pat.search(code).group(0)}"


def test_options_round_trip_and_reject_unknown_ids():
    lib = L.load()
    for name, oid in L.OPTIONS.items():
        v = C.c_int32(-7)
        assert lib.dnx_get_option(oid, C.byref(v)) == 0 and v.value >= 0, name
        assert lib.dnx_set_option(oid, v.value + 1) == 0
        w = C.c_int32()
        lib.dnx_get_option(oid, C.byref(w))
        assert w.value == v.value + 1
        assert lib.dnx_set_option(oid, v.value) == 0
    assert lib.dnx_set_option(99, 1) == -1 and b"unknown option" in lib.dnx_last_error()
    assert lib.dnx_get_option(-1, C.byref(C.c_int32())) == -1


def _fake_grid(shape, code):
    # host-side queries never dereference the coefficient pointers
    return L.Grid3D(shape[0], shape[1], shape[2], code, 0.2, 16, 16, 16, 16, 16)


def test_adjoint_workspace_query():
    lib = L.load()
    g = _fake_grid((256, 256, 256), L.DNX_F32)
    m = L.Monitor(L.Box(0, 256, 0, 256, 236, 237), 2, 5, 16, None, 16)
    assert lib.dnx_fdtd3d_adjoint_workspace_bytes(C.byref(g), C.byref(m), 1) == 256 * 256 * 4
    m3 = L.Monitor(L.Box(0, 4, 0, 5, 6, 9), 7, 2, 16, None, 16)
    assert lib.dnx_fdtd3d_adjoint_workspace_bytes(C.byref(g), C.byref(m3), 1) == 3 * 60 * 4
    assert lib.dnx_fdtd3d_adjoint_workspace_bytes(C.byref(g), None, 0) == 0
    pair = (L.Monitor * 2)(m, m3)   # more monitors than the two-step kernels take
    assert lib.dnx_fdtd3d_adjoint_workspace_bytes(C.byref(g), pair, 2) == 0
    g64 = _fake_grid((256, 256, 256), L.DNX_F64)  # the float64 path has no two-step kernels
    assert lib.dnx_fdtd3d_adjoint_workspace_bytes(C.byref(g64), C.byref(m), 1) == 0
    bad = _fake_grid((0, 4, 4), L.DNX_F32)
    assert lib.dnx_fdtd3d_adjoint_workspace_bytes(C.byref(bad), None, 0) < 0


@pytest.mark.parametrize("dim", [2, 3])
def test_non_finite_eps_is_flagged_before_any_launch(dim):
    from diffnano_amd import FDTDSolver2D, FDTDSolver3D
    shape = (6, 7, 8) if dim == 3 else (7, 8)
    eps = torch.full(shape, 2.0, dtype=torch.float64)
    eps[(1,) * dim] = float("nan")
    mk = (lambda **kw: FDTDSolver3D(grid_shape=shape, pml_layers=1, n_steps=4, **kw)) if dim == 3 \
        else (lambda **kw: FDTDSolver2D(grid_shape=shape, pml_layers=1, n_steps=4, **kw))
    with pytest.raises(ValueError, match="non-finite"):
        mk(check_eps="raise").forward(eps)
    eps[(1,) * dim] = float("inf")
    with pytest.raises(ValueError, match="1 non-finite"):
        mk(check_eps="raise").forward(eps)
    if not torch.cuda.is_available():
        # "warn" warns, then reaches the device selection (no GPU here)
        with pytest.warns(RuntimeWarning, match="non-finite"):
            with pytest.raises(RuntimeError, match="no CPU fallback"):
                mk().forward(eps)
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            mk(check_eps="off").forward(eps)
    with pytest.raises(ValueError, match="check_eps"):
        mk(check_eps="sometimes")


def test_batched_backward_refuses_a_stale_forward():
    """The batched plan keeps one set of checkpoints: a backward after a
    later forward on the same batch raises instead of returning a gradient
    of the wrong forward (host logic of _BatchedFn, with a stand-in batch)."""
    from diffnano_amd.batched import BatchedFDTD3D

    class FakeBatch:
        out_device = torch.device("cpu")
        forward = BatchedFDTD3D.forward
        _forward_detached = BatchedFDTD3D._forward_detached
        _next_gen = BatchedFDTD3D._next_gen
        _current_gen = BatchedFDTD3D._current_gen

        def __init__(self, owner=None):
            self._gen = 0
            self._gen_owner = owner if owner is not None else self
            self.accs = [torch.zeros(1, 1, 1, 1, 2, 2, dtype=torch.float64)]
            self.calls = []

        def _set_eps(self, eps_list):
            self.accs[0].fill_(float(eps_list[0].sum()))

        def _run_forward(self, record):
            self.calls.append(("fwd", record))

        def _run_backward(self, grads):
            self.calls.append(("bwd",))
            return [torch.ones(2, 2, 2, dtype=torch.float64)]

    b = FakeBatch()
    e = torch.ones(2, 2, 2, dtype=torch.float64, requires_grad=True)
    out = b.forward([e])
    out[0].abs().sum().backward()           # fresh: fine
    assert torch.equal(e.grad, torch.ones(2, 2, 2, dtype=torch.float64))
    out1 = b.forward([e])
    with torch.no_grad():
        b.forward([e])                       # an evaluation forward in between
    with pytest.raises(RuntimeError, match="another forward ran"):
        out1[0].abs().sum().backward()
    assert b.calls.count(("bwd",)) == 1
    assert ("fwd", False) in b.calls and ("fwd", True) in b.calls
    # plans sharing their checkpoint stores share the counter: a forward of
    # the other plan invalidates this one's backward, its own backward is fine
    class Owner:
        _gen = 0
    own = Owner()
    b1, b2 = FakeBatch(own), FakeBatch(own)
    o1 = b1.forward([e])
    o2 = b2.forward([e])
    o2[0].abs().sum().backward()
    with pytest.raises(RuntimeError, match="another forward ran"):
        o1[0].abs().sum().backward()


def _designs(shapes, dtype=L.DNX_F32):
    arr = (L.Design3D * len(shapes))()
    for d, (nz, ny, nx) in enumerate(shapes):
        arr[d].grid.nz, arr[d].grid.ny, arr[d].grid.nx = nz, ny, nx
        arr[d].grid.dtype = dtype
        arr[d].src.comp = -1
    return arr


def _batch_tile_shape(lib, kernel):
    ty, zc = C.c_int32(), C.c_int32()
    assert lib.dnx_fdtd3d_get_tiles(L.DNX_F32, kernel, C.byref(ty), C.byref(zc)) == 0
    t = ty.value if ty.value > 0 else 2
    return (min(t, 2) if kernel == 3 else t), zc.value


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("kernel", [2, 3])
def test_batched_block_map_covers_every_tile_once_and_balances_the_xcds(seed, kernel):
    """The two-step batched launches' work-group map: every tile of every
    design exactly once, idle entries (-1, -1) only, a design's tiles on one
    XCD in increasing order, and the per-XCD cost within one run (1/16 of
    the batch, 1/64 below 16 designs, or one tile) of the
    longest-processing-time bound, so that
    even a batch of one design spreads over all 8 XCDs (work-group b runs
    on XCD b % 8)."""
    lib = L.load()
    rng = np.random.default_rng(seed)
    classes = [(96, 96, 96), (128, 128, 128), (96, 160, 160), (128, 128, 192), (10, 12, 40),
               (37, 21, 70)]
    n = int(rng.integers(1, 60)) if seed else 1
    shapes = [classes[i] for i in rng.integers(0, len(classes), n)]
    arr = _designs(shapes)
    cnt = lib.dnx_batch3d_block_map(arr, n, kernel, None, 0)
    assert cnt > 0 and cnt % 8 == 0
    buf = (C.c_int32 * (2 * cnt))()
    assert lib.dnx_batch3d_block_map(arr, n, kernel, buf, cnt - 1) == -1
    assert lib.dnx_batch3d_block_map(arr, n, kernel, buf, cnt) == cnt
    m = np.frombuffer(buf, dtype=np.int32).reshape(-1, 2)
    ty, zc = _batch_tile_shape(lib, kernel)
    tiles = [-(-ny // ty) * -(-nz // zc) for nz, ny, nx in shapes]
    cost = [(-(-nx // 64) * 64) * ty * (min(zc, nz) + 2) for nz, ny, nx in shapes]
    live = m[m[:, 0] >= 0]
    assert np.all(m[m[:, 0] < 0] == -1)
    seen = sorted(map(tuple, live.tolist()))
    assert seen == sorted((d, t) for d in range(n) for t in range(tiles[d]))
    load = np.zeros(8)
    for k in range(8):
        col = m[k::8]
        col = col[col[:, 0] >= 0]
        for d in set(col[:, 0].tolist()):
            assert np.all(np.diff(col[col[:, 0] == d][:, 1]) > 0)
        load[k] = sum(cost[d] for d in col[:, 0])
    total = sum(c * t for c, t in zip(cost, tiles))
    runs = 64 if n < 16 else 16
    run = max(total // runs, max(cost)) + max(cost)
    assert load.max() <= total / 8 + run
    if sum(tiles) >= 8 * runs:
        assert np.all(load > 0)


def test_batched_queries_reject_grids_the_batch_cannot_take():
    lib = L.load()
    for bad in [(8, 8, 300), (0, 8, 8), (8, -1, 8), (1 << 20, 1 << 12, 256)]:
        arr = _designs([(16, 16, 16), bad])
        assert lib.dnx_batch3d_workspace_bytes(arr, 2) == -1, bad
        assert lib.dnx_batch3d_side_elems(arr, 2) == -1, bad
        assert lib.dnx_batch3d_block_map(arr, 2, 2, None, 0) == -1, bad
    arr = _designs([(16, 16, 16)])
    assert lib.dnx_batch3d_block_map(arr, 1, 1, None, 0) == -1     # one-step kernels: no map
    assert lib.dnx_batch3d_workspace_bytes(arr, 1) > 0
    arr[0].mon.n_freq = 1
    arr[0].mon.box = L.Box(0, 4, 0, 4, 0, 99)                      # box past the grid
    assert lib.dnx_batch3d_side_elems(arr, 1) == -1
    arr[0].mon.box = L.Box(0, 4, 0, 4, 0, 5)
    arr[0].mon.comp_mask = 5
    assert lib.dnx_batch3d_side_elems(arr, 1) == 2 * 2 * 4 * 4 * 5

"""CPU: the C ABI's host code under AddressSanitizer and UBSan.

Builds the library with ``-fsanitize=address,undefined`` on the host pass
only (the device code is built as usual and never launched here), into
``_build/asan/``, and drives every entry point whose work is
host-only in a child process that preloads the ASan runtime: version and
error strings, options, tile shapes, the adjoint and batched workspace
queries, the batched block maps over random design lists, and the
argument checks of the run, adjoint, batch-plan and FNO entry points, which
return before any HIP call.  Any heap overflow, use after free or undefined
behaviour aborts the child.
"""

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "diffnano_amd", "csrc")
OUT = os.path.join(ROOT, "_build", "asan")
LIB = os.path.join(OUT, "libdnx_fdtd_asan.so")
HIPCC = "/opt/rocm/bin/hipcc"

# host-pass sanitizers only
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host",
       "-fno-omit-frame-pointer"]


def _runtime():
    hits = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    return hits[0] if hits else None


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-O1", "-g", "-std=c++17", "-fPIC", "--offload-arch=gfx950", *SAN,
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-ffp-contract=fast-honor-pragmas", "-fvisibility=hidden", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return obj


def _build():
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-shared-libsan",
            *SAN[:4], "-o", LIB, *objs]
    r = subprocess.run(link, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


CHILD = r'''
import ctypes as C, importlib.util, sys, types
import numpy as np
sys.modules["torch"] = types.ModuleType("torch")      # the bindings use torch only at call time
spec = importlib.util.spec_from_file_location("dnx_lib", sys.argv[1] + "/diffnano_amd/_lib.py")
L = importlib.util.module_from_spec(spec); spec.loader.exec_module(L)
lib = L.load(sys.argv[2])
assert lib.dnx_abi_version() == L.ABI_VERSION
# options and tiles: valid round trips, rejected ids and values
v = C.c_int32()
for oid in range(-2, 12):
    rc = lib.dnx_get_option(oid, C.byref(v))
    if rc == 0:
        assert lib.dnx_set_option(oid, v.value) == 0
    else:
        assert lib.dnx_set_option(oid, 1) != 0 and lib.dnx_last_error()
ty, zc = C.c_int32(), C.c_int32()
for dt in (-1, 0, 1, 2):
    for k in range(-1, 6):
        if lib.dnx_fdtd3d_get_tiles(dt, k, C.byref(ty), C.byref(zc)) == 0:
            lib.dnx_fdtd3d_set_tiles(dt, k, 0, 0)
            lib.dnx_fdtd3d_set_tiles(dt, k, 3, 1 << 30)
            assert lib.dnx_fdtd3d_set_tiles(dt, k, ty.value, zc.value) == 0
        lib.dnx_last_error()
# adjoint workspace query over grids and monitor lists
rng = np.random.default_rng(0)
for _ in range(300):
    g = L.Grid3D(); g.nz, g.ny, g.nx = (int(x) for x in rng.integers(-2, 300, 3))
    g.dtype = int(rng.integers(-1, 3))
    nm = int(rng.integers(0, 4))
    ms = (L.Monitor * max(nm, 1))()
    for m in ms[:nm]:
        m.box = L.Box(*(int(x) for x in rng.integers(-3, 300, 6)))
        m.comp_mask = int(rng.integers(-1, 9)); m.n_freq = int(rng.integers(-1, 6))
    lib.dnx_fdtd3d_adjoint_workspace_bytes(C.byref(g), ms if nm else None, nm)
lib.dnx_fdtd3d_adjoint_workspace_bytes(None, None, 0)
# batched queries, block maps and plan argument checks over random design lists
classes = [(96, 96, 96), (128, 128, 128), (96, 160, 160), (128, 128, 192), (10, 12, 40),
           (0, 8, 8), (8, 8, 300), (3, 1, 1)]
for _ in range(200):
    n = int(rng.integers(1, 80))
    arr = (L.Design3D * n)()
    for d in range(n):
        arr[d].grid.nz, arr[d].grid.ny, arr[d].grid.nx = classes[int(rng.integers(0, len(classes)))]
        arr[d].grid.dtype = int(rng.integers(0, 2)); arr[d].src.comp = -1
        if rng.random() < 0.5:
            arr[d].mon.n_freq = int(rng.integers(1, 4)); arr[d].mon.comp_mask = int(rng.integers(1, 8))
            arr[d].mon.box = L.Box(*(int(x) for x in rng.integers(0, 20, 6)))
    ws = lib.dnx_batch3d_workspace_bytes(arr, n)
    lib.dnx_batch3d_side_elems(arr, n)
    for k in (2, 3):
        cnt = lib.dnx_batch3d_block_map(arr, n, k, None, 0)
        if cnt > 0:
            buf = (C.c_int32 * (2 * cnt))()
            assert lib.dnx_batch3d_block_map(arr, n, k, buf, cnt) == cnt
            assert lib.dnx_batch3d_block_map(arr, n, k, buf, cnt - 1) == -1
    plan = C.c_void_p()
    pools = L.Pools3D()
    # null pools / odd K / too small a workspace: rejected before any HIP call
    assert lib.dnx_batch3d_create(arr, n, 10, 8, None, None, 0, 0, None, C.byref(plan)) != 0
    assert lib.dnx_batch3d_create(arr, n, 10, 7, C.byref(pools), 1, 1 << 40, 0, None,
                                  C.byref(plan)) != 0
    assert lib.dnx_batch3d_create(arr, n, 10, 8, C.byref(pools), 1, max(ws - 1, 0), 0, None,
                                  C.byref(plan)) != 0
    assert not plan.value
# a plan rejected in its design loop (allocated, then freed)
arr = (L.Design3D * 3)()
for d, shp in enumerate([(16, 16, 16), (16, 16, 16), (16, 16, 300)]):
    arr[d].grid.nz, arr[d].grid.ny, arr[d].grid.nx = shp; arr[d].src.comp = -1
pools = L.Pools3D(); pools.state[0] = pools.state[1] = pools.lam[0] = pools.lam[1] = 1
pools.ckpt = pools.estore = pools.S = 1
arr2 = (L.Design3D * 2)(arr[0], arr[1]); arr2[1].grid.ce = None
ws = lib.dnx_batch3d_workspace_bytes(arr2, 2)
plan = C.c_void_p()
assert lib.dnx_batch3d_create(arr2, 2, 10, 8, C.byref(pools), 1, ws, 0, None, C.byref(plan)) != 0
assert lib.dnx_batch3d_create(arr, 3, 10, 8, C.byref(pools), 1, 1 << 40, 0, None,
                              C.byref(plan)) != 0
assert lib.dnx_batch3d_destroy(None) == 0
# run / adjoint / 2D / FNO argument checks
g = L.Grid3D(); g.nz = g.ny = g.nx = 8; g.dtype = 7
fin = C.c_int32()
assert lib.dnx_fdtd3d_run(None, None, None, 0, 1, None, 0, None, 0, None, 0, None, -1, None, None,
                          C.byref(fin)) != 0
assert lib.dnx_fdtd3d_run(C.byref(g), 1, None, 0, 1, None, 0, None, 0, None, 0, None, -1, None,
                          None, C.byref(fin)) != 0
g2 = L.Grid2D(); g2.ny = g2.nx = 8; g2.te = 5
assert lib.dnx_fdtd2d_run(C.byref(g2), 1, None, 0, 1, None, 0, None, 0, None, 0, None, -1, None,
                          None, C.byref(fin)) != 0
assert lib.dnx_prepare_coef(None, None, 0, 1.0, 0, None) != 0
ga = L.GemmArgs(); ga.M = -1; ga.A = ga.B = ga.C = 1
assert lib.dnx_gemm_bf16(C.byref(ga), None) != 0
assert lib.dnx_gemm_bf16(None, None) != 0
print("sanitized host checks passed")
'''


@pytest.mark.skipif(not (os.path.exists(HIPCC) and _runtime()),
                    reason="needs hipcc and the ASan runtime of the ROCm LLVM")
def test_c_abi_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    _build()
    syms = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert "__asan_report_load" in syms and "__ubsan_handle_" in syms   # instrumented
    env = dict(os.environ)
    env["LD_PRELOAD"] = _runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    script = tmp_path / "child.py"
    script.write_text(CHILD)
    r = subprocess.run([sys.executable, str(script), ROOT, LIB], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0 and "sanitized host checks passed" in r.stdout, \
        (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]

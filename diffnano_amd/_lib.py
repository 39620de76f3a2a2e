"""ctypes binding of libdnx_fdtd.so (the C ABI declared in include/dnx_fdtd.h).

The library is required: there is no CPU fallback.  ``lib()`` raises a
RuntimeError naming the missing file when it has not been built
(``python -m diffnano_amd._build``).
"""

from __future__ import annotations

import contextlib
import ctypes as C
import os
import threading

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdnx_fdtd.so")
ABI_VERSION = 5

DNX_F32 = 0
DNX_F64 = 1

_lock = threading.Lock()
_lib = None


class Box(C.Structure):
    _fields_ = [("z0", C.c_int32), ("z1", C.c_int32), ("y0", C.c_int32), ("y1", C.c_int32),
                ("x0", C.c_int32), ("x1", C.c_int32)]


class Grid3D(C.Structure):
    _fields_ = [("nz", C.c_int32), ("ny", C.c_int32), ("nx", C.c_int32), ("dtype", C.c_int32),
                ("kh", C.c_double), ("ce", C.c_void_p), ("by", C.c_void_p), ("ay", C.c_void_p),
                ("bx", C.c_void_p), ("ax", C.c_void_p)]


class Grid2D(C.Structure):
    _fields_ = [("ny", C.c_int32), ("nx", C.c_int32), ("dtype", C.c_int32), ("te", C.c_int32),
                ("kh", C.c_double), ("ce", C.c_void_p), ("by", C.c_void_p), ("ay", C.c_void_p),
                ("bx", C.c_void_p), ("ax", C.c_void_p)]


class Source(C.Structure):
    _fields_ = [("comp", C.c_int32), ("box", Box), ("profile", C.c_void_p), ("wave", C.c_void_p)]


class Monitor(C.Structure):
    _fields_ = [("box", Box), ("comp_mask", C.c_int32), ("n_freq", C.c_int32),
                ("phase", C.c_void_p), ("acc", C.c_void_p), ("grad", C.c_void_p)]


class Design3D(C.Structure):
    _fields_ = [("grid", Grid3D), ("src", Source), ("mon", Monitor)]


class Pools3D(C.Structure):
    _fields_ = [("state", C.c_void_p * 2), ("lam", C.c_void_p * 2), ("ckpt", C.c_void_p),
                ("estore", C.c_void_p), ("S", C.c_void_p), ("side", C.c_void_p),
                ("etail", C.c_void_p), ("a_tail", C.c_int32)]


class GemmArgs(C.Structure):
    """dnx_gemm_args (include/dnx_fno.h)."""
    _fields_ = [("M", C.c_int32), ("N", C.c_int32), ("K", C.c_int32), ("batch", C.c_int32),
                ("A", C.c_void_p), ("sAm", C.c_int64), ("sAk", C.c_int64), ("sAb", C.c_int64),
                ("B", C.c_void_p), ("sBk", C.c_int64), ("sBn", C.c_int64), ("sBb", C.c_int64),
                ("C", C.c_void_p), ("sCm", C.c_int64), ("sCn", C.c_int64), ("sCb", C.c_int64),
                ("Cin", C.c_void_p), ("sIm", C.c_int64), ("sIn", C.c_int64), ("sIb", C.c_int64),
                ("alpha", C.c_float), ("beta", C.c_float), ("bias", C.c_void_p),
                ("bias_mode", C.c_int32), ("act", C.c_int32), ("pre", C.c_void_p),
                ("a_bf16", C.c_int32), ("b_bf16", C.c_int32), ("c_bf16", C.c_void_p)]


_P = C.c_void_p
_I = C.c_int32
_L = C.c_int64

_SIGS = {
    "dnx_abi_version": ([], _I),
    "dnx_last_error": ([], C.c_char_p),
    "dnx_set_option": ([_I, _I], _I),
    "dnx_get_option": ([_I, C.POINTER(_I)], _I),
    "dnx_timing_enable": ([_I], _I),
    "dnx_timing_read": ([C.c_char_p, C.POINTER(C.c_double), C.POINTER(_L)], _I),
    "dnx_prepare_coef": ([_P, _P, _L, C.c_double, _I, _P], _I),
    "dnx_finalize_grad": ([_P, _P, _P, _L, _I, _P], _I),
    "dnx_fdtd3d_run": ([C.POINTER(Grid3D), _P, _P, _I, _I, C.POINTER(Source), _I,
                        C.POINTER(Monitor), _I, _P, _I, _P, _L, _P, _P, C.POINTER(_I)], _I),
    "dnx_fdtd3d_run_from": ([C.POINTER(Grid3D), _P, _P, _P, _I, _I, C.POINTER(Source), _I,
                             C.POINTER(Monitor), _I, _P, _I, _P, _L, _P, _P, C.POINTER(_I)], _I),
    "dnx_fdtd3d_adjoint_init": ([C.POINTER(Grid3D), _P, _P, C.POINTER(Monitor), _I, _I, _P, _P],
                                _I),
    "dnx_fdtd3d_adjoint_workspace_bytes": ([C.POINTER(Grid3D), C.POINTER(Monitor), _I], _L),
    "dnx_fdtd3d_adjoint_run": ([C.POINTER(Grid3D), _P, _P, _I, _I, _P, _P, C.POINTER(Source), _I,
                                C.POINTER(Monitor), _I, _P, _P, _L, _P, C.POINTER(_I)], _I),
    "dnx_fdtd3d_adjoint_static_scratch_bytes": ([], _L),
    "dnx_fdtd3d_adjoint_run_noalloc": ([C.POINTER(Grid3D), _P, _P, _I, _I, _P, _P,
                                        C.POINTER(Source), _I, C.POINTER(Monitor), _I, _P, _P,
                                        C.POINTER(_I)], _I),
    "dnx_fdtd3d_set_tiles": ([_I, _I, _I, _I], _I),
    "dnx_fdtd3d_get_tiles": ([_I, _I, C.POINTER(_I), C.POINTER(_I)], _I),
    "dnx_fdtd2d_tm_batch": ([C.POINTER(Grid2D), _I, _P, _I, C.POINTER(Source), C.POINTER(Box),
                             _I, _P, _P, _P], _I),
    "dnx_gemm_bf16": ([C.POINTER(GemmArgs), _P], _I),
    "dnx_gelu_bwd": ([_P, _P, _P, _P, _L, _P], _I),
    "dnx_fno_proj": ([_P, _I, _P, _P, _P, _P, _P, _I, _I, _I, _L, _P], _I),
    "dnx_fno_inverse_pw": ([_P, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P, _I, _I, _I, _I, _P],
                           _I),
    "dnx_sum_partials": ([_P, _I, _L, _P, C.c_float, _P], _I),
    "dnx_fno_mode_mix": ([_P, _P, _P, _P, _I, _I, _I, _L, _I, _I, _P], _I),
    "dnx_fno_mode_wgrad": ([_P, _P, _P, _P, _I, _I, _L, _I, _I, _P], _I),
    "dnx_fno_wgrad": ([_P, _P, _P, _P, _I, _I, _L, _I, _P], _I),
    "dnx_batch3d_workspace_bytes": ([C.POINTER(Design3D), _I], _L),
    "dnx_batch3d_side_elems": ([C.POINTER(Design3D), _I], _L),
    "dnx_stream_probe": ([_I, _P, _P, _L, _P], _I),
    "dnx_batch3d_block_map": ([C.POINTER(Design3D), _I, _I, C.POINTER(C.c_int32), _L], _L),
    "dnx_batch3d_create": ([C.POINTER(Design3D), _I, _I, _I, C.POINTER(Pools3D), _P, _L, _I, _P,
                            C.POINTER(_P)], _I),
    "dnx_batch3d_forward": ([_P, _I, _P], _I),
    "dnx_batch3d_backward": ([_P, _P], _I),
    "dnx_batch3d_uses_graph": ([_P], _I),
    "dnx_batch3d_final_slot": ([_P], _I),
    "dnx_batch3d_uses_pairs": ([_P], _I),
    "dnx_batch3d_destroy": ([_P], _I),
    "dnx_fdtd2d_run": ([C.POINTER(Grid2D), _P, _P, _I, _I, C.POINTER(Source), _I,
                        C.POINTER(Monitor), _I, _P, _I, _P, _L, _P, _P, C.POINTER(_I)], _I),
    "dnx_fdtd2d_adjoint_init": ([C.POINTER(Grid2D), _P, _P, C.POINTER(Monitor), _I, _I, _P, _P],
                                _I),
    "dnx_fdtd2d_adjoint_run": ([C.POINTER(Grid2D), _P, _P, _I, _I, _P, _P, C.POINTER(Source), _I,
                                C.POINTER(Monitor), _I, _P, _P, C.POINTER(_I)], _I),
}

EXPORTED = tuple(_SIGS)


def load(path=LIB_PATH):
    """Load the library and declare every signature (no GPU needed)."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"libdnx_fdtd.so not found at {path}: build it with `python -m diffnano_amd._build` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    v = lib.dnx_abi_version()
    if v != ABI_VERSION:
        raise RuntimeError(f"libdnx_fdtd ABI version {v}, expected {ABI_VERSION}")
    return lib


def lib():
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                _lib = load()
    return _lib


# Kernel-selection options (DNX_OPT_* in include/dnx_fdtd.h).  They pick
# kernel variants, never results.
OPTIONS = {"simple": 0, "one_step": 1, "adj2": 2, "fwd2_pf": 3, "adj2_occ": 4,
           "gemm_generic": 5, "gemm_pw_valu": 6, "gemm_small": 7, "gemm_big_reg": 8,
           "fdtd2d_per_step": 9, "fdtd2d_separate": 10,
           "fdtd2d_tb_steps": 11, "fdtd2d_tb_region": 12}


def get_option(name):
    v = C.c_int32()
    check(lib().dnx_get_option(OPTIONS[name], C.byref(v)))
    return v.value


def set_option(name, value):
    check(lib().dnx_set_option(OPTIONS[name], int(value)))


@contextlib.contextmanager
def options(**values):
    """Set kernel-selection options for the duration of a block, then
    restore the previous values."""
    saved = {k: get_option(k) for k in values}
    try:
        for k, v in values.items():
            set_option(k, v)
        yield
    finally:
        for k, v in saved.items():
            set_option(k, v)


def check(rc):
    if rc != 0:
        msg = lib().dnx_last_error()
        raise RuntimeError(f"libdnx_fdtd error {rc}: {msg.decode() if msg else '?'}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else t.data_ptr()


def stream_handle(device):
    return torch.cuda.current_stream(device).cuda_stream


def dtype_code(dtype):
    if dtype == torch.float32:
        return DNX_F32
    if dtype == torch.float64:
        return DNX_F64
    raise ValueError(f"compute dtype must be float32 or float64, got {dtype}")


def compute_device(device=None):
    """The HIP device the kernels run on.

    ``device`` is the caller's result device; when it is a GPU it is used,
    otherwise the current GPU is.  Raises when no GPU is visible: the
    product has no CPU path.
    """
    if not torch.cuda.is_available():
        raise RuntimeError("diffnano_amd runs its FDTD kernels on an AMD GPU (gfx950) and none is "
                           "visible; there is no CPU fallback")
    if device is not None:
        d = torch.device(device)
        if d.type == "cuda":
            return d if d.index is not None else torch.device("cuda", torch.cuda.current_device())
    return torch.device("cuda", torch.cuda.current_device())

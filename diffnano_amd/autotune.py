"""Tile-shape autotuner for the fused 3D kernels.

For a grid shape and dtype, times every tile shape (rows per work-group ty,
z planes per work-group zc in {8, 16, 32, 64}) of the one-step forward and
adjoint kernels (ty in {2, 4, 8}, run with the two-step path off) and, for
float32, of the two-step forward (ty in {2, 4}) and adjoint (ty in {1, 2,
4}) kernels, over the same number of steps.  Per direction it keeps the
fastest one-step shape and, when the fastest two-step shape beats it, that
shape for the two-step kernel; otherwise the two-step kernel is switched
off for the grid (tile height 0).  Results are applied
(``dnx_fdtd3d_set_tiles``) and cached in a JSON file keyed by device name,
dtype, kernel and grid.  A cache file that cannot be parsed is discarded
and the shape is re-tuned.  Tile shape changes speed only: every shape
gives bitwise-identical results.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import threading

import numpy as np
import torch

from . import _lib as L

TY_CHOICES = (2, 4, 8)
ZC_CHOICES = (8, 16, 32, 64)
TY2_CHOICES = {2: (2, 4), 3: (1, 2, 4)}  # two-step forward / adjoint
TY2_BATCH = (2, 4)                       # batched two-step kernels
NAMES = {0: "forward", 1: "adjoint", 2: "forward2", 3: "adjoint2"}
OFF = (0, 1)  # tile of a two-step kernel that is switched off
_lock = threading.Lock()


def _kernels(dtype):
    return (0, 1, 2, 3) if dtype == torch.float32 else (0, 1)


def _pick(res, dtype):
    """Best one-step shape per direction, and the best two-step shape where
    it is faster than that (else OFF).  res: kernel -> {(ty, zc): ms}."""
    out = {}
    for k in (0, 1):
        out[k] = min(res[k].items(), key=lambda kv: kv[1])
    if dtype == torch.float32:
        for k in (2, 3):
            best2 = min(res[k].items(), key=lambda kv: kv[1]) if res.get(k) else None
            out[k] = best2 if best2 and best2[1] < out[k - 2][1] else (OFF, out[k - 2][1])
    return out


def cache_path():
    return os.environ.get("DNX_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "diffnano_amd",
                                       "tiles.json"))


def _load_cache(path):
    try:
        with open(path) as f:
            data = json.load(f)
        if not isinstance(data, dict):
            raise ValueError("cache root is not an object")
        return data
    except FileNotFoundError:
        return {}
    except (ValueError, OSError):
        # Corrupt or unreadable cache: start over (and overwrite it on save).
        return {}


def _save_cache(path, data):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _key(dev, dtype, kernel, shape):
    name = torch.cuda.get_device_name(dev).replace(" ", "_")
    return f"{name}|{'f64' if dtype == torch.float64 else 'f32'}|{kernel}|" + "x".join(
        str(int(v)) for v in shape)


def get_tiles(dtype, kernel):
    ty, zc = C.c_int32(), C.c_int32()
    L.check(L.lib().dnx_fdtd3d_get_tiles(L.dtype_code(dtype), kernel, C.byref(ty), C.byref(zc)))
    return ty.value, zc.value


def set_tiles(dtype, kernel, ty, zc):
    L.check(L.lib().dnx_fdtd3d_set_tiles(L.dtype_code(dtype), kernel, int(ty), int(zc)))


def _time_configs(shape, dtype, dev, steps):
    """ms for `steps` steps of every candidate shape of every kernel (steps
    even, so the two-step kernels run pairs only)."""
    from .fdtd3d import FDTDSolver3D

    steps += steps & 1
    N = int(np.prod(shape))
    s = FDTDSolver3D(grid_shape=shape, dl=25.0, pml_layers=8, n_steps=steps, courant=0.4,
                     dtype=dtype)
    eng = s._engine(dev)
    rng = np.random.default_rng(0)
    eng.set_eps(torch.tensor(2.0 + 10 * rng.random(shape), device=dev))
    state = torch.randn(6 * N, device=dev, dtype=dtype) * 1e-3
    work = torch.empty_like(state)
    estore = torch.empty(steps * 3 * N, device=dev, dtype=dtype)
    S = torch.zeros(N, device=dev, dtype=dtype)
    lam = torch.randn(6 * N, device=dev, dtype=dtype) * 1e-3
    lw = torch.empty_like(lam)
    none = (None, 0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kernels = _kernels(dtype)
    results = {k: {} for k in kernels}
    saved = {k: get_tiles(dtype, k) for k in kernels}

    def cands(k):
        if k < 2:
            return [(ty, zc) for ty in TY_CHOICES for zc in ZC_CHOICES
                    if not (dtype == torch.float64 and ty == 8)]  # f64 ty 8 spills
        return [(ty, zc) for ty in TY2_CHOICES[k] for zc in ZC_CHOICES]

    try:
        for kernel in kernels:
            with L.options(one_step=1 if kernel < 2 else 0):
                for ty, zc in cands(kernel):
                    set_tiles(dtype, kernel, ty, zc)
                    best = float("inf")
                    for rep in range(3):
                        e0.record()
                        if kernel in (0, 2):
                            eng.run(state, 0, steps, none, none, work=work)
                        else:  # steps m = steps .. 1 (E^m from the store)
                            eng.adjoint_run(lam, steps + 1, 0, None, estore, none, none, S, lw)
                        e1.record()
                        torch.cuda.synchronize()
                        if rep:
                            best = min(best, e0.elapsed_time(e1))
                    results[kernel][(ty, zc)] = best
                set_tiles(dtype, kernel, *saved[kernel])
    finally:
        for k, (ty, zc) in saved.items():
            set_tiles(dtype, k, ty, zc)
    return results


def _store(cache, keys, picked, steps):
    for k, ((ty, zc), ms) in picked.items():
        cache[keys[NAMES[k]]] = {"ty": ty, "zc": zc, "ms_per_step": ms / steps}


def apply_tiles(dtype, out):
    """Apply an autotuner result (the kernels it names)."""
    for k in _kernels(dtype):
        if NAMES[k] in out:
            set_tiles(dtype, k, *out[NAMES[k]])


_apply = apply_tiles


def tune(shape, dtype=torch.float32, device=None, steps=8, force=False, apply=True):
    """Return {"forward": (ty, zc), "adjoint": ..., and for float32
    "forward2", "adjoint2" (two-step kernels; (0, 1) = off)} for this grid,
    from the cache or by measuring; apply it unless ``apply=False``."""
    dev = L.compute_device(device)
    shape = tuple(int(v) for v in shape)
    path = cache_path()
    names = [NAMES[k] for k in _kernels(dtype)]
    with _lock:
        cache = _load_cache(path)
        keys = {k: _key(dev, dtype, k, shape) for k in names}
        hit = all(k in cache and _valid(cache[k]) for k in keys.values())
        if force or not hit:
            res = _time_configs(shape, dtype, dev, steps)
            _store(cache, keys, _pick(res, dtype), steps)
            _save_cache(path, cache)
        out = {name: (cache[keys[name]]["ty"], cache[keys[name]]["zc"]) for name in names}
    if apply:
        _apply(dtype, out)
    return out


def tune_batch(shapes, dtype=torch.float32, device=None, steps=8, force=False, apply=True):
    """Tile shapes for a batched plan over these grids (measured on the
    batched kernels themselves, cached under the sorted shape list): the
    one-step shapes with the two-step path off, then the two-step forward
    shape (forward time) and, with it, the two-step adjoint shape (backward
    time), each kept only where it beats the one-step kernels."""
    from .batched import BatchedFDTD3D

    dev = L.compute_device(device)
    shapes = [tuple(int(v) for v in s) for s in shapes]
    steps += steps & 1
    tag = "batch:" + "+".join("x".join(map(str, s)) for s in sorted(shapes))
    path = cache_path()
    kernels = _kernels(dtype)
    names = [NAMES[k] for k in kernels]
    with _lock:
        cache = _load_cache(path)
        keys = {k: _key(dev, dtype, k, ()) + "|" + tag for k in names}
        hit = all(k in cache and _valid(cache[k]) for k in keys.values())
        if force or not hit:
            saved = {k: get_tiles(dtype, k) for k in kernels}
            srcs = [{"type": "gaussian_pulse", "pos": [s[0] // 2, s[1] // 2, s[2] // 2]}
                    for s in shapes]
            mons = [{"box": (0, s[0], 0, s[1], s[2] - 2, s[2] - 1)} for s in shapes]
            rng = np.random.default_rng(0)
            eps = [torch.tensor(2.0 + rng.random(s), device=dev) for s in shapes]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            times = {k: {} for k in kernels}

            def measure(one_step):
                with L.options(one_step=one_step):
                    b = BatchedFDTD3D(shapes, srcs, mons, n_steps=steps, checkpoint_every=steps,
                                      pml_layers=2, device=dev, dtype=dtype)
                b._set_eps(eps)
                best = [float("inf"), float("inf")]
                for rep in range(3):
                    e0.record()
                    b._run_forward(record=True)
                    e1.record()
                    torch.cuda.synchronize()
                    tf = e0.elapsed_time(e1)
                    e0.record()
                    b._run_backward([None] * len(shapes))
                    e1.record()
                    torch.cuda.synchronize()
                    if rep:
                        best = [min(best[0], tf), min(best[1], e0.elapsed_time(e1))]
                del b
                return best

            try:
                for ty in TY_CHOICES:
                    if dtype == torch.float64 and ty == 8:
                        continue
                    for zc in ZC_CHOICES:
                        for k in (0, 1):
                            set_tiles(dtype, k, ty, zc)
                        tf, tb = measure(1)
                        times[0][(ty, zc)], times[1][(ty, zc)] = tf, tb
                if dtype == torch.float32:
                    picked = _pick({0: times[0], 1: times[1]}, torch.float64)
                    for k in (0, 1):
                        set_tiles(dtype, k, *picked[k][0])
                    for k in (2, 3):
                        for ty in (TY2_BATCH if k == 2 else (2,)):
                            for zc in ZC_CHOICES:
                                set_tiles(dtype, k, ty, zc)
                                tf, tb = measure(0)
                                times[k][(ty, zc)] = tf if k == 2 else tb
                        best = min(times[k].items(), key=lambda kv: kv[1])[0]
                        set_tiles(dtype, k, *best)
            finally:
                for k, (ty, zc) in saved.items():
                    set_tiles(dtype, k, ty, zc)
            _store(cache, keys, _pick(times, dtype), steps)
            _save_cache(path, cache)
        out = {name: (cache[keys[name]]["ty"], cache[keys[name]]["zc"]) for name in names}
    if apply:
        _apply(dtype, out)
    return out


def _valid(entry):
    return (isinstance(entry, dict) and entry.get("ty") in (0, 1) + TY_CHOICES
            and isinstance(entry.get("zc"), int) and 1 <= entry["zc"] <= 4096)

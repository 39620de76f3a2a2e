"""Tile-shape autotuner for the fused 3D kernels.

For a grid shape and dtype, times the forward and the adjoint kernel over a
few steps for every tile shape (rows per work-group ty in {2, 4, 8}, z
planes per work-group zc in {8, 16, 32, 64}), keeps the fastest per kernel,
applies it (``dnx_fdtd3d_set_tiles``) and caches it in a JSON file keyed by
device name, dtype, kernel and grid.  A cache file that cannot be parsed is
discarded and the shape is re-tuned.  Tile shape changes speed only: every
shape gives bitwise-identical results.
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
_lock = threading.Lock()


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
    from .fdtd3d import FDTDSolver3D

    N = int(np.prod(shape))
    s = FDTDSolver3D(grid_shape=shape, dl=25.0, pml_layers=8, n_steps=steps, courant=0.4,
                     dtype=dtype)
    eng = s._engine(dev)
    rng = np.random.default_rng(0)
    eng.set_eps(torch.tensor(2.0 + 10 * rng.random(shape), device=dev))
    state = torch.randn(6 * N, device=dev, dtype=dtype) * 1e-3
    work = torch.empty_like(state)
    estore = torch.empty((steps - 1) * 3 * N, device=dev, dtype=dtype)
    S = torch.zeros(N, device=dev, dtype=dtype)
    lam = torch.randn(6 * N, device=dev, dtype=dtype) * 1e-3
    lw = torch.empty_like(lam)
    none = (None, 0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    results = {0: {}, 1: {}}
    saved = [get_tiles(dtype, k) for k in (0, 1)]
    try:
        for kernel in (0, 1):
            for ty in TY_CHOICES:
                if dtype == torch.float64 and ty == 8:
                    continue  # spills registers
                for zc in ZC_CHOICES:
                    set_tiles(dtype, kernel, ty, zc)
                    best = float("inf")
                    for rep in range(3):
                        e0.record()
                        if kernel == 0:
                            eng.run(state, 0, steps, none, none, work=work)
                        else:
                            eng.adjoint_run(lam, steps, 0, None, estore, none, none, S, lw)
                        e1.record()
                        torch.cuda.synchronize()
                        if rep:
                            best = min(best, e0.elapsed_time(e1))
                    results[kernel][(ty, zc)] = best
    finally:
        for k, (ty, zc) in enumerate(saved):
            set_tiles(dtype, k, ty, zc)
    return results


def tune(shape, dtype=torch.float32, device=None, steps=8, force=False, apply=True):
    """Return {"forward": (ty, zc), "adjoint": (ty, zc)} for this grid, from the
    cache or by measuring; apply it unless ``apply=False``."""
    dev = L.compute_device(device)
    shape = tuple(int(v) for v in shape)
    path = cache_path()
    with _lock:
        cache = _load_cache(path)
        keys = {k: _key(dev, dtype, k, shape) for k in ("forward", "adjoint")}
        hit = all(k in cache and _valid(cache[k]) for k in keys.values())
        if force or not hit:
            res = _time_configs(shape, dtype, dev, steps)
            for kernel, name in ((0, "forward"), (1, "adjoint")):
                (ty, zc), ms = min(res[kernel].items(), key=lambda kv: kv[1])
                cache[keys[name]] = {"ty": ty, "zc": zc, "ms_per_step": ms / steps}
            _save_cache(path, cache)
        out = {name: (cache[keys[name]]["ty"], cache[keys[name]]["zc"])
               for name in ("forward", "adjoint")}
    if apply:
        set_tiles(dtype, 0, *out["forward"])
        set_tiles(dtype, 1, *out["adjoint"])
    return out


def tune_batch(shapes, dtype=torch.float32, device=None, steps=8, force=False, apply=True):
    """Tile shapes for a batched plan over these grids (measured on the
    batched kernels themselves, cached under the sorted shape list)."""
    from .batched import BatchedFDTD3D

    dev = L.compute_device(device)
    shapes = [tuple(int(v) for v in s) for s in shapes]
    tag = "batch:" + "+".join("x".join(map(str, s)) for s in sorted(shapes))
    path = cache_path()
    with _lock:
        cache = _load_cache(path)
        keys = {k: _key(dev, dtype, k, ()) + "|" + tag for k in ("forward", "adjoint")}
        hit = all(k in cache and _valid(cache[k]) for k in keys.values())
        if force or not hit:
            saved = [get_tiles(dtype, k) for k in (0, 1)]
            srcs = [{"type": "gaussian_pulse", "pos": [s[0] // 2, s[1] // 2, s[2] // 2]}
                    for s in shapes]
            mons = [{"box": (0, s[0], 0, s[1], s[2] - 2, s[2] - 1)} for s in shapes]
            rng = np.random.default_rng(0)
            eps = [torch.tensor(2.0 + rng.random(s), device=dev) for s in shapes]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            times = {0: {}, 1: {}}
            try:
                for ty in TY_CHOICES:
                    if dtype == torch.float64 and ty == 8:
                        continue
                    for zc in ZC_CHOICES:
                        for k in (0, 1):
                            set_tiles(dtype, k, ty, zc)
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
                        times[0][(ty, zc)] = best[0]
                        times[1][(ty, zc)] = best[1]
                        del b
            finally:
                for k, (ty, zc) in enumerate(saved):
                    set_tiles(dtype, k, ty, zc)
            for kernel, name in ((0, "forward"), (1, "adjoint")):
                (ty, zc), ms = min(times[kernel].items(), key=lambda kv: kv[1])
                cache[keys[name]] = {"ty": ty, "zc": zc, "ms_per_step": ms / steps}
            _save_cache(path, cache)
        out = {name: (cache[keys[name]]["ty"], cache[keys[name]]["zc"])
               for name in ("forward", "adjoint")}
    if apply:
        set_tiles(dtype, 0, *out["forward"])
        set_tiles(dtype, 1, *out["adjoint"])
    return out


def _valid(entry):
    return (isinstance(entry, dict) and entry.get("ty") in TY_CHOICES
            and isinstance(entry.get("zc"), int) and 1 <= entry["zc"] <= 4096)

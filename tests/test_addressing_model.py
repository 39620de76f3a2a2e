"""CPU: a host model of the two-step 3D kernels' memory addressing.

``k3_fused2_fwd`` / ``k3_fused2_adj`` (and their batched forms, which run the
same tile bodies on each design) address every field array through a
buffer descriptor whose base is one z-plane of the array, a scalar
component offset and the lane's clamped in-plane byte offset:

    byte = plane * P * 4 + (clamp(y0 - 2 + r, 0, ny - 1) * W + xc) * 4 + c * N * 4

with xc = x for lanes inside the grid and W - 1 for the padding lanes of the
last wave, rows r = 0 .. TY + 3 of a tile at (y0, z0), and the planes the
z-march touches (stage A at p < nz, its next plane p + 1 clamped to p at the
top, stage B at q = p - 1 >= zs).  This model enumerates every tile, plane,
row, lane and component of every load and store of both kernels for the
test grids and the benchmark grids and checks each byte offset against its
array's extent and the 2^31-byte limit of the 32-bit offsets.

It also models the reach of the removed prefetching variant of the first
two-step adjoint (64-bit row pointers, E^m, S and E^{m-1} of the owned rows
loaded at the top of every plane iteration, before the barriers, instead of
inside the in-grid branch after them).  Its loads, at unclamped indices,
run past the arrays at the last plane of every z-chunk (p = z1 = nz, where
stage A does not run), before them at the first (q = zs - 1 = -1), and past
them for the padding lanes and rows of ragged grids: the out-of-bounds
reads that explain the fault seen with that variant.
"""

import numpy as np
import pytest

LIMIT = 1 << 31


def _edge(starts):
    """The first two and last two tile origins: every offset is monotone in
    the plane and the clamped row, so the extremes of every access lie in
    the tiles at the grid's edges (interior tiles sit between them)."""
    s = list(starts)
    return sorted(set(s[:2] + s[-2:]))


def _tiles(shape, ty, zc, every=False):
    nz, ny, nx = shape
    zs = range(0, nz, zc)
    ys = range(0, ny, ty)
    for z0 in (zs if every else _edge(zs)):
        for y0 in (ys if every else _edge(ys)):
            yield y0, z0


def _voff(shape, ty, y0, threads):
    """[R, lanes] in-plane byte offsets (clamped rows and columns)."""
    _, ny, W = shape
    rows = np.clip(y0 - 2 + np.arange(ty + 4), 0, ny - 1)
    x = np.arange(threads)
    xc = np.where(x < W, x, W - 1)
    return (rows[:, None] * W + xc[None, :]) * 4


def _planes(shape, zc, z0):
    """(stage-A planes p < nz, their next planes, stage-B planes q, zs)."""
    nz = shape[0]
    zs = max(z0 - 2, 0)
    z1 = min(nz, z0 + zc)
    a = [p for p in range(zs, z1 + 1) if p < nz]
    nxt = [p + 1 if p + 1 < nz else p for p in a]
    b = [p - 1 for p in range(zs, z1 + 1) if p - 1 >= zs]
    return a, nxt, b, zs


def _check(shape, ty, zc, kernel):
    """Every access of one kernel over the whole grid stays inside its array
    (bytes), and the descriptor's record count stays positive."""
    nz, ny, W = shape
    N, P = nz * ny * W, ny * W
    threads = -(-W // 64) * 64
    assert 6 * N * 4 < LIMIT, "the host keeps these grids on the one-step kernels"
    comps = {"in": 6, "out": 6, "ce": 1, "S": 1, "em": 3, "em2": 3, "es": 3, "es2": 3}
    worst = {}
    for y0, z0 in _tiles(shape, ty, zc, every=N <= 64 ** 3):
        vo = _voff(shape, ty, y0, threads)
        a, nxt, b, zs = _planes(shape, zc, z0)
        if kernel == "fwd":
            sites = [("in", [zs], range(0, 3)), ("in", nxt, range(0, 3)),
                     ("in", a, range(3, 6)), ("ce", a, [0]), ("es", a, range(3)),
                     ("out", b, range(6)), ("es2", b, range(3))]
        else:
            sites = [("ce", [zs], [0]), ("in", [zs], range(0, 3)), ("ce", nxt, [0]),
                     ("in", nxt, range(0, 3)), ("in", a, range(3, 6)), ("em", a, range(3)),
                     ("S", a, [0]), ("em2", b, range(3)), ("out", b, range(6)),
                     ("S", b, [0])]
        for arr, planes, cs in sites:
            for plane in planes:
                assert 0 <= plane < nz
                base = plane * P * 4
                records = comps[arr] * N * 4 - base
                assert records > 0 and base < LIMIT
                for c in cs:
                    off = vo + c * N * 4             # relative to the descriptor base
                    hi = int(off.max())
                    assert off.min() >= 0 and hi < records, (arr, plane, c)
                    assert base + hi < comps[arr] * N * 4 and base + hi < LIMIT
                    worst[arr] = max(worst.get(arr, 0), base + hi)
    return worst, N


GRIDS = [(37, 21, 70), (33, 19, 96), (17, 11, 66), (14, 11, 40), (24, 40, 64), (12, 14, 40),
         (20, 24, 96), (10, 12, 40), (96, 96, 96), (128, 128, 128), (96, 160, 160),
         (128, 128, 192), (256, 256, 256)]


@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("kernel,tys", [("fwd", (2, 4)), ("adj", (1, 2, 4))])
def test_two_step_kernel_accesses_stay_inside_their_arrays(shape, kernel, tys):
    for ty in tys:
        for zc in sorted({3, 5, 7, 8, 32, 64, shape[0]}):
            worst, N = _check(shape, ty, zc, kernel)
            # the last component of the last plane is reached, nothing past it
            assert worst["in" if kernel == "fwd" else "S"] >= 0


def _removed_prefetch_reach(shape, ty, zc):
    """Element index range of the removed variant's prefetches: E^m / S of
    the owned rows at every loop plane p in [zs, z1], E^{m-1} at q = p - 1,
    index p * P + y * W + x with y = y0 - 2 + r (r = 2 .. TY + 1) and every
    lane x of the work-group, no clamp."""
    nz, ny, W = shape
    P = ny * W
    threads = -(-W // 64) * 64
    lo, hi = 0, 0
    for y0, z0 in _tiles(shape, ty, zc):
        zs, z1 = max(z0 - 2, 0), min(nz, z0 + zc)
        ys = y0 + np.arange(ty)             # owned rows
        for p in range(zs, z1 + 1):
            for plane in (p, p - 1):
                idx = plane * P + ys[:, None] * W + np.arange(threads)[None, :]
                lo, hi = min(lo, int(idx.min())), max(hi, int(idx.max()))
    return lo, hi


@pytest.mark.parametrize("shape,ty,zc", [((17, 11, 66), 1, 5), ((17, 11, 66), 2, 7),
                                         ((33, 19, 96), 2, 3), ((24, 40, 64), 2, 8),
                                         ((256, 256, 256), 2, 64)])
def test_removed_prefetch_variant_reads_outside_the_arrays(shape, ty, zc):
    nz, ny, W = shape
    N = nz * ny * W
    lo, hi = _removed_prefetch_reach(shape, ty, zc)
    # S has N elements, each E component N: plane nz lies one plane past the
    # end, plane -1 one plane before the start
    assert hi >= N and lo < 0
    assert hi - N >= ny * W - W            # at least one plane minus a row past the end
    # the clamped scheme of the current kernel covers the same cells in bounds
    worst, N2 = _check(shape, ty, zc, "adj")
    assert worst["S"] < 4 * N2 and worst["em"] < 12 * N2

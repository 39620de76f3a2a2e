// Two-step (temporally blocked) fused 3D Yee kernels on gfx950.
//
// The one-step fused kernels move 52 bytes per cell-update (forward) and
// 72 (adjoint) and already run at the device's streaming bandwidth, so the
// remaining lever is to touch the state once per two steps.  A work-group
// owns TY full rows of a chunk of z-planes (one thread per x column, as in
// the one-step kernels) and marches z upward computing
//
//   stage A at plane p:   H1(p) from E0(p), E0(p+1), H0(p)       rows -2..TY
//                         E1(p) from H1(p), H1(p-1), E0(p)       rows -1..TY
//   stage B at plane p-1: H2(p-1) from E1(p-1), E1(p), H1(p-1)   rows -1..TY-1
//                         E2(p-1) from H2(p-1), H2(p-2), E1(p-1) rows  0..TY-1
//
// (rows relative to the tile).  The intermediate state (E1, H1) never
// leaves the chip: its halo rows are recomputed by both neighbouring tiles
// (bitwise the same arithmetic), its z-neighbours stay in registers, and
// x-neighbours go through LDS: four exchanges per plane, two barriers
// (E0 and E1 double-buffered by plane parity).  A chunk starting at z0 > 0
// starts its march at z0 - 2; values of the planes below z0 that depend on
// data the march has not seen are computed but never stored.
//
// Per sub-step the cell arithmetic is exactly that of the one-step kernels
// (fdtd3d_fused.hip), including the fused soft source and the DFT monitor
// accumulation, so the two-step path reproduces the one-step path bitwise.
#include <algorithm>

#include "fdtd3d_fused.h"

#include <cstdlib>

// No multiply-add contraction: the same rounding as the one-step kernels.
#pragma clang fp contract(off)

namespace dnx {

template <typename T>
__device__ __forceinline__ bool in_box(const Box& b, int z, int y, int x, int64_t& i) {
  const int dz = z - b.z0, dy = y - b.y0, dx = x - b.x0;
  if (dz < 0 || dz >= b.nz || dy < 0 || dy >= b.ny || dx < 0 || dx >= b.nx) return false;
  i = (int64_t(dz) * b.ny + dy) * b.nx + dx;
  return true;
}

// Stage the monitor components of E for sub-step `sub` (0, 1) into the
// compact buffer side = [2][ncomp][cells] (plain stores; launch_monitor_pair
// accumulates the DFT afterwards).
template <typename T>
__device__ __forceinline__ void mon_stage(const FusedMon& m, T* side, int sub, int z, int y,
                                          int x, const T (&e)[3]) {
  int64_t i;
  if (!in_box<T>(m.box, z, y, x, i)) return;
  const int64_t cells = m.box.cells();
  int ci = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!((m.mask >> c) & 1)) continue;
    side[(int64_t(sub) * m.ncomp + ci) * cells + i] = e[c];
    ++ci;
  }
}

// A box membership test split by axis for the tile loops: the x part once
// per thread, the y part once per row of the tile (rows r <-> y = yrow0 + r),
// the z part once per plane, so the per-cell cost is one AND.
template <int R>
struct RowBox {
  int dx;
  bool yin[R];
  int dy[R];
  __device__ void init(const Box& b, int x, int yrow0, bool enabled) {
    dx = x - b.x0;
    const bool xin = enabled && dx >= 0 && dx < b.nx;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      dy[r] = yrow0 + r - b.y0;
      yin[r] = xin && dy[r] >= 0 && dy[r] < b.ny;
    }
  }
  __device__ static bool zin(const Box& b, int z, int& dz) {
    dz = z - b.z0;
    return dz >= 0 && dz < b.nz;
  }
  __device__ int64_t index(const Box& b, int dz, int r) const {
    return (int64_t(dz) * b.ny + dy[r]) * b.nx + dx;
  }
};

// The same test with the per-row flags packed into one uniform word and the
// column test kept per lane (fewer scalar registers in the forward tile).
struct RowBits {
  int dx, dy0;    // lane column and tile row 0 relative to the box corner
  bool xin;       // the lane's column lies inside the box
  uint32_t yin;   // bit r: row r lies inside the box (uniform)
  __device__ void init(const Box& b, int x, int yrow0, int rows, bool enabled) {
    dx = x - b.x0;
    dy0 = yrow0 - b.y0;
    xin = enabled && dx >= 0 && dx < b.nx;
    yin = 0;
    for (int r = 0; r < rows; ++r)
      if (dy0 + r >= 0 && dy0 + r < b.ny) yin |= 1u << r;
  }
  __device__ bool row(int r) const { return xin && ((yin >> r) & 1u); }
  __device__ int64_t index(const Box& b, int dz, int r) const {
    return (int64_t(dz) * b.ny + dy0 + r) * b.nx + dx;
  }
};

// Branch-free access to the box-compact arrays (source profile, monitor
// adjoint terms) from the tiles: the byte offset of cell (z, y, x) of box b
// is lane + row + plane, each part replaced by kBoxOut when that coordinate
// lies outside the box.  kBoxOut is past the end of every such array (the
// host keeps them below 2^29 bytes) and three of them stay below 2^31, so a
// load for a cell outside the box reads 0 through the descriptor's range
// check, with no branch and no per-row lane mask.
constexpr int kBoxOut = 1 << 29;
__device__ __forceinline__ int box_lane(const Box& b, int x, bool on) {
  return (on && x >= b.x0 && x < b.x0 + b.nx) ? (x - b.x0) * 4 : kBoxOut;
}
__device__ __forceinline__ int box_row(const Box& b, int y) {
  return (y >= b.y0 && y < b.y0 + b.ny) ? (y - b.y0) * b.nx * 4 : kBoxOut;
}
__device__ __forceinline__ int box_plane(const Box& b, int z) {
  return (z >= b.z0 && z < b.z0 + b.nz) ? (z - b.z0) * b.ny * b.nx * 4 : kBoxOut;
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t box_rsrc(const void* p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, p ? int(bytes) : 0,
                                           0x00020000);
}
__device__ __forceinline__ float box_ld(const __amdgpu_buffer_rsrc_t& rs, int off, int soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off, soff, 0));
}

// e[comp] += w * prof[i] (the fused soft source, one component).
template <typename T>
__device__ __forceinline__ void add_src_at(int comp, T w, const T* prof, int64_t i, T (&e)[3]) {
  const T s = w * prof[i];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
    if (cc == comp) e[cc] += s;
}

// e[comp] += s (the fused soft source with its profile value prefetched).
template <typename T>
__device__ __forceinline__ void add_src_val(int comp, T s, T (&e)[3]) {
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
    if (cc == comp) e[cc] += s;
}

// Monitor staging at a known in-box cell index i (see mon_stage).
template <typename T>
__device__ __forceinline__ void mon_stage_at(const FusedMon& m, T* side, int sub, int64_t i,
                                             const T (&e)[3]) {
  const int64_t cells = m.box.cells();
  int ci = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!((m.mask >> c) & 1)) continue;
    side[(int64_t(sub) * m.ncomp + ci) * cells + i] = e[c];
    ++ci;
  }
}

// Soft source of step `step`: e[comp] += wave[step] * prof(z, y, x).
template <typename T>
__device__ __forceinline__ void add_source(const Fused3Args<T>& a, T w, int z, int y, int x,
                                           T (&e)[3]) {
  int64_t i;
  if (!in_box<T>(a.src_box, z, y, x, i)) return;
  const T s = w * a.src_prof[i];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
    if (cc == a.src_comp) e[cc] += s;
}

// A uniform value kept in a vector register.  The per-row profile factors
// are VALU operands on every plane; held in scalar registers they push the
// tile kernels past the scalar register file, and the compiler then parks
// scalars in vector lanes and reads them back with v_readlane every plane.
__device__ __forceinline__ float vreg(float v) {
  float r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(v));
  return r;
}

// x-neighbour read from an LDS row at a clamped lane, issued unconditionally
// and selected afterwards (a read under the lane condition becomes an
// exec-masked branch per read).
template <typename T>
__device__ __forceinline__ T lds_nb(bool on, const T* row, int xn) {
  const T v = row[xn];
  return on ? v : T(0);
}

// MON: stage the monitor samples of both sub-steps from inside the tile
// (the batched kernel).  The single-grid kernel leaves them to
// k_monitor_pair_fwd, which recomputes the intermediate sample after the
// launch: staging from inside the tile cost 41 us of a 242 us launch at
// 256^3 (one partial-line store per monitor cell and sub-step, whose
// completion the next loads' waits include), the recompute 24 us.  With
// the batch's many monitor planes the recompute costs more than the
// staging, so the batched kernel stages.
// FOLD: the y-edge tests folded into per-row ay factors (the single-grid
// kernel) or applied as selects where the factors are used (the batched
// kernel, whose per-design arguments leave no scalar registers for them:
// folded, it spills).  The forms differ at most in the sign of a zero.
// NT: non-temporal stores (the output state is not read again by this
// launch), for grids whose state does not fit the last-level cache; on
// cache-resident grids they cost the next launch its cache hits.
// LW: the LDS row pitch, a compile-time constant (immediate LDS offsets) at
// least the work-group width (every lane writes its column): 128 for grids
// up to 128 columns, 256 otherwise (the smaller pitch keeps the LDS of
// narrow grids from limiting the work-groups per CU).
// MONX: the DFT of monitor 0, an x-normal plane, inside the tile: the lane
// of the plane's column writes both samples of its owned cells to LDS
// during the march, and after it the work-group accumulates them (FP64, in
// the order of k_monitor_pair_fwd) into the monitor's accumulators; each
// cell belongs to exactly one tile.  The samples are the tile's own values.
template <typename T, int TY, bool PF, bool MON, bool FOLD = true, bool NT = false,
          bool MONX = false, int LW = FUSED_MAX_W>
__device__ __forceinline__ void fwd2_tile(const Fused2Args<T>& A2, const int y0, const int z0,
                                          T* sm) {
  const Fused3Args<T>& a = A2.a;
  const int W = a.nx;
  static_assert(LW % 64 == 0 && LW <= FUSED_MAX_W, "LDS row pitch");  // >= the work-group width
  constexpr int R = TY + 4;  // rows r = 0 .. TY+3  <->  y = y0 - 2 + r
  T* sA = sm;                           // E0 (Ey, Ez), rows 0..TY+2, by parity
  T* sB = sA + 4 * (TY + 3) * LW;        // H1 (Hy, Hz), rows 1..TY+2
  T* sC = sB + 2 * (TY + 2) * LW;        // E1 (Ey, Ez), rows 1..TY+1, by parity
  T* sD = sC + 4 * (TY + 1) * LW;        // H2 (Hy, Hz), rows 2..TY+1
  T* sM = sD + 2 * TY * LW;              // MONX: samples [zc][TY][2][3]
#define LA(p, c, r) sA[(((p) * 2 + (c)) * (TY + 3) + (r)) * LW]
#define LB(c, r) sB[((c) * (TY + 2) + (r) - 1) * LW]
#define LC(p, c, r) sC[(((p) * 2 + (c)) * (TY + 1) + (r) - 1) * LW]
#define LD(c, r) sD[((c) * TY + (r) - 2) * LW]

  const int z1 = min(a.nz, z0 + a.zc);
  const int x = threadIdx.x;
  const bool xv = x < W;
  const int xp1 = min(x + 1, W - 1), xm1 = max(x - 1, 0);  // clamped neighbour lanes
  const int64_t N = a.N, P = a.P;

  const T kh = a.kh;
  const T bxv = xv ? a.bx[x] : T(0);
  const T axv = xv ? a.ax[x] : T(0);
  const bool xf = xv && x >= 1, xl = x + 1 < W;

  T byr[R], ayr[R];
  // ay of row r for its y differences: 0 where the difference crosses the
  // grid edge (y + 1 = ny for ayl, y = 0 for ayf), so no per-row edge test
  // (a lane mask per row, kept live across the march) is needed
  T ayl[R], ayf[R];
  uint32_t rvm = 0;  // bit r: row r lies inside the grid
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int y = y0 - 2 + r;
    const bool in = y >= 0 && y < a.ny;
    if (in) rvm |= 1u << r;
    const T ay = in ? a.ay[y] : T(0);
    // (vector registers for the single-grid kernel; the batched kernel
    // keeps them scalar, its vector registers are full)
    byr[r] = FOLD ? vreg(in ? a.by[y] : T(0)) : (in ? a.by[y] : T(0));
    ayr[r] = FOLD ? vreg(ay) : ay;
    if constexpr (FOLD) {
      ayl[r] = y + 1 < a.ny ? ay : T(0);
      ayf[r] = y >= 1 ? ay : T(0);
    }
  }
  auto rv = [&](int r) { return ((rvm >> r) & 1u) != 0; };
  // the y-difference factors of row r (see FOLD)
  auto fl = [&](int r) -> T {
    if constexpr (FOLD) return ayl[r];
    else return (y0 - 2 + r + 1 < a.ny) ? ayr[r] : T(0);
  };
  auto ff = [&](int r) -> T {
    if constexpr (FOLD) return ayf[r];
    else return (y0 - 2 + r >= 1) ? ayr[r] : T(0);
  };
  // Loads use a clamped row and column so that every one is in bounds and
  // none needs a branch.  Values of rows or columns outside the grid are
  // never combined into a cell of the grid (every difference across the
  // grid edge is selected away), so they are not masked.  An access is a
  // uniform plane base plus the lane's unsigned 32-bit byte offset within
  // the plane; stores (rows and columns inside the grid only) use the same
  // offsets, where clamped and true coordinates agree.
  const int xc = xv ? x : W - 1;
  uint32_t voff[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    voff[r] = uint32_t((min(max(y0 - 2 + r, 0), a.ny - 1) * W + xc) * int(sizeof(T)));
  // Buffer accesses: the descriptor base is the plane of an array (uniform),
  // the component offset is scalar, the lane's in-plane offset the vector
  // part.  Every offset stays below 2^31 bytes (checked on the host).
  static_assert(sizeof(T) == 4, "the two-step forward kernel is float32");
  const int cbytes = int(N * int64_t(sizeof(T)));  // one component
  const int pbytes = int(P * int64_t(sizeof(T)));   // one plane
  // 32-bit byte arithmetic (every offset is below 2^31): fewer scalar
  // registers and no 64-bit multiplies per plane
  auto rsrc = [&](const T* base, int plane, int ncomp) {
    const int off = plane * pbytes;
    return __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char*>(reinterpret_cast<const char*>(base)) + off, 0, ncomp * cbytes - off,
        0x00020000);
  };
  auto ld = [&](const __amdgpu_buffer_rsrc_t& rs, int c, int r) -> T {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, int(voff[r]), c * cbytes, 0));
  };
  // Stores of lanes outside the grid go to an offset past every record count
  // (bit 31 set; in-grid offsets stay below 2^31) and are dropped by the
  // descriptor's range check, so no store needs a lane mask.
  const uint32_t lane_oob = xv ? 0u : 0x80000000u;
  auto st = [&](const __amdgpu_buffer_rsrc_t& rs, int c, int r, T v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rs, int(voff[r] | lane_oob),
                                          c * cbytes, NT ? 2 : 0);
  };

  const bool has_src = a.src_comp >= 0;
  const T w1 = has_src ? a.src_wave[A2.n] : T(0);
  const T w2 = has_src ? a.src_wave[A2.n + 1] : T(0);
  RowBits mbx;  // monitor box test (staging, MON only)
  // source profile: branch-free loads (0 outside the box; see box_lane)
  const int s_lane = box_lane(a.src_box, x, has_src && xv);
  const auto rsp = box_rsrc(has_src ? a.src_prof : nullptr, a.src_box.cells() * 4);
  RowBits sbx;  // the batched kernel's source test (FOLD = false)
  if constexpr (!FOLD) sbx.init(a.src_box, x, y0 - 2, R, has_src && xv);
  mbx.init(A2.mon[0].box, x, y0 - 2, R, MON && A2.n_mon > 0 && xv);
  // MONX: the owned cells of the monitor plane in this tile: planes
  // [mzlo, mzhi), tile rows in the bit mask mrows (uniform; 0 when the tile
  // holds none), the plane's column in lane mon_lane.  All three
  // components of both samples of a cell are staged, [zc][TY][2][3], so
  // that a staging store is one per-plane offset plus an immediate.
  int mzlo = 0, mzhi = 0;
  uint32_t mrows = 0;
  bool mon_lane = false;
  if constexpr (MONX) {
    const Box& mxb = A2.mon[0].box;
    mzlo = max(z0, mxb.z0);
    mzhi = min(z1, mxb.z0 + mxb.nz);
    const int mylo = max(y0, mxb.y0), myhi = min(min(y0 + TY, a.ny), mxb.y0 + mxb.ny);
#pragma unroll
    for (int r = 2; r <= TY + 1; ++r)
      if (mzlo < mzhi && y0 - 2 + r >= mylo && y0 - 2 + r < myhi) mrows |= 1u << r;
    mon_lane = x == mxb.x0;
  }
  auto mstage = [&](int sub, int z, int r, const T (&e)[3]) {
    T* cell = sM + (z - z0) * (TY * 6) + (r - 2) * 6 + sub * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) cell[c] = e[c];
  };

  // The march keeps two register sets of every plane-carried value and the
  // loop body runs twice per trip with their roles swapped (A: the older
  // plane, B: the newer), so no value is copied from one plane's set to the
  // next.  Index = tile row for e0 and h1, row - 1 for the others.
  T eA[3][R], eB[3][R];            // E0 at planes p (C) and p+1 (N)
  T h1A[3][TY + 3], h1B[3][TY + 3];  // H1 at planes p-1 (F) and p (N)
  T e1A[3][TY + 2], e1B[3][TY + 2];  // E1 at planes p-1 (Q) and p (P)
  T h2A[3][TY + 1], h2B[3][TY + 1];  // H2 at planes p-2 (P) and p-1 (N)
  T ceA[TY + 2], ceB[TY + 2];        // ce at planes p-1 (Q) and p (P)
  T spA[TY + 2], spB[TY + 2];        // source profile at planes p-1 (Q) and p (P)
  const int zs = z0 >= 2 ? z0 - 2 : 0;
  {
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) eA[c][r] = ld(rsrc(a.in, zs, 6), c, r);
    }
    {
#pragma unroll
      for (int r = 0; r <= TY + 2; ++r) {
        (&LA(zs & 1, 0, r))[x] = eA[1][r];
        (&LA(zs & 1, 1, r))[x] = eA[2][r];
      }
    }
#pragma unroll
    for (int r = 0; r < TY + 3; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) h1A[c][r] = T(0);
    }
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) {
      ceA[r] = spA[r] = T(0);
#pragma unroll
      for (int c = 0; c < 3; ++c) e1A[c][r] = T(0);
    }
#pragma unroll
    for (int r = 0; r < TY + 1; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) h2A[c][r] = T(0);
    }
  }
  // PF: the loads of stage A (E at p+1, H and ce at p) are issued one plane
  // ahead, right after stage A of the previous plane, so that they are in
  // flight during stage B and the barrier.
  T pfE[3][R], pfH[3][TY + 3], pfC[TY + 2];
  auto prefetch = [&](int pn) {
    const int pe = pn + 1 < a.nz ? pn + 1 : pn;
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) pfE[c][r] = ld(rsrc(a.in, pe, 6), c, r);
    }
#pragma unroll
    for (int r = 0; r <= TY + 2; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) pfH[c][r] = ld(rsrc(a.in, pn, 6), 3 + c, r);
    }
#pragma unroll
    for (int r = 1; r <= TY + 2; ++r) pfC[r - 1] = ld(rsrc(a.ce, pn, 1), 0, r);
  };
  if constexpr (PF) {
    if (zs < a.nz) prefetch(zs);
  }
  __syncthreads();

  // One plane of the march: reads the older set (e0C, h1F, e1Q, h2P, ceQ,
  // spQ), writes the newer (e0N, h1N, e1P, h2N, ceP, spP).
  auto plane = [&](const int p, T(&e0C)[3][R], T(&e0N)[3][R], T(&h1F)[3][TY + 3],
                   T(&h1N)[3][TY + 3], T(&e1Q)[3][TY + 2], T(&e1P)[3][TY + 2],
                   T(&h2P)[3][TY + 1], T(&h2N)[3][TY + 1], T(&ceQ)[TY + 2], T(&ceP)[TY + 2],
                   T(&spQ)[TY + 2], T(&spP)[TY + 2]) {
    const bool doA = p < a.nz;       // stage A at plane p
    const bool doB = p - 1 >= zs;    // stage B at plane q = p - 1
    const int q = p - 1;
    const bool zl = p + 1 < a.nz, zf = p >= 1;
    // ---------------- stage A, part 1: loads and H1(p)
    if (doA) {
      T h0[3][TY + 3];
      if constexpr (PF) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
          for (int c = 0; c < 3; ++c) e0N[c][r] = pfE[c][r];
        }
#pragma unroll
        for (int r = 0; r <= TY + 2; ++r) {
#pragma unroll
          for (int c = 0; c < 3; ++c) h0[c][r] = pfH[c][r];
        }
#pragma unroll
        for (int r = 1; r <= TY + 2; ++r) ceP[r - 1] = pfC[r - 1];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
          for (int c = 0; c < 3; ++c) e0N[c][r] = ld(rsrc(a.in, zl ? p + 1 : p, 6), c, r);
        }
#pragma unroll
        for (int r = 0; r <= TY + 2; ++r) {
#pragma unroll
          for (int c = 0; c < 3; ++c) h0[c][r] = ld(rsrc(a.in, p, 6), 3 + c, r);
        }
#pragma unroll
        for (int r = 1; r <= TY + 2; ++r) ceP[r - 1] = ld(rsrc(a.ce, p, 1), 0, r);
      }
      {
        if constexpr (FOLD) {
          const int sp = s_lane + box_plane(a.src_box, p);
#pragma unroll
          for (int r = 1; r <= TY + 2; ++r)
            spP[r - 1] = box_ld(rsp, sp + box_row(a.src_box, y0 - 2 + r), 0);
        } else {
          int dzs;
          const bool sz = RowBox<R>::zin(a.src_box, p, dzs);
#pragma unroll
          for (int r = 1; r <= TY + 2; ++r)
            spP[r - 1] = (sz && sbx.row(r)) ? a.src_prof[sbx.index(a.src_box, dzs, r)] : T(0);
        }
      }
      const int pb = p & 1;
#pragma unroll
      for (int r = 0; r <= TY + 2; ++r) {
        const int y = y0 - 2 + r;
        const T ex = e0C[0][r], ey = e0C[1][r], ez = e0C[2][r];
        const T ey_x1 = lds_nb<T>(xl, &LA(pb, 0, r), xp1);
        const T ez_x1 = lds_nb<T>(xl, &LA(pb, 1, r), xp1);
        const T dEz_dy = e0C[2][r + 1] - ez;
        const T dEx_dy = e0C[0][r + 1] - ex;
        const T dEy_dz = zl ? e0N[1][r] - ey : T(0);
        const T dEx_dz = zl ? e0N[0][r] - ex : T(0);
        const T dEz_dx = xl ? ez_x1 - ez : T(0);
        const T dEy_dx = xl ? ey_x1 - ey : T(0);
        h1N[0][r] = byr[r] * h0[0][r] - kh * (fl(r) * dEz_dy - axv * dEy_dz);
        h1N[1][r] = bxv * h0[1][r] - kh * (axv * dEx_dz - ayr[r] * dEz_dx);
        h1N[2][r] = bxv * h0[2][r] - kh * (axv * dEy_dx - fl(r) * dEx_dy);
      }
      {
#pragma unroll
        for (int r = 1; r <= TY + 2; ++r) {
          (&LB(0, r))[x] = h1N[1][r];
          (&LB(1, r))[x] = h1N[2][r];
        }
#pragma unroll
        for (int r = 0; r <= TY + 2; ++r) {
          (&LA(pb ^ 1, 0, r))[x] = e0N[1][r];
          (&LA(pb ^ 1, 1, r))[x] = e0N[2][r];
        }
      }
    }
    __syncthreads();
    // ---------------- stage A, part 2: E1(p) (all rows the tile needs)
    if (doA) {
      const bool own = p >= z0 && p < z1;
      int dzmA;
      const bool mzA = MON && RowBox<R>::zin(A2.mon[0].box, p, dzmA);
      const uint32_t mxA = (MONX && p >= mzlo && p < mzhi) ? mrows : 0u;  // (uniform)
#pragma unroll
      for (int r = 1; r <= TY + 2; ++r) {
        const int y = y0 - 2 + r;
        const T hx = h1N[0][r], hy = h1N[1][r], hz = h1N[2][r];
        const T hy_x0 = lds_nb<T>(xf, &LB(0, r), xm1);
        const T hz_x0 = lds_nb<T>(xf, &LB(1, r), xm1);
        const T dHz_dy = hz - h1N[2][r - 1];
        const T dHx_dy = hx - h1N[0][r - 1];
        const T dHy_dz = zf ? hy - h1F[1][r] : T(0);
        const T dHx_dz = zf ? hx - h1F[0][r] : T(0);
        const T dHz_dx = xf ? hz - hz_x0 : T(0);
        const T dHy_dx = xf ? hy - hy_x0 : T(0);
        const T c = ceP[r - 1];
        const T bxy = bxv * byr[r];
        T e[3];
        e[0] = bxy * e0C[0][r] + c * (ff(r) * dHz_dy - axv * dHy_dz);
        e[1] = bxy * e0C[1][r] + c * (axv * dHx_dz - ayr[r] * dHz_dx);
        e[2] = bxy * e0C[2][r] + c * (axv * dHy_dx - ff(r) * dHx_dy);
        add_src_val<T>(a.src_comp, w1 * spP[r - 1], e);  // spP is 0 outside the box
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) e1P[cc][r - 1] = e[cc];
        if (own && r >= 2 && r <= TY + 1 && rv(r)) {  // (uniform; st drops padding lanes)
          if (a.es) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) st(rsrc(a.es, p, 3), cc, r, e[cc]);
          }
          if (mzA && mbx.row(r))  // (the row test includes the lane's)
            mon_stage_at<T>(A2.mon[0], A2.side[0], 0, mbx.index(A2.mon[0].box, dzmA, r), e);
          if constexpr (MONX) {
            if ((mxA >> r) & 1u) {
              if (mon_lane) mstage(0, p, r, e);
            }
          }
        }
      }
    }
    if constexpr (PF) {
      if (p + 1 < a.nz) prefetch(p + 1);
    }
    // ---------------- stage B, part 1: H2(q)
    const bool zlq = q + 1 < a.nz, zfq = q >= 1;
    if (doB) {
      const int qb = q & 1;
#pragma unroll
      for (int r = 1; r <= TY + 1; ++r) {
        const int y = y0 - 2 + r;
        const T ex = e1Q[0][r - 1], ey = e1Q[1][r - 1], ez = e1Q[2][r - 1];
        const T ey_x1 = lds_nb<T>(xl, &LC(qb, 0, r), xp1);
        const T ez_x1 = lds_nb<T>(xl, &LC(qb, 1, r), xp1);
        const T dEz_dy = e1Q[2][r] - ez;
        const T dEx_dy = e1Q[0][r] - ex;
        const T dEy_dz = zlq ? e1P[1][r - 1] - ey : T(0);
        const T dEx_dz = zlq ? e1P[0][r - 1] - ex : T(0);
        const T dEz_dx = xl ? ez_x1 - ez : T(0);
        const T dEy_dx = xl ? ey_x1 - ey : T(0);
        h2N[0][r - 1] = byr[r] * h1F[0][r] - kh * (fl(r) * dEz_dy - axv * dEy_dz);
        h2N[1][r - 1] = bxv * h1F[1][r] - kh * (axv * dEx_dz - ayr[r] * dEz_dx);
        h2N[2][r - 1] = bxv * h1F[2][r] - kh * (axv * dEy_dx - fl(r) * dEx_dy);
      }
      {
#pragma unroll
        for (int r = 2; r <= TY + 1; ++r) {
          (&LD(0, r))[x] = h2N[1][r - 1];
          (&LD(1, r))[x] = h2N[2][r - 1];
        }
      }
    }
    if (doA) {
#pragma unroll
      for (int r = 1; r <= TY + 1; ++r) {
        (&LC(p & 1, 0, r))[x] = e1P[1][r - 1];
        (&LC(p & 1, 1, r))[x] = e1P[2][r - 1];
      }
    }
    __syncthreads();
    // ---------------- stage B, part 2: E2(q), store
    if (doB) {
      const bool own = q >= z0;
      int dzmB;
      const bool mzB = MON && RowBox<R>::zin(A2.mon[0].box, q, dzmB);
      const uint32_t mxB = (MONX && q >= mzlo && q < mzhi) ? mrows : 0u;  // (uniform)
#pragma unroll
      for (int r = 2; r <= TY + 1; ++r) {
        const int y = y0 - 2 + r;
        const T hx = h2N[0][r - 1], hy = h2N[1][r - 1], hz = h2N[2][r - 1];
        const T hy_x0 = lds_nb<T>(xf, &LD(0, r), xm1);
        const T hz_x0 = lds_nb<T>(xf, &LD(1, r), xm1);
        const T dHz_dy = hz - h2N[2][r - 2];
        const T dHx_dy = hx - h2N[0][r - 2];
        const T dHy_dz = zfq ? hy - h2P[1][r - 1] : T(0);
        const T dHx_dz = zfq ? hx - h2P[0][r - 1] : T(0);
        const T dHz_dx = xf ? hz - hz_x0 : T(0);
        const T dHy_dx = xf ? hy - hy_x0 : T(0);
        const T c = ceQ[r - 1];
        const T bxy = bxv * byr[r];
        T e[3];
        e[0] = bxy * e1Q[0][r - 1] + c * (ff(r) * dHz_dy - axv * dHy_dz);
        e[1] = bxy * e1Q[1][r - 1] + c * (axv * dHx_dz - ayr[r] * dHz_dx);
        e[2] = bxy * e1Q[2][r - 1] + c * (axv * dHy_dx - ff(r) * dHx_dy);
        add_src_val<T>(a.src_comp, w2 * spQ[r - 1], e);
        if (own && rv(r)) {  // (uniform; st drops padding lanes)
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            st(rsrc(a.out, q, 6), cc, r, e[cc]);
            st(rsrc(a.out, q, 6), 3 + cc, r, h2N[cc][r - 1]);
          }
          if (A2.es2) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) st(rsrc(A2.es2, q, 3), cc, r, e[cc]);
          }
          if (mzB && mbx.row(r))
            mon_stage_at<T>(A2.mon[0], A2.side[0], 1, mbx.index(A2.mon[0].box, dzmB, r), e);
          if constexpr (MONX) {
            if ((mxB >> r) & 1u) {
              if (mon_lane) mstage(1, q, r, e);
            }
          }
        }
      }
    } else {
      // (first plane of the march: H2 of plane zs - 1 reads as zero next)
#pragma unroll
      for (int r = 0; r < TY + 1; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) h2N[c][r] = T(0);
      }
    }
  };
  int p = zs;
#pragma unroll 1
  for (; p + 1 <= z1; p += 2) {
    plane(p, eA, eB, h1A, h1B, e1A, e1B, h2A, h2B, ceA, ceB, spA, spB);
    plane(p + 1, eB, eA, h1B, h1A, e1B, e1A, h2B, h2A, ceB, ceA, spB, spA);
  }
  if (p <= z1) plane(p, eA, eB, h1A, h1B, e1A, e1B, h2A, h2B, ceA, ceB, spA, spB);
  if constexpr (MONX) {
    if (mrows != 0u) {  // (uniform)
      __syncthreads();
      // (the monitor's parameters are read here, not kept live through the
      // march: volatile reads are neither hoisted nor merged with the ones
      // above)
      auto vget = [](const auto& v) {
        return *reinterpret_cast<const volatile std::remove_reference_t<decltype(v)>*>(&v);
      };
      const FusedMon& mmv = A2.mon[0];
      const Box mxb{vget(mmv.box.z0), vget(mmv.box.y0), vget(mmv.box.x0),
                    vget(mmv.box.nz), vget(mmv.box.ny), vget(mmv.box.nx)};
      const int mmask = vget(mmv.mask), ncomp = vget(mmv.ncomp), nf = vget(A2.mnf);
      const double* ph = vget(A2.mph) + int64_t(A2.n) * nf * 2;
      double* const macc = vget(A2.macc);
      const int mylo = max(y0, mxb.y0), myhi = min(min(y0 + TY, a.ny), mxb.y0 + mxb.ny);
      const int nzm = mzhi - mzlo, nym = myhi - mylo;
      const int items = nzm * nym * ncomp * nf;
      const int64_t cells = mxb.cells();
      for (int t = int(threadIdx.x); t < items; t += int(blockDim.x)) {
        int u = t;
        const int dy = u % nym;
        u /= nym;
        const int dz = u % nzm;
        u /= nzm;
        const int ci = u % ncomp, f = u / ncomp;
        int c = 0;  // the ci-th monitored component
        for (int k = 0, seen = 0; k < 3; ++k)
          if ((mmask >> k) & 1) {
            if (seen == ci) c = k;
            ++seen;
          }
        const int z = mzlo + dz, y = mylo + dy;
        const T* cell = sM + (z - z0) * (TY * 6) + (y - y0) * 6;
        const double v1 = double(cell[c]), v2 = double(cell[3 + c]);
        const int64_t i = (int64_t(z - mxb.z0) * mxb.ny + (y - mxb.y0)) * mxb.nx;
        double* qa = macc + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
        double re = qa[0], im = qa[1];
        re += v1 * ph[2 * f];
        im -= v1 * ph[2 * f + 1];
        re += v2 * ph[2 * nf + 2 * f];
        im -= v2 * ph[2 * nf + 2 * f + 1];
        qa[0] = re;
        qa[1] = im;
      }
    }
  }
#undef LA
#undef LB
#undef LC
#undef LD
}

template <int TY>
__device__ inline void block_tile2(int nblocks, int ty_tiles, int zc, int& y0, int& z0,
                                   bool& live) {
  const int nb = gridDim.x;  // multiple of 8
  const int b = blockIdx.x;
  const int logical = (b & 7) * (nb >> 3) + (b >> 3);
  live = logical < nblocks;
  y0 = (logical % ty_tiles) * TY;
  z0 = (logical / ty_tiles) * zc;
}

template <typename T, int TY, bool PF, bool NT, bool MONX, int LW>
__global__ __launch_bounds__(256, 2) void k3_fused2_fwd(Fused2Args<T> a) {
  extern __shared__ unsigned char smem_raw[];
  int y0, z0;
  bool live;
  block_tile2<TY>(a.a.nblocks, a.a.ty_tiles, a.a.zc, y0, z0, live);
  if (!live) return;
  fwd2_tile<T, TY, PF, false, true, NT, MONX, LW>(a, y0, z0, reinterpret_cast<T*>(smem_raw));
}

// MONX staging space: both samples of the 3 components per owned cell of
// an x-normal plane (zc planes x TY rows per tile), at most 8 KB.
constexpr int kMonxBytes = 8192;
bool monitor_in_tile(const dnx_monitor& m, int ty, int zc) {
  const Box b = make_box(m.box);
  return m.n_freq > 0 && b.nx == 1 && (m.comp_mask & 7) != 0 &&
         int64_t(2) * 3 * zc * ty * int64_t(sizeof(float)) <= kMonxBytes &&
         b.cells() < (int64_t(1) << 31);
}

// Whether a grid's stores bypass the cache: one state (6 N floats) above
// half of the 256 MB last-level cache.
inline bool stream_stores(int64_t N) { return N * 6 * 4 > (int64_t(128) << 20); }

template <int TY>
constexpr int fwd2_lds_elems() {
  return 4 * (TY + 3) + 2 * (TY + 2) + 4 * (TY + 1) + 2 * TY;
}

template <typename T, int TY, bool PF, bool NT, bool MONX, int LW>
static int launch_fwd2_ty(const Fused2Args<T>& a, hipStream_t s) {
  const int threads = ((a.a.nx + 63) / 64) * 64;
  const int nb = ((a.a.nblocks + 7) / 8) * 8;
  DNX_REQUIRE(threads <= LW, "fused2: LDS pitch %d below the work-group width %d", LW, threads);
  const size_t smem = sizeof(T) * size_t(fwd2_lds_elems<TY>()) * LW +
                      (MONX ? sizeof(T) * 2 * 3 * a.a.zc * TY : 0);
  static bool attr_set = false;
  if (!attr_set) {
    DNX_CHECK_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k3_fused2_fwd<T, TY, PF, NT, MONX, LW>),
        hipFuncAttributeMaxDynamicSharedMemorySize,
        int(sizeof(T) * fwd2_lds_elems<TY>() * LW + (MONX ? kMonxBytes : 0))));
    attr_set = true;
  }
  hipLaunchKernelGGL((k3_fused2_fwd<T, TY, PF, NT, MONX, LW>), dim3(nb), dim3(threads), smem, s,
                     a);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

// ---- DFT monitor of a forward pair without staging --------------------
//
// After a two-step forward launch (state n in a.in, n + 2 in a.out), the
// sample after step n (E^{n+1}) of each monitor cell is recomputed from
// a.in with the arithmetic of fwd2_tile's stage A (H1 at the cell and at
// its three lower neighbours, then E1 with the fused source of step n), and
// the sample after step n + 1 is read from a.out; both are accumulated in
// the order of k_monitor_pair.  The recomputed values are the tile's, up to
// the sign of a zero where a difference is cut at the lower y edge.
template <typename T>
struct HVec {
  T x, y, z;
};

template <typename T>
__device__ HVec<T> h1_cell(const Fused3Args<T>& a, int z, int y, int x) {
  const T* in = a.in;
  const int64_t N = a.N, P = a.P;
  const int nx = a.nx;
  auto at = [&](int c, int zz, int yy, int xx) {
    return in[int64_t(c) * N + int64_t(zz) * P + int64_t(yy) * nx + xx];
  };
  const bool zl = z + 1 < a.nz, xl = x + 1 < nx;
  // clamped neighbours (every load in bounds, no branch); a difference
  // across the grid edge is selected away afterwards
  const int yc = min(y + 1, a.ny - 1), zc = min(z + 1, a.nz - 1), xc = min(x + 1, nx - 1);
  const T ex = at(0, z, y, x), ey = at(1, z, y, x), ez = at(2, z, y, x);
  const T ey_x1 = at(1, z, y, xc), ez_x1 = at(2, z, y, xc);
  const T ey_z1 = at(1, zc, y, x), ex_z1 = at(0, zc, y, x);
  const T dEz_dy = at(2, z, yc, x) - ez;
  const T dEx_dy = at(0, z, yc, x) - ex;
  const T dEy_dz = zl ? ey_z1 - ey : T(0);
  const T dEx_dz = zl ? ex_z1 - ex : T(0);
  const T dEz_dx = xl ? ez_x1 - ez : T(0);
  const T dEy_dx = xl ? ey_x1 - ey : T(0);
  const T ay = a.ay[y], ayl = y + 1 < a.ny ? ay : T(0);
  const T bx = a.bx[x], ax = a.ax[x], kh = a.kh;
  HVec<T> h;
  h.x = a.by[y] * at(3, z, y, x) - kh * (ayl * dEz_dy - ax * dEy_dz);
  h.y = bx * at(4, z, y, x) - kh * (ax * dEx_dz - ay * dEz_dx);
  h.z = bx * at(5, z, y, x) - kh * (ax * dEy_dx - ayl * dEx_dy);
  return h;
}

// MASK: the monitored components (compile-time, so that the loads and
// arithmetic of the components not monitored are dropped).
template <typename T, int MASK>
__device__ void monitor_pair_cell(const Fused3Args<T>& a, int n, const Box& mb, int nf,
                                  const double* __restrict__ ph, double* __restrict__ acc,
                                  int64_t i) {
  constexpr int comp_mask = MASK;
  const int64_t cells = mb.cells();
  // 32-bit decode: the launcher admits boxes of fewer than 2^31 cells
  const uint32_t ui = uint32_t(i), bnx = uint32_t(mb.nx), bny = uint32_t(mb.ny);
  const uint32_t iy = ui / bnx;
  const int dx = int(ui - iy * bnx), dy = int(iy % bny), dz = int(iy / bny);
  const int x = mb.x0 + dx, y = mb.y0 + dy, z = mb.z0 + dz;
  const int64_t c0 = int64_t(z) * a.P + int64_t(y) * a.nx + x;
  const HVec<T> h = h1_cell(a, z, y, x);
  const bool zf = z >= 1, xf = x >= 1;
  // lower neighbours at clamped indices: at the grid edge the clamp lands
  // on the cell itself (= h), with no branch around the loads
  const HVec<T> hym = h1_cell(a, z, max(y - 1, 0), x);
  const HVec<T> hzm = h1_cell(a, max(z - 1, 0), y, x);
  const HVec<T> hxm = h1_cell(a, z, y, max(x - 1, 0));
  const T hy_x0 = xf ? hxm.y : T(0), hz_x0 = xf ? hxm.z : T(0);
  const T dHz_dy = h.z - hym.z;
  const T dHx_dy = h.x - hym.x;
  const T dHy_dz = zf ? h.y - hzm.y : T(0);
  const T dHx_dz = zf ? h.x - hzm.x : T(0);
  const T dHz_dx = xf ? h.z - hz_x0 : T(0);
  const T dHy_dx = xf ? h.y - hy_x0 : T(0);
  const T c = a.ce[c0];
  const T ay = a.ay[y], ayf = y >= 1 ? ay : T(0);
  const T ax = a.ax[x];
  const T bxy = a.bx[x] * a.by[y];
  const T* in = a.in;
  T e[3];
  e[0] = bxy * in[c0] + c * (ayf * dHz_dy - ax * dHy_dz);
  e[1] = bxy * in[a.N + c0] + c * (ax * dHx_dz - ay * dHz_dx);
  e[2] = bxy * in[2 * a.N + c0] + c * (ax * dHy_dx - ayf * dHx_dy);
  int64_t si;
  if (a.src_comp >= 0 && in_box<T>(a.src_box, z, y, x, si))
    add_src_at<T>(a.src_comp, a.src_wave[n], a.src_prof, si, e);
  int ci = 0;
  for (int comp = 0; comp < 3; ++comp) {
    if (!((comp_mask >> comp) & 1)) continue;
    const double v1 = double(e[comp]);
    const double v2 = double(a.out[int64_t(comp) * a.N + c0]);
    constexpr int ncomp = ((MASK >> 0) & 1) + ((MASK >> 1) & 1) + ((MASK >> 2) & 1);
    for (int f = 0; f < nf; ++f) {
      double* q = acc + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
      q[0] += v1 * ph[2 * f];
      q[1] -= v1 * ph[2 * f + 1];
      q[0] += v2 * ph[2 * nf + 2 * f];
      q[1] -= v2 * ph[2 * nf + 2 * f + 1];
    }
    ++ci;
  }
}

template <typename T, int MASK>
__global__ void k_monitor_pair_fwd(Fused3Args<T> a, int n, Box mb, int nf,
                                   const double* __restrict__ ph, double* __restrict__ acc) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < mb.cells()) monitor_pair_cell<T, MASK>(a, n, mb, nf, ph, acc, i);
}

template <typename T>
int launch_monitor_pair_fwd(const Fused2Args<T>& f, const dnx_monitor& m, int n, hipStream_t s) {
  const Box b = make_box(m.box);
  if (b.cells() >= (int64_t(1) << 31)) {
    set_error("monitor: %lld cells in one box (at most 2^31 - 1)", (long long)b.cells());
    return DNX_EINVAL;
  }
  const dim3 grid(unsigned((b.cells() + 255) / 256)), block(256);
  const double* ph = m.phase + int64_t(n) * m.n_freq * 2;
  switch (m.comp_mask) {
#define DNX_MPF(K)                                                                           \
  case K:                                                                                  \
    hipLaunchKernelGGL((k_monitor_pair_fwd<T, K>), grid, block, 0, s, f.a, n, b, m.n_freq, ph, \
                       m.acc);                                                             \
    break;
    DNX_MPF(1) DNX_MPF(2) DNX_MPF(3) DNX_MPF(4) DNX_MPF(5) DNX_MPF(6) DNX_MPF(7)
#undef DNX_MPF
    default:
      set_error("monitor: component mask %d", m.comp_mask);
      return DNX_EINVAL;
  }
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

// Variant choice: the prefetch variant (opt-in) at pitch 256 with cached
// stores; otherwise pitch 128 for grids up to 128 columns (whose state is
// cache-resident, cached stores), pitch 256 with non-temporal stores when
// the state exceeds the cache; the in-tile monitor where the caller set it.
template <typename T, int TY>
static int launch_fwd2_pick(const Fused2Args<T>& a, hipStream_t s) {
  const bool monx = a.mnf > 0;
  if (option(DNX_OPT_FDTD3D_FWD2_PF) != 0) {
    DNX_REQUIRE(!monx, "fused2: the prefetch variant has no in-tile monitor");
    return launch_fwd2_ty<T, TY, true, false, false, 256>(a, s);
  }
  if (a.a.nx <= 128)
    return monx ? launch_fwd2_ty<T, TY, false, false, true, 128>(a, s)
                : launch_fwd2_ty<T, TY, false, false, false, 128>(a, s);
  if (stream_stores(a.a.N))
    return monx ? launch_fwd2_ty<T, TY, false, true, true, 256>(a, s)
                : launch_fwd2_ty<T, TY, false, true, false, 256>(a, s);
  return monx ? launch_fwd2_ty<T, TY, false, false, true, 256>(a, s)
              : launch_fwd2_ty<T, TY, false, false, false, 256>(a, s);
}

template <typename T>
int launch_fused2_fwd(const Fused2Args<T>& a, int ty, hipStream_t s) {
  DNX_REQUIRE(a.mnf == 0 || a.n_mon == 1, "fused2: in-tile monitor needs exactly one monitor");
  if (ty == 4) return launch_fwd2_pick<T, 4>(a, s);
  if (ty == 2) return launch_fwd2_pick<T, 2>(a, s);
  set_error("fused2: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template int launch_fused2_fwd<float>(const Fused2Args<float>&, int, hipStream_t);
template int launch_monitor_pair_fwd<float>(const Fused2Args<float>&, const dnx_monitor&, int,
                                            hipStream_t);

// ---------------------------------------------------------------------------
// Adjoint, two steps: lam^{m+1} -> lam^m -> lam^{m-1}.  Same march and row
// sets as the forward, with u = ce lamE in the role of E, muH in the role
// of H.  Per sub-step the arithmetic is that of adj_tile followed by the
// monitor adjoint (k_monitor_adj): lamE^m += Re(g e^{i phi}) over the
// monitor box (on every row the tile computes, halos included, since the
// next sub-step reads them) and S += E^m . d on owned cells.  S is read in
// stage A and written in stage B, accumulating in the one-step order.
// ---------------------------------------------------------------------------

// Prefetch of the precomputed adjoint source d = [ncomp][cells] of monitor
// m at cell (z, y, x): dv[c] for the components in the mask (0 elsewhere);
// returns whether the cell is in the box.
template <typename T>
__device__ __forceinline__ bool mon_fetch(const FusedMon& m, const T* d, int z, int y, int x,
                                          T (&dv)[3]) {
  int64_t i;
  dv[0] = dv[1] = dv[2] = T(0);
  if (!in_box<T>(m.box, z, y, x, i)) return false;
  const int64_t cells = m.box.cells();
  int ci = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!((m.mask >> c) & 1)) continue;
    dv[c] = d[int64_t(ci) * cells + i];
    ++ci;
  }
  return true;
}

// k_monitor_adj with d prefetched: lam_c += d_c; returns sum_c em_c d_c in
// the same order.
template <typename T>
__device__ __forceinline__ T mon_apply(int mask, const T (&dv)[3], T (&lam)[3],
                                       const T (&em)[3]) {
  T sacc = T(0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!((mask >> c) & 1)) continue;
    lam[c] += dv[c];
    sacc += em[c] * dv[c];
  }
  return sacc;
}

// FOLD: the y-edge tests folded into per-row ay factors (fewer scalar
// registers; the single-grid kernel), or as selects on each difference (the
// batched kernel, whose per-design arguments leave no room for the factors).
// The two forms differ at most in the sign of a zero.
template <typename T, int TY, bool FOLD, bool NT = false, int LW = FUSED_MAX_W>
__device__ __forceinline__ void adj2_tile(const Fused2Args<T>& A2, const int y0, const int z0,
                                          T* sm) {
  const Fused3Args<T>& a = A2.a;
  const int W = a.nx;
  static_assert(LW % 64 == 0 && LW <= FUSED_MAX_W, "LDS row pitch");  // >= the work-group width
  constexpr int R = TY + 4;  // rows r = 0 .. TY+3  <->  y = y0 - 2 + r
  T* sA = sm;                           // u0 (uy, uz), rows 0..TY+2, by parity
  T* sB = sA + 4 * (TY + 3) * LW;        // mu1 (y, z), rows 1..TY+2
  T* sC = sB + 2 * (TY + 2) * LW;        // u1 (uy, uz), rows 1..TY+1, by parity
  T* sD = sC + 4 * (TY + 1) * LW;        // mu2 (y, z), rows 2..TY+1
#define LA(p, c, r) sA[(((p) * 2 + (c)) * (TY + 3) + (r)) * LW]
#define LB(c, r) sB[((c) * (TY + 2) + (r) - 1) * LW]
#define LC(p, c, r) sC[(((p) * 2 + (c)) * (TY + 1) + (r) - 1) * LW]
#define LD(c, r) sD[((c) * TY + (r) - 2) * LW]

  const int z1 = min(a.nz, z0 + a.zc);
  const int x = threadIdx.x;
  const bool xv = x < W;
  const int xp1 = min(x + 1, W - 1), xm1 = max(x - 1, 0);  // clamped neighbour lanes
  const int64_t N = a.N, P = a.P;

  const T kh = a.kh;
  const T bxv = xv ? a.bx[x] : T(0);
  const T axv = xv ? a.ax[x] : T(0);
  const T axp = (x + 1 < W) ? a.ax[x + 1] : T(0);
  const T axm = (xv && x >= 1) ? a.ax[x - 1] : T(0);
  const bool xf = xv && x >= 1, xl = x + 1 < W, xin = x <= W - 2;

  T byr[R], ayr[R];
  uint32_t rvm = 0;  // bit r: row r lies inside the grid
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int y = y0 - 2 + r;
    const bool in = y >= 0 && y < a.ny;
    if (in) rvm |= 1u << r;
    byr[r] = vreg(in ? a.by[y] : T(0));
    ayr[r] = vreg(in ? a.ay[y] : T(0));
  }
  // ay factors of the y differences with the grid-edge tests folded in (0
  // where the difference crosses the edge), so no per-row edge test (a lane
  // mask per row, kept live across the march) is needed: af = [y >= 1]
  // ay(y), al1 = [y + 1 < ny] ay(y + 1), afm1 = [y >= 1] ay(y - 1),
  // ain = [y <= ny - 2] ay(y)
  T af[R], al1[R], afm1[R], ain[R];
#pragma unroll
  for (int r = 0; r < R && FOLD; ++r) {
    const int y = y0 - 2 + r;
    auto ayv = [&](int yy) { return (yy >= 0 && yy < a.ny) ? a.ay[yy] : T(0); };
    af[r] = vreg(y >= 1 ? ayv(y) : T(0));
    al1[r] = y + 1 < a.ny ? ayv(y + 1) : T(0);
    afm1[r] = y >= 1 ? ayv(y - 1) : T(0);
    ain[r] = vreg(y <= a.ny - 2 ? ayv(y) : T(0));
  }
  auto rv = [&](int r) { return ((rvm >> r) & 1u) != 0; };
  // Accesses as in fwd2_tile: clamped row and column (every load in bounds,
  // no branch), a uniform buffer descriptor per array plane, the component
  // as scalar offset and the lane's in-plane byte offset as vector offset.
  // Planes are always in [0, nz): stage A touches plane p < nz, stage B
  // plane q = p - 1 >= 0.  Values of rows or columns outside the grid are
  // never combined into a cell of the grid (each difference across the grid
  // edge is selected away); stores cover grid cells only.
  const int xc = xv ? x : W - 1;
  uint32_t voff[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    voff[r] = uint32_t((min(max(y0 - 2 + r, 0), a.ny - 1) * W + xc) * int(sizeof(T)));
  static_assert(sizeof(T) == 4, "the two-step adjoint kernel is float32");
  const int cbytes = int(N * int64_t(sizeof(T)));
  const int pbytes = int(P * int64_t(sizeof(T)));
  auto rsrc = [&](const T* base, int plane, int ncomp) {
    const int off = plane * pbytes;
    return __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char*>(reinterpret_cast<const char*>(base)) + off, 0, ncomp * cbytes - off,
        0x00020000);
  };
  auto ld = [&](const __amdgpu_buffer_rsrc_t& rs, int c, int r) -> T {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, int(voff[r]), c * cbytes, 0));
  };
  // Stores of lanes outside the grid go to an offset past every record count
  // (bit 31 set; in-grid offsets stay below 2^31) and are dropped by the
  // descriptor's range check, so no store needs a lane mask.
  const uint32_t lane_oob = xv ? 0u : 0x80000000u;
  auto st = [&](const __amdgpu_buffer_rsrc_t& rs, int c, int r, T v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rs, int(voff[r] | lane_oob),
                                          c * cbytes, NT ? 2 : 0);
  };

  const int m = A2.n;  // sub-steps m (stage A) and m-1 (stage B)
  const bool has_src = a.src_comp >= 0;
  const T w1 = has_src ? a.src_wave[m] : T(0);
  const T w2 = has_src ? a.src_wave[m - 1] : T(0);
  // source profile and monitor adjoint terms: branch-free loads (0 outside
  // their boxes; see box_lane)
  const int s_lane = box_lane(a.src_box, x, has_src && xv);
  const auto rsp = box_rsrc(has_src ? a.src_prof : nullptr, a.src_box.cells() * 4);
  int m_lane[FUSED2_MAX_MON];
  __amdgpu_buffer_rsrc_t rmd[FUSED2_MAX_MON];
#pragma unroll
  for (int j = 0; j < FUSED2_MAX_MON; ++j) {
    const bool on = j < A2.n_mon;
    m_lane[j] = box_lane(A2.mon[j].box, x, on && xv);
    rmd[j] = box_rsrc(on ? A2.side[j] : nullptr, on ? A2.mon[j].box.cells() * A2.mon[j].ncomp * 4
                                                    : 0);
  }
  T spQ[TY];  // source profile at plane p-1, owned rows 2..TY+1 (0 outside the box)
#pragma unroll
  for (int r = 0; r < TY; ++r) spQ[r] = T(0);

  T uC[3][R], uN[3][R];   // u = ce lamE^{m+1} at planes p, p+1
  T lC[3][TY + 2], lN[3][TY + 2];  // raw lamE^{m+1}, rows 1..TY+2 (index r-1)
  T ceC[TY + 2], ceN[TY + 2];      // ce at planes p, p+1, rows 1..TY+2
  T m1F[3][TY + 2];       // mu^m at plane p-1, rows 1..TY+2
  T l1Q[3][TY + 2];       // lamE^m (after the monitor term) at plane p-1, rows 1..TY+2
  T u1Q[3][TY + 2];       // ce lamE^m at plane p-1
  T m2P[2][TY];           // mu^{m-1} (x, y) at plane p-2, rows 2..TY+1
  T sQ[TY];               // S of plane p-1 after stage A (owned rows 2..TY+1)
  const int zs = z0 >= 2 ? z0 - 2 : 0;
  {
    const auto rc = rsrc(a.ce, zs, 1);
    const auto rl = rsrc(a.in, zs, 6);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const T c = ld(rc, 0, r);
      if (r >= 1 && r <= TY + 2) ceC[r - 1] = c;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T l = ld(rl, k, r);
        uC[k][r] = c * l;
        if (r >= 1 && r <= TY + 2) lC[k][r - 1] = l;
      }
    }
    {
#pragma unroll
      for (int r = 0; r <= TY + 2; ++r) {
        (&LA(zs & 1, 0, r))[x] = uC[1][r];
        (&LA(zs & 1, 1, r))[x] = uC[2][r];
      }
    }
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) {
#pragma unroll
      for (int k = 0; k < 3; ++k) m1F[k][r] = l1Q[k][r] = u1Q[k][r] = T(0);
    }
#pragma unroll
    for (int r = 0; r < TY; ++r) m2P[0][r] = m2P[1][r] = sQ[r] = T(0);
  }
  __syncthreads();

#pragma unroll 1
  for (int p = zs; p <= z1; ++p) {
    const bool doA = p < a.nz;
    const bool doB = p - 1 >= zs;
    const int q = p - 1;
    const bool zl = p + 1 < a.nz, zf = p >= 1, zin = p <= a.nz - 2;
    T m1N[3][TY + 3];  // mu^m at plane p, rows 0..TY+2
    T l1P[3][TY + 2];  // lamE^m at plane p, rows 1..TY+2 (index r-1)
    T u1P[3][TY + 2];
    T sP[TY];
    T spP[TY];  // source profile at plane p, owned rows 2..TY+1
    T dpre[FUSED2_MAX_MON][TY + 2][3];  // monitor adjoint source, rows 1..TY+2
    // E^m, S at plane p and E^{m-1} at plane q (owned rows 2..TY+1) are
    // loaded here and consumed after the barriers.
    T emA[3][TY], sIn[TY], emB[3][TY];
    // ---------------- stage A, part 1: loads and mu^m(p)
    if (doA) {
      T m0[3][TY + 3];
      {
        const int pn = zl ? p + 1 : p;  // (the values of plane p+1 are unused when p+1 = nz)
        const auto rc = rsrc(a.ce, pn, 1);
        const auto rl = rsrc(a.in, pn, 6);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const T c = ld(rc, 0, r);
          if (r >= 1 && r <= TY + 2) ceN[r - 1] = c;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const T l = ld(rl, k, r);
            uN[k][r] = c * l;
            if (r >= 1 && r <= TY + 2) lN[k][r - 1] = l;
          }
        }
      }
      {
        const auto rm = rsrc(a.in, p, 6);
#pragma unroll
        for (int r = 0; r <= TY + 2; ++r) {
#pragma unroll
          for (int k = 0; k < 3; ++k) m0[k][r] = ld(rm, 3 + k, r);
        }
        const auto re = rsrc(a.em, p, 3);
        const auto rs = rsrc(a.S, p, 1);
#pragma unroll
        for (int r = 2; r <= TY + 1; ++r) {
#pragma unroll
          for (int k = 0; k < 3; ++k) emA[k][r - 2] = ld(re, k, r);
          sIn[r - 2] = ld(rs, 0, r);
        }
      }
      {
        const int sp = s_lane + box_plane(a.src_box, p);
#pragma unroll
        for (int r = 2; r <= TY + 1; ++r)
          spP[r - 2] = box_ld(rsp, sp + box_row(a.src_box, y0 - 2 + r), 0);
      }
      // adjoint source of sample m-1 for lamE^m(p), rows 1..TY+2 (0 outside
      // the monitor box)
#pragma unroll
      for (int j = 0; j < FUSED2_MAX_MON; ++j) {
        const FusedMon& mj = A2.mon[j];
        const int mp = m_lane[j] + box_plane(mj.box, p);
        const int cb = int(mj.box.cells()) * 4;
#pragma unroll
        for (int r = 1; r <= TY + 2; ++r) {
          const int off = mp + box_row(mj.box, y0 - 2 + r);
          int ci = 0;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const bool mc = j < A2.n_mon && ((mj.mask >> c) & 1);
            dpre[j][r - 1][c] = mc ? box_ld(rmd[j], off, ci * cb) : T(0);
            ci += mc ? 1 : 0;
          }
        }
      }
      const int pb = p & 1;
#pragma unroll
      for (int r = 0; r <= TY + 2; ++r) {
        const int y = y0 - 2 + r;
        const T ux = uC[0][r], uy = uC[1][r], uz = uC[2][r];
        const T uy_x1 = lds_nb<T>(xl, &LA(pb, 0, r), xp1);
        const T uz_x1 = lds_nb<T>(xl, &LA(pb, 1, r), xp1);
        const T tz_y = (zf ? uy : T(0)) - (zl ? uN[1][r] : T(0));
        const T tz_x = (zf ? ux : T(0)) - (zl ? uN[0][r] : T(0));
        T ty_z, ty_x;
        if constexpr (FOLD) {
          ty_z = af[r] * uz - al1[r] * uC[2][r + 1];
          ty_x = af[r] * ux - al1[r] * uC[0][r + 1];
        } else {
          const bool yf = y >= 1, yl = y + 1 < a.ny;
          ty_z = (yf ? ayr[r] * uz : T(0)) - (yl ? ayr[r + 1] * uC[2][r + 1] : T(0));
          ty_x = (yf ? ayr[r] * ux : T(0)) - (yl ? ayr[r + 1] * uC[0][r + 1] : T(0));
        }
        const T tx_z = (xf ? axv * uz : T(0)) - (xl ? axp * uz_x1 : T(0));
        const T tx_y = (xf ? uy : T(0)) - (xl ? uy_x1 : T(0));
        m1N[0][r] = byr[r] * m0[0][r] + axv * tz_y - ty_z;
        m1N[1][r] = bxv * m0[1][r] + tx_z - axv * tz_x;
        m1N[2][r] = bxv * m0[2][r] + ty_x - ayr[r] * tx_y;
      }
      {
#pragma unroll
        for (int r = 1; r <= TY + 2; ++r) {
          (&LB(0, r))[x] = m1N[1][r];
          (&LB(1, r))[x] = m1N[2][r];
        }
#pragma unroll
        for (int r = 0; r <= TY + 2; ++r) {
          (&LA(pb ^ 1, 0, r))[x] = uN[1][r];
          (&LA(pb ^ 1, 1, r))[x] = uN[2][r];
        }
      }
    }
    if (doB) {
      const auto re = rsrc(A2.em2, q, 3);
#pragma unroll
      for (int r = 2; r <= TY + 1; ++r) {
#pragma unroll
        for (int k = 0; k < 3; ++k) emB[k][r - 2] = ld(re, k, r);
      }
    }
    __syncthreads();
    // ---------------- stage A, part 2: lamE^m(p), S partial, monitor term
    if (doA) {
      const bool own = p >= z0 && p < z1;
#pragma unroll
      for (int r = 1; r <= TY + 2; ++r) {
        const int y = y0 - 2 + r;
        const bool mine = own && r >= 2 && r <= TY + 1 && rv(r);  // uniform (padding lanes: stores dropped)
        const T mx = m1N[0][r], my = m1N[1][r], mz = m1N[2][r];
        const T my_x0 = lds_nb<T>(xf, &LB(0, r), xm1);
        const T mz_x0 = lds_nb<T>(xf, &LB(1, r), xm1);
        T y_mz, y_mx;
        if constexpr (FOLD) {
          y_mz = afm1[r] * m1N[2][r - 1] - ain[r] * mz;
          y_mx = afm1[r] * m1N[0][r - 1] - ain[r] * mx;
        } else {
          const bool yf = y >= 1, yin = y <= a.ny - 2;
          y_mz = (yf ? ayr[r - 1] * m1N[2][r - 1] : T(0)) - (yin ? ayr[r] * mz : T(0));
          y_mx = (yf ? ayr[r - 1] * m1N[0][r - 1] : T(0)) - (yin ? ayr[r] * mx : T(0));
        }
        const T z_my = (zf ? m1F[1][r - 1] : T(0)) - (zin ? my : T(0));
        const T z_mx = (zf ? m1F[0][r - 1] : T(0)) - (zin ? mx : T(0));
        const T x_mz = (xf ? axm * mz_x0 : T(0)) - (xin ? axv * mz : T(0));
        const T x_my = (xf ? my_x0 : T(0)) - (xin ? my : T(0));
        T t[3];
        t[0] = kh * (y_mz - axv * z_my);
        t[1] = kh * (axv * z_mx - x_mz);
        t[2] = kh * (ayr[r] * x_my - y_mx);
        const T bxy = bxv * byr[r];
        T l[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) l[k] = bxy * lC[k][r - 1] + t[k];
        T em[3] = {T(0), T(0), T(0)};
        T s = T(0);
        if (r >= 2 && r <= TY + 1) {
          if (mine) {
#pragma unroll
            for (int k = 0; k < 3; ++k) em[k] = emA[k][r - 2];
            s = em[0] * t[0] + em[1] * t[1] + em[2] * t[2];
            {
              const T w = w1 * spP[r - 2];  // 0 outside the source box
#pragma unroll
              for (int cc = 0; cc < 3; ++cc)
                if (cc == a.src_comp) s -= w * lC[cc][r - 1];
            }
            s = sIn[r - 2] + s;
          }
        }
#pragma unroll
        for (int j = 0; j < FUSED2_MAX_MON; ++j) {
          if (j < A2.n_mon) {  // dpre is 0 outside the monitor box
            const T sacc = mon_apply<T>(A2.mon[j].mask, dpre[j][r - 1], l, em);
            if (mine) s += sacc;
          }
        }
        if (r >= 2 && r <= TY + 1) sP[r - 2] = s;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          l1P[k][r - 1] = l[k];
          u1P[k][r - 1] = ceC[r - 1] * l[k];
        }
      }
    }
    // ---------------- stage B, part 1: mu^{m-1}(q)
    T m2N[3][TY + 1];  // rows 1..TY+1 (index r-1)
    const bool zlq = q + 1 < a.nz, zfq = q >= 1, zinq = q <= a.nz - 2;
    if (doB) {
      const int qb = q & 1;
#pragma unroll
      for (int r = 1; r <= TY + 1; ++r) {
        const int y = y0 - 2 + r;
        const T ux = u1Q[0][r - 1], uy = u1Q[1][r - 1], uz = u1Q[2][r - 1];
        const T uy_x1 = lds_nb<T>(xl, &LC(qb, 0, r), xp1);
        const T uz_x1 = lds_nb<T>(xl, &LC(qb, 1, r), xp1);
        const T tz_y = (zfq ? uy : T(0)) - (zlq ? u1P[1][r - 1] : T(0));
        const T tz_x = (zfq ? ux : T(0)) - (zlq ? u1P[0][r - 1] : T(0));
        T ty_z, ty_x;
        if constexpr (FOLD) {
          ty_z = af[r] * uz - al1[r] * u1Q[2][r];
          ty_x = af[r] * ux - al1[r] * u1Q[0][r];
        } else {
          const bool yf = y >= 1, yl = y + 1 < a.ny;
          ty_z = (yf ? ayr[r] * uz : T(0)) - (yl ? ayr[r + 1] * u1Q[2][r] : T(0));
          ty_x = (yf ? ayr[r] * ux : T(0)) - (yl ? ayr[r + 1] * u1Q[0][r] : T(0));
        }
        const T tx_z = (xf ? axv * uz : T(0)) - (xl ? axp * uz_x1 : T(0));
        const T tx_y = (xf ? uy : T(0)) - (xl ? uy_x1 : T(0));
        m2N[0][r - 1] = byr[r] * m1F[0][r - 1] + axv * tz_y - ty_z;
        m2N[1][r - 1] = bxv * m1F[1][r - 1] + tx_z - axv * tz_x;
        m2N[2][r - 1] = bxv * m1F[2][r - 1] + ty_x - ayr[r] * tx_y;
      }
      {
#pragma unroll
        for (int r = 2; r <= TY + 1; ++r) {
          (&LD(0, r))[x] = m2N[1][r - 1];
          (&LD(1, r))[x] = m2N[2][r - 1];
        }
      }
    }
    if (doA) {
#pragma unroll
      for (int r = 1; r <= TY + 1; ++r) {
        (&LC(p & 1, 0, r))[x] = u1P[1][r - 1];
        (&LC(p & 1, 1, r))[x] = u1P[2][r - 1];
      }
    }
    __syncthreads();
    // ---------------- stage B, part 2: lamE^{m-1}(q), S, store
    if (doB) {
      const bool own = q >= z0;
      const auto ro = rsrc(a.out, q, 6);
      const auto rs = rsrc(a.S, q, 1);
#pragma unroll
      for (int r = 2; r <= TY + 1; ++r) {
        const int y = y0 - 2 + r;
        const bool mine = own && rv(r);  // uniform (padding lanes: stores dropped)
        const T mx = m2N[0][r - 1], my = m2N[1][r - 1], mz = m2N[2][r - 1];
        const T my_x0 = lds_nb<T>(xf, &LD(0, r), xm1);
        const T mz_x0 = lds_nb<T>(xf, &LD(1, r), xm1);
        T y_mz, y_mx;
        if constexpr (FOLD) {
          y_mz = afm1[r] * m2N[2][r - 2] - ain[r] * mz;
          y_mx = afm1[r] * m2N[0][r - 2] - ain[r] * mx;
        } else {
          const bool yf = y >= 1, yin = y <= a.ny - 2;
          y_mz = (yf ? ayr[r - 1] * m2N[2][r - 2] : T(0)) - (yin ? ayr[r] * mz : T(0));
          y_mx = (yf ? ayr[r - 1] * m2N[0][r - 2] : T(0)) - (yin ? ayr[r] * mx : T(0));
        }
        const T z_my = (zfq ? m2P[1][r - 2] : T(0)) - (zinq ? my : T(0));
        const T z_mx = (zfq ? m2P[0][r - 2] : T(0)) - (zinq ? mx : T(0));
        const T x_mz = (xf ? axm * mz_x0 : T(0)) - (xin ? axv * mz : T(0));
        const T x_my = (xf ? my_x0 : T(0)) - (xin ? my : T(0));
        T t[3];
        t[0] = kh * (y_mz - axv * z_my);
        t[1] = kh * (axv * z_mx - x_mz);
        t[2] = kh * (ayr[r] * x_my - y_mx);
        const T bxy = bxv * byr[r];
        T l[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) l[k] = bxy * l1Q[k][r - 1] + t[k];
        if (mine) {
          T em[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) em[k] = emB[k][r - 2];
          T s = em[0] * t[0] + em[1] * t[1] + em[2] * t[2];
          {
            const T w = w2 * spQ[r - 2];  // 0 outside the source box
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
              if (cc == a.src_comp) s -= w * l1Q[cc][r - 1];
          }
          s = sQ[r - 2] + s;  // the last sample's monitor term follows the launch
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            st(ro, k, r, l[k]);
            st(ro, 3 + k, r, m2N[k][r - 1]);
          }
          st(rs, 0, r, s);
        }
      }
#pragma unroll
      for (int r = 2; r <= TY + 1; ++r) {
        m2P[0][r - 2] = m2N[0][r - 1];
        m2P[1][r - 2] = m2N[1][r - 1];
      }
    }
    // ---------------- advance the march
    if (doA) {
#pragma unroll
      for (int r = 1; r <= TY + 2; ++r) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          m1F[k][r - 1] = m1N[k][r];
          l1Q[k][r - 1] = l1P[k][r - 1];
          u1Q[k][r - 1] = u1P[k][r - 1];
        }
      }
#pragma unroll
      for (int r = 0; r < TY; ++r) {
        sQ[r] = sP[r];
        spQ[r] = spP[r];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int k = 0; k < 3; ++k) uC[k][r] = uN[k][r];
      }
#pragma unroll
      for (int r = 0; r < TY + 2; ++r) {
        ceC[r] = ceN[r];
#pragma unroll
        for (int k = 0; k < 3; ++k) lC[k][r] = lN[k][r];
      }
    }
  }
#undef LA
#undef LB
#undef LC
#undef LD
}

template <typename T, int TY, int MINW, bool NT, int LW>
__global__ __launch_bounds__(256, MINW) void k3_fused2_adj(Fused2Args<T> a) {
  extern __shared__ unsigned char smem_raw[];
  int y0, z0;
  bool live;
  block_tile2<TY>(a.a.nblocks, a.a.ty_tiles, a.a.zc, y0, z0, live);
  if (!live) return;
  adj2_tile<T, TY, true, NT, LW>(a, y0, z0, reinterpret_cast<T*>(smem_raw));
}

template <typename T, int TY, int MINW, bool NT = false, int LW = 256>
static int launch_adj2_ty(const Fused2Args<T>& a, hipStream_t s) {
  const int threads = ((a.a.nx + 63) / 64) * 64;
  const int nb = ((a.a.nblocks + 7) / 8) * 8;
  DNX_REQUIRE(threads <= LW, "fused2: LDS pitch %d below the work-group width %d", LW, threads);
  const size_t smem = sizeof(T) * size_t(fwd2_lds_elems<TY>()) * LW;
  static bool attr_set = false;
  if (!attr_set) {
    DNX_CHECK_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k3_fused2_adj<T, TY, MINW, NT, LW>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(T) * fwd2_lds_elems<TY>() * LW)));
    attr_set = true;
  }
  hipLaunchKernelGGL((k3_fused2_adj<T, TY, MINW, NT, LW>), dim3(nb), dim3(threads), smem, s, a);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_fused2_adj(const Fused2Args<T>& a, int ty, hipStream_t s) {
  // occupancy 2 (launch bounds for two work-groups per CU) unless
  // DNX_FDTD3D_ADJ2_OCC=1; TY = 4 fits the register file only at occupancy 1
  const bool two = option(DNX_OPT_FDTD3D_ADJ2_OCC) != 1;
  // (at the default occupancy: pitch 128 up to 128 columns, non-temporal
  // stores beyond the cache)
  const bool nt = stream_stores(a.a.N), narrow = a.a.nx <= 128;
  if (ty == 4) return launch_adj2_ty<T, 4, 1>(a, s);
  if (ty == 2)
    return !two ? launch_adj2_ty<T, 2, 1>(a, s)
                : narrow ? launch_adj2_ty<T, 2, 2, false, 128>(a, s)
                : nt     ? launch_adj2_ty<T, 2, 2, true>(a, s)
                         : launch_adj2_ty<T, 2, 2>(a, s);
  if (ty == 1)
    return !two ? launch_adj2_ty<T, 1, 1>(a, s)
                : narrow ? launch_adj2_ty<T, 1, 2, false, 128>(a, s)
                : nt     ? launch_adj2_ty<T, 1, 2, true>(a, s)
                         : launch_adj2_ty<T, 1, 2>(a, s);
  set_error("fused2: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template int launch_fused2_adj<float>(const Fused2Args<float>&, int, hipStream_t);

// ---------------------------------------------------------------------------
// Batched two-step launches: one launch advances every design of a batch
// by two steps, one work-group per tile (tile shapes of kernel 2 / 3), the
// work-group -> (design, tile) assignment read from the plan's block map
// (XCD-balanced, heavy tiles first; batch_block_map).  The tile bodies are
// the single-grid kernels' with folded edge factors; the forward stages
// its monitor samples (a recompute after the launch would cost more here:
// the batch's monitor planes are many and strided).  Buffers
// follow layout 1 of batch_in_slot; E^m of the adjoint comes from the
// checkpoint (m % K == 0), the final segments' store or the replay store
// (batch_em), as in k3_batch_adj.
// ---------------------------------------------------------------------------

template <typename T>
__device__ inline void batch2_mon(const Batch3Design<T>& Dd, bool on, Fused2Args<T>& f) {
  f.n_mon = (on && Dd.m_nf > 0) ? 1 : 0;
  f.mon[0].box = Dd.mbox;
  f.mon[0].mask = Dd.m_mask;
  f.mon[0].ncomp = (Dd.m_mask & 1) + ((Dd.m_mask >> 1) & 1) + ((Dd.m_mask >> 2) & 1);
  f.side[0] = Dd.m_side;
}

template <typename T, int TY, int LW, bool NT>
__global__ __launch_bounds__(256, 2) void k3_batch2_fwd(Batch3Launch L) {
  extern __shared__ unsigned char smem_raw[];
  const int2 bt = L.block_tile[blockIdx.x];
  if (bt.x < 0) return;
  const Batch3Design<T>& Dd = static_cast<const Batch3Design<T>*>(L.designs)[bt.x];
  const int n = *L.step;
  const int ib = batch_in_slot(1, n);
  Fused2Args<T> f{};
  f.a = Dd.a;
  f.a.in = Dd.buf[ib];
  f.a.out = Dd.buf[ib ^ 1];
  f.a.es = batch_es(Dd, L.store, n, L.K);
  f.es2 = batch_es(Dd, L.store, n + 1, L.K);
  f.a.src_wave = Dd.wave;
  f.n = n;
  const int lt = bt.y;
  batch2_mon(Dd, L.store != 1, f);
  fwd2_tile<T, TY, false, true, true, NT, false, LW>(f, (lt % f.a.ty_tiles) * TY,
                                       (lt / f.a.ty_tiles) * f.a.zc,
                                       reinterpret_cast<T*>(smem_raw));
}

template <typename T, int TY, int LW, bool NT>
__global__ __launch_bounds__(256, 2) void k3_batch2_adj(Batch3Launch L) {
  extern __shared__ unsigned char smem_raw[];
  const int2 bt = L.block_tile[blockIdx.x];
  if (bt.x < 0) return;
  const Batch3Design<T>& Dd = static_cast<const Batch3Design<T>*>(L.designs)[bt.x];
  const int m = *L.step;  // steps m and m-1
  const int ib = batch_in_slot(1, m + 1);
  Fused2Args<T> f{};
  f.a = Dd.a;
  f.a.in = Dd.buf[ib];
  f.a.out = Dd.buf[ib ^ 1];
  f.a.em = batch_em(Dd, m, L.K);
  f.em2 = batch_em(Dd, m - 1, L.K);
  f.a.src_wave = Dd.wave;
  f.n = m;
  batch2_mon(Dd, true, f);
  const int lt = bt.y;
  adj2_tile<T, TY, true, NT, LW>(f, (lt % f.a.ty_tiles) * TY, (lt / f.a.ty_tiles) * f.a.zc,
                         reinterpret_cast<T*>(smem_raw));
}

// After a forward pair at counter n: acc += (v1 at phase row n) then (v2 at
// row n+1), the arithmetic of k_monitor_pair.  Before an adjoint pair at
// counter m: m_side = T(sum_f Re(g e^{i phase[m-1]})), that of k_monitor_d.
template <typename T>
__global__ void k_batch_monitor2(const Batch3Design<T>* D, const int* step, int adjoint) {
  const Batch3Design<T>& dz = D[blockIdx.y];
  if (dz.m_nf <= 0) return;
  const int64_t cells = dz.mbox.cells();
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  const int c = *step;
  const int mask = dz.m_mask;
  const int ncomp = (mask & 1) + ((mask >> 1) & 1) + ((mask >> 2) & 1);
  const int nf = dz.m_nf;
  if (!adjoint) {
    const double* ph = dz.m_phase + int64_t(c) * nf * 2;
    for (int ci = 0; ci < ncomp; ++ci) {
      const double v1 = double(dz.m_side[int64_t(ci) * cells + i]);
      const double v2 = double(dz.m_side[(int64_t(ncomp) + ci) * cells + i]);
      for (int f = 0; f < nf; ++f) {
        double* a = dz.m_acc + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
        a[0] += v1 * ph[2 * f];
        a[1] -= v1 * ph[2 * f + 1];
        a[0] += v2 * ph[2 * nf + 2 * f];
        a[1] -= v2 * ph[2 * nf + 2 * f + 1];
      }
    }
  } else {
    const double* ph = dz.m_phase + int64_t(c - 1) * nf * 2;
    for (int ci = 0; ci < ncomp; ++ci) {
      double sv = 0.0;
      for (int f = 0; f < nf; ++f) {
        const double* gg = dz.m_grad + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
        sv += gg[0] * ph[2 * f] - gg[1] * ph[2 * f + 1];
      }
      dz.m_side[int64_t(ci) * cells + i] = T(sv);
    }
  }
}

// LDS pitch of a batch: 128 when every design has at most 128 columns.
inline int batch2_pitch(int max_w) { return max_w <= 128 ? 128 : 256; }

template <typename T, int TY>
static size_t batch2_smem(int max_w) {
  return sizeof(T) * size_t(fwd2_lds_elems<TY>()) * size_t(batch2_pitch(max_w));
}

// The batched two-step adjoint exists for 2-row tiles only (with 4 rows
// its per-design arguments make it spill to scratch).
template <typename T, int TY, int LW, bool NT>
static const void* batch2_kernel(bool adjoint) {
  if constexpr (TY == 2) {
    if (adjoint) return reinterpret_cast<const void*>(&k3_batch2_adj<T, TY, LW, NT>);
  }
  return adjoint ? nullptr : reinterpret_cast<const void*>(&k3_batch2_fwd<T, TY, LW, NT>);
}

template <typename T, int TY, bool NT>
static const void* batch2_kernel_w(bool adjoint, int max_w) {
  return batch2_pitch(max_w) == 128 ? batch2_kernel<T, TY, 128, NT>(adjoint)
                                    : batch2_kernel<T, TY, 256, NT>(adjoint);
}

template <typename T, int TY>
static int prepare_batch2_ty(bool adjoint, int max_w, int* blocks_per_cu) {
  const void* f = batch2_kernel_w<T, TY, false>(adjoint, max_w);
  DNX_REQUIRE(f != nullptr, "batch3 pair: no %d-row batched two-step adjoint", TY);
  DNX_CHECK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(batch2_smem<T, TY>(max_w))));
  DNX_CHECK_HIP(hipFuncSetAttribute(batch2_kernel_w<T, TY, true>(adjoint, max_w),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(batch2_smem<T, TY>(max_w))));
  const int threads = ((max_w + 63) / 64) * 64;
  int nb = 0;
  DNX_CHECK_HIP(
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, threads, batch2_smem<T, TY>(max_w)));
  *blocks_per_cu = nb > 0 ? nb : 1;
  return DNX_OK;
}

template <typename T, int TY, int LW, bool NT>
static void launch_batch2_nt(const Batch3Launch& L, bool adjoint, int threads, size_t smem,
                             hipStream_t s) {
  if constexpr (TY == 2) {
    if (adjoint) {
      hipLaunchKernelGGL((k3_batch2_adj<T, TY, LW, NT>), dim3(L.blocks), dim3(threads), smem, s,
                         L);
      return;
    }
  }
  hipLaunchKernelGGL((k3_batch2_fwd<T, TY, LW, NT>), dim3(L.blocks), dim3(threads), smem, s, L);
}

// Non-temporal output stores when the batch's state exceeds half the
// last-level cache (as for a single grid, stream_stores).
template <typename T, int TY, int LW>
static void launch_batch2_lw(const Batch3Launch& L, bool adjoint, int threads, size_t smem,
                             hipStream_t s) {
  if (L.nt)
    launch_batch2_nt<T, TY, LW, true>(L, adjoint, threads, smem, s);
  else
    launch_batch2_nt<T, TY, LW, false>(L, adjoint, threads, smem, s);
}

template <typename T, int TY>
static int launch_batch2_ty(const Batch3Launch& L, bool adjoint, int max_w, int grid,
                            hipStream_t s) {
  (void)grid;  // one work-group per block-map entry
  const int threads = ((max_w + 63) / 64) * 64;
  const size_t smem = batch2_smem<T, TY>(max_w);
  const void* kern = batch2_kernel_w<T, TY, false>(adjoint, max_w);
  DNX_REQUIRE(kern != nullptr, "batch3 pair: no %d-row batched two-step adjoint", TY);
  if (batch2_pitch(max_w) == 128)
    launch_batch2_lw<T, TY, 128>(L, adjoint, threads, smem, s);
  else
    launch_batch2_lw<T, TY, 256>(L, adjoint, threads, smem, s);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int prepare_batch3_pair(int ty, bool adjoint, int max_w, int* blocks_per_cu) {
  if (ty == 4) return prepare_batch2_ty<T, 4>(adjoint, max_w, blocks_per_cu);
  if (ty == 2) return prepare_batch2_ty<T, 2>(adjoint, max_w, blocks_per_cu);
  set_error("batch3 pair: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template <typename T>
int launch_batch3_pair(const Batch3Launch& L, bool adjoint, int ty, int max_w, int grid,
                       hipStream_t s) {
  if (ty == 4) return launch_batch2_ty<T, 4>(L, adjoint, max_w, grid, s);
  if (ty == 2) return launch_batch2_ty<T, 2>(L, adjoint, max_w, grid, s);
  set_error("batch3 pair: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template <typename T>
int launch_batch_monitor_pair(const void* designs, int n_designs, int64_t max_cells,
                              const int* step, bool adjoint, hipStream_t s) {
  if (max_cells <= 0) return DNX_OK;
  hipLaunchKernelGGL(k_batch_monitor2<T>, dim3(cdiv(max_cells, 256), n_designs), dim3(256), 0,
                     s, static_cast<const Batch3Design<T>*>(designs), step, adjoint ? 1 : 0);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

// Each design's tiles are cut into runs of consecutive tiles of at most
// about 1/16 of the batch's tile cost, 1/64 for a batch of fewer than 16
// designs (so a batch of a few designs spreads over all 8 XCDs, while the
// designs of a batch of 16 or more similar ones stay whole: measured
// faster for both), and the runs are
// packed onto the 8 XCDs (work-group b runs on XCD b % 8) longest-
// processing-time first.  On each XCD the runs go in decreasing tile cost,
// then design, then first tile, so the runs of one design that share an XCD
// follow each other in tile order (neighbouring tiles share halo rows in
// that XCD's L2).  Entry s * 8 + k is the s-th tile of XCD k; shorter lists
// are padded with idle entries.
void batch_block_map(const int* tiles, const int64_t* tile_cost, int n, std::vector<int2>& out) {
  struct Run {
    int d, t0, t1;
    int64_t cost;
  };
  const int64_t kRunsPerBatch = n < 16 ? 64 : 16;
  int64_t total = 0;
  for (int d = 0; d < n; ++d) total += tile_cost[d] * tiles[d];
  const int64_t unit = std::max<int64_t>(1, total / kRunsPerBatch);
  std::vector<Run> runs;
  for (int d = 0; d < n; ++d) {
    if (tiles[d] <= 0) continue;
    const int64_t c = tile_cost[d] * tiles[d];
    const int64_t k = std::max<int64_t>(1, std::min<int64_t>(tiles[d], (c + unit - 1) / unit));
    const int len = int((tiles[d] + k - 1) / k);
    for (int t = 0; t < tiles[d]; t += len) {
      const int e = std::min(tiles[d], t + len);
      runs.push_back({d, t, e, tile_cost[d] * (e - t)});
    }
  }
  std::stable_sort(runs.begin(), runs.end(),
                   [](const Run& a, const Run& b) { return a.cost > b.cost; });
  std::vector<std::vector<Run>> bin(8);
  int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (const Run& r : runs) {
    int k = 0;
    for (int j = 1; j < 8; ++j)
      if (load[j] < load[k]) k = j;
    bin[k].push_back(r);
    load[k] += r.cost;
  }
  size_t rows = 0;
  std::vector<std::vector<int2>> lists(8);
  for (int k = 0; k < 8; ++k) {
    std::sort(bin[k].begin(), bin[k].end(), [&](const Run& a, const Run& b) {
      if (tile_cost[a.d] != tile_cost[b.d]) return tile_cost[a.d] > tile_cost[b.d];
      if (a.d != b.d) return a.d < b.d;
      return a.t0 < b.t0;
    });
    for (const Run& r : bin[k])
      for (int t = r.t0; t < r.t1; ++t) lists[k].push_back(make_int2(r.d, t));
    rows = std::max(rows, lists[k].size());
  }
  out.assign(rows * 8, make_int2(-1, -1));
  for (int k = 0; k < 8; ++k)
    for (size_t r = 0; r < lists[k].size(); ++r) out[r * 8 + k] = lists[k][r];
}

template int prepare_batch3_pair<float>(int, bool, int, int*);
template int launch_batch3_pair<float>(const Batch3Launch&, bool, int, int, int, hipStream_t);
template int launch_batch_monitor_pair<float>(const void*, int, int64_t, const int*, bool,
                                              hipStream_t);

}  // namespace dnx

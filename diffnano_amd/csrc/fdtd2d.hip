// 2D Yee FDTD on gfx950 (DiffNano FDTDSolver2D): TM and TE leapfrog steps
// and their exact discrete adjoints.
//
// Rows (axis 0, ny) carry the profile (by, ay = 1 + cy); columns (axis 1, nx)
// carry (bx, ax).  kh = dt/dl, ce = dt/(eps dl).  D+r / D-r differ along
// rows (axis 0), D+c / D-c along columns (axis 1); zero at the truncated end.
// TM (fdtd2d.py:233-267), state [Ez, Hx, Hy]:
//   Hx' = by Hx - kh ay D+c Ez
//   Hy' = bx Hy + kh ax D+r Ez
//   Ez' = bxby Ez + ce (ax D-r Hy' - ay D-c Hx')
// TE (fdtd2d.py:269-303), state [Hz, Ex, Ey], E first, reference Hz sign:
//   Ex' = bx Ex + ce ax D-c Hz
//   Ey' = by Ey - ce ay D-r Hz
//   Hz' = bxby Hz + kh (ax D+r Ey' - ay D+c Ex')
#include <algorithm>
#include <atomic>
#include <utility>

#include "dnx_aux.h"

// No multiply-add contraction: the batched and single-design kernels round
// identically.
#pragma clang fp contract(off)

namespace dnx {

template <typename T>
struct G2 {
  int ny, nx;
  int64_t n;
  T kh;
  const T* __restrict__ ce;
  const T* __restrict__ by;
  const T* __restrict__ ay;
  const T* __restrict__ bx;
  const T* __restrict__ ax;
};

constexpr int BX2 = 64;
constexpr int BY2 = 4;

template <typename T>
static G2<T> make_g2(const dnx_grid2d* g) {
  G2<T> r;
  r.ny = g->ny;
  r.nx = g->nx;
  r.n = int64_t(g->ny) * g->nx;
  r.kh = T(g->kh);
  r.ce = static_cast<const T*>(g->ce);
  r.by = static_cast<const T*>(g->by);
  r.ay = static_cast<const T*>(g->ay);
  r.bx = static_cast<const T*>(g->bx);
  r.ax = static_cast<const T*>(g->ax);
  return r;
}

#define DNX_2D_PROLOGUE                                  \
  const int c = blockIdx.x * BX2 + threadIdx.x;         \
  const int r = blockIdx.y * BY2 + threadIdx.y;         \
  if (c >= g.nx || r >= g.ny) return;                    \
  const int64_t N = g.n, nx = g.nx;                      \
  const int64_t i = int64_t(r) * nx + c;

// ---- TM forward ----------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2tm_h(G2<T> g, T* __restrict__ st) {
  DNX_2D_PROLOGUE
  const T* Ez = st;
  T* Hx = st + N;
  T* Hy = st + 2 * N;
  const T ez = Ez[i];
  const T dc = (c + 1 < g.nx) ? Ez[i + 1] - ez : T(0);
  const T dr = (r + 1 < g.ny) ? Ez[i + nx] - ez : T(0);
  Hx[i] = g.by[r] * Hx[i] - g.kh * (g.ay[r] * dc);
  Hy[i] = g.bx[c] * Hy[i] + g.kh * (g.ax[c] * dr);
}

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2tm_e(G2<T> g, T* __restrict__ st,
                                                   T* __restrict__ es) {
  DNX_2D_PROLOGUE
  T* Ez = st;
  const T* Hx = st + N;
  const T* Hy = st + 2 * N;
  const T dHy = r > 0 ? Hy[i] - Hy[i - nx] : T(0);
  const T dHx = c > 0 ? Hx[i] - Hx[i - 1] : T(0);
  const T ez = (g.bx[c] * g.by[r]) * Ez[i] + g.ce[i] * (g.ax[c] * dHy - g.ay[r] * dHx);
  Ez[i] = ez;
  if (es) es[i] = ez;
}

// ---- TM with the source and monitors fused ------------------------------
//
// The single-design TM step with its (at most one) soft source and (at most
// kMaxMon2) Ez DFT monitors done by the E kernel, and the adjoint step with
// the source's adjoint term in the H kernel and the monitors' adjoint terms
// in the E kernel: two launches per step instead of four.  Per element the
// operations and their order are those of the separate kernels (k_source,
// k_monitor, k_source_adj, k_monitor_adj), so results are bitwise theirs.

constexpr int kMaxMon2 = 4;

template <typename T>
struct Side2 {
  int has_src;
  Box sb;
  const T* sprof;
  const T* swave;  // waveform table (indexed by the step)
  int n_mon;
  Box mb[kMaxMon2];
  int nf[kMaxMon2];
  const double* ph[kMaxMon2];  // phase tables [step][f][2]
  double* acc[kMaxMon2];       // forward: accumulators [f][1][cells][2]
  const double* grd[kMaxMon2];  // adjoint: their gradients, same layout
};

__device__ __forceinline__ bool in_box2(const Box& b, int r, int c, int64_t& q) {
  const int dr = r - b.y0, dc = c - b.x0;
  if (dr < 0 || dr >= b.ny || dc < 0 || dc >= b.nx) return false;
  q = int64_t(dr) * b.nx + dc;
  return true;
}

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2tm_e_side(G2<T> g, T* __restrict__ st,
                                                        T* __restrict__ es, Side2<T> f, int n) {
  DNX_2D_PROLOGUE
  T* Ez = st;
  const T* Hx = st + N;
  const T* Hy = st + 2 * N;
  const T dHy = r > 0 ? Hy[i] - Hy[i - nx] : T(0);
  const T dHx = c > 0 ? Hx[i] - Hx[i - 1] : T(0);
  T ez = (g.bx[c] * g.by[r]) * Ez[i] + g.ce[i] * (g.ax[c] * dHy - g.ay[r] * dHx);
  int64_t q;
  if (f.has_src && in_box2(f.sb, r, c, q)) ez += f.swave[n] * f.sprof[q];
  Ez[i] = ez;
  if (es) es[i] = ez;
  for (int j = 0; j < f.n_mon; ++j) {
    if (!in_box2(f.mb[j], r, c, q)) continue;
    const int64_t cells = f.mb[j].cells();
    const double* ph = f.ph[j] + int64_t(n) * f.nf[j] * 2;
    const double v = double(ez);
    for (int k = 0; k < f.nf[j]; ++k) {
      double* a = f.acc[j] + (int64_t(k) * cells + q) * 2;
      a[0] += v * ph[2 * k];
      a[1] -= v * ph[2 * k + 1];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2tm_adj_h_side(G2<T> g, T* __restrict__ lam,
                                                            T* __restrict__ S, Side2<T> f,
                                                            int m) {
  DNX_2D_PROLOGUE
  const T* L = lam;
  T* Mx = lam + N;
  T* My = lam + 2 * N;
  int64_t q;
  if (f.has_src && in_box2(f.sb, r, c, q)) S[i] -= f.swave[m] * f.sprof[q] * L[i];
  const T u = g.ce[i] * L[i];
  const T tc = (c >= 1 ? u : T(0)) - (c + 1 < g.nx ? g.ce[i + 1] * L[i + 1] : T(0));
  const T tr = (r >= 1 ? u : T(0)) - (r + 1 < g.ny ? g.ce[i + nx] * L[i + nx] : T(0));
  Mx[i] = g.by[r] * Mx[i] - g.ay[r] * tc;
  My[i] = g.bx[c] * My[i] + g.ax[c] * tr;
}

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2tm_adj_e_side(G2<T> g, T* __restrict__ lam,
                                                            const T* __restrict__ em,
                                                            T* __restrict__ S, Side2<T> f,
                                                            int m) {
  DNX_2D_PROLOGUE
  T* L = lam;
  const T* Mx = lam + N;
  const T* My = lam + 2 * N;
  const T r_my = (r >= 1 ? My[i - nx] : T(0)) - (r <= g.ny - 2 ? My[i] : T(0));
  const T c_mx = (c >= 1 ? Mx[i - 1] : T(0)) - (c <= g.nx - 2 ? Mx[i] : T(0));
  const T t = g.kh * (g.ax[c] * r_my - g.ay[r] * c_mx);
  L[i] = (g.bx[c] * g.by[r]) * L[i] + t;
  S[i] += em[i] * t;
  for (int j = 0; j < f.n_mon; ++j) {
    int64_t q;
    if (!in_box2(f.mb[j], r, c, q)) continue;
    const int64_t cells = f.mb[j].cells();
    const double* ph = f.ph[j] + int64_t(m - 1) * f.nf[j] * 2;
    double d = 0.0;
    for (int k = 0; k < f.nf[j]; ++k) {
      const double* gg = f.grd[j] + (int64_t(k) * cells + q) * 2;
      d += gg[0] * ph[2 * k] - gg[1] * ph[2 * k + 1];
    }
    const T dv = T(d);
    L[i] += dv;
    T sacc = T(0);
    sacc += em[i] * dv;
    S[i] += sacc;
  }
}

// Whether the TM step takes the fused kernels: TM, at most one source, at
// most kMaxMon2 monitors, each of Ez only, and not switched off.
static bool side2_ok(const dnx_grid2d* gd, int n_src, const dnx_monitor* mons, int n_mon) {
  if (gd->te || n_src > 1 || n_mon > kMaxMon2 || option(DNX_OPT_FDTD2D_SEPARATE_SIDE) != 0)
    return false;
  for (int j = 0; j < n_mon; ++j)
    if (mons[j].comp_mask != 1) return false;
  return true;
}

template <typename T>
static Side2<T> make_side2(const dnx_source* srcs, int n_src, const dnx_monitor* mons,
                           int n_mon, bool adjoint) {
  Side2<T> f{};
  f.has_src = n_src > 0 ? 1 : 0;
  if (n_src > 0) {
    f.sb = make_box(srcs[0].box);
    f.sprof = static_cast<const T*>(srcs[0].profile);
    f.swave = static_cast<const T*>(srcs[0].wave);
  }
  f.n_mon = n_mon;
  for (int j = 0; j < n_mon; ++j) {
    f.mb[j] = make_box(mons[j].box);
    f.nf[j] = mons[j].n_freq;
    f.ph[j] = mons[j].phase;
    f.acc[j] = adjoint ? nullptr : mons[j].acc;
    f.grd[j] = adjoint ? mons[j].grad : nullptr;
  }
  return f;
}

// ---- TM, several steps per launch (temporal blocking) --------------------
//
// One 1024-thread work-group per tile of RW - 2 HS cells square.  It loads
// the tile with a halo of HS cells on each side (a region of RW x RW cells;
// each thread's cells in registers, LDS for the values its neighbours read,
// RW * RW / 1024 rows per thread) and advances the whole region by ks <= HS
// steps.  A
// region cell whose neighbour lies outside the region (but inside the grid)
// takes its own value in the neighbour's place; the error that starts there
// moves one cell per step (the H update reads the +1 neighbours, the E update
// the -1 neighbours), so after ks steps every cell of the tile is exact.  The
// work-group writes the tile's cells only, into the other state buffer: the
// neighbouring tiles read their halos from the input buffer.  The E store
// and the monitors take the tile's cells at every step; the source is
// applied to every region cell it covers.  Per cell and step the arithmetic
// and its order are those of k2tm_h + k2tm_e_side (forward) and
// k2tm_adj_h_side + k2tm_adj_e_side (adjoint), so results are bitwise theirs.
//
// The step loop reads global memory only for the adjoint's E^m (loaded one
// step ahead): the waveform values and the DFT phases of the launch's steps
// are staged in LDS, and so are the DFT accumulators of the tile's monitor
// cells (forward) and the monitor gradients of the region's cells (adjoint)
// when they fit in TB_SLOTS complex values (otherwise those go to global
// memory at every step, as in the step kernels).

// Region shapes: RW x RW cells (RW = 64: four rows per thread; RW = 32:
// one), RW threads per row block, 1024 / RW row blocks.
constexpr int TBS = 16;               // largest halo (= most steps per launch)
constexpr int TB_NF = 8;              // most frequencies per monitor on this path
constexpr int TB_SLOTS = 1024;        // LDS slots for accumulators / gradients

struct TbTile {
  int ri0, ci0;            // region origin (row, column)
  int or0, or1, oc0, oc1;  // the tile: rows [or0, or1), columns [oc0, oc1)
};

// The tile of this work-group for a halo of HS cells in RW x RW regions
// (tile edge RW - 2 HS).
template <int HS, int RW>
__device__ __forceinline__ TbTile tb_tile(int ny, int nx) {
  constexpr int TBW = RW, TBH = RW;
  constexpr int TI = TBW - 2 * HS;
  static_assert(TI > 0, "the halo must leave a tile");
  TbTile t;
  const int a = blockIdx.y * TI, b = blockIdx.x * TI;
  t.ri0 = max(0, min(a - HS, ny - TBH));
  t.ci0 = max(0, min(b - HS, nx - TBW));
  t.or0 = a;
  t.or1 = min(a + TI, ny);
  t.oc0 = b;
  t.oc1 = min(b + TI, nx);
  return t;
}

// Monitor m's box clipped to a rectangle, with the first of its LDS slots:
// slot(k, r, c) = base + (k * (r1 - r0) + r - r0) * (c1 - c0) + c - c0.
struct TbClip {
  int r0, r1, c0, c1, base;
  int nf, y0, x0, bnx;  // the monitor's frequencies and box origin / width
};

template <typename T>
__device__ __forceinline__ int tb_clip(const Side2<T>& f, int rr0, int rr1, int cc0, int cc1,
                                       TbClip* cl) {
  int n = 0;
  for (int m = f.n_mon; m < kMaxMon2; ++m) cl[m] = TbClip{0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int m = 0; m < f.n_mon; ++m) {
    const Box& b = f.mb[m];
    TbClip k;
    k.r0 = max(b.y0, rr0);
    k.r1 = max(k.r0, min(b.y0 + b.ny, rr1));
    k.c0 = max(b.x0, cc0);
    k.c1 = max(k.c0, min(b.x0 + b.nx, cc1));
    k.base = n;
    k.nf = f.nf[m];
    k.y0 = b.y0;
    k.x0 = b.x0;
    k.bnx = b.nx;
    cl[m] = k;
    n += f.nf[m] * (k.r1 - k.r0) * (k.c1 - k.c0);
  }
  return n;
}

// Global offset (in doubles) of LDS slot `sl` in monitor m's [f][cells][2]
// array; m is found from the bases.
template <typename T>
__device__ __forceinline__ int64_t tb_slot_global(const Side2<T>& f, const TbClip* cl, int sl,
                                                  int& m) {
  m = 0;
  while (m + 1 < f.n_mon && sl >= cl[m + 1].base) ++m;
  const TbClip& k = cl[m];
  const int w = k.c1 - k.c0, h = k.r1 - k.r0;
  const int o = sl - k.base;
  const int kf = o / (w * h), rem = o - kf * w * h;
  const int r = k.r0 + rem / w, c = k.c0 + rem % w;
  const Box& b = f.mb[m];
  const int64_t q = int64_t(r - b.y0) * b.nx + (c - b.x0);
  return (int64_t(kf) * b.cells() + q) * 2;
}

// This thread's cells inside the monitors' clipped boxes (among the rows
// set in `rows`): bit j * kMaxMon2 + m for row j and monitor m.
template <int TBR>
__device__ __forceinline__ unsigned tb_mon_cells(const TbClip* clip, int n_mon, unsigned rows,
                                                 int r0, int c) {
  unsigned bits = 0;
  for (int m = 0; m < n_mon; ++m) {
    const TbClip k = clip[m];
#pragma unroll
    for (int j = 0; j < TBR; ++j) {
      const int r = r0 + j;
      const bool in = ((rows >> j) & 1u) && r >= k.r0 && r < k.r1 && c >= k.c0 && c < k.c1;
      bits |= unsigned(in) << (j * kMaxMon2 + m);
    }
  }
  return bits;
}

template <typename T, int HS, int RW>
__global__ __launch_bounds__(1024) void k2tm_tb_fwd(G2<T> g, const T* __restrict__ sin,
                                                    T* __restrict__ sout, T* __restrict__ es,
                                                    Side2<T> f, int n0, int ks) {
  constexpr int TBW = RW, TBH = RW, TBRB = 1024 / RW, TBR = RW * RW / 1024;
  static_assert(TBR >= 1 && TBRB * TBR == TBH, "square regions of 32 or 64 cells");
  __shared__ T A[TBH * TBW];     // Ez
  __shared__ T B[TBH * TBW];     // Hx
  __shared__ T Ch[TBRB * TBW];   // Hy of each row block's last row
  __shared__ double PH[kMaxMon2 * HS * TB_NF * 2];
  __shared__ double LA[TB_SLOTS * 2];
  __shared__ T WV[HS];
  const int ny = g.ny, nx = g.nx;
  const int64_t N = g.n;
  const int tid = threadIdx.x;
  const TbTile t = tb_tile<HS, RW>(ny, nx);
  const int cl = tid % TBW, rb = tid / TBW;
  const int rl0 = rb * TBR;
  const int c = t.ci0 + cl;
  const bool cv = c < nx;
  const bool cown = c >= t.oc0 && c < t.oc1;
  const int dxr = cl + 1 < TBW ? 1 : 0, dxl = cl > 0 ? -1 : 0;
  const T bxv = cv ? g.bx[c] : T(0), axv = cv ? g.ax[c] : T(0);
  const T kh = g.kh;
  // staging: waveform, phases, the tile's accumulators
  if (tid < ks) WV[tid] = f.has_src ? f.swave[n0 + tid] : T(0);
  for (int m = 0; m < f.n_mon; ++m) {
    const int nf = f.nf[m];
    for (int e = tid; e < ks * nf * 2; e += 1024) {
      const int sidx = e / (nf * 2), o = e - sidx * nf * 2;
      PH[(m * HS + sidx) * TB_NF * 2 + o] = f.ph[m][int64_t(n0 + sidx) * nf * 2 + o];
    }
  }
  __shared__ TbClip clip[kMaxMon2];
  __shared__ int nslots_s;
  if (tid == 0) nslots_s = tb_clip(f, t.or0, t.or1, t.oc0, t.oc1, clip);
  __syncthreads();
  const int nslots = nslots_s;
  const bool lds_acc = nslots <= TB_SLOTS;
  if (lds_acc) {
    for (int sl = tid; sl < nslots; sl += 1024) {
      int m;
      const int64_t o = tb_slot_global(f, clip, sl, m);
      LA[2 * sl] = f.acc[m][o];
      LA[2 * sl + 1] = f.acc[m][o + 1];
    }
  }
  // this thread's cells in registers; A and B carry the values its
  // neighbours read
  T ezr[TBR], hxr[TBR], hy[TBR], cer[TBR], byr[TBR], ayr[TBR], spr[TBR];
  unsigned own = 0, srcm = 0;
  T* const Ar = A + rl0 * TBW + cl;
  T* const Br = B + rl0 * TBW + cl;
#pragma unroll
  for (int j = 0; j < TBR; ++j) {
    const int r = t.ri0 + rl0 + j;
    const bool v = cv && r < ny;
    const int64_t i = int64_t(r) * nx + c;
    ezr[j] = v ? sin[i] : T(0);
    hxr[j] = v ? sin[N + i] : T(0);
    Ar[j * TBW] = ezr[j];
    hy[j] = v ? sin[2 * N + i] : T(0);
    cer[j] = v ? g.ce[i] : T(0);
    byr[j] = r < ny ? g.by[r] : T(0);
    ayr[j] = r < ny ? g.ay[r] : T(0);
    own |= unsigned(v && cown && r >= t.or0 && r < t.or1) << j;
    int64_t q;
    spr[j] = T(0);
    if (f.has_src && v && in_box2(f.sb, r, c, q)) {
      srcm |= 1u << j;
      spr[j] = f.sprof[q];
    }
  }
  const unsigned monb = tb_mon_cells<TBR>(clip, f.n_mon, own, t.ri0 + rl0, c);
  const T* const Anext = A + min(rl0 + TBR, TBH - 1) * TBW + cl;
  // every load of the prologue lands here: the step loop then waits on no
  // memory counter (a wait there would also wait for the E-store writes)
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int s = 0; s < ks; ++s) {
    // H from E
#pragma unroll
    for (int j = 0; j < TBR; ++j) {
      const int r = t.ri0 + rl0 + j;
      const T e = ezr[j];
      const T er = dxr ? Ar[j * TBW + 1] : e;
      const T eu = j + 1 < TBR ? ezr[j + 1] : *Anext;
      const T dc = (c + 1 < nx) ? er - e : T(0);
      const T dr = (r + 1 < ny) ? eu - e : T(0);
      hxr[j] = byr[j] * hxr[j] - kh * (ayr[j] * dc);
      Br[j * TBW] = hxr[j];
      hy[j] = bxv * hy[j] + kh * (axv * dr);
    }
    Ch[rb * TBW + cl] = hy[TBR - 1];
    __syncthreads();
    // E from H, the source, the E store and the monitors
    const T hbelow = Ch[max(rb - 1, 0) * TBW + cl];
    const T w = WV[s];
#pragma unroll
    for (int j = 0; j < TBR; ++j) {
      const int r = t.ri0 + rl0 + j;
      const T hb = j > 0 ? hy[j - 1] : hbelow;
      const T hxo = hxr[j];
      const T hl = dxl ? Br[j * TBW - 1] : hxo;
      const T dHy = r > 0 ? hy[j] - hb : T(0);
      const T dHx = c > 0 ? hxo - hl : T(0);
      T ez = (bxv * byr[j]) * ezr[j] + cer[j] * (axv * dHy - ayr[j] * dHx);
      if ((srcm >> j) & 1u) ez += w * spr[j];
      ezr[j] = ez;
      Ar[j * TBW] = ez;
      if ((own >> j) & 1u) {
        const int64_t i = int64_t(r) * nx + c;
        if (es) es[int64_t(s) * N + i] = ez;
        if ((monb >> (j * kMaxMon2)) & ((1u << kMaxMon2) - 1u)) {
          const double v = double(ez);
#pragma unroll 1
          for (int m = 0; m < f.n_mon; ++m) {
            if (!((monb >> (j * kMaxMon2 + m)) & 1u)) continue;
            const TbClip k = clip[m];
            const double* ph = PH + (m * HS + s) * TB_NF * 2;
            // two loops, so that the LDS one compiles to LDS (not flat) accesses
            if (lds_acc) {
              const int w_ = k.c1 - k.c0, hw = (k.r1 - k.r0) * w_;
              const int sl = k.base + (r - k.r0) * w_ + (c - k.c0);
              for (int kf = 0; kf < k.nf; ++kf) {
                double* a = LA + 2 * (sl + kf * hw);
                a[0] += v * ph[2 * kf];
                a[1] -= v * ph[2 * kf + 1];
              }
            } else {
              const int64_t cells = f.mb[m].cells();
              const int64_t q = int64_t(r - k.y0) * k.bnx + (c - k.x0);
              for (int kf = 0; kf < k.nf; ++kf) {
                double* a = f.acc[m] + (int64_t(kf) * cells + q) * 2;
                a[0] += v * ph[2 * kf];
                a[1] -= v * ph[2 * kf + 1];
              }
            }
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < TBR; ++j) {
    if (!((own >> j) & 1u)) continue;
    const int r = t.ri0 + rl0 + j;
    const int64_t i = int64_t(r) * nx + c;
    sout[i] = ezr[j];
    sout[N + i] = hxr[j];
    sout[2 * N + i] = hy[j];
  }
  if (lds_acc) {
    for (int sl = tid; sl < nslots; sl += 1024) {
      int m;
      const int64_t o = tb_slot_global(f, clip, sl, m);
      f.acc[m][o] = LA[2 * sl];
      f.acc[m][o + 1] = LA[2 * sl + 1];
    }
  }
}

// Adjoint steps m_top, m_top - 1, ..., m_top - ks + 1 (all >= 1).  E^m of
// step m comes from e_lo (m == m_lo) or es slot m - m_lo - 1.
template <typename T, int HS, int RW>
__global__ __launch_bounds__(1024) void k2tm_tb_adj(G2<T> g, const T* __restrict__ lin,
                                                    T* __restrict__ lout, T* __restrict__ S,
                                                    const T* __restrict__ e_lo,
                                                    const T* __restrict__ es, int m_lo,
                                                    Side2<T> f, int m_top, int ks) {
  constexpr int TBW = RW, TBH = RW, TBRB = 1024 / RW, TBR = RW * RW / 1024;
  static_assert(TBR >= 1 && TBRB * TBR == TBH, "square regions of 32 or 64 cells");
  __shared__ T U[TBH * TBW];     // ce * lambda_E
  __shared__ T X[TBH * TBW];     // mu_x
  __shared__ T Cm[TBRB * TBW];   // mu_y of each row block's last row
  __shared__ double PH[kMaxMon2 * HS * TB_NF * 2];
  __shared__ double GR[TB_SLOTS * 2];
  __shared__ T WV[HS];
  const int ny = g.ny, nx = g.nx;
  const int64_t N = g.n;
  const int tid = threadIdx.x;
  const TbTile t = tb_tile<HS, RW>(ny, nx);
  const int cl = tid % TBW, rb = tid / TBW;
  const int rl0 = rb * TBR;
  const int c = t.ci0 + cl;
  const bool cv = c < nx;
  const bool cown = c >= t.oc0 && c < t.oc1;
  const int dxr = cl + 1 < TBW ? 1 : 0, dxl = cl > 0 ? -1 : 0;
  const T bxv = cv ? g.bx[c] : T(0), axv = cv ? g.ax[c] : T(0);
  const T kh = g.kh;
  // staging: waveform, phases (row m - 1 of step m), the region's gradients
  if (tid < ks) WV[tid] = f.has_src ? f.swave[m_top - tid] : T(0);
  for (int m = 0; m < f.n_mon; ++m) {
    const int nf = f.nf[m];
    for (int e = tid; e < ks * nf * 2; e += 1024) {
      const int sidx = e / (nf * 2), o = e - sidx * nf * 2;
      PH[(m * HS + sidx) * TB_NF * 2 + o] = f.ph[m][int64_t(m_top - sidx - 1) * nf * 2 + o];
    }
  }
  __shared__ TbClip clip[kMaxMon2];
  __shared__ int nslots_s;
  if (tid == 0)
    nslots_s = tb_clip(f, t.ri0, min(t.ri0 + TBH, ny), t.ci0, min(t.ci0 + TBW, nx), clip);
  __syncthreads();
  const int nslots = nslots_s;
  const bool lds_grd = nslots <= TB_SLOTS;
  if (lds_grd) {
    for (int sl = tid; sl < nslots; sl += 1024) {
      int m;
      const int64_t o = tb_slot_global(f, clip, sl, m);
      GR[2 * sl] = f.grd[m][o];
      GR[2 * sl + 1] = f.grd[m][o + 1];
    }
  }
  T L[TBR], mx[TBR], my[TBR], Sr[TBR], cer[TBR], byr[TBR], ayr[TBR], spr[TBR], emr[TBR];
  unsigned own = 0, srcm = 0, valid = 0;
  T* const Ur = U + rl0 * TBW + cl;
  T* const Xr = X + rl0 * TBW + cl;
  {
    const T* em0 = (m_top == m_lo) ? e_lo : es + int64_t(m_top - m_lo - 1) * N;
#pragma unroll
    for (int j = 0; j < TBR; ++j) {
      const int r = t.ri0 + rl0 + j;
      const bool v = cv && r < ny;
      const int64_t i = int64_t(r) * nx + c;
      L[j] = v ? lin[i] : T(0);
      mx[j] = v ? lin[N + i] : T(0);
      my[j] = v ? lin[2 * N + i] : T(0);
      cer[j] = v ? g.ce[i] : T(0);
      byr[j] = r < ny ? g.by[r] : T(0);
      ayr[j] = r < ny ? g.ay[r] : T(0);
      const bool o = v && cown && r >= t.or0 && r < t.or1;
      own |= unsigned(o) << j;
      valid |= unsigned(v) << j;
      Sr[j] = o ? S[i] : T(0);
      emr[j] = o ? em0[i] : T(0);
      int64_t q;
      spr[j] = T(0);
      if (f.has_src && v && in_box2(f.sb, r, c, q)) {
        srcm |= 1u << j;
        spr[j] = f.sprof[q];
      }
    }
  }
  const unsigned monb = tb_mon_cells<TBR>(clip, f.n_mon, valid, t.ri0 + rl0, c);
  const T* const Unext = U + min(rl0 + TBR, TBH - 1) * TBW + cl;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int s = 0; s < ks; ++s) {
    const int m = m_top - s;
    // the source's adjoint term (tile cells), then ce * lambda for the H update
    const T w = WV[s];
    T ur[TBR];
#pragma unroll
    for (int j = 0; j < TBR; ++j) {
      if (((srcm & own) >> j) & 1u) Sr[j] -= w * spr[j] * L[j];
      ur[j] = cer[j] * L[j];
      Ur[j * TBW] = ur[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < TBR; ++j) {
      const int r = t.ri0 + rl0 + j;
      const T u = ur[j];
      const T uc = dxr ? Ur[j * TBW + 1] : u;
      const T uu = j + 1 < TBR ? ur[j + 1] : *Unext;
      const T tc = (c >= 1 ? u : T(0)) - (c + 1 < nx ? uc : T(0));
      const T tr = (r >= 1 ? u : T(0)) - (r + 1 < ny ? uu : T(0));
      mx[j] = byr[j] * mx[j] - ayr[j] * tc;
      Xr[j * TBW] = mx[j];
      my[j] = bxv * my[j] + axv * tr;
    }
    Cm[rb * TBW + cl] = my[TBR - 1];
    __syncthreads();
    const T mbelow = Cm[max(rb - 1, 0) * TBW + cl];
#pragma unroll
    for (int j = 0; j < TBR; ++j) {
      const int r = t.ri0 + rl0 + j;
      const T mb = j > 0 ? my[j - 1] : mbelow;
      const T r_my = (r >= 1 ? mb : T(0)) - (r <= ny - 2 ? my[j] : T(0));
      const T xl = dxl ? Xr[j * TBW - 1] : mx[j];
      const T c_mx = (c >= 1 ? xl : T(0)) - (c <= nx - 2 ? mx[j] : T(0));
      const T tt = kh * (axv * r_my - ayr[j] * c_mx);
      L[j] = (bxv * byr[j]) * L[j] + tt;
      const bool o = (own >> j) & 1u;
      const T emv = emr[j];
      if (o) Sr[j] += emv * tt;
      if (!((monb >> (j * kMaxMon2)) & ((1u << kMaxMon2) - 1u))) continue;
#pragma unroll 1
      for (int mm = 0; mm < f.n_mon; ++mm) {
        if (!((monb >> (j * kMaxMon2 + mm)) & 1u)) continue;
        const TbClip k = clip[mm];
        const double* ph = PH + (mm * HS + s) * TB_NF * 2;
        double d = 0.0;
        if (lds_grd) {
          const int w_ = k.c1 - k.c0, hw = (k.r1 - k.r0) * w_;
          const int sl = k.base + (r - k.r0) * w_ + (c - k.c0);
          for (int kf = 0; kf < k.nf; ++kf) {
            const double* gg = GR + 2 * (sl + kf * hw);
            d += gg[0] * ph[2 * kf] - gg[1] * ph[2 * kf + 1];
          }
        } else {
          const int64_t cells = f.mb[mm].cells();
          const int64_t q = int64_t(r - k.y0) * k.bnx + (c - k.x0);
          for (int kf = 0; kf < k.nf; ++kf) {
            const double* gg = f.grd[mm] + (int64_t(kf) * cells + q) * 2;
            d += gg[0] * ph[2 * kf] - gg[1] * ph[2 * kf + 1];
          }
        }
        const T dv = T(d);
        L[j] += dv;
        if (o) {
          T sacc = T(0);
          sacc += emv * dv;
          Sr[j] += sacc;
        }
      }
    }
    // E^m of the next step, ahead of its use
    if (s + 1 < ks) {
      const int mn = m - 1;
      const T* emn = (mn == m_lo) ? e_lo : es + int64_t(mn - m_lo - 1) * N;
#pragma unroll
      for (int j = 0; j < TBR; ++j) {
        const int r = t.ri0 + rl0 + j;
        emr[j] = ((own >> j) & 1u) ? emn[int64_t(r) * nx + c] : T(0);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < TBR; ++j) {
    if (!((own >> j) & 1u)) continue;
    const int r = t.ri0 + rl0 + j;
    const int64_t i = int64_t(r) * nx + c;
    lout[i] = L[j];
    lout[N + i] = mx[j];
    lout[2 * N + i] = my[j];
    S[i] = Sr[j];
  }
}

// Launch shape of the temporally blocked TM kernels: steps per launch, halo
// and region width, or steps = 0 when the step kernels run instead (TE,
// more than one source or kMaxMon2 monitors, a monitor of more than TB_NF
// frequencies, a probe, no second buffer, or DNX_OPT_FDTD2D_TB_STEPS = 1).
// The launches are latency-bound while the tiles fit the chip once, so by
// default: 32 x 32 regions (16 x 16 tiles, 8 steps) when at most 256 tiles
// cover the grid, else 64 x 64 regions with 16-cell halos (16 steps) on up
// to 512 tiles of 32 x 32, else 64 x 64 regions with 8-cell halos (8 steps,
// 48 x 48 tiles: 1.8x recompute instead of 4x).  DNX_OPT_FDTD2D_TB_STEPS = k
// in 2..16 forces 64 x 64 regions and k steps (halo 8 up to 8 steps, 16
// above); DNX_OPT_FDTD2D_TB_REGION = 32 or 64 forces the region width.
struct TbShape {
  int steps, hs, rw;
};

static TbShape tb_shape(const dnx_grid2d* gd, int n_src, const dnx_monitor* mons, int n_mon,
                        int64_t probe, const void* a, const void* b) {
  const int opt = option(DNX_OPT_FDTD2D_TB_STEPS);
  const int region = option(DNX_OPT_FDTD2D_TB_REGION);
  if (opt == 1 || probe >= 0 || b == nullptr || a == b || !side2_ok(gd, n_src, mons, n_mon))
    return TbShape{0, 0, 0};
  for (int j = 0; j < n_mon; ++j)
    if (mons[j].n_freq > TB_NF) return TbShape{0, 0, 0};
  if (opt > 1) {
    const int k = std::min(opt, TBS);
    return TbShape{k, k > 8 ? 16 : 8, 64};
  }
  const int64_t t16 = int64_t(cdiv(gd->nx, 16)) * cdiv(gd->ny, 16);
  const int64_t t32 = int64_t(cdiv(gd->nx, 32)) * cdiv(gd->ny, 32);
  if (region == 32 || (region != 64 && t16 <= 256)) return TbShape{8, 8, 32};
  return t32 <= 512 ? TbShape{16, 16, 64} : TbShape{8, 8, 64};
}

static dim3 tb_grid(const dnx_grid2d* g, int hs, int rw) {
  return dim3(cdiv(g->nx, rw - 2 * hs), cdiv(g->ny, rw - 2 * hs));
}

// ---- TE forward ----------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2te_e(G2<T> g, T* __restrict__ st) {
  DNX_2D_PROLOGUE
  const T* Hz = st;
  T* Ex = st + N;
  T* Ey = st + 2 * N;
  const T hz = Hz[i];
  const T dc = c > 0 ? hz - Hz[i - 1] : T(0);
  const T dr = r > 0 ? hz - Hz[i - nx] : T(0);
  const T ce = g.ce[i];
  Ex[i] = g.bx[c] * Ex[i] + ce * (g.ax[c] * dc);
  Ey[i] = g.by[r] * Ey[i] - ce * (g.ay[r] * dr);
}

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2te_h(G2<T> g, T* __restrict__ st,
                                                   T* __restrict__ es) {
  DNX_2D_PROLOGUE
  T* Hz = st;
  const T* Ex = st + N;
  const T* Ey = st + 2 * N;
  const T dEy = (r + 1 < g.ny) ? Ey[i + nx] - Ey[i] : T(0);
  const T dEx = (c + 1 < g.nx) ? Ex[i + 1] - Ex[i] : T(0);
  const T hz = (g.bx[c] * g.by[r]) * Hz[i] + g.kh * (g.ax[c] * dEy - g.ay[r] * dEx);
  Hz[i] = hz;
  if (es) es[i] = hz;
}

// ---- TM adjoint: lam = [lamEz, muHx, muHy] -------------------------------

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2tm_adj_h(G2<T> g, T* __restrict__ lam) {
  DNX_2D_PROLOGUE
  const T* L = lam;
  T* Mx = lam + N;
  T* My = lam + 2 * N;
  const T u = g.ce[i] * L[i];
  // muHx = by muHx - ay D-c^T(u);  muHy = bx muHy + D-r^T(ax u)
  const T tc = (c >= 1 ? u : T(0)) - (c + 1 < g.nx ? g.ce[i + 1] * L[i + 1] : T(0));
  const T tr = (r >= 1 ? u : T(0)) - (r + 1 < g.ny ? g.ce[i + nx] * L[i + nx] : T(0));
  Mx[i] = g.by[r] * Mx[i] - g.ay[r] * tc;
  My[i] = g.bx[c] * My[i] + g.ax[c] * tr;
}

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2tm_adj_e(G2<T> g, T* __restrict__ lam,
                                                       const T* __restrict__ em,
                                                       T* __restrict__ S) {
  DNX_2D_PROLOGUE
  T* L = lam;
  const T* Mx = lam + N;
  const T* My = lam + 2 * N;
  // t = kh (ax D+r^T(muHy) - ay D+c^T(muHx))
  const T r_my = (r >= 1 ? My[i - nx] : T(0)) - (r <= g.ny - 2 ? My[i] : T(0));
  const T c_mx = (c >= 1 ? Mx[i - 1] : T(0)) - (c <= g.nx - 2 ? Mx[i] : T(0));
  const T t = g.kh * (g.ax[c] * r_my - g.ay[r] * c_mx);
  L[i] = (g.bx[c] * g.by[r]) * L[i] + t;
  S[i] += em[i] * t;
}

// ---- TE adjoint: lam = [lamHz, muEx, muEy] -------------------------------

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2te_adj_e(G2<T> g, T* __restrict__ lam,
                                                       const T* __restrict__ hzm,
                                                       T* __restrict__ S) {
  DNX_2D_PROLOGUE
  const T* L = lam;
  T* Mx = lam + N;
  T* My = lam + 2 * N;
  // muEx = bx muEx - kh ay D+c^T(lamHz); muEy = by muEy + kh ax D+r^T(lamHz)
  const T l = L[i];
  const T tc = (c >= 1 ? L[i - 1] : T(0)) - (c <= g.nx - 2 ? l : T(0));
  const T tr = (r >= 1 ? L[i - nx] : T(0)) - (r <= g.ny - 2 ? l : T(0));
  const T mx = g.bx[c] * Mx[i] - g.kh * g.ay[r] * tc;
  const T my = g.by[r] * My[i] + g.kh * g.ax[c] * tr;
  Mx[i] = mx;
  My[i] = my;
  if (hzm) {
    // S += ce (muEx ax D-c Hz^m - muEy ay D-r Hz^m)
    const T h = hzm[i];
    const T dc = c > 0 ? h - hzm[i - 1] : T(0);
    const T dr = r > 0 ? h - hzm[i - nx] : T(0);
    S[i] += g.ce[i] * (mx * g.ax[c] * dc - my * g.ay[r] * dr);
  }
}

template <typename T>
__global__ __launch_bounds__(BX2* BY2) void k2te_adj_h(G2<T> g, T* __restrict__ lam) {
  DNX_2D_PROLOGUE
  T* L = lam;
  const T* Mx = lam + N;
  const T* My = lam + 2 * N;
  // lamHz = bxby lamHz + D-c^T(ax ce muEx) - D-r^T(ay ce muEy)
  const T vx = g.ce[i] * Mx[i];
  const T vy = g.ce[i] * My[i];
  const T tc = (c >= 1 ? g.ax[c] * vx : T(0)) -
               (c + 1 < g.nx ? g.ax[c + 1] * g.ce[i + 1] * Mx[i + 1] : T(0));
  const T tr = (r >= 1 ? g.ay[r] * vy : T(0)) -
               (r + 1 < g.ny ? g.ay[r + 1] * g.ce[i + nx] * My[i + nx] : T(0));
  L[i] = (g.bx[c] * g.by[r]) * L[i] + tc - tr;
}

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------

static int check_grid2(const dnx_grid2d* g) {
  DNX_REQUIRE(g != nullptr, "grid2d: null");
  DNX_REQUIRE(g->ny > 0 && g->nx > 0, "grid2d: bad shape %d x %d", g->ny, g->nx);
  DNX_REQUIRE(g->dtype == DNX_F32 || g->dtype == DNX_F64, "grid2d: bad dtype %d", g->dtype);
  DNX_REQUIRE(g->te == 0 || g->te == 1, "grid2d: te must be 0 or 1");
  DNX_REQUIRE(g->ce && g->by && g->ay && g->bx && g->ax, "grid2d: null coefficient array");
  return DNX_OK;
}

static dim3 grid2_launch(const dnx_grid2d* g) { return dim3(cdiv(g->nx, BX2), cdiv(g->ny, BY2)); }

template <typename T>
static int run2(const dnx_grid2d* gd, T* st, T* work, int n0, int n_steps,
                const dnx_source* srcs, int n_src, const dnx_monitor* mons, int n_mon, T* ckpt,
                int ckpt_every, T* es, int64_t probe, T* probe_out, hipStream_t s,
                int32_t* final_in_work) {
  G2<T> g = make_g2<T>(gd);
  const dim3 grid = grid2_launch(gd), block(BX2, BY2);
  const int64_t N = g.n;
  Comps<T> F{{st, nullptr, nullptr}};
  const bool fuse = side2_ok(gd, n_src, mons, n_mon);
  const Side2<T> side = fuse ? make_side2<T>(srcs, n_src, mons, n_mon, false) : Side2<T>{};
  const TbShape tbs = tb_shape(gd, n_src, mons, n_mon, probe, st, work);
  if (const int tb = tbs.steps) {
    // temporally blocked: launches of up to tb steps, ping-ponging between
    // st and work, none crossing a checkpoint
    T* cur = st;
    T* nxt = work;
    for (int k = 0; k < n_steps;) {
      if (ckpt && k % ckpt_every == 0) {
        DNX_CHECK_HIP(hipMemcpyAsync(ckpt + int64_t(k / ckpt_every) * 3 * N, cur,
                                     sizeof(T) * 3 * N, hipMemcpyDeviceToDevice, s));
      }
      int ks = std::min(tb, n_steps - k);
      if (ckpt) ks = std::min(ks, ckpt_every - k % ckpt_every);
      T* es_k = es ? es + int64_t(k) * N : nullptr;
      const dim3 grid_tb = tb_grid(gd, tbs.hs, tbs.rw);
      if (tbs.rw == 32)
        hipLaunchKernelGGL((k2tm_tb_fwd<T, 8, 32>), grid_tb, dim3(1024), 0, s, g, cur, nxt,
                           es_k, side, n0 + k, ks);
      else if (tbs.hs == 16)
        hipLaunchKernelGGL((k2tm_tb_fwd<T, 16, 64>), grid_tb, dim3(1024), 0, s, g, cur, nxt,
                           es_k, side, n0 + k, ks);
      else
        hipLaunchKernelGGL((k2tm_tb_fwd<T, 8, 64>), grid_tb, dim3(1024), 0, s, g, cur, nxt,
                           es_k, side, n0 + k, ks);
      DNX_CHECK_HIP(hipGetLastError());
      std::swap(cur, nxt);
      k += ks;
    }
    if (final_in_work) *final_in_work = cur != st ? 1 : 0;
    return DNX_OK;
  }
  for (int k = 0; k < n_steps; ++k) {
    const int n = n0 + k;
    if (ckpt && k % ckpt_every == 0) {
      DNX_CHECK_HIP(hipMemcpyAsync(ckpt + int64_t(k / ckpt_every) * 3 * N, st,
                                   sizeof(T) * 3 * N, hipMemcpyDeviceToDevice, s));
    }
    T* slot = es ? es + int64_t(k) * N : nullptr;
    if (fuse) {
      hipLaunchKernelGGL(k2tm_h<T>, grid, block, 0, s, g, st);
      hipLaunchKernelGGL(k2tm_e_side<T>, grid, block, 0, s, g, st, slot, side, n);
      DNX_CHECK_HIP(hipGetLastError());
      if (probe >= 0) {
        DNX_CHECK_HIP(hipMemcpyAsync(probe_out + k, st + probe, sizeof(T),
                                     hipMemcpyDeviceToDevice, s));
      }
      continue;
    }
    if (gd->te) {
      hipLaunchKernelGGL(k2te_e<T>, grid, block, 0, s, g, st);
      hipLaunchKernelGGL(k2te_h<T>, grid, block, 0, s, g, st, slot);
    } else {
      hipLaunchKernelGGL(k2tm_h<T>, grid, block, 0, s, g, st);
      hipLaunchKernelGGL(k2tm_e<T>, grid, block, 0, s, g, st, slot);
    }
    DNX_CHECK_HIP(hipGetLastError());
    for (int j = 0; j < n_src; ++j) {
      int rc = launch_source<T>(st, g.ny, g.nx, srcs[j], n, s);
      if (rc) return rc;
      if (slot && (rc = launch_source<T>(slot, g.ny, g.nx, srcs[j], n, s))) return rc;
    }
    if (probe >= 0) {
      DNX_CHECK_HIP(hipMemcpyAsync(probe_out + k, st + probe, sizeof(T), hipMemcpyDeviceToDevice,
                                   s));
    }
    for (int j = 0; j < n_mon; ++j) {
      int rc = launch_monitor<T>(F, g.ny, g.nx, mons[j], n, s);
      if (rc) return rc;
    }
  }
  return DNX_OK;
}

template <typename T>
static int adj_init2(const dnx_grid2d* gd, T* lam, const T* eT, const dnx_monitor* mons,
                     int n_mon, int n_total, T* S, hipStream_t s) {
  G2<T> g = make_g2<T>(gd);
  const int64_t N = g.n;
  Comps<T> L{{lam, nullptr, nullptr}};
  Comps<T> none{{nullptr, nullptr, nullptr}};
  for (int j = 0; j < n_mon; ++j) {
    int rc = launch_monitor_adj<T>(L, none, nullptr, g.ny, g.nx, mons[j], n_total - 1, s);
    if (rc) return rc;
  }
  int rc = launch_zero<T>(lam + N, 2 * N, s);
  if (rc) return rc;
  if (gd->te) return launch_zero<T>(S, N, s);
  Comps<T> Ev{{const_cast<T*>(eT), nullptr, nullptr}};
  return launch_dot_init<T>(S, Ev, L, 1, N, s);
}

template <typename T>
static int adj_run2(const dnx_grid2d* gd, T* lam, T* work, int m_hi, int m_lo, const T* e_lo,
                    const T* es, const dnx_source* srcs, int n_src, const dnx_monitor* mons,
                    int n_mon, T* S, hipStream_t s, int32_t* final_in_work) {
  G2<T> g = make_g2<T>(gd);
  const dim3 grid = grid2_launch(gd), block(BX2, BY2);
  const int64_t N = g.n;
  const bool fuse = side2_ok(gd, n_src, mons, n_mon);
  const Side2<T> side = fuse ? make_side2<T>(srcs, n_src, mons, n_mon, true) : Side2<T>{};
  const TbShape tbs = tb_shape(gd, n_src, mons, n_mon, -1, lam, work);
  if (const int tb = tbs.steps) {
    T* cur = lam;
    T* nxt = work;
    const int m_stop = std::max(m_lo, 1);
    for (int m = m_hi - 1; m >= m_stop;) {
      const int ks = std::min(tb, m - m_stop + 1);
      const dim3 grid_tb = tb_grid(gd, tbs.hs, tbs.rw);
      if (tbs.rw == 32)
        hipLaunchKernelGGL((k2tm_tb_adj<T, 8, 32>), grid_tb, dim3(1024), 0, s, g, cur, nxt, S,
                           e_lo, es, m_lo, side, m, ks);
      else if (tbs.hs == 16)
        hipLaunchKernelGGL((k2tm_tb_adj<T, 16, 64>), grid_tb, dim3(1024), 0, s, g, cur, nxt, S,
                           e_lo, es, m_lo, side, m, ks);
      else
        hipLaunchKernelGGL((k2tm_tb_adj<T, 8, 64>), grid_tb, dim3(1024), 0, s, g, cur, nxt, S,
                           e_lo, es, m_lo, side, m, ks);
      DNX_CHECK_HIP(hipGetLastError());
      std::swap(cur, nxt);
      m -= ks;
    }
    if (m_lo == 0) {
      for (int j = 0; j < n_src; ++j) {
        int rc = launch_source_adj<T>(S, cur, g.ny, g.nx, srcs[j], 0, s);
        if (rc) return rc;
      }
    }
    if (final_in_work) *final_in_work = cur != lam ? 1 : 0;
    return DNX_OK;
  }
  Comps<T> L{{lam, nullptr, nullptr}};
  for (int m = m_hi - 1; m >= m_lo && m >= 1; --m) {
    const T* em = (m == m_lo) ? e_lo : es + int64_t(m - m_lo - 1) * N;
    Comps<T> Em{{const_cast<T*>(em), nullptr, nullptr}};
    if (fuse) {
      hipLaunchKernelGGL(k2tm_adj_h_side<T>, grid, block, 0, s, g, lam, S, side, m);
      hipLaunchKernelGGL(k2tm_adj_e_side<T>, grid, block, 0, s, g, lam, em, S, side, m);
      DNX_CHECK_HIP(hipGetLastError());
      continue;
    }
    if (gd->te) {
      hipLaunchKernelGGL(k2te_adj_e<T>, grid, block, 0, s, g, lam, em, S);
      hipLaunchKernelGGL(k2te_adj_h<T>, grid, block, 0, s, g, lam);
      DNX_CHECK_HIP(hipGetLastError());
      Comps<T> none{{nullptr, nullptr, nullptr}};
      for (int j = 0; j < n_mon; ++j) {
        int rc = launch_monitor_adj<T>(L, none, nullptr, g.ny, g.nx, mons[j], m - 1, s);
        if (rc) return rc;
      }
    } else {
      for (int j = 0; j < n_src; ++j) {
        int rc = launch_source_adj<T>(S, lam, g.ny, g.nx, srcs[j], m, s);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(k2tm_adj_h<T>, grid, block, 0, s, g, lam);
      hipLaunchKernelGGL(k2tm_adj_e<T>, grid, block, 0, s, g, lam, em, S);
      DNX_CHECK_HIP(hipGetLastError());
      for (int j = 0; j < n_mon; ++j) {
        int rc = launch_monitor_adj<T>(L, Em, S, g.ny, g.nx, mons[j], m - 1, s);
        if (rc) return rc;
      }
    }
  }
  if (m_lo == 0 && !gd->te) {
    for (int j = 0; j < n_src; ++j) {
      int rc = launch_source_adj<T>(S, lam, g.ny, g.nx, srcs[j], 0, s);
      if (rc) return rc;
    }
  }
  return DNX_OK;
}

// ---------------------------------------------------------------------------
// Batched TM forward over many designs of one shape (blockIdx.z = design),
// source shared, with a running DFT over a box accumulated in FP32 inside
// the E kernel.  Used to generate surrogate training data and to evaluate
// candidate designs; forward only.
// ---------------------------------------------------------------------------

struct Tm2Batch {
  int ny, nx;
  int64_t n;
  float kh;
  const float* __restrict__ ce;  // [B][n]
  const float* __restrict__ by;
  const float* __restrict__ ay;
  const float* __restrict__ bx;
  const float* __restrict__ ax;
  float* __restrict__ st;  // [B][3][n]
  int has_src;
  Box sbox;
  const float* __restrict__ sprof;
  const float* __restrict__ wave;
  Box mbox;
  int nf;
  const float* __restrict__ phase;  // [T][nf][2] (cos, sin)
  float* __restrict__ acc;          // [B][nf][2][mbox cells]
};

__global__ __launch_bounds__(BX2* BY2) void k2tm_batch_h(Tm2Batch a) {
  const int c = blockIdx.x * BX2 + threadIdx.x;
  const int r = blockIdx.y * BY2 + threadIdx.y;
  if (c >= a.nx || r >= a.ny) return;
  const int64_t n = a.n, nx = a.nx;
  float* st = a.st + int64_t(blockIdx.z) * 3 * n;
  const int64_t i = int64_t(r) * nx + c;
  const float ez = st[i];
  const float dc = (c + 1 < a.nx) ? st[i + 1] - ez : 0.f;
  const float dr = (r + 1 < a.ny) ? st[i + nx] - ez : 0.f;
  st[n + i] = a.by[r] * st[n + i] - a.kh * (a.ay[r] * dc);
  st[2 * n + i] = a.bx[c] * st[2 * n + i] + a.kh * (a.ax[c] * dr);
}

__global__ __launch_bounds__(BX2* BY2) void k2tm_batch_e(Tm2Batch a, int step) {
  const int c = blockIdx.x * BX2 + threadIdx.x;
  const int r = blockIdx.y * BY2 + threadIdx.y;
  if (c >= a.nx || r >= a.ny) return;
  const int64_t n = a.n, nx = a.nx;
  const int b = blockIdx.z;
  float* st = a.st + int64_t(b) * 3 * n;
  const int64_t i = int64_t(r) * nx + c;
  const float* Hx = st + n;
  const float* Hy = st + 2 * n;
  const float dHy = r > 0 ? Hy[i] - Hy[i - nx] : 0.f;
  const float dHx = c > 0 ? Hx[i] - Hx[i - 1] : 0.f;
  float ez = (a.bx[c] * a.by[r]) * st[i] +
             a.ce[int64_t(b) * n + i] * (a.ax[c] * dHy - a.ay[r] * dHx);
  if (a.has_src) {
    const Box& s = a.sbox;
    const int dr = r - s.y0, dc = c - s.x0;
    if (dr >= 0 && dr < s.ny && dc >= 0 && dc < s.nx) ez += a.wave[step] * a.sprof[dr * s.nx + dc];
  }
  st[i] = ez;
  const Box& m = a.mbox;
  const int mr = r - m.y0, mc = c - m.x0;
  if (a.nf > 0 && mr >= 0 && mr < m.ny && mc >= 0 && mc < m.nx) {
    const int64_t cells = int64_t(m.ny) * m.nx;
    const int64_t j = int64_t(mr) * m.nx + mc;
    const float* ph = a.phase + int64_t(step) * a.nf * 2;
    float* acc = a.acc + int64_t(b) * a.nf * 2 * cells;
    for (int f = 0; f < a.nf; ++f) {
      acc[(2 * f) * cells + j] += ez * ph[2 * f];
      acc[(2 * f + 1) * cells + j] -= ez * ph[2 * f + 1];
    }
  }
}

// ---------------------------------------------------------------------------
// Resident batched TM kernel for grids up to 128 x 128: one work-group of
// 1024 threads per design advances it through every step of the call (thread
// = one column c < 128 and a block of 16 consecutive rows), so a step costs
// two barriers instead of two kernel launches over the whole batch.  Ez (A)
// and Hx (B) of the whole design live in LDS, Hy and the coefficient in
// registers; the last Hy row of every row block (C) is exchanged through
// LDS.  H phase: read Ez and its x+1 / y+1 neighbours from A, update B and
// Hy in place; E phase: read Hx and its x-1 neighbour from B, Hy of the row
// below (register or C), update A in place.  The arithmetic and its order
// are k2tm_batch_h / k2tm_batch_e's (bitwise their results), including the
// FP32 DFT accumulation.
// ---------------------------------------------------------------------------

constexpr int RES_NX = 128;              // LDS row pitch and largest nx
constexpr int RES_R = 16;                // rows per thread
constexpr int RES_RB = 1024 / RES_NX;    // row blocks (8): largest ny = 128
constexpr int RES_NY = RES_RB * RES_R;
constexpr int RES_MON_SLOTS = 2048;      // DFT accumulators (complex) staged in LDS
constexpr int RES_SRC_CELLS = 512;       // source profile values staged in LDS
constexpr size_t RES_SMEM =
    sizeof(float) * (size_t(2) * RES_NY * RES_NX + RES_RB * RES_NX + 2 * RES_NY +
                     2 * RES_MON_SLOTS + RES_SRC_CELLS);

__global__ __launch_bounds__(1024) void k2tm_batch_res(Tm2Batch a, int n_steps) {
  extern __shared__ float rsm[];
  float* A = rsm;                            // Ez [row][RES_NX]
  float* Bh = A + RES_NY * RES_NX;           // Hx [row][RES_NX]
  float* Ch = Bh + RES_NY * RES_NX;          // Hy of each row block's last row
  float* Py = Ch + RES_RB * RES_NX;          // (by, ay) per row
  float* MLA = Py + 2 * RES_NY;              // the design's DFT accumulators (re, im)
  float* SPL = MLA + 2 * RES_MON_SLOTS;      // the source profile
  const int ny = a.ny, nx = a.nx;
  const int64_t n = a.n;
  const int b = blockIdx.x;
  const int c = threadIdx.x % RES_NX;
  const int rb = threadIdx.x / RES_NX;
  const int r0 = rb * RES_R;
  const bool cv = c < nx;
  const int cr = min(c + 1, RES_NX - 1), cl = max(c - 1, 0);  // clamped LDS columns
  float* st = a.st + int64_t(b) * 3 * n;
  const float* ce = a.ce + int64_t(b) * n;
  for (int i = threadIdx.x; i < RES_NY; i += 1024) {
    Py[2 * i] = i < ny ? a.by[i] : 0.f;
    Py[2 * i + 1] = i < ny ? a.ay[i] : 0.f;
  }
  const float bxv = cv ? a.bx[c] : 0.f, axv = cv ? a.ax[c] : 0.f;
  const float kh = a.kh;
  float hy[RES_R], cer[RES_R];
#pragma unroll
  for (int j = 0; j < RES_R; ++j) {
    const int r = r0 + j;
    const bool v = cv && r < ny;
    const int64_t i = int64_t(r) * nx + c;
    A[r * RES_NX + c] = v ? st[i] : 0.f;
    Bh[r * RES_NX + c] = v ? st[n + i] : 0.f;
    hy[j] = v ? st[2 * n + i] : 0.f;
    cer[j] = v ? ce[i] : 0.f;
  }
  const bool src = a.has_src != 0;
  const Box sb = a.sbox, mb = a.mbox;
  const int64_t mcells = int64_t(mb.ny) * mb.nx;
  float* acc = a.nf > 0 ? a.acc + int64_t(b) * a.nf * 2 * mcells : nullptr;
  // whether this thread owns any source or monitor cell
  const bool own_src = src && cv && c >= sb.x0 && c < sb.x0 + sb.nx && r0 < sb.y0 + sb.ny &&
                       r0 + RES_R > sb.y0;
  const bool own_mon = a.nf > 0 && cv && c >= mb.x0 && c < mb.x0 + mb.nx &&
                       r0 < mb.y0 + mb.ny && r0 + RES_R > mb.y0;
  // rows of this thread that hold source / monitor cells (empty ranges
  // when the thread owns none)
  const int sj0 = own_src ? max(sb.y0 - r0, 0) : 0;
  const int sj1 = own_src ? min(sb.y0 + sb.ny - r0, RES_R) : 0;
  const int mj0 = own_mon ? max(mb.y0 - r0, 0) : 0;
  const int mj1 = own_mon ? min(mb.y0 + mb.ny - r0, RES_R) : 0;
  float* const Ar = A + r0 * RES_NX + c;    // this thread's column, row r0
  float* const Br = Bh + r0 * RES_NX + c;
  const int dx = cr - c, dxl = cl - c;      // 1 / -1, or 0 at a clamped edge
  // row below the block (clamped into the array; only used where r + 1 < ny)
  const float* const Anext = A + min(r0 + RES_R, RES_NY - 1) * RES_NX + c;
  // The DFT accumulators and the source profile live in LDS for the whole
  // call when they fit (loaded here, accumulators stored after the last
  // step), so that the step loop reads no global memory for them; per
  // accumulator the additions are those of the global-memory path, in the
  // same order.
  const bool lds_mon = a.nf > 0 && int64_t(a.nf) * mcells <= RES_MON_SLOTS;
  const bool lds_src = src && sb.cells() <= RES_SRC_CELLS;
  if (lds_mon) {
    for (int e = threadIdx.x; e < a.nf * mcells; e += 1024) {
      const int f = e / int(mcells), q = e - f * int(mcells);
      MLA[2 * e] = acc[(2 * f) * mcells + q];
      MLA[2 * e + 1] = acc[(2 * f + 1) * mcells + q];
    }
  }
  if (lds_src) {
    for (int e = threadIdx.x; e < sb.cells(); e += 1024) SPL[e] = a.sprof[e];
  }
  __syncthreads();
  for (int k = 0; k < n_steps; ++k) {
    // H from E (k2tm_batch_h)
#pragma unroll
    for (int j = 0; j < RES_R; ++j) {
      const int r = r0 + j;
      const float by = Py[2 * r], ay = Py[2 * r + 1];
      const float e = Ar[j * RES_NX];
      const float er = Ar[j * RES_NX + dx];
      const float eu = j + 1 < RES_R ? Ar[(j + 1) * RES_NX] : *Anext;
      const float dc = (c + 1 < nx) ? er - e : 0.f;
      const float dr = (r + 1 < ny) ? eu - e : 0.f;
      Br[j * RES_NX] = by * Br[j * RES_NX] - kh * (ay * dc);
      hy[j] = bxv * hy[j] + kh * (axv * dr);
    }
    Ch[rb * RES_NX + c] = hy[RES_R - 1];
    __syncthreads();
    // E from H (k2tm_batch_e)
    const float hbelow = Ch[max(rb - 1, 0) * RES_NX + c];
#pragma unroll
    for (int j = 0; j < RES_R; ++j) {
      const int r = r0 + j;
      const float by = Py[2 * r], ay = Py[2 * r + 1];
      const float hb = j > 0 ? hy[j - 1] : hbelow;
      const float hxo = Br[j * RES_NX];
      const float hl = Br[j * RES_NX + dxl];
      const float dHy = r > 0 ? hy[j] - hb : 0.f;
      const float dHx = c > 0 ? hxo - hl : 0.f;
      Ar[j * RES_NX] = (bxv * by) * Ar[j * RES_NX] + cer[j] * (axv * dHy - ay * dHx);
    }
    // the soft source, then the DFT of the new Ez (own cells only)
    if (sj0 < sj1) {
      const float w = a.wave[k];
      if (lds_src) {
        for (int j = sj0; j < sj1; ++j)
          Ar[j * RES_NX] += w * SPL[(r0 + j - sb.y0) * sb.nx + (c - sb.x0)];
      } else {
        for (int j = sj0; j < sj1; ++j)
          Ar[j * RES_NX] += w * a.sprof[(r0 + j - sb.y0) * sb.nx + (c - sb.x0)];
      }
    }
    if (mj0 < mj1) {
      const float* ph = a.phase + int64_t(k) * a.nf * 2;
      // two loops, so that the LDS one compiles to LDS (not flat) accesses
      if (lds_mon) {
        for (int j = mj0; j < mj1; ++j) {
          const float e = Ar[j * RES_NX];
          const int q = (r0 + j - mb.y0) * mb.nx + (c - mb.x0);
          for (int f = 0; f < a.nf; ++f) {
            MLA[2 * (f * int(mcells) + q)] += e * ph[2 * f];
            MLA[2 * (f * int(mcells) + q) + 1] -= e * ph[2 * f + 1];
          }
        }
      } else {
        for (int j = mj0; j < mj1; ++j) {
          const float e = Ar[j * RES_NX];
          const int64_t q = int64_t(r0 + j - mb.y0) * mb.nx + (c - mb.x0);
          for (int f = 0; f < a.nf; ++f) {
            acc[(2 * f) * mcells + q] += e * ph[2 * f];
            acc[(2 * f + 1) * mcells + q] -= e * ph[2 * f + 1];
          }
        }
      }
    }
    __syncthreads();
  }
  if (lds_mon) {
    for (int e = threadIdx.x; e < a.nf * mcells; e += 1024) {
      const int f = e / int(mcells), q = e - f * int(mcells);
      acc[(2 * f) * mcells + q] = MLA[2 * e];
      acc[(2 * f + 1) * mcells + q] = MLA[2 * e + 1];
    }
  }
#pragma unroll
  for (int j = 0; j < RES_R; ++j) {
    const int r = r0 + j;
    if (cv && r < ny) {
      const int64_t i = int64_t(r) * nx + c;
      st[i] = A[r * RES_NX + c];
      st[n + i] = Bh[r * RES_NX + c];
      st[2 * n + i] = hy[j];
    }
  }
}

}  // namespace dnx

using namespace dnx;

extern "C" int32_t dnx_fdtd2d_run(const dnx_grid2d* g, void* state, void* work, int32_t n0,
                                  int32_t n_steps, const dnx_source* srcs, int32_t n_src,
                                  const dnx_monitor* mons, int32_t n_mon, void* ckpt,
                                  int32_t ckpt_every, void* estore, int64_t probe_index,
                                  void* probe_out, void* stream, int32_t* final_in_work) {
  int rc = check_grid2(g);
  if (rc) return rc;
  DNX_REQUIRE(state != nullptr, "fdtd2d_run: null state");
  DNX_REQUIRE(n0 >= 0 && n_steps >= 0, "fdtd2d_run: bad step range");
  DNX_REQUIRE(ckpt == nullptr || ckpt_every > 0, "fdtd2d_run: ckpt_every must be > 0");
  if ((rc = check_sources(srcs, n_src, 1, g->ny, g->nx, 0))) return rc;
  if ((rc = check_monitors(mons, n_mon, 1, g->ny, g->nx, 1, false))) return rc;
  DNX_REQUIRE(probe_index < 0 || (probe_out != nullptr &&
                                   probe_index < int64_t(3) * g->ny * g->nx),
              "run: bad probe index %lld", (long long)probe_index);
  if (final_in_work) *final_in_work = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g->dtype == DNX_F32)
    return run2<float>(g, static_cast<float*>(state), static_cast<float*>(work), n0, n_steps,
                       srcs, n_src, mons, n_mon, static_cast<float*>(ckpt), ckpt_every,
                       static_cast<float*>(estore), probe_index,
                       static_cast<float*>(probe_out), s, final_in_work);
  return run2<double>(g, static_cast<double*>(state), static_cast<double*>(work), n0, n_steps,
                      srcs, n_src, mons, n_mon, static_cast<double*>(ckpt), ckpt_every,
                      static_cast<double*>(estore), probe_index,
                      static_cast<double*>(probe_out), s, final_in_work);
}

extern "C" int32_t dnx_fdtd2d_adjoint_init(const dnx_grid2d* g, void* lam, const void* e_T,
                                           const dnx_monitor* mons, int32_t n_mon,
                                           int32_t n_total, void* S, void* stream) {
  int rc = check_grid2(g);
  if (rc) return rc;
  DNX_REQUIRE(lam && S && (g->te || e_T), "fdtd2d_adjoint_init: null pointer");
  DNX_REQUIRE(n_total >= 1, "fdtd2d_adjoint_init: n_total must be >= 1");
  if ((rc = check_monitors(mons, n_mon, 1, g->ny, g->nx, 1, true))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g->dtype == DNX_F32)
    return adj_init2<float>(g, static_cast<float*>(lam), static_cast<const float*>(e_T), mons,
                            n_mon, n_total, static_cast<float*>(S), s);
  return adj_init2<double>(g, static_cast<double*>(lam), static_cast<const double*>(e_T), mons,
                           n_mon, n_total, static_cast<double*>(S), s);
}

extern "C" int32_t dnx_fdtd2d_adjoint_run(const dnx_grid2d* g, void* lam, void* work,
                                          int32_t m_hi, int32_t m_lo, const void* e_lo,
                                          const void* estore, const dnx_source* srcs,
                                          int32_t n_src, const dnx_monitor* mons, int32_t n_mon,
                                          void* S, void* stream, int32_t* final_in_work) {
  int rc = check_grid2(g);
  if (rc) return rc;
  DNX_REQUIRE(lam && S, "fdtd2d_adjoint_run: null pointer");
  DNX_REQUIRE(m_lo >= 0 && m_hi >= m_lo, "fdtd2d_adjoint_run: bad range [%d, %d)", m_lo, m_hi);
  DNX_REQUIRE(!(m_hi - 1 > m_lo && m_hi - 1 >= 1) || estore != nullptr,
              "fdtd2d_adjoint_run: estore required");
  DNX_REQUIRE(!(m_lo >= 1 && m_hi > m_lo) || e_lo != nullptr,
              "fdtd2d_adjoint_run: e_lo required");
  if ((rc = check_sources(srcs, n_src, 1, g->ny, g->nx, 0))) return rc;
  if ((rc = check_monitors(mons, n_mon, 1, g->ny, g->nx, 1, true))) return rc;
  if (final_in_work) *final_in_work = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g->dtype == DNX_F32)
    return adj_run2<float>(g, static_cast<float*>(lam), static_cast<float*>(work), m_hi, m_lo,
                           static_cast<const float*>(e_lo), static_cast<const float*>(estore),
                           srcs, n_src, mons, n_mon, static_cast<float*>(S), s, final_in_work);
  return adj_run2<double>(g, static_cast<double*>(lam), static_cast<double*>(work), m_hi, m_lo,
                          static_cast<const double*>(e_lo), static_cast<const double*>(estore),
                          srcs, n_src, mons, n_mon, static_cast<double*>(S), s, final_in_work);
}

extern "C" int32_t dnx_fdtd2d_tm_batch(const dnx_grid2d* g, int32_t batch, void* state,
                                       int32_t n_steps, const dnx_source* src,
                                       const dnx_box* mbox, int32_t n_freq, const float* phase,
                                       float* acc, void* stream) {
  int rc = check_grid2(g);
  if (rc) return rc;
  DNX_REQUIRE(g->dtype == DNX_F32 && g->te == 0, "tm_batch: float32 TM only");
  DNX_REQUIRE(batch >= 1 && batch <= 65535 && state && n_steps >= 0,
              "tm_batch: bad batch/state/steps");
  if (src && (rc = check_sources(src, 1, 1, g->ny, g->nx, 0))) return rc;
  DNX_REQUIRE(n_freq == 0 || (mbox && phase && acc && box_ok(*mbox, 1, g->ny, g->nx)),
              "tm_batch: bad monitor");
  Tm2Batch a{};
  a.ny = g->ny;
  a.nx = g->nx;
  a.n = int64_t(g->ny) * g->nx;
  a.kh = float(g->kh);
  a.ce = static_cast<const float*>(g->ce);
  a.by = static_cast<const float*>(g->by);
  a.ay = static_cast<const float*>(g->ay);
  a.bx = static_cast<const float*>(g->bx);
  a.ax = static_cast<const float*>(g->ax);
  a.st = static_cast<float*>(state);
  a.has_src = src ? 1 : 0;
  if (src) {
    a.sbox = make_box(src->box);
    a.sprof = static_cast<const float*>(src->profile);
    a.wave = static_cast<const float*>(src->wave);
  }
  a.nf = n_freq;
  if (n_freq > 0) {
    a.mbox = make_box(*mbox);
    a.phase = phase;
    a.acc = acc;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // The resident kernel runs each design on one CU: with few designs and a
  // large DFT box its per-step accumulator traffic (per cell and frequency
  // 16 bytes) is bound by that CU's memory bandwidth (128 x 128 x 3 DFT:
  // 53 ms resident against 15 ms spread over the chip by per-step launches
  // for 8 designs; equal near 256 designs), so it takes small DFT boxes
  // (line monitors) at any batch size and large ones from 256 designs on.
  const int opt = option(DNX_OPT_FDTD2D_LAUNCH_PER_STEP);
  const int64_t dft = n_freq > 0 ? a.mbox.cells() * n_freq : 0;
  const bool fits = g->nx <= RES_NX && g->ny <= RES_RB * RES_R;
  if (fits && (opt == 2 || (opt == 0 && (dft <= 8192 || batch >= 256)))) {
    // the whole call in one launch, each design resident in one work-group
    static std::atomic<bool> attr{false};
    if (!attr.load()) {
      DNX_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k2tm_batch_res),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        int(RES_SMEM)));
      attr = true;
    }
    if (n_steps > 0)
      hipLaunchKernelGGL(k2tm_batch_res, dim3(batch), dim3(1024), RES_SMEM, s, a, n_steps);
    DNX_CHECK_HIP(hipGetLastError());
    return DNX_OK;
  }
  const dim3 grid(cdiv(g->nx, BX2), cdiv(g->ny, BY2), batch), block(BX2, BY2);
  for (int k = 0; k < n_steps; ++k) {
    hipLaunchKernelGGL(k2tm_batch_h, grid, block, 0, s, a);
    hipLaunchKernelGGL(k2tm_batch_e, grid, block, 0, s, a, k);
  }
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

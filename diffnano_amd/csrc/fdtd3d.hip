// 3D Yee FDTD on gfx950: forward leapfrog, replay with E storage, and the
// exact discrete adjoint (reverse sweep) of DiffNano's FDTDSolver3D.
//
// Semantics (fdtd3d.py:378-465), with kh = dt/dl, ce = dt/(eps dl), the
// axis-1 profile (by, ay = 1 + cy) and the axis-2 profile (bx, ax = 1 + cx):
//   Hx' = by  Hx - kh (ay D+y Ez - ax D+z Ey)
//   Hy' = bx  Hy - kh (ax D+z Ex - ay D+x Ez)
//   Hz' = bx  Hz - kh (ax D+x Ey - ay D+y Ex)
//   Ex' = bxby Ex + ce (ay D-y Hz' - ax D-z Hy')
//   Ey' = bxby Ey + ce (ax D-z Hx' - ay D-x Hz')
//   Ez' = bxby Ez + ce (ax D-x Hy' - ay D-y Hx')
// D+ is the forward difference, zero at the last index; D- the backward
// difference, zero at index 0.  The reference applies its z-named profile
// along axis 2 and its x-named profile (in two terms) along axis 1, so no
// profile varies along axis 0; on cubic grids all named profiles are equal.
//
// Kernel layout: a 64x4 thread tile in (x, y), each thread marching a chunk
// of z so that the z-neighbour comes from a register; x is the fastest axis
// so every wave's loads are 256-byte coalesced rows.
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "fdtd3d_fused.h"

#include <type_traits>

// No multiply-add contraction (same rounding in every FDTD kernel variant).
#pragma clang fp contract(off)

namespace dnx {

template <typename T>
struct G3 {
  int nz, ny, nx;
  int64_t plane, n;
  T kh;
  const T* __restrict__ ce;
  const T* __restrict__ by;
  const T* __restrict__ ay;
  const T* __restrict__ bx;
  const T* __restrict__ ax;
};

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int ZC = 16;  // z planes per thread

template <typename T>
static G3<T> make_g3(const dnx_grid3d* g) {
  G3<T> r;
  r.nz = g->nz;
  r.ny = g->ny;
  r.nx = g->nx;
  r.plane = int64_t(g->ny) * g->nx;
  r.n = r.plane * g->nz;
  r.kh = T(g->kh);
  r.ce = static_cast<const T*>(g->ce);
  r.by = static_cast<const T*>(g->by);
  r.ay = static_cast<const T*>(g->ay);
  r.bx = static_cast<const T*>(g->bx);
  r.ax = static_cast<const T*>(g->ax);
  return r;
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(TX* TY) void k3_update_h(G3<T> g, T* __restrict__ st) {
  const int x = blockIdx.x * TX + threadIdx.x;
  const int y = blockIdx.y * TY + threadIdx.y;
  if (x >= g.nx || y >= g.ny) return;
  const int z0 = blockIdx.z * ZC;
  const int z1 = min(g.nz, z0 + ZC);
  const int64_t N = g.n, P = g.plane, nx = g.nx;
  const T* Ex = st;
  const T* Ey = st + N;
  const T* Ez = st + 2 * N;
  T* Hx = st + 3 * N;
  T* Hy = st + 4 * N;
  T* Hz = st + 5 * N;
  const T by = g.by[y], ay = g.ay[y], bx = g.bx[x], ax = g.ax[x], kh = g.kh;
  const bool yl = y + 1 < g.ny, xl = x + 1 < g.nx;
  int64_t i = z0 * P + int64_t(y) * nx + x;
  T ex = Ex[i], ey = Ey[i];
  for (int z = z0; z < z1; ++z, i += P) {
    const bool zl = z + 1 < g.nz;
    const T exn = zl ? Ex[i + P] : T(0);
    const T eyn = zl ? Ey[i + P] : T(0);
    const T ez = Ez[i];
    const T dEz_dy = yl ? Ez[i + nx] - ez : T(0);
    const T dEy_dz = zl ? eyn - ey : T(0);
    const T dEx_dz = zl ? exn - ex : T(0);
    const T dEz_dx = xl ? Ez[i + 1] - ez : T(0);
    const T dEy_dx = xl ? Ey[i + 1] - ey : T(0);
    const T dEx_dy = yl ? Ex[i + nx] - ex : T(0);
    Hx[i] = by * Hx[i] - kh * (ay * dEz_dy - ax * dEy_dz);
    Hy[i] = bx * Hy[i] - kh * (ax * dEx_dz - ay * dEz_dx);
    Hz[i] = bx * Hz[i] - kh * (ax * dEy_dx - ay * dEx_dy);
    ex = exn;
    ey = eyn;
  }
}

template <typename T>
__global__ __launch_bounds__(TX* TY) void k3_update_e(G3<T> g, T* __restrict__ st,
                                                      T* __restrict__ es) {
  const int x = blockIdx.x * TX + threadIdx.x;
  const int y = blockIdx.y * TY + threadIdx.y;
  if (x >= g.nx || y >= g.ny) return;
  const int z0 = blockIdx.z * ZC;
  const int z1 = min(g.nz, z0 + ZC);
  const int64_t N = g.n, P = g.plane, nx = g.nx;
  T* Ex = st;
  T* Ey = st + N;
  T* Ez = st + 2 * N;
  const T* Hx = st + 3 * N;
  const T* Hy = st + 4 * N;
  const T* Hz = st + 5 * N;
  const T by = g.by[y], ay = g.ay[y], bx = g.bx[x], ax = g.ax[x];
  const T bxy = bx * by;
  const bool yf = y > 0, xf = x > 0;
  int64_t i = z0 * P + int64_t(y) * nx + x;
  T hxp = z0 > 0 ? Hx[i - P] : T(0);
  T hyp = z0 > 0 ? Hy[i - P] : T(0);
  for (int z = z0; z < z1; ++z, i += P) {
    const bool zf = z > 0;
    const T hx = Hx[i], hy = Hy[i], hz = Hz[i];
    const T dHz_dy = yf ? hz - Hz[i - nx] : T(0);
    const T dHy_dz = zf ? hy - hyp : T(0);
    const T dHx_dz = zf ? hx - hxp : T(0);
    const T dHz_dx = xf ? hz - Hz[i - 1] : T(0);
    const T dHy_dx = xf ? hy - Hy[i - 1] : T(0);
    const T dHx_dy = yf ? hx - Hx[i - nx] : T(0);
    const T c = g.ce[i];
    const T ex = bxy * Ex[i] + c * (ay * dHz_dy - ax * dHy_dz);
    const T ey = bxy * Ey[i] + c * (ax * dHx_dz - ay * dHz_dx);
    const T ez = bxy * Ez[i] + c * (ax * dHy_dx - ay * dHx_dy);
    Ex[i] = ex;
    Ey[i] = ey;
    Ez[i] = ez;
    if (es) {
      es[i] = ex;
      es[N + i] = ey;
      es[2 * N + i] = ez;
    }
    hxp = hx;
    hyp = hy;
  }
}

// ---------------------------------------------------------------------------
// Adjoint.  lam = [lamEx, lamEy, lamEz, muHx, muHy, muHz].
//   kernel A:  muH <- B_H muH + K_E^T (ce lamE)       (in place on muH)
//   kernel B:  t = kh K_H^T-part(muH);  lamE <- bxby lamE + t;  S += E^m . t
// with the exact transposes of the truncated differences:
//   (D-^T g)[j] = [j>=1] g[j] - [j+1<n] g[j+1]
//   (D+^T g)[j] = [j>=1] g[j-1] - [j<=n-2] g[j]
// and each profile factor applied at the index where g lives.
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(TX* TY) void k3_adj_h(G3<T> g, T* __restrict__ lam) {
  const int x = blockIdx.x * TX + threadIdx.x;
  const int y = blockIdx.y * TY + threadIdx.y;
  if (x >= g.nx || y >= g.ny) return;
  const int z0 = blockIdx.z * ZC;
  const int z1 = min(g.nz, z0 + ZC);
  const int64_t N = g.n, P = g.plane, nx = g.nx;
  const T* Lx = lam;
  const T* Ly = lam + N;
  const T* Lz = lam + 2 * N;
  T* Mx = lam + 3 * N;
  T* My = lam + 4 * N;
  T* Mz = lam + 5 * N;
  const T by = g.by[y], ay = g.ay[y], bx = g.bx[x], ax = g.ax[x];
  const bool my = y >= 1, myp = y + 1 < g.ny, mx = x >= 1, mxp = x + 1 < g.nx;
  const T ayp = myp ? g.ay[y + 1] : T(0);
  const T axp = mxp ? g.ax[x + 1] : T(0);
  int64_t i = z0 * P + int64_t(y) * nx + x;
  // u = ce * lamE at the current plane (x, y) for the z-direction terms
  T cez = g.ce[i];
  T ux = cez * Lx[i], uy = cez * Ly[i];
  for (int z = z0; z < z1; ++z, i += P) {
    const bool mz = z >= 1, mzp = z + 1 < g.nz;
    T uxn = T(0), uyn = T(0);
    if (mzp) {
      const T cn = g.ce[i + P];
      uxn = cn * Lx[i + P];
      uyn = cn * Ly[i + P];
    }
    const T c = g.ce[i];
    const T uz = c * Lz[i];
    const T uz_y1 = myp ? g.ce[i + nx] * Lz[i + nx] : T(0);
    const T uz_x1 = mxp ? g.ce[i + 1] * Lz[i + 1] : T(0);
    const T ux_y1 = myp ? g.ce[i + nx] * Lx[i + nx] : T(0);
    const T uy_x1 = mxp ? g.ce[i + 1] * Ly[i + 1] : T(0);
    const T tz_y = (mz ? uy : T(0)) - (mzp ? uyn : T(0));  // D-z^T (uEy)
    const T tz_x = (mz ? ux : T(0)) - (mzp ? uxn : T(0));  // D-z^T (uEx)
    const T ty_z = (my ? ay * uz : T(0)) - ayp * uz_y1;    // D-y^T (ay uEz)
    const T tx_z = (mx ? ax * uz : T(0)) - axp * uz_x1;    // D-x^T (ax uEz)
    const T ty_x = (my ? ay * ux : T(0)) - ayp * ux_y1;    // D-y^T (ay uEx)
    const T tx_y = (mx ? uy : T(0)) - uy_x1;               // D-x^T (uEy)
    Mx[i] = by * Mx[i] + ax * tz_y - ty_z;
    My[i] = bx * My[i] + tx_z - ax * tz_x;
    Mz[i] = bx * Mz[i] + ty_x - ay * tx_y;
    ux = uxn;
    uy = uyn;
  }
}

template <typename T>
__global__ __launch_bounds__(TX* TY) void k3_adj_e(G3<T> g, T* __restrict__ lam,
                                                   const T* __restrict__ em, T* __restrict__ S) {
  const int x = blockIdx.x * TX + threadIdx.x;
  const int y = blockIdx.y * TY + threadIdx.y;
  if (x >= g.nx || y >= g.ny) return;
  const int z0 = blockIdx.z * ZC;
  const int z1 = min(g.nz, z0 + ZC);
  const int64_t N = g.n, P = g.plane, nx = g.nx;
  T* Lx = lam;
  T* Ly = lam + N;
  T* Lz = lam + 2 * N;
  const T* Mx = lam + 3 * N;
  const T* My = lam + 4 * N;
  const T* Mz = lam + 5 * N;
  const T by = g.by[y], ay = g.ay[y], bx = g.bx[x], ax = g.ax[x], kh = g.kh;
  const T bxy = bx * by;
  const bool yf = y >= 1, yl = y <= g.ny - 2, xf = x >= 1, xl = x <= g.nx - 2;
  const T aym = yf ? g.ay[y - 1] : T(0);
  const T axm = xf ? g.ax[x - 1] : T(0);
  int64_t i = z0 * P + int64_t(y) * nx + x;
  T mxp = z0 > 0 ? Mx[i - P] : T(0);
  T myp = z0 > 0 ? My[i - P] : T(0);
  for (int z = z0; z < z1; ++z, i += P) {
    const bool zf = z >= 1, zl = z <= g.nz - 2;
    const T mx = Mx[i], my = My[i], mz = Mz[i];
    // D+^T terms
    const T y_mz = (yf ? aym * Mz[i - nx] : T(0)) - (yl ? ay * mz : T(0));  // D+y^T(ay muHz)
    const T z_my = (zf ? myp : T(0)) - (zl ? my : T(0));                     // D+z^T(muHy)
    const T z_mx = (zf ? mxp : T(0)) - (zl ? mx : T(0));                     // D+z^T(muHx)
    const T x_mz = (xf ? axm * Mz[i - 1] : T(0)) - (xl ? ax * mz : T(0));   // D+x^T(ax muHz)
    const T x_my = (xf ? My[i - 1] : T(0)) - (xl ? my : T(0));               // D+x^T(muHy)
    const T y_mx = (yf ? aym * Mx[i - nx] : T(0)) - (yl ? ay * mx : T(0));  // D+y^T(ay muHx)
    const T tx = kh * (y_mz - ax * z_my);
    const T ty = kh * (ax * z_mx - x_mz);
    const T tz = kh * (ay * x_my - y_mx);
    Lx[i] = bxy * Lx[i] + tx;
    Ly[i] = bxy * Ly[i] + ty;
    Lz[i] = bxy * Lz[i] + tz;
    S[i] += em[i] * tx + em[N + i] * ty + em[2 * N + i] * tz;
    mxp = mx;
    myp = my;
  }
}

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------

static int check_grid3(const dnx_grid3d* g) {
  DNX_REQUIRE(g != nullptr, "grid3d: null");
  DNX_REQUIRE(g->nz > 0 && g->ny > 0 && g->nx > 0, "grid3d: bad shape %d x %d x %d", g->nz,
              g->ny, g->nx);
  DNX_REQUIRE(g->dtype == DNX_F32 || g->dtype == DNX_F64, "grid3d: bad dtype %d", g->dtype);
  DNX_REQUIRE(g->ce && g->by && g->ay && g->bx && g->ax, "grid3d: null coefficient array");
  return DNX_OK;
}

static dim3 grid3_launch(const dnx_grid3d* g) {
  return dim3(cdiv(g->nx, TX), cdiv(g->ny, TY), cdiv(g->nz, ZC));
}

// Kernel variant: the fused single-pass kernels for nx <= 256 (default), or
// the two-kernel in-place variant ("simple") when DNX_OPT_FDTD3D_SIMPLE is
// set, no ping-pong buffer is given, or nx > 256.
static bool use_fused(const dnx_grid3d* g, const void* work) {
  if (work == nullptr || g->nx > FUSED_MAX_W) return false;
  return option(DNX_OPT_FDTD3D_SIMPLE) == 0;
}

template <typename T>
static Fused3Args<T> fused_args(const dnx_grid3d* gd, const Fused3Tiles& t) {
  Fused3Args<T> a{};
  a.nz = gd->nz;
  a.ny = gd->ny;
  a.nx = gd->nx;
  a.ty_tiles = int(cdiv(gd->ny, t.ty));
  a.zc = t.zc;
  a.nblocks = a.ty_tiles * int(cdiv(gd->nz, t.zc));
  a.P = int64_t(gd->ny) * gd->nx;
  a.N = a.P * gd->nz;
  a.kh = T(gd->kh);
  a.ce = static_cast<const T*>(gd->ce);
  a.by = static_cast<const T*>(gd->by);
  a.ay = static_cast<const T*>(gd->ay);
  a.bx = static_cast<const T*>(gd->bx);
  a.ax = static_cast<const T*>(gd->ax);
  a.src_comp = -1;
  return a;
}

// Two-step kernels: float32, fused variant, at most one source (fused) and
// FUSED2_MAX_MON monitors (fused), no probe.  DNX_OPT_FDTD3D_ONE_STEP forces
// the one-step kernels.
template <typename T>
static bool use_pairs(const dnx_grid3d* gd, bool fused, int n_src, int n_mon, int64_t probe,
                      int kernel) {
  if (!std::is_same<T, float>::value || !fused || gd->dtype != DNX_F32) return false;
  if (n_src > 1 || n_mon > FUSED2_MAX_MON || probe >= 0) return false;
  // tile height 0 for kernel 2 / 3: two-step path off (the autotuner's
  // choice where the one-step kernels are faster)
  if (fused3_tiles(gd->dtype, kernel).ty <= 0) return false;
  // the two-step kernels address the six fields through one 32-bit buffer
  // offset: 6 * N * 4 bytes must stay below 2^31
  if (int64_t(gd->nx) * gd->ny * gd->nz * 6 * int64_t(sizeof(T)) >= (int64_t(1) << 31)) return false;
  return option(DNX_OPT_FDTD3D_ONE_STEP) == 0;
}

static int device_cus() {
  static const int cus = [] {
    int dev = 0, n = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      n = 256;
    return n;
  }();
  return cus;
}

// The tile shape of `kernel` for this grid: the configured one, with the
// z-chunk halved (down to 8 planes) while the grid would have fewer than
// two work-groups per CU, so that a default shape tuned at 256^3 still
// fills the device on smaller grids.  Results do not depend on tiles.
static Fused3Tiles grid_tiles(const dnx_grid3d* gd, int kernel) {
  Fused3Tiles t = fused3_tiles(gd->dtype, kernel);
  if (t.ty <= 0) return t;
  const int64_t target = 2 * int64_t(device_cus());
  const int64_t ty_tiles = cdiv(gd->ny, t.ty);
  while (t.zc > 8 && ty_tiles * cdiv(gd->nz, t.zc) < target) t.zc = (t.zc + 1) / 2;
  return t;
}

template <typename T>
static Fused2Args<T> fused2_args(const dnx_grid3d* gd, const Fused3Tiles& t,
                                 const dnx_source* srcs, int n_src, const dnx_monitor* mons,
                                 int n_mon, T* side, int subs) {
  Fused2Args<T> f{};
  f.a = fused_args<T>(gd, t);
  if (n_src > 0) {
    f.a.src_comp = srcs[0].comp;
    f.a.src_box = make_box(srcs[0].box);
    f.a.src_prof = static_cast<const T*>(srcs[0].profile);
    f.a.src_wave = static_cast<const T*>(srcs[0].wave);
  }
  f.n_mon = n_mon;
  int64_t off = 0;
  for (int j = 0; j < n_mon; ++j) {
    FusedMon& m = f.mon[j];
    m.box = make_box(mons[j].box);
    m.mask = mons[j].comp_mask;
    m.ncomp = (m.mask & 1) + ((m.mask >> 1) & 1) + ((m.mask >> 2) & 1);
    f.side[j] = side + off;
    off += int64_t(subs) * m.ncomp * m.box.cells();
  }
  return f;
}

// Elements of the compact monitor buffers of the two-step kernels.
static int64_t side_elems(const dnx_monitor* mons, int n_mon, int subs) {
  int64_t n = 0;
  for (int j = 0; j < n_mon; ++j) {
    const int mk = mons[j].comp_mask;
    n += int64_t(subs) * ((mk & 1) + ((mk >> 1) & 1) + ((mk >> 2) & 1)) *
         make_box(mons[j].box).cells();
  }
  return n;
}

// Bytes of caller-provided workspace the adjoint sweep needs: the compact
// monitor adjoint buffers of the float32 two-step kernels ([ncomp][cells]
// per monitor).  An upper bound that does not depend on the kernel-selection
// options, so that callers can size it once per plan.
static int64_t adj3_workspace_bytes(const dnx_grid3d* gd, const dnx_monitor* mons, int n_mon) {
  if (gd->dtype != DNX_F32 || n_mon > FUSED2_MAX_MON) return 0;
  const int64_t n = side_elems(mons, n_mon, 1) * int64_t(sizeof(float));
  return n < (int64_t(1) << 29) ? n : 0;
}

template <typename T>
static int run3(const dnx_grid3d* gd, T* st, T* work, int n0, int n_steps,
                const dnx_source* srcs, int n_src, const dnx_monitor* mons, int n_mon, T* ckpt,
                int ckpt_every, T* es, int64_t probe, T* probe_out, hipStream_t s, int* fin,
                const T* init = nullptr) {
  G3<T> g = make_g3<T>(gd);
  const int64_t N = g.n;
  const bool fused = use_fused(gd, work);
  const Fused3Tiles tiles = grid_tiles(gd, 0);
  Fused3Args<T> fa = fused ? fused_args<T>(gd, tiles) : Fused3Args<T>{};
  const int ty = tiles.ty;
  int j0 = 0;  // sources handled by the side kernel start at j0
  if (fused && n_src > 0) {
    fa.src_comp = srcs[0].comp;
    fa.src_box = make_box(srcs[0].box);
    fa.src_prof = static_cast<const T*>(srcs[0].profile);
    j0 = 1;
  }
  // The fused kernels ping-pong between st and work ("pp" buffers); a launch
  // whose output is a checkpointed step writes straight into its
  // checkpoint slot (no copy), and the next launch reads it from there.
  // init (if given) is read by the first launch in place of st.
  if (init && (!fused || n_steps == 0)) {
    DNX_CHECK_HIP(hipMemcpyAsync(st, init, sizeof(T) * 6 * N, hipMemcpyDeviceToDevice, s));
    init = nullptr;
  }
  T* cur = init ? const_cast<T*>(init) : st;
  T* nxt = fused ? work : st;
  auto ckpt_slot = [&](int k) { return ckpt + int64_t(k / ckpt_every) * 6 * N; };
  // output of a fused launch that produces step k_out (relative)
  auto out_for = [&](int k_out) -> T* {
    if (fused && ckpt && k_out < n_steps && k_out % ckpt_every == 0) return ckpt_slot(k_out);
    return nxt;
  };
  auto advance = [&](T* out) {
    cur = out;
    if (out == st || out == work) nxt = (out == st) ? work : st;
  };
  const dim3 grid = grid3_launch(gd), block(TX, TY);
  const bool pairs = use_pairs<T>(gd, fused, n_src, n_mon, probe, 2);
  Fused2Args<T> f2{};
  int ty2 = 0;
  bool monx = false;  // monitor 0's DFT inside the forward pair's tiles
  if (pairs) {
    // An x-normal monitor plane is accumulated by the tiles themselves;
    // otherwise k_monitor_pair_fwd recomputes the intermediate sample after
    // each launch.
    const Fused3Tiles t2 = grid_tiles(gd, 2);
    monx = n_mon == 1 && monitor_in_tile(mons[0], t2.ty, t2.zc);
    f2 = fused2_args<T>(gd, t2, srcs, n_src, mons, monx ? 1 : 0, nullptr, 2);
    if (monx) {
      f2.mph = mons[0].phase;
      f2.macc = mons[0].acc;
      f2.mnf = mons[0].n_freq;
    }
    ty2 = t2.ty;
  }
  for (int k = 0; k < n_steps; ++k) {
    const int n = n0 + k;
    T* slot = es ? es + int64_t(k) * 3 * N : nullptr;
    if (ckpt && k % ckpt_every == 0 && cur != ckpt_slot(k)) {
      DNX_CHECK_HIP(hipMemcpyAsync(ckpt_slot(k), cur, sizeof(T) * 6 * N,
                                   hipMemcpyDeviceToDevice, s));
    }
    if (pairs && k + 1 < n_steps && !(ckpt && (k + 1) % ckpt_every == 0)) {
      T* out = out_for(k + 2);
      if constexpr (std::is_same<T, float>::value) {
        f2.a.in = cur;
        f2.a.out = out;
        f2.a.es = slot;
        f2.es2 = slot ? slot + 3 * N : nullptr;
        f2.n = n;
        const char* tname = slot ? "fdtd3d_fwd2_store" : "fdtd3d_fwd2";
        TimedLaunch tl = timed_begin(tname, s);
        int rc = launch_fused2_fwd<T>(f2, ty2, s);
        timed_end(tname, s, tl);
        if (rc) return rc;
        for (int j = monx ? 1 : 0; j < n_mon; ++j)
          if ((rc = launch_monitor_pair_fwd<T>(f2, mons[j], n, s))) return rc;
      }
      advance(out);
      ++k;  // the pair covered steps k and k+1
      continue;
    }
    T* out = out_for(k + 1);  // (the in-place simple kernels write cur == nxt)
    if (fused) {
      fa.in = cur;
      fa.out = out;
      fa.es = slot;
      if (fa.src_comp >= 0) fa.src_wave = static_cast<const T*>(srcs[0].wave) + n;
      const char* tname = slot ? "fdtd3d_fwd_store" : "fdtd3d_fwd";
      TimedLaunch tl = timed_begin(tname, s);
      int rc = launch_fused3_fwd<T>(fa, ty, s);
      timed_end(tname, s, tl);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL(k3_update_h<T>, grid, block, 0, s, g, cur);
      hipLaunchKernelGGL(k3_update_e<T>, grid, block, 0, s, g, cur, slot);
      DNX_CHECK_HIP(hipGetLastError());
    }
    for (int j = j0; j < n_src; ++j) {
      int rc = launch_source<T>(out + srcs[j].comp * N, g.ny, g.nx, srcs[j], n, s);
      if (rc) return rc;
      if (slot && (rc = launch_source<T>(slot + srcs[j].comp * N, g.ny, g.nx, srcs[j], n, s)))
        return rc;
    }
    if (probe >= 0) {
      DNX_CHECK_HIP(hipMemcpyAsync(probe_out + k, out + probe, sizeof(T),
                                   hipMemcpyDeviceToDevice, s));
    }
    Comps<T> E{{out, out + N, out + 2 * N}};
    for (int j = 0; j < n_mon; ++j) {
      int rc = launch_monitor<T>(E, g.ny, g.nx, mons[j], n, s);
      if (rc) return rc;
    }
    if (fused) advance(out);
  }
  *fin = (cur != st) ? 1 : 0;  // the last output is never a checkpoint slot
  return DNX_OK;
}

template <typename T>
static int adj_init3(const dnx_grid3d* gd, T* lam, const T* eT, const dnx_monitor* mons,
                     int n_mon, int n_total, T* S, hipStream_t s) {
  G3<T> g = make_g3<T>(gd);
  const int64_t N = g.n;
  Comps<T> L{{lam, lam + N, lam + 2 * N}};
  Comps<T> none{{nullptr, nullptr, nullptr}};
  for (int j = 0; j < n_mon; ++j) {
    int rc = launch_monitor_adj<T>(L, none, nullptr, g.ny, g.nx, mons[j], n_total - 1, s);
    if (rc) return rc;
  }
  int rc = launch_zero<T>(lam + 3 * N, 3 * N, s);
  if (rc) return rc;
  Comps<T> Ev{{const_cast<T*>(eT), const_cast<T*>(eT) + N, const_cast<T*>(eT) + 2 * N}};
  return launch_dot_init<T>(S, Ev, L, 3, N, s);
}

template <typename T>
static int adj_run3(const dnx_grid3d* gd, T* lam, T* work, int m_hi, int m_lo, const T* e_lo,
                    const T* es, const dnx_source* srcs, int n_src, const dnx_monitor* mons,
                    int n_mon, T* S, void* ws, int64_t ws_bytes, hipStream_t s, int* fin) {
  G3<T> g = make_g3<T>(gd);
  const int64_t N = g.n;
  const bool fused = use_fused(gd, work);
  const Fused3Tiles tiles = grid_tiles(gd, 1);
  Fused3Args<T> fa = fused ? fused_args<T>(gd, tiles) : Fused3Args<T>{};
  const int ty = tiles.ty;
  int j0 = 0;
  if (fused && n_src > 0) {
    fa.src_comp = srcs[0].comp;
    fa.src_box = make_box(srcs[0].box);
    fa.src_prof = static_cast<const T*>(srcs[0].profile);
    j0 = 1;
  }
  fa.S = S;
  T* cur = lam;
  T* nxt = fused ? work : lam;
  const dim3 grid = grid3_launch(gd), block(TX, TY);
  // Step pairs run the two-step adjoint (DNX_OPT_FDTD3D_ADJ2 = 0 turns it
  // off; DNX_OPT_FDTD3D_ONE_STEP turns off both two-step kernels).
  // (the tiles read the monitor adjoint terms through one descriptor whose
  // out-of-box offsets start at 2^29 bytes)
  const bool pairs = use_pairs<T>(gd, fused, n_src, n_mon, -1, 3) &&
                     option(DNX_OPT_FDTD3D_ADJ2) != 0 &&
                     side_elems(mons, n_mon, 1) * int64_t(sizeof(T)) < (int64_t(1) << 29);
  Fused2Args<T> f2{};
  int ty2 = 0;
  if (pairs) {
    const Fused3Tiles t2 = grid_tiles(gd, 3);
    const int64_t need = side_elems(mons, n_mon, 1) * int64_t(sizeof(T));
    DNX_REQUIRE(need == 0 || (ws != nullptr && ws_bytes >= need),
                "fdtd3d_adjoint_run: workspace of %lld bytes required, %lld given (size it with "
                "dnx_fdtd3d_adjoint_workspace_bytes)", (long long)need, (long long)ws_bytes);
    f2 = fused2_args<T>(gd, t2, srcs, n_src, mons, n_mon, static_cast<T*>(ws), 1);
    f2.a.S = S;
    ty2 = t2.ty;
  }
  auto estep = [&](int mm) -> const T* {
    return (mm == m_lo) ? e_lo : es + int64_t(mm - m_lo - 1) * 3 * N;
  };
  // The monitor adjoint of a pair's last sample waits for the next pair's
  // adjoint-source launch and shares it (k_monitor_adj_d); it is applied
  // before anything else reads lambda or S.
  Comps<T> pend_l{}, pend_e{};
  int pend_row = -1;
  auto flush = [&]() -> int {
    if (pend_row < 0) return DNX_OK;
    for (int j = 0; j < n_mon; ++j) {
      int rc = launch_monitor_adj<T>(pend_l, pend_e, S, g.ny, g.nx, mons[j], pend_row, s);
      if (rc) return rc;
    }
    pend_row = -1;
    return DNX_OK;
  };
  for (int m = m_hi - 1; m >= m_lo && m >= 1; --m) {
    const T* em = estep(m);
    if (pairs && m - 1 >= m_lo && m - 1 >= 1) {
      if constexpr (std::is_same<T, float>::value) {
        f2.a.in = cur;
        f2.a.out = nxt;
        f2.a.em = em;
        f2.em2 = estep(m - 1);
        f2.n = m;
        int rc;
        for (int j = 0; j < n_mon; ++j) {
          rc = pend_row >= 0 ? launch_monitor_adj_d<T>(pend_l, pend_e, S, g.ny, g.nx, mons[j],
                                                       pend_row, f2.side[j], m - 1, s)
                             : launch_monitor_d<T>(f2.side[j], mons[j], m - 1, s);
          if (rc) return rc;
        }
        pend_row = -1;
        TimedLaunch tl = timed_begin("fdtd3d_adj2", s);
        rc = launch_fused2_adj<T>(f2, ty2, s);
        timed_end("fdtd3d_adj2", s, tl);
        if (rc) return rc;
        T* e2 = const_cast<T*>(f2.em2);
        pend_l = Comps<T>{{nxt, nxt + N, nxt + 2 * N}};
        pend_e = Comps<T>{{e2, e2 + N, e2 + 2 * N}};
        pend_row = n_mon > 0 ? m - 2 : -1;
      }
      T* t = cur;
      cur = nxt;
      nxt = t;
      --m;  // the pair covered steps m and m-1
      continue;
    }
    if (int rc = flush()) return rc;
    for (int j = j0; j < n_src; ++j) {
      int rc = launch_source_adj<T>(S, cur + srcs[j].comp * N, g.ny, g.nx, srcs[j], m, s);
      if (rc) return rc;
    }
    if (fused) {
      fa.in = cur;
      fa.out = nxt;
      fa.em = em;
      if (fa.src_comp >= 0) fa.src_wave = static_cast<const T*>(srcs[0].wave) + m;
      TimedLaunch tl = timed_begin("fdtd3d_adj", s);
      int rc = launch_fused3_adj<T>(fa, ty, s);
      timed_end("fdtd3d_adj", s, tl);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL(k3_adj_h<T>, grid, block, 0, s, g, cur);
      hipLaunchKernelGGL(k3_adj_e<T>, grid, block, 0, s, g, cur, em, S);
      DNX_CHECK_HIP(hipGetLastError());
    }
    Comps<T> L{{nxt, nxt + N, nxt + 2 * N}};
    Comps<T> Em{{const_cast<T*>(em), const_cast<T*>(em) + N, const_cast<T*>(em) + 2 * N}};
    for (int j = 0; j < n_mon; ++j) {
      int rc = launch_monitor_adj<T>(L, Em, S, g.ny, g.nx, mons[j], m - 1, s);
      if (rc) return rc;
    }
    if (fused) {
      T* t = cur;
      cur = nxt;
      nxt = t;
    }
  }
  if (int rc = flush()) return rc;
  if (m_lo == 0) {
    for (int j = 0; j < n_src; ++j) {
      int rc = launch_source_adj<T>(S, cur + srcs[j].comp * N, g.ny, g.nx, srcs[j], 0, s);
      if (rc) return rc;
    }
  }
  *fin = (cur != lam) ? 1 : 0;
  return DNX_OK;
}

}  // namespace dnx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

using namespace dnx;

extern "C" int32_t dnx_fdtd3d_run_from(const dnx_grid3d* g, const void* init, void* state,
                                       void* work, int32_t n0, int32_t n_steps,
                                       const dnx_source* srcs, int32_t n_src,
                                       const dnx_monitor* mons, int32_t n_mon, void* ckpt,
                                       int32_t ckpt_every, void* estore, int64_t probe_index,
                                       void* probe_out, void* stream, int32_t* final_in_work) {
  int rc = check_grid3(g);
  if (rc) return rc;
  DNX_REQUIRE(state != nullptr, "fdtd3d_run: null state");
  DNX_REQUIRE(n0 >= 0 && n_steps >= 0, "fdtd3d_run: bad step range n0=%d n_steps=%d", n0,
              n_steps);
  DNX_REQUIRE(ckpt == nullptr || ckpt_every > 0, "fdtd3d_run: ckpt_every must be > 0");
  if ((rc = check_sources(srcs, n_src, g->nz, g->ny, g->nx, 2))) return rc;
  if ((rc = check_monitors(mons, n_mon, g->nz, g->ny, g->nx, 7, false))) return rc;
  DNX_REQUIRE(probe_index < 0 || (probe_out != nullptr &&
                                   probe_index < int64_t(6) * g->nz * g->ny * g->nx),
              "run: bad probe index %lld", (long long)probe_index);
  DNX_REQUIRE(work != state, "fdtd3d_run: work must differ from state");
  DNX_REQUIRE(init != state && init != work, "fdtd3d_run_from: init must differ from state and "
              "work");
  int fin = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g->dtype == DNX_F32)
    rc = run3<float>(g, static_cast<float*>(state), static_cast<float*>(work), n0, n_steps, srcs,
                     n_src, mons, n_mon, static_cast<float*>(ckpt), ckpt_every,
                     static_cast<float*>(estore), probe_index, static_cast<float*>(probe_out), s,
                     &fin, static_cast<const float*>(init));
  else
    rc = run3<double>(g, static_cast<double*>(state), static_cast<double*>(work), n0, n_steps,
                      srcs, n_src, mons, n_mon, static_cast<double*>(ckpt), ckpt_every,
                      static_cast<double*>(estore), probe_index, static_cast<double*>(probe_out),
                      s, &fin, static_cast<const double*>(init));
  if (final_in_work) *final_in_work = fin;
  return rc;
}

extern "C" int32_t dnx_fdtd3d_run(const dnx_grid3d* g, void* state, void* work, int32_t n0,
                                  int32_t n_steps, const dnx_source* srcs, int32_t n_src,
                                  const dnx_monitor* mons, int32_t n_mon, void* ckpt,
                                  int32_t ckpt_every, void* estore, int64_t probe_index,
                                  void* probe_out, void* stream, int32_t* final_in_work) {
  const void* init = nullptr;
  return dnx_fdtd3d_run_from(g, init, state, work, n0, n_steps, srcs, n_src, mons, n_mon, ckpt,
                             ckpt_every, estore, probe_index, probe_out, stream, final_in_work);
}

extern "C" int32_t dnx_fdtd3d_adjoint_init(const dnx_grid3d* g, void* lam, const void* e_T,
                                           const dnx_monitor* mons, int32_t n_mon,
                                           int32_t n_total, void* S, void* stream) {
  int rc = check_grid3(g);
  if (rc) return rc;
  DNX_REQUIRE(lam && e_T && S, "fdtd3d_adjoint_init: null pointer");
  DNX_REQUIRE(n_total >= 1, "fdtd3d_adjoint_init: n_total must be >= 1");
  if ((rc = check_monitors(mons, n_mon, g->nz, g->ny, g->nx, 7, true))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g->dtype == DNX_F32)
    return adj_init3<float>(g, static_cast<float*>(lam), static_cast<const float*>(e_T), mons,
                            n_mon, n_total, static_cast<float*>(S), s);
  return adj_init3<double>(g, static_cast<double*>(lam), static_cast<const double*>(e_T), mons,
                           n_mon, n_total, static_cast<double*>(S), s);
}

extern "C" int64_t dnx_fdtd3d_adjoint_workspace_bytes(const dnx_grid3d* g,
                                                     const dnx_monitor* mons, int32_t n_mon) {
  if (check_grid3(g) || n_mon < 0 || (n_mon > 0 && mons == nullptr)) return -1;
  return adj3_workspace_bytes(g, mons, n_mon);
}

extern "C" int32_t dnx_fdtd3d_adjoint_run(const dnx_grid3d* g, void* lam, void* work,
                                          int32_t m_hi, int32_t m_lo, const void* e_lo,
                                          const void* estore, const dnx_source* srcs,
                                          int32_t n_src, const dnx_monitor* mons, int32_t n_mon,
                                          void* S, void* workspace, int64_t ws_bytes,
                                          void* stream, int32_t* final_in_work) {
  int rc = check_grid3(g);
  if (rc) return rc;
  DNX_REQUIRE(lam && S, "fdtd3d_adjoint_run: null pointer");
  DNX_REQUIRE(work != lam, "fdtd3d_adjoint_run: work must differ from lam");
  DNX_REQUIRE(m_lo >= 0 && m_hi >= m_lo, "fdtd3d_adjoint_run: bad range [%d, %d)", m_lo, m_hi);
  DNX_REQUIRE(!(m_hi - 1 > m_lo && m_hi - 1 >= 1) || estore != nullptr,
              "fdtd3d_adjoint_run: estore required");
  DNX_REQUIRE(!(m_lo >= 1 && m_hi > m_lo) || e_lo != nullptr,
              "fdtd3d_adjoint_run: e_lo required");
  if ((rc = check_sources(srcs, n_src, g->nz, g->ny, g->nx, 2))) return rc;
  if ((rc = check_monitors(mons, n_mon, g->nz, g->ny, g->nx, 7, true))) return rc;
  int fin = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (g->dtype == DNX_F32)
    rc = adj_run3<float>(g, static_cast<float*>(lam), static_cast<float*>(work), m_hi, m_lo,
                         static_cast<const float*>(e_lo), static_cast<const float*>(estore), srcs,
                         n_src, mons, n_mon, static_cast<float*>(S), workspace, ws_bytes,
                         s, &fin);
  else
    rc = adj_run3<double>(g, static_cast<double*>(lam), static_cast<double*>(work), m_hi, m_lo,
                          static_cast<const double*>(e_lo), static_cast<const double*>(estore),
                          srcs, n_src, mons, n_mon, static_cast<double*>(S), workspace,
                          ws_bytes, s, &fin);
  if (final_in_work) *final_in_work = fin;
  return rc;
}

// ---------------------------------------------------------------------------
// Adjoint sweep without caller workspace: the two-step kernels' monitor
// adjoint terms live in a fixed device array of this library
// ---------------------------------------------------------------------------

// Every byte of `scratch` is written by k_monitor_d before a two-step
// adjoint launch reads it, so nothing has to start at zero.
static constexpr int64_t kAdjScratchBytes = int64_t(4) << 20;  // 4 MiB
static __device__ __align__(256) unsigned char scratch[kAdjScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).  The array's name
// is only a host-side stand-in: kernels receive this address.
int scratch_address(void** out) {
  int dev = -1;
  DNX_CHECK_HIP(hipGetDevice(&dev));
  DNX_REQUIRE(dev >= 0 && dev < kMaxDevices, "adjoint scratch: device %d beyond %d", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    DNX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DNX_REQUIRE(p != nullptr, "adjoint scratch: no device address on device %d", dev);
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return DNX_OK;
}

}  // namespace

extern "C" int64_t dnx_fdtd3d_adjoint_static_scratch_bytes(void) { return kAdjScratchBytes; }

extern "C" int32_t dnx_fdtd3d_adjoint_run_noalloc(const dnx_grid3d* g, void* lam, void* work,
                                                  int32_t m_hi, int32_t m_lo, const void* e_lo,
                                                  const void* estore, const dnx_source* srcs,
                                                  int32_t n_src, const dnx_monitor* mons,
                                                  int32_t n_mon, void* S, void* stream,
                                                  int32_t* final_in_work) {
  int rc = check_grid3(g);
  if (rc) return rc;
  DNX_REQUIRE(n_mon >= 0 && (n_mon == 0 || mons != nullptr), "adjoint_run_noalloc: monitors");
  const int64_t need = adj3_workspace_bytes(g, mons, n_mon);
  if (need > kAdjScratchBytes) {
    set_error("fdtd3d_adjoint_run_noalloc: the monitors need %lld bytes of scratch, beyond the "
              "library's fixed %lld (use dnx_fdtd3d_adjoint_run with a workspace)",
              (long long)need, (long long)kAdjScratchBytes);
    return DNX_ERANGE;
  }
  void* ws = nullptr;
  if (need > 0 && (rc = scratch_address(&ws))) return rc;
  return dnx_fdtd3d_adjoint_run(g, lam, work, m_hi, m_lo, e_lo, estore, srcs, n_src, mons, n_mon,
                                S, ws, need, stream, final_in_work);
}

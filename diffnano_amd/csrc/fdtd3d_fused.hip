// Fused single-pass 3D Yee step on gfx950, forward and adjoint.
//
// Work decomposition.  A work-group owns TY full rows (all nx <= 256 x
// columns, one thread per column, waves of 64 consecutive x) of a chunk of
// zc z-planes and marches z upward.  Per plane it loads the "E-like" input
// of plane z+1 (the z+1 neighbour of this plane and the own value of the
// next), the "H-like" input and ce of plane z, computes the H-like update
// for rows y0-1 .. y0+TY-1 (the row y0-1 is a halo recomputed locally, so
// no work-group reads another's new values), exchanges x neighbours through
// LDS (double-buffered by plane parity: one barrier per plane), then
// computes the E-like update of rows y0 .. y0+TY-1 and stores both.  y and
// z neighbours live in registers.  Reads and writes are whole 256-byte rows
// per wave; input and output are separate buffers (ping-pong), so adjacent
// work-groups never race.  A chunk starting at z0 > 0 recomputes the
// H-like plane z0-1 first.
//
// Work-group ids are remapped so that consecutive tiles (y-neighbours that
// share halo rows) run on the same XCD and share its L2: hardware assigns
// block b to XCD b % 8.
//
// Forward (FDTDSolver3D._time_step, fdtd3d.py:378-465; source fdtd3d.py:502):
//   H' = bH H - kh (..D+ E..);   E' = bxby E + ce (..D- H'..);   E'_c += w s
// Adjoint (transpose of one forward step; lam = dL/dE, mu = dL/dH'):
//   mu' = bH mu + K_E^T (ce lam)  ... uses +1 neighbours of u = ce lam
//   t   = kh (D+^T terms of mu')  ... uses -1 neighbours of mu'
//   lam' = bxby lam + t;  S += E^m . t  - w_m s lam_src
#include "fdtd3d_fused.h"

#include <atomic>

// Every FDTD kernel rounds each operation as written (no fused
// multiply-add contraction), so the one-step, two-step, simple and batched
// kernels produce bitwise the same fields whatever code surrounds them.
#pragma clang fp contract(off)

namespace dnx {

// [dtype][kernel: 0 forward, 1 adjoint] -> (ty, zc).  Defaults from a sweep
// at 256^3 on MI355X (float32: forward 4 x 32 ~5.0 TB/s algorithmic,
// adjoint 2 x 16 ~4.6 TB/s); float64 uses ty = 2 to stay within registers.
// Kernels 2 and 3 are the two-step forward and adjoint (float32 only):
// forward 4 x 32 (~273 us per step pair at 256^3), adjoint 2 x 64 (~388 us
// per step pair).
static std::atomic<int> g_ty[2][4] = {{{4}, {2}, {4}, {2}}, {{2}, {2}, {2}, {2}}};
static std::atomic<int> g_zc[2][4] = {{{32}, {16}, {32}, {64}}, {{32}, {16}, {64}, {16}}};

Fused3Tiles fused3_tiles(int dtype, int kernel) {
  const int d = dtype == DNX_F64 ? 1 : 0;
  const int k = (kernel >= 0 && kernel < 4) ? kernel : 0;
  return {g_ty[d][k].load(), g_zc[d][k].load()};
}

template <int TY>
__device__ inline void block_tile(int nblocks, int ty_tiles, int zc, int& y0, int& z0,
                                  bool& live) {
  // XCD-aware remap: blocks b, b+8, b+16 ... run on one XCD; give that XCD a
  // contiguous run of logical tiles.
  const int nb = gridDim.x;  // multiple of 8
  const int b = blockIdx.x;
  const int logical = (b & 7) * (nb >> 3) + (b >> 3);
  live = logical < nblocks;
  const int yt = logical % ty_tiles;
  const int zt = logical / ty_tiles;
  y0 = yt * TY;
  z0 = zt * zc;
}

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------

// One work-group tile of the forward step: rows y0 .. y0+TY-1, planes
// z0 .. z0+zc-1.  Must be called by the whole work-group (it synchronises).
template <typename T, int TY>
__device__ __forceinline__ void fwd_tile(const Fused3Args<T>& a, const int y0, const int z0,
                                         T* sm) {
  const int W = a.nx;
  T* sE = sm;                          // [2][2: Ey, Ez][TY+1][W]
  T* sH = sm + 2 * 2 * (TY + 1) * W;   // [2][2: Hy, Hz][TY][W]
#define SE(buf, c, r) sE[(((buf) * 2 + (c)) * (TY + 1) + (r)) * W]
#define SH(buf, c, r) sH[(((buf) * 2 + (c)) * TY + (r)) * W]

  const int z1 = min(a.nz, z0 + a.zc);
  const int x = threadIdx.x;
  const bool xv = x < W;
  const int64_t N = a.N, P = a.P;
  const T* __restrict__ Ein = a.in;
  const T* __restrict__ Hin = a.in + 3 * N;
  T* __restrict__ Eout = a.out;
  T* __restrict__ Hout = a.out + 3 * N;

  const T kh = a.kh;
  const T bxv = xv ? a.bx[x] : T(0);
  const T axv = xv ? a.ax[x] : T(0);
  const bool xf = xv && x >= 1, xl = x + 1 < W;

  // Row r <-> y = y0 - 1 + r, r = 0 .. TY+1.
  T byr[TY + 2], ayr[TY + 2];
  bool rv[TY + 2];
#pragma unroll
  for (int r = 0; r < TY + 2; ++r) {
    const int y = y0 - 1 + r;
    rv[r] = y >= 0 && y < a.ny;
    byr[r] = rv[r] ? a.by[y] : T(0);
    ayr[r] = rv[r] ? a.ay[y] : T(0);
  }
  const int64_t colbase = int64_t(y0 - 1) * W + x;  // row r at colbase + r*W

  T eC[3][TY + 2], eN[3][TY + 2];
  T hp[2][TY];
  const int zs = z0 > 0 ? z0 - 1 : z0;
  {
    const int64_t pb = int64_t(zs) * P + colbase;
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
        eC[c][r] = (rv[r] && xv) ? Ein[c * N + pb + r * W] : T(0);
    }
    if (xv) {
#pragma unroll
      for (int r = 0; r <= TY; ++r) {
        (&SE(zs & 1, 0, r))[x] = eC[1][r];
        (&SE(zs & 1, 1, r))[x] = eC[2][r];
      }
    }
#pragma unroll
    for (int r = 0; r < TY; ++r) hp[0][r] = hp[1][r] = T(0);
  }
  __syncthreads();

  const bool has_src = a.src_comp >= 0;
  const T wsrc = has_src ? a.src_wave[0] : T(0);

  for (int z = zs; z < z1; ++z) {
    const int buf = z & 1;
    const bool zl = z + 1 < a.nz;
    const bool zf = z >= 1;
    const bool doB = z >= z0;
    // ---- loads: E(z+1) rows 0..TY+1, H(z) rows 0..TY, ce(z) rows 1..TY
    T h[3][TY + 1];
    T cez[TY];
    {
      const int64_t pn = int64_t(z + 1) * P + colbase;
      const int64_t pz = int64_t(z) * P + colbase;
#pragma unroll
      for (int r = 0; r < TY + 2; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          eN[c][r] = (zl && rv[r] && xv) ? Ein[c * N + pn + r * W] : T(0);
      }
#pragma unroll
      for (int r = 0; r < TY + 1; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          h[c][r] = (rv[r] && xv) ? Hin[c * N + pz + r * W] : T(0);
      }
#pragma unroll
      for (int r = 1; r <= TY; ++r)
        cez[r - 1] = (doB && rv[r] && xv) ? a.ce[pz + r * W] : T(0);
    }
    // ---- H-like update, rows 0..TY
    T hn[3][TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) {
      const int y = y0 - 1 + r;
      const bool yl = y + 1 < a.ny;
      const T ex = eC[0][r], ey = eC[1][r], ez = eC[2][r];
      const T ey_x1 = xl ? (&SE(buf, 0, r))[x + 1] : T(0);
      const T ez_x1 = xl ? (&SE(buf, 1, r))[x + 1] : T(0);
      const T dEz_dy = yl ? eC[2][r + 1] - ez : T(0);
      const T dEx_dy = yl ? eC[0][r + 1] - ex : T(0);
      const T dEy_dz = zl ? eN[1][r] - ey : T(0);
      const T dEx_dz = zl ? eN[0][r] - ex : T(0);
      const T dEz_dx = xl ? ez_x1 - ez : T(0);
      const T dEy_dx = xl ? ey_x1 - ey : T(0);
      hn[0][r] = byr[r] * h[0][r] - kh * (ayr[r] * dEz_dy - axv * dEy_dz);
      hn[1][r] = bxv * h[1][r] - kh * (axv * dEx_dz - ayr[r] * dEz_dx);
      hn[2][r] = bxv * h[2][r] - kh * (axv * dEy_dx - ayr[r] * dEx_dy);
    }
    if (xv) {
#pragma unroll
      for (int r = 1; r <= TY; ++r) {
        (&SH(buf, 0, r - 1))[x] = hn[1][r];
        (&SH(buf, 1, r - 1))[x] = hn[2][r];
      }
#pragma unroll
      for (int r = 0; r <= TY; ++r) {
        (&SE(buf ^ 1, 0, r))[x] = eN[1][r];
        (&SE(buf ^ 1, 1, r))[x] = eN[2][r];
      }
    }
    __syncthreads();
    // ---- E-like update, rows 1..TY
    if (doB) {
      const int64_t pz = int64_t(z) * P + colbase;
#pragma unroll
      for (int r = 1; r <= TY; ++r) {
        const int y = y0 - 1 + r;
        const bool yf = y >= 1;
        const T hx = hn[0][r], hy = hn[1][r], hz = hn[2][r];
        const T hy_x0 = xf ? (&SH(buf, 0, r - 1))[x - 1] : T(0);
        const T hz_x0 = xf ? (&SH(buf, 1, r - 1))[x - 1] : T(0);
        const T dHz_dy = yf ? hz - hn[2][r - 1] : T(0);
        const T dHx_dy = yf ? hx - hn[0][r - 1] : T(0);
        const T dHy_dz = zf ? hy - hp[1][r - 1] : T(0);
        const T dHx_dz = zf ? hx - hp[0][r - 1] : T(0);
        const T dHz_dx = xf ? hz - hz_x0 : T(0);
        const T dHy_dx = xf ? hy - hy_x0 : T(0);
        const T c = cez[r - 1];
        const T bxy = bxv * byr[r];
        T e[3];
        e[0] = bxy * eC[0][r] + c * (ayr[r] * dHz_dy - axv * dHy_dz);
        e[1] = bxy * eC[1][r] + c * (axv * dHx_dz - ayr[r] * dHz_dx);
        e[2] = bxy * eC[2][r] + c * (axv * dHy_dx - ayr[r] * dHx_dy);
        if (has_src) {
          const Box& b = a.src_box;
          const int dz = z - b.z0, dy = y - b.y0, dx = x - b.x0;
          if (dz >= 0 && dz < b.nz && dy >= 0 && dy < b.ny && dx >= 0 && dx < b.nx) {
            const T s = wsrc * a.src_prof[(int64_t(dz) * b.ny + dy) * b.nx + dx];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
              if (cc == a.src_comp) e[cc] += s;
          }
        }
        if (rv[r] && xv) {
          const int64_t i = pz + r * W;
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            Eout[cc * N + i] = e[cc];
            Hout[cc * N + i] = hn[cc][r];
          }
          if (a.es) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) a.es[cc * N + i] = e[cc];
          }
        }
      }
    }
#pragma unroll
    for (int r = 1; r <= TY; ++r) {
      hp[0][r - 1] = hn[0][r];
      hp[1][r - 1] = hn[1][r];
    }
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) eC[c][r] = eN[c][r];
    }
  }
#undef SE
#undef SH
}

template <typename T, int TY>
__global__ __launch_bounds__(256) void k3_fused_fwd(Fused3Args<T> a) {
  extern __shared__ unsigned char smem_raw[];
  int y0, z0;
  bool live;
  block_tile<TY>(a.nblocks, a.ty_tiles, a.zc, y0, z0, live);
  if (!live) return;  // uniform per work-group
  fwd_tile<T, TY>(a, y0, z0, reinterpret_cast<T*>(smem_raw));
}

// ---------------------------------------------------------------------------
// Adjoint.  Input lam = (lamE^{m+1}, muH^{m+1}); output (lamE^m, muH^m).
//   u = ce lamE (E-like input scaled by its own ce)
//   muHx' = by muHx + ax ([z>=1] uy - [z+1<nz] uy(z+1))
//                   - ([y>=1] ay uz - [y+1<ny] ay(y+1) uz(y+1))
//   muHy' = bx muHy + ([x>=1] ax uz - [x+1<nx] ax(x+1) uz(x+1))
//                   - ax ([z>=1] ux - [z+1<nz] ux(z+1))
//   muHz' = bx muHz + ([y>=1] ay ux - [y+1<ny] ay(y+1) ux(y+1))
//                   - ay ([x>=1] uy - [x+1<nx] uy(x+1))
//   tx = kh (D+y^T(ay muHz') - ax D+z^T(muHy'))
//   ty = kh (ax D+z^T(muHx') - D+x^T(ax muHz'))
//   tz = kh (ay D+x^T(muHy') - D+y^T(ay muHx'))
//   with (D+^T g)[j] = [j>=1] g[j-1] - [j<=n-2] g[j].
// ---------------------------------------------------------------------------

template <typename T, int TY>
__device__ __forceinline__ void adj_tile(const Fused3Args<T>& a, const int y0, const int z0,
                                         T* sm) {
  const int W = a.nx;
  T* sE = sm;                          // [2][2: uy, uz][TY+1][W]
  T* sH = sm + 2 * 2 * (TY + 1) * W;   // [2][2: muHy, muHz][TY][W]
#define SE(buf, c, r) sE[(((buf) * 2 + (c)) * (TY + 1) + (r)) * W]
#define SH(buf, c, r) sH[(((buf) * 2 + (c)) * TY + (r)) * W]

  const int z1 = min(a.nz, z0 + a.zc);
  const int x = threadIdx.x;
  const bool xv = x < W;
  const int64_t N = a.N, P = a.P;
  const T* __restrict__ Lin = a.in;
  const T* __restrict__ Min = a.in + 3 * N;
  T* __restrict__ Lout = a.out;
  T* __restrict__ Mout = a.out + 3 * N;

  const T kh = a.kh;
  const T bxv = xv ? a.bx[x] : T(0);
  const T axv = xv ? a.ax[x] : T(0);
  const T axp = (x + 1 < W) ? a.ax[x + 1] : T(0);
  const T axm = (xv && x >= 1) ? a.ax[x - 1] : T(0);
  const bool xf = xv && x >= 1, xl = x + 1 < W, xin = x <= W - 2;

  T byr[TY + 2], ayr[TY + 2];
  bool rv[TY + 2];
#pragma unroll
  for (int r = 0; r < TY + 2; ++r) {
    const int y = y0 - 1 + r;
    rv[r] = y >= 0 && y < a.ny;
    byr[r] = rv[r] ? a.by[y] : T(0);
    ayr[r] = rv[r] ? a.ay[y] : T(0);
  }
  const int64_t colbase = int64_t(y0 - 1) * W + x;
  // Branch-free loads: clamped row / column (always in bounds), masked after.
  const int xc = xv ? x : W - 1;
  int yrow[TY + 2];
#pragma unroll
  for (int r = 0; r < TY + 2; ++r) yrow[r] = min(max(y0 - 1 + r, 0), a.ny - 1);
  auto ld = [&](const T* base, int64_t plane, int r) -> T {
    return (base + plane * P + int64_t(yrow[r]) * W)[xc];
  };

  T uC[3][TY + 2], uN[3][TY + 2];
  T lC[3][TY], lN[3][TY];  // lamE own rows 1..TY at planes z, z+1
  T mp[2][TY];             // muHx', muHy' of plane z-1, rows 1..TY
  const int zs = z0 > 0 ? z0 - 1 : z0;
  {
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) {
      const bool ok = rv[r] && xv;
      const T cv = ld(a.ce, zs, r);
      const T c = ok ? cv : T(0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T lv = ld(Lin + k * N, zs, r);
        const T l = ok ? lv : T(0);
        uC[k][r] = c * l;
        if (r >= 1 && r <= TY) lC[k][r - 1] = l;
      }
    }
    if (xv) {
#pragma unroll
      for (int r = 0; r <= TY; ++r) {
        (&SE(zs & 1, 0, r))[x] = uC[1][r];
        (&SE(zs & 1, 1, r))[x] = uC[2][r];
      }
    }
#pragma unroll
    for (int r = 0; r < TY; ++r) mp[0][r] = mp[1][r] = T(0);
  }
  __syncthreads();

  const bool has_src = a.src_comp >= 0;
  const T wsrc = has_src ? a.src_wave[0] : T(0);

  for (int z = zs; z < z1; ++z) {
    const int buf = z & 1;
    const bool zl = z + 1 < a.nz;
    const bool zf = z >= 1;
    const bool zin = z <= a.nz - 2;
    const bool doB = z >= z0;
    T m[3][TY + 1];
    T em[3][TY];
    {
      const int zn = zl ? z + 1 : z;
#pragma unroll
      for (int r = 0; r < TY + 2; ++r) {
        const bool ok = zl && rv[r] && xv;
        const T cv = ld(a.ce, zn, r);
        const T c = ok ? cv : T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const T lv = ld(Lin + k * N, zn, r);
          const T l = ok ? lv : T(0);
          uN[k][r] = c * l;
          if (r >= 1 && r <= TY) lN[k][r - 1] = l;
        }
      }
#pragma unroll
      for (int r = 0; r < TY + 1; ++r) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const T v = ld(Min + k * N, z, r);
          m[k][r] = (rv[r] && xv) ? v : T(0);
        }
      }
      if (doB) {
#pragma unroll
        for (int r = 1; r <= TY; ++r) {
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const T v = ld(a.em + k * N, z, r);
            em[k][r - 1] = (rv[r] && xv) ? v : T(0);
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < TY; ++r) em[0][r] = em[1][r] = em[2][r] = T(0);
      }
    }
    // ---- mu update, rows 0..TY
    T mn[3][TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) {
      const int y = y0 - 1 + r;
      const bool yf = y >= 1, yl = y + 1 < a.ny;
      const T ux = uC[0][r], uy = uC[1][r], uz = uC[2][r];
      const T uy_x1 = xl ? (&SE(buf, 0, r))[x + 1] : T(0);
      const T uz_x1 = xl ? (&SE(buf, 1, r))[x + 1] : T(0);
      const T tz_y = (zf ? uy : T(0)) - (zl ? uN[1][r] : T(0));
      const T tz_x = (zf ? ux : T(0)) - (zl ? uN[0][r] : T(0));
      const T ty_z = (yf ? ayr[r] * uz : T(0)) - (yl ? ayr[r + 1] * uC[2][r + 1] : T(0));
      const T ty_x = (yf ? ayr[r] * ux : T(0)) - (yl ? ayr[r + 1] * uC[0][r + 1] : T(0));
      const T tx_z = (xf ? axv * uz : T(0)) - (xl ? axp * uz_x1 : T(0));
      const T tx_y = (xf ? uy : T(0)) - (xl ? uy_x1 : T(0));
      mn[0][r] = byr[r] * m[0][r] + axv * tz_y - ty_z;
      mn[1][r] = bxv * m[1][r] + tx_z - axv * tz_x;
      mn[2][r] = bxv * m[2][r] + ty_x - ayr[r] * tx_y;
    }
    if (xv) {
#pragma unroll
      for (int r = 1; r <= TY; ++r) {
        (&SH(buf, 0, r - 1))[x] = mn[1][r];
        (&SH(buf, 1, r - 1))[x] = mn[2][r];
      }
#pragma unroll
      for (int r = 0; r <= TY; ++r) {
        (&SE(buf ^ 1, 0, r))[x] = uN[1][r];
        (&SE(buf ^ 1, 1, r))[x] = uN[2][r];
      }
    }
    __syncthreads();
    // ---- lam update + gradient, rows 1..TY
    if (doB) {
      const int64_t pz = int64_t(z) * P + colbase;
#pragma unroll
      for (int r = 1; r <= TY; ++r) {
        const int y = y0 - 1 + r;
        const bool yf = y >= 1, yin = y <= a.ny - 2;
        const T mx = mn[0][r], my = mn[1][r], mz = mn[2][r];
        const T my_x0 = xf ? (&SH(buf, 0, r - 1))[x - 1] : T(0);
        const T mz_x0 = xf ? (&SH(buf, 1, r - 1))[x - 1] : T(0);
        const T y_mz = (yf ? ayr[r - 1] * mn[2][r - 1] : T(0)) - (yin ? ayr[r] * mz : T(0));
        const T y_mx = (yf ? ayr[r - 1] * mn[0][r - 1] : T(0)) - (yin ? ayr[r] * mx : T(0));
        const T z_my = (zf ? mp[1][r - 1] : T(0)) - (zin ? my : T(0));
        const T z_mx = (zf ? mp[0][r - 1] : T(0)) - (zin ? mx : T(0));
        const T x_mz = (xf ? axm * mz_x0 : T(0)) - (xin ? axv * mz : T(0));
        const T x_my = (xf ? my_x0 : T(0)) - (xin ? my : T(0));
        T t[3];
        t[0] = kh * (y_mz - axv * z_my);
        t[1] = kh * (axv * z_mx - x_mz);
        t[2] = kh * (ayr[r] * x_my - y_mx);
        const T bxy = bxv * byr[r];
        T s = em[0][r - 1] * t[0] + em[1][r - 1] * t[1] + em[2][r - 1] * t[2];
        if (has_src) {
          const Box& b = a.src_box;
          const int dz = z - b.z0, dy = y - b.y0, dx = x - b.x0;
          if (dz >= 0 && dz < b.nz && dy >= 0 && dy < b.ny && dx >= 0 && dx < b.nx) {
            const T w = wsrc * a.src_prof[(int64_t(dz) * b.ny + dy) * b.nx + dx];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
              if (cc == a.src_comp) s -= w * lC[cc][r - 1];
          }
        }
        if (rv[r] && xv) {
          const int64_t i = pz + r * W;
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            Lout[cc * N + i] = bxy * lC[cc][r - 1] + t[cc];
            Mout[cc * N + i] = mn[cc][r];
          }
          a.S[i] += s;
        }
      }
    }
#pragma unroll
    for (int r = 1; r <= TY; ++r) {
      mp[0][r - 1] = mn[0][r];
      mp[1][r - 1] = mn[1][r];
    }
#pragma unroll
    for (int r = 0; r < TY + 2; ++r) {
#pragma unroll
      for (int k = 0; k < 3; ++k) uC[k][r] = uN[k][r];
    }
#pragma unroll
    for (int r = 0; r < TY; ++r) {
#pragma unroll
      for (int k = 0; k < 3; ++k) lC[k][r] = lN[k][r];
    }
  }
#undef SE
#undef SH
}

template <typename T, int TY>
__global__ __launch_bounds__(256) void k3_fused_adj(Fused3Args<T> a) {
  extern __shared__ unsigned char smem_raw[];
  int y0, z0;
  bool live;
  block_tile<TY>(a.nblocks, a.ty_tiles, a.zc, y0, z0, live);
  if (!live) return;
  adj_tile<T, TY>(a, y0, z0, reinterpret_cast<T*>(smem_raw));
}

// ---------------------------------------------------------------------------
// Batched step over many designs of different grids (one launch per step).
// Each work-group walks the global tile list with stride gridDim.x (tiles of
// design d are [tile_start[d], tile_start[d+1])), XCD-remapped like the
// single-grid kernels.  The step index comes from a device counter, so the
// same launch can be replayed from a captured graph for every segment.
// A launch at counter n (forward) reads the state from
// buf[batch_in_slot(layout, n)] and writes the other buffer; an adjoint
// launch at m reads lam^{m+1} from buf[batch_in_slot(layout, m + 1)].
// ---------------------------------------------------------------------------

template <typename T>
__device__ inline int batch_design(const Batch3Design<T>* D, int n_designs, int t) {
  int d = 0;
  while (d + 1 < n_designs && D[d + 1].tile_start <= t) ++d;
  return d;
}

template <typename T, int TY>
__global__ __launch_bounds__(256) void k3_batch_fwd(Batch3Launch L) {
  extern __shared__ unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  const Batch3Design<T>* D = static_cast<const Batch3Design<T>*>(L.designs);
  const int n = *L.step;
  const int ib = batch_in_slot(L.layout, n);
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int lw = (b & 7) * (G >> 3) + (b >> 3);
  for (int t = lw; t < L.total_tiles; t += G) {
    const int d = batch_design(D, L.n_designs, t);
    Fused3Args<T> a = D[d].a;
    a.in = D[d].buf[ib];
    a.out = D[d].buf[ib ^ 1];
    a.es = batch_es(D[d], L.store, n, L.K);
    if (a.src_comp >= 0) a.src_wave = D[d].wave + n;
    const int lt = t - D[d].tile_start;
    fwd_tile<T, TY>(a, (lt % a.ty_tiles) * TY, (lt / a.ty_tiles) * a.zc, sm);
    __syncthreads();  // LDS is reused by the next tile
  }
}

template <typename T, int TY>
__global__ __launch_bounds__(256) void k3_batch_adj(Batch3Launch L) {
  extern __shared__ unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  const Batch3Design<T>* D = static_cast<const Batch3Design<T>*>(L.designs);
  const int m = *L.step;
  const int ib = batch_in_slot(L.layout, m + 1);
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int lw = (b & 7) * (G >> 3) + (b >> 3);
  for (int t = lw; t < L.total_tiles; t += G) {
    const int d = batch_design(D, L.n_designs, t);
    Fused3Args<T> a = D[d].a;
    a.in = D[d].buf[ib];
    a.out = D[d].buf[ib ^ 1];
    a.em = batch_em(D[d], m, L.K);
    if (a.src_comp >= 0) a.src_wave = D[d].wave + m;
    const int lt = t - D[d].tile_start;
    adj_tile<T, TY>(a, (lt % a.ty_tiles) * TY, (lt / a.ty_tiles) * a.zc, sm);
    __syncthreads();
  }
}

__global__ void k_counter_add(int* c, int delta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c += delta;
}

__global__ void k_counter_set(int* c, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = v;
}

// Batched DFT monitors: blockIdx.y = design.  adjoint = 0: after forward step
// n, acc += E^{n+1} (cos, -sin)(phase[n]).  adjoint = 1: after adjoint step
// m, d = Re(g e^{i phase[m-1]}); lam^m_c += d; S += E^m_c d.
template <typename T>
__global__ void k_batch_monitor(const Batch3Design<T>* D, int K, const int* step, int adjoint,
                                int koff, int layout) {
  const Batch3Design<T>& dz = D[blockIdx.y];
  const Box b = dz.mbox;
  if (dz.m_nf <= 0) return;
  const int64_t cells = b.cells();
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  // sample k of the launch at counter c (k = c + koff; the adjoint pair's
  // second sample has koff = -1), read from that launch's output buffer
  const int c = *step;
  const int k = c + koff;
  const int obuf = batch_in_slot(layout, adjoint ? c + 1 : c) ^ 1;
  const int bx = int(i % b.nx);
  const int64_t rr = i / b.nx;
  const int by = int(rr % b.ny);
  const int bz = int(rr / b.ny);
  const int ny = dz.a.ny, nx = dz.a.nx;
  const int64_t N = dz.a.N;
  const int64_t idx = (int64_t(b.z0 + bz) * ny + (b.y0 + by)) * nx + (b.x0 + bx);
  const int mask = dz.m_mask;
  const int ncomp = (mask & 1) + ((mask >> 1) & 1) + ((mask >> 2) & 1);
  const int nf = dz.m_nf;
  if (!adjoint) {
    const T* E = dz.buf[obuf];
    const double* ph = dz.m_phase + int64_t(k) * nf * 2;
    int ci = 0;
    for (int c = 0; c < 3; ++c) {
      if (!((mask >> c) & 1)) continue;
      const double v = double(E[c * N + idx]);
      for (int f = 0; f < nf; ++f) {
        double* acc = dz.m_acc + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
        acc[0] += v * ph[2 * f];
        acc[1] -= v * ph[2 * f + 1];
      }
      ++ci;
    }
  } else {
    T* Lm = dz.buf[obuf];
    const T* Em = batch_em(dz, k, K);
    const double* ph = dz.m_phase + int64_t(k - 1) * nf * 2;
    int ci = 0;
    T s = T(0);
    for (int c = 0; c < 3; ++c) {
      if (!((mask >> c) & 1)) continue;
      double dv = 0.0;
      for (int f = 0; f < nf; ++f) {
        const double* g = dz.m_grad + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
        dv += g[0] * ph[2 * f] - g[1] * ph[2 * f + 1];
      }
      const T dt = T(dv);
      Lm[c * N + idx] += dt;
      s += Em[c * N + idx] * dt;
      ++ci;
    }
    dz.a.S[idx] += s;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <typename T, int TY>
static int launch_fwd_ty(const Fused3Args<T>& a, hipStream_t s) {
  const int threads = ((a.nx + 63) / 64) * 64;
  const int nb = ((a.nblocks + 7) / 8) * 8;
  const size_t smem = sizeof(T) * size_t(2 * 2 * (TY + 1) + 2 * 2 * TY) * a.nx;
  static bool attr_set = false;
  if (!attr_set) {
    DNX_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k3_fused_fwd<T, TY>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int(sizeof(T) * (2 * 2 * (TY + 1) + 2 * 2 * TY) *
                                          FUSED_MAX_W)));
    attr_set = true;
  }
  hipLaunchKernelGGL((k3_fused_fwd<T, TY>), dim3(nb), dim3(threads), smem, s, a);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T, int TY>
static int launch_adj_ty(const Fused3Args<T>& a, hipStream_t s) {
  const int threads = ((a.nx + 63) / 64) * 64;
  const int nb = ((a.nblocks + 7) / 8) * 8;
  const size_t smem = sizeof(T) * size_t(2 * 2 * (TY + 1) + 2 * 2 * TY) * a.nx;
  static bool attr_set = false;
  if (!attr_set) {
    DNX_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k3_fused_adj<T, TY>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int(sizeof(T) * (2 * 2 * (TY + 1) + 2 * 2 * TY) *
                                          FUSED_MAX_W)));
    attr_set = true;
  }
  hipLaunchKernelGGL((k3_fused_adj<T, TY>), dim3(nb), dim3(threads), smem, s, a);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_fused3_fwd(const Fused3Args<T>& a, int ty, hipStream_t s) {
  if (ty == 8) return launch_fwd_ty<T, 8>(a, s);
  if (ty == 4) return launch_fwd_ty<T, 4>(a, s);
  if (ty == 2) return launch_fwd_ty<T, 2>(a, s);
  set_error("fused3: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template <typename T>
int launch_fused3_adj(const Fused3Args<T>& a, int ty, hipStream_t s) {
  if (ty == 8) return launch_adj_ty<T, 8>(a, s);
  if (ty == 4) return launch_adj_ty<T, 4>(a, s);
  if (ty == 2) return launch_adj_ty<T, 2>(a, s);
  set_error("fused3: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template <typename T, int TY>
static int prepare_batch_ty(bool adjoint, int max_w, int* blocks_per_cu) {
  const int smem_max = int(sizeof(T) * (2 * 2 * (TY + 1) + 2 * 2 * TY) * FUSED_MAX_W);
  const void* f = adjoint ? reinterpret_cast<const void*>(&k3_batch_adj<T, TY>)
                          : reinterpret_cast<const void*>(&k3_batch_fwd<T, TY>);
  DNX_CHECK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, smem_max));
  const int threads = ((max_w + 63) / 64) * 64;
  const size_t smem = sizeof(T) * size_t(2 * 2 * (TY + 1) + 2 * 2 * TY) * max_w;
  int nb = 0;
  DNX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, threads, smem));
  *blocks_per_cu = nb > 0 ? nb : 1;
  return DNX_OK;
}

// Sets the kernel attributes (never called while a stream is captured) and
// reports how many work-groups of this kernel fit on one CU, which sizes
// the persistent grid.
template <typename T>
int prepare_batch3(int ty, bool adjoint, int max_w, int* blocks_per_cu) {
  if (ty == 8) return prepare_batch_ty<T, 8>(adjoint, max_w, blocks_per_cu);
  if (ty == 4) return prepare_batch_ty<T, 4>(adjoint, max_w, blocks_per_cu);
  if (ty == 2) return prepare_batch_ty<T, 2>(adjoint, max_w, blocks_per_cu);
  set_error("batch3: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template <typename T, int TY>
static int launch_batch_ty(const Batch3Launch& L, bool adjoint, int max_w, int grid,
                           hipStream_t s) {
  const int threads = ((max_w + 63) / 64) * 64;
  const size_t smem = sizeof(T) * size_t(2 * 2 * (TY + 1) + 2 * 2 * TY) * max_w;
  const int nb = ((grid + 7) / 8) * 8;
  if (adjoint) {
    hipLaunchKernelGGL((k3_batch_adj<T, TY>), dim3(nb), dim3(threads), smem, s, L);
  } else {
    hipLaunchKernelGGL((k3_batch_fwd<T, TY>), dim3(nb), dim3(threads), smem, s, L);
  }
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

// Launch one batched step; prepare_batch3 must have been called for (ty, kind).
template <typename T>
int launch_batch3(const Batch3Launch& L, bool adjoint, int ty, int max_w, int grid,
                  hipStream_t s) {
  if (ty == 8) return launch_batch_ty<T, 8>(L, adjoint, max_w, grid, s);
  if (ty == 4) return launch_batch_ty<T, 4>(L, adjoint, max_w, grid, s);
  if (ty == 2) return launch_batch_ty<T, 2>(L, adjoint, max_w, grid, s);
  set_error("batch3: unsupported tile height %d", ty);
  return DNX_EINVAL;
}

template <typename T>
int launch_batch_monitor(const void* designs, int n_designs, int64_t max_cells, int K,
                         const int* step, bool adjoint, int koff, int layout, hipStream_t s) {
  if (max_cells <= 0) return DNX_OK;
  hipLaunchKernelGGL(k_batch_monitor<T>, dim3(cdiv(max_cells, 256), n_designs), dim3(256), 0, s,
                     static_cast<const Batch3Design<T>*>(designs), K, step, adjoint ? 1 : 0,
                     koff, layout);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

int launch_counter_add(int* c, int delta, hipStream_t s) {
  hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, s, c, delta);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

int launch_counter_set(int* c, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_counter_set, dim3(1), dim3(64), 0, s, c, v);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template int launch_fused3_fwd<float>(const Fused3Args<float>&, int, hipStream_t);
template int launch_fused3_fwd<double>(const Fused3Args<double>&, int, hipStream_t);
template int launch_fused3_adj<float>(const Fused3Args<float>&, int, hipStream_t);
template int launch_fused3_adj<double>(const Fused3Args<double>&, int, hipStream_t);
template int prepare_batch3<float>(int, bool, int, int*);
template int prepare_batch3<double>(int, bool, int, int*);
template int launch_batch3<float>(const Batch3Launch&, bool, int, int, int, hipStream_t);
template int launch_batch3<double>(const Batch3Launch&, bool, int, int, int, hipStream_t);
template int launch_batch_monitor<float>(const void*, int, int64_t, int, const int*, bool, int,
                                         int, hipStream_t);
template int launch_batch_monitor<double>(const void*, int, int64_t, int, const int*, bool, int,
                                          int, hipStream_t);

}  // namespace dnx

extern "C" int32_t dnx_fdtd3d_set_tiles(int32_t dtype, int32_t kernel, int32_t ty,
                                        int32_t zc) {
  using namespace dnx;
  DNX_REQUIRE(dtype == DNX_F32 || dtype == DNX_F64, "set_tiles: bad dtype %d", dtype);
  DNX_REQUIRE(kernel >= 0 && kernel <= 3,
              "set_tiles: kernel must be 0 (forward), 1 (adjoint), 2 (two-step forward) or 3 "
              "(two-step adjoint)");
  DNX_REQUIRE(ty == 2 || ty == 4 || (ty == 8 && kernel < 2) || (ty == 1 && kernel == 3) ||
                  (ty == 0 && kernel >= 2),
              "set_tiles: ty must be 2, 4 or 8 (2 or 4 for the two-step forward, 1, 2 or 4 "
              "for the two-step adjoint, 0 to switch a two-step kernel off; got %d)", ty);
  DNX_REQUIRE(zc >= 1 && zc <= 4096, "set_tiles: zc must be in [1, 4096] (got %d)", zc);
  const int d = dtype == DNX_F64 ? 1 : 0;
  g_ty[d][kernel] = ty;
  g_zc[d][kernel] = zc;
  return DNX_OK;
}

extern "C" int32_t dnx_fdtd3d_get_tiles(int32_t dtype, int32_t kernel, int32_t* ty,
                                        int32_t* zc) {
  using namespace dnx;
  DNX_REQUIRE(ty && zc, "get_tiles: null output");
  DNX_REQUIRE(dtype == DNX_F32 || dtype == DNX_F64, "get_tiles: bad dtype %d", dtype);
  DNX_REQUIRE(kernel >= 0 && kernel <= 3, "get_tiles: kernel must be 0-3 (got %d)", kernel);
  const Fused3Tiles t = fused3_tiles(dtype, kernel);
  *ty = t.ty;
  *zc = t.zc;
  return DNX_OK;
}

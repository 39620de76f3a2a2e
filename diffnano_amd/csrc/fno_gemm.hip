// BF16 MFMA GEMM with FP32 accumulation for the FNO surrogate (gfx950).
//
//   C[b] = act( alpha * A[b] B[b] + beta * Cin[b] + bias )     (per batch b)
//
// A (M x K), B (K x N), C (M x N) are addressed through element strides, so
// transposed operands cost nothing: A[m][k] = A + b*sAb + m*sAm + k*sAk.
// Inputs are FP32 or BF16 in memory; both are staged in LDS as BF16 (round
// to nearest even) and multiplied with v_mfma_f32_16x16x32_bf16; the
// accumulators and the epilogue are FP32.  Optional epilogue: bias per row
// or per column, GELU (erf form), storing the pre-activation for the
// backward pass.
//
// Tiling: 256 threads = 4 waves in a 2 x 2 arrangement over a 128 x 128
// output tile; each wave owns 64 x 64 = 4 x 4 MFMA tiles of 16 x 16 and
// steps K by 32.  LDS holds the A tile as [m][k] and the B tile as [n][k]
// (k contiguous, padded), so each lane's 8 consecutive k values are one
// 16-byte LDS read.  MFMA 16x16x32 lane layout: A lane l holds row l%16,
// k = (l/16)*8 .. +8; B lane l holds column l%16, the same k; the result
// lane l holds rows (l/16)*4 .. +4 of column l%16.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdlib>
#include <type_traits>

#include "../../include/dnx_fno.h"
#include "dnx_common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace dnx {

using GemmArgs = dnx_gemm_args;

constexpr int GBM = 128, GBN = 128, GBK = 32;
constexpr int LDK = GBK + 8;  // padded row (bf16 elements)

__device__ inline unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ inline unsigned short load_bf(const void* p, int64_t i, int is_bf16) {
  if (is_bf16) return static_cast<const unsigned short*>(p)[i];
  return f2bf(static_cast<const float*>(p)[i]);
}

__device__ inline float gelu(float v) { return 0.5f * v * (1.0f + erff(v * 0.70710678118654752f)); }

// Elementwise helpers of the FNO (FP32).
// GELU backward: dx = dy * gelu'(pre)
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  __shared__ unsigned short As[GBM * LDK];
  __shared__ unsigned short Bs[GBN * LDK];
  const int b = blockIdx.z;
  const int m0 = blockIdx.y * GBM;
  const int n0 = blockIdx.x * GBN;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = (wave >> 1) * 64;  // wave's row offset in the tile
  const int wn = (wave & 1) * 64;

  const int64_t aoff = int64_t(b) * g.sAb;
  const int64_t boff = int64_t(b) * g.sBb;
  // Loader orientation: walk the operand's contiguous axis across threads.
  const bool a_kfast = g.sAk == 1;
  const bool b_nfast = g.sBn == 1;

  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += GBK) {
    // A tile: GBM x GBK = 4096 elements, 16 per thread.
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = r * 256 + t;
      int mm, kk;
      if (a_kfast) {
        mm = e / GBK;
        kk = e % GBK;
      } else {
        mm = e % GBM;
        kk = e / GBM;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      unsigned short v = 0;
      if (gm < g.M && gk < g.K) v = load_bf(g.A, aoff + gm * g.sAm + gk * g.sAk, g.a_bf16);
      As[mm * LDK + kk] = v;
    }
    // B tile: GBK x GBN, stored as [n][k].
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = r * 256 + t;
      int nn, kk;
      if (b_nfast) {
        nn = e % GBN;
        kk = e / GBN;
      } else {
        nn = e / GBK;
        kk = e % GBK;
      }
      const int gn = n0 + nn, gk = k0 + kk;
      unsigned short v = 0;
      if (gn < g.N && gk < g.K) v = load_bf(g.B, boff + gk * g.sBk + gn * g.sBn, g.b_bf16);
      Bs[nn * LDK + kk] = v;
    }
    __syncthreads();
    const int kq = (lane >> 4) * 8;
    bf16x8 af[4], bfr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const unsigned short* pa = &As[(wm + i * 16 + (lane & 15)) * LDK + kq];
      const unsigned short* pb = &Bs[(wn + i * 16 + (lane & 15)) * LDK + kq];
      af[i] = *reinterpret_cast<const bf16x8*>(pa);
      bfr[i] = *reinterpret_cast<const bf16x8*>(pb);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  // Epilogue.
  const int64_t coff = int64_t(b) * g.sCb;
  const int64_t ioff = int64_t(b) * g.sIb;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + wn + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        if (m >= g.M || n >= g.N) continue;
        float v = g.alpha * acc[i][j][r];
        if (g.beta != 0.f) v += g.beta * g.Cin[ioff + m * g.sIm + n * g.sIn];
        if (g.bias_mode == 1) v += g.bias[n];
        else if (g.bias_mode == 2) v += g.bias[m];
        const int64_t ci = coff + m * g.sCm + n * g.sCn;
        if (g.pre) g.pre[ci] = v;
        if (g.act == 1) v = gelu(v);
        g.C[ci] = v;
        if (g.c_bf16) g.c_bf16[ci] = f2bf(v);
      }
    }
  }
}

// ---- NT fast path: both operands contiguous along K ----------------------
//
// A[m][k] = A[m*sAm + k], B[k][n] = B[n*sBn + k] (B stored as [n][k]).  Tile
// 128 x 128 x 64, 256 threads (4 waves, 2 x 2, 64 x 64 each).  Each K step
// loads the next tile into registers with 16-byte global loads (FP32 inputs
// are rounded to BF16 on the way), computes 2 x 16 MFMAs per wave from the
// current LDS buffer, then writes the registers into the other LDS buffer:
// one barrier per K step.  Preconditions (checked by the dispatcher): K % 64
// == 0; sAm, sBn, sAb, sBb multiples of 8 elements; A, B 16-byte aligned.

constexpr int NBK = 64;
constexpr int NLD = NBK + 8;  // padded LDS row: 144 bytes

template <bool BF>
struct TileLoad;

template <>
struct TileLoad<false> {  // FP32 in memory: 128 rows x 64 k = 2048 float4, 8 per thread
  float4 v[8];
  __device__ void load(const float* base, int64_t ld, int row0, int nrows, int k0, int t) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 4, c4 = e & 15;
      if (row0 + row < nrows)
        v[r] = *reinterpret_cast<const float4*>(base + int64_t(row0 + row) * ld + k0 + c4 * 4);
      else
        v[r] = float4{0.f, 0.f, 0.f, 0.f};
    }
  }
  __device__ void store(unsigned short* lds, int t) const {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 4, c4 = e & 15;
      uint2 p;
      p.x = uint32_t(f2bf(v[r].x)) | (uint32_t(f2bf(v[r].y)) << 16);
      p.y = uint32_t(f2bf(v[r].z)) | (uint32_t(f2bf(v[r].w)) << 16);
      *reinterpret_cast<uint2*>(lds + row * NLD + c4 * 4) = p;
    }
  }
};

template <>
struct TileLoad<true> {  // BF16 in memory: 128 rows x 64 k = 1024 x 16 B, 4 per thread
  uint4 v[4];
  __device__ void load(const unsigned short* base, int64_t ld, int row0, int nrows, int k0,
                       int t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 3, c8 = e & 7;
      if (row0 + row < nrows)
        v[r] = *reinterpret_cast<const uint4*>(base + int64_t(row0 + row) * ld + k0 + c8 * 8);
      else
        v[r] = uint4{0u, 0u, 0u, 0u};
    }
  }
  __device__ void store(unsigned short* lds, int t) const {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 3, c8 = e & 7;
      *reinterpret_cast<uint4*>(lds + row * NLD + c8 * 8) = v[r];
    }
  }
};

template <bool ABF, bool BBF>
__global__ __launch_bounds__(256) void k_gemm_nt(GemmArgs g) {
  __shared__ unsigned short As[2][GBM * NLD];
  __shared__ unsigned short Bs[2][GBN * NLD];
  using TA = typename std::conditional<ABF, unsigned short, float>::type;
  using TB = typename std::conditional<BBF, unsigned short, float>::type;
  const int b = blockIdx.z;
  const int m0 = blockIdx.y * GBM;
  const int n0 = blockIdx.x * GBN;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = (wave >> 1) * 64;
  const int wn = (wave & 1) * 64;
  const TA* A = static_cast<const TA*>(g.A) + int64_t(b) * g.sAb;
  const TB* B = static_cast<const TB*>(g.B) + int64_t(b) * g.sBb;

  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  TileLoad<ABF> la;
  TileLoad<BBF> lb;
  la.load(A, g.sAm, m0, g.M, 0, t);
  lb.load(B, g.sBn, n0, g.N, 0, t);
  la.store(As[0], t);
  lb.store(Bs[0], t);
  __syncthreads();
  const int nk = g.K / NBK;
  const int kq = (lane >> 4) * 8;
  const int lr = lane & 15;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      la.load(A, g.sAm, m0, g.M, (kt + 1) * NBK, t);
      lb.load(B, g.sBn, n0, g.N, (kt + 1) * NBK, t);
    }
#pragma unroll
    for (int ks = 0; ks < NBK; ks += 32) {
      bf16x8 af[4], bfr[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        af[i] = *reinterpret_cast<const bf16x8*>(&As[cur][(wm + i * 16 + lr) * NLD + ks + kq]);
        bfr[i] = *reinterpret_cast<const bf16x8*>(&Bs[cur][(wn + i * 16 + lr) * NLD + ks + kq]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    if (more) {
      la.store(As[cur ^ 1], t);
      lb.store(Bs[cur ^ 1], t);
    }
    __syncthreads();
  }

  const int64_t coff = int64_t(b) * g.sCb;
  const int64_t ioff = int64_t(b) * g.sIb;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + wn + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        if (m >= g.M || n >= g.N) continue;
        float v = g.alpha * acc[i][j][r];
        if (g.beta != 0.f) v += g.beta * g.Cin[ioff + m * g.sIm + n * g.sIn];
        if (g.bias_mode == 1) v += g.bias[n];
        else if (g.bias_mode == 2) v += g.bias[m];
        const int64_t ci = coff + m * g.sCm + n * g.sCn;
        if (g.pre) g.pre[ci] = v;
        if (g.act == 1) v = gelu(v);
        g.C[ci] = v;
      }
    }
  }
}

// ---- Layout-specialised tile kernel --------------------------------------
//
// Same 128 x 128 x 64 tile, LDS layout ([row][k], padded), register
// prefetch and one barrier per K step as k_gemm_nt, for any operand layout:
// an operand contiguous along K is loaded 16 bytes at a time along K; an
// operand contiguous along its row index (M for A, N for B) is loaded as
// 8 scalar loads along K per thread, consecutive lanes reading consecutive
// rows (so every load instruction is one contiguous 256-byte span per wave),
// and packed to 8 BF16 = one 16-byte LDS store.  K need not be a multiple
// of 64 (the tail is zero-filled); a K-contiguous operand needs K, its row
// stride and batch stride to be multiples of 8 and a 16-byte aligned base.

template <bool BF, bool KFAST>
struct TileLoadG;

template <bool BF>
struct TileLoadG<BF, true> {  // contiguous along K: 8 k per 16 bytes (BF16) or 2 x float4
  using TE = typename std::conditional<BF, unsigned short, float>::type;
  uint4 v[4];
  __device__ void load(const TE* base, int64_t ld, int64_t /*sk*/, int row0, int nrows, int k0,
                       int K, int t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 3, c8 = e & 7;
      const int k = k0 + c8 * 8;
      const bool ok = row0 + row < nrows && k < K;
      const TE* p = base + int64_t(row0 + row) * ld + k;
      if constexpr (BF) {
        v[r] = ok ? *reinterpret_cast<const uint4*>(p) : uint4{0u, 0u, 0u, 0u};
      } else {
        float4 a = float4{0.f, 0.f, 0.f, 0.f}, b = a;
        if (ok) {
          a = reinterpret_cast<const float4*>(p)[0];
          b = reinterpret_cast<const float4*>(p)[1];
        }
        v[r].x = uint32_t(f2bf(a.x)) | (uint32_t(f2bf(a.y)) << 16);
        v[r].y = uint32_t(f2bf(a.z)) | (uint32_t(f2bf(a.w)) << 16);
        v[r].z = uint32_t(f2bf(b.x)) | (uint32_t(f2bf(b.y)) << 16);
        v[r].w = uint32_t(f2bf(b.z)) | (uint32_t(f2bf(b.w)) << 16);
      }
    }
  }
  __device__ void store(unsigned short* lds, int t) const {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = r * 256 + t;
      *reinterpret_cast<uint4*>(lds + (e >> 3) * NLD + (e & 7) * 8) = v[r];
    }
  }
};

template <bool BF>
struct TileLoadG<BF, false> {  // contiguous along rows: lanes walk rows, 8 loads along K
  using TE = typename std::conditional<BF, unsigned short, float>::type;
  uint4 v[4];
  __device__ void load(const TE* base, int64_t /*ld*/, int64_t sk, int row0, int nrows, int k0,
                       int K, int t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = r * 256 + t;
      const int row = e & 127, kg = e >> 7;  // kg = 0..7
      const int k = k0 + kg * 8;
      const bool rok = row0 + row < nrows;
      const TE* p = base + (row0 + row) + int64_t(k) * sk;
      unsigned short h[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool ok = rok && k + j < K;
        if constexpr (BF) {
          h[j] = ok ? p[int64_t(j) * sk] : (unsigned short)0;
        } else {
          h[j] = ok ? f2bf(p[int64_t(j) * sk]) : (unsigned short)0;
        }
      }
      v[r].x = uint32_t(h[0]) | (uint32_t(h[1]) << 16);
      v[r].y = uint32_t(h[2]) | (uint32_t(h[3]) << 16);
      v[r].z = uint32_t(h[4]) | (uint32_t(h[5]) << 16);
      v[r].w = uint32_t(h[6]) | (uint32_t(h[7]) << 16);
    }
  }
  __device__ void store(unsigned short* lds, int t) const {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = r * 256 + t;
      *reinterpret_cast<uint4*>(lds + (e & 127) * NLD + (e >> 7) * 8) = v[r];
    }
  }
};

template <bool ABF, bool BBF, bool AK, bool BK>
__global__ __launch_bounds__(256) void k_gemm_tile(GemmArgs g) {
  __shared__ unsigned short As[2][GBM * NLD];
  __shared__ unsigned short Bs[2][GBN * NLD];
  using TA = typename std::conditional<ABF, unsigned short, float>::type;
  using TB = typename std::conditional<BBF, unsigned short, float>::type;
  const int b = blockIdx.z;
  const int m0 = blockIdx.y * GBM;
  const int n0 = blockIdx.x * GBN;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = (wave >> 1) * 64;
  const int wn = (wave & 1) * 64;
  const TA* A = static_cast<const TA*>(g.A) + int64_t(b) * g.sAb;
  const TB* B = static_cast<const TB*>(g.B) + int64_t(b) * g.sBb;
  // row stride / k stride of each operand in its [row][k] view
  const int64_t a_ld = g.sAm, a_sk = g.sAk;
  const int64_t b_ld = g.sBn, b_sk = g.sBk;

  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  TileLoadG<ABF, AK> la;
  TileLoadG<BBF, BK> lb;
  la.load(A, a_ld, a_sk, m0, g.M, 0, g.K, t);
  lb.load(B, b_ld, b_sk, n0, g.N, 0, g.K, t);
  la.store(As[0], t);
  lb.store(Bs[0], t);
  __syncthreads();
  const int nk = (g.K + NBK - 1) / NBK;
  const int kq = (lane >> 4) * 8;
  const int lr = lane & 15;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      la.load(A, a_ld, a_sk, m0, g.M, (kt + 1) * NBK, g.K, t);
      lb.load(B, b_ld, b_sk, n0, g.N, (kt + 1) * NBK, g.K, t);
    }
#pragma unroll
    for (int ks = 0; ks < NBK; ks += 32) {
      bf16x8 af[4], bfr[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        af[i] = *reinterpret_cast<const bf16x8*>(&As[cur][(wm + i * 16 + lr) * NLD + ks + kq]);
        bfr[i] = *reinterpret_cast<const bf16x8*>(&Bs[cur][(wn + i * 16 + lr) * NLD + ks + kq]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    if (more) {
      la.store(As[cur ^ 1], t);
      lb.store(Bs[cur ^ 1], t);
    }
    __syncthreads();
  }

  const int64_t coff = int64_t(b) * g.sCb;
  const int64_t ioff = int64_t(b) * g.sIb;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + wn + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        if (m >= g.M || n >= g.N) continue;
        float v = g.alpha * acc[i][j][r];
        if (g.beta != 0.f) v += g.beta * g.Cin[ioff + m * g.sIm + n * g.sIn];
        if (g.bias_mode == 1) v += g.bias[n];
        else if (g.bias_mode == 2) v += g.bias[m];
        const int64_t ci = coff + m * g.sCm + n * g.sCn;
        if (g.pre) g.pre[ci] = v;
        if (g.act == 1) v = gelu(v);
        g.C[ci] = v;
      }
    }
  }
}

// Dispatch of the tile kernel; returns false when an operand's layout fits
// neither specialisation (then the element-wise kernel runs).
template <bool ABF, bool BBF>
static bool launch_tile(const GemmArgs& g, dim3 grid, hipStream_t s) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  const bool ak = g.sAk == 1 && g.K % 8 == 0 && g.sAm % 8 == 0 && g.sAb % 8 == 0 && al(g.A);
  const bool am = g.sAm == 1;
  const bool bk = g.sBk == 1 && g.K % 8 == 0 && g.sBn % 8 == 0 && g.sBb % 8 == 0 && al(g.B);
  const bool bn = g.sBn == 1;
  if (!(ak || am) || !(bk || bn)) return false;
  if (ak && bk) hipLaunchKernelGGL((k_gemm_tile<ABF, BBF, true, true>), grid, dim3(256), 0, s, g);
  else if (ak) hipLaunchKernelGGL((k_gemm_tile<ABF, BBF, true, false>), grid, dim3(256), 0, s, g);
  else if (bk) hipLaunchKernelGGL((k_gemm_tile<ABF, BBF, false, true>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_gemm_tile<ABF, BBF, false, false>), grid, dim3(256), 0, s, g);
  return true;
}

// ---- Channel-mixing products (small M, K; huge N) ------------------------
//
// C[M][n] = act(alpha A[M][K] X[K][n] + beta Cin + bias) with M, K <= 128 and
// n = batch x pixels in the millions: the FNO's lifting, pointwise and
// projection layers and their data gradients.  They move O(n (M + K)) bytes
// for O(n M K) flops with M K <= 4096.  This kernel runs them on the vector
// ALU in FP32 (operands rounded to BF16 first, as on the matrix cores): each
// thread owns 4 consecutive columns (16-byte loads and stores), A sits in
// LDS and is read by broadcast.  The shapes with K <= 32 run on the matrix
// cores instead (k_pw, below); this kernel takes the K = 128 ones and the
// cross-check.

__device__ inline float bfr(float v) { return __uint_as_float(uint32_t(f2bf(v)) << 16); }

template <int M, int K, bool BBF>
__global__ __launch_bounds__(256) void k_cmix(GemmArgs g) {
  __shared__ float Ws[M * K];
  {
    const bool abf = g.a_bf16 != 0;
    for (int e = threadIdx.x; e < M * K; e += 256) {
      const int m = e / K, k = e % K;
      const int64_t i = int64_t(m) * g.sAm + int64_t(k) * g.sAk;
      Ws[e] = abf ? __uint_as_float(uint32_t(static_cast<const unsigned short*>(g.A)[i]) << 16)
                  : bfr(static_cast<const float*>(g.A)[i]);
    }
  }
  __syncthreads();
  const int64_t n = (int64_t(blockIdx.x) * 256 + threadIdx.x) * 4;
  if (n >= g.N) return;
  auto loadx = [&](int k) -> float4 {
    if constexpr (BBF) {
      const uint2 u = *reinterpret_cast<const uint2*>(static_cast<const unsigned short*>(g.B) +
                                                        int64_t(k) * g.sBk + n);
      return float4{__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                    __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u)};
    } else {
      const float4 v =
          *reinterpret_cast<const float4*>(static_cast<const float*>(g.B) + int64_t(k) * g.sBk + n);
      return float4{bfr(v.x), bfr(v.y), bfr(v.z), bfr(v.w)};
    }
  };
  auto emit = [&](int m, float4 a) {
    float v[4] = {g.alpha * a.x, g.alpha * a.y, g.alpha * a.z, g.alpha * a.w};
    const int64_t ci = int64_t(m) * g.sCm + n;
    if (g.beta != 0.f) {
      const float4 c = *reinterpret_cast<const float4*>(g.Cin + int64_t(m) * g.sIm + n);
      v[0] += g.beta * c.x;
      v[1] += g.beta * c.y;
      v[2] += g.beta * c.z;
      v[3] += g.beta * c.w;
    }
    if (g.bias_mode == 2) {
      const float b = g.bias[m];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] += b;
    } else if (g.bias_mode == 1) {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] += g.bias[n + j];
    }
    if (g.pre) *reinterpret_cast<float4*>(g.pre + ci) = float4{v[0], v[1], v[2], v[3]};
    if (g.act == 1) {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = gelu(v[j]);
    }
    *reinterpret_cast<float4*>(g.C + ci) = float4{v[0], v[1], v[2], v[3]};
    if (g.c_bf16)
      *reinterpret_cast<uint2*>(g.c_bf16 + ci) =
          uint2{uint32_t(f2bf(v[0])) | (uint32_t(f2bf(v[1])) << 16),
                uint32_t(f2bf(v[2])) | (uint32_t(f2bf(v[3])) << 16)};
  };
  // Small K: all K input rows are loaded first (K loads in flight per
  // thread), then each output row is summed and stored.  Large K with few
  // outputs: the outputs stay in registers while K streams.  The sum over k
  // runs in the same order either way.
  if constexpr (K > 32 && 4 * M <= 128) {
    float4 acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int k = 0; k < K; ++k) {
      const float4 x = loadx(k);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const float w = Ws[m * K + k];
        acc[m].x += w * x.x;
        acc[m].y += w * x.y;
        acc[m].z += w * x.z;
        acc[m].w += w * x.w;
      }
    }
#pragma unroll
    for (int m = 0; m < M; ++m) emit(m, acc[m]);
  } else {
    float4 x[K];
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = loadx(k);
#pragma unroll 2
    for (int m = 0; m < M; ++m) {
      float4 a = float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float w = Ws[m * K + k];
        a.x += w * x[k].x;
        a.y += w * x[k].y;
        a.z += w * x[k].z;
        a.w += w * x[k].w;
      }
      emit(m, a);
    }
  }
}

template <int M, int K>
static bool cmix_mk(const GemmArgs& g, hipStream_t s) {
  if (g.M != M || g.K != K) return false;
  const dim3 grid(cdiv(g.N, 1024));
  if (g.b_bf16) hipLaunchKernelGGL((k_cmix<M, K, true>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_cmix<M, K, false>), grid, dim3(256), 0, s, g);
  return true;
}

// The channel-mixing shapes of the FNO (width 32, hidden 128, 9 inputs,
// 6 outputs); other shapes take the tile kernel.
static bool launch_cmix(const GemmArgs& g, hipStream_t s) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  if (g.batch != 1 || g.sBn != 1 || g.sCn != 1 || g.N % 4 != 0) return false;
  if (g.sBk % 4 != 0 || g.sCm % 4 != 0 || !al(g.B) || !al(g.C)) return false;
  if (g.beta != 0.f && (g.sIn != 1 || g.sIm % 4 != 0 || !al(g.Cin))) return false;
  if (g.pre && !al(g.pre)) return false;
  if (g.c_bf16 && (reinterpret_cast<uintptr_t>(g.c_bf16) & 7u) != 0) return false;
  if (g.b_bf16 && g.sBk % 8 != 0) return false;
  return cmix_mk<32, 9>(g, s) || cmix_mk<32, 32>(g, s) || cmix_mk<128, 32>(g, s) ||
         cmix_mk<6, 128>(g, s) || cmix_mk<128, 6>(g, s) || cmix_mk<32, 128>(g, s);
}


static bool nt_ok(const GemmArgs& g) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  return g.sAk == 1 && g.sBk == 1 && g.K % NBK == 0 && g.sAm % 8 == 0 && g.sBn % 8 == 0 &&
         g.sAb % 8 == 0 && g.sBb % 8 == 0 && al(g.A) && al(g.B);
}

// ---- Large NT kernel: 256 x 256 x 64 tiles on 32x32x16 MFMA -------------
//
// For the big spectral products with long reductions (M, N >= 256,
// K >= 2048, both operands contiguous along K, K % 64 == 0).  256 threads = 4 waves in 2 x 2, each owning a
// 128 x 128 block of C as 4 x 4 tiles of v_mfma_f32_32x32x16_bf16 (256 FP32
// accumulators per lane), so each 16-byte LDS read feeds four MFMAs of
// 32 x 32 x 16: half the LDS bytes per flop of the 64 x 64 x 16x16 layout.
// LDS: A and B tiles [row][k] (row padded to 72 elements, so 8 lanes
// reading 8 consecutive rows hit 8 disjoint bank groups), double-buffered
// (144 KiB, one work-group per CU); the next K step is prefetched into
// registers (FP32 inputs converted to BF16 with v_cvt_pk_bf16_f32, round
// to nearest even, on the LDS store) with one barrier per K step.  Blocks
// are remapped so that consecutive logical tiles, which share the A panel
// (all N tiles of one M panel), run on one XCD and share its L2.  MFMA
// 32x32x16 lane layout: A lane l holds row l%32, k = (l/32)*8 .. +8; B lane
// l holds column l%32, the same k; accumulator element v of lane l is row
// 8*(v/4) + 4*(l/32) + v%4, column l%32.

constexpr int XBM = 256, XBN = 256, XBK = 64, XLD = XBK + 8;
// Used for long reductions only: with one work-group per CU and a one-step
// register prefetch, short K (few steps to hide the load latency and a
// large epilogue share) runs faster on the 128 x 128 kernel (measured at
// 8192 x 16384 x 1024: 704 vs 485 us; at 8192 x 1024 x 16384: 466 vs 604).
constexpr int XBIG_K = 2048;
constexpr size_t XSMEM = size_t(2) * (XBM + XBN) * XLD * sizeof(unsigned short);

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <bool BF>
struct BigLoad;

template <>
struct BigLoad<true> {  // 256 rows x 64 k BF16 = 2048 x 16 B, 8 per thread
  u32x4 v[8];
  __device__ __forceinline__ void load(const unsigned short* base, int64_t ld, int row0,
                                       int nrows, int k0, int t) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 3, c8 = e & 7;
      const int gr = min(row0 + row, nrows - 1);  // clamped: rows past the edge are unused
      v[r] = *reinterpret_cast<const u32x4*>(base + int64_t(gr) * ld + k0 + c8 * 8);
    }
  }
  __device__ __forceinline__ void store(unsigned short* lds, int t) const {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 3, c8 = e & 7;
      *reinterpret_cast<u32x4*>(lds + row * XLD + c8 * 8) = v[r];
    }
  }
};

template <>
struct BigLoad<false> {  // 256 rows x 64 k FP32 = 4096 x 16 B, 16 per thread
  f32x4 v[16];
  __device__ __forceinline__ void load(const float* base, int64_t ld, int row0, int nrows,
                                       int k0, int t) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 4, c4 = e & 15;
      const int gr = min(row0 + row, nrows - 1);
      v[r] = *reinterpret_cast<const f32x4*>(base + int64_t(gr) * ld + k0 + c4 * 4);
    }
  }
  __device__ __forceinline__ void store(unsigned short* lds, int t) const {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int e = r * 256 + t;
      const int row = e >> 4, c4 = e & 15;
      bf16x2 lo, hi;
      lo[0] = (__bf16)v[r][0];
      lo[1] = (__bf16)v[r][1];
      hi[0] = (__bf16)v[r][2];
      hi[1] = (__bf16)v[r][3];
      u32x2 p;
      p[0] = __builtin_bit_cast(uint32_t, lo);
      p[1] = __builtin_bit_cast(uint32_t, hi);
      *reinterpret_cast<u32x2*>(lds + row * XLD + c4 * 4) = p;
    }
  }
};

// Epilogue of four consecutive columns n .. n+3 of row m: vector loads and
// stores when C (and Cin) are contiguous along n and the run is inside the
// matrix, element by element otherwise.
__device__ __forceinline__ void big_epilogue(const GemmArgs& g, int m, int n, int64_t coff,
                                          int64_t ioff, f32x4 a) {
  if (m >= g.M || n >= g.N) return;
  const int64_t ci = coff + m * g.sCm + n * g.sCn;
  const bool vec = n + 3 < g.N && g.sCn == 1 && (ci & 3) == 0 &&
                   (g.beta == 0.f || (g.sIn == 1 && ((ioff + m * g.sIm + n) & 3) == 0));
  if (vec) {
    f32x4 x = a * g.alpha;
    if (g.beta != 0.f) x += g.beta * *reinterpret_cast<const f32x4*>(g.Cin + ioff + m * g.sIm + n);
    if (g.bias_mode == 1) x += f32x4{g.bias[n], g.bias[n + 1], g.bias[n + 2], g.bias[n + 3]};
    else if (g.bias_mode == 2) x += g.bias[m];
    if (g.pre) *reinterpret_cast<f32x4*>(g.pre + ci) = x;
    if (g.act == 1) x = f32x4{gelu(x[0]), gelu(x[1]), gelu(x[2]), gelu(x[3])};
    *reinterpret_cast<f32x4*>(g.C + ci) = x;
    return;
  }
  for (int e = 0; e < 4 && n + e < g.N; ++e) {
    float x = g.alpha * a[e];
    if (g.beta != 0.f) x += g.beta * g.Cin[ioff + m * g.sIm + (n + e) * g.sIn];
    if (g.bias_mode == 1) x += g.bias[n + e];
    else if (g.bias_mode == 2) x += g.bias[m];
    const int64_t c = coff + m * g.sCm + (n + e) * g.sCn;
    if (g.pre) g.pre[c] = x;
    if (g.act == 1) x = gelu(x);
    g.C[c] = x;
  }
}

template <bool ABF, bool BBF>
__global__ __launch_bounds__(256, 1) void k_gemm_big(GemmArgs g, int tiles_n, int tiles) {
  extern __shared__ unsigned short xsm[];
  // XCD-aware remap (the grid is a multiple of 8; block b runs on XCD b % 8)
  const int nb = gridDim.x, bx = blockIdx.x;
  const int logical = (bx & 7) * (nb >> 3) + (bx >> 3);
  if (logical >= tiles) return;
  const int m0 = (logical / tiles_n) * XBM;
  const int n0 = (logical % tiles_n) * XBN;
  const int b = blockIdx.z;
  unsigned short* As = xsm;
  unsigned short* Bs = xsm + 2 * XBM * XLD;
  using TA = typename std::conditional<ABF, unsigned short, float>::type;
  using TB = typename std::conditional<BBF, unsigned short, float>::type;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = (wave >> 1) * 128;
  const int wn = (wave & 1) * 128;
  const TA* A = static_cast<const TA*>(g.A) + int64_t(b) * g.sAb;
  const TB* B = static_cast<const TB*>(g.B) + int64_t(b) * g.sBb;

  f32x16 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

  BigLoad<ABF> la;
  BigLoad<BBF> lb;
  la.load(A, g.sAm, m0, g.M, 0, t);
  lb.load(B, g.sBn, n0, g.N, 0, t);
  la.store(As, t);
  lb.store(Bs, t);
  __syncthreads();
  const int nk = g.K / XBK;
  const int kq = (lane >> 5) * 8;
  const int lr = lane & 31;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      la.load(A, g.sAm, m0, g.M, (kt + 1) * XBK, t);
      lb.load(B, g.sBn, n0, g.N, (kt + 1) * XBK, t);
    }
    const unsigned short* Ac = As + cur * XBM * XLD;
    const unsigned short* Bc = Bs + cur * XBN * XLD;
#pragma unroll
    for (int ks = 0; ks < XBK; ks += 16) {
      bf16x8 af[4], bfr[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        af[i] = *reinterpret_cast<const bf16x8*>(&Ac[(wm + i * 32 + lr) * XLD + ks + kq]);
        bfr[i] = *reinterpret_cast<const bf16x8*>(&Bc[(wn + i * 32 + lr) * XLD + ks + kq]);
      }
      // operands swapped: each accumulator holds a 32 x 32 tile of C^T, so a
      // lane's values are 4-column runs of one row of C (vector stores)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bfr[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (more) {
      la.store(As + (cur ^ 1) * XBM * XLD, t);
      lb.store(Bs + (cur ^ 1) * XBN * XLD, t);
    }
    __syncthreads();
  }

  const int64_t coff = int64_t(b) * g.sCb;
  const int64_t ioff = int64_t(b) * g.sIb;
  // (written out per accumulator so that every index is a literal and the
  // accumulators stay in registers)
  auto tile = [&](const f32x16& a, int i, int j) {
    const int m = m0 + wm + i * 32 + lr;
    const int n = n0 + wn + j * 32 + 4 * (lane >> 5);
    big_epilogue(g, m, n, coff, ioff, f32x4{a[0], a[1], a[2], a[3]});
    big_epilogue(g, m, n + 8, coff, ioff, f32x4{a[4], a[5], a[6], a[7]});
    big_epilogue(g, m, n + 16, coff, ioff, f32x4{a[8], a[9], a[10], a[11]});
    big_epilogue(g, m, n + 24, coff, ioff, f32x4{a[12], a[13], a[14], a[15]});
  };
  tile(acc[0][0], 0, 0); tile(acc[0][1], 0, 1); tile(acc[0][2], 0, 2); tile(acc[0][3], 0, 3);
  tile(acc[1][0], 1, 0); tile(acc[1][1], 1, 1); tile(acc[1][2], 1, 2); tile(acc[1][3], 1, 3);
  tile(acc[2][0], 2, 0); tile(acc[2][1], 2, 1); tile(acc[2][2], 2, 2); tile(acc[2][3], 2, 3);
  tile(acc[3][0], 3, 0); tile(acc[3][1], 3, 1); tile(acc[3][2], 3, 2); tile(acc[3][3], 3, 3);
}

template <bool ABF, bool BBF>
static int launch_big(const GemmArgs& g, hipStream_t s) {
  static std::atomic<bool> attr{false};
  if (!attr.load()) {
    DNX_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gemm_big<ABF, BBF>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(XSMEM)));
    attr = true;
  }
  const int tn = int(cdiv(g.N, XBN)), tiles = int(cdiv(g.M, XBM)) * tn;
  const int nb = ((tiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k_gemm_big<ABF, BBF>), dim3(nb, 1, g.batch), dim3(256), XSMEM, s, g, tn,
                     tiles);
  return DNX_OK;
}

// ---- Long-K BF16 kernel with operands loaded straight into LDS ------------
//
// The products of k_gemm_big for BF16 A and B (the truncated DFTs with
// K = 2048-16384): the same 256 x 256 x 64 tile, XCD remap, operand swap
// and MFMA sequence per accumulator (so bitwise the same results), but
//  * 8 waves (4 along M x 2 along N, 64 x 128 of C each, 128 accumulators
//    per lane): two waves per SIMD, so one wave's LDS reads and barrier
//    waits overlap the other's MFMAs;
//  * global -> LDS without registers: global_load_lds_dwordx4 writes 16
//    bytes per lane to LDS at a wave-uniform base + 16 * lane, a 1 KB run per
//    instruction (8 rows of 64 k); two stages of 64 KB, the next K step in
//    flight while the current one is multiplied, one barrier per K step;
//  * unpadded rows (128 bytes) with the 16-byte chunk c of row r stored at
//    chunk c ^ ((r ^ (r >> 3)) & 7): the lanes of an MFMA operand read
//    (32 rows, one chunk) fall on distinct banks, and the DMA writes stay
//    contiguous (each lane loads the global chunk its LDS slot holds).
// Measured on 8192 x 1024 x 4096 (x4 split) against k_gemm_big: 311-330 vs
// 472-499 us with FP32 A, 397-404 us with BF16 A.

constexpr int LBK = 64;                      // K per stage
constexpr int LDS_MIN_K = 1024;              // shortest reduction it takes
constexpr int LSTAGE = 256 * LBK;            // BF16 elements of one operand's stage
constexpr size_t LSMEM = size_t(2) * 2 * LSTAGE * sizeof(unsigned short);  // 128 KiB

// 16 bytes per lane from global memory into LDS at lds + 16 * lane (lds
// wave-uniform).  A plain function, so the host pass never instantiates the
// device builtin inside a template.
__device__ __forceinline__ void lds_dma16(const void* gp, void* lds) {
  __builtin_amdgcn_global_load_lds(gp, lds, 16, 0, 0);
}

__device__ __forceinline__ int lds_swz(int r) { return (r ^ (r >> 3)) & 7; }

__global__ __launch_bounds__(512, 1) void k_gemm_lds(GemmArgs g, int tiles_n, int tiles) {
  constexpr int IPO = 4;  // load instructions per wave per operand per stage (32 / 8 waves)
  extern __shared__ unsigned short lsm[];
  const int nb = gridDim.x, bx = blockIdx.x;
  const int logical = (bx & 7) * (nb >> 3) + (bx >> 3);
  if (logical >= tiles) return;
  const int m0 = (logical / tiles_n) * 256;
  const int n0 = (logical % tiles_n) * 256;
  const int b = blockIdx.z;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 128;
  const unsigned short* A = static_cast<const unsigned short*>(g.A) + int64_t(b) * g.sAb;
  const unsigned short* B = static_cast<const unsigned short*>(g.B) + int64_t(b) * g.sBb;
  // Instruction j of this wave fills rows 8q .. 8q+7 (q = wave + 8 j) of a
  // stage; lane i writes slot (row 8q + i/8, chunk i%8), i.e. the global
  // chunk (i%8) ^ swz(row) of that row.  Rows past the matrix are clamped
  // (their products are never stored).
  int64_t ga[IPO], gb[IPO];
#pragma unroll
  for (int j = 0; j < IPO; ++j) {
    const int row = (wave + 8 * j) * 8 + (lane >> 3);
    const int c = (lane & 7) ^ lds_swz(row);
    ga[j] = int64_t(min(m0 + row, g.M - 1)) * g.sAm + c * 8;
    gb[j] = int64_t(min(n0 + row, g.N - 1)) * g.sBn + c * 8;
  }
  const int nk = g.K / LBK;
  auto issue = [&](int kt) {
    unsigned short* sa = lsm + (kt & 1) * 2 * LSTAGE;
    unsigned short* sb = sa + LSTAGE;
#pragma unroll
    for (int j = 0; j < IPO; ++j) {
      const int q = wave + 8 * j;
      lds_dma16(A + ga[j] + int64_t(kt) * LBK, sa + q * 512);
      lds_dma16(B + gb[j] + int64_t(kt) * LBK, sb + q * 512);
    }
  };
  f32x16 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
  const int lr = lane & 31, kh = lane >> 5;
  issue(0);
  for (int kt = 0; kt < nk; ++kt) {
    // this wave's loads of step kt have landed (vmcnt counts the DMA loads;
    // expcnt and lgkmcnt are left at their maxima), then every wave's have,
    // and every wave is done with the stage that step kt + 1 overwrites
    __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8));
    __builtin_amdgcn_s_barrier();
    if (kt + 1 < nk) issue(kt + 1);
    const unsigned short* Ac = lsm + (kt & 1) * 2 * LSTAGE;
    const unsigned short* Bc = Ac + LSTAGE;
#pragma unroll
    for (int ks = 0; ks < LBK; ks += 16) {
      const int c = ks / 8 + kh;
      bf16x8 af[2], bfr[4];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = wm + i * 32 + lr;
        af[i] = *reinterpret_cast<const bf16x8*>(&Ac[r * LBK + ((c ^ lds_swz(r)) << 3)]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = wn + j * 32 + lr;
        bfr[j] = *reinterpret_cast<const bf16x8*>(&Bc[r * LBK + ((c ^ lds_swz(r)) << 3)]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bfr[j], af[i], acc[i][j], 0, 0, 0);
    }
  }
  const int64_t coff = int64_t(b) * g.sCb;
  const int64_t ioff = int64_t(b) * g.sIb;
  auto tile = [&](const f32x16& a, int i, int j) {
    const int m = m0 + wm + i * 32 + lr;
    const int n = n0 + wn + j * 32 + 4 * kh;
    big_epilogue(g, m, n, coff, ioff, f32x4{a[0], a[1], a[2], a[3]});
    big_epilogue(g, m, n + 8, coff, ioff, f32x4{a[4], a[5], a[6], a[7]});
    big_epilogue(g, m, n + 16, coff, ioff, f32x4{a[8], a[9], a[10], a[11]});
    big_epilogue(g, m, n + 24, coff, ioff, f32x4{a[12], a[13], a[14], a[15]});
  };
  tile(acc[0][0], 0, 0); tile(acc[0][1], 0, 1); tile(acc[0][2], 0, 2); tile(acc[0][3], 0, 3);
  tile(acc[1][0], 1, 0); tile(acc[1][1], 1, 1); tile(acc[1][2], 1, 2); tile(acc[1][3], 1, 3);
}

// An FNO layer's inverse DFT and its pointwise part in one kernel:
//   out[c][b P + p] = act(sum_k Y[c B + b][k] GT[p][k]
//                         + sum_c' W[c][c'] h16[c'][b P + p] + bias[c])
// with 32 channels c, B samples, P cells and K = 2 x modes; Y, GT, h16 BF16,
// W rounded to BF16, FP32 accumulation.  The product runs as k_gemm_lds
// does (256 x 256 x 64 tiles, 8 waves, operands straight into LDS) over
// virtual rows v = 32 b + c, so that a tile's 256 rows are the 32 channels
// of 8 whole samples; a wave's 32 x 32 accumulator sub-tile is then one
// sample's 32 channels at 32 cells, which is also the shape of the
// pointwise product (W times the 32 channels of h16 at those cells): two
// more MFMAs add it into the same accumulators before the bias, the
// activation and the stores.  Against the inverse-DFT product followed by
// the pointwise kernel, the FP32 sum of the spectral branch is neither
// written nor read back, and the two sums are added inside the
// accumulators (FP32 rounding differs from the two-kernel path).
struct InvPwArgs {
  const unsigned short* Y;   // [32 B][K]
  const unsigned short* GT;  // [P][K]
  const unsigned short* h16; // [32][B P], the layer's input
  const float* W;            // [32][32]
  const float* bias;         // [32]
  float* out;                // [32][B P] or null
  float* pre;                // before the activation, or null
  unsigned short* out16;     // out rounded to BF16, or null
  const float* gpre;         // backward: out = (sum) * GELU'(gpre), or null
  int act, transW, B, P, K;  // transW: W read as [c'][c] (W^T)
};

__global__ __launch_bounds__(512, 1) void k_inv_pw(InvPwArgs a, int tiles_n, int tiles) {
  constexpr int C = 32, IPO = 4;
  extern __shared__ unsigned short lsm[];
  const int nb = gridDim.x, bx = blockIdx.x;
  const int logical = (bx & 7) * (nb >> 3) + (bx >> 3);
  if (logical >= tiles) return;
  const int m0 = (logical / tiles_n) * 256;  // virtual rows (32 b + c)
  const int n0 = (logical % tiles_n) * 256;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 128;
  // as k_gemm_lds: instruction j of this wave fills rows 8q .. 8q+7
  // (q = wave + 8 j) of a stage, lane i the chunk (i%8) ^ swz(row); the
  // virtual row 32 b + c reads Y's row c B + b (every tile is inside the
  // matrix: 32 B and P are multiples of 256)
  int64_t ga[IPO], gb[IPO];
#pragma unroll
  for (int j = 0; j < IPO; ++j) {
    const int row = (wave + 8 * j) * 8 + (lane >> 3);
    const int c = (lane & 7) ^ lds_swz(row);
    const int v = m0 + row;
    ga[j] = (int64_t(v % C) * a.B + v / C) * a.K + c * 8;
    gb[j] = int64_t(n0 + row) * a.K + c * 8;
  }
  const int nk = a.K / LBK;
  auto issue = [&](int kt) {
    unsigned short* sa = lsm + (kt & 1) * 2 * LSTAGE;
    unsigned short* sb = sa + LSTAGE;
#pragma unroll
    for (int j = 0; j < IPO; ++j) {
      const int q = wave + 8 * j;
      lds_dma16(a.Y + ga[j] + int64_t(kt) * LBK, sa + q * 512);
      lds_dma16(a.GT + gb[j] + int64_t(kt) * LBK, sb + q * 512);
    }
  };
  f32x16 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
  const int lr = lane & 31, kh = lane >> 5;
  issue(0);
  for (int kt = 0; kt < nk; ++kt) {
    __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8));
    __builtin_amdgcn_s_barrier();
    if (kt + 1 < nk) issue(kt + 1);
    const unsigned short* Ac = lsm + (kt & 1) * 2 * LSTAGE;
    const unsigned short* Bc = Ac + LSTAGE;
#pragma unroll
    for (int ks = 0; ks < LBK; ks += 16) {
      const int c = ks / 8 + kh;
      bf16x8 af[2], bfr[4];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = wm + i * 32 + lr;
        af[i] = *reinterpret_cast<const bf16x8*>(&Ac[r * LBK + ((c ^ lds_swz(r)) << 3)]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = wn + j * 32 + lr;
        bfr[j] = *reinterpret_cast<const bf16x8*>(&Bc[r * LBK + ((c ^ lds_swz(r)) << 3)]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bfr[j], af[i], acc[i][j], 0, 0, 0);
    }
  }
  // the pointwise product into the same accumulators: lane holds channel
  // lr of W (k = c' in chunk kh) and cell lr of the sub-tile (h16 rows c')
  bf16x8 wf[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    u16x8 u;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = ks * 16 + kh * 8 + e;
      u[e] = f2bf(a.transW ? a.W[k * C + lr] : a.W[lr * C + k]);
    }
    wf[ks] = __builtin_bit_cast(bf16x8, u);
  }
  const int64_t BP = int64_t(a.B) * a.P;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int bb = (m0 + wm) / C + i;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = int64_t(bb) * a.P + n0 + wn + j * 32 + lr;
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        u16x8 u;
#pragma unroll
        for (int e = 0; e < 8; ++e) u[e] = a.h16[int64_t(ks * 16 + kh * 8 + e) * BP + col];
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, u),
                                                            wf[ks], acc[i][j], 0, 0, 0);
      }
    }
  }
  // accumulator element v of a sub-tile: channel lr, cell
  // 8 (v / 4) + 4 kh + v % 4 -- four consecutive cells per v / 4
  const float bc = a.bias ? a.bias[lr] : 0.f;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int bb = (m0 + wm) / C + i;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int qv = 0; qv < 4; ++qv) {
        const int64_t idx =
            int64_t(lr) * BP + int64_t(bb) * a.P + n0 + wn + j * 32 + 8 * qv + 4 * kh;
        f32x4 x = f32x4{acc[i][j][4 * qv], acc[i][j][4 * qv + 1], acc[i][j][4 * qv + 2],
                        acc[i][j][4 * qv + 3]} + bc;
        if (a.pre) *reinterpret_cast<f32x4*>(a.pre + idx) = x;
        if (a.act == 1) x = f32x4{gelu(x[0]), gelu(x[1]), gelu(x[2]), gelu(x[3])};
        if (a.gpre) {  // as k_gelu_bwd: dy * GELU'(pre)
          const f32x4 g = *reinterpret_cast<const f32x4*>(a.gpre + idx);
          x = f32x4{x[0] * gelu_grad(g[0]), x[1] * gelu_grad(g[1]), x[2] * gelu_grad(g[2]),
                    x[3] * gelu_grad(g[3])};
        }
        if (a.out) *reinterpret_cast<f32x4*>(a.out + idx) = x;
        if (a.out16)
          *reinterpret_cast<uint2*>(a.out16 + idx) =
              uint2{uint32_t(f2bf(x[0])) | (uint32_t(f2bf(x[1])) << 16),
                    uint32_t(f2bf(x[2])) | (uint32_t(f2bf(x[3])) << 16)};
      }
    }
  }
}

static int launch_lds(const GemmArgs& g, hipStream_t s) {
  static std::atomic<bool> attr{false};
  if (!attr.load()) {
    DNX_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_gemm_lds),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(LSMEM)));
    attr = true;
  }
  const int tn = int(cdiv(g.N, 256)), tiles = int(cdiv(g.M, 256)) * tn;
  const int nb = ((tiles + 7) / 8) * 8;
  hipLaunchKernelGGL(k_gemm_lds, dim3(nb, 1, g.batch), dim3(512), LSMEM, s, g, tn, tiles);
  return DNX_OK;
}

// out[i] = sum_j part[j*n + i] + beta * out[i]
__global__ void k_sum_partials(const float* __restrict__ part, int s, int64_t n,
                               float* __restrict__ out, float beta) {
  const int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
  if (i >= n) return;
  if (i + 4 <= n && (n & 3) == 0) {
    float4 a = beta != 0.f ? *reinterpret_cast<const float4*>(out + i) : float4{0.f, 0.f, 0.f, 0.f};
    if (beta != 0.f) a = float4{a.x * beta, a.y * beta, a.z * beta, a.w * beta};
    for (int j = 0; j < s; ++j) {
      const float4 p = *reinterpret_cast<const float4*>(part + int64_t(j) * n + i);
      a.x += p.x;
      a.y += p.y;
      a.z += p.z;
      a.w += p.w;
    }
    *reinterpret_cast<float4*>(out + i) = a;
    return;
  }
  for (int64_t q = i; q < i + 4 && q < n; ++q) {
    float a = beta != 0.f ? beta * out[q] : 0.f;
    for (int j = 0; j < s; ++j) a += part[int64_t(j) * n + q];
    out[q] = a;
  }
}


// four elements per thread (16-byte accesses when all three arrays are
// 16-byte aligned and n % 4 == 0; element by element otherwise)
__global__ void k_gelu_bwd(const float* dy, const float* __restrict__ pre, float* dx,
                           unsigned short* __restrict__ dx16, int64_t n,
                           int vec) {  // dx may be dy (in place)
  const int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
  if (i >= n) return;
  if (vec) {
    const float4 x = *reinterpret_cast<const float4*>(pre + i);
    const float4 d = *reinterpret_cast<const float4*>(dy + i);
    const float4 r = float4{d.x * gelu_grad(x.x), d.y * gelu_grad(x.y), d.z * gelu_grad(x.z),
                            d.w * gelu_grad(x.w)};
    *reinterpret_cast<float4*>(dx + i) = r;
    if (dx16)
      *reinterpret_cast<uint2*>(dx16 + i) = uint2{uint32_t(f2bf(r.x)) | (uint32_t(f2bf(r.y)) << 16),
                                                  uint32_t(f2bf(r.z)) | (uint32_t(f2bf(r.w)) << 16)};
    return;
  }
  for (int64_t q = i; q < i + 4 && q < n; ++q) {
    const float r = dy[q] * gelu_grad(pre[q]);
    dx[q] = r;
    if (dx16) dx16[q] = f2bf(r);
  }
}

// ---- Channel-mixing products on the matrix cores ------------------------
//
// The same products as k_cmix (C[m][n] = act(alpha sum_k W[m][k] X[k][n] +
// beta Cin[m][n] + bias), n = batch x pixels), with W and X rounded to BF16
// as k_cmix does and multiplied on v_mfma_f32_32x32x16_bf16 (FP32
// accumulation; the order of the k sum differs from k_cmix's, so results
// agree to FP32 rounding, not bit for bit).  On the vector ALU the products
// cost 2 M K flops per cell (M K = 4096: 0.44 ms of FMA per 2^22 cells);
// here the matrix cores take them and the kernel streams X, Cin and C
// (measured per call at 2^22 / 2^21 cells, against k_cmix: 32 x 32 261 vs
// 292 us average, 128 x 32 770 vs 834, 32 x 9 167 vs 209).
//
// MFMA operands: A = W, 32 rows per tile (MT tiles, rows >= M zero), kept in
// registers for the whole launch; B = X, 32 cells, K padded to 16 KS with
// zeros.  Lane l holds W row l%32 and cell l%32 with k = (l/32)*8 .. +8; the
// accumulator element v of lane l is output row 8(v/4) + 4(l/32) + v%4 at
// cell l%32, so every load and store of X, Cin, pre and C covers 32
// consecutive cells of a row (128-byte segments of FP32).  Each wave walks
// groups of 32 cells.

template <int M, int K, bool BBF>
__global__ __launch_bounds__(256) void k_pw(GemmArgs g) {
  constexpr int MT = (M + 31) / 32, KS = (K + 15) / 16;
  __shared__ float bs[128];
  const int lane = threadIdx.x & 63;
  const int lr = lane & 31, half = lane >> 5, kq = half * 8;
  const int N = g.N;
  if (threadIdx.x < 128)
    bs[threadIdx.x] = (g.bias_mode == 2 && int(threadIdx.x) < M) ? g.bias[threadIdx.x] : 0.f;
  bf16x8 wa[MT][KS];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int m = mt * 32 + lr;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      u16x8 u;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = ks * 16 + kq + j;
        u[j] = (m < M && k < K) ? load_bf(g.A, int64_t(m) * g.sAm + int64_t(k) * g.sAk, g.a_bf16)
                                : (unsigned short)0;
      }
      wa[mt][ks] = __builtin_bit_cast(bf16x8, u);
    }
  }
  __syncthreads();
  // Buffer accesses: a descriptor per block of rows (16 of X, 32 of C, pre
  // and Cin) with its base in scalar registers, the row within the block as
  // scalar offset and the cell as the lane's vector offset; byte offsets
  // within a block stay below 2^31 (checked by the launcher).
  constexpr int xb_bytes = BBF ? 2 : 4;
  const int sBk = int(g.sBk) * xb_bytes, sCm = int(g.sCm) * 4, sIm = int(g.sIm) * 4;
  auto rsrc = [](const void* base, int64_t off) {
    return __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char*>(static_cast<const char*>(base)) + off, 0, 0x7fffffff, 0x00020000);
  };
  const int groups = (N + 31) / 32;
  const int stride = int(gridDim.x) * 4;
  for (int gi = int(blockIdx.x) * 4 + int(threadIdx.x >> 6); gi < groups; gi += stride) {
    const int n = gi * 32 + lr;
    const bool nv = n < N;
    const int nc = nv ? n : N - 1;
    bf16x8 xb[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      u16x8 u;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = ks * 16 + kq + j;
        const auto rx = rsrc(g.B, int64_t(ks * 16) * sBk);
        unsigned short v = 0;
        if (k < K) {
          // lane-dependent part (cell, k block of the lane half) in the
          // vector offset, the row within it in the scalar offset
          if constexpr (BBF)
            v = __builtin_amdgcn_raw_buffer_load_b16(rx, nc * 2 + kq * sBk, j * sBk, 0);
          else
            v = f2bf(__uint_as_float(
                __builtin_amdgcn_raw_buffer_load_b32(rx, nc * 4 + kq * sBk, j * sBk, 0)));
        }
        u[j] = v;
      }
      xb[ks] = __builtin_bit_cast(bf16x8, u);
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int mb = mt * 32 + 4 * half;  // row of element v: mb + 8(v/4) + v%4
      float cin[16];
      if (g.beta != 0.f) {
        const auto ri = rsrc(g.Cin, int64_t(mt * 32) * sIm);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int m = mb + 8 * (v >> 2) + (v & 3);
          cin[v] = m < M ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                               ri, nc * 4 + 4 * half * sIm, (8 * (v >> 2) + (v & 3)) * sIm, 0))
                         : 0.f;
        }
      }
      f32x16 acc;
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa[mt][ks], xb[ks], acc, 0, 0, 0);
      if (!nv) continue;
      const int vo = n * 4 + 4 * half * sCm;
      const auto rc = rsrc(g.C, int64_t(mt * 32) * sCm);
      const auto rp = rsrc(g.pre ? g.pre : g.C, int64_t(mt * 32) * sCm);
      // BF16 copy: the same element offsets at half the bytes
      const auto r16 = rsrc(g.c_bf16 ? static_cast<void*>(g.c_bf16) : static_cast<void*>(g.C),
                            int64_t(mt * 32) * (sCm / 2));
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int m = mb + 8 * (v >> 2) + (v & 3);
        if (m >= M) continue;
        const int so = (8 * (v >> 2) + (v & 3)) * sCm;
        float r = g.alpha * acc[v];
        if (g.beta != 0.f) r += g.beta * cin[v];
        if (g.bias_mode == 2) r += bs[m];
        else if (g.bias_mode == 1) r += g.bias[n];
        if (g.pre) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(r), rp, vo, so, 0);
        if (g.act == 1) r = gelu(r);
        if (g.C) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(r), rc, vo, so, 0);
        if (g.c_bf16) __builtin_amdgcn_raw_buffer_store_b16(f2bf(r), r16, vo / 2, so / 2, 0);
      }
    }
  }
}

// The FNO's projection head in one pass, for inference:
//   out[o][n] = b2[o] + sum_j W2[o][j] bf(gelu(b1[j] + sum_i W1[j][i] bf(h[i][n])))
// with I = 32 inputs, J = 128 hidden channels and O <= 32 outputs.  The
// first product is k_pw's (same operand layout and MFMA order, so the hidden
// values are bitwise those k_pw<128, 32> writes); the hidden values never
// leave the registers: rounded to BF16 (as k_cmix rounds its input), they
// are the B operand of the second product, whose k index runs over the
// hidden rows in the order the first product's accumulators hold them (W2's
// columns are permuted to match).  Against k_pw + k_cmix, this saves
// writing and reading J FP32 values per cell; the second product's sum runs
// in the MFMA's order instead of k_cmix's sequential one (FP32 rounding).
template <int O, bool XBF>
__global__ __launch_bounds__(256) void k_proj(const void* __restrict__ hin,
                                              const float* __restrict__ W1,
                                              const float* __restrict__ b1,
                                              const float* __restrict__ W2,
                                              const float* __restrict__ b2,
                                              float* __restrict__ out, int N) {
  constexpr int I = 32, J = 128, JT = J / 32;
  static_assert(O >= 1 && O <= 32, "at most 32 outputs");
  __shared__ float bs1[J];
  __shared__ float bs2[32];
  // MFMA A fragments in LDS, one 16-byte entry per (k-step, lane): W1 as k_pw
  // holds it (row 32 jt + lane % 32, k = 16 ks + 8 (lane / 32) + 0..7); W2
  // for k-step (jt, s): entry i of lane half hf is hidden row
  // 32 jt + 8 (v / 4) + 4 hf + v % 4 with v = 8 s + i (the row the first
  // product's accumulator element v holds), output row lane % 32 (zero past O)
  __shared__ u16x8 f1[JT * 2][64];
  __shared__ u16x8 f2[JT * 2][64];
  const int lane = threadIdx.x & 63;
  const int lr = lane & 31, half = lane >> 5, kq = half * 8;
  if (threadIdx.x < J) bs1[threadIdx.x] = b1[threadIdx.x];
  if (threadIdx.x < 32) bs2[threadIdx.x] = int(threadIdx.x) < O ? b2[threadIdx.x] : 0.f;
  for (int e = threadIdx.x; e < JT * 2 * 64; e += 256) {
    const int step = e / 64, l = e % 64;
    const int jt = step / 2, ks = step % 2, r = l & 31, hf = l >> 5;
    u16x8 u1, u2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      u1[i] = f2bf(W1[(jt * 32 + r) * I + ks * 16 + hf * 8 + i]);
      const int v = 8 * ks + i;
      const int jrow = jt * 32 + 8 * (v >> 2) + 4 * hf + (v & 3);
      u2[i] = r < O ? f2bf(W2[r * J + jrow]) : (unsigned short)0;
    }
    f1[step][l] = u1;
    f2[step][l] = u2;
  }
  __syncthreads();
  const int groups = (N + 31) / 32;
  const int stride = int(gridDim.x) * 4;
  for (int gi = int(blockIdx.x) * 4 + int(threadIdx.x >> 6); gi < groups; gi += stride) {
    const int n = gi * 32 + lr;
    const bool nv = n < N;
    const int nc = nv ? n : N - 1;
    bf16x8 xb[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      u16x8 u;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t e = int64_t(ks * 16 + kq + j) * N + nc;
        if constexpr (XBF) u[j] = static_cast<const unsigned short*>(hin)[e];
        else u[j] = f2bf(static_cast<const float*>(hin)[e]);
      }
      xb[ks] = __builtin_bit_cast(bf16x8, u);
    }
    f32x16 acc2;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc2[v] = 0.f;
#pragma unroll 1
    for (int jt = 0; jt < JT; ++jt) {
      f32x16 acc;
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            __builtin_bit_cast(bf16x8, f1[jt * 2 + ks][lane]), xb[ks], acc, 0, 0, 0);
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        u16x8 u;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int v = 8 * s2 + i;
          const int jrow = jt * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
          u[i] = f2bf(gelu(1.0f * acc[v] + bs1[jrow]));
        }
        acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
            __builtin_bit_cast(bf16x8, f2[jt * 2 + s2][lane]), __builtin_bit_cast(bf16x8, u),
            acc2, 0, 0, 0);
      }
    }
    if (!nv) continue;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int o = 8 * (v >> 2) + 4 * half + (v & 3);
      if (o < O) out[int64_t(o) * N + n] = 1.0f * acc2[v] + bs2[o];
    }
  }
}

template <int M, int K>
static bool pw_mk(const GemmArgs& g, hipStream_t s) {
  if (g.M != M || g.K != K) return false;
  const int64_t groups = (int64_t(g.N) + 31) / 32;
  const int blocks = int(std::min<int64_t>((groups + 3) / 4, 4096));
  if (g.b_bf16) hipLaunchKernelGGL((k_pw<M, K, true>), dim3(blocks), dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_pw<M, K, false>), dim3(blocks), dim3(256), 0, s, g);
  return true;
}

// The FNO's channel-mixing shapes with K <= 32 inputs (M outputs x K inputs,
// compile-time so that the row and k bounds fold away), cells contiguous.
// The K = 128 shapes (the projection's data gradient and output layer)
// stay on k_cmix, which measured faster for them (420 vs 462 us, 352 vs
// 370 us at 2^22 and 2^21 cells).  DNX_GEMM_PW_VALU=1 keeps every shape on
// k_cmix (cross-check).
static bool launch_pw(const GemmArgs& g, hipStream_t s) {
  if (g.batch != 1 || g.sBn != 1 || g.sCn != 1) return false;
  if (g.beta != 0.f && g.sIn != 1) return false;
  if (g.M > 128 || g.K > 128 || g.N < 32) return false;
  if (g.c_bf16 && (reinterpret_cast<uintptr_t>(g.c_bf16) & 1u) != 0) return false;
  // byte offsets within a block of 16 rows of X or 32 rows of C, pre, Cin
  // fit in 31 bits
  const int64_t lim = int64_t(1) << 31;
  if (g.sBk < 0 || g.sCm < 0 || g.sIm < 0) return false;
  if ((16 * g.sBk + g.N) * 4 >= lim || (32 * g.sCm + g.N) * 4 >= lim) return false;
  if (g.beta != 0.f && (32 * g.sIm + g.N) * 4 >= lim) return false;
  return pw_mk<32, 9>(g, s) || pw_mk<32, 32>(g, s) || pw_mk<128, 32>(g, s) ||
         pw_mk<128, 6>(g, s);
}

}  // namespace dnx

using namespace dnx;

extern "C" int32_t dnx_gemm_bf16(const GemmArgs* g, void* stream) {
  DNX_REQUIRE(g && g->A && g->B && (g->C || g->c_bf16), "gemm: null operand");
  DNX_REQUIRE(g->M > 0 && g->N > 0 && g->K > 0 && g->batch > 0, "gemm: bad sizes");
  DNX_REQUIRE(g->beta == 0.f || g->Cin, "gemm: beta needs Cin");
  DNX_REQUIRE(g->bias_mode == 0 || g->bias, "gemm: bias missing");
  DNX_REQUIRE(g->batch <= 65535, "gemm: batch > 65535");
  DNX_REQUIRE(!g->c_bf16 || (g->M <= 128 && g->K <= 128),
              "gemm: a BF16 copy of C is written by the channel-mixing products only");
  dim3 grid(cdiv(g->N, GBN), cdiv(g->M, GBM), g->batch);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool generic = option(DNX_OPT_GEMM_GENERIC) != 0;
  bool done = false;
  if (!generic && option(DNX_OPT_GEMM_PW_VALU) == 0 && launch_pw(*g, s)) {
    done = true;
  } else if (!g->C) {
    DNX_REQUIRE(false, "gemm: C may be NULL (BF16 copy only) on the pointwise MFMA kernels "
                       "only (M, K of 32 x 9, 32 x 32, 128 x 32 or 128 x 6, one batch, "
                       "contiguous cells)");
  } else if (!generic && launch_cmix(*g, s)) {
    done = true;
  } else if (g->c_bf16) {
    // a layout the mixing kernels do not take: the element-wise kernel
    // below writes the BF16 copy too
  } else if (!generic && nt_ok(*g) && g->M >= XBM && g->N >= XBN &&
             g->K >= ((g->a_bf16 && g->b_bf16 && option(DNX_OPT_GEMM_BIG_REG) == 0) ? LDS_MIN_K
                                                                                 : XBIG_K) &&
             option(DNX_OPT_GEMM_SMALL) == 0) {
    int rc;
    if (g->a_bf16 && g->b_bf16 && option(DNX_OPT_GEMM_BIG_REG) == 0) rc = launch_lds(*g, s);
    else if (g->a_bf16 && g->b_bf16) rc = launch_big<true, true>(*g, s);
    else if (g->a_bf16) rc = launch_big<true, false>(*g, s);
    else if (g->b_bf16) rc = launch_big<false, true>(*g, s);
    else rc = launch_big<false, false>(*g, s);
    if (rc) return rc;
    done = true;
  } else if (!generic && nt_ok(*g)) {
    if (g->a_bf16 && g->b_bf16) hipLaunchKernelGGL((k_gemm_nt<true, true>), grid, dim3(256), 0, s, *g);
    else if (g->a_bf16) hipLaunchKernelGGL((k_gemm_nt<true, false>), grid, dim3(256), 0, s, *g);
    else if (g->b_bf16) hipLaunchKernelGGL((k_gemm_nt<false, true>), grid, dim3(256), 0, s, *g);
    else hipLaunchKernelGGL((k_gemm_nt<false, false>), grid, dim3(256), 0, s, *g);
    done = true;
  } else if (!generic) {
    if (g->a_bf16 && g->b_bf16) done = launch_tile<true, true>(*g, grid, s);
    else if (g->a_bf16) done = launch_tile<true, false>(*g, grid, s);
    else if (g->b_bf16) done = launch_tile<false, true>(*g, grid, s);
    else done = launch_tile<false, false>(*g, grid, s);
  }
  if (!done) hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, *g);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

// Few outputs, many partials (weight-gradient reductions): one output per
// thread, the partials of consecutive outputs read by consecutive lanes.
// Few outputs (the weight and bias gradients), many partials: a work-group
// of 1024 threads per 64 outputs, 16 threads per output each summing every
// 16th partial, then the 16 sums added in a fixed order (deterministic; 16x
// the loads in flight of one thread per output).
__global__ __launch_bounds__(1024) void k_sum_partials1(const float* __restrict__ part, int s,
                                                        int64_t n, float* __restrict__ out,
                                                        float beta) {
  __shared__ float red[16][64];
  const int t = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t i = int64_t(blockIdx.x) * 64 + t;
  float a = 0.f;
  if (i < n) {
#pragma unroll 4
    for (int j = g; j < s; j += 16) a += part[int64_t(j) * n + i];
  }
  red[g][t] = a;
  __syncthreads();
  if (g == 0 && i < n) {
    float r = beta != 0.f ? beta * out[i] : 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) r += red[k][t];
    out[i] = r;
  }
}

extern "C" int32_t dnx_sum_partials(const float* part, int32_t s, int64_t n, float* out,
                                    float beta, void* stream) {
  DNX_REQUIRE(part && out && s > 0 && n > 0, "sum_partials: bad arguments");
  if (n < (int64_t(1) << 16)) {
    hipLaunchKernelGGL(k_sum_partials1, dim3(cdiv(n, 64)), dim3(1024), 0,
                       static_cast<hipStream_t>(stream), part, s, n, out, beta);
    DNX_CHECK_HIP(hipGetLastError());
    return DNX_OK;
  }
  const int64_t threads = cdiv(n, 4);
  hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(threads, 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), part, s, n, out, beta);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_fno_inverse_pw(const void* Y, const void* GT, const void* h16,
                                      const float* W, int32_t transW, const float* bias,
                                      float* out, float* pre, unsigned short* out16,
                                      int32_t act, const float* gelu_pre, int32_t C, int32_t B,
                                      int32_t P, int32_t K, void* stream) {
  DNX_REQUIRE(Y && GT && h16 && W && (out || out16), "fno_inverse_pw: null operand");
  DNX_REQUIRE(!gelu_pre || (act == 0 && !pre), "fno_inverse_pw: gelu_pre excludes act and pre");
  DNX_REQUIRE(C == 32 && B > 0 && B % 8 == 0 && P > 0 && P % 256 == 0 && K >= LBK &&
                  K % LBK == 0 && int64_t(32) * B * P < (int64_t(1) << 31) * 4,
              "fno_inverse_pw: needs 32 channels, B %% 8 == 0, P %% 256 == 0, K %% 64 == 0 "
              "(C=%d B=%d P=%d K=%d)", C, B, P, K);
  auto al = [](const void* p, uintptr_t m) { return (reinterpret_cast<uintptr_t>(p) & m) == 0; };
  DNX_REQUIRE(al(Y, 15) && al(GT, 15) && al(h16, 1) && al(out, 15) && al(pre, 15) &&
                  al(out16, 7) && al(gelu_pre, 15),
              "fno_inverse_pw: misaligned operand");
  InvPwArgs a{static_cast<const unsigned short*>(Y), static_cast<const unsigned short*>(GT),
              static_cast<const unsigned short*>(h16), W, bias, out, pre, out16, gelu_pre,
              act, transW ? 1 : 0, B, P, K};
  static std::atomic<bool> attr{false};
  if (!attr.load()) {
    DNX_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_inv_pw),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(LSMEM)));
    attr = true;
  }
  const int tn = P / 256, tiles = (B / 8) * tn;
  const int nb = ((tiles + 7) / 8) * 8;
  hipLaunchKernelGGL(k_inv_pw, dim3(nb), dim3(512), LSMEM, static_cast<hipStream_t>(stream), a,
                     tn, tiles);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_fno_proj(const void* h, int32_t h_bf16, const float* W1,
                                const float* b1, const float* W2, const float* b2, float* out,
                                int32_t I, int32_t J, int32_t O, int64_t n, void* stream) {
  DNX_REQUIRE(h && W1 && b1 && W2 && b2 && out, "fno_proj: null operand");
  DNX_REQUIRE(I == 32 && J == 128 && O >= 1 && O <= 32 && n >= 1 && n <= 0x7fffffff,
              "fno_proj: needs 32 inputs, 128 hidden channels, 1-32 outputs (I=%d J=%d O=%d "
              "n=%lld)", I, J, O, (long long)n);
  const int64_t groups = (n + 31) / 32;
  const int blocks = int(std::min<int64_t>((groups + 3) / 4, 4096));
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (O) {
#define DNX_PROJ_CASE(o)                                                                      \
  case o:                                                                                  \
    if (h_bf16)                                                                            \
      hipLaunchKernelGGL((k_proj<o, true>), dim3(blocks), dim3(256), 0, s, h, W1, b1, W2,  \
                         b2, out, int(n));                                                 \
    else                                                                                   \
      hipLaunchKernelGGL((k_proj<o, false>), dim3(blocks), dim3(256), 0, s, h, W1, b1, W2, \
                         b2, out, int(n));                                                 \
    break;
    DNX_PROJ_CASE(1) DNX_PROJ_CASE(2) DNX_PROJ_CASE(3) DNX_PROJ_CASE(4) DNX_PROJ_CASE(5)
    DNX_PROJ_CASE(6) DNX_PROJ_CASE(8) DNX_PROJ_CASE(16) DNX_PROJ_CASE(32)
#undef DNX_PROJ_CASE
    default:
      DNX_REQUIRE(false, "fno_proj: %d outputs (built for 1-6, 8, 16, 32)", O);
  }
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_gelu_bwd(const float* dy, const float* pre, float* dx,
                                unsigned short* dx_bf16, int64_t n, void* stream) {
  DNX_REQUIRE(dy && pre && dx && n > 0, "gelu_bwd: bad arguments");
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  const int vec = (n % 4 == 0 && al(dy) && al(pre) && al(dx) &&
                   (reinterpret_cast<uintptr_t>(dx_bf16) & 7u) == 0) ? 1 : 0;
  hipLaunchKernelGGL(k_gelu_bwd, dim3(cdiv(cdiv(n, 4), 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), dy, pre, dx, dx_bf16, n, vec);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

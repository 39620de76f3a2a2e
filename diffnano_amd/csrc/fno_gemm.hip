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

#include "../../include/dnx_fno.h"
#include "dnx_common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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
      }
    }
  }
}

// Elementwise helpers of the FNO (FP32).
// GELU backward: dx = dy * gelu'(pre)
__global__ void k_gelu_bwd(const float* __restrict__ dy, const float* __restrict__ pre,
                           float* __restrict__ dx, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = pre[i];
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  dx[i] = dy[i] * (cdf + x * pdf);
}

}  // namespace dnx

using namespace dnx;

extern "C" int32_t dnx_gemm_bf16(const GemmArgs* g, void* stream) {
  DNX_REQUIRE(g && g->A && g->B && g->C, "gemm: null operand");
  DNX_REQUIRE(g->M > 0 && g->N > 0 && g->K > 0 && g->batch > 0, "gemm: bad sizes");
  DNX_REQUIRE(g->beta == 0.f || g->Cin, "gemm: beta needs Cin");
  DNX_REQUIRE(g->bias_mode == 0 || g->bias, "gemm: bias missing");
  DNX_REQUIRE(g->batch <= 65535, "gemm: batch > 65535");
  dim3 grid(cdiv(g->N, GBN), cdiv(g->M, GBM), g->batch);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, static_cast<hipStream_t>(stream), *g);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_gelu_bwd(const float* dy, const float* pre, float* dx, int64_t n,
                                void* stream) {
  DNX_REQUIRE(dy && pre && dx && n > 0, "gelu_bwd: bad arguments");
  hipLaunchKernelGGL(k_gelu_bwd, dim3(cdiv(n, 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), dy, pre, dx, n);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

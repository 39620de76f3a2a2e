// Spectral channel mixing of the FNO surrogate (per Fourier mode, complex,
// FP32 on the vector ALU).  The products are tiny (per mode a B x 32 by
// 32 x 32 complex product, K = 32) and the data are the truncated spectra
// of all channels and samples, so the kernels are bound by memory traffic,
// not arithmetic; what they fix is the access pattern: spectra rows (c, b)
// hold the real parts of all modes, then the imaginary parts, so one mode's
// values of different rows are 2M elements apart.  A work-group takes 16
// consecutive modes (64-byte runs of a row) and a chunk of 8 samples,
// stages that spectra tile in LDS, transposed to [sample][mode][channel],
// and each thread owns one (output channel, mode) pair.
//
//   k_mode_mix:   Y[q][b][mu] = sum_p X[p][b][mu] (Wr + i s Wi)[p][q][mu]
//                 (s = +1; s = -1 gives the conjugate, for dX = dY W^H)
//   k_mode_wgrad: dW[p][q][mu] = sum_b conj(X[p][b][mu]) dY[q][b][mu]
//                 (partial sums over sample chunks)
//
// Weights are passed mode-fastest ([p][q][mu], real and imaginary planes)
// so that the 16 lanes of a mode run read 64 contiguous bytes.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/dnx_fno.h"
#include "dnx_common.h"

namespace dnx {

constexpr int MIX_C = 32;    // channels
constexpr int MIX_MB = 16;   // modes per work-group
constexpr int MIX_BB = 8;    // samples per chunk (LDS tile 36 KiB)
constexpr int MIX_LD = 36;   // padded channel run in LDS (rows 4 banks apart)

// Stage X[p][b0 + b][r*M + mu0 + m] (p < 32, b < 8, m < 16, r = 0, 1) as
// Xs[r][b][m][p]: 2048 float4 reads, 4 per thread of 512.
__device__ __forceinline__ void mix_stage(const float* __restrict__ X, int B, int M, int64_t ld,
                                          int mu0, int b0, float* Xs, int t) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int idx = j * 512 + t;
    const int r = idx >> 10;  // real / imaginary plane
    const int rem = idx & 1023;
    const int row = rem >> 2;  // p * 8 + b
    const int m4 = rem & 3;
    const int p = row >> 3, b = row & 7;
    const float4 v = *reinterpret_cast<const float4*>(
        X + (int64_t(p) * B + b0 + b) * ld + int64_t(r) * M + mu0 + m4 * 4);
    float* dst = Xs + ((r * MIX_BB + b) * MIX_MB + m4 * 4) * MIX_LD + p;
    dst[0] = v.x;
    dst[MIX_LD] = v.y;
    dst[2 * MIX_LD] = v.z;
    dst[3 * MIX_LD] = v.w;
  }
}

// TY = float, or unsigned short for BF16 output (round to nearest even: the
// rounding the GEMM that consumes Y would apply to an FP32 operand).
template <typename TY>
__global__ __launch_bounds__(512) void k_mode_mix(const float* __restrict__ X,
                                                  const float* __restrict__ Wr,
                                                  const float* __restrict__ Wi,
                                                  TY* __restrict__ Y, int B, int M, int64_t ld,
                                                  float s) {
  __shared__ float Xs[2 * MIX_BB * MIX_MB * MIX_LD];
  const int t = threadIdx.x;
  const int m = t & 15, q = t >> 4;
  const int mu0 = blockIdx.x * MIX_MB, mu = mu0 + m;
  // this thread's weights over p (coalesced: the 16 lanes of a mode run)
  float wr[MIX_C], wi[MIX_C];
#pragma unroll
  for (int p = 0; p < MIX_C; ++p) {
    wr[p] = Wr[(int64_t(p) * MIX_C + q) * M + mu];
    wi[p] = s * Wi[(int64_t(p) * MIX_C + q) * M + mu];
  }
  for (int b0 = blockIdx.y * MIX_BB; b0 < B; b0 += gridDim.y * MIX_BB) {
    __syncthreads();  // the previous chunk's reads are done
    mix_stage(X, B, M, ld, mu0, b0, Xs, t);
    __syncthreads();
    for (int b = 0; b < MIX_BB; ++b) {
      const float* xr = Xs + ((0 * MIX_BB + b) * MIX_MB + m) * MIX_LD;
      const float* xi = Xs + ((1 * MIX_BB + b) * MIX_MB + m) * MIX_LD;
      float yr = 0.f, yi = 0.f;
#pragma unroll
      for (int p4 = 0; p4 < MIX_C; p4 += 4) {
        const float4 a = *reinterpret_cast<const float4*>(xr + p4);
        const float4 c = *reinterpret_cast<const float4*>(xi + p4);
        const float ar[4] = {a.x, a.y, a.z, a.w};
        const float ai[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          yr += ar[e] * wr[p4 + e] - ai[e] * wi[p4 + e];
          yi += ar[e] * wi[p4 + e] + ai[e] * wr[p4 + e];
        }
      }
      TY* yrow = Y + (int64_t(q) * B + b0 + b) * ld;
      if constexpr (std::is_same<TY, float>::value) {
        yrow[mu] = yr;
        yrow[M + mu] = yi;
      } else {
        yrow[mu] = __builtin_bit_cast(unsigned short, (__bf16)yr);
        yrow[M + mu] = __builtin_bit_cast(unsigned short, (__bf16)yi);
      }
    }
  }
}

__global__ __launch_bounds__(512) void k_mode_wgrad(const float* __restrict__ X,
                                                    const float* __restrict__ dY,
                                                    float* __restrict__ dWr,
                                                    float* __restrict__ dWi, int B, int M,
                                                    int64_t ld, int chunks_per_split) {
  __shared__ float Xs[2 * MIX_BB * MIX_MB * MIX_LD];
  const int t = threadIdx.x;
  const int m = t & 15, q = t >> 4;
  const int mu0 = blockIdx.x * MIX_MB, mu = mu0 + m;
  float ar_[MIX_C], ai_[MIX_C];
#pragma unroll
  for (int p = 0; p < MIX_C; ++p) ar_[p] = ai_[p] = 0.f;
  const int c0 = blockIdx.y * chunks_per_split;
  for (int c = c0; c < c0 + chunks_per_split && c * MIX_BB < B; ++c) {
    const int b0 = c * MIX_BB;
    __syncthreads();
    mix_stage(X, B, M, ld, mu0, b0, Xs, t);
    __syncthreads();
    for (int b = 0; b < MIX_BB; ++b) {
      const float* drow = dY + (int64_t(q) * B + b0 + b) * ld;
      const float dr = drow[mu], di = drow[M + mu];
      const float* xr = Xs + ((0 * MIX_BB + b) * MIX_MB + m) * MIX_LD;
      const float* xi = Xs + ((1 * MIX_BB + b) * MIX_MB + m) * MIX_LD;
#pragma unroll
      for (int p4 = 0; p4 < MIX_C; p4 += 4) {
        const float4 a = *reinterpret_cast<const float4*>(xr + p4);
        const float4 c4 = *reinterpret_cast<const float4*>(xi + p4);
        const float xa[4] = {a.x, a.y, a.z, a.w};
        const float xb[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ar_[p4 + e] += xa[e] * dr + xb[e] * di;
          ai_[p4 + e] += xa[e] * di - xb[e] * dr;
        }
      }
    }
  }
  const int64_t base = int64_t(blockIdx.y) * MIX_C * MIX_C * M;
#pragma unroll
  for (int p = 0; p < MIX_C; ++p) {
    dWr[base + (int64_t(p) * MIX_C + q) * M + mu] = ar_[p];
    dWi[base + (int64_t(p) * MIX_C + q) * M + mu] = ai_[p];
  }
}

// Weight and bias gradients of a pointwise (1x1) layer over n = B*H*W
// columns, in partial sums over column splits:
//   part[s][o][i] = sum_{n in split s} dY[o][n] X[i][n],  part_b[s][o] = sum dY[o][n]
// A work-group owns one split and an OT x IT (o, i) tile (32 x 32, or
// 128 x 32 so that such a layer reads each operand once); per 64-column
// step it stages dY[o][k] and X[i][k] (transposed to [k][i]) in LDS, and
// each of 256 threads accumulates OT/32 x IT/32 groups of 4 outputs (rows
// o + 32a, four consecutive i at i4 + 32b).  FP32; per output the sum runs
// over k in order, whatever the tile.  The transposed X rows are
// XOR-swizzled in 4-element groups, Xs[k][i ^ wsw(k)], so that the
// transposing stores of a wave (16 k values x 4 i values) fall in 64
// distinct banks; the float4 reads of four consecutive i stay contiguous.
__device__ __forceinline__ int wsw(int k) { return ((k >> 4) & 3) << 2; }

template <int OT, int IT>
__global__ __launch_bounds__(256) void k_wgrad(const float* __restrict__ dY,
                                               const float* __restrict__ X,
                                               float* __restrict__ part,
                                               float* __restrict__ part_b, int O, int I,
                                               int64_t n, int64_t chunk) {
  static_assert((OT == 32 || OT == 128) && IT == 32, "tiles of 32 x 32 or 128 x 32");
  constexpr int OE = OT / 32, IE = IT / 32;
  constexpr int DL = OT / 16, XL = IT / 16;  // float4 loads per thread and step
  __shared__ float Ds[OT][68];
  __shared__ float Xs[64][IT + 4];
  const int t = threadIdx.x;
  const int o = t >> 3, i4 = (t & 7) * 4;
  const int o0 = blockIdx.y * OT, i0 = blockIdx.z * IT;
  const int64_t k_begin = int64_t(blockIdx.x) * chunk;
  const int64_t k_end = min(n, k_begin + chunk);
  float acc[OE][IE][4], accb[OE];
#pragma unroll
  for (int a = 0; a < OE; ++a) {
    accb[a] = 0.f;
#pragma unroll
    for (int b = 0; b < IE; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[a][b][e] = 0.f;
  }
  // register prefetch of the next 64-column step (loads in flight during
  // the current step's arithmetic); n % 64 == 0 and 64-aligned splits, so a
  // step is either entirely inside the split or past its end
  float4 dv[DL], xv[XL];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < DL; ++j) {
      const int idx = j * 256 + t;
      const int r = idx >> 4, c4 = (idx & 15) * 4;
      dv[j] = float4{0.f, 0.f, 0.f, 0.f};
      if (k0 < k_end && o0 + r < O)
        dv[j] = *reinterpret_cast<const float4*>(dY + int64_t(o0 + r) * n + k0 + c4);
    }
#pragma unroll
    for (int j = 0; j < XL; ++j) {
      const int idx = j * 256 + t;
      const int r = idx >> 4, c4 = (idx & 15) * 4;
      xv[j] = float4{0.f, 0.f, 0.f, 0.f};
      if (k0 < k_end && i0 + r < I)
        xv[j] = *reinterpret_cast<const float4*>(X + int64_t(i0 + r) * n + k0 + c4);
    }
  };
  fetch(k_begin);
  for (int64_t k0 = k_begin; k0 < k_end; k0 += 64) {
    __syncthreads();  // the previous step's LDS reads are done
#pragma unroll
    for (int j = 0; j < DL; ++j) {
      const int idx = j * 256 + t;
      const int r = idx >> 4, c4 = (idx & 15) * 4;
      *reinterpret_cast<float4*>(&Ds[r][c4]) = dv[j];
    }
#pragma unroll
    for (int j = 0; j < XL; ++j) {
      const int idx = j * 256 + t;
      const int r = idx >> 4, c4 = (idx & 15) * 4;
      Xs[c4][r ^ wsw(c4)] = xv[j].x;
      Xs[c4 + 1][r ^ wsw(c4 + 1)] = xv[j].y;
      Xs[c4 + 2][r ^ wsw(c4 + 2)] = xv[j].z;
      Xs[c4 + 3][r ^ wsw(c4 + 3)] = xv[j].w;
    }
    __syncthreads();
    fetch(k0 + 64);
#pragma unroll 8
    for (int k = 0; k < 64; ++k) {
      float d[OE];
#pragma unroll
      for (int a = 0; a < OE; ++a) d[a] = Ds[o + 32 * a][k];
#pragma unroll
      for (int b = 0; b < IE; ++b) {
        const float4 x = *reinterpret_cast<const float4*>(&Xs[k][(i4 + 32 * b) ^ wsw(k)]);
#pragma unroll
        for (int a = 0; a < OE; ++a) {
          acc[a][b][0] += d[a] * x.x;
          acc[a][b][1] += d[a] * x.y;
          acc[a][b][2] += d[a] * x.z;
          acc[a][b][3] += d[a] * x.w;
        }
      }
#pragma unroll
      for (int a = 0; a < OE; ++a) accb[a] += d[a];
    }
  }
  const int64_t base = int64_t(blockIdx.x) * O * I;
#pragma unroll
  for (int a = 0; a < OE; ++a) {
    const int oo = o0 + o + 32 * a;
    if (oo >= O) continue;
#pragma unroll
    for (int b = 0; b < IE; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ii = i0 + i4 + 32 * b + e;
        if (ii < I) part[base + int64_t(oo) * I + ii] = acc[a][b][e];
      }
    if (part_b && blockIdx.z == 0 && i4 == 0) part_b[int64_t(blockIdx.x) * O + oo] = accb[a];
  }
}

}  // namespace dnx

using namespace dnx;

static bool mix_ok(const void* X, int32_t B, int32_t M, int64_t ld, int32_t c) {
  return c == MIX_C && B > 0 && B % MIX_BB == 0 && M > 0 && M % MIX_MB == 0 && ld >= 2 * M &&
         ld % 4 == 0 && (reinterpret_cast<uintptr_t>(X) & 15u) == 0;
}

extern "C" int32_t dnx_fno_mode_mix(const float* X, const float* Wr, const float* Wi, void* Y,
                                    int32_t y_bf16, int32_t B, int32_t M, int64_t ld, int32_t c,
                                    int32_t conj, void* stream) {
  DNX_REQUIRE(X && Wr && Wi && Y, "fno_mode_mix: null operand");
  DNX_REQUIRE(mix_ok(X, B, M, ld, c),
              "fno_mode_mix: needs 32 channels, B %% 8 == 0, M %% 16 == 0, ld >= 2M, ld %% 4 "
              "== 0, 16-byte aligned X (B=%d M=%d ld=%lld c=%d)", B, M, (long long)ld, c);
  const int chunks = B / MIX_BB;
  const int gy = chunks < 32 ? chunks : 32;
  if (y_bf16)
    hipLaunchKernelGGL(k_mode_mix<unsigned short>, dim3(M / MIX_MB, gy), dim3(512), 0,
                       static_cast<hipStream_t>(stream), X, Wr, Wi,
                       static_cast<unsigned short*>(Y), B, M, ld, conj ? -1.f : 1.f);
  else
    hipLaunchKernelGGL(k_mode_mix<float>, dim3(M / MIX_MB, gy), dim3(512), 0,
                       static_cast<hipStream_t>(stream), X, Wr, Wi, static_cast<float*>(Y), B, M,
                       ld, conj ? -1.f : 1.f);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_fno_mode_wgrad(const float* X, const float* dY, float* dWr, float* dWi,
                                      int32_t B, int32_t M, int64_t ld, int32_t c,
                                      int32_t nsplit, void* stream) {
  DNX_REQUIRE(X && dY && dWr && dWi, "fno_mode_wgrad: null operand");
  DNX_REQUIRE(mix_ok(X, B, M, ld, c), "fno_mode_wgrad: bad shape (B=%d M=%d c=%d)", B, M, c);
  const int chunks = B / MIX_BB;
  DNX_REQUIRE(nsplit >= 1 && nsplit <= chunks && chunks % nsplit == 0,
              "fno_mode_wgrad: nsplit must divide B / 8 (got %d)", nsplit);
  hipLaunchKernelGGL(k_mode_wgrad, dim3(M / MIX_MB, nsplit), dim3(512), 0,
                     static_cast<hipStream_t>(stream), X, dY, dWr, dWi, B, M, ld,
                     chunks / nsplit);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_fno_wgrad(const float* dY, const float* X, float* part, float* part_b,
                                 int32_t O, int32_t I, int64_t n, int32_t nsplit, void* stream) {
  DNX_REQUIRE(dY && X && part, "fno_wgrad: null operand");
  DNX_REQUIRE(O > 0 && I > 0 && n > 0 && n % 64 == 0 && nsplit > 0 && nsplit <= 65535 &&
                  (reinterpret_cast<uintptr_t>(dY) & 15u) == 0 &&
                  (reinterpret_cast<uintptr_t>(X) & 15u) == 0,
              "fno_wgrad: needs n %% 64 == 0 and 16-byte aligned operands (O=%d I=%d n=%lld)", O,
              I, (long long)n);
  const int64_t chunk = ((n / nsplit + 63) / 64) * 64;
  const int splits = int((n + chunk - 1) / chunk);
  DNX_REQUIRE(splits == nsplit, "fno_wgrad: nsplit %d does not split n = %lld into 64-column "
              "multiples", nsplit, (long long)n);
  // 33-128 outputs over at most 32 inputs run as one tile, which reads each
  // operand once (469 -> 369 us on 128 x 32 over 2M columns).  The mirrored
  // shape keeps 32 x 32 tiles: its 128-row transposed X tile costs more LDS
  // stores than the second read of dY saves (666 against 504 us).
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (O > 32 && O <= 128 && I <= 32)
    hipLaunchKernelGGL((k_wgrad<128, 32>), dim3(nsplit, 1, 1), dim3(256), 0, s, dY, X, part,
                       part_b, O, I, n, chunk);
  else
    hipLaunchKernelGGL((k_wgrad<32, 32>), dim3(nsplit, (O + 31) / 32, (I + 31) / 32), dim3(256),
                       0, s, dY, X, part, part_b, O, I, n, chunk);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

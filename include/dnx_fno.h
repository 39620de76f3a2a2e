/*
 * dnx_fno.h -- C ABI of the BF16 MFMA kernels behind the Fourier-neural-
 * operator surrogate (libdnx_fdtd.so).  No reference counterpart: DiffNano's
 * surrogate (diffnano/solvers/surrogate.py:23-75) is a PyTorch CNN for RCWA;
 * its prescreen-then-verify pattern (proxy_prescreen.py:180-210) is what the
 * FNO prescreen of the FDTD path follows.
 *
 * Same conventions as dnx_fdtd.h: device pointers, asynchronous on the given
 * HIP stream, 0 on success, negative code + dnx_last_error() on failure.
 */
#ifndef DNX_FNO_H
#define DNX_FNO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

/* C[b] = act(alpha * A[b] B[b] + beta * Cin[b] + bias), b < batch.
 * A[m][k] at A + b*sAb + m*sAm + k*sAk (element strides; any layout, so
 * transposes are free); likewise B[k][n], C[m][n], Cin[m][n].  A and B are
 * FP32 (a_bf16/b_bf16 = 0) or BF16 (= 1) in memory and are multiplied in
 * BF16 on MFMA with FP32 accumulation; C, Cin, bias, pre are FP32.
 * bias_mode: 0 none, 1 per column, 2 per row.  act: 0 none, 1 GELU (erf).
 * pre: if not NULL, receives the value before the activation.
 * c_bf16: if not NULL, also receives C rounded to BF16 (nearest even), at
 * C's element offsets -- the operand a later BF16 product reads instead of
 * rounding C itself (same values, half the bytes); channel-mixing
 * products only (M, K <= 128), others are rejected.  C may then be NULL
 * (only the BF16 copy is written) on the pointwise MFMA shapes (M x K of
 * 32 x 9, 32 x 32, 128 x 32 or 128 x 6, batch 1, contiguous cells); other
 * products with a NULL C are rejected. */
typedef struct dnx_gemm_args {
  int32_t M, N, K, batch;
  const void* A;
  int64_t sAm, sAk, sAb;
  const void* B;
  int64_t sBk, sBn, sBb;
  float* C;
  int64_t sCm, sCn, sCb;
  const float* Cin;
  int64_t sIm, sIn, sIb;
  float alpha, beta;
  const float* bias;
  int32_t bias_mode;
  int32_t act;
  float* pre;
  int32_t a_bf16, b_bf16;
  unsigned short* c_bf16;
} dnx_gemm_args;

int32_t dnx_gemm_bf16(const dnx_gemm_args* args, void* stream);

/* out[i] = sum_{j<s} part[j*n + i] + beta * out[i], FP32: the reduction
 * step of a split-K product (partial products written as a batch). */
int32_t dnx_sum_partials(const float* part, int32_t s, int64_t n, float* out, float beta,
                         void* stream);

/* The projection head of the FNO in one pass (inference; no reference
 * counterpart beyond the surrogate's output layers, surrogate.py:23-75):
 *   out[o][n] = b2[o] + sum_j W2[o][j] bf(gelu(b1[j] + sum_i W1[j][i] bf(h[i][n])))
 * h [I][n] FP32 (h_bf16 = 0) or BF16 (= 1, used as is), W1 [J][I],
 * W2 [O][J], out [O][n], FP32 row-major; bf() rounds to BF16 (nearest
 * even), products on the matrix cores with FP32 accumulation.  I = 32,
 * J = 128, O in 1-6, 8, 16 or 32. */
int32_t dnx_fno_proj(const void* h, int32_t h_bf16, const float* W1, const float* b1,
                     const float* W2, const float* b2, float* out, int32_t I, int32_t J,
                     int32_t O, int64_t n, void* stream);

/* One FNO layer's inverse DFT and pointwise part in one pass (no reference
 * counterpart, as above):
 *   z[c][b P + p] = sum_k Y[c B + b][k] GT[p][k] + sum_c' W[c][c'] h16[c'][b P + p]
 *                   + bias[c]
 *   out = act(z), or out = z * GELU'(gelu_pre) (the backward's form)
 * Y [C B][K], GT [P][K], h16 [C][B P] BF16 (the layer's input rounded);
 * W [C][C] (read transposed if transW), bias [C] (optional) FP32, W rounded
 * to BF16; out, pre (z) FP32 and out16 (out rounded to BF16) each
 * optional, at least one of out and out16; gelu_pre excludes act and pre.
 * FP32 accumulation, the two sums added in the accumulators.  C = 32, B a
 * multiple of 8, P of 256, K of 64; Y, GT, out, pre, gelu_pre 16-byte and
 * out16 8-byte aligned. */
int32_t dnx_fno_inverse_pw(const void* Y, const void* GT, const void* h16, const float* W,
                           int32_t transW, const float* bias, float* out, float* pre,
                           unsigned short* out16, int32_t act, const float* gelu_pre,
                           int32_t C, int32_t B, int32_t P, int32_t K, void* stream);

/* dx[i] = dy[i] * GELU'(pre[i]) (erf form), FP32; dx_bf16, if not NULL,
 * also receives dx rounded to BF16 (nearest even). */
int32_t dnx_gelu_bwd(const float* dy, const float* pre, float* dx, unsigned short* dx_bf16,
                     int64_t n, void* stream);

/* Spectral channel mixing of an FNO layer, per Fourier mode, complex, FP32:
 *   Y[q][b][r*M + mu] (r = 0 real, 1 imaginary part)
 *     = sum_p X[p][b][.] (Wr + i s Wi)[p][q][mu],   s = -1 if conj else +1
 * X rows (p, b) and Y rows (q, b) have length ld (>= 2M, multiple of 4; X
 * 16-byte aligned); Wr, Wi are [p][q][mu] (mode fastest).  Y is FP32, or
 * BF16 (round to nearest even) if y_bf16 -- the rounding a BF16 MFMA
 * product applies to an FP32 operand anyway.  32 channels,
 * B % 8 == 0, M % 16 == 0.  (The spectral convolution of an FNO layer,
 * no reference counterpart.) */
int32_t dnx_fno_mode_mix(const float* X, const float* Wr, const float* Wi, void* Y,
                         int32_t y_bf16, int32_t B, int32_t M, int64_t ld, int32_t c,
                         int32_t conj, void* stream);
/* Its weight gradient in nsplit partial sums over sample chunks:
 *   dWr[s][p][q][mu] + i dWi[s][p][q][mu] = sum_{b in split s} conj(X[p][b][mu]) dY[q][b][mu]
 * (nsplit divides B / 8). */
int32_t dnx_fno_mode_wgrad(const float* X, const float* dY, float* dWr, float* dWi, int32_t B,
                           int32_t M, int64_t ld, int32_t c, int32_t nsplit, void* stream);
/* Weight and bias gradients of a pointwise layer in nsplit partial sums over
 * the n columns (n % 64 == 0, split into equal 64-column multiples):
 *   part[s][o][i] = sum_{n in split s} dY[o][n] X[i][n],  part_b[s][o] = sum dY[o][n]
 * (part_b may be NULL).  dY [O][n], X [I][n], FP32. */
int32_t dnx_fno_wgrad(const float* dY, const float* X, float* part, float* part_b, int32_t O,
                      int32_t I, int64_t n, int32_t nsplit, void* stream);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* DNX_FNO_H */

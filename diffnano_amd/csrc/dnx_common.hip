// libdnx_fdtd: error state, launch timing, coefficient preparation, and the
// per-step side kernels (soft sources, running DFT monitors and adjoints).
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "dnx_aux.h"

// No multiply-add contraction (same rounding in every FDTD kernel variant).
#pragma clang fp contract(off)

namespace dnx {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

const char* last_error() { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Kernel-selection options: initial values from the environment, read once
// ---------------------------------------------------------------------------

static std::atomic<int> g_opt[DNX_OPT_COUNT];
static std::once_flag g_opt_once;

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return (v && *v) ? atoi(v) : dflt;
}

static void init_options() {
  const char* var = getenv("DNX_FDTD3D_VARIANT");
  g_opt[DNX_OPT_FDTD3D_SIMPLE] = (var && strcmp(var, "simple") == 0) ? 1 : 0;
  g_opt[DNX_OPT_FDTD3D_ONE_STEP] = env_int("DNX_FDTD3D_STEPS", 2) == 1 ? 1 : 0;
  g_opt[DNX_OPT_FDTD3D_ADJ2] = env_int("DNX_FDTD3D_ADJ2", 1) != 0 ? 1 : 0;
  g_opt[DNX_OPT_FDTD3D_FWD2_PF] = env_int("DNX_FDTD3D_FWD2_PF", 0) != 0 ? 1 : 0;
  g_opt[DNX_OPT_FDTD3D_ADJ2_OCC] = env_int("DNX_FDTD3D_ADJ2_OCC", 2) == 1 ? 1 : 2;
  g_opt[DNX_OPT_GEMM_GENERIC] = env_int("DNX_GEMM_GENERIC", 0) != 0 ? 1 : 0;
  g_opt[DNX_OPT_GEMM_PW_VALU] = env_int("DNX_GEMM_PW_VALU", 0) != 0 ? 1 : 0;
  g_opt[DNX_OPT_GEMM_SMALL] = env_int("DNX_GEMM_SMALL", 0) != 0 ? 1 : 0;
  g_opt[DNX_OPT_GEMM_BIG_REG] = env_int("DNX_GEMM_BIG_REG", 0) != 0 ? 1 : 0;
  g_opt[DNX_OPT_FDTD2D_LAUNCH_PER_STEP] = std::min(2, std::max(0, env_int("DNX_FDTD2D_PER_STEP", 0)));
  g_opt[DNX_OPT_FDTD2D_SEPARATE_SIDE] = env_int("DNX_FDTD2D_SEPARATE", 0) != 0 ? 1 : 0;
  g_opt[DNX_OPT_FDTD2D_TB_STEPS] = std::max(0, env_int("DNX_FDTD2D_TB", 0));
  g_opt[DNX_OPT_FDTD2D_TB_REGION] = env_int("DNX_FDTD2D_TB_REGION", 0);
}

int option(int id) {
  std::call_once(g_opt_once, init_options);
  return (id >= 0 && id < DNX_OPT_COUNT) ? g_opt[id].load(std::memory_order_relaxed) : 0;
}

// ---------------------------------------------------------------------------
// Per-launch timing registry
// ---------------------------------------------------------------------------

// g_timing_stride: 0 off, k >= 1 one launch in every k of each name timed
static std::atomic<int> g_timing_stride{0};
static std::mutex g_timing_mu;
static std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> g_timing;
static std::map<std::string, int64_t> g_timing_seen;

TimedLaunch timed_begin(const char* name, hipStream_t s) {
  TimedLaunch t;
  const int stride = g_timing_stride.load(std::memory_order_relaxed);
  if (stride <= 0) return t;
  if (stride > 1) {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    if (g_timing_seen[name]++ % stride != 0) return t;
  }
  if (hipEventCreate(&t.start) != hipSuccess) {
    t.start = nullptr;
    return t;
  }
  if (hipEventRecord(t.start, s) != hipSuccess) {
    (void)hipEventDestroy(t.start);
    t.start = nullptr;
  }
  return t;
}

void timed_end(const char* name, hipStream_t s, TimedLaunch& t) {
  if (t.start == nullptr) return;
  hipEvent_t stop = nullptr;
  if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, s) != hipSuccess) {
    if (stop) (void)hipEventDestroy(stop);
    (void)hipEventDestroy(t.start);
    t.start = nullptr;
    return;
  }
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing[name].emplace_back(t.start, stop);
  t.start = nullptr;
}

int check_sources(const dnx_source* s, int n, int nz, int ny, int nx, int max_comp) {
  DNX_REQUIRE(n >= 0 && (n == 0 || s != nullptr), "sources: null array with n_src=%d", n);
  for (int i = 0; i < n; ++i) {
    DNX_REQUIRE(s[i].comp >= 0 && s[i].comp <= max_comp, "source %d: bad component %d", i,
                s[i].comp);
    DNX_REQUIRE(box_ok(s[i].box, nz, ny, nx), "source %d: box outside the grid", i);
    DNX_REQUIRE(s[i].profile && s[i].wave, "source %d: null profile or waveform", i);
  }
  return DNX_OK;
}

int check_monitors(const dnx_monitor* m, int n, int nz, int ny, int nx, int comp_bits,
                   bool need_grad) {
  DNX_REQUIRE(n >= 0 && (n == 0 || m != nullptr), "monitors: null array with n_mon=%d", n);
  for (int i = 0; i < n; ++i) {
    DNX_REQUIRE(box_ok(m[i].box, nz, ny, nx), "monitor %d: box outside the grid", i);
    DNX_REQUIRE(m[i].comp_mask > 0 && (m[i].comp_mask & ~comp_bits) == 0,
                "monitor %d: bad component mask %d", i, m[i].comp_mask);
    DNX_REQUIRE(m[i].n_freq > 0 && m[i].phase, "monitor %d: no frequencies", i);
    if (need_grad) {
      DNX_REQUIRE(m[i].grad != nullptr, "monitor %d: null gradient", i);
    } else {
      DNX_REQUIRE(m[i].acc != nullptr, "monitor %d: null accumulator", i);
    }
  }
  return DNX_OK;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_prepare_coef(const double* __restrict__ eps, T* __restrict__ ce, int64_t n,
                               double k) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) ce[i] = T(k / eps[i]);
}

template <typename T>
__global__ void k_finalize_grad(const T* __restrict__ S, const double* __restrict__ eps,
                                double* __restrict__ g, int64_t n) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) g[i] = -double(S[i]) / eps[i];
}

template <typename T>
__global__ void k_zero(T* __restrict__ p, int64_t n) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = T(0);
}

template <typename T>
__global__ void k_dot_init(T* __restrict__ S, Comps<T> a, Comps<T> b, int ncomp, int64_t n) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T s = T(0);
  for (int c = 0; c < ncomp; ++c) s += a.c[c][i] * b.c[c][i];
  S[i] = s;
}

__device__ inline int64_t box_index(const Box& b, int64_t i, int ny, int nx) {
  int bx = int(i % b.nx);
  int64_t r = i / b.nx;
  int by = int(r % b.ny);
  int bz = int(r / b.ny);
  return (int64_t(b.z0 + bz) * ny + (b.y0 + by)) * nx + (b.x0 + bx);
}

// f[cell] += wave[n] * prof[cell in box]
template <typename T>
__global__ void k_source(T* __restrict__ f, int ny, int nx, Box b, const T* __restrict__ prof,
                         const T* __restrict__ wave_n) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.cells()) return;
  f[box_index(b, i, ny, nx)] += wave_n[0] * prof[i];
}

template <typename T>
__global__ void k_source_adj(T* __restrict__ S, const T* __restrict__ lam, int ny, int nx, Box b,
                             const T* __restrict__ prof, const T* __restrict__ wave_m) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.cells()) return;
  int64_t idx = box_index(b, i, ny, nx);
  S[idx] -= wave_m[0] * prof[i] * lam[idx];
}

template <typename T>
__global__ void k_monitor(Comps<T> E, int ny, int nx, Box b, int comp_mask, int nf,
                          const double* __restrict__ ph, double* __restrict__ acc) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t cells = b.cells();
  if (i >= cells) return;
  int64_t idx = box_index(b, i, ny, nx);
  int ncomp = ((comp_mask >> 0) & 1) + ((comp_mask >> 1) & 1) + ((comp_mask >> 2) & 1);
  int ci = 0;
  for (int c = 0; c < 3; ++c) {
    if (!((comp_mask >> c) & 1)) continue;
    double v = double(E.c[c][idx]);
    for (int f = 0; f < nf; ++f) {
      double* a = acc + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
      a[0] += v * ph[2 * f];
      a[1] -= v * ph[2 * f + 1];
    }
    ++ci;
  }
}

// One cell (box-local index i < cells) of the monitor adjoint at one sample:
// lam += dv, S += E . dv with dv = T(sum_f Re(g e^{i phase})).
template <typename T>
__device__ __forceinline__ void monitor_adj_cell(Comps<T> lam, Comps<T> E, T* __restrict__ S,
                                                 int use_s, int ny, int nx, const Box& b,
                                                 int comp_mask, int nf,
                                                 const double* __restrict__ ph,
                                                 const double* __restrict__ g, int64_t i) {
  int64_t cells = b.cells();
  int64_t idx = box_index(b, i, ny, nx);
  int ncomp = ((comp_mask >> 0) & 1) + ((comp_mask >> 1) & 1) + ((comp_mask >> 2) & 1);
  int ci = 0;
  T sacc = T(0);
  for (int c = 0; c < 3; ++c) {
    if (!((comp_mask >> c) & 1)) continue;
    double d = 0.0;
    for (int f = 0; f < nf; ++f) {
      const double* gg = g + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
      d += gg[0] * ph[2 * f] - gg[1] * ph[2 * f + 1];
    }
    T dv = T(d);
    lam.c[c][idx] += dv;
    if (use_s) sacc += E.c[c][idx] * dv;
    ++ci;
  }
  if (use_s) S[idx] += sacc;
}

template <typename T>
__global__ void k_monitor_adj(Comps<T> lam, Comps<T> E, T* __restrict__ S, int use_s, int ny,
                              int nx, Box b, int comp_mask, int nf,
                              const double* __restrict__ ph, const double* __restrict__ g) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.cells()) return;
  monitor_adj_cell<T>(lam, E, S, use_s, ny, nx, b, comp_mask, nf, ph, g, i);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <typename T>
int launch_source(T* comp_base, int ny, int nx, const dnx_source& s, int n, hipStream_t st) {
  Box b = make_box(s.box);
  hipLaunchKernelGGL(k_source<T>, dim3(cdiv(b.cells(), 256)), dim3(256), 0, st, comp_base, ny,
                     nx, b, static_cast<const T*>(s.profile), static_cast<const T*>(s.wave) + n);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_source_adj(T* S, const T* lam_comp, int ny, int nx, const dnx_source& s, int m,
                      hipStream_t st) {
  Box b = make_box(s.box);
  hipLaunchKernelGGL(k_source_adj<T>, dim3(cdiv(b.cells(), 256)), dim3(256), 0, st, S, lam_comp,
                     ny, nx, b, static_cast<const T*>(s.profile),
                     static_cast<const T*>(s.wave) + m);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_monitor(Comps<T> E, int ny, int nx, const dnx_monitor& m, int n, hipStream_t st) {
  Box b = make_box(m.box);
  hipLaunchKernelGGL(k_monitor<T>, dim3(cdiv(b.cells(), 256)), dim3(256), 0, st, E, ny, nx, b,
                     m.comp_mask, m.n_freq, m.phase + int64_t(n) * m.n_freq * 2, m.acc);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_monitor_adj(Comps<T> lam, Comps<T> E, T* S, int ny, int nx, const dnx_monitor& m,
                       int sample_row, hipStream_t st) {
  Box b = make_box(m.box);
  int use_s = (E.c[0] != nullptr || E.c[1] != nullptr || E.c[2] != nullptr) && S != nullptr;
  hipLaunchKernelGGL(k_monitor_adj<T>, dim3(cdiv(b.cells(), 256)), dim3(256), 0, st, lam, E, S,
                     use_s, ny, nx, b, m.comp_mask, m.n_freq,
                     m.phase + int64_t(sample_row) * m.n_freq * 2, m.grad);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_zero(T* p, int64_t n, hipStream_t st) {
  if (n <= 0) return DNX_OK;
  hipLaunchKernelGGL(k_zero<T>, dim3(cdiv(n, 256)), dim3(256), 0, st, p, n);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_dot_init(T* S, Comps<T> a, Comps<T> b, int ncomp, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_init<T>, dim3(cdiv(n, 256)), dim3(256), 0, st, S, a, b, ncomp, n);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

// Two samples (steps n, n+1) staged compactly by the two-step kernels:
// side = [2][ncomp][cells] in mask order.  Per accumulator element the
// order of additions is that of two k_monitor launches.
template <typename T>
__global__ void k_monitor_pair(const T* __restrict__ side, int64_t cells, int comp_mask, int nf,
                               const double* __restrict__ ph, double* __restrict__ acc) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  int ncomp = ((comp_mask >> 0) & 1) + ((comp_mask >> 1) & 1) + ((comp_mask >> 2) & 1);
  for (int ci = 0; ci < ncomp; ++ci) {
    const double v1 = double(side[int64_t(ci) * cells + i]);
    const double v2 = double(side[(int64_t(ncomp) + ci) * cells + i]);
    for (int f = 0; f < nf; ++f) {
      double* a = acc + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
      a[0] += v1 * ph[2 * f];
      a[1] -= v1 * ph[2 * f + 1];
      a[0] += v2 * ph[2 * nf + 2 * f];
      a[1] -= v2 * ph[2 * nf + 2 * f + 1];
    }
  }
}

// d[ci][i] = T(sum_f Re(g e^{i phi})) at one sample (the adjoint source
// k_monitor_adj adds), for the two-step adjoint kernel to apply.
template <typename T>
__device__ __forceinline__ void monitor_d_cell(T* __restrict__ d, int64_t cells, int comp_mask,
                                               int nf, const double* __restrict__ ph,
                                               const double* __restrict__ g, int64_t i) {
  int ncomp = ((comp_mask >> 0) & 1) + ((comp_mask >> 1) & 1) + ((comp_mask >> 2) & 1);
  for (int ci = 0; ci < ncomp; ++ci) {
    double s = 0.0;
    for (int f = 0; f < nf; ++f) {
      const double* gg = g + ((int64_t(f) * ncomp + ci) * cells + i) * 2;
      s += gg[0] * ph[2 * f] - gg[1] * ph[2 * f + 1];
    }
    d[int64_t(ci) * cells + i] = T(s);
  }
}

template <typename T>
__global__ void k_monitor_d(T* __restrict__ d, int64_t cells, int comp_mask, int nf,
                            const double* __restrict__ ph, const double* __restrict__ g) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  monitor_d_cell<T>(d, cells, comp_mask, nf, ph, g, i);
}

// k_monitor_adj at one sample and k_monitor_d at another in one launch (the
// two touch different buffers): the monitor adjoint after one two-step
// adjoint launch and the adjoint source the next one reads.
template <typename T>
__global__ void k_monitor_adj_d(Comps<T> lam, Comps<T> E, T* __restrict__ S, int use_s, int ny,
                                int nx, Box b, int comp_mask, int nf,
                                const double* __restrict__ ph_adj, const double* __restrict__ g,
                                T* __restrict__ d, const double* __restrict__ ph_d) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.cells()) return;
  monitor_adj_cell<T>(lam, E, S, use_s, ny, nx, b, comp_mask, nf, ph_adj, g, i);
  monitor_d_cell<T>(d, b.cells(), comp_mask, nf, ph_d, g, i);
}

template <typename T>
int launch_monitor_pair(const T* side, const dnx_monitor& m, int n, hipStream_t st) {
  Box b = make_box(m.box);
  hipLaunchKernelGGL(k_monitor_pair<T>, dim3(cdiv(b.cells(), 256)), dim3(256), 0, st, side,
                     b.cells(), m.comp_mask, m.n_freq, m.phase + int64_t(n) * m.n_freq * 2,
                     m.acc);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_monitor_d(T* d, const dnx_monitor& m, int sample_row, hipStream_t st) {
  Box b = make_box(m.box);
  hipLaunchKernelGGL(k_monitor_d<T>, dim3(cdiv(b.cells(), 256)), dim3(256), 0, st, d, b.cells(),
                     m.comp_mask, m.n_freq, m.phase + int64_t(sample_row) * m.n_freq * 2,
                     m.grad);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

template <typename T>
int launch_monitor_adj_d(Comps<T> lam, Comps<T> E, T* S, int ny, int nx, const dnx_monitor& m,
                         int adj_row, T* d, int d_row, hipStream_t st) {
  Box b = make_box(m.box);
  int use_s = (E.c[0] != nullptr || E.c[1] != nullptr || E.c[2] != nullptr) && S != nullptr;
  hipLaunchKernelGGL(k_monitor_adj_d<T>, dim3(cdiv(b.cells(), 256)), dim3(256), 0, st, lam, E, S,
                     use_s, ny, nx, b, m.comp_mask, m.n_freq,
                     m.phase + int64_t(adj_row) * m.n_freq * 2, m.grad, d,
                     m.phase + int64_t(d_row) * m.n_freq * 2);
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

#define DNX_INST(T)                                                                        \
  template int launch_monitor_adj_d<T>(Comps<T>, Comps<T>, T*, int, int, const dnx_monitor&, \
                                       int, T*, int, hipStream_t);                       \
  template int launch_monitor_pair<T>(const T*, const dnx_monitor&, int, hipStream_t);   \
  template int launch_monitor_d<T>(T*, const dnx_monitor&, int, hipStream_t);            \
  template int launch_source<T>(T*, int, int, const dnx_source&, int, hipStream_t);      \
  template int launch_source_adj<T>(T*, const T*, int, int, const dnx_source&, int,      \
                                    hipStream_t);                                        \
  template int launch_monitor<T>(Comps<T>, int, int, const dnx_monitor&, int, hipStream_t); \
  template int launch_monitor_adj<T>(Comps<T>, Comps<T>, T*, int, int, const dnx_monitor&, \
                                     int, hipStream_t);                                  \
  template int launch_zero<T>(T*, int64_t, hipStream_t);                                 \
  template int launch_dot_init<T>(T*, Comps<T>, Comps<T>, int, int64_t, hipStream_t);
DNX_INST(float)
DNX_INST(double)
#undef DNX_INST

}  // namespace dnx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" int32_t dnx_abi_version(void) { return DNX_ABI_VERSION; }

extern "C" const char* dnx_last_error(void) { return dnx::last_error(); }

extern "C" int32_t dnx_set_option(int32_t id, int32_t value) {
  DNX_REQUIRE(id >= 0 && id < DNX_OPT_COUNT, "set_option: unknown option %d", id);
  (void)dnx::option(id);  // the environment defaults are read before any change
  dnx::g_opt[id].store(value, std::memory_order_relaxed);
  return DNX_OK;
}

extern "C" int32_t dnx_get_option(int32_t id, int32_t* value) {
  DNX_REQUIRE(id >= 0 && id < DNX_OPT_COUNT && value, "get_option: bad option %d", id);
  *value = dnx::option(id);
  return DNX_OK;
}

// ---- streaming-bandwidth probes (the ceiling the FDTD kernels are held to) ----
namespace dnx {
typedef float v4f __attribute__((ext_vector_type(4)));
// kind 0: a copy of n floats, 16 bytes per lane, non-temporal loads and stores
__global__ void k_stream_copy(const v4f* __restrict__ a, v4f* __restrict__ b, int64_t n4) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
// kind 1: the streams of one FDTD step on n cells, one dword per lane and
// array: six fields and a coefficient in (7 n floats), six fields out
__global__ void k_stream_step(const float* __restrict__ in, float* __restrict__ out, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float k = in[6 * n + i];
#pragma unroll
  for (int f = 0; f < 6; ++f) out[f * n + i] = in[f * n + i] * k;
}
}  // namespace dnx

extern "C" int32_t dnx_stream_probe(int32_t kind, const void* src, void* dst, int64_t n,
                                    void* stream) {
  using namespace dnx;
  DNX_REQUIRE(src && dst && n > 0, "stream_probe: null buffer or empty size");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (kind == 0) {
    DNX_REQUIRE(n % 4 == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(dst) & 15) == 0,
                "stream_probe: the copy needs 16-byte aligned buffers of a multiple of 4 floats");
    const int64_t n4 = n / 4;
    hipLaunchKernelGGL(k_stream_copy, dim3(unsigned(cdiv(n4, 256))), dim3(256), 0, s,
                       static_cast<const v4f*>(src), static_cast<v4f*>(dst), n4);
  } else if (kind == 1) {
    hipLaunchKernelGGL(k_stream_step, dim3(unsigned(cdiv(n, 256))), dim3(256), 0, s,
                       static_cast<const float*>(src), static_cast<float*>(dst), n);
  } else {
    set_error("stream_probe: kind must be 0 (copy) or 1 (FDTD step streams), got %d", kind);
    return DNX_EINVAL;
  }
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_timing_enable(int32_t on) {
  DNX_REQUIRE(on >= 0, "timing_enable: negative stride %d", on);
  {
    std::lock_guard<std::mutex> lk(dnx::g_timing_mu);
    dnx::g_timing_seen.clear();
  }
  dnx::g_timing_stride = on;
  return DNX_OK;
}

extern "C" int32_t dnx_timing_read(const char* name, double* total_ms, int64_t* launches) {
  using namespace dnx;
  DNX_REQUIRE(name && total_ms && launches, "timing_read: null argument");
  std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    auto it = g_timing.find(name);
    if (it != g_timing.end()) evs.swap(it->second);
  }
  double tot = 0.0;
  int rc = DNX_OK;
  for (auto& p : evs) {
    float ms = 0.f;
    if (rc == DNX_OK) {
      hipError_t e = hipEventSynchronize(p.second);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.first, p.second);
      if (e != hipSuccess) {
        set_error("timing_read: %s", hipGetErrorString(e));
        rc = DNX_EHIP;
      }
    }
    tot += ms;
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  *total_ms = tot;
  *launches = int64_t(evs.size());
  return rc;
}

extern "C" int32_t dnx_prepare_coef(const double* eps, void* ce, int64_t n, double dt_over_dl,
                                    int32_t dtype, void* stream) {
  using namespace dnx;
  DNX_REQUIRE(eps && ce && n > 0, "prepare_coef: null pointer or n=%lld", (long long)n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == DNX_F32) {
    hipLaunchKernelGGL(k_prepare_coef<float>, dim3(cdiv(n, 256)), dim3(256), 0, st, eps,
                       static_cast<float*>(ce), n, dt_over_dl);
  } else if (dtype == DNX_F64) {
    hipLaunchKernelGGL(k_prepare_coef<double>, dim3(cdiv(n, 256)), dim3(256), 0, st, eps,
                       static_cast<double*>(ce), n, dt_over_dl);
  } else {
    DNX_REQUIRE(false, "prepare_coef: bad dtype %d", dtype);
  }
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

extern "C" int32_t dnx_finalize_grad(const void* S, const double* eps, double* grad_eps,
                                     int64_t n, int32_t dtype, void* stream) {
  using namespace dnx;
  DNX_REQUIRE(S && eps && grad_eps && n > 0, "finalize_grad: null pointer or n=%lld",
              (long long)n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == DNX_F32) {
    hipLaunchKernelGGL(k_finalize_grad<float>, dim3(cdiv(n, 256)), dim3(256), 0, st,
                       static_cast<const float*>(S), eps, grad_eps, n);
  } else if (dtype == DNX_F64) {
    hipLaunchKernelGGL(k_finalize_grad<double>, dim3(cdiv(n, 256)), dim3(256), 0, st,
                       static_cast<const double*>(S), eps, grad_eps, n);
  } else {
    DNX_REQUIRE(false, "finalize_grad: bad dtype %d", dtype);
  }
  DNX_CHECK_HIP(hipGetLastError());
  return DNX_OK;
}

// Launchers for the per-step side kernels shared by the 2D and 3D paths:
// soft-source injection, DFT monitor accumulation and their adjoints.
#pragma once

#include "dnx_common.h"

namespace dnx {

// Field views: up to three component base pointers of one grid.
template <typename T>
struct Comps {
  T* c[3];
};

// After step n: comp += wave[n] * profile over the box.
template <typename T>
int launch_source(T* comp_base, int ny, int nx, const dnx_source& s, int n, hipStream_t st);

// Adjoint of the source at step m: S -= wave[m] * profile * lam_comp.
template <typename T>
int launch_source_adj(T* S, const T* lam_comp, int ny, int nx, const dnx_source& s, int m,
                      hipStream_t st);

// After step n: acc += E (cos, -sin)(phase[n]) for the components in the mask.
template <typename T>
int launch_monitor(Comps<T> E, int ny, int nx, const dnx_monitor& m, int n, hipStream_t st);

// Adjoint of sample index m (the sample taken after step m-1, phase row m-1):
// d = Re(g e^{i phi}); lam_c += d; S += E^m_c * d.  When E is null only lam
// is updated (used at m = T, where S is initialised separately).
template <typename T>
int launch_monitor_adj(Comps<T> lam, Comps<T> E, T* S, int ny, int nx, const dnx_monitor& m,
                       int sample_row, hipStream_t st);

// Two consecutive samples (rows n, n+1) staged by the two-step kernels in
// side = [2][ncomp][cells] (mask order): acc += both, in step order.
template <typename T>
int launch_monitor_pair(const T* side, const dnx_monitor& m, int n, hipStream_t st);

// d = [ncomp][cells]: the adjoint source of sample `sample_row` (what
// launch_monitor_adj adds to lam), for the two-step adjoint kernel.
template <typename T>
int launch_monitor_d(T* d, const dnx_monitor& m, int sample_row, hipStream_t st);
// launch_monitor_adj at adj_row and launch_monitor_d at d_row, one launch.
template <typename T>
int launch_monitor_adj_d(Comps<T> lam, Comps<T> E, T* S, int ny, int nx, const dnx_monitor& m,
                         int adj_row, T* d, int d_row, hipStream_t st);

// Fill n elements with zero (a kernel, so it is safe inside captured graphs).
template <typename T>
int launch_zero(T* p, int64_t n, hipStream_t st);

// S[i] = sum_c a_c[i] * b_c[i] over ncomp components of length n.
template <typename T>
int launch_dot_init(T* S, Comps<T> a, Comps<T> b, int ncomp, int64_t n, hipStream_t st);

}  // namespace dnx

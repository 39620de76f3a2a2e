// Fused single-pass 3D step kernels (forward leapfrog and adjoint), shared
// declarations.  One launch advances the whole grid by one time step:
// reads (E, H, ce) once and writes (E', H') once.
#pragma once

#include <vector>

#include "dnx_aux.h"

namespace dnx {

constexpr int FUSED_MAX_W = 256;  // full-row tiles: one thread per x column

template <typename T>
struct Fused3Args {
  int nz, ny, nx;
  int ty_tiles;   // ceil(ny / TY)
  int zc;         // z planes per work-group
  int nblocks;    // ty_tiles * z-chunks (logical work-groups)
  int64_t P, N;   // plane and grid sizes
  T kh;
  const T* __restrict__ ce;
  const T* __restrict__ by;
  const T* __restrict__ ay;
  const T* __restrict__ bx;
  const T* __restrict__ ax;
  const T* __restrict__ in;  // forward: state at n; adjoint: (lamE^{m+1}, muH^{m+1})
  T* __restrict__ out;       // forward: state at n+1; adjoint: (lamE^m, muH^m)
  T* __restrict__ es;        // forward: E store slot (3N) or null
  const T* __restrict__ em;  // adjoint: E^m (3N)
  T* __restrict__ S;         // adjoint: gradient accumulator (N)
  // one fused soft source (comp < 0: none)
  int src_comp;
  Box src_box;
  const T* __restrict__ src_prof;
  const T* __restrict__ src_wave;  // points at wave[n] (forward) / wave[m] (adjoint)
};

// One design of a batched launch (device-resident table entry).
template <typename T>
struct Batch3Design {
  Fused3Args<T> a;        // geometry, coefficients, fused source, tiles; a.S set
  T* buf[2];              // ping-pong buffers (forward state or adjoint), 6N each
  T* ckpt;                // checkpoint of step k*K at ckpt + k*ckpt_stride (6N)
  T* estore;              // replay E of step kK+1+j at estore + j*estore_stride (3N)
  T* etail;               // E of step a_tail+1+j at etail + j*estore_stride (3N), kept by
  int a_tail;             //   the forward for the final segments (a_tail = T: none)
  int64_t ckpt_stride;
  int64_t estore_stride;
  const T* wave;          // source waveform table
  int tile_start;         // first global tile of this design
  // one DFT monitor (m_nf == 0: none)
  int m_mask;
  int m_nf;
  Box mbox;
  const double* m_phase;
  double* m_acc;
  const double* m_grad;
  T* m_side;              // two-step path: staged monitor values / adjoint source
};

// Buffer layouts of the batched state and adjoint pools.  Layout 0 (one-step
// path): state n in buf[n & 1], lam^m in buf[m & 1].  Layout 1 (two-step
// path, launches of two steps starting at even n): state n in
// buf[((n + 1) >> 1) & 1] and lam^m in buf[((m + 1) >> 1) & 1]; a launch at
// counter c reads buf[in] and writes buf[in ^ 1] with in = batch_in_slot(c)
// (forward) or batch_in_slot(c + 1) (adjoint, which reads lam^{c+1}).
__host__ __device__ inline int batch_in_slot(int layout, int j) {
  return layout ? ((j + 1) >> 1) & 1 : j & 1;
}

// E^m for the adjoint at step m: the checkpoint (m % K == 0), the final
// segments' store (m > a_tail) or the replay store of the current segment.
template <typename T>
__device__ inline const T* batch_em(const Batch3Design<T>& D, int m, int K) {
  const int r = m % K;
  if (r == 0) return D.ckpt + int64_t(m / K) * D.ckpt_stride;
  if (m > D.a_tail) return D.etail + int64_t(m - D.a_tail - 1) * D.estore_stride;
  return D.estore + int64_t(r - 1) * D.estore_stride;
}

// Where the forward step from n stores E^{n+1}: store 1 = the replay store
// (slot n % K), 2 = the final segments' store (slot n - a_tail), 0 = nowhere.
template <typename T>
__device__ inline T* batch_es(const Batch3Design<T>& D, int store, int n, int K) {
  if (store == 2) return D.etail + int64_t(n - D.a_tail) * D.estore_stride;
  if (store == 1) return D.estore + int64_t(n % K) * D.estore_stride;
  return nullptr;
}

struct Batch3Launch {
  const void* designs;  // device array of Batch3Design<T>
  int n_designs;
  int total_tiles;
  int K;                // checkpoint interval
  int store;            // forward: also store E (1: replay store, 2: final segments' store,
                        // monitors accumulated as without a store)
  const int* step;      // device step counter
  int layout;           // buffer layout (batch_in_slot)
  // two-step launches: one work-group per entry, (design, tile of that
  // design) or (-1, -1) for an idle slot; `blocks` entries (a multiple of 8)
  const int2* block_tile;
  int blocks;
  int nt;               // two-step launches: non-temporal output stores
};

// Block -> tile map of a two-step batched launch (see batch_block_map).
void batch_block_map(const int* tiles, const int64_t* tile_cost, int n_designs,
                     std::vector<int2>& out);

template <typename T>
int prepare_batch3(int ty, bool adjoint, int max_w, int* blocks_per_cu);
template <typename T>
int launch_batch3(const Batch3Launch& L, bool adjoint, int ty, int max_w, int grid,
                  hipStream_t s);
template <typename T>
int launch_batch_monitor(const void* designs, int n_designs, int64_t max_cells, int K,
                         const int* step, bool adjoint, int koff, int layout, hipStream_t s);
int launch_counter_add(int* c, int delta, hipStream_t s);
int launch_counter_set(int* c, int v, hipStream_t s);

struct Fused3Tiles {
  int ty;  // rows per work-group (4 or 8)
  int zc;  // z planes per work-group
};

// Tile configuration of the fused kernels (kernel 0 forward, 1 adjoint),
// set by dnx_fdtd3d_set_tiles.
Fused3Tiles fused3_tiles(int dtype, int kernel);

// ---- Two-step (temporally blocked) kernels --------------------------------
// One launch advances two time steps: state (n) -> (n+2) for the forward,
// lam^{m+1} -> lam^{m-1} for the adjoint, reading and writing the state
// once.  One soft source is fused.  DFT monitors (at most FUSED2_MAX_MON)
// go through compact side buffers so that no kernel waits on them: the
// forward stores both sub-steps' monitor-box values ([2][ncomp][cells]),
// accumulated afterwards by launch_monitor_pair; the adjoint reads the
// intermediate sample's adjoint source precomputed by launch_monitor_d
// ([ncomp][cells]) and adds it in-kernel (the next sub-step needs it); the
// last sample's term is added by launch_monitor_adj after the launch.

constexpr int FUSED2_MAX_MON = 1;

struct FusedMon {
  Box box;
  int mask, ncomp;
};

template <typename T>
struct Fused2Args {
  Fused3Args<T> a;  // geometry, tiles, in/out, ce, source (a.src_wave = wave base);
                    // forward: a.es = E store of sub-step 1; adjoint: a.em = E^m, a.S
  int n;            // forward: first step index n; adjoint: m (steps m, m-1)
  T* es2;           // forward: E store of sub-step 2 (or null)
  const T* em2;     // adjoint: E^{m-1}
  int n_mon;
  FusedMon mon[FUSED2_MAX_MON];
  T* side[FUSED2_MAX_MON];  // compact monitor buffers (see above)
  // single-grid forward with the DFT of monitor 0 done in the tile (an
  // x-normal monitor plane; mnf = 0: not used): phase table, accumulators
  const double* mph;
  double* macc;
  int mnf;
};

// Two-step tile configuration (kernel 2 forward, 3 adjoint); float32 only.
template <typename T>
int launch_fused2_fwd(const Fused2Args<T>& a, int ty, hipStream_t s);
template <typename T>
int launch_fused2_adj(const Fused2Args<T>& a, int ty, hipStream_t s);
// Whether the forward pair kernel can run monitor m's DFT in its tiles (an
// x-normal plane whose per-tile samples fit the staging space).
bool monitor_in_tile(const dnx_monitor& m, int ty, int zc);
// DFT monitor m of the forward pair just launched with f (steps n, n+1):
// the intermediate sample recomputed from f.a.in, the second read from
// f.a.out (the single-grid forward kernel stages no monitor samples).
template <typename T>
int launch_monitor_pair_fwd(const Fused2Args<T>& f, const dnx_monitor& m, int n, hipStream_t s);

// Batched two-step launches over a design table whose tiles follow kernel 2
// (forward) or 3 (adjoint); float32 only.  prepare_batch3_pair sets the
// kernel attributes (never during capture) and reports work-groups per CU.
template <typename T>
int prepare_batch3_pair(int ty, bool adjoint, int max_w, int* blocks_per_cu);
template <typename T>
int launch_batch3_pair(const Batch3Launch& L, bool adjoint, int ty, int max_w, int grid,
                       hipStream_t s);
// Monitors of the batched two-step path: after a forward pair at counter n,
// accumulate the two staged samples (phase rows n, n+1); before an adjoint
// pair at counter m, precompute the adjoint source of sample m-1 (phase row
// m-1) into each design's m_side.
template <typename T>
int launch_batch_monitor_pair(const void* designs, int n_designs, int64_t max_cells,
                              const int* step, bool adjoint, hipStream_t s);


template <typename T>
int launch_fused3_fwd(const Fused3Args<T>& a, int ty, hipStream_t s);

template <typename T>
int launch_fused3_adj(const Fused3Args<T>& a, int ty, hipStream_t s);

}  // namespace dnx

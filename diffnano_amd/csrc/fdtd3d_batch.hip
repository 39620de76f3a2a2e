// Batched 3D FDTD over many designs with different grids: one persistent
// launch per time step covers every design's tiles; the step index lives in
// a device counter, so a segment of K steps is captured once into a HIP
// graph and replayed for every checkpoint segment of the forward, the
// replay and the adjoint.  Semantics per design are exactly those of
// dnx_fdtd3d_run / dnx_fdtd3d_adjoint_run (same tile bodies).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fdtd3d_fused.h"

// No multiply-add contraction (same rounding in every FDTD kernel variant).
#pragma clang fp contract(off)

struct dnx_batch3d_plan {
  int dtype = 0, n = 0, T = 0, K = 0;
  int ty_f = 0, ty_a = 0, max_w = 0, grid_f = 0, grid_a = 0;
  // two-step path (float32): forward / adjoint in step pairs, buffer layout 1
  bool fpairs = false, apairs = false;
  int ty_f2 = 0, ty_a2 = 0, grid_f2 = 0, grid_a2 = 0;
  int lam1_slot = 1;  // adjoint pool holding lam^1 after the backward sweep
  int64_t max_mon_cells = 0, sum6 = 0, sum3 = 0, sumN = 0;
  std::vector<unsigned char> host_tables;  // staging of the four descriptor tables
  void* d_fwd = nullptr;
  void* d_adj = nullptr;
  void* d_fwd2 = nullptr;
  void* d_adj2 = nullptr;
  void* side_mem = nullptr;  // pools.side when the two-step path stages monitors
  int* d_ctr = nullptr;
  dnx::Batch3Launch Lf{}, La{}, Lf2{}, La2{};
  dnx_batch3d_pools pools{};
  std::vector<dnx_design3d> designs;
  std::vector<int64_t> off;  // sum_{e<d} N_e
  std::vector<int64_t> side_off;  // element offset of design d's monitor buffer
  bool use_graph = false;
  hipGraphExec_t g_fwd = nullptr, g_rep = nullptr, g_adj = nullptr, g_ftail = nullptr;
  std::vector<int2> map_f2, map_a2;  // block maps of the two-step launches (host copies)
  int2* d_map_f2 = nullptr;
  int2* d_map_a2 = nullptr;
};

namespace dnx {

// Tile shape of a batched launch kind: kernel 0 / 1 (one-step forward /
// adjoint) or 2 / 3 (two-step).  The batched two-step adjoint is built for
// 2-row tiles only (with 4 rows its per-design arguments make it spill).
static Fused3Tiles batch_tiles(int dtype, int kernel) {
  Fused3Tiles t = fused3_tiles(dtype, kernel);
  if (t.ty <= 0) t.ty = 2;  // two-step kernel switched off: table unused
  if (kernel == 3 && t.ty > 2) t.ty = 2;
  return t;
}

// Work-groups of a two-step launch kind over these designs (block map rows
// x 8), for the workspace query and the plan.
static void block_map_for(const dnx_design3d* ds, int n, int dtype, int kernel,
                          std::vector<int2>& out) {
  const Fused3Tiles t = batch_tiles(dtype, kernel);
  std::vector<int> tiles(n);
  std::vector<int64_t> cost(n);
  for (int d = 0; d < n; ++d) {
    const dnx_grid3d& g = ds[d].grid;
    tiles[d] = int(cdiv(g.ny, t.ty)) * int(cdiv(g.nz, t.zc));
    // lanes x rows x planes (plus the two restart planes of a z-chunk)
    cost[d] = int64_t(cdiv(g.nx, 64) * 64) * t.ty * (std::min(t.zc, int(g.nz)) + 2);
  }
  batch_block_map(tiles.data(), cost.data(), n, out);
}

// Descriptor table of one launch kind, whose tile shape sets the tiles per
// design.
template <typename T>
static void fill_table(dnx_batch3d_plan* p, const dnx_design3d* ds, int kernel,
                       Batch3Design<T>* out, int* total_tiles) {
  const bool adjoint = kernel == 1 || kernel == 3;
  const Fused3Tiles t = batch_tiles(p->dtype, kernel);
  int start = 0;
  for (int d = 0; d < p->n; ++d) {
    const dnx_design3d& de = ds[d];
    Batch3Design<T> b{};
    const dnx_grid3d& g = de.grid;
    b.a.nz = g.nz;
    b.a.ny = g.ny;
    b.a.nx = g.nx;
    b.a.ty_tiles = int(cdiv(g.ny, t.ty));
    b.a.zc = t.zc;
    b.a.nblocks = b.a.ty_tiles * int(cdiv(g.nz, t.zc));
    b.a.P = int64_t(g.ny) * g.nx;
    b.a.N = b.a.P * g.nz;
    b.a.kh = T(g.kh);
    b.a.ce = static_cast<const T*>(g.ce);
    b.a.by = static_cast<const T*>(g.by);
    b.a.ay = static_cast<const T*>(g.ay);
    b.a.bx = static_cast<const T*>(g.bx);
    b.a.ax = static_cast<const T*>(g.ax);
    b.a.S = static_cast<T*>(p->pools.S) + p->off[d];
    b.a.src_comp = de.src.comp;
    if (de.src.comp >= 0) {
      b.a.src_box = make_box(de.src.box);
      b.a.src_prof = static_cast<const T*>(de.src.profile);
      b.wave = static_cast<const T*>(de.src.wave);
    }
    void* const* bufs = adjoint ? p->pools.lam : p->pools.state;
    b.buf[0] = static_cast<T*>(bufs[0]) + 6 * p->off[d];
    b.buf[1] = static_cast<T*>(bufs[1]) + 6 * p->off[d];
    b.ckpt = static_cast<T*>(p->pools.ckpt) + 6 * p->off[d];
    b.estore = p->pools.estore ? static_cast<T*>(p->pools.estore) + 3 * p->off[d] : nullptr;
    b.etail = p->pools.etail ? static_cast<T*>(p->pools.etail) + 3 * p->off[d] : nullptr;
    b.a_tail = p->pools.etail ? p->pools.a_tail : p->T;
    b.ckpt_stride = p->sum6;
    b.estore_stride = p->sum3;
    b.tile_start = start;
    start += b.a.nblocks;
    b.m_nf = de.mon.n_freq;
    if (de.mon.n_freq > 0) {
      b.m_mask = de.mon.comp_mask;
      b.mbox = make_box(de.mon.box);
      b.m_phase = de.mon.phase;
      b.m_acc = de.mon.acc;
      b.m_grad = de.mon.grad;
    }
    b.m_side = p->side_mem ? static_cast<T*>(p->side_mem) + p->side_off[d] : nullptr;
    out[d] = b;
  }
  *total_tiles = start;
}

// Element offset of design d's monitor staging in pools.side (two samples
// of its monitor-box components); d = n gives the total.
static int64_t side_offset(const dnx_design3d* ds, int d) {
  int64_t acc = 0;
  for (int e = 0; e < d; ++e) {
    const dnx_monitor& m = ds[e].mon;
    if (m.n_freq > 0) {
      const int mk = m.comp_mask;
      acc += 2 * int64_t((mk & 1) + ((mk >> 1) & 1) + ((mk >> 2) & 1)) * box_cells(m.box);
    }
  }
  return acc;
}

template <typename T>
static int batch_setup(dnx_batch3d_plan* p, const dnx_design3d* ds, void* ws, hipStream_t s) {
  // Two-step path: float32, not switched off, every design addressable by
  // the two-step kernels' 32-bit buffer offsets (6 N 4 bytes < 2^31), and
  // tile heights the batched two-step kernels are built for.
  if (std::is_same<T, float>::value) {
    bool ok = option(DNX_OPT_FDTD3D_ONE_STEP) == 0;
    for (int d = 0; d < p->n && ok; ++d) {
      const dnx_grid3d& g = ds[d].grid;
      ok = int64_t(g.nz) * g.ny * g.nx * 6 * int64_t(sizeof(T)) < (int64_t(1) << 31);
      // the tiles read the monitor adjoint terms through one descriptor
      // whose out-of-box offsets start at 2^29 bytes
      if (ok && ds[d].mon.n_freq > 0) {
        const int mk = ds[d].mon.comp_mask;
        const int nc = (mk & 1) + ((mk >> 1) & 1) + ((mk >> 2) & 1);
        ok = make_box(ds[d].mon.box).cells() * nc * int64_t(sizeof(T)) < (int64_t(1) << 29);
      }
    }
    const int tf2 = fused3_tiles(p->dtype, 2).ty, ta2 = fused3_tiles(p->dtype, 3).ty;
    p->fpairs = ok && (tf2 == 2 || tf2 == 4);
    p->apairs = ok && ta2 > 0 && option(DNX_OPT_FDTD3D_ADJ2) != 0;
  }
  if ((p->fpairs || p->apairs) && p->max_mon_cells > 0) {
    if (!p->pools.side) {
      p->fpairs = p->apairs = false;  // no staging buffer: one-step kernels
    } else {
      for (int d = 0; d < p->n; ++d) p->side_off.push_back(side_offset(ds, d));
      p->side_mem = p->pools.side;
    }
  }
  const size_t tb = sizeof(Batch3Design<T>) * p->n;
  const size_t tbr = ((tb + 255) / 256) * 256;
  p->host_tables.assign(4 * tb, 0);
  Batch3Design<T>* tab[4];
  int tot[4] = {0, 0, 0, 0};
  auto* base = static_cast<unsigned char*>(ws);
  void** dst[4] = {&p->d_fwd, &p->d_adj, &p->d_fwd2, &p->d_adj2};
  for (int k = 0; k < 4; ++k) {
    tab[k] = reinterpret_cast<Batch3Design<T>*>(p->host_tables.data() + k * tb);
    fill_table<T>(p, ds, k, tab[k], &tot[k]);
    *dst[k] = base + k * tbr;
    DNX_CHECK_HIP(hipMemcpyAsync(*dst[k], tab[k], tb, hipMemcpyHostToDevice, s));
  }
  p->d_ctr = reinterpret_cast<int*>(base + 4 * tbr);
  // block maps of the two-step launches, after the counter
  unsigned char* maps = base + 4 * tbr + 256;
  for (int k = 2; k < 4; ++k) {
    std::vector<int2>& m = k == 2 ? p->map_f2 : p->map_a2;
    block_map_for(ds, p->n, p->dtype, k, m);
    void* dm = maps;
    maps += ((m.size() * sizeof(int2) + 255) / 256) * 256;
    DNX_CHECK_HIP(hipMemcpyAsync(dm, m.data(), m.size() * sizeof(int2), hipMemcpyHostToDevice,
                                 s));
    (k == 2 ? p->d_map_f2 : p->d_map_a2) = static_cast<int2*>(dm);
  }
  int dev = 0, cus = 256;
  DNX_CHECK_HIP(hipGetDevice(&dev));
  DNX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  p->ty_f = fused3_tiles(p->dtype, 0).ty;
  p->ty_a = fused3_tiles(p->dtype, 1).ty;
  // Kernel attributes are set here, never while a stream is being captured;
  // the persistent grid is as many work-groups as fit on the device at once.
  int bf = 1, ba = 1;
  int rc = prepare_batch3<T>(p->ty_f, false, p->max_w, &bf);
  if (rc == DNX_OK) rc = prepare_batch3<T>(p->ty_a, true, p->max_w, &ba);
  if (rc) return rc;
  p->grid_f = std::min(tot[0], bf * cus);
  p->grid_a = std::min(tot[1], ba * cus);
  const int fl = p->fpairs ? 1 : 0, al = p->apairs ? 1 : 0;
  p->Lf = Batch3Launch{p->d_fwd, p->n, tot[0], p->K, 0, p->d_ctr, fl, nullptr, 0};
  p->La = Batch3Launch{p->d_adj, p->n, tot[1], p->K, 0, p->d_ctr, al, nullptr, 0};
  if constexpr (std::is_same<T, float>::value) {
    // non-temporal stores when one state of the whole batch exceeds half of
    // the 256 MB last-level cache (the single-grid rule)
    const int nt = p->sum6 * int64_t(sizeof(T)) > (int64_t(128) << 20) ? 1 : 0;
    if (p->fpairs) {
      p->ty_f2 = batch_tiles(p->dtype, 2).ty;
      int b2 = 1;
      if ((rc = prepare_batch3_pair<T>(p->ty_f2, false, p->max_w, &b2))) return rc;
      p->grid_f2 = int(p->map_f2.size());
      p->Lf2 = Batch3Launch{p->d_fwd2, p->n, tot[2], p->K, 0, p->d_ctr, 1, p->d_map_f2,
                            int(p->map_f2.size()), nt};
    }
    if (p->apairs) {
      p->ty_a2 = batch_tiles(p->dtype, 3).ty;
      int b2 = 1;
      if ((rc = prepare_batch3_pair<T>(p->ty_a2, true, p->max_w, &b2))) return rc;
      p->grid_a2 = int(p->map_a2.size());
      p->La2 = Batch3Launch{p->d_adj2, p->n, tot[3], p->K, 0, p->d_ctr, 1, p->d_map_a2,
                            int(p->map_a2.size()), nt};
    }
  }
  return DNX_OK;
}

// ---- step sequences (each reads the device counter) ----------------------

// Forward: step pairs on the two-step path, a single step when one
// remains (only at the end of a sequence, so every pair starts at an even
// step).  store: 0 none, 1 the replay (E into the replay store, no
// monitors), 2 a recording forward over the final segments (E into their
// store, monitors accumulated).
template <typename T>
static int seq_forward(dnx_batch3d_plan* p, int steps, int store, hipStream_t s,
                       bool timed = true) {
  static const char* names2[3] = {"batch3d_fwd2", "batch3d_fwd2_store", "batch3d_fwd2_etail"};
  static const char* names1[3] = {"batch3d_fwd", "batch3d_fwd_store", "batch3d_fwd_etail"};
  const bool mon = store != 1;
  int k = 0;
  int rc;
  if constexpr (std::is_same<T, float>::value) {
    if (p->fpairs) {
      Batch3Launch L = p->Lf2;
      L.store = store;
      for (; k + 1 < steps; k += 2) {
        TimedLaunch tl = timed ? timed_begin(names2[store], s) : TimedLaunch{};
        rc = launch_batch3_pair<T>(L, false, p->ty_f2, p->max_w, p->grid_f2, s);
        timed_end(names2[store], s, tl);
        if (rc) return rc;
        if (mon && (rc = launch_batch_monitor_pair<T>(p->d_fwd2, p->n, p->max_mon_cells,
                                                       p->d_ctr, false, s)))
          return rc;
        if ((rc = launch_counter_add(p->d_ctr, 2, s))) return rc;
      }
    }
  }
  Batch3Launch L = p->Lf;
  L.store = store;
  for (; k < steps; ++k) {
    TimedLaunch tl = timed ? timed_begin(names1[store], s) : TimedLaunch{};
    rc = launch_batch3<T>(L, false, p->ty_f, p->max_w, p->grid_f, s);
    timed_end(names1[store], s, tl);
    if (rc) return rc;
    if (mon && (rc = launch_batch_monitor<T>(p->d_fwd, p->n, p->max_mon_cells, p->K,
                                              p->d_ctr, false, 0, L.layout, s)))
      return rc;
    if ((rc = launch_counter_add(p->d_ctr, 1, s))) return rc;
  }
  return DNX_OK;
}

// Adjoint from the counter m downward: pairs (m, m-1) on the two-step path
// (adjoint source of sample m-1 precomputed, that of m-2 applied after the
// launch), a single step when one remains.
template <typename T>
static int seq_adjoint(dnx_batch3d_plan* p, int steps, hipStream_t s, bool timed = true) {
  int k = 0;
  int rc;
  if constexpr (std::is_same<T, float>::value) {
    if (p->apairs) {
      for (; k + 1 < steps; k += 2) {
        if ((rc = launch_batch_monitor_pair<T>(p->d_adj2, p->n, p->max_mon_cells, p->d_ctr, true,
                                               s)))
          return rc;
        TimedLaunch tl = timed ? timed_begin("batch3d_adj2", s) : TimedLaunch{};
        rc = launch_batch3_pair<T>(p->La2, true, p->ty_a2, p->max_w, p->grid_a2, s);
        timed_end("batch3d_adj2", s, tl);
        if (rc) return rc;
        if ((rc = launch_batch_monitor<T>(p->d_adj2, p->n, p->max_mon_cells, p->K, p->d_ctr,
                                          true, -1, 1, s)))
          return rc;
        if ((rc = launch_counter_add(p->d_ctr, -2, s))) return rc;
      }
    }
  }
  for (; k < steps; ++k) {
    TimedLaunch tl = timed ? timed_begin("batch3d_adj", s) : TimedLaunch{};
    rc = launch_batch3<T>(p->La, true, p->ty_a, p->max_w, p->grid_a, s);
    timed_end("batch3d_adj", s, tl);
    if (rc) return rc;
    if ((rc = launch_batch_monitor<T>(p->d_adj, p->n, p->max_mon_cells, p->K, p->d_ctr, true, 0,
                                      p->La.layout, s)))
      return rc;
    if ((rc = launch_counter_add(p->d_ctr, -1, s))) return rc;
  }
  return DNX_OK;
}

template <typename T>
static int capture(dnx_batch3d_plan* p, int which, int steps, hipGraphExec_t* out) {
  hipStream_t cs = nullptr;
  DNX_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  int rc = DNX_OK;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = which == 2 ? seq_adjoint<T>(p, steps, cs, false)
                    : seq_forward<T>(p, steps, which == 3 ? 2 : which, cs, false);
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (rc == DNX_OK && e2 != hipSuccess) {
      set_error("batch3d: graph capture failed: %s", hipGetErrorString(e2));
      rc = DNX_EHIP;
    }
  } else {
    set_error("batch3d: begin capture failed: %s", hipGetErrorString(e));
    rc = DNX_EHIP;
  }
  if (rc == DNX_OK) {
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_error("batch3d: graph instantiate failed: %s", hipGetErrorString(e));
      rc = DNX_EHIP;
      *out = nullptr;
    }
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  return rc;
}

// Run `steps` steps of a sequence (which: 0 forward, 1 replay, 2 adjoint,
// 3 recording forward over the final segments): from the graph when the
// count matches a captured full segment, else by direct launches.
template <typename T>
static int run_seq(dnx_batch3d_plan* p, int which, int steps, hipStream_t s) {
  if (steps <= 0) return DNX_OK;
  hipGraphExec_t* slot = which == 0 ? &p->g_fwd : which == 1 ? &p->g_rep
                         : which == 2 ? &p->g_adj : &p->g_ftail;
  const int full = which == 1 ? p->K - 1 : p->K;
  if (p->use_graph && steps == full) {
    if (*slot == nullptr) {
      int rc = capture<T>(p, which, steps, slot);
      if (rc) {
        // Capture failure: fall back to eager launches for this plan.
        p->use_graph = false;
      }
    }
    if (*slot) {
      DNX_CHECK_HIP(hipGraphLaunch(*slot, s));
      return DNX_OK;
    }
  }
  return which == 2 ? seq_adjoint<T>(p, steps, s)
                    : seq_forward<T>(p, steps, which == 3 ? 2 : which, s);
}

template <typename T>
static int batch_forward(dnx_batch3d_plan* p, bool record, hipStream_t s) {
  int rc = launch_counter_set(p->d_ctr, 0, s);
  if (rc) return rc;
  for (int k0 = 0; k0 < p->T; k0 += p->K) {
    const int len = std::min(p->K, p->T - k0);
    if (record) {
      DNX_CHECK_HIP(hipMemcpyAsync(static_cast<T*>(p->pools.ckpt) + int64_t(k0 / p->K) * p->sum6,
                                   p->pools.state[batch_in_slot(p->fpairs, k0)],
                                   sizeof(T) * p->sum6, hipMemcpyDeviceToDevice, s));
    }
    const bool tail = record && p->pools.etail && k0 >= p->pools.a_tail;
    if ((rc = run_seq<T>(p, tail ? 3 : 0, len, s))) return rc;
  }
  return DNX_OK;
}

template <typename T>
static int batch_backward(dnx_batch3d_plan* p, hipStream_t s) {
  const int T_ = p->T;
  // lam^T: zero both adjoint pools, then per design the monitor terms of
  // sample T and S = E^T . lam^T (E^T in state buffer T & 1).
  int rc = launch_zero<T>(static_cast<T*>(p->pools.lam[0]), p->sum6, s);
  if (rc) return rc;
  if ((rc = launch_zero<T>(static_cast<T*>(p->pools.lam[1]), p->sum6, s))) return rc;
  for (int d = 0; d < p->n; ++d) {
    const dnx_design3d& de = p->designs[d];
    T* lam = static_cast<T*>(p->pools.lam[batch_in_slot(p->apairs, T_)]) + 6 * p->off[d];
    const T* eT =
        static_cast<const T*>(p->pools.state[batch_in_slot(p->fpairs, T_)]) + 6 * p->off[d];
    T* S = static_cast<T*>(p->pools.S) + p->off[d];
    rc = dnx_fdtd3d_adjoint_init(&de.grid, lam, eT, de.mon.n_freq > 0 ? &de.mon : nullptr,
                                 de.mon.n_freq > 0 ? 1 : 0, T_, S, s);
    if (rc) return rc;
  }
  const int nseg = (T_ + p->K - 1) / p->K;
  for (int k = nseg - 1; k >= 0; --k) {
    const int a = k * p->K;
    const int len = std::min(p->K, T_ - a);
    // segments after a_tail: E of every step was kept by the forward
    if (len > 1 && !(p->pools.etail && a >= p->pools.a_tail)) {
      DNX_CHECK_HIP(hipMemcpyAsync(p->pools.state[batch_in_slot(p->fpairs, a)],
                                   static_cast<const T*>(p->pools.ckpt) + int64_t(k) * p->sum6,
                                   sizeof(T) * p->sum6, hipMemcpyDeviceToDevice, s));
      if ((rc = launch_counter_set(p->d_ctr, a, s))) return rc;
      if ((rc = run_seq<T>(p, 1, len - 1, s))) return rc;
    }
    if ((rc = launch_counter_set(p->d_ctr, a + len - 1, s))) return rc;
    const int steps = k == 0 ? len - 1 : len;  // m runs down to max(a, 1)
    if ((rc = run_seq<T>(p, 2, steps, s))) return rc;
    if (k == 0) {
      // lam^1 sits where the last launch wrote it: layout 1 puts a final
      // single step (m = 1, odd) in the other pool than a final pair
      p->lam1_slot = (p->apairs && steps > 0 && (steps & 1)) ? 0 : 1;
    }
  }
  // step-0 source term with lam^1 (in adjoint buffer 1)
  for (int d = 0; d < p->n; ++d) {
    const dnx_design3d& de = p->designs[d];
    if (de.src.comp < 0) continue;
    T* lam1 = static_cast<T*>(p->pools.lam[p->lam1_slot]) + 6 * p->off[d];
    rc = launch_source_adj<T>(static_cast<T*>(p->pools.S) + p->off[d],
                              lam1 + int64_t(de.src.comp) * (int64_t(de.grid.nz) * de.grid.ny *
                                                             de.grid.nx),
                              de.grid.ny, de.grid.nx, de.src, 0, s);
    if (rc) return rc;
  }
  return DNX_OK;
}

}  // namespace dnx

using namespace dnx;

// Grid shapes the batched queries can size (the plan checks the rest):
// positive extents, nx within one work-group, at most 2^31 cells per design
// and 2^28 tiles of the smallest tile shape in the batch.
static bool batch_shapes_ok(const dnx_design3d* ds, int n) {
  int64_t tiles = 0;
  for (int d = 0; d < n; ++d) {
    const dnx_grid3d& g = ds[d].grid;
    if (g.nz <= 0 || g.ny <= 0 || g.nx <= 0 || g.nx > FUSED_MAX_W) return false;
    if (int64_t(g.nz) * g.ny * g.nx > (int64_t(1) << 31)) return false;
    tiles += int64_t(g.ny) * cdiv(g.nz, 8);
    if (tiles > (int64_t(1) << 28)) return false;
  }
  return true;
}

extern "C" int64_t dnx_batch3d_side_elems(const dnx_design3d* designs, int32_t n) {
  if (!designs || n <= 0) return 0;
  if (!batch_shapes_ok(designs, n)) return -1;
  for (int d = 0; d < n; ++d) {
    const dnx_monitor& m = designs[d].mon;
    if (m.n_freq > 0 && !box_ok(m.box, designs[d].grid.nz, designs[d].grid.ny,
                                designs[d].grid.nx))
      return -1;
  }
  return dnx::side_offset(designs, n);
}

extern "C" int64_t dnx_batch3d_block_map(const dnx_design3d* designs, int32_t n_designs,
                                         int32_t kernel, int32_t* out, int64_t capacity) {
  if (!designs || n_designs <= 0 || (kernel != 2 && kernel != 3) || capacity < 0 ||
      !batch_shapes_ok(designs, n_designs))
    return -1;
  std::vector<int2> m;
  block_map_for(designs, n_designs, designs[0].grid.dtype, kernel, m);
  const int64_t n = int64_t(m.size());
  if (out) {
    if (capacity < n) return -1;
    for (int64_t i = 0; i < n; ++i) {
      out[2 * i] = m[i].x;
      out[2 * i + 1] = m[i].y;
    }
  }
  return n;
}

extern "C" int64_t dnx_batch3d_workspace_bytes(const dnx_design3d* designs, int32_t n_designs) {
  if (!designs || n_designs <= 0 || !batch_shapes_ok(designs, n_designs)) return -1;
  const size_t tb = sizeof(Batch3Design<double>) * size_t(n_designs);
  int64_t bytes = int64_t(4 * ((tb + 255) / 256) * 256 + 256);  // four tables + counter
  const int dtype = designs[0].grid.dtype;
  for (int k = 2; k < 4; ++k) {  // the two-step block maps at the current tile shapes
    std::vector<int2> m;
    block_map_for(designs, n_designs, dtype, k, m);
    bytes += int64_t(((m.size() * sizeof(int2) + 255) / 256) * 256);
  }
  return bytes;
}

extern "C" int32_t dnx_batch3d_create(const dnx_design3d* designs, int32_t n, int32_t n_steps,
                                      int32_t ckpt_every, const dnx_batch3d_pools* pools,
                                      void* workspace, int64_t ws_bytes, int32_t use_graph,
                                      void* stream, dnx_batch3d_plan** out) {
  DNX_REQUIRE(out != nullptr, "batch3d_create: null output");
  *out = nullptr;
  DNX_REQUIRE(designs && n > 0 && pools && workspace, "batch3d_create: null argument");
  DNX_REQUIRE(n_steps >= 1, "batch3d_create: n_steps must be >= 1");
  DNX_REQUIRE(ckpt_every >= 2 && ckpt_every % 2 == 0,
              "batch3d_create: ckpt_every must be even and >= 2 (got %d)", ckpt_every);
  const int64_t ws_need = dnx_batch3d_workspace_bytes(designs, n);
  DNX_REQUIRE(ws_need >= 0,
              "batch3d_create: a design's grid is empty, wider than %d or too large",
              FUSED_MAX_W);
  DNX_REQUIRE(ws_bytes >= ws_need,
              "batch3d_create: workspace too small (size it with dnx_batch3d_workspace_bytes "
              "after setting the tile shapes)");
  DNX_REQUIRE(pools->state[0] && pools->state[1] && pools->lam[0] && pools->lam[1] &&
                  pools->ckpt && pools->S && (ckpt_every < 2 || pools->estore),
              "batch3d_create: null pool");
  DNX_REQUIRE(!pools->etail || (pools->a_tail >= 0 && pools->a_tail < n_steps &&
                                pools->a_tail % ckpt_every == 0),
              "batch3d_create: a_tail must be a multiple of ckpt_every in [0, n_steps) (got %d)",
              pools->a_tail);
  const int dtype = designs[0].grid.dtype;
  auto* p = new dnx_batch3d_plan();
  p->dtype = dtype;
  p->n = n;
  p->T = n_steps;
  p->K = ckpt_every;
  p->pools = *pools;
  p->use_graph = use_graph != 0;
  p->designs.assign(designs, designs + n);
  int64_t acc = 0;
  for (int d = 0; d < n; ++d) {
    const dnx_design3d& de = designs[d];
    const dnx_grid3d& g = de.grid;
    bool ok = g.dtype == dtype && g.nz > 0 && g.ny > 0 && g.nx > 0 && g.nx <= FUSED_MAX_W &&
              g.ce && g.by && g.ay && g.bx && g.ax;
    if (ok && de.src.comp >= 0)
      ok = de.src.comp <= 2 && box_ok(de.src.box, g.nz, g.ny, g.nx) && de.src.profile &&
           de.src.wave;
    if (ok && de.mon.n_freq > 0)
      ok = box_ok(de.mon.box, g.nz, g.ny, g.nx) && de.mon.comp_mask > 0 &&
           (de.mon.comp_mask & ~7) == 0 && de.mon.phase && de.mon.acc && de.mon.grad;
    if (!ok) {
      delete p;
      set_error("batch3d_create: design %d invalid (dtype, shape, nx <= %d, source or monitor)",
                d, FUSED_MAX_W);
      return DNX_EINVAL;
    }
    p->off.push_back(acc);
    acc += int64_t(g.nz) * g.ny * g.nx;
    p->max_w = std::max(p->max_w, int(g.nx));
    if (de.mon.n_freq > 0) p->max_mon_cells = std::max(p->max_mon_cells, box_cells(de.mon.box));
  }
  p->sumN = acc;
  p->sum6 = 6 * acc;
  p->sum3 = 3 * acc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = dtype == DNX_F32 ? batch_setup<float>(p, designs, workspace, s)
                            : batch_setup<double>(p, designs, workspace, s);
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return DNX_OK;
}

extern "C" int32_t dnx_batch3d_forward(dnx_batch3d_plan* p, int32_t record_ckpt, void* stream) {
  DNX_REQUIRE(p != nullptr, "batch3d_forward: null plan");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == DNX_F32 ? batch_forward<float>(p, record_ckpt != 0, s)
                             : batch_forward<double>(p, record_ckpt != 0, s);
}

extern "C" int32_t dnx_batch3d_backward(dnx_batch3d_plan* p, void* stream) {
  DNX_REQUIRE(p != nullptr, "batch3d_backward: null plan");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == DNX_F32 ? batch_backward<float>(p, s) : batch_backward<double>(p, s);
}

extern "C" int32_t dnx_batch3d_uses_graph(const dnx_batch3d_plan* p) {
  return (p && p->use_graph) ? 1 : 0;
}

extern "C" int32_t dnx_batch3d_final_slot(const dnx_batch3d_plan* p) {
  return p ? batch_in_slot(p->fpairs ? 1 : 0, p->T) : 0;
}

extern "C" int32_t dnx_batch3d_uses_pairs(const dnx_batch3d_plan* p) {
  return p ? (p->fpairs ? 1 : 0) | (p->apairs ? 2 : 0) : 0;
}

extern "C" int32_t dnx_batch3d_destroy(dnx_batch3d_plan* p) {
  if (!p) return DNX_OK;
  if (p->g_fwd) (void)hipGraphExecDestroy(p->g_fwd);
  if (p->g_rep) (void)hipGraphExecDestroy(p->g_rep);
  if (p->g_adj) (void)hipGraphExecDestroy(p->g_adj);
  if (p->g_ftail) (void)hipGraphExecDestroy(p->g_ftail);
  delete p;
  return DNX_OK;
}

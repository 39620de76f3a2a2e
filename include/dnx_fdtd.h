/*
 * dnx_fdtd.h -- C ABI of the MI355X (gfx950) Yee-FDTD library libdnx_fdtd.so.
 *
 * The library replaces the arithmetic of DiffNano's FDTD solvers
 * (diffnano/solvers/fdtd3d.py, fdtd2d.py).  Those are pure PyTorch, so the
 * reference has no FFI of its own; each entry point below names the
 * reference function whose semantics it implements.  The Python host
 * (diffnano_amd/_lib.py) binds them with ctypes.
 *
 * Conventions
 *  - Arrays are row-major [axis0][axis1][axis2] = [z][y][x] (3D) or [y][x]
 *    (2D), x fastest, exactly the reference's (D, H, W) / (H, W) tensors.
 *  - Every pointer inside a struct or argument is a DEVICE pointer unless
 *    the comment says otherwise; the structs themselves live in host memory.
 *  - dtype: DNX_F32 or DNX_F64 selects the element type of every field,
 *    coefficient, profile, waveform and gradient array of the call.
 *    DFT accumulators are always float64.
 *  - Calls are asynchronous on the HIP stream passed in (NULL = the null
 *    stream); the library never synchronises.
 *  - The library allocates no device memory: every buffer, including the
 *    scratch ("workspace") a call needs, is the caller's, sized by the
 *    *_workspace_bytes queries.  (The HIP runtime itself creates the event
 *    and graph objects of timed launches and of batched plans.)
 *  - Return 0 on success, a negative DNX_E* code on error; the message is in
 *    dnx_last_error() (thread-local).  No C++ exception crosses the ABI.
 */
#ifndef DNX_FDTD_H
#define DNX_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The library is built with hidden visibility; only these symbols export. */
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#define DNX_ABI_VERSION 5

enum { DNX_F32 = 0, DNX_F64 = 1 };
enum {
  DNX_OK = 0,
  DNX_EINVAL = -1,  /* bad argument (shape, dtype, box, null pointer) */
  DNX_EHIP = -2,    /* HIP runtime error (launch, copy) */
  DNX_ERANGE = -3   /* size beyond a documented limit */
};

/* Axis-aligned box [z0,z1) x [y0,y1) x [x0,x1) (2D: z0=0, z1=1). */
typedef struct dnx_box {
  int32_t z0, z1, y0, y1, x0, x1;
} dnx_box;

/* 3D grid and coefficients.  Profiles: b and a = 1 + c of the reference's
 * memoryless CPML (fdtd3d.py:75-85); (by, ay) vary along axis 1 (length ny),
 * (bx, ax) along axis 2 (length nx); axis 0 is undamped (fdtd3d.py:402-463). */
typedef struct dnx_grid3d {
  int32_t nz, ny, nx;
  int32_t dtype;       /* DNX_F32 / DNX_F64 */
  double kh;           /* dt / dl */
  const void* ce;      /* [nz*ny*nx]  dt / (eps * dl) */
  const void* by;      /* [ny] */
  const void* ay;      /* [ny] */
  const void* bx;      /* [nx] */
  const void* ax;      /* [nx] */
} dnx_grid3d;

/* 2D grid (fdtd2d.py): axis 0 (ny rows) carries (by, ay), axis 1 (nx
 * columns) carries (bx, ax).  te = 0: TM state [Ez, Hx, Hy]; te = 1: TE
 * state [Hz, Ex, Ey] with the reference's Hz sign (fdtd2d.py:299-301). */
typedef struct dnx_grid2d {
  int32_t ny, nx;
  int32_t dtype;
  int32_t te;
  double kh;
  const void* ce;      /* [ny*nx] dt / (eps * dl) */
  const void* by;
  const void* ay;
  const void* bx;
  const void* ax;
} dnx_grid2d;

/* Soft source: after step n, field[comp] += wave[n] * profile over box
 * (fdtd3d.py:145,502 with a 0/1 mask and comp = Ez; here any component
 * and any amplitude profile).  3D comp: 0 Ex, 1 Ey, 2 Ez; 2D comp must be 0
 * (Ez in TM, Hz in TE). */
typedef struct dnx_source {
  int32_t comp;
  dnx_box box;
  const void* profile; /* [box cells], x fastest */
  const void* wave;    /* [total steps] */
} dnx_source;

/* Running DFT monitor (no reference counterpart; the nearest is the probe
 * series of fdtd3d.py:652-700).  After step n it adds
 *   acc[f][c][cell] += E_c(cell) * (cos(phi), -sin(phi)),  phi = phase[n][f]
 * for every component c in comp_mask (3D bit c = Ex/Ey/Ez; 2D bit 0).
 * phase holds (cos, sin) pairs: phase[(n*n_freq + f)*2 + {0,1}].
 * acc layout: [n_freq][n_comp][box cells][2] (re, im), float64.
 * grad (adjoint only): dL/dRe, dL/dIm in the same layout. */
typedef struct dnx_monitor {
  dnx_box box;
  int32_t comp_mask;
  int32_t n_freq;
  const double* phase;
  double* acc;
  const double* grad;
} dnx_monitor;

int32_t dnx_abi_version(void);
const char* dnx_last_error(void);

/* Process-wide kernel-selection options.  They change which kernels run,
 * never results (every variant is bitwise identical).  Initial values come
 * from the environment variable in brackets, read once at the first query;
 * dnx_set_option overrides them for later calls.  No reference counterpart
 * (the reference has no kernels). */
enum {
  DNX_OPT_FDTD3D_SIMPLE = 0,   /* 1: two-kernel in-place 3D step [DNX_FDTD3D_VARIANT=simple] */
  DNX_OPT_FDTD3D_ONE_STEP = 1, /* 1: one-step 3D kernels only [DNX_FDTD3D_STEPS=1] */
  DNX_OPT_FDTD3D_ADJ2 = 2,     /* 0: two-step 3D adjoint off (default 1) [DNX_FDTD3D_ADJ2] */
  DNX_OPT_FDTD3D_FWD2_PF = 3,  /* 1: two-step forward loads one plane ahead [DNX_FDTD3D_FWD2_PF] */
  DNX_OPT_FDTD3D_ADJ2_OCC = 4, /* waves per SIMD of the two-step adjoint, 1 or 2 (default 2)
                                  [DNX_FDTD3D_ADJ2_OCC] */
  DNX_OPT_GEMM_GENERIC = 5,    /* 1: FNO products on the element-wise kernel [DNX_GEMM_GENERIC] */
  DNX_OPT_GEMM_PW_VALU = 6,    /* 1: FNO 1x1 layers on the vector ALU [DNX_GEMM_PW_VALU] */
  DNX_OPT_GEMM_SMALL = 7,      /* 1: FNO products on the 128x128 tile kernel [DNX_GEMM_SMALL] */
  DNX_OPT_GEMM_BIG_REG = 8,    /* 1: long-K BF16 products stage operands through registers instead
                                  of loading them straight into LDS [DNX_GEMM_BIG_REG] */
  DNX_OPT_FDTD2D_LAUNCH_PER_STEP = 9, /* batched 2D TM: 0 picks, 1 two launches per step, 2 one
                                  resident launch per call where the grid fits
                                  [DNX_FDTD2D_PER_STEP] */
  DNX_OPT_FDTD2D_SEPARATE_SIDE = 10, /* 1: 2D TM source and monitors as separate kernels
                                  instead of inside the step kernels [DNX_FDTD2D_SEPARATE] */
  DNX_OPT_FDTD2D_TB_STEPS = 11,  /* 2D TM steps per launch of the temporally blocked kernels
                                  (0: 8, the most; 1: the one-step kernels) [DNX_FDTD2D_TB] */
  DNX_OPT_FDTD2D_TB_REGION = 12, /* region width of the 2D blocked kernels: 0 picks, 32 or
                                  64 forces it [DNX_FDTD2D_TB_REGION] */
  DNX_OPT_COUNT = 13
};
int32_t dnx_set_option(int32_t id, int32_t value);
int32_t dnx_get_option(int32_t id, int32_t* value);

/* Per-launch timing of the hot kernels with HIP events on the launch
 * stream (no reference counterpart; measurement only).  Names:
 * "fdtd3d_fwd", "fdtd3d_fwd_store" (replay, also stores E), "fdtd3d_adj".
 * dnx_timing_enable(k): 0 off, k >= 1 times one launch in every k of each
 * name (counted from this call), so that the events' own cost stays out of
 * the wall time of long runs.  dnx_timing_read waits for the recorded
 * events, returns the summed duration and the count of timed launches, and
 * clears that name. */
int32_t dnx_timing_enable(int32_t on);
/* Streaming-bandwidth probe on `stream` (no reference counterpart: the
 * measured ceiling the kernels are compared with).  kind 0: copy n floats
 * src -> dst (16-byte lanes, non-temporal; n a multiple of 4, buffers
 * 16-byte aligned); kind 1: the streams of one FDTD step on n cells, src =
 * 7 n floats (six fields and a coefficient), dst = 6 n floats, one dword
 * per lane and array. */
int32_t dnx_stream_probe(int32_t kind, const void* src, void* dst, int64_t n, void* stream);
int32_t dnx_timing_read(const char* name, double* total_ms, int64_t* launches);

/* ce[i] = dt / (eps[i] * dl) computed in float64, stored in dtype.
 * eps is float64.  Replaces the reference's per-step (dt / eps_r)
 * (fdtd3d.py:435; fdtd2d.py:262). */
int32_t dnx_prepare_coef(const double* eps, void* ce, int64_t n, double dt_over_dl,
                         int32_t dtype, void* stream);

/* grad_eps[i] = -S[i] / eps[i] (float64 out); S from the adjoint. */
int32_t dnx_finalize_grad(const void* S, const double* eps, double* grad_eps, int64_t n,
                          int32_t dtype, void* stream);

/* ---- 3D -------------------------------------------------------------- */

/* Advance `state` ([Ex,Ey,Ez,Hx,Hy,Hz], 6*N) by n_steps leapfrog steps
 * starting at absolute step n0: FDTDSolver3D._time_step (fdtd3d.py:378-465)
 * then source injection (fdtd3d.py:502) then monitors, per step.
 * work: second 6*N state buffer used for ping-pong (may equal NULL only if
 *   the implementation selected needs none; always pass one).
 * ckpt: NULL or (ceil(n_steps/ckpt_every)) slots of 6*N: slot k receives the
 *   state at step n0 + k*ckpt_every before that step runs.
 * estore: NULL or n_steps slots of 3*N: slot i receives (Ex,Ey,Ez) after
 *   step n0 + i.
 * probe_index >= 0: after step n0 + i, probe_out[i] = state[probe_index]
 *   (flat index into the 6*N state; FDTDSolver3D.time_series, fdtd3d.py:652-700).
 * *final_in_work (host int): 1 if the final state is in `work`, else 0. */
int32_t dnx_fdtd3d_run(const dnx_grid3d* g, void* state, void* work, int32_t n0,
                       int32_t n_steps, const dnx_source* srcs, int32_t n_src,
                       const dnx_monitor* mons, int32_t n_mon, void* ckpt,
                       int32_t ckpt_every, void* estore, int64_t probe_index,
                       void* probe_out, void* stream, int32_t* final_in_work);
/* As dnx_fdtd3d_run, with the initial state read from init (6*N, left
 * unchanged) instead of state: the replay of a checkpoint segment starts
 * from the checkpoint slot without copying it (init NULL: as
 * dnx_fdtd3d_run).  Same reference functions as dnx_fdtd3d_run. */
int32_t dnx_fdtd3d_run_from(const dnx_grid3d* g, const void* init, void* state, void* work,
                            int32_t n0, int32_t n_steps, const dnx_source* srcs, int32_t n_src,
                            const dnx_monitor* mons, int32_t n_mon, void* ckpt,
                            int32_t ckpt_every, void* estore, int64_t probe_index,
                            void* probe_out, void* stream, int32_t* final_in_work);

/* Adjoint initialisation at m = T: lam[0:3N] holds dL/dE^T on entry;
 * adds the monitor terms of sample T (n_total = T), clears lam[3N:6N]
 * (the H adjoint) and sets S = sum_c E^T_c * lam_c.  e_T: E^T (3*N). */
int32_t dnx_fdtd3d_adjoint_init(const dnx_grid3d* g, void* lam, const void* e_T,
                                const dnx_monitor* mons, int32_t n_mon, int32_t n_total,
                                void* S, void* stream);

/* Reverse sweep over adjoint steps m = m_hi-1 ... max(m_lo, 1), each
 * mapping (lam_E^{m+1}, mu_H^{m+1}) in `lam` to (lam_E^m, mu_H^m) and
 * accumulating S += sum_c E^m_c (lam^m_c - b lam^{m+1}_c) - s_m lam^{m+1};
 * if m_lo == 0 it finally adds the step-0 source term.  This is the exact
 * discrete transpose of dnx_fdtd3d_run, i.e. the gradient the reference
 * obtains by autograd (fdtd3d.py:166-219).  dL/deps = -S / eps.
 * e_lo: E^{m_lo} (3*N); estore: E^{m_lo+1..m_hi-1} (slot j = E^{m_lo+1+j}).
 * workspace: device scratch of at least
 *   dnx_fdtd3d_adjoint_workspace_bytes(g, mons, n_mon) bytes (the two-step
 *   kernels' monitor adjoint terms); may be NULL when that is 0.  Contents
 *   on entry do not matter; calls sharing one workspace must not overlap. */
int64_t dnx_fdtd3d_adjoint_workspace_bytes(const dnx_grid3d* g, const dnx_monitor* mons,
                                           int32_t n_mon);
int32_t dnx_fdtd3d_adjoint_run(const dnx_grid3d* g, void* lam, void* work, int32_t m_hi,
                               int32_t m_lo, const void* e_lo, const void* estore,
                               const dnx_source* srcs, int32_t n_src,
                               const dnx_monitor* mons, int32_t n_mon, void* S,
                               void* workspace, int64_t ws_bytes, void* stream,
                               int32_t* final_in_work);

/* dnx_fdtd3d_adjoint_run without a workspace argument.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * The scratch the two-step kernels need (the monitor adjoint terms, the
 * bytes dnx_fdtd3d_adjoint_workspace_bytes reports) is a fixed device array
 * of this library, dnx_fdtd3d_adjoint_static_scratch_bytes() = 4 MiB per
 * device, held until the process ends; monitors needing more are rejected
 * with DNX_ERANGE.  Calls on one device share that array, so calls that
 * overlap on one device (other streams, other host threads) race; nothing
 * in it has to start at zero.  Everything else as dnx_fdtd3d_adjoint_run;
 * no hipMalloc, hipFree or allocating library routine is called. */
int64_t dnx_fdtd3d_adjoint_static_scratch_bytes(void);
int32_t dnx_fdtd3d_adjoint_run_noalloc(const dnx_grid3d* g, void* lam, void* work, int32_t m_hi,
                                       int32_t m_lo, const void* e_lo, const void* estore,
                                       const dnx_source* srcs, int32_t n_src,
                                       const dnx_monitor* mons, int32_t n_mon, void* S,
                                       void* stream, int32_t* final_in_work);

/* Tile shape of the fused 3D kernels for one dtype and kernel (0 forward/
 * replay, 1 adjoint, 2 two-step forward/replay, 3 two-step adjoint; the
 * two-step kernels run for float32 only): ty rows (2, 4 or 8; 2 or 4 for
 * the two-step forward; 1, 2 or 4 for the two-step adjoint, 2 or 4 in
 * batched plans; 0 switches a two-step kernel off, so the one-step kernel
 * runs every step) and zc z-planes per work-group.  Process-wide; set by
 * the autotuner.  No reference counterpart (the reference has no kernels);
 * it changes performance only, never results. */
int32_t dnx_fdtd3d_set_tiles(int32_t dtype, int32_t kernel, int32_t ty, int32_t zc);
int32_t dnx_fdtd3d_get_tiles(int32_t dtype, int32_t kernel, int32_t* ty, int32_t* zc);

/* ---- batched 3D (many designs, uneven grids, one launch per step) ------ */

/* One design of a batch: its grid (dtype shared by the batch, nx <= 256),
 * at most one soft source (src.comp < 0: none) and one DFT monitor
 * (mon.n_freq == 0: none; acc and grad both required).  No reference
 * counterpart: the reference solves designs one at a time
 * (fdtd3d.py:471-504 per design). */
typedef struct dnx_design3d {
  dnx_grid3d grid;
  dnx_source src;
  dnx_monitor mon;
} dnx_design3d;

/* Device pools shared by all designs; design d starts at element offset
 * 6*Nd (state, lam), 3*Nd (estore slot), Nd (S), Nd = sum_{e<d} N_e.
 * ckpt: ceil(T/K) slots of sum 6N; estore: K-1 slots of sum 3N.  side:
 * dnx_batch3d_side_elems() elements (monitor staging of the float32
 * two-step kernels); null runs the one-step kernels when a design has a
 * monitor.  etail (may be null): T - a_tail slots of sum 3N in which the
 * recording forward keeps E of every step after a_tail (a multiple of K,
 * 0 <= a_tail < T), so that the backward runs those segments' adjoint
 * without replaying them. */
typedef struct dnx_batch3d_pools {
  void* state[2];
  void* lam[2];
  void* ckpt;
  void* estore;
  void* S;
  void* side;
  void* etail;
  int32_t a_tail;
} dnx_batch3d_pools;

typedef struct dnx_batch3d_plan dnx_batch3d_plan;  /* opaque host object */

/* Device workspace the plan needs: descriptor tables, the step counter and
 * the block maps of the two-step launches (one work-group per tile, the
 * designs packed onto the XCDs by cost), at the tile shapes set now (set
 * them before this query; the plan takes them when it is created).
 * Returns -1 on bad arguments. */
int64_t dnx_batch3d_workspace_bytes(const dnx_design3d* designs, int32_t n_designs);
/* Elements (of the designs' dtype) of the pools.side buffer: two samples of
 * every design's monitor-box components (-1: a grid or monitor box the
 * batch cannot take). */
int64_t dnx_batch3d_side_elems(const dnx_design3d* designs, int32_t n_designs);
/* The work-group -> (design, tile) map of the two-step launches (kernel 2
 * forward, 3 adjoint) at the current tile shapes, as the plan would build
 * it: entry b (out[2b], out[2b+1]) is what work-group b runs, (-1, -1) idle;
 * work-group b runs on XCD b % 8.  Returns the entry count (out may be
 * null to query it), -1 on bad arguments or too small a capacity.  Host
 * only.  No reference counterpart (introspection of the batched launch). */
int64_t dnx_batch3d_block_map(const dnx_design3d* designs, int32_t n_designs, int32_t kernel,
                              int32_t* out, int64_t capacity);
/* Build a plan (uploads descriptor tables into `workspace` on `stream`).
 * ckpt_every must be even.  use_graph: capture one K-step segment of each
 * sequence (forward, replay, adjoint) into a HIP graph on first use and
 * replay it per segment (falls back to direct launches if capture fails). */
int32_t dnx_batch3d_create(const dnx_design3d* designs, int32_t n, int32_t n_steps,
                           int32_t ckpt_every, const dnx_batch3d_pools* pools,
                           void* workspace, int64_t ws_bytes, int32_t use_graph,
                           void* stream, dnx_batch3d_plan** out);
/* Forward T steps from the state in pools.state[0] (caller zeroes it and
 * the monitor accumulators); record_ckpt: store checkpoints.  Final state
 * of every design in pools.state[dnx_batch3d_final_slot(plan)]. */
int32_t dnx_batch3d_forward(dnx_batch3d_plan* plan, int32_t record_ckpt, void* stream);
/* Gradient of a loss on the monitors (grads in each design's mon.grad):
 * replay + adjoint for every design; S in pools.S (dL/deps = -S/eps). */
int32_t dnx_batch3d_backward(dnx_batch3d_plan* plan, void* stream);
int32_t dnx_batch3d_uses_graph(const dnx_batch3d_plan* plan);
/* Which state pool (0 or 1) holds the state after the forward's last step:
 * T & 1 on the one-step path; the float32 two-step path (launches of two
 * steps) uses its own ping-pong order.  (Batched form of the reference's
 * per-design state, fdtd3d.py:497-505; no direct counterpart.) */
int32_t dnx_batch3d_final_slot(const dnx_batch3d_plan* plan);
/* Whether the plan runs the two-step kernels (bit 0 forward, bit 1 adjoint). */
int32_t dnx_batch3d_uses_pairs(const dnx_batch3d_plan* plan);
int32_t dnx_batch3d_destroy(dnx_batch3d_plan* plan);

/* ---- 2D -------------------------------------------------------------- */

/* As dnx_fdtd3d_run for the 2D state (3*N): TM FDTDSolver2D._time_step_tm
 * (fdtd2d.py:233-267), TE _time_step_te (fdtd2d.py:269-303), source into
 * Ez/Hz (fdtd2d.py:333,351).  estore slots hold the first component (Ez or
 * Hz), N each.  probe: as 3D (fdtd2d.py:335-337, :486-524). */
int32_t dnx_fdtd2d_run(const dnx_grid2d* g, void* state, void* work, int32_t n0,
                       int32_t n_steps, const dnx_source* srcs, int32_t n_src,
                       const dnx_monitor* mons, int32_t n_mon, void* ckpt,
                       int32_t ckpt_every, void* estore, int64_t probe_index,
                       void* probe_out, void* stream, int32_t* final_in_work);

/* 2D adjoint, same contract as the 3D pair; e_T / e_lo / estore hold the
 * first state component (N each).  TE accumulates S from the stored Hz
 * (the E update reads the old Hz). */
int32_t dnx_fdtd2d_adjoint_init(const dnx_grid2d* g, void* lam, const void* e_T,
                                const dnx_monitor* mons, int32_t n_mon, int32_t n_total,
                                void* S, void* stream);
int32_t dnx_fdtd2d_adjoint_run(const dnx_grid2d* g, void* lam, void* work, int32_t m_hi,
                               int32_t m_lo, const void* e_lo, const void* estore,
                               const dnx_source* srcs, int32_t n_src,
                               const dnx_monitor* mons, int32_t n_mon, void* S,
                               void* stream, int32_t* final_in_work);

/* Batched 2D TM forward (float32) over `batch` designs of one shape: g->ce
 * points at batch*N coefficients, state at batch*3*N ([Ez, Hx, Hy] per
 * design, zeroed by the caller), one shared soft source (src may be NULL)
 * and a running DFT over mbox with n_freq frequencies accumulated in
 * float32: acc[b][f][re, im][mbox cells] += Ez (cos, -sin)(phase[n][f]),
 * phase float32 [(n*n_freq + f)*2 + {0,1}].  Semantics per design are
 * those of dnx_fdtd2d_run (FDTDSolver2D._time_step_tm, fdtd2d.py:233-267);
 * no reference counterpart for the batching (surrogate data / candidate
 * judging). */
int32_t dnx_fdtd2d_tm_batch(const dnx_grid2d* g, int32_t batch, void* state, int32_t n_steps,
                            const dnx_source* src, const dnx_box* mbox, int32_t n_freq,
                            const float* phase, float* acc, void* stream);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* DNX_FDTD_H */

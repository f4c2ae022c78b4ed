/*
 * licom_sw.h — C-ABI of the MI355X (gfx950) barotropic shallow-water step.
 *
 * All six cube faces of one resolution live on one GPU as (6, N, N) arrays,
 * N = nx + 6 (halo 3), C-contiguous, j (last axis) fastest.  Device pointers
 * are plain addresses of device memory; `stream` is a hipStream_t (or null
 * for the default stream).  Every entry point returns 0 on success or a
 * negative SW_E* code; sw_last_error() then describes the failure (the text
 * is thread-local).  No entry point takes torch types.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   sw_rk2 / sw_rk3      momentum/barotr.py:41-61 `_barotr_rk2_core` and
 *                        :64-87 `_barotr_rk3_core` (the RK substep loop,
 *                        stages from momentum/barotr_pure.py:100-168)
 *   sw_stage             momentum/barotr_pure.py:100-168 `_step_rk_logic`
 *   sw_halo_scalar       kernel/cube.py:130-134 `ext_scalar`
 *                        (= communication.py:387-480 `update_domain` with
 *                        px = py = 1, then cube.py:40-127 `cube_rmp`)
 *   sw_halo_vector       kernel/cube.py:137-153 `ext_vector`
 *   sw_edge_avg          kernel/communication.py:485-570
 *                        `boundary_communication`
 *   sw_routing_table     kernel/communication.py:57-87, 277-284 (the
 *                        `_routing` table built by `Communication.configure`)
 */
#ifndef LICOM_SW_H
#define LICOM_SW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sw_ctx sw_ctx;

enum {
  SW_F64 = 0,  /* state and stencil arithmetic in binary64 (the drop-in) */
  SW_F32 = 1   /* state and stencil arithmetic in binary32 */
};

enum {
  SW_OK = 0,
  SW_E_ARG = -1,     /* invalid argument (null pointer, bad size, bad dtype) */
  SW_E_HIP = -2,     /* a HIP runtime call failed */
  SW_E_CONST = -3,   /* consts violate an assumption (see sw_last_error) */
  SW_E_STATE = -4    /* precondition on the state failed */
};

/* Flags for sw_rk2 / sw_rk3. */
enum {
  SW_SKIP_DIAG = 1,  /* do not write the 10 diagnostic outputs */
  SW_GRAPH = 2,      /* replay recorded hipGraphs of substep pairs (same
                        results; the first call per dtb records them) */
  SW_INPLACE = 4     /* out[0..2] may alias in[] (the inputs are read by the
                        first substep only): an in-place update */
};

/* Number of consts and their order: mymodule/schedule.py:95-116
 *   0 dzph_x   1 dzph_y   2 rdx   3 rdy   4 a_f
 *   5 k2e_coef (6,N,N,2)  6 loc_i_local (int32)  7 loc_j_local (int32)
 *   8 a_c2l (6,N,N,2,2)   9 a_l2c (6,N,N,2,2)  10 inner  11 outer
 *  12 a_gct (6,N,N,2,2)  13 a_sina  14 rda  15 c_dy  16 d_dx
 * Float consts are binary64 device arrays; they are copied (and, for SW_F32,
 * rounded) into the context, so the caller may free them after sw_create. */
#define SW_N_CONSTS 17

/* Create a context for resolution C{nx} (6 faces of (nx+6)^2) whose state is
 * of `dtype`.  Validates the index tables (loc values in range) and the
 * masks (inner = 1 on the interior and 0 on the halo, outer = 1 - inner). */
int sw_create(sw_ctx** out, int nx, int dtype, const void* const consts[SW_N_CONSTS],
              void* stream);
void sw_destroy(sw_ctx* ctx);

/* Bind the stream later launches go to. */
int sw_set_stream(sw_ctx* ctx, void* stream);

/* nbb RK2-midpoint substeps of length dtb (barotr.py:41-61).
 *   in[6]  = h0, h0p, ub, vb, ubp, vbp          (read only; may alias)
 *   out[13]= h0, ub, vb, celerity_x, celerity_y, ub_ct, vb_ct, ub_cx, ub_cy,
 *            vb_cx, vb_cy, advx, advy          (distinct from in[])
 * With SW_SKIP_DIAG, out[3..12] are not written and may be null.
 * The reference returns h0p = h0, ubp = ub, vbp = vb and passes the coupling
 * fields through; callers alias those.  nbb = 0 copies in -> out (h0, ub, vb)
 * and needs SW_SKIP_DIAG: the reference's empty loop returns the input's
 * diagnostics, which in[] does not carry, so the caller keeps its own
 * (nbb = 0 without SW_SKIP_DIAG returns SW_E_ARG). */
int sw_rk2(sw_ctx* ctx, const void* const in[6], void* const out[13], int nbb,
           double dtb, int flags);
/* nbb RK3 substeps (barotr.py:64-87): stages dt/3, dt/2, dt with beta = 1. */
int sw_rk3(sw_ctx* ctx, const void* const in[6], void* const out[13], int nbb,
           double dtb, int flags);

/* One RK stage (barotr_pure.py:100-168); out[13] as sw_rk2, in the order of
 * the reference's 13-tuple.  beta is the forward-backward blend weight. */
int sw_stage(sw_ctx* ctx, const void* const in[6], void* const out[13], double dt,
             double beta, int is_last);

/* In-place duo-grid halo fill of `ncomp` stacked fields laid out
 * (6, ncomp, N, N) (ncomp = 1 for (6, N, N)). */
int sw_halo_scalar(sw_ctx* ctx, void* field, int ncomp);
/* In-place vector halo fill of covariant (u, v) via lat-lon rotation. */
int sw_halo_vector(sw_ctx* ctx, void* u, void* v);
/* In-place C-grid edge averaging: u holds the axis -2 normal component on the
 * N/S edges, v the axis -1 normal component on the E/W edges. */
int sw_edge_avg(sw_ctx* ctx, void* u, void* v);

/* Wait for all work the context launched. */
int sw_sync(sw_ctx* ctx);

/* Host-only: the cube routing table R[t][d] = (neighbour face, its edge,
 * transform 0=ID / 1=FLIP), directions E=0, W=1, N=2, S=3. */
int sw_routing_table(int32_t out[6 * 4 * 3]);

/* Eager RK2 run of nbb substeps (as sw_rk2 with flags 0) on the context's
 * stream, each launch repeated `reps` times (every launch is idempotent) and
 * timed by start / stop hipEvents attached to its own dispatch
 * (hipExtLaunchKernelGGL); waits for completion.
 * ms[k] / counts[k] = total milliseconds / launches of kernel class k:
 * 0 edge lines, 1 height stage kernel, 2 height halo fill, 3 velocity stage
 * kernel, 4 velocity halo fill, 5 diagnostics, 6 fused stage kernel (height
 * and deep-cell velocity), 7 frame velocity with the height halo fill.
 * Classes 1, 2 and 3 run when fusion is off (sw_set_tuning rows_f = 0), 6 and
 * 7 when it is on.  A measurement helper with no counterpart in the
 * reference. */
#define SW_N_KERNEL_CLASSES 8
int sw_profile(sw_ctx* ctx, const void* const in[6], void* const out[13], int nbb,
               double dtb, int reps, double ms[SW_N_KERNEL_CLASSES], int counts[SW_N_KERNEL_CLASSES]);

/* The fused stage kernel as it runs in the recorded substep-pair graphs of
 * sw_rk2's SW_GRAPH path (which event nodes inside a graph cannot time): the
 * first RK2 substep runs eagerly (untimed) from in[] into the context's
 * buffer, then a graph of `pairs` substep pairs is recorded with the fused
 * kernel's launch stamps on (lane 0 of every wave writes the device wall
 * clock at entry and exit; a launch lasts from its first entry to its last
 * exit) and replayed: `replays` measured replays, each after 3 unmeasured
 * ones queued back to back (so the device runs at the clocks of a sustained
 * sequence of replays), the host reading the stamps after each measured one.
 * ms[6] / counts[6] (the fused class of sw_profile) sum the launch durations
 * over the measured replays, the other classes stay 0; *graph_ms sums the
 * measured replays (events around each graph launch).  out[0..2] receive
 * h, u, v after 1 + 2 * pairs * 4 * replays substeps.  Needs the fused form (rows_f > 0).  A
 * measurement helper with no counterpart in the reference. */
int sw_profile_graph(sw_ctx* ctx, const void* const in[6], void* const out[13], int pairs, int replays,
                     double dtb, double ms[SW_N_KERNEL_CLASSES], int counts[SW_N_KERNEL_CLASSES],
                     double* graph_ms);

/* Global integrals of a state in binary64 with compensated summation over
 * the interior cells, whatever the state precision:
 *   out[0] mass   = sum (dzph + h) * da
 *   out[1] energy = sum da * (1/2 (dzph + h) u_i G^ij u_j + 1/2 * 9.8 * h^2)
 * with da = 1 / rda and G^ij = a_gct.  Waits for the result.  Build-defined
 * diagnostics with no counterpart in the reference (whose diagnostics are
 * interior max / min / mean, mymodule/diag.py:45-75). */
int sw_integrals(sw_ctx* ctx, const void* h, const void* u, const void* v, double out[2]);

/* sw_integrals without allocating: the same kernels and the same result bits,
 * with the per-block partial sums in a library-owned static device array
 * (1024 blocks x 4 doubles = 32 KiB per GPU, held until the process ends,
 * whatever the grid size: the kernel strides over the cells) and the two
 * results written to `out`, device memory of 2 doubles, asynchronously on the
 * context's stream (no host synchronisation).  Calls no hipMalloc / hipFree,
 * so it may run under hipGraph stream capture or in a loop that must not
 * allocate.  Calls that overlap on one GPU share the array and race.  No
 * counterpart in the reference; for callers that cannot allocate (graph
 * capture, serving loops). */
int sw_integrals_noalloc(sw_ctx* ctx, const void* h, const void* u, const void* v, double* out);

/* Rows per wavefront band of the stage kernels: rows_a / rows_b for the
 * separate height and velocity kernels (default 32), rows_f for the fused
 * kernel, which runs when rows_f > 0; rows_f = 0 selects the separate
 * kernels.  sw_create sets rows_f from the resolution, the precision and the
 * device's CU count (the band height that keeps the fused waves within few
 * rounds of resident waves).  Performance only: every setting in [4, 4096] gives
 * bit-identical results.  A batch (sw_batch_run) uses the fused kernel when
 * every member's rows_f > 0.  Build-defined, no counterpart in the
 * reference. */
int sw_set_tuning(sw_ctx* ctx, int rows_a, int rows_b, int rows_f);

/* The context's current band heights (sw_set_tuning).  Build-defined. */
int sw_get_tuning(const sw_ctx* ctx, int* rows_a, int* rows_b, int* rows_f);

/* Split bands of the fused kernel (sw_rk2 / sw_rk3 / sw_stage on one
 * context; batched runs keep uniform bands).  A launch fills the device in
 * one round of waves, and the older of two waves sharing a SIMD gets about
 * twice the issue slots, so with uniform bands the second block on every CU
 * finishes far later than the first.  With rows_short > 0 the columns of
 * each face alternate bands of rows_f rows (marched by the first blocks
 * dispatched, at most one per CU) and rows_short rows (the later blocks), so
 * both finish together.  0 restores uniform bands.  Performance only:
 * results are bit-identical for every setting.  Build-defined. */
int sw_set_band_split(sw_ctx* ctx, int rows_short);
int sw_get_band_split(const sw_ctx* ctx, int* rows_short);

/* Form of the launch that follows the fused kernel in every stage (sw_rk2 /
 * sw_rk3 / sw_stage on one context).  1: one cube-edge launch evaluates the
 * frame cells, the velocity halo fill and the next stage's averaged edge
 * values, a block per segment of a cube edge holding both faces' frame cells
 * in LDS (the default when every halo cell's interpolation sources lie
 * within 3 along-edge positions of its own, which sw_create checks); 0: the
 * frame launch, then a launch for the velocity halo fill and the edge pass.
 * SW_E_ARG when 1 is asked of a grid that fails the check.  Performance
 * only: results are bit-identical.  Build-defined. */
int sw_set_frame_form(sw_ctx* ctx, int cube_edge);
int sw_get_frame_form(const sw_ctx* ctx, int* cube_edge);

/* Host-only: what sw_create detected in the consts.  Bit 0: dzph_x/dzph_y
 * uniform (read as scalars); bit 1: rdx, rdy, rda, c_dy, d_dx identical on all
 * faces (read from face 0); bit 2: a_gct symmetric with equal diagonal. */
int sw_layout_flags(const sw_ctx* ctx, int* flags);

/* A context for another member of the same resolution: it shares base's
 * device copy of the consts (kept alive until the last context using it is
 * destroyed) and has its own workspace.  For ensembles and served solves; no
 * counterpart in the reference, which runs one field per process. */
int sw_create_like(sw_ctx** out, const sw_ctx* base);

/* One member of a batched run: its own context, the six input fields and the
 * 13 outputs (as for sw_rk2), its substep count and substep length. */
typedef struct sw_member {
  sw_ctx* ctx;
  const void* in[6];
  void* out[13];
  int nbb;
  double dtb;
} sw_member;

typedef struct sw_batch sw_batch;

/* Reusable device tables of batched runs. */
int sw_batch_create(sw_batch** out);
void sw_batch_destroy(sw_batch* b);

/* Advance `count` members (any resolutions, one dtype, distinct contexts) by
 * their own nbb RK substeps (rk = 2 or 3), every kernel class launched once
 * per stage for all members still stepping, on `stream`.  Each member's
 * outputs equal those of sw_rk2 / sw_rk3 on its context alone (flags:
 * SW_SKIP_DIAG; SW_GRAPH replays recorded graphs of substep pairs, one per
 * number of members still stepping).  A member with nbb = 0 needs
 * SW_SKIP_DIAG, as for sw_rk2.  Generalises the reference's one-field loop
 * (momentum/barotr.py:41-87) to many fields per launch.  A batch object may be
 * reused; its tables are replaced at every run. */
int sw_batch_run(sw_batch* b, const sw_member* members, int count, int rk, int flags, void* stream);

/* Host-only: resolution and dtype of a context. */
int sw_info(const sw_ctx* ctx, int* nx, int* dtype);

const char* sw_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LICOM_SW_H */

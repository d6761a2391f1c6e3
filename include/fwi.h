/* fwi.h -- batched FWI gradient API (this library's own entry points).
 *
 * The reference computes one shot per MPI rank through the per-step launchers of
 * include/fwi_serialcalls.h (src/main.c:396-538): nt x launchPropagation with a full-wavefield copy to
 * pinned host memory, the residual, nt x launchBackPropagation, then an MPI float reduction of the
 * gradient.  The functions below replace that whole per-iteration block (main.c:396-538) and the
 * forward-only surveys of main.c:334-366 and :695-737 with on-device shot batching: all shots of a batch
 * advance together in one launch per time step, the adjoint pass recomputes the forward field from
 * checkpoints every K ~ sqrt(nt) steps instead of storing it, and the cross-shot sum, misfit and
 * gradient are reduced on the device in FP64.
 *
 * All array arguments are device pointers unless marked host.  Fields are x fastest (ix + nx*iz).
 * Status-returning functions return 0 on success and a nonzero code with fwi_error_message() set on
 * failure; they never fall back to another implementation.  Before any device work they refuse a null
 * plan, optimizer or required array, and a device-array argument that the runtime reports as pageable
 * host memory (a kernel would fault on it); device, managed and pinned host memory are accepted.
 */
#ifndef FWI_H
#define FWI_H

#ifdef __cplusplus
extern "C" {
#endif

/* One shot in window-local coordinates: its grid is model columns [x0, x0+nx) with its own C-PML on the
 * window's left/right edges (and the bottom); source at (isx, isz); receivers on row irz at columns
 * [ir0, ir0+nrec).  The reference's shot (propagator_F, src/parallelcalls.cu:107-116) is x0=0, nx=model
 * width, isx=sx-1, isz=irz=sz-1, ir0=CPML, nrec=nx-2*CPML.  Source and receivers must lie in the updated
 * interior (4 <= i <= n-5). */
typedef struct fwi_shot {
  int x0, nx;
  int isx, isz;
  int irz, ir0, nrec;
} fwi_shot;

typedef struct fwi_survey {
  int nx, nz;            /* model grid (cells) */
  int nt;                /* time steps */
  int cpml;              /* C-PML width in cells (main.c:38), > 4 */
  float dh, dt;          /* grid spacing [m], time step [s] */
  int nshots;
  const fwi_shot* shots; /* host array of nshots */
  const float* wavelet;  /* host, nt samples (one source wavelet for every shot, main.c:328-333) */
  const float* cpml_a;   /* host, cpml entries: C-PML a by distance from the edge (fwi_cpml_table) */
  const float* cpml_b;   /* host, cpml entries: C-PML b */
} fwi_survey;

enum { FWI_GRAD_REFCOMPAT = 0, FWI_GRAD_EXACT = 1 };
/* Stencil arithmetic: FP32 in the reference's expression order; MIXED = sums in FP64 as the
 * reference's double literals make them; FP32D = FP32 with second derivatives in difference form
 * sum_k c_k((u_{+k}-u_0)+(u_{-k}-u_0)), which avoids the cancellation of the reference form (up to 30x
 * smaller round-off in isotropic media; little change in strongly anisotropic ones); FP32I (default) =
 * FP32 with the leapfrog in increment form, forward and exact adjoint: the state is (P^n, P^n - P^{n-1})
 * instead of two levels, so the time difference keeps FP32 relative precision (at 150+ samples per period
 * the two-level form amplifies rounding several hundred times), for 8 more bytes per cell on every step;
 * and, by default, the near-surface band of the forward (surface_rows) in FP64.  On the 1600x400 survey at
 * 8 Hz this takes the exact gradient from ~3e-3 to ~2e-5 (relative L2 against FP64, full grid). */
enum { FWI_ARITH_FP32 = 0, FWI_ARITH_MIXED = 1, FWI_ARITH_FP32D = 2, FWI_ARITH_FP32I = 3 };

/* Checkpoint storage: FP32, or 16 bits after an exact power-of-two scale per field, shot and snapshot (half
 * the checkpoint memory; needs FWI_ARITH_FP32I): BF16 (8-bit mantissa) or FP16 with the field maximum
 * scaled to [2^14, 2^15) (11-bit mantissa).  Recomputed segments then start from rounded states: the
 * gradient moves by ~1e-2 (BF16) or ~1e-3 (FP16) relative. */
enum { FWI_CKPT_FP32 = 0, FWI_CKPT_BF16 = 1, FWI_CKPT_FP16 = 2 };

typedef struct fwi_options {
  int grad_mode;       /* FWI_GRAD_EXACT (true discrete gradient) or FWI_GRAD_REFCOMPAT (the reference's) */
  int arith;           /* FWI_ARITH_FP32I (default), FWI_ARITH_FP32, FWI_ARITH_FP32D or FWI_ARITH_MIXED */
  int ckpt_interval;   /* K: 0 = ceil(sqrt(nt)); >= nt stores every step (no recomputation) */
  int max_batch;       /* shots per batch, 0 = as many as mem_budget allows */
  double mem_budget;   /* bytes of device memory the plan may use, 0 = 85 % of free memory */
  int rows_per_item;   /* z rows per wave in the stencil kernels, 0 = chosen by the tuner */
  int item_order;      /* work-item order 0..4 (see plan.hip build_items), -1 = chosen by the tuner */
  int grad_group;      /* exact mode: up to this many consecutive shots with identical windows share one
                          FP32 gradient slot, their per-step imaging terms added in shot order before one
                          read-modify-write (0 = 4, the default; 1 = one slot per shot) */
  int ckpt_format;     /* FWI_CKPT_FP32 (default), FWI_CKPT_BF16 or FWI_CKPT_FP16 */
  int store_segments;  /* trailing checkpoint segments whose imaging terms the forward pass keeps (they are
                          neither checkpointed nor recomputed): 0 = as many as mem_budget allows with FP32
                          checkpoints (the result is the same for any count), 1 with 16-bit checkpoints */
  int use_graph;       /* 1 (default): each batch's device work of fwi_forward / fwi_misfit / fwi_gradient /
                          fwi_gradient_src is captured once as a HIP graph per set of caller buffers and
                          replayed on later calls with the same buffers; 0: launched eagerly */
  int step_pairs;      /* increment form: the core of every window (4+ cells from the C-PML and below the FP64 band)
                          advances two steps per launch, the rest by single steps; 1 (default) = rows per core work
                          item tuned, k > 1 = k rows; 0: single steps everywhere.  The result is the same either
                          way. */
  int surface_rows;    /* FWI_ARITH_FP32I: rows 4..surface_rows-1 of every shot (the source, receivers and near-surface
                          gradient) advance in FP64 in the forward and recomputation passes.  0 (default) = 24 rows
                          where nz - cpml - 4 >= 24 and the plan allows it (FP32 checkpoints), none otherwise;
                          -1 = none; k > 4 = k rows (k <= nz - cpml - 4; FP32 checkpoints). */
  int adj_rows_per_item; /* z rows per wave in the adjoint kernels: 0 = rows_per_item when that is set, else tuned */
  int adj_item_order;    /* adjoint work-item order 0..3 (1 means 0 with gradient groups), -1 = derived from item_order
                            when that is set (1 -> 0 with groups, 4 -> 3), else tuned */
  int adj_step_pairs;    /* exact gradient, FP32 increment form, gradient groups: the core of every group (8+ cells from
                            the C-PML, 4+ below the FP64 band) advances two adjoint steps per launch, the rest by
                            single steps; 1 (default) = rows per core work item tuned and kept only where faster than
                            single steps, k > 1 = k rows; 0: single steps.  The result is the same either way. */
} fwi_options;

typedef struct fwi_plan fwi_plan;

/* Fill *opt with defaults (exact gradient, FP32 increment form, K = ceil(sqrt(nt)), automatic batch, tile and
 * order, gradient groups of 4). */
void fwi_options_default(fwi_options* opt);

/* Validate the survey, size the batches, allocate the device workspace. */
int fwi_plan_create(const fwi_survey* survey, const fwi_options* opt, fwi_plan** plan);
void fwi_plan_destroy(fwi_plan* plan);

/* Replace the wavelet (host, nt) or the C-PML table (host, cpml entries), e.g. for a new frequency stage. */
int fwi_plan_set_wavelet(fwi_plan* plan, const float* wavelet);
int fwi_plan_set_cpml(fwi_plan* plan, const float* cpml_a, const float* cpml_b);

/* Size of the gather array of the whole survey: sum over shots of nt*nrec (floats); layout [shot][it][rec]. */
long long fwi_plan_gather_size(const fwi_plan* plan);
/* Shots in the first batch and number of batches chosen by the plan (every batch runs the same launches). */
int fwi_plan_batch_size(const fwi_plan* plan);
int fwi_plan_batch_count(const fwi_plan* plan);
/* Device bytes held by the plan. */
double fwi_plan_workspace_bytes(const fwi_plan* plan);

/* Forward modelling of every shot (observed data: main.c:334-366).  v_hor, v_po, v_nmo: model grid. */
int fwi_forward(fwi_plan* plan, const float* v_hor, const float* v_po, const float* v_nmo, float* gathers,
                void* stream);

/* Misfit only (a line-search trial: main.c:695-737): *misfit = 1/2 sum (syn - obs)^2 in FP64 (host out). */
int fwi_misfit(fwi_plan* plan, const float* v_hor, const float* v_po, const float* v_nmo, const float* obs,
               double* misfit, void* stream);

/* One FWI gradient (main.c:396-538): forward with checkpoints, residual, FP64 misfit (host out), adjoint
 * with recomputation, imaging, FP64 cross-shot sum.  grad (device, FP64, 3*nx*nz) receives, in order, the
 * gradient with respect to v_hor (the reference's gcx), v_po (gcz) and v_nmo (gdx); overwritten.
 * syn (optional, may be NULL): the synthetic gathers.  In FWI_GRAD_REFCOMPAT mode the result reproduces
 * the reference's formula (serialcalls.cu:68-90, a dt-scaled imaging with a one-step lag). */
int fwi_gradient(fwi_plan* plan, const float* v_hor, const float* v_po, const float* v_nmo, const float* obs,
                 double* grad, double* misfit, float* syn, void* stream);

/* Gradient of <adj_src, gathers(model)> (the adjoint driven by an arbitrary adjoint source instead of the
 * residual): adj_src has the survey gather layout.  With adj_src = dL/dsyn this is dL/d(v_hor, v_po, v_nmo)
 * for any loss L of the gathers; fwi_gradient is the case adj_src = syn - obs.  No counterpart in the
 * reference, whose adjoint source is always the residual (parallelcalls.cu:159-163). */
int fwi_gradient_src(fwi_plan* plan, const float* v_hor, const float* v_po, const float* v_nmo, const float* adj_src,
                     double* grad, float* syn, void* stream);

/* Per-shot misfits 1/2 sum res^2 (FP64, host out[nshots]) of the last fwi_misfit or fwi_gradient call.  Each
 * is a fixed-order reduction over its own gather, and the total misfit adds them in shot order, so neither
 * depends on how the shots were batched.  The reference keeps only the rank sums (main.c:448-454,498-510). */
int fwi_plan_shot_misfits(fwi_plan* plan, double* out, void* stream);

/* Chain rule to (vp, epsilon, delta) for v_po = vp^2, v_hor = vp^2(1+2 eps), v_nmo = vp^2(1+2 delta):
 * out = [dJ/dvp, dJ/deps, dJ/ddelta] (FP64, 3*n) from grad = [dJ/dv_hor, dJ/dv_po, dJ/dv_nmo]. */
int fwi_chain_rule(const float* vp, const float* eps, const float* delta, const double* grad, double* out, long long n,
                   void* stream);

/* ---------------------------------------------------------------------------------------------------
 * Optimizer (main.c:539-800) on the device in FP64.  Three blocks in the order of fwi_gradient's output:
 * v_hor, v_po, v_nmo (n = nx*nz each; grad is FP64 3*n, the model three FP32 fields).  Model updates touch
 * rows iz > 15 only (include/parallelcalls.h:6) and end with v_hor >= v_nmo (parallelcalls.cu:282,400).
 * --------------------------------------------------------------------------------------------------- */
typedef struct fwi_lbfgs fwi_lbfgs;
/* History of m (s, y) pairs per block (the reference uses m = 5, main.c:59). */
int fwi_lbfgs_create(int nx, int nz, int m, fwi_lbfgs** out);
void fwi_lbfgs_destroy(fwi_lbfgs* h);
/* First iteration (main.c:539-568, NORMAL_h + UpdateModel_h): back up m and g, then
 * m -= alpha_k * g_k / ||g_k|| (FP64 norm over the whole grid); alpha is host double[3]. */
int fwi_lbfgs_sd_step(fwi_lbfgs* h, float* v_hor, float* v_po, float* v_nmo, const double* grad, const double* alpha,
                      void* stream);
/* Later iterations (main.c:569-689): push the pair (m - m_back, g - g_back), run the two-loop recursion
 * per block with FP64 dot products (H0 = gamma I, gamma = s.y / y.y of the newest pair) into the
 * direction r, then back up m and g. */
int fwi_lbfgs_direction(fwi_lbfgs* h, const float* v_hor, const float* v_po, const float* v_nmo, const double* grad,
                        void* stream);
/* Trial model of the line search (update_LBFGS_h with flags 1,1,1,0,0): m = m_back - alpha_k r_k. */
int fwi_lbfgs_trial(fwi_lbfgs* h, float* v_hor, float* v_po, float* v_nmo, const double* alpha, void* stream);
/* Restore the backed-up model (main.c:777-781). */
int fwi_lbfgs_restore(fwi_lbfgs* h, float* v_hor, float* v_po, float* v_nmo, void* stream);
/* Copy the current direction r (FP64 3*n) and the number of stored pairs. */
int fwi_lbfgs_get_direction(fwi_lbfgs* h, double* out, void* stream);
int fwi_lbfgs_pairs(const fwi_lbfgs* h);
/* Physical bounds for a driver (no counterpart in the reference, whose only bound is v_hor >= v_nmo): clamp
 * every squared velocity of the three fields (n cells each) to [lo, hi], then v_hor >= v_nmo.  An L-BFGS step
 * from few curvature pairs can otherwise drive a squared velocity negative, and the coupled system then
 * grows without bound. */
int fwi_model_clamp(float* v_hor, float* v_po, float* v_nmo, long long n, float lo, float hi, void* stream);
/* Deterministic FP64 dot products of three blocks of n: out (host double[3]). */
int fwi_dot3(const double* a, const double* b, long long n, double* out, void* stream);
/* fwi_dot3 without any allocation -- no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  out is device or
 * pinned host memory (double[3]), written by the stream's last kernel; the call does not synchronize.  The same
 * block partition as fwi_dot3, so the same bits.  Scratch: a fixed 12 KiB device array per GPU (512 x 3 partial
 * sums, enough for any n), held until the process ends and shared by every call on that GPU: calls that overlap
 * on one GPU (other streams or threads) race on it.  n is refused beyond 2^40. */
int fwi_dot3_noalloc(const double* a, const double* b, long long n, double* out, void* stream);

/* Instrumentation (no counterpart in the reference, whose only timer is clock() around whole phases,
 * main.c:323,368-372).  With stride > 0, every stride-th launch of each fused step kernel is bracketed by
 * HIP events on the stream it runs on; fwi_plan_kernel_time returns the summed event time and the
 * number of timed launches of a kind (0 forward step, 1 forward step storing imaging terms -- each one launch with
 * the FP64 near-surface band --, 2 adjoint step, 3 / 4 step-pair launch without / with stored imaging terms (two
 * steps of the core, one of the band and the shell), 8 / 9 the launch that completes a step pair without / with
 * stored imaging terms (the band's and the shell's second step); 5 the adjoint step-pair launch (two steps of every
 * gradient group's core, one of the shell), 7 the launch completing it (the shell's second step); 6 is no longer
 * used: the exact adjoint computes its C-PML terms inside the adjoint step) since the last fwi_plan_set_timing call; fwi_plan_kernel_launches the number of
 * launches of a kind since then (timed or not; counted while stride > 0). */
void fwi_plan_set_timing(fwi_plan* plan, int stride);
int fwi_plan_kernel_time(fwi_plan* plan, int kind, double* total_ms, long long* count);
long long fwi_plan_kernel_launches(const fwi_plan* plan, int kind);
/* Cells of all shot windows that step pairs advance two steps per launch (the cores; 0 without step pairs). */
long long fwi_plan_core_cells(const fwi_plan* plan);
/* Cells of all shot windows (nx_s*nz summed over shots) and, in *strip_cells, their C-PML cells. */
long long fwi_plan_cells(const fwi_plan* plan, long long* strip_cells);
/* Checkpoint interval K, z rows per work item and work-item order in use (after the first call): the
 * forward kernels' and, through fwi_plan_adj_tiling, the adjoint kernels' (grouped = 1 when the first
 * batch accumulates gradients per group of shots). */
int fwi_plan_ckpt_interval(const fwi_plan* plan);
/* Time steps whose imaging terms the forward pass keeps (the rest are recomputed from checkpoints). */
int fwi_plan_kept_steps(const fwi_plan* plan);
int fwi_plan_rows_per_item(const fwi_plan* plan);
int fwi_plan_item_order(const fwi_plan* plan);
int fwi_plan_adj_tiling(const fwi_plan* plan, int* rows_per_item, int* item_order, int* grouped);
/* Rows per fused core item of the step pairs in use (0 when the plan takes single steps only). */
int fwi_plan_pair_rows(const fwi_plan* plan);
/* Rows per core item of the adjoint step pairs in use (0 when the adjoint takes single steps only), and the cells
 * of the shots' windows that those pairs advance two steps per launch (0 likewise). */
int fwi_plan_adj_pair_rows(const fwi_plan* plan);
long long fwi_plan_adj_core_cells(const fwi_plan* plan);
/* End row of the FP64 near-surface band (rows 4..end-1 of every shot) of the forward in use (0: none). */
int fwi_plan_surface_rows(const fwi_plan* plan);
/* Graph captures and replays so far (use_graph). */
int fwi_plan_graph_stats(const fwi_plan* plan, long long* captures, long long* replays);
/* Kernel nodes of the captured graphs: the largest graph so far, and all graphs of the last capture (a capture is
 * cut into consecutive graphs of at most 4096 time steps each). */
int fwi_plan_graph_nodes(const fwi_plan* plan, long long* largest, long long* last);

/* The backtracking line search of the reference (main.c:690-786) as a host state machine, shared by the drivers
 * (no device work).  fwi_line_search_init starts a stage: step lengths (v_hor, v_po, v_nmo) = 1, no failures, at
 * most 5.  After each trial model's misfit, fwi_line_search_step compares it with the iteration's misfit phi_n:
 * a NaN trial counts as 1e30 (written back) and as a failure; a trial above phi_n halves every step length and
 * counts as a failure; the step lengths and the failure count carry over to the stage's later iterations.
 * Returns 0 (accept: trial <= phi_n), 1 (try again with the halved step lengths), 2 (more than max_failures
 * failures in the stage: restore the backed-up model and end the stage), -1 on a null argument. */
typedef struct fwi_line_search {
  double alpha[3];
  int count;
  int max_failures;
} fwi_line_search;
void fwi_line_search_init(fwi_line_search* ls);
int fwi_line_search_step(fwi_line_search* ls, double phi_n, double* trial);

/* Host helpers. */
void fwi_cpml_table(int cpml, float vmax, float dt, float dh, float f, float d0, float L, float* a, float* b);
const char* fwi_error_message(void);

#ifdef __cplusplus
}
#endif

#endif /* FWI_H */

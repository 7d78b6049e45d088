/*
 * starmap_hip.h -- C ABI of the MI355X (gfx950) StaRMAP time-stepping core.
 *
 * The reference runs the whole solver inside one MATLAB function,
 * `output = starmap_solver(par)` (starmap_solver.m:1).  This library replaces
 * its time loop (starmap_solver.m:159-475); everything above the loop -- the
 * problem struct of starmap_init.m, the closures starmap_closure_pn.m /
 * starmap_closure_spn.m, the evaluation of the user's function handles on the
 * staggered grids -- stays with the caller (a MEX gateway, the Python mirror in
 * starmap_amd/solver.py, or C/C++), which hands plain arrays down.
 *
 * Conventions
 *   - Every call returns an SMP_* status; on failure smp_last_error() returns
 *     a message (thread-local).  No C++ exception crosses this boundary.
 *   - Host buffers belong to the caller and are only read/written during the
 *     call.  The handle owns all device memory.
 *   - A handle is not thread-safe.  Separate handles may run concurrently.
 *   - Component indices j are 0-based (the reference's j-1).
 *   - Field arrays use the reference's layout of U{j}: column-major over the
 *     solver axes (x fastest) with the component's own extents, i.e. n(d)
 *     cells on a cell-centred axis, n(d) faces (periodic) or n(d)+1 faces
 *     (extrapolation) on a staggered axis (starmap_solver.m:113-126).
 */
#ifndef STARMAP_HIP_H
#define STARMAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SMP_OK          0
#define SMP_E_INVALID   1   /* bad argument or descriptor                       */
#define SMP_E_HIP       2   /* a HIP runtime call failed                         */
#define SMP_E_OOM       3   /* device or host allocation failed                  */
#define SMP_E_STATE     4   /* call not valid in the handle's current state      */
#define SMP_E_UNSTABLE  5   /* non-finite values detected in the state (smp_step,
                               smp_interp, smp_check_finite); the reference
                               carries on with them                            */

/* state precision */
#define SMP_FP64 0
#define SMP_FP32 1

/* filter position, par.f_position (starmap_solver.m:62-66, :355, :417) */
#define SMP_FILTER_NONE    0
#define SMP_FILTER_SUBSTEP 1
#define SMP_FILTER_FULL    2

/* coefficient kinds for smp_set_coeff */
#define SMP_COEF_SIGMA_A1 0   /* sA1: decay rate of the zeroth moment, grid 111  */
#define SMP_COEF_SIGMA_T  1   /* sT1{gt,order}: decay rate of the other moments  */
#define SMP_COEF_SOURCE   2   /* Q{j}                                             */
#define SMP_COEF_FILTER   3   /* filterTerm{gt,f_order,slot}                      */

typedef struct smp_problem smp_problem;
typedef struct smp_batch smp_batch;

typedef struct smp_desc {
    int32_t n_sys;            /* number of moment components, length(par.Mz)    */
    int32_t n[3];             /* cells per solver axis, par.n after lifting     */
    int32_t bc[3];            /* 0 periodic, 1 extrapolation, par.bc            */
    double  h[3];             /* grid spacing, (ax_hi-ax_lo)./n                 */
    /* coupling matrices Mx, My, Mz in CSR over rows (targets); row_ptr has
     * n_sys+1 entries; columns ascending within a row.  A direction with no
     * coupling may pass NULL pointers.  Non-symmetric matrices are allowed. */
    const int32_t *row_ptr[3];
    const int32_t *col[3];
    const double  *val[3];
    const int32_t *mom_order; /* order used to look up sT (par.mom_order after
                                 the isotropic override of :68-70)              */
    const int32_t *f_order;   /* filter order per component, or NULL           */
    int32_t f_position;       /* SMP_FILTER_*                                   */
    int32_t precision;        /* SMP_FP64 or SMP_FP32                           */
    int32_t device;           /* HIP device ordinal; -1 creates a host-only
                                 handle (validation, grid types and schedules
                                 without touching a GPU; device calls fail
                                 with SMP_E_STATE)                             */
    int32_t tile[3];          /* cells per workgroup tile per kernel axis (the
                                 axes that carry cells, fastest first); zeros
                                 take the library default.  tile[0]+1 must be
                                 a power of two, the tile at most 256 cells   */
    int32_t blocks_per_cu;    /* LDS budget: workgroups per CU the schedule is
                                 made for (1..8); 0 takes SMP_BLOCKS_PER_CU
                                 from the environment, else 1                 */
    int32_t cell_waves;       /* waves updating the cells of a tile: 4, or 8
                                 (two per SIMD, each with half of a group's
                                 targets); 0 takes SMP_CELL_WAVES from the
                                 environment, else 8                          */
} smp_desc;

/* Last error message of the calling thread ("" if none). */
const char *smp_last_error(void);

/* Library version string. */
const char *smp_version(void);

/* Create a solver handle: validates the descriptor, computes the grid types
 * of every component (starmap_solver.m:76-98), builds the substep schedules,
 * allocates the state (zeroed) and uploads the coupling tables once.
 * Replaces the setup part of starmap_solver.m:58-157. */
int smp_create(const smp_desc *desc, smp_problem **out);

/* Release all device memory of the handle. */
int smp_destroy(smp_problem *h);

/* Number of moment components of the handle, length(par.Mz). */
int smp_num_components(const smp_problem *h, int32_t *n_sys);

/* Grid type (1 = cell centres, 2 = faces) of each component per solver axis,
 * gt[3*j+d] -- gtx/gty/gtz of starmap_solver.m:95-98. */
int smp_grid_types(const smp_problem *h, int32_t *gt);

/* Extents of component j per solver axis (the size of U{j}). */
int smp_component_shape(const smp_problem *h, int32_t j, int64_t shape[3]);

/* Copy component j in / out (count = product of its shape), FP64 on the host
 * whatever the state precision.  Replaces U{j} = ... (:140-146) and the
 * final read-out of U (:484-499). */
int smp_set_state(smp_problem *h, int32_t j, const double *src, int64_t count);
int smp_get_state(smp_problem *h, int32_t j, double *dst, int64_t count);

/* Set a coefficient (starmap_solver.m:214-252, :278-305).
 *   SIGMA_A1: gt, order, slot ignored
 *   SIGMA_T : gt[3] and order select sT1{gt,order}
 *   SOURCE  : order = component j; field NULL clears the source
 *   FILTER  : gt[3], order = filter order, slot 1/2 ('substep') or 0 ('full')
 * field == NULL means the value is the scalar; otherwise field holds count
 * values on the grid gt (for SOURCE: on the grid of component j).
 * The decay factors E = expm1div(-sigma*dt/2) (:267-274) are recomputed on
 * the device by the next smp_step whenever a sigma or dt changed. */
int smp_set_coeff(smp_problem *h, int32_t kind, const int32_t gt[3], int32_t order,
                  int32_t slot, const double *field, double scalar, int64_t count);

/* Advance n_steps full time steps of size dt: the three Strang substeps
 * [1 2 1] and the optional full filter (starmap_solver.m:307-421).  Kernels
 * are queued on the handle's stream; the call returns after they finish.
 * Once at least 64 steps (SMP_CHECK_EVERY in the environment at smp_create;
 * 0 turns it off) have run since the last scan, the call ends with one pass
 * over the state that returns SMP_E_UNSTABLE if any value is NaN or
 * infinite -- the monitor of a time step beyond the stability limit, which
 * the reference runs through silently. */
int smp_step(smp_problem *h, double dt, int64_t n_steps);

/* Scan the whole state once: *nonfinite = 1 and SMP_E_UNSTABLE if any value
 * is NaN or infinite, else 0 and SMP_OK.  The Python mirror calls it after
 * the last step of a solve.  No counterpart in the reference. */
int smp_check_finite(smp_problem *h, int32_t *nonfinite);

/* Output accumulation (starmap_solver.m:308, :425-428):
 * smp_snapshot keeps U0 = U(comps) on the device; smp_interp writes
 * (1-lambda)*U0 + lambda*U(comps) to the host, components concatenated in the
 * reference layout (count = total elements), and returns SMP_E_UNSTABLE
 * (after writing) if a written value is NaN or infinite. */
int smp_snapshot(smp_problem *h, const int32_t *comps, int32_t n_comps);
int smp_interp(smp_problem *h, double lambda, double *out, int64_t count);

/* FP64 reductions over the current state, with wavefront-shuffle partial sums
 * per workgroup: sums[j] = sum of U{j}; sq[j] = sum over cells of w*U{j}^2
 * with the staggered trapezoid weights of starmap_2d_ex_l2norm.m:93-99
 * (times prod(h)).  Either pointer may be NULL.  Also reports whether any
 * value is non-finite (*nonfinite = 1). */
int smp_observables(smp_problem *h, double *sums, double *sq, int32_t *nonfinite);

/* Timing of the last smp_step call: elapsed milliseconds of all its kernels
 * measured with HIP events on the handle's stream; the time of the substep
 * kernels alone, their launch count and their algorithmic HBM bytes (each
 * staged source component read once, each target component read and written
 * once).  The substep launches run back to back, and their time is the
 * elapsed time; with the 'full' filter, whose scale kernels run between
 * them, it is the sum of an event pair around each substep launch (the
 * first 12288). */
int smp_last_step_timing(const smp_problem *h, double *total_ms, double *substep_ms,
                         int64_t *substep_launches, double *substep_bytes);

/* The HIP stream (hipStream_t) the handle launches on. */
void *smp_stream(const smp_problem *h);

/* Device bytes the handle allocated. */
int64_t smp_device_bytes(const smp_problem *h);

/* Schedule statistics for inspection and tests: for substep kind 0 (odd
 * targets) and 1 (even targets): number of groups, LDS slots, source tile
 * loads per tile and distinct sources. */
int smp_plan_stats(const smp_problem *h, int32_t kind, int32_t *n_groups, int32_t *n_slots,
                   int32_t *n_loads, int32_t *n_sources);

/* ---- schedule (no counterpart in the reference) ------------------------ */

/* Re-make the substep schedules of the handle for another tile shape (cells
 * per workgroup per kernel axis; NULL or zeros: library default), LDS budget
 * (workgroups per CU, 1..8; 0 keeps the current one) and cell-wave count (4
 * or 8; 0 keeps the current one).  The state and coefficients are kept; on
 * failure the previous schedule stays in force.  Used by the autotuner. */
int smp_set_schedule(smp_problem *h, const int32_t *tile, int32_t blocks_per_cu,
                     int32_t cell_waves);
int smp_get_schedule(const smp_problem *h, int32_t *tile, int32_t *blocks_per_cu,
                     int32_t *cell_waves);

/* ---- state checkpoints in device memory (no counterpart in the reference) */

/* Copy the whole state into device checkpoint slot `slot` (0..4096,
 * allocated on first use, kept until smp_checkpoint_free or smp_destroy), or
 * back.  Used for recomputation between checkpoints and by the autotuner.
 * smp_checkpoint_free releases one slot, or every slot for slot = -1. */
int smp_checkpoint_save(smp_problem *h, int32_t slot);
int smp_checkpoint_load(smp_problem *h, int32_t slot);
int smp_checkpoint_free(smp_problem *h, int32_t slot);

/* ---- batched stepping (no counterpart in the reference) ----------------- */

/* Group n handles (same device and precision; no filter or the substep
 * filter) for batched stepping.  The handles stay owned by the caller and
 * must outlive the batch; each keeps its own state, coefficients and
 * schedule. */
int smp_batch_create(smp_problem *const *handles, int32_t n, smp_batch **out);
int smp_batch_destroy(smp_batch *b);

/* Advance problem i by n_steps[i] time steps of size dt[i] -- the same
 * arithmetic as smp_step(handles[i], dt[i], n_steps[i]), bit for bit -- with
 * all problems in lockstep: one persistent launch per substep position runs
 * the substeps of every problem that has steps left, its workgroups taking
 * (problem, tile) units from a counter, heaviest problems first.
 * use_graph != 0 captures the launches in a hipGraph and replays it; the
 * graph is kept and replayed by later calls with the same launch structure. */
int smp_batch_run(smp_batch *b, const double *dt, const int64_t *n_steps, int32_t use_graph);

/* Timing of the last smp_batch_run: elapsed milliseconds (HIP events on the
 * batch's stream), the sum of the per-launch times (equal to the elapsed time
 * when a graph ran), launches, algorithmic bytes of all launches (each
 * staged source read once, each target read and written once), moment-cell
 * updates, and the persistent grid size. */
int smp_batch_timing(const smp_batch *b, double *total_ms, double *kernel_ms, int64_t *launches,
                     double *bytes, int64_t *mcu, int32_t *grid);

/* ---- discrete adjoint (no counterpart in the reference) ----------------- */

/* Zero the adjoint state lambda and the gradient accumulators. */
int smp_adjoint_reset(smp_problem *h);

/* Set component j of the adjoint state lambda (reference layout, as
 * smp_set_state): the seed dJ/dU(T) of an objective J of the final state. */
int smp_adjoint_seed(smp_problem *h, int32_t j, const double *src, int64_t count);

/* Discrete adjoint of n_steps time steps of size dt from the current state
 * (the arithmetic of smp_step(h, dt, n_steps), starmap_solver.m:307-421):
 * runs the steps forward -- the state ends at the final time, as after
 * smp_step -- keeping a copy of the state every checkpoint_every steps, then
 * sweeps back from the seeded lambda(T), recomputing each segment from its
 * checkpoint.  Afterwards lambda holds dJ/dU(0) and, accumulated in FP64
 * over the steps, dJ/dsigma_a and dJ/dsigma_s0 at every point of every
 * staggered grid type (sigma_t = sigma_a + sigma_s0 - sigma_sm for j > 1,
 * sigma_a for j = 1, starmap_solver.m:214-252).  Filters must be off or at
 * 'substep' position and constant in space; sources must be fields; each
 * component slab must be below 2 GiB.  The recomputation runs on a second
 * stream of the handle (default priority); the call returns after both
 * streams are done. */
int smp_adjoint(smp_problem *h, double dt, int64_t n_steps, int32_t checkpoint_every);

/* Component j of the adjoint state (what = 0), or dJ/dsigma_a (what = 1) or
 * dJ/dsigma_s0 (what = 2) on the grid type of component j; reference
 * layout. */
int smp_adjoint_get(smp_problem *h, int32_t what, int32_t j, double *dst, int64_t count);

/* Milliseconds of the last smp_adjoint: forward pass and reverse sweep
 * (HIP events on the handle's stream). */
int smp_adjoint_timing(const smp_problem *h, double *forward_ms, double *reverse_ms);

/* Device memory smp_adjoint may hold for checkpoints and recorded launch
 * outputs, in bytes (0, the default: a checkpoint at the start of every
 * segment of checkpoint_every steps and the records of two segments).  Under
 * a budget the schedule records the last segments during the forward pass
 * (they are never recomputed) and keeps checkpoints for the others; when
 * fewer checkpoints fit than there are segments left, it places them by
 * binomial checkpointing and re-advances segments from them.  A budget
 * below what one segment needs fails with SMP_E_OOM. */
int smp_adjoint_set_memory(smp_problem *h, int64_t bytes);

/* The schedule of the last smp_adjoint: stats[0..7] = segments, steps per
 * segment, segments recorded in the forward pass, checkpoint slots,
 * segments recomputed, further segment advances from checkpoints,
 * state-sized buffers and bytes at the peak. */
int smp_adjoint_stats(const smp_problem *h, int64_t stats[8]);

/* Host-only planner behind smp_adjoint: the schedule of n_steps steps in
 * segments of `every` steps within `buffers` state-sized buffers (-1: no
 * budget), a step's records taking rec_factor buffers (1: the launches'
 * states; 2: and the W+Q the launches stored, which smp_adjoint records
 * unless coefficients are read from global memory).  stats as
 * smp_adjoint_stats (bytes 0).  If actions is not null it
 * receives 6 int32 per action {kind, first segment, end segment, source
 * buffer, destination buffer, record range}: the forward pass (*n_fwd
 * actions), then the reverse sweep (*n_actions in all).  Kinds: 0 run in
 * place, 1 copy the state to a buffer, 2 record a segment, 3 final state from
 * records, 4 copy and re-advance, 5 recompute into records, 6 reverse.
 * Buffers: >= 0 checkpoint slots, -2 scratch state, -3 the state at the
 * start of the recorded segments, -4 the previous segment's records. */
int smp_adjoint_schedule(int64_t n_steps, int32_t every, int64_t buffers, int32_t rec_factor,
                         int64_t stats[8], int32_t *actions, int64_t cap, int64_t *n_fwd,
                         int64_t *n_actions);

/* Moment norms of the current state, per component j in the reference's
 * output layout: sumsq[j] = sum of U_j^2 over its grid points, diffsq[3*j+e]
 * = sum of squared differences of neighbours along output dimension e+1
 * (MATLAB diff(U,[],e+1)); the terms of the moment-order norms of
 * starmap_2d_ex_lattice_filter_convergence.m:66-76 (grouping by moment order
 * and the square roots stay with the caller).  diffsq has 3*n_sys entries,
 * zero for dimensions the problem does not have. */
int smp_moment_norms(smp_problem *h, double *sumsq, double *diffsq);

/* Every component linearly interpolated at one point on its own staggered
 * grid, as the interp2 loop of starmap_2d_ex_linesource_filter_reconstruction.m
 * :118-121 (NaN outside the grid).  lo and point have 3 entries in the
 * solver's axis order of smp_desc.n (lo: the lower domain corner
 * par.ax(1:2:end)); out has n_sys entries. */
int smp_probe(smp_problem *h, const double *lo, const double *point, double *out);

/* ---- allocation-free variant (no counterpart in the reference) --------- */

/* Largest n_sys smp_probe_noalloc accepts. */
#define SMP_PROBE_NOALLOC_MAX_COMPS 4096

/* smp_probe without any device allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving or
 * monitoring loops, harnesses that forbid hipMalloc): the per-call component
 * table and values live in one fixed device array of the library (about
 * 590 KB per GPU, held until the process ends) instead of the two buffers
 * smp_probe allocates and frees on every call.  Same arguments, results and
 * errors as smp_probe; SMP_E_INVALID for a problem with more than
 * SMP_PROBE_NOALLOC_MAX_COMPS components.  Every call on one GPU shares that
 * array, so calls that overlap on the same GPU (from several threads) race;
 * the call is synchronous, so calls in sequence do not. */
int smp_probe_noalloc(smp_problem *h, const double *lo, const double *point, double *out);

#ifdef __cplusplus
}
#endif

#endif /* STARMAP_HIP_H */

/*
 * zamr — C-ABI of the MI355X per-box RHS / time-advance library.
 *
 * One shared library is built per generated system (Model.H + Numerics.H +
 * MarchConstants.H, exactly what zoomy_amrex's printers emit), the way the
 * reference compiles one executable per system (zoomy_amrex/run_case.py:
 * 210-289, 390-398).  Every entry point returns 0 on success and a non-zero
 * ZAMR_E* code on failure; zamr_last_error() gives the message.  No exception
 * crosses the ABI.  A context is driven by one host thread; device work is
 * asynchronous on the context's stream and entry points that return host
 * data synchronise it.
 *
 * Array convention: dense float64 arrays of shape (ncomp, ny, nx), i fastest
 * — the raster order of Source/init_solution.cpp:63 (idx = i + nx*j).
 *
 * The reference has no C-ABI; its in-process analogue is
 *   ZoomyAmr solver; solver.InitData(); solver.Evolve();   (Source/main.cpp:23-39)
 * and each entry below names the member function it replaces.
 */
#ifndef ZAMR_H
#define ZAMR_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zamr_ctx zamr_ctx;

enum {
    ZAMR_OK = 0,
    ZAMR_EINVAL = 1,    /* bad argument / deck value */
    ZAMR_ETAG = 2,      /* boundary tag not declared by the model (ZoomyAmr.cpp:361-372) */
    ZAMR_ERASTER = 3,   /* raster missing or too small (init_solution.cpp:20-40) */
    ZAMR_EHIP = 4,      /* HIP runtime failure */
    ZAMR_EUNSUP = 5,    /* a deck option outside this build (e.g. amr.max_level > 3) */
    ZAMR_EIO = 6        /* plotfile / file write failure */
};

/* One box of an independent box list (batched mode).  Each box is its own
 * physical domain: lo/hi are its valid cells, dx/prob_lo its geometry, group
 * the time-step group it shares dt with (an AMR level or an ensemble member),
 * bc the boundary tag names of its x_lo, x_hi, y_lo, y_hi sides ("" = tag 0). */
typedef struct zamr_box {
    int lo[2], hi[2];
    double dx[2], prob_lo[2];
    int group;
    const char* bc[4];
} zamr_box;

/* Model introspection: Model::n_dof_q, n_dof_qaux, n_parameters, dimension,
 * n_boundary_tags (to_amrex.py:348-360). */
int zamr_model_info(int* n_dof_q, int* n_dof_qaux, int* n_parameters, int* dimension, int* n_tags);
/* Model::get_boundary_tags()[k] / parameter_names()[k] / default_parameters()[k]. */
int zamr_tag_name(int k, char* buf, int cap);
int zamr_param_name(int k, char* buf, int cap);
int zamr_default_param(int k, double* value);
/* ZoomyConstants::c_cfl_dimension and c_cfl_degree_factor compiled in (_constants.py:59-88). */
int zamr_march_constants(int* c_cfl_dimension, int* c_cfl_degree_factor);

/* ZoomyAmr::ZoomyAmr() (ZoomyAmr.cpp:27-107): reads the ParmParse deck text
 * written by build.py:write_inputs (build.py:55-120).  amr.max_level = 1..3
 * sets up the hierarchy (AmrCore with amr.ref_ratio 2 or 4, a 2-D model; the
 * amr.blocking_factor / n_error_buf / n_proper / grid_eff / max_grid_size and
 * tagging.* keys apply).  solver.implicit_global = true solves the source globally
 * after each stage (Source/imex_solver.H: Newton + GMRES; with MOOD and on every
 * AMR level too).  Keys outside this build (amr.max_level > 3, ref_ratio not 2
 * or 4, implicit_global in a box list) fail with ZAMR_EUNSUP. */
int zamr_create(const char* inputs_text, zamr_ctx** out);
/* Batched mode: nbox independent boxes sharing one launch per RK stage.
 * inputs_text supplies the solver.* / params.* keys (geometry keys ignored). */
int zamr_create_boxlist(const zamr_box* boxes, int nbox, const char* inputs_text, zamr_ctx** out);
int zamr_destroy(zamr_ctx* ctx);
const char* zamr_last_error(const zamr_ctx* ctx);

/* ZoomyAmr::InitData (ZoomyAmr.cpp:112-156): init_solution on Q and Qaux
 * (init_solution.cpp:78-115), then the init.* rasters, then (AMR) the level
 * loop of :137-155 that tags level 0 and fills level 1 from it. */
int zamr_init_data(zamr_ctx* ctx);

/* Geometry of a level: its domain in its own index space and its box count
 * (in batched mode, box `lev`). */
int zamr_level_shape(const zamr_ctx* ctx, int lev, int* nx, int* ny, int* nboxes);

/* Whole-level state copy in/out (field 0 = Q, 1 = Qaux), dense (ncomp, ny, nx)
 * over the level's domain; on a refined level only the cells its boxes cover
 * are read or written (the rest of the array is left as it is).
 * In batched mode lev is the box index. n = number of doubles in host. */
int zamr_set_state(zamr_ctx* ctx, int lev, int field, const double* host, size_t n);
int zamr_get_state(zamr_ctx* ctx, int lev, int field, double* host, size_t n);

/* ZoomyAmr::ComputeDt (ZoomyAmr.cpp:448-515) over the state and aux as stored. */
int zamr_compute_dt(zamr_ctx* ctx, int lev, double* dt);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): zamr_compute_dt's step size, written to the device double
 * dt_device on the context's stream, with no allocation and no host
 * synchronisation.  Its CFL reduction runs in a fixed scratch array of this
 * library on the current GPU (one per GPU, held until the process ends):
 * contexts with more than 4096 time-step groups are rejected (ZAMR_EINVAL),
 * and calls that overlap on one GPU race on that array. */
int zamr_compute_dt_noalloc(zamr_ctx* ctx, int lev, double* dt_device);
/* ZoomyAmr::UpdateState (ZoomyAmr.cpp:293-334). */
int zamr_update_state(zamr_ctx* ctx, int lev, double time);
/* ZoomyAmr::FillPatch (ZoomyAmr.cpp:258-277, single level): same-level ghost
 * exchange plus physical boundary conditions, written into the ghost cells. */
int zamr_fill_patch(zamr_ctx* ctx, int lev, double time);
/* One RK stage, do_stage(order) of ZoomyAmr::Advance (ZoomyAmr.cpp:561-662),
 * applied to the current state in place. */
int zamr_stage(zamr_ctx* ctx, int lev, int order, double time, double dt);
/* ZoomyAmr::Advance (ZoomyAmr.cpp:538-731) with a caller-given dt. */
int zamr_advance(zamr_ctx* ctx, int lev, double time, double dt);
/* ZoomyAmr::TimeStep (ZoomyAmr.cpp:733-762): (AMR: regrid when iteration > 0)
 * ComputeDt, time_end cap, Advance (AMR: then ref_ratio substeps of the next
 * level, each the same TimeStep one level up, and average_down).  On return *dt_taken holds the level-0 step size. */
int zamr_time_step(zamr_ctx* ctx, int lev, double time, int iteration, double* dt_taken);
/* Up to nsteps iterations of the Evolve loop body (ZoomyAmr.cpp:782-810) with
 * dt computed and applied on the device; steps are captured in a hipGraph.
 * Stops early when time_end or max_step is reached. */
int zamr_run_steps(zamr_ctx* ctx, int nsteps, int* steps_done, double* time);
/* ZoomyAmr::Evolve (ZoomyAmr.cpp:767-816) including plotfiles and the
 * "Initial total mass" / "Step" / "Total evolution time" lines on stdout. */
int zamr_evolve(zamr_ctx* ctx);
/* ZoomyAmr::ComputeTotalMass (ZoomyAmr.cpp:517-536): every level, level-0
 * cells under level 1 left out. */
int zamr_total_mass(zamr_ctx* ctx, double* mass);
/* Time, step count and last dt of level 0 (of group `lev` in batched mode). */
int zamr_get_time(const zamr_ctx* ctx, int lev, double* time, int* step, double* dt);
/* Reset time/step of every group to zero (a fresh Evolve on the current state). */
int zamr_reset_clock(zamr_ctx* ctx);

/* ZoomyAmr::WritePlotFile (ZoomyAmr.cpp:821-882): AMReX plotfile
 * plt_<identifier><step,5 digits> under dir (NULL = current directory). */
int zamr_write_plotfile(zamr_ctx* ctx, const char* dir, int step, double time);

/* Raw device pointer of a field of box `box` of level `lev` (field 0 = Q,
 * 1 = Qaux), the grown-box corner of an AMReX Array4 layout, with its strides
 * in elements, and the context's HIP stream (for the PyTorch wrapper). */
int zamr_device_ptr(zamr_ctx* ctx, int lev, int box, int field, void** ptr, long long* row_stride,
                    long long* comp_stride, int* ngrow, void** stream);
/* Run on a caller-owned HIP stream (hipStream_t) instead of the context's own. */
int zamr_set_stream(zamr_ctx* ctx, void* stream);

/* Tuning: rows per work item for the stage kernels in every box (0 = per box
 * size class: the table zamr_autotune measured, else an analytic rule over
 * the whole box list: the deepest of 32/16/8/4 rows that still gives every
 * SIMD two waves), per-box launches instead of one batched launch (baseline
 * for the batching speed-up), and XCD-aware item order.  No counterpart in
 * the reference, whose launch shape is AMReX's ParallelFor. */
int zamr_set_tuning(zamr_ctx* ctx, int rows_per_item, int per_box_launch, int xcd_remap);
/* Measured launch-shape tuner (single-level and box-list contexts; AMR gets
 * ZAMR_EUNSUP).  Times the stage kernels of a step on the current state with
 * HIP events (the shortest of `reps` launches per candidate) and keeps the
 * fastest rows per work item in {4, 8, 16, 32, 64} for each box size class
 * present (the larger side <= 32, 64, 128, 256, 512, 1024, 2048, larger: 8
 * classes) and the fastest grid (one block per four-item slot, or 1x / 2x /
 * 4x the blocks the device holds at once, walking the slots) of the
 * first-stage and of the step-ending stage kernel.  Writes the table to class_rows[8] (0 = class
 * absent or the analytic rule kept), the grids to grid_blocks[2] (0 = one
 * block per slot) and the stage-kernel time per step (ms) before and after
 * to ms[2]; any may be NULL.  The state is left as it was, and no result
 * depends on the shape chosen.  No counterpart in the reference. */
int zamr_autotune(zamr_ctx* ctx, int reps, int* class_rows, int* grid_blocks, double* ms);
/* Set the launch shape zamr_autotune would set (a table measured earlier on
 * the same box list): class_rows[8], grid_blocks[2]. */
int zamr_set_launch_shape(zamr_ctx* ctx, const int* class_rows, const int* grid_blocks);
/* Capture chunks of steps of zamr_run_steps into hipGraphs (on = 1, the
 * default) or launch every kernel eagerly (on = 0). */
int zamr_set_graph(zamr_ctx* ctx, int on);
/* Average device time (ms) per launch of each stage kernel over the last
 * zamr_run_steps call with profiling on: out[0] stage 1, out[1] stage 2,
 * out[2] per-step bookkeeping; returns the number of timed steps in *nsteps. */
int zamr_set_profiling(zamr_ctx* ctx, int on);
int zamr_kernel_times(zamr_ctx* ctx, double* out_ms, int* nsteps);

/* solver.positivity = mood (ZoomyAmr.cpp:699-720): troubled cells of the last
 * step's order-2 candidate in time-step group `group` (0 for a single level,
 * the level index with AMR) and the number of steps that group has redone at
 * order 1 since creation. */
int zamr_mood_stats(zamr_ctx* ctx, int group, int* troubled, long long* redos);

/* ---- AMR hierarchy (amr.max_level = 1..3) ------------------------------- */
/* finest_level (AmrCore): the highest level that has boxes. */
int zamr_finest_level(zamr_ctx* ctx, int* finest);
/* BoxArray of level lev (boxArray(lev)) as (lo0, lo1, hi0, hi1) quadruples in
 * the level's index space, up to cap boxes; *nboxes receives the count. */
int zamr_level_boxes(zamr_ctx* ctx, int lev, int* lohi, int cap, int* nboxes);
/* The level loop of InitData (ZoomyAmr.cpp:137-155) on the current level-0
 * state, for a state set with zamr_set_state after zamr_init_data. */
int zamr_init_levels(zamr_ctx* ctx);
/* AmrCore::regrid(0, time) (ZoomyAmr.cpp:736): ErrorEst from the finest
 * level that may tag down to level 0, new grids for every level above 0
 * (proper nesting), RemakeLevel / MakeNewLevelFromCoarse / ClearLevel. */
int zamr_regrid(zamr_ctx* ctx, double time);
/* ZoomyAmr::ErrorEst (ZoomyAmr.cpp:225-253) on level 0 as it stands: one byte
 * per cell (1 = tagged), raster order, n = nx * ny. */
int zamr_error_est(zamr_ctx* ctx, unsigned char* tags, size_t n);
/* Regrids run, levels remade, and the host time spent regridding (ms), since
 * creation. */
int zamr_amr_stats(zamr_ctx* ctx, long long* regrids, long long* remakes, double* regrid_ms);

/* solver.implicit_global (Source/imex_solver.H): global source solves run,
 * and the Newton and GMRES iterations they took, since creation. */
int zamr_jfnk_stats(zamr_ctx* ctx, long long* solves, long long* newton, long long* gmres);

/* (dt, time after the step) of the steps taken on the device clock since the
 * last zamr_reset_clock, up to cap entries — the values the reference prints
 * per step (ZoomyAmr.cpp:804-809).  *n receives the count. */
int zamr_dt_history(zamr_ctx* ctx, int cap, double* dts, double* times, int* n);

/* ---- Adjoint (gradients) -------------------------------------------------
 * No counterpart in the reference, which has no gradients.  Systems that ship
 * an Adjoint.H beside their Model.H (swe2d_nf: per-cell Manning field in aux
 * row 1) support these; other systems return ZAMR_EUNSUP.  Scope: single-level
 * decks (any number of boxes; the adjoint runs the domain as one box, whose
 * forward is the boxes' own bit for bit), order 2, point-implicit source,
 * with or without the well-balanced (hydrostatic) reconstruction and
 * UpdateState's positivity clamp; MOOD, solver.implicit_global, AMR and box
 * lists give ZAMR_EUNSUP.
 *
 * zamr_adj_setup: nsteps SSP-RK2 steps with the frozen step sizes dts[nsteps]
 * (time t_k = dts[0] + ... + dts[k-1]), a checkpoint every seg_steps steps,
 * and ngauges distinct gauge cells (gi[g], gj[g]) whose depth h is recorded
 * after every step.
 * zamr_adj_forward: advances the current state by the nsteps frozen steps
 * (the solver clock is not moved) and writes the gauge depths to
 * gauge_h[k * ngauges + g] (after step k), if gauge_h is not NULL.
 * zamr_adj_backward: given the adjoint of every gauge depth (same layout),
 * recomputes each segment from its checkpoint and runs the reverse stages.
 * zamr_adj_gradient: the adjoint of the initial state (n_dof_q * ny * nx)
 * and of the friction aux row (ny * nx), raster order; either may be NULL. */
int zamr_adj_setup(zamr_ctx* ctx, int nsteps, int seg_steps, const double* dts, int ngauges, const int* gi,
                   const int* gj);
int zamr_adj_forward(zamr_ctx* ctx, double* gauge_h);
int zamr_adj_backward(zamr_ctx* ctx, const double* gauge_hbar);
int zamr_adj_gradient(zamr_ctx* ctx, double* qbar0, size_t n_qbar, double* auxbar, size_t n_auxbar);
/* Frozen limiter, for derivative checks: with on = 1 the next
 * zamr_adj_forward records which argument every minmod of every face and
 * stage returned, and later forwards replay those choices instead of
 * comparing, so that finite differences see the same one-sided derivative
 * the adjoint takes at slope-limiter switching points.  on = 0 turns it off.
 * A backward must follow a forward that computed its limiter choices. */
int zamr_adj_freeze_limiter(zamr_ctx* ctx, int on);
/* Device-resident adjoint I/O, for callers that keep their tensors on the GPU
 * (the PyTorch wrapper).  With on = 1, zamr_adj_forward's gauge_h,
 * zamr_adj_backward's gauge_hbar and zamr_adj_gradient's qbar0 / auxbar are
 * device pointers (same layouts), and the three calls queue their kernels
 * and copies on the context's stream (zamr_set_stream) and return without
 * waiting for it.  The caller writes the initial state and the friction row
 * in place through zamr_device_ptr (fields 0 and 1, valid cells) on that
 * stream, and reads results in stream order.  on = 0 restores host arrays and
 * waiting calls.  No counterpart in the reference. */
int zamr_adj_device_io(zamr_ctx* ctx, int on);

#ifdef __cplusplus
}
#endif

#endif /* ZAMR_H */

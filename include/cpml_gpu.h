/*
 * cpml_gpu -- C ABI of the MI355X (gfx950) C-PML time-loop library.
 *
 * The reference programs (geodynamics/seismic_cpml) have no library API: the
 * time loop is inline in each Fortran `program`, configured by compile-time
 * `parameter`s.  These entry points replace those inner loops; their argument
 * names are the reference's own (NX, NY, NZ, DELTAX, DELTAT, NPOINTS_PML,
 * ISOURCE, JSOURCE, ix_rec, iy_rec, sisvx, ...).  The Fortran module
 * `cpml_gpu` (seismic_cpml_amd/fortran/cpml_gpu.f90) binds them with
 * ISO_C_BINDING.
 *
 * Conventions
 *  - All functions return a status code (CPML_OK == 0).  No C++ exception
 *    crosses the ABI.  cpml_gpu_last_error() returns the message.
 *  - Grid arrays are passed in Fortran order (x fastest), interior points only
 *    (NX*NY or NX*NY*NZ values) unless cpml_gpu_set_host_layout declares
 *    ghost cells, in the handle's precision (prec = 4: float, 8: double).
 *    Host arrays are borrowed for the duration of the call.
 *  - Indices (ISOURCE, ix_rec, ...) are 1-based as in the reference.
 *  - A handle is single-threaded; it owns one HIP stream, its device buffers
 *    and its captured step graphs.  Distinct handles may run concurrently.
 */
#ifndef CPML_GPU_H
#define CPML_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cpml_solver cpml_solver;

/* status codes */
#define CPML_OK 0
#define CPML_ERR_ARG 1      /* invalid argument */
#define CPML_ERR_CFL 2      /* Courant number > 1 (2D2:513-516, 3D:712-717) */
#define CPML_ERR_STENCIL 3  /* grid too small for the stencil / slab layout */
#define CPML_ERR_OOM 4      /* device memory budget exceeded */
#define CPML_ERR_HIP 5      /* HIP runtime error */
#define CPML_ERR_BLOWUP 6   /* max |v| above threshold (2D2:716-726, 3D:1183-1195) */
#define CPML_ERR_STATE 7    /* call out of order (e.g. run before set_medium) */
#define CPML_ERR_NODEVICE 8 /* no GPU visible */

/* field / memory-variable selectors for get_field / set_field / get_memory */
enum {
  CPML_VX = 0, CPML_VY, CPML_VZ, CPML_SIGMAXX, CPML_SIGMAYY, CPML_SIGMAZZ,
  CPML_SIGMAXY, CPML_SIGMAXZ, CPML_SIGMAYZ
};
/* memory variables, named as in the reference (3D:222-240) */
enum {
  CPML_MEM_DVX_DX = 0, CPML_MEM_DVX_DY, CPML_MEM_DVX_DZ, CPML_MEM_DVY_DX, CPML_MEM_DVY_DY,
  CPML_MEM_DVY_DZ, CPML_MEM_DVZ_DX, CPML_MEM_DVZ_DY, CPML_MEM_DVZ_DZ, CPML_MEM_DSIGMAXX_DX,
  CPML_MEM_DSIGMAYY_DY, CPML_MEM_DSIGMAZZ_DZ, CPML_MEM_DSIGMAXY_DX, CPML_MEM_DSIGMAXY_DY,
  CPML_MEM_DSIGMAXZ_DX, CPML_MEM_DSIGMAXZ_DZ, CPML_MEM_DSIGMAYZ_DY, CPML_MEM_DSIGMAYZ_DZ
};
/* kernels timed by cpml_gpu_kernel_time (the fused 3D step counts as STRESS) */
enum { CPML_KERNEL_STRESS = 0, CPML_KERNEL_VELOCITY = 1 };

/* ---- lifecycle ------------------------------------------------------- */

/* Replaces the L1 parameter block of 2D2:137-218 / 2D4:131-219 / 3D:123-218
 * (/ 3D-visco:152-245 = seismic_CPML_3D_viscoelastic_MPI.f90 for 3D order 4).
 * ndim 2 (NZ ignored) or 3; order 2, 4 or 8 in 2D (2D2 / 2D4 / Holberg
 * operator of ADE8:263-266), 2 or 4 in 3D (order 4: the elastic update of
 * 3D-visco:977-1398 on one domain, with that program's loop ranges, Lame
 * coefficients and two-plane rigid faces; the source goes at KSOURCE as
 * there at the cut plane); prec 4 or 8.  NSTEP sizes the source and
 * seismogram tables.  Validates everything on the host before any HIP call;
 * returns CPML_ERR_NODEVICE when no GPU is visible. */
int cpml_gpu_init(cpml_solver **h, int ndim, int order, int prec, int NX, int NY, int NZ,
                  double DELTAX, double DELTAY, double DELTAZ, double DELTAT, int NPOINTS_PML,
                  int NSTEP);
int cpml_gpu_finalize(cpml_solver *h);

/* ---- set-up (replaces L2 uploads) ------------------------------------ */

/* Lame parameters and density per point: 2D2:204, 468-474 (arrays);
 * 3D:139-145 (scalars in the reference; arrays here).  Checks rho, mu > 0 and
 * the Courant number with cp = max sqrt((lambda+2mu)/rho) (2D2:513-516). */
int cpml_gpu_set_medium(cpml_solver *h, const void *lambda, const void *mu, const void *rho);
/* Same from device arrays (interior-only Fortran order, handle precision). */
int cpml_gpu_set_medium_device(cpml_solver *h, const void *d_lambda, const void *d_mu,
                               const void *d_rho);

/* Damping profiles a, b, K at integer and half points per axis, as built by
 * 2D2:283-465 / 3D:399-667 (z arrays ignored in 2D).  The absorbing slabs are
 * the points where a or a_half is non-zero or K or K_half differs from 1. */
int cpml_gpu_set_pml(cpml_solver *h, const void *a_x, const void *b_x, const void *K_x,
                     const void *a_x_half, const void *b_x_half, const void *K_x_half,
                     const void *a_y, const void *b_y, const void *K_y, const void *a_y_half,
                     const void *b_y_half, const void *K_y_half, const void *a_z,
                     const void *b_z, const void *K_z, const void *a_z_half,
                     const void *b_z_half, const void *K_z_half);

/* Point force at (ISOURCE, JSOURCE[, KSOURCE]); force_x/force_y[it-1] is the
 * force at step it (2D2:644-667, 3D:1055-1083).  ISOURCE <= 0 removes it. */
int cpml_gpu_set_source(cpml_solver *h, int ISOURCE, int JSOURCE, int KSOURCE,
                        const void *force_x, const void *force_y);

/* Receivers at grid points (2D2:683-686, 3D:1123-1129); iz_rec ignored in 2D. */
int cpml_gpu_set_receivers(cpml_solver *h, int NREC, const int *ix_rec, const int *iy_rec,
                           const int *iz_rec);

/* Energy curve every `cadence` steps (0 = off) over the window
 * window[0..5] = i0,i1,j0,j1,k0,k1 (1-based, inclusive; NULL = the
 * reference's window: 2D2:695 / 2D4:696 / 3D:1137-1139).  flags bit 0: use the
 * textbook 3D potential energy instead of the reference formula, which counts
 * eps_yy*sigma_yy twice and omits eps_zz*sigma_zz (3D:1170-1172). */
int cpml_gpu_set_energy(cpml_solver *h, int cadence, const int *window, int flags);

/* Stability check inside run(): at it % IT_DISPLAY == 0 and it == 5 the max
 * norm of v is compared with threshold (reference: 1e25); IT_DISPLAY <= 0
 * disables it. */
int cpml_gpu_set_stability(cpml_solver *h, int IT_DISPLAY, double threshold);

/* ---- time loop ------------------------------------------------------- */

/* Runs steps it_begin..it_end (1-based, inclusive) of `do it = 1,NSTEP`
 * (2D2:550-735, 2D4:551-736, 3D:802-1245): stress + C-PML, velocity +
 * C-PML, source, Dirichlet edges, receivers, energy.  Blocks until done. */
int cpml_gpu_run(cpml_solver *h, int it_begin, int it_end);

/* ---- results --------------------------------------------------------- */

/* sisvx(NSTEP,NREC) etc. in Fortran order; sisvz may be NULL. */
int cpml_gpu_get_seismograms(cpml_solver *h, void *sisvx, void *sisvy, void *sisvz);
/* total_energy_kinetic(NSTEP), total_energy_potential(NSTEP) (double). */
int cpml_gpu_get_energy(cpml_solver *h, double *kinetic, double *potential);
/* One field (CPML_VX ...) as an interior Fortran-order array. */
int cpml_gpu_get_field(cpml_solver *h, int which, void *out);
int cpml_gpu_set_field(cpml_solver *h, int which, const void *in);
/* One memory variable (CPML_MEM_...) expanded to the full grid (zero outside
 * its absorbing slab), as the reference stores it. */
int cpml_gpu_get_memory(cpml_solver *h, int which, void *out);
/* The inverse: a full-grid memory variable (its slab points are stored, the
 * rest is the reference's exact zero).  With cpml_gpu_set_field, restores a
 * state saved by get_field / get_memory exactly, so that a run resumed from
 * a checkpoint continues as the uninterrupted run does, bit for bit (the
 * reference keeps these arrays as plain program variables, 3D:222-240). */
int cpml_gpu_set_memory(cpml_solver *h, int which, const void *in);
/* max over the grid of sqrt(vx^2 + vy^2 [+ vz^2]) (2D2:719, 3D:1185). */
int cpml_gpu_velocity_norm(cpml_solver *h, double *vmax);

/* The same norm without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory when they call (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_vx, d_vy (and d_vz in 3D) are device arrays of the NX*NY[*NZ] interior
 * points (Fortran order, precision prec); the result (a double) is written to
 * device memory at d_vmax on `stream` (a hipStream_t; NULL = the default
 * stream) without synchronising.  The per-block partial maxima live in a
 * fixed 8 KiB device array of the library (one per GPU, held until the
 * process ends); grids above 2^36 points are refused (CPML_ERR_ARG).  Calls
 * that overlap on one GPU share that array and race: serialise them.  On
 * error, cpml_gpu_last_error(NULL, ...) gives the message. */
int cpml_velocity_norm_noalloc(int prec, int ndim, int NX, int NY, int NZ, const void *d_vx,
                               const void *d_vy, const void *d_vz, double *d_vmax,
                               void *stream);

/* ---- execution options and measurement ------------------------------- */

/* Steps captured per HIP graph replay (default 16; 0 = launch eagerly). */
int cpml_gpu_set_graph_steps(cpml_solver *h, int steps);
/* Tiling: planes per block along z (3D, default 8) and rows of 64 points per
 * tile (4, 8 or 16; default 4; 2D also 32 or 64, walked by 16-row blocks; the
 * fused 3D step owns 62-point rows, 6 of them per block for 4 or 8 and 14 for
 * 16).  Results do not depend on either. */
int cpml_gpu_set_kchunk(cpml_solver *h, int kchunk);
int cpml_gpu_set_tile(cpml_solver *h, int tile_y);
/* 3D only: run each time step as ONE fused kernel (stress and velocity in a
 * single pass over HBM, state kept in two ping-pong buffers: 9 more field
 * arrays).  Results are identical to the two-kernel path. */
int cpml_gpu_set_fused(cpml_solver *h, int on);
/* Advance `steps` full leapfrog steps per launch and per pass over HBM
 * (time blocking of the loops 2D2:550-686, 2D4:551-687, 3D:836-1121: every
 * intermediate time level stays in registers (and, in 3D, LDS); the state
 * ping-pongs between two buffers: one more set of field arrays and memory
 * variables).  steps 0 turns it off.  2D (orders 2, 4, 8): at most 4 / 3 / 2
 * steps in float, 2 in double; one 64-lane wavefront per strip of
 * strip_width columns marching through chunks of chunk_rows rows (0: 40).
 * 3D (order 2): at most 2 steps; a workgroup owns strip_width columns x
 * 12 (float; double: 4) rows for 2 steps (14 / 6 for 1) and marches through
 * z chunks of chunk_rows planes (0: 64).  strip_width 0 = every column the
 * ring of steps*(order-1) columns per side leaves (3D: rounded down to a
 * multiple of 8).  Used
 * on the elastic step without a learned boundary; with the energy curve on
 * (set_energy), in 2D at any cadence (the kernel sums the energy terms of
 * every level it computes; the curve equals the single-step one to
 * rounding), in 3D when the cadence is a multiple of `steps` (the energy of
 * each block's last step); run() ends a block at every stability check.  Results
 * are identical to the single-step kernels. */
int cpml_gpu_set_time_block(cpml_solver *h, int steps, int strip_width, int chunk_rows);
/* No counterpart in the reference (it has one loop structure): times this
 * handle's step with single steps and with every steps-per-launch x chunk
 * depth of cpml_gpu_set_time_block (2D: 16, 24, 40, 64 rows; 3D: 32, 64
 * planes; default strips) over about trial_steps steps, and sets the fastest.
 * The state, seismograms, energy curve and kernel timers are saved before the
 * trials and restored after them (temporary device copies), so it may be
 * called at any point of a run.  With energy on in 3D, only steps per launch
 * that divide the cadence are tried.  A step without a time-blocked form
 * (attenuation, the RK4 scheme, a learned boundary, the fused 3D step, 3D
 * order 4) is left as it is (best_ms = -1).  Outputs (may be NULL): steps per launch (0: single
 * steps), chunk depth (0 for single steps), time of trial_steps steps in ms. */
int cpml_gpu_autotune(cpml_solver *h, int trial_steps, int *steps_per_launch, int *chunk,
                      double *best_ms);
/* When on, run() launches eagerly with a HIP event pair around every stress
 * and velocity kernel on the handle's stream; see cpml_gpu_kernel_time. */
int cpml_gpu_set_kernel_timing(cpml_solver *h, int on);
int cpml_gpu_kernel_time(cpml_solver *h, int kernel, int *launches, double *total_ms);
/* Device time of the last run() (HIP events around it on the handle's stream). */
int cpml_gpu_elapsed_ms(cpml_solver *h, double *ms);
/* Device bytes allocated by the handle. */
int cpml_gpu_device_bytes(cpml_solver *h, double *bytes);
/* Widths of the absorbing slabs found by set_pml: w[0..5] = x min/max,
 * y min/max, z min/max. */
int cpml_gpu_slab_widths(cpml_solver *h, int *w);

/* CPML_OK when arrays of `kind` bytes per value (4 or 8) match the handle's
 * precision, else CPML_ERR_ARG with the message; the Fortran module calls it
 * before handing over real(4) or real(8) arrays. */
int cpml_gpu_check_kind(cpml_solver *h, int kind);
/* Grid of the handle (NZ = 1 in 2D). */
int cpml_gpu_grid(cpml_solver *h, int *NX, int *NY, int *NZ);
/* Host layout of the grid arrays handed to set_medium, get_field, set_field,
 * get_memory and set_memory: dimensioned (1-lo : N+hi) along every grid axis,
 * x fastest -- the reference's explicit-lower-bound arrays, e.g.
 * 2D4:205-225 `(0:NX+1,0:NY+1)` is lo = hi = 1 and ADE8:272-301
 * `(-4:NX+4,-4:NY+4)` is lo = 5, hi = 4.  Only the interior 1..N is read or
 * written (the ghost cells of the host array are left as they are), straight
 * from the caller's array: no interior copy on the host.  Default 0, 0;
 * 0 <= lo, hi <= 8.  The 1D profile, source and seismogram arrays keep their
 * own bounds (the reference declares its profiles (NX), 2D4:244). */
int cpml_gpu_set_host_layout(cpml_solver *h, int lo, int hi);
/* CPML_OK when a host grid array of `kind` bytes per value, rank `rank` and
 * extents shape[0..rank-1] fits the handle's grid and host layout (rank ==
 * ndim: every extent; other ranks: the number of values), else CPML_ERR_ARG
 * naming both shapes; the Fortran module calls it on every grid array. */
int cpml_gpu_check_grid_array(cpml_solver *h, int kind, int rank, const int64_t *shape);

/* Debug aid against out-of-grid accesses.  Guard cells are the words of the
 * device field, medium and memory-variable arrays that belong neither to the
 * grid nor to its ghost points (the row padding that aligns every row to 256
 * bytes).  op 1 fills every guard cell with quiet NaN and sets *count to
 * their number; op 0 sets *count to the number of guard cells that no longer
 * hold a NaN (a store outside the grid).  A run after op 1 whose fields and
 * seismograms are unchanged shows that no kernel lets a guard cell reach a
 * result.  The guard cells of arrays allocated later (set_pml, the second
 * state of set_fused / set_time_block) are not poisoned: call it after the
 * set-up. */
int cpml_gpu_guard_cells(cpml_solver *h, int op, long long *count);

/* Message of the last error on this handle (h == NULL: of the last failed
 * cpml_gpu_init or cpml_velocity_norm_noalloc on this thread).  Always
 * NUL-terminated. */
int cpml_gpu_last_error(cpml_solver *h, char *buf, int len);

/* ---- batched multi-shot 2D ------------------------------------------- */

/* The reference runs one shot per program execution (2D2:137-218 fixes one
 * source and one model); a survey of many shots is a loop of such runs.
 * cpml_batch advances nshots independent 2D shots together on one GPU, one
 * launch per phase for all of them; each shot has its own grid size, spacing,
 * time step, NPOINTS_PML, medium, profiles, source and receivers, and its
 * results equal a cpml_solver run of that shot alone, bit for bit.  Per-shot
 * arrays are indexed 0..nshots-1; every other convention is cpml_solver's.
 * Energy and the stability check are per shot (cpml_batch_set_energy,
 * cpml_batch_set_stability). */
typedef struct cpml_batch cpml_batch;

int cpml_batch_init(cpml_batch **b, int order, int prec, int nshots, const int *NX,
                    const int *NY, const double *DELTAX, const double *DELTAY,
                    const double *DELTAT, const int *NPOINTS_PML, int NSTEP);
int cpml_batch_finalize(cpml_batch *b);
/* as cpml_gpu_set_medium / set_pml / set_source / set_receivers for one shot */
int cpml_batch_set_medium(cpml_batch *b, int shot, const void *lambda, const void *mu,
                          const void *rho);
int cpml_batch_set_pml(cpml_batch *b, int shot, const void *a_x, const void *b_x,
                       const void *K_x, const void *a_x_half, const void *b_x_half,
                       const void *K_x_half, const void *a_y, const void *b_y, const void *K_y,
                       const void *a_y_half, const void *b_y_half, const void *K_y_half);
int cpml_batch_set_source(cpml_batch *b, int shot, int ISOURCE, int JSOURCE,
                          const void *force_x, const void *force_y);
int cpml_batch_set_receivers(cpml_batch *b, int shot, int NREC, const int *ix_rec,
                             const int *iy_rec);
/* Launch shape: rows per tile (4, 8, 16, or 32 / 64 walked by 16-row
 * workgroups); blocks_per_cu 0 = one workgroup per tile, else a persistent
 * grid of blocks_per_cu * CUs workgroups; steps per captured graph (0: no
 * graph). */
int cpml_batch_set_tiling(cpml_batch *b, int tile_y, int blocks_per_cu, int graph_steps);
/* Times every launch shape (single steps: rows per tile x persistent grid;
 * time-blocked: steps per launch x chunk rows x persistent grid) over about
 * trial_steps steps of this batch, keeps the fastest per step, then returns
 * the batch to rest (fields, memory variables and seismograms zero); the
 * whole choice is read back with cpml_batch_launch_shape.  cache_path (may
 * be NULL) is a text file of earlier choices keyed by device, order,
 * precision and batch shape; a hit skips the trials (best_ms = -1). */
int cpml_batch_autotune(cpml_batch *b, int trial_steps, const char *cache_path, int *tile_y,
                        int *blocks_per_cu, double *best_ms);
/* Time blocking of the whole batch: `steps` leapfrog steps of every shot per
 * launch (the kernel of cpml_gpu_set_time_block over a flat list of (shot,
 * strip, chunk) items; slab items beside the others on a second stream, each
 * covering half a chunk when chunks have 48 rows or more;
 * state ping-ponging between two pools), 0 = single steps.  strip_width,
 * chunk_rows: as cpml_gpu_set_time_block (0 = defaults).  With energy on,
 * blocks are used when the cadence is a multiple of `steps`. */
int cpml_batch_set_time_block(cpml_batch *b, int steps, int strip_width, int chunk_rows);
/* The launch shape in use (after set_tiling / set_time_block / autotune). */
int cpml_batch_launch_shape(cpml_batch *b, int *tile_y, int *blocks_per_cu, int *time_block,
                            int *chunk_rows);
/* Per-shot energy every `cadence` steps (0 = off) over each shot's reference
 * window (2D2:695 / 2D4:696), 2D2:694-713; kinetic(NSTEP), potential(NSTEP)
 * of one shot (double). */
int cpml_batch_set_energy(cpml_batch *b, int cadence);
int cpml_batch_get_energy(cpml_batch *b, int shot, double *kinetic, double *potential);
/* Per-shot stability check inside run (2D2:716-726): at it % IT_DISPLAY == 0
 * and it == 5 the max norm of v of every shot is compared with threshold; the
 * first shot above it (or NaN) fails the run with CPML_ERR_BLOWUP and its
 * index in the message.  IT_DISPLAY <= 0 disables it. */
int cpml_batch_set_stability(cpml_batch *b, int IT_DISPLAY, double threshold);
/* Steps it_begin..it_end of every shot (2D2:545-735 per shot). */
int cpml_batch_run(cpml_batch *b, int it_begin, int it_end);
int cpml_batch_get_seismograms(cpml_batch *b, int shot, void *sisvx, void *sisvy);
/* which: CPML_VX, CPML_VY, CPML_SIGMAXX, CPML_SIGMAYY or CPML_SIGMAXY */
int cpml_batch_get_field(cpml_batch *b, int shot, int which, void *out);
int cpml_batch_elapsed_ms(cpml_batch *b, double *ms);
/* as cpml_gpu_check_kind for a batch */
int cpml_batch_check_kind(cpml_batch *b, int kind);
/* grid of one shot */
int cpml_batch_grid(cpml_batch *b, int shot, int *NX, int *NY);
int cpml_batch_last_error(cpml_batch *b, char *buf, int len);

/* ---- learned thin absorbing boundary (3D, FP32) ------------------------ */

/* The reference absorbs outgoing waves with a C-PML of NPOINTS_PML cells
 * (3D:399-667 profiles, 3D:836-1052 memory variables).  The learned form
 * replaces it by a 4-cell strip on every face: the grid is the interior plus
 * the strip, updated either without C-PML (profiles with a = 0, K = 1) or
 * with a 4-cell C-PML of its own (NPOINTS_PML = 4), and after each step
 * (after the source and the rigid edges, before the receivers) a per-face
 * corrector adds a correction to the outermost cells the face owns (the
 * nearest face; ties to the lowest of x-, x+, y-, y+, z-, z+).  Over a thin
 * C-PML the corrector learns what that C-PML misses against a wide one.
 *
 * Corrector of one face: three linear 3x3x3 convolutions 9 -> 16 -> 16 -> 9
 * channels (no bias, no activation) over a 4-cell-deep box along the face,
 * evaluated at depth 0,
 * inputs = the 9 fields divided by scales[9]; its output times the scales is
 * added to the fields.  Weights: float[6 faces][3 layers][27 taps][16 in]
 * [16 out], tap = (dd+1)*9 + (da+1)*3 + (db+1) for offsets along the inward
 * depth and the face's two tangential axes in increasing axis order; input
 * channels 9..15 of layer 1 and output channels 9..15 of layer 3 are padding
 * (keep them 0).  Inference uses BF16 weights and activations with FP32
 * accumulation (MFMA); training runs in FP32. */
typedef struct cpml_abc_net cpml_abc_net;
#define CPML_BOUNDARY_CPML 0
#define CPML_BOUNDARY_LEARNED 1
#define CPML_ABC_WEIGHTS (6 * 3 * 27 * 16 * 16)

/* weights may be NULL (all zero: the corrector then leaves the fields as the
 * plain step made them). */
int cpml_abc_net_create(cpml_abc_net **net, const float *weights, const float *scales);
int cpml_abc_net_destroy(cpml_abc_net *net);
int cpml_abc_net_get_weights(cpml_abc_net *net, float *weights);
int cpml_abc_net_set_weights(cpml_abc_net *net, const float *weights);
/* mode CPML_BOUNDARY_LEARNED attaches net (not owned; keep it alive while
 * attached) to a 3D FP32 handle (the strip with or without its own thin
 * C-PML); CPML_BOUNDARY_CPML detaches it. */
int cpml_gpu_set_boundary(cpml_solver *h, int mode, cpml_abc_net *net);
/* Adam optimiser state (zeroed); one Adam step every `batch` training steps. */
int cpml_abc_train_init(cpml_abc_net *net, double lr, double beta1, double beta2, double eps,
                        int batch);
/* One training step on one time step of a teacher run: the student grid
 * (no corrector, no source; no C-PML or the strip's own thin C-PML) is the
 * window of the teacher starting at 0-based cell (i0, j0, k0).  Loads the
 * teacher's current fields into the student (the student keeps its own
 * memory variables), steps the student once (P), steps the teacher once
 * (step `it`, T) and adds the gradient of this sample to the mini-batch; the
 * batch's
 * Adam step minimises, per face, sum (y - target)^2 over the owned strip
 * cells of its samples (target = (T - P) / scales) divided by the batch's
 * target energy, floored at 1% of its input energy.  loss6 (may be NULL):
 * this sample's per-face normalised error before the update. */
int cpml_abc_train_step(cpml_abc_net *net, cpml_solver *student, cpml_solver *teacher, int i0,
                        int j0, int k0, int it, double *loss6);
/* Input noise of train_step: each sample's box gets rel_noise times its RMS
 * of white noise (target unchanged), penalising the corrector's response to
 * patterns the data never shows (default 0). */
int cpml_abc_train_noise(cpml_abc_net *net, double rel_noise);
/* Per-face normalised error and its gradient (float[CPML_ABC_WEIGHTS]) for the
 * current student state (taken as P) against the teacher's current state
 * (taken as T); no step, no update (a partial mini-batch is dropped).  Needs
 * cpml_abc_train_init. */
int cpml_abc_grad(cpml_abc_net *net, cpml_solver *student, cpml_solver *teacher, int i0, int j0,
                  int k0, double *loss6, float *grad);
/* Corrector output for the current state of h, before ownership and scaling:
 * float[face][field 9][depth 4][b][a] (each face's block 9*4*na*nb; depths
 * 1..3 are 0),
 * computed by the FP32 training path (use_mfma = 0) or the BF16 MFMA
 * inference kernel (use_mfma = 1). */
int cpml_abc_correction(cpml_abc_net *net, cpml_solver *h, int use_mfma, float *out);
int cpml_abc_last_error(cpml_abc_net *net, char *buf, int len);

/* ---- 2D viscoelastic attenuation (2D-visco2 =
 * seismic_CPML_2D_velocity_and_stress_second_order_viscoelastic.f90) ----- */

/* Switches a 2D handle to the viscoelastic step of 2D-visco2 (order 2) or of
 * seismic_CPML_2D_velocity_and_stress_fourth_order_viscoelastic.f90 (order 4:
 * the same step with the (9/8, -1/24) difference operator on its ghost-padded
 * arrays) (VISCOELASTIC_ATTENUATION, 2D-visco2:736-929): N_SLS standard linear solids
 * per modulus with the relaxation times the reference's SolvOpt fit gives
 * (tau_epsilon/tau_sigma for Qp: nu1, for Qs: nu2; 2D-visco2:360-372), memory
 * variables R_e1, R_e11, R_e13 (zeroed here), derivatives multiplied by
 * ONE_OVER_DELTA as that program does, and the pressure seismogram of
 * 2D-visco2:958-990 recorded beside vx, vy.  The medium passed to
 * cpml_gpu_set_medium is the unrelaxed one (lambda_unrelaxed, mu_unrelaxed,
 * rho); the force table is the program's force / (DELTAX*DELTAY)
 * (2D-visco2:899).  N_SLS = 0 returns to the elastic step.  N_SLS <= 8. */
int cpml_gpu_set_attenuation(cpml_solver *h, int N_SLS, const double *tau_epsilon_nu1,
                             const double *tau_sigma_nu1, const double *tau_epsilon_nu2,
                             const double *tau_sigma_nu2);
/* R_e1 (which 0), R_e11 (1) or R_e13 (2): N_SLS planes of NX*NY (x fastest),
 * the memory_variable_R_*_dot arrays of 2D-visco2:324-326. */
int cpml_gpu_get_attenuation_memory(cpml_solver *h, int which, void *out);
/* sispressure(NSTEP, NREC) of 2D-visco2:297, 990 (viscoelastic step only). */
int cpml_gpu_get_pressure(cpml_solver *h, void *sispressure);

/* ---- ADE-PML RK4, 8th order (ADE8 = seismic_ADEPML_2D_elastic_RK4_eighth_order.f90) */

/* Switches a 2D order-8 handle to the time scheme of ADE8 (ADE8:692-906):
 * four Runge-Kutta stages per step (rk41 = 0, 1/2, 1/2, 1; rk42 = 1/6, 2/6,
 * 2/6, 1/6), the Holberg operator divided by DELTA, the auxiliary-
 * differential-equation PML with the stage tables a_x(4, NX), b_x(4, NX),
 * a_x_half, b_x_half, a_y(4, NY) ... of ADE8:410-640 (Fortran layout, stage
 * fastest; the K profiles are those given to cpml_gpu_set_pml) applied where
 * i <= NPOINTS_PML+2 or i >= NX-NPOINTS_PML-2 (same in j), the force at each
 * stage time force_x4(4, NSTEP) / force_y4(4, NSTEP) (ADE8:786-796), and vx
 * recorded as the mean of the four nodes (i..i+1, j..j+1) (ADE8:877-884).
 * All stage and memory arrays start at zero. */
int cpml_gpu_set_ade_rk4(cpml_solver *h, const void *a_x, const void *b_x,
                         const void *a_x_half, const void *b_x_half, const void *a_y,
                         const void *b_y, const void *a_y_half, const void *b_y_half,
                         const void *force_x4, const void *force_y4);

/* ---- PNM snapshots (2D2:716-731, 861-999; 3D:1231-1234) ----------------- */

/* Colour image of the current Vx (field_number 1) or Vy (2), computed on the
 * device as the reference's create_color_image does (2D2:861-999): pixel
 * values from v / max|v| (power 0.3, red positive, blue negative, white below
 * 1% of the maximum), the source cross, a two-pixel black frame, the inner
 * edges of the PML layers in use and the receiver squares, in the handle's
 * precision of arithmetic.  rgb receives NX*NY*3 bytes, top row (j = NY)
 * first; for order 4, (NX+2)*(NY+2)*3 bytes: the 4th-order program images its
 * ghost-padded arrays (2D4:729-731).  3D handles image the plane k = NZ/2
 * (the top plane of the reference's cut-plane rank, 3D:346, 1231-1234), with
 * every receiver drawn at its (i, j).  Replaces the image half of
 * 2D2:716-731 and 3D:1231-1234. */
int cpml_gpu_color_image(cpml_solver *h, int field_number, unsigned char *rgb);

/* ---- host-side set-up helpers (no GPU needed) -------------------------- */

/* Write rgb (NX*NY*3 bytes, top row first) as the reference's PNM P3 text
 * file (2D2:900-903, 989: 'P3', ' NX NY', ' 255', one 'RRR GGG BBB' line per
 * pixel). */
int cpml_pnm_write(const char *path, int nx, int ny, const unsigned char *rgb);

/* Write ntr traces of ns samples as the SU file of the 8th-order program
 * (ADE8:987-1003 for seismograms, ADE8:935-953 for snapshots): per trace a
 * 240-byte header of 120 native int16 words, all zero but word 58 = ns and
 * word 59 = dt_value truncated toward zero, then the ns samples as native
 * float32.  traces: ntr blocks of ns values (the Fortran array
 * traces(ns, ntr)), float (prec 4) or double (prec 8, rounded to float).
 * dt_value is what the header carries: DELTAT*1e6 (microseconds) for
 * seismograms, DELTAY*1e3 for snapshots.  The reference stores both words
 * through integer(2) assignments; values outside 0..32767 (ns >= 1) are
 * refused here (CPML_ERR_ARG) rather than written wrapped. */
int cpml_su_write(const char *path, int ntr, int ns, double dt_value, const void *traces,
                  int prec);


/* C-PML profile along one axis (2D2:283-394 for x; 2D2:396-465 for y;
 * 3D:399-667): d0 = -(NPOWER+1) cp ln(Rcoef) / (2 L), NPOWER = 2,
 * K_MAX_PML = 1, ALPHA_MAX_PML = 2 pi (f0/2), Rcoef = 1e-3.
 * origin_max_uses_n reproduces 2D4:401 (N*DELTA - L), clamp_alpha the
 * x-only clamp of 2D2:385-386.  Output arrays of length n in precision prec. */
int cpml_pml_profile(int prec, int n, double delta, double deltat, int npoints_pml, double cp,
                     double f0, int use_min, int use_max, int origin_max_uses_n,
                     int clamp_alpha, void *a, void *b, void *K, void *a_half, void *b_half,
                     void *K_half);
/* Force tables (2D2:644-657): kind 0 Gaussian, 1 first derivative of a
 * Gaussian (what the reference programs run), 2 Ricker; ANGLE_FORCE in
 * degrees clockwise from the vertical y axis. */
int cpml_source_time_function(int prec, int nstep, double deltat, double f0, double t0,
                              double factor, double angle_force, int kind, void *force_x,
                              void *force_y);
/* Nearest grid point of NREC receivers evenly spaced from (xdeb,ydeb) to
 * (xfin,yfin) (2D2:483-509).  1-based indices. */
int cpml_nearest_receivers(int prec, int NX, int NY, double DELTAX, double DELTAY, int NREC,
                           double xdeb, double ydeb, double xfin, double yfin, int *ix_rec,
                           int *iy_rec);
/* Courant number cp*DELTAT*sqrt(1/DELTAX^2 + 1/DELTAY^2 [+ 1/DELTAZ^2])
 * (2D2:513, 3D:712). */
int cpml_courant_number(int ndim, double cp, double deltat, double deltax, double deltay,
                        double deltaz, double *courant);

#ifdef __cplusplus
}
#endif
#endif

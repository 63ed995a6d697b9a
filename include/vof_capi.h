/*
 * vof_capi.h — C ABI of the MI355X (gfx950) VOF timestep library `libvof_hip.so`.
 *
 * The reference solver (voftrial1.cu) has no plugin or FFI surface; its step is a
 * fixed sequence of host wrappers over `DeviceFields` with the parameters in
 * `__constant__ Params cpar`.  Each entry point below replaces one piece of that
 * sequence; the comment on each names the reference code it stands in for.
 *
 * Conventions
 *  - The library owns all device memory; the caller owns host buffers and no
 *    pointer is retained after a call returns.
 *  - Every call returns a vof_status; nothing inside the library calls exit().
 *    vof_last_error() gives a message for the last failure on the calling thread.
 *  - One handle = one HIP stream; a handle is not thread-safe, separate handles
 *    may be driven from separate threads.  vof_step is asynchronous; the
 *    download/diagnostics/vtk calls synchronise the handle's stream.
 *  - Field layout in host buffers is the reference's: (nx+2)(ny+2)(nz+2)
 *    values, one ghost layer, x fastest, idx = (k*(ny+2) + j)*(nx+2) + i
 *    (voftrial1.cu:66-70).  Element type is float for precision 4 and double
 *    for precision 8.
 */
#ifndef VOF_CAPI_H
#define VOF_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vof_status {
    VOF_OK = 0,
    VOF_ERR_INVALID = 1,     /* bad argument / size / precision */
    VOF_ERR_HIP = 2,         /* a HIP runtime call failed */
    VOF_ERR_ALLOC = 3,       /* device or host allocation failed */
    VOF_ERR_UNSUPPORTED = 4, /* feature not available in this build */
    VOF_ERR_NUMERICS = 5     /* watchdog: NaN/Inf or F outside [0,1] */
} vof_status;

/* Fields a caller can move across the boundary (names follow DeviceFields,
 * voftrial1.cu:1176-1198).  RHO and NU are derived from F and are read-only. */
typedef enum vof_field {
    VOF_FIELD_F = 0,
    VOF_FIELD_U = 1,
    VOF_FIELD_V = 2,
    VOF_FIELD_W = 3,
    VOF_FIELD_P = 4,
    VOF_FIELD_USTAR = 5,
    VOF_FIELD_VSTAR = 6,
    VOF_FIELD_WSTAR = 7,
    VOF_FIELD_RHS = 8,
    VOF_FIELD_RHO = 9,
    VOF_FIELD_NU = 10,
    VOF_FIELD_KAPPA = 11 /* interface curvature of the last step (VOF_FLAG_CSF only; read-only) */
} vof_field;

/* Extension flags (no counterpart in the reference; all default off). */
#define VOF_FLAG_REMOVE_P_MEAN 0x1u /* subtract the FP64 interior mean of p after the Poisson sweeps */
/* Continuum-surface-force surface tension: after the predictor, every interior
 * face gains dt * sigma * kappa_f * dF/dn / rho_f, kappa = -div(grad F/|grad F|)
 * from vertex normals.  sigma = 0 gives the parity step. */
#define VOF_FLAG_CSF           0x2u
/* Density-weighted viscous term in the predictor (the reference's predictor
 * takes rho but never reads it, voftrial1.cu:233, :290, :347, and scales the
 * Laplacian by the nu of one cell, :277-280): every interior face gains
 * dt * div(mu grad u)/rho_f in flux form, mu = rho nu per cell, edge
 * viscosities the mean of the four cells around the edge, rho_f the mean of
 * the face's two cells.  It conserves momentum across density jumps (the edge
 * fluxes telescope) and keeps a two-layer shear steady; with one fluid it is
 * the reference's term up to rounding.  Not supported in reverse mode. */
#define VOF_FLAG_RHO_PREDICTOR 0x4u
/* No-slip walls: every tangential velocity ghost holds the negated interior
 * value (the reference copies it: free slip, :959-1051), so the velocity
 * vanishes on the wall halfway between ghost and interior face; normal wall
 * faces stay 0 as in the reference.  Not supported in reverse mode. */
#define VOF_FLAG_NO_SLIP       0x8u
#define VOF_FLAG_WATCHDOG      0x10u/* after every vof_step call: F in [0,1] (up to clamp01's half-ulp
                                         leaks) and u,v,w,p,F finite, else VOF_ERR_NUMERICS */

/* Solver parameters.  Replaces `struct Params` + `cpar_host` defaults
 * (voftrial1.cu:27-41, :1258-1291).  Physical inputs are double; they are
 * rounded to the field precision before any derived value (dx = L/n,
 * dxi = 1/dx, vol) is computed in that precision, as the reference does in float. */
typedef struct vof_params {
    int32_t nx, ny, nz;         /* interior cells */
    int32_t precision;          /* 4 = FP32 fields, 8 = FP64 fields */
    double Lx, Ly, Lz;          /* domain extents [m] */
    double rho_l, rho_g;        /* densities [kg/m^3] */
    double nu_l, nu_g;          /* kinematic viscosities [m^2/s] */
    double sigma;               /* surface tension [N/m]; used only with VOF_FLAG_CSF */
    double gx, gy, gz;          /* gravity [m/s^2] */
    double dt;                  /* time step [s] */
    int32_t jacobi_sweeps;      /* Jacobi sweeps per step (reference: 10, voftrial1.cu:1371) */
    uint32_t flags;             /* VOF_FLAG_* */
    int32_t jacobi_tblock;      /* Jacobi sweeps per launch, 1 .. 5 (temporal blocking); 0 = auto */
    int32_t reserved;
} vof_params;

/* FP64 global diagnostics (no counterpart in the reference, which has no reductions). */
typedef struct vof_diag {
    double liquid_volume;   /* vol * sum_interior F  [m^3] */
    double p_mean;          /* mean of p over interior cells */
    double residual_l2;     /* || rhs - A p ||_2 over interior cells, after the last step's sweeps */
    double rhs_l2;          /* || rhs ||_2 (for a relative residual) */
    double f_min, f_max;    /* min/max of F over interior cells */
    int64_t nonfinite;      /* count of NaN/Inf in u,v,w,p,F (interior+ghost) */
} vof_diag;

typedef struct vof_solver vof_solver;

/* Reference defaults of voftrial1.cu:1258-1283 (200^3, 0.1 m cube, rho 1000/50,
 * nu 1e-6/1.5e-5, sigma 0.007, g = (0,-5,0), dt = 4e-6, 10 sweeps, FP32). */
void vof_params_default(vof_params* p);

/* Allocate and zero every device field (replaces alloc_field / the 20 cudaMalloc +
 * cudaMemset(0) calls, voftrial1.cu:1200-1203, :1303-1326; the zero fill is
 * load-bearing: never-written ghost entries must read 0). */
vof_status vof_create(const vof_params* p, vof_solver** out);
void vof_destroy(vof_solver* s);                      /* replaces free_all, :1205-1226 */

/* Dam-break initial condition, then the F ghost fill.  ic == 1 is the
 * reference predicate (set_init_F_kernel + apply_bc_F, :72-93, :1328-1336),
 * evaluated in float exactly as the reference does; ic 2 and 3 leave F = 0
 * (the reference accepts them and sets no liquid). */
vof_status vof_init_ic(vof_solver* s, int ic);

/* Extension IC: liquid in the box x < bx, y < by, z < bz [m] by the lower-face
 * rule with strict '<', evaluated in exact integer arithmetic (no reference
 * counterpart; the reference box is fixed at Lx/3, Ly/2, Lz/3 inclusive). */
vof_status vof_init_box(vof_solver* s, double bx, double by, double bz);

/* Advance nsteps time steps.  first_istep is the 1-based index of the first
 * step (the reference's ++istep, :1362); it selects the FCT sweep rotation
 * istep%3 (:1384-1396).  Replaces the loop body :1362-1401.  A negative
 * first_istep or nsteps returns VOF_ERR_INVALID (as in every entry point that
 * takes a step index). */
vof_status vof_step(vof_solver* s, int64_t first_istep, int nsteps);

/* Same as vof_step, but replays one captured hipGraph per step (captured on
 * first use per rotation/buffer parity). Results are identical to vof_step. */
vof_status vof_step_graph(vof_solver* s, int64_t first_istep, int nsteps);

/* Host <-> device copies of a whole field in the reference layout.  count is
 * the number of elements and must equal (nx+2)(ny+2)(nz+2).  Replaces the
 * cudaMemcpy of F at :1405 (and generalises it to every state field). */
vof_status vof_download(vof_solver* s, vof_field f, void* host, size_t count);
vof_status vof_upload(vof_solver* s, vof_field f, const void* host, size_t count);

/* FP64 diagnostics of the current state (synchronises). */
vof_status vof_diagnostics(vof_solver* s, vof_diag* d);

/* Diagnostics that allocate nothing (no counterpart in the reference, whose
 * step has no reductions).  For callers that cannot allocate device memory:
 * code under HIP stream capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The same FP64 record as vof_diagnostics, computed from the
 * caller's device fields in the reference layout ((nx+2)(ny+2)(nz+2) values
 * of the field precision each: F, u, v, w, p and the Poisson right-hand side
 * rhs), written to `out`, a vof_diag in device memory, by two kernels
 * enqueued on `stream` (a hipStream_t; NULL = the null stream); the call does
 * not wait for them.  No handle is needed and nothing is allocated: the
 * per-work-group partial sums live in a static device array of 65536 x 7
 * doubles (3.7 MB) that each GPU holds until the process ends, shared by
 * every call on that GPU, so calls that overlap on one GPU race.  Grids
 * whose reduction needs more than 65536 work-groups even with one chunk of
 * all planes (a stored plane of more than 16,777,216 cells) are refused
 * with VOF_ERR_INVALID.  Results equal
 * vof_diagnostics' bit for bit on grids of up to 65536 work-groups with
 * 8-plane chunks (e.g. 384^3); larger grids use longer chunks (same sums,
 * other rounding order). */
vof_status vof_diagnostics_noalloc(const vof_params* p, const void* F, const void* u, const void* v, const void* w,
                                   const void* pres, const void* rhs, vof_diag* out, void* stream);

/* Legacy ASCII VTK of F, byte-identical to write_legacy_vtk (:907-941). */
vof_status vof_write_vtk(vof_solver* s, const char* path);

/* The same file written without stalling the steps (no counterpart in the
 * reference, whose loop writes synchronously, :1403-1411): F as of the calls
 * enqueued so far is copied on the handle's stream into a device staging slot,
 * then to pinned host memory on a second stream while later steps run, and a
 * host thread writes the file.  Two slots; a third pending file waits for the
 * oldest.  Staging buffers are allocated on the first call.  Errors of a
 * writer (e.g. the file cannot be opened) are returned by the next
 * vof_write_vtk_async or by vof_output_wait; vof_destroy waits for every
 * pending file. */
vof_status vof_write_vtk_async(vof_solver* s, const char* path);
vof_status vof_output_wait(vof_solver* s);

/* Restart file (no counterpart in the reference, whose VTK output cannot
 * restart a run): the solver state after step `istep` -- both F buffers, u,
 * v, w and p -- in a binary file with a versioned header (grid, precision).
 * vof_load_state into a solver created with the same parameters and
 * continuing with vof_step(s, *istep + 1, ...) gives the bits of the
 * uninterrupted run.  Single solvers only. */
vof_status vof_save_state(vof_solver* s, const char* path, int64_t istep);
vof_status vof_load_state(vof_solver* s, const char* path, int64_t* istep);

/* Host-only VTK writer on a caller buffer (same bytes as vof_write_vtk; no
 * device access).  precision 4: data is float*, 8: double*. */
vof_status vof_write_vtk_host(const char* path, const void* data, int precision, int nxTot, int nyTot,
                              int nzTot);

/* Test hook of the VTK writer (no reference counterpart): n >= 0 lets at
 * most n formatting threads start per batch of slabs, as if the system could
 * start no more; the writer then formats the remaining slabs on the calling
 * thread and the file is the same bytes.  n < 0 restores the default. */
void vof_debug_vtk_thread_limit(int n);

/* Attainable device bandwidth: a plain 16-byte-per-lane copy of `bytes`
 * bytes repeated `reps` times; *gbs = (read + written bytes) / time. */
vof_status vof_copy_bandwidth(int64_t bytes, int reps, double* gbs);

/* Attainable time of a stream mix (extension, a measurement with no
 * counterpart in the reference): a plain kernel reads `nread` arrays (1..6)
 * and writes `nwrite` arrays (1..3) of `field_bytes` bytes each (1 MiB ..
 * 4 GiB), 16 bytes per lane, every written element formed from the elements
 * read at its index.  *us_per_pass receives the average microseconds of one
 * pass over `reps` passes.  resident = 0 rotates the passes over enough array
 * sets that each pass's inputs come from HBM (as a step kernel's inputs
 * mostly do); resident = 1 repeats one set (inputs cached when they fit).
 * With a step kernel's compulsory fields (e.g. 4 read, 3 written for the
 * predictor) it is the time that kernel's traffic needs with no arithmetic or
 * stencil to hide.  Allocates and frees its arrays. */
vof_status vof_stream_bandwidth(int64_t field_bytes, int nread, int nwrite, int resident, int reps,
                                double* us_per_pass);

/* Synchronise the handle's stream. */
vof_status vof_sync(vof_solver* s);

/* Time the plain step with HIP events on the handle's stream: runs `warmup`
 * untimed and `steps` timed steps starting at first_istep and returns the
 * milliseconds in *ms_total.  When stage_ms is not NULL the steps run eagerly
 * with an event between stages and stage_ms[VOF_NSTAGE_TIMES] receives the
 * device time per stage summed over the timed steps, in the order
 * predictor, rhs, jacobi (all sweeps), p-mean removal, correction+BC,
 * fct x, fct y, fct z, F ghosts, fused fct z+x and fused fct x+y (steps whose
 * sweep pairs run in one launch: see fct_zx_fused and fct_xy_fused). */
#define VOF_NSTAGE_TIMES 11
vof_status vof_time_steps(vof_solver* s, int64_t first_istep, int warmup, int steps,
                          int use_graph, double* ms_total, double* stage_ms);

/* Launch-shape tuning (no reference counterpart).  kz_* = planes each wave
 * walks in the column kernels and the x sweep (kz_predict 4 or 8 in FP32 with
 * every instance's ny a multiple of 4: the predictor holds 4 rows per lane;
 * kz_jacobi 20 or 25 in FP32 on grids of at least that many planes: the
 * two-sweep Jacobi streams its planes through LDS), fct_chunk = cells per lane in
 * the y/z FCT marches, jacobi_tblock = Jacobi sweeps fused per launch (1 .. 5:
 * T >= 3 runs T sweeps per launch while T are left, then pairs, then one;
 * another value is VOF_ERR_INVALID),
 * fct_zx_fused = FCT sweeps per launch where the z sweep is directly followed
 * by the x sweep (sweep orders y-z-x and z-x-y; 2 = one launch with kz_fct_x
 * planes per chunk, 1 = two launches), fct_xy_fused the same for the x and y
 * sweeps of the order x-y-z (2 = one launch with fct_chunk rows per chunk);
 * the taped steps of reverse mode always launch the sweeps one by one.
 * Values <= 0 keep the current setting.  Changing them drops captured graphs. */
typedef struct vof_tuning {
    int32_t kz_predict, kz_rhs, kz_jacobi, kz_correct, fct_chunk, jacobi_tblock, kz_fct_x, fct_zx_fused;
    int32_t fct_xy_fused;
} vof_tuning;
vof_status vof_set_tuning(vof_solver* s, const vof_tuning* t);
vof_status vof_get_tuning(const vof_solver* s, vof_tuning* t);

/* Autotuner (no reference counterpart): if cache_path holds an entry for
 * (GPU arch, grid, precision, instance count, sweeps) it is applied; otherwise
 * candidate z-chunk lengths, march chunk lengths and Jacobi temporal-block
 * depths are timed stage by stage over steps_per_trial steps each, rounded up
 * to whole rotations of the sweep order and at the same phase of it (state
 * snapshotted and restored on the device, so the run is unaffected), the
 * fastest per stage is applied and written to cache_path (JSON; NULL: no cache). */
vof_status vof_autotune(vof_solver* s, int64_t first_istep, int steps_per_trial, const char* cache_path,
                        vof_tuning* out);

/* Test hook: run one stage on the current state.  stage: 0 predictor,
 * 1 rhs, 2 one Jacobi sweep, 3 correction + velocity BC, 4/5/6 FCT sweep
 * x/y/z, 7 clamp + F ghost fill, 8 two Jacobi sweeps in one launch, 9 CSF
 * smoothing + curvature + face force (VOF_FLAG_CSF only), 10 jacobi_tblock
 * Jacobi sweeps in one launch (jacobi_tblock 3 .. 5 only).  Not part of the
 * reference surface. */
vof_status vof_debug_stage(vof_solver* s, int stage);

/* Number of interior cells n_x n_y n_z and of stored cells (n+2)^3. */
int64_t vof_interior_cells(const vof_solver* s);
int64_t vof_total_cells(const vof_solver* s);

const char* vof_status_str(vof_status st);
const char* vof_last_error(void);

/* ---------------------------------------------------------------------------
 * Batched ensemble: many instances of uneven grid size and parameters advance
 * together, one launch per stage for the whole ensemble (no reference
 * counterpart; the reference runs one fixed 200^3 case).
 * All instances share the precision, flags, Jacobi sweep count and Jacobi
 * temporal-block depth of params[0]; vof_batch_create returns VOF_ERR_INVALID
 * when any instance's precision, flags, jacobi_sweeps or jacobi_tblock differ.
 */
typedef struct vof_batch vof_batch;
vof_status vof_batch_create(const vof_params* params, int ninst, vof_batch** out);
void vof_batch_destroy(vof_batch* b);
/* per-instance IC box [m], lower-face strict '<' rule; bx/by/bz arrays of ninst */
vof_status vof_batch_init_box(vof_batch* b, const double* bx, const double* by, const double* bz);
vof_status vof_batch_init_ic(vof_batch* b, int ic);
vof_status vof_batch_step(vof_batch* b, int64_t first_istep, int nsteps, int use_graph);
vof_status vof_batch_download(vof_batch* b, int inst, vof_field f, void* host, size_t count);
vof_status vof_batch_upload(vof_batch* b, int inst, vof_field f, const void* host, size_t count);
vof_status vof_batch_diagnostics(vof_batch* b, int inst, vof_diag* d);
vof_status vof_batch_sync(vof_batch* b);
vof_status vof_batch_time_steps(vof_batch* b, int64_t first_istep, int warmup, int steps,
                                int use_graph, double* ms_total);
vof_status vof_batch_set_tuning(vof_batch* b, const vof_tuning* t);
vof_status vof_batch_autotune(vof_batch* b, int64_t first_istep, int steps_per_trial, const char* cache_path,
                              vof_tuning* out);
int64_t vof_batch_interior_cells(const vof_batch* b);
int vof_batch_size(const vof_batch* b);

/* ---------------------------------------------------------------------------
 * Reverse mode (no reference counterpart): the forward/backward pair behind a
 * torch.autograd.Function.  The objective is the liquid-centroid height
 * J = sum(F z) / sum(F) over interior cells after the last step (z = cell
 * centre).  Adjoint arithmetic is FP64 whatever the field precision.
 *
 * vof_adj_forward advances nsteps steps from the solver's current state
 * (exactly as vof_step does) keeping a copy of the state every
 * checkpoint_every steps, and returns J.  vof_adj_backward then runs the
 * steps in reverse, recomputing each checkpoint interval forward once (taping
 * each step's u*, v*, w* and the F between sweeps), and leaves the solver at
 * the state vof_adj_forward ended on.  The seed of the final state is
 * gJ * dJ/d(state) plus any field set with vof_adj_seed between the two calls
 * (vof_adj_forward clears them), so gJ = 0 with seeds gives a general
 * vector-Jacobian product.  The F ghost layer is treated as what the solver
 * keeps it, a copy of the nearest interior cell: the adjoint of F is reported
 * on interior cells (ghost entries 0), that of p likewise (ghost pressures
 * only meet zero coefficients).
 * Supported: flags 0, VOF_FLAG_CSF, VOF_FLAG_WATCHDOG; any other flag returns
 * VOF_ERR_UNSUPPORTED. */
typedef struct vof_adj_scalars {
    double d_rho_l, d_rho_g, d_nu_l, d_nu_g; /* dJ/d parameter */
    double d_gx, d_gy, d_gz;
    double d_sigma; /* 0 unless VOF_FLAG_CSF: sigma does not enter the reference physics */
} vof_adj_scalars;
vof_status vof_adj_forward(vof_solver* s, int64_t first_istep, int nsteps, int checkpoint_every, double* J);
/* Seed of a final-state field (F, U, V, W or P), count = stored cells, FP64. */
vof_status vof_adj_seed(vof_solver* s, vof_field f, const double* host, size_t count);
vof_status vof_adj_backward(vof_solver* s, double gJ, vof_adj_scalars* out);
/* Adjoint of the initial state after vof_adj_backward: F, U, V, W or P. */
vof_status vof_adj_download(vof_solver* s, vof_field f, double* host, size_t count);
/* Device milliseconds of the last vof_adj_forward / vof_adj_backward. */
vof_status vof_adj_times(const vof_solver* s, double* fwd_ms, double* bwd_ms);
/* Free the checkpoints and adjoint arrays. */
vof_status vof_adj_release(vof_solver* s);

#ifdef __cplusplus
}
#endif
#endif /* VOF_CAPI_H */

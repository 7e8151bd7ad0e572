/* npde: MI355X-native 2D Euler finite-volume step (MUSCL + HLLC/Rusanov + SSP-RK2, binary64).
 *
 * C-ABI of libnpde.so.  Plain C types only; every call returns an npde_status and never exits or
 * throws.  One handle drives one instance or a batch of instances of any sizes; device memory
 * belongs to the library, host arrays to the caller.  A handle is not re-entrant: use one per host
 * thread.
 *
 * State layouts accepted by set/get:
 *   NPDE_LAYOUT_PLANES     4 x ny x nx binary64: rho, rho*u, rho*v, E over the interior cells
 *   NPDE_LAYOUT_GHOSTED    (ny+4) x (nx+4) cells of 4 binary64 {rho, rho*u, rho*v, E}, the reference's
 *                          std::vector<Conserved> with ng = 2 (grid2d.hpp:12-28); ghosts are ignored
 *                          on input and filled as apply_bcs would (grid2d.hpp:43-66) on output
 */
#ifndef NPDE_H
#define NPDE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NPDE_OK = 0,
    NPDE_ERR_INVALID = 1,       /* bad argument (sizes, enum values, save sizes that do not divide) */
    NPDE_ERR_NONPHYSICAL = 2,   /* rho <= 0 or p <= 0 met when a step was needed (solver2d.hpp:75) */
    NPDE_ERR_DEVICE = 3,        /* HIP runtime error */
    NPDE_ERR_NOMEM = 4,         /* device allocation failed */
    NPDE_ERR_STATE = 5          /* call out of order (e.g. trajectory requested before a run) */
} npde_status;

typedef enum { NPDE_BC_TRANSMISSIVE = 0, NPDE_BC_PERIODIC = 1 } npde_bc;        /* grid2d.hpp:7 */
typedef enum { NPDE_FLUX_RUSANOV = 0, NPDE_FLUX_HLLC = 1 } npde_flux;           /* riemann2d.hpp:8 */
typedef enum {
    NPDE_LIM_NONE = 0, NPDE_LIM_UNLIMITED = 1, NPDE_LIM_MINMOD = 2, NPDE_LIM_VANLEER = 3
} npde_limiter;                                                                 /* limiters.hpp:5 */
typedef enum { NPDE_LAYOUT_PLANES = 0, NPDE_LAYOUT_GHOSTED = 1 } npde_layout;
/* Arithmetic of the stage kernels.  EXACT (default): every operation is the IEEE binary64 one the
 * reference performs, in its order, without contraction -- results are bit-identical to the reference
 * solver.  FAST: reciprocal-based divisions, a shortened square root and fused multiply-adds; each
 * step differs from EXACT by rounding-level amounts.  No counterpart in the reference. */
typedef enum { NPDE_ARITH_EXACT = 0, NPDE_ARITH_FAST = 1 } npde_arith;
/* Task layout of the stage kernels (from npde_tuning and the instance widths; every layout gives the same
 * bits): 60-column strips with a halo, row groups, or 64-column strips that solve their seams first. */
typedef enum { NPDE_TASKS_STRIPS = 0, NPDE_TASKS_ROW_GROUPS = 1, NPDE_TASKS_SEAMS = 2 } npde_task_layout;

/* Grid of one instance: replaces make_grid(nx, ny, x_min, x_max, y_min, y_max, bc_x, bc_y)
 * (grid2d.hpp:30-39). */
typedef struct {
    int nx, ny;
    double x_min, x_max, y_min, y_max;
    int bc_x, bc_y; /* npde_bc */
} npde_grid_desc;

/* First non-physical cell, as Solver::report_bad_cell finds it (row-major scan, solver2d.hpp:116-126). */
typedef struct {
    int instance, i, j;
    double rho, p;
} npde_bad_cell;

/* Kernel launch configuration (chosen by npde_autotune or set by hand). */
typedef struct {
    int rows_per_task;    /* band height of one wavefront's task; 0 = automatic (4..64, enough tasks to fill the GPU) */
    int waves_per_block;  /* wavefronts per workgroup: 1, 2, 4 (or 8 within the 256-lane bound) */
    int steps_per_graph;  /* RK2 steps captured per hipGraph launch in npde_run / npde_advance_to */
    int use_graph;        /* 0: plain stream launches */
    int row_groups;       /* when every instance is periodic in x with the same width 64*W, one workgroup of
                             W wavefronts per band exchanges edge columns through LDS instead of recomputing
                             a halo: 1 (default) for W = 1, 2 for W <= 8; 0: never (row groups take
                             precedence over seams where both apply) */
    int seams;            /* when every instance's width is a multiple of 64, row groups do not apply and bands
                             are at most 64 rows, each wavefront owns a 64-column strip and first solves the
                             two strip boundaries of its band (from edge columns the previous stage kernel
                             wrote): 1 (default) for batches of at least 2^18 cells, 2 at any size; 0: never
                             (60-column strips with a 2-column halo) */
    int waves_per_simd;   /* occupancy cap of the stage kernels: at most this many resident wavefronts per SIMD,
                             enforced by reserving LDS per workgroup (0: as many as registers and LDS allow) */
    int net_period;       /* flux-network runs: the network kernel is launched in every net_period-th iteration
                             of the device loop and in the last iteration of every launch (1: every iteration;
                             1..64, default 4).  Instances that land on a frame hold until it runs, so the
                             period changes only time, never a result */
} npde_tuning;

typedef struct npde_solver npde_solver;

/* ---- lifetime: replaces Solver(const Grid&, FluxFn, Limiter, double cfl) (solver2d.hpp:61-64) ---- */
npde_status npde_create(const npde_grid_desc* grid, int flux, int limiter, double cfl, npde_solver** out);
/* A batch of n independent instances of (possibly) different sizes, stepped in the same launches,
 * each with its own dt, time and frame landing. */
npde_status npde_batch_create(int n, const npde_grid_desc* grids, int flux, int limiter, double cfl,
                              npde_solver** out);
npde_status npde_destroy(npde_solver* s);
int npde_num_instances(const npde_solver* s);
const char* npde_last_error(const npde_solver* s);

/* ---- state: the caller's std::vector<Conserved> U (solver2d.hpp:67, 86) ---- */
/* Setting a state also resets that instance's time and step counter to 0. */
npde_status npde_set_state(npde_solver* s, int instance, const double* U, int layout);
npde_status npde_get_state(npde_solver* s, int instance, double* U, int layout);
/* Device-resident variants (4 x ny x nx planes already in device memory, e.g. a torch tensor).  The
 * handle's stream does not wait for other streams: the work that wrote dU must be complete before
 * npde_set_state_device is called (synchronise its stream, or write dU on npde_stream(s)).  Both calls
 * return after their copy has completed, so dU may be freed or reused at once. */
npde_status npde_set_state_device(npde_solver* s, int instance, const double* dU);
npde_status npde_get_state_device(npde_solver* s, int instance, double* dU);
npde_status npde_get_time(npde_solver* s, int instance, double* t, long long* steps);
/* The same for every instance at once (arrays of npde_num_instances entries; either may be NULL). */
npde_status npde_get_times(npde_solver* s, double* t, long long* steps);
/* Where the library keeps an instance's state: *U = device address of its 4 planes, *plane = elements
 * between planes (nx*ny rounded up to a multiple of 64), *instance_stride = elements from this instance's
 * state to the next instance's.  For callers that read (or, between calls, write) the state in place
 * without copies; valid until npde_destroy.  The library launches on npde_stream(s), which does not wait
 * for other streams: a write through this pointer must be issued on npde_stream(s), or be complete
 * (its stream synchronised) before the next call on the handle.  No counterpart in the reference (its U
 * is host memory). */
npde_status npde_state_pointer(npde_solver* s, int instance, double** U, long long* plane, long long* instance_stride);

/* ---- the hot path: Solver::compute_dt (solver2d.hpp:67-83) and Solver::step (solver2d.hpp:86-103) ---- */
/* dt of every instance (array of n); on a non-physical state returns NPDE_ERR_NONPHYSICAL and fills bad. */
npde_status npde_compute_dt(npde_solver* s, double* dt, npde_bad_cell* bad);
/* One SSP-RK2 step of every instance with the given dt (array of n; a negative dt skips that instance). */
npde_status npde_step(npde_solver* s, const double* dt);

/* ---- the driver's time loop (main2d.cpp:224-246), device-resident ---- */
/* Advance every instance until t >= t_target, each landing exactly with dt = min(dt_cfl, t_target - t). */
npde_status npde_advance_to(npde_solver* s, double t_target, npde_bad_cell* bad);
/* As npde_advance_to, but every instance also stops after nsteps more steps (fixed-work timing). */
npde_status npde_advance_steps(npde_solver* s, long long nsteps, double t_target, npde_bad_cell* bad);
/* Pre-roll to t_start, then nout frames at t_start + (t_end - t_start) k / nout, recording the block
 * average of the conserved variables at each of the n_sizes sizes (trajectory.hpp:31-51). */
npde_status npde_run_trajectory(npde_solver* s, double t_start, double t_end, int nout, const int* sizes,
                                int n_sizes, npde_bad_cell* bad);
/* Frames of the last npde_run_trajectory: (nout+1) x 4 x (ny/f) x m float32 for size m. */
npde_status npde_get_trajectory(npde_solver* s, int instance, int m, float* out);
npde_status npde_get_frame_times(npde_solver* s, int instance, double* t);

/* Block average of the current state (trajectory.hpp:39-48): 4 x (ny/f) x m float32, f = nx/m. */
npde_status npde_block_average(npde_solver* s, int instance, int m, float* out);
/* npde_block_average without a device allocation: the values are staged in a static device array of
 * 4 x 512 x 512 float32 (4 MiB per GPU, held until the process ends) instead of a buffer allocated per
 * call.  Results are identical; a block average of more than 4 x 512 x 512 values returns
 * NPDE_ERR_INVALID.  Every handle on a GPU shares that GPU's array, so calls that overlap on one GPU
 * (handles driven from several host threads) race: serialise them.  No counterpart in the reference;
 * for callers that cannot allocate device memory at call time (tight output loops, harnesses that
 * forbid hipMalloc after set-up). */
npde_status npde_block_average_noalloc(npde_solver* s, int instance, int m, float* out);

/* ---- learned flux correction (coarse solver) ---- */
/* Device pointers (4 x ny x nx binary64 each) holding dimensionless per-face factors: every subsequent
 * stage adds to the face flux of conserved component q the factor times the face's flux scale (rho a,
 * rho a^2, rho a^2, rho a^3 for q = 0..3, with the smaller density and the larger sound speed of the two
 * cells sharing the face, from that stage's input state).  corr_x[q][j][i] belongs to the face left of
 * cell (i, j), corr_y[q][j][i] to the face below it (on a transmissive boundary the outermost faces get
 * none).  Pass NULLs to switch the
 * correction off for that instance.  The pointers are read by the kernels until changed: the caller
 * keeps the buffers alive and finishes writing them before the next step.  nx*ny must be a multiple
 * of 64.  No counterpart in the reference. */
npde_status npde_set_flux_correction(npde_solver* s, int instance, const double* corr_x, const double* corr_y);

/* ---- learned flux-correction network run by the library (no counterpart in the reference) ----
 * The fused inference kernel of the flux-correction network (8 features -> four 3x3 periodic convolutions
 * of 32 channels with tanh-GELU -> a 3x3 head of 8 channels -> out_scale * tanh) writes dimensionless
 * factors into correction planes the library allocates for every instance, in the layout and with the
 * meaning of npde_set_flux_correction.  npde_run_trajectory evaluates it at every frame landing, from the
 * instances' states at that time (before their next step): an instance that lands holds until every
 * running instance has landed on that frame, one launch evaluates them all, and they move on together.
 * npde_advance_to keeps the factors of the last evaluation.  weights: device pointer to the 69 packed bfloat16 MFMA fragments of
 * the five layers (64 lanes x 8 values each: first layer 3 k-blocks x 2 n-tiles, three hidden layers 9 x 2,
 * head 9 x 1; npde.learned.model.FluxCorrector.pack_fused writes them); biases: device pointer to
 * 32 + 32 + 32 + 32 + 8 float32.  Both stay the caller's and must stay alive while the network is set;
 * the library's stream does not wait for the caller's streams, so their contents must be complete on the
 * device before this call (e.g. the caller synchronises the stream that wrote them).
 * Every instance must be periodic with nx and ny multiples of 16.  weights = NULL switches it off. */
npde_status npde_set_flux_network(npde_solver* s, const void* weights, const float* biases, double out_scale);
/* In npde_run_trajectory with the network set: switch the network off for an instance (its factors are
 * zero from then on, i.e. the plain coarse solver) at the first frame landing where its CFL signal-speed
 * sum max(|u|+a)/dx + max(|v|+a)/dy exceeds `ratio` times its value at the run's first frame -- a
 * corrected flow draining a cell toward vacuum at finite pressure, whose shrinking steps would hold back
 * the whole batch.  0: never.  Default 2.5. */
npde_status npde_set_flux_network_limit(npde_solver* s, double ratio);
/* Per instance (array of npde_num_instances), 1 if the last run switched its network off. */
npde_status npde_flux_network_switched_off(npde_solver* s, int* off);
/* Evaluate the network now for every instance (factors from the current states). */
npde_status npde_flux_network_eval(npde_solver* s);
/* The error indicator of the learned solver: per instance (host array of n), the largest density the network
 * has seen at its evaluations since the last reset (0 if none); reset != 0 clears it after the read.  Non-
 * finite and non-physical states are the solver's own check (NPDE_ERR_NONPHYSICAL). */
npde_status npde_flux_network_peak(npde_solver* s, float* peak, int reset);
/* Device addresses of an instance's correction planes written by the network (4 x ny x nx binary64 each). */
npde_status npde_flux_network_planes(npde_solver* s, int instance, double** corr_x, double** corr_y);

/* ---- built-in problems of the driver: sod_x, sod_y, lax_liu3, sine2d, kh (main2d.cpp:52-135) ---- */
/* Fills grid and t_end; if U is non-NULL also the initial state as 4 x ny x nx planes.  Call once
 * with U = NULL to learn the grid size. */
npde_status npde_builtin_problem(const char* test, int n, npde_grid_desc* grid, double* t_end, double* U);

/* ---- learned-correction network kernels (no counterpart in the reference: its networks are JAX;
 *      closest analog the periodic 3x3 convolution of surrogate/models.py:16-29) ----
 * Device pointers; activations NHWC float32, values rounded to bfloat16 at the MFMA inputs, FP32
 * accumulation.  npde_conv3x3: input channels a multiple of 8 (x and wpack 16-byte aligned), output
 * channels 16 to 128 or 192 in steps of 16, W a multiple of 8; wpack: bfloat16 weights in the MFMA
 * A-operand order [tap][cin/32][cout/16][64][8] (zero past cin and cout).  epilogue: 0 linear (+bias),
 * 1 +bias then tanh-GELU (pre-activation to `pre` if non-NULL), 2 multiply by GELU'(gpre) (data gradient).
 * npde_conv3x3_wgrad: input channels a multiple of 16 up to 256, output channels up to 128, W a multiple
 * of 4, rows * W a multiple of 16.  Return 0 on success, 1 on shapes or pointers outside these limits. */
int npde_conv3x3(const float* x, int B, int H, int W, int cin, const void* wpack, const float* bias, int cout,
                 int cout_store, float* y, float* pre, int epilogue, const float* gpre, void* stream);
/* Weight/bias gradient of a 3x3 periodic convolution: dw[tap][ci][co], db[co], summed over bands of
 * `rows` image rows in a fixed order (part, bpart: scratch of (B*H/rows)*9*cin*cout and (B*H/rows)*cout). */
int npde_conv3x3_wgrad(const float* x, const float* g, int B, int H, int W, int cin, int cout, int rows,
                       float* part, float* bpart, float* dw, float* db, void* stream);
/* One 16x16x16 BF16 MFMA on row-major 16x16 float inputs (operand-layout check). */
int npde_mfma_probe(const float* a, const float* b, float* d, void* stream);
/* One 16x16x32 BF16 MFMA, D (16 x 16) = A (16 x 32) B (32 x 16), row-major floats (operand-layout check
 * of the fused network kernel). */
int npde_mfma32_probe(const float* a, const float* b, float* d, void* stream);

/* ---- diagnostics (no counterpart in the reference) ---- */
/* Element-wise on the device: q_fast = x / y by the kernels' shared-reciprocal sequence, q_ieee = x / y
 * by the compiler's division, root = sqrt(x); host arrays of length n.  Lets a test prove, on the GPU
 * it runs on, that both quotients are the correctly rounded IEEE ones. */
npde_status npde_selftest_arith(const double* x, const double* y, long long n, double* q_fast, double* q_ieee,
                                double* root);

/* ---- performance: the timing calls replace the reference's per-phase Profile (solver2d.hpp:29-52,
 * printed by main2d.cpp:257-261); the tuning calls have no counterpart (the reference's OpenMP loops
 * have no launch shape) ---- */
npde_status npde_get_tuning(const npde_solver* s, npde_tuning* t);
npde_status npde_set_tuning(npde_solver* s, const npde_tuning* t);
/* Time candidate launch configurations on this handle's own instances (state is saved and restored)
 * and keep the fastest; results are cached per (sizes, bc, flux, limiter) in cache_path if non-NULL. */
npde_status npde_autotune(npde_solver* s, const char* cache_path, npde_tuning* chosen);
/* Select the stage kernels' arithmetic (npde_arith) for the following steps of this handle. */
npde_status npde_set_arithmetic(npde_solver* s, int arith);
npde_status npde_get_arithmetic(const npde_solver* s, int* arith);
/* The task layout (npde_task_layout) the current tuning gives this handle's instances. */
npde_status npde_get_task_layout(const npde_solver* s, int* layout);
/* Resident wavefronts per SIMD the stage-2 kernel reaches under the current tuning (HIP's occupancy
 * calculator on its registers, its LDS including an occupancy cap's reservation, and wavefront slots). */
npde_status npde_get_occupancy(const npde_solver* s, int* waves_per_simd);
/* Average device milliseconds of the stage-1 and stage-2 kernel launches over the next nsteps steps of
 * every instance, each with the CFL dt of its state as in npde_advance_steps, launched one at a time with
 * HIP events on the handle's stream around each stage kernel (the state advances). */
npde_status npde_time_stages(npde_solver* s, int nsteps, double* ms_stage1, double* ms_stage2);
/* Milliseconds of device time of the last npde_step / npde_advance_to / npde_run_trajectory. */
npde_status npde_last_elapsed_ms(const npde_solver* s, double* ms);
/* Per-kernel profile: replaces the reference's per-phase Profile (solver2d.hpp:29-52, printed by
 * main2d.cpp:261).  While profiling is on, the handle launches its kernels one at a time on its stream
 * (no hipGraphs) with HIP events around each launch, so the profiled run is slower, as the reference's
 * timed run is.  Switching it on clears the accumulated times. */
typedef enum {
    NPDE_PROF_STAGE1 = 0,    /* stage 1: slopes, face fluxes, U1 */
    NPDE_PROF_STAGE2 = 1,    /* stage 2: the same from U1, U^{n+1}, and the next dt's wave-speed partials */
    NPDE_PROF_FINISH = 2,    /* per-instance dt reduction and run-state decision */
    NPDE_PROF_FRAME = 3,     /* frame block averages and landings */
    NPDE_PROF_WAVESPEED = 4, /* wave speeds of a fresh state (start of a run, npde_compute_dt) */
    NPDE_PROF_NETWORK = 5,   /* flux-network evaluations */
    NPDE_PROF_KINDS = 6
} npde_profile_kind;
npde_status npde_set_profiling(npde_solver* s, int on);
/* Accumulated device milliseconds and launch counts per npde_profile_kind (arrays of NPDE_PROF_KINDS;
 * either may be NULL). */
npde_status npde_get_profile(npde_solver* s, double* ms, long long* launches);
/* Stream the handle launches on (hipStream_t), for callers that order their own work after it. */
void* npde_stream(npde_solver* s);

#ifdef __cplusplus
}
#endif
#endif /* NPDE_H */

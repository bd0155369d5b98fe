/*
 * se_gpu.h -- device-resident Spectral Ewald engine of libse_fgg_amd.so.
 *
 * No reference counterpart: the reference has no device boundary.  This is the
 * interface the fast path runs behind.  A plan owns one grid geometry, its
 * rocFFT plans and all scratch in HBM; particle arrays are DEVICE pointers in the
 * reference's layout (x SoA: x[0..N) y[N..2N) z[2N..3N); q[N]; phi[N]; F SoA[3N])
 * and every call is asynchronous on the plan's stream (an FP32 P = 16 plan also owns
 * one internal stream, forked from and joined back to the plan's stream inside the
 * gather, so the call stays ordered on the plan's stream and can be graph-captured).
 * Functions return 0 or a
 * negative SE_ERR_* code; se_gpu_last_error() returns the last message.
 *
 * The solve is the reference driver (spectral_ewald.c:83-165) re-staged for the GPU:
 *   se_spread  = expand + spread + wrap   (particles bucketed by grid tile, P^3
 *                factors built in registers, LDS sub-grids, fixed-order halo merge)
 *   se_kspace  = FFT, Green's-function scaling, inverse FFT (rocFFT R2C/C2R)
 *   se_interp  = extend + interpolate     (LDS-staged P^3 gather per particle)
 */
#ifndef SE_FGG_AMD_SE_GPU_H
#define SE_FGG_AMD_SE_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SE_OK 0
#define SE_ERR_ARG -1
#define SE_ERR_HIP -2
#define SE_ERR_FFT -3
#define SE_ERR_CAPACITY -4
#define SE_ERR_RANGE -5

#define SE_PREC_FP64 0
#define SE_PREC_FP32 1

#define SE_Q_POTENTIAL 1
#define SE_Q_FORCE 2

typedef struct se_plan se_plan;

typedef struct {
    int periodicity;   /* 3 (SE3P), 2 (SE2P: periodic in x and y, free in z) or
                          1 (SE1P: periodic in z, free in x and y) */
    int M;             /* grid points per period of the periodic directions (x and y for SE2P,
                          z for SE1P, all three for SE3P); h = box / M */
    int P;             /* Gaussian support, 1..32 */
    double box;        /* period L (SE3P: the cube side; SE2P: the periodic square's side;
                          SE1P: the period in z) */
    double xi;         /* Ewald parameter */
    double m;          /* Gaussian shape; 0 selects the driver's 0.95*sqrt(pi*P) */
    int precision;     /* SE_PREC_FP64 or SE_PREC_FP32 (spread, FFT and gather arithmetic) */
    int n_max;         /* particle capacity */
    int tile;          /* tile side in grid cells (at most 16), 0 = automatic */
    int chunk;         /* max particles per work item, 0 = automatic */
    int block;         /* threads per work-group (<= 512) of every kernel (the P = 16 MFMA gather
                          uses at most 256; the P = 16 column spread and gather always run
                          512); 0 = each kernel's own default: 768 for the FP32
                          register-ring spread, 256 for the MFMA gather, 512 otherwise */
    int flags;         /* SE_PLAN_* bits */
    double slab;       /* SE2P: particles have 0 <= z < slab; SE1P: particles have
                          0 <= x, y < slab (ignored for SE3P) */
    int Mz;            /* grid points of the zero-padded free direction(s), 0 = automatic (the
                          smallest FFT-friendly size that holds the particles' range, the support
                          and the kernel truncation radius): z for SE2P, x and y for SE1P;
                          ignored for SE3P */
} se_plan_desc;

/* Use the generic one-particle-per-wave kernels even where the FP32 register-ring
 * (pencil) kernels exist (for A/B comparisons and tests). */
#define SE_PLAN_GENERIC_KERNELS 1
/* Keep the P = 16 FP32 gather on the vector register-ring kernel instead of the matrix
 * cores (MFMA); for A/B comparisons and tests. */
#define SE_PLAN_NO_MFMA 2
/* Spread sparse tiles with the pencil register ring too, instead of the plane kernel that
 * keeps each (x, z) column of the tile's sub-grid in registers (for A/B comparisons and
 * tests). */
#define SE_PLAN_NO_PLANE_SPREAD 4
/* Run the dense work items of the P = 16 FP32 path on the 4 x 1 pencil kernels instead of
 * the 1 x 1 column kernels, whose matrix-core products take sixteen particles of one
 * window-start column at a time (for A/B comparisons and tests).  Without this flag the
 * column kernels take each work item whose column groups hold 3.5 particles or more on
 * average (decided on the device per item); the column gather forms its products from
 * fp16 (hi, lo) pairs of the FP32 operands, good to about 2^-22 each. */
#define SE_PLAN_NO_COLUMN 8

int se_plan_create(se_plan** plan, const se_plan_desc* desc, void* hip_stream);
int se_plan_destroy(se_plan* plan);
int se_plan_set_stream(se_plan* plan, void* hip_stream);
/* Bytes of HBM the plan holds (for admission control). */
size_t se_plan_bytes(const se_plan* plan);
/* Bytes a plan with this description would hold, without creating it. */
size_t se_plan_bytes_for(const se_plan_desc* desc);

/*
 * SE2P: the potential and force are the SE2P Ewald k-space sum (SE_direct.c:256-338:
 * the k_par != 0 terms plus the k_par = 0 term).  The free direction is zero-padded to
 * Mz points and uses the Coulomb kernel truncated at a radius that covers the slab and
 * the Gaussian supports, which the padded FFT convolves exactly; z grid index 0 sits
 * at z0 = a + h/2 with a = -(P_half + 1) h (the reference's SE2P grid origin
 * convention, SE_fgg.c:753-764).  Particles with z outside [0, slab) are left out and
 * counted as discarded.
 */

/*
 * SE1P: the potential and force are the SE1P Ewald k-space sum (SE_direct.c:90-176: the
 * k_z != 0 terms plus the k_z = 0 term, whose constant follows the reference's
 * gamma + ln(rho^2 xi^2) + E1(rho^2 xi^2) also for a net charge).  x and y are zero-padded
 * to Mz points each and use the in-plane Coulomb kernel of each z mode truncated at a
 * radius that covers the square's diagonal and the Gaussian supports; x and y grid index
 * 0 sits at a + h/2 with a = -(P_half + 1) h (the reference's SE1P grid origin convention,
 * SE_fgg.c:836-846).  Particles with x or y outside [0, slab) are left out and counted as
 * discarded.
 */

/* Grid shape of the plan, [x][y][z] with z fastest: dims = {M, M, Mz} for SE3P and SE2P,
 * {Mz, Mz, M} for SE1P; *z0 = z coordinate of grid index 0 in z (0 for SE3P and SE1P). */
int se_plan_grid_shape(const se_plan* plan, int dims[3], double* z0);

/* Full k-space solve.  phi and/or force may be NULL (not computed). */
int se_solve(se_plan* plan, int N, const double* d_x, const double* d_q, double* d_phi, double* d_force);

/* se_solve replayed from a HIP graph: the first call with a given (N, d_x, d_q, d_phi,
 * d_force) captures the solve on the plan's stream, later calls with the same arguments
 * launch the captured graph (one launch instead of one per kernel).  The plan keeps the
 * 32 most recently used graphs.  The arrays' contents may change between calls; the
 * arguments may not (a different argument set captures a new graph).  Falls back to
 * se_solve while timing is enabled. */
int se_solve_graph(se_plan* plan, int N, const double* d_x, const double* d_q, double* d_phi, double* d_force);

/*
 * Grouped solves (no reference counterpart: the reference solves one system per call).
 * A group plan holds up to `max_systems` independent systems of the plan's geometry
 * (desc->n_max = particle capacity summed over the systems) and solves them together:
 * each stage is one launch over all systems (sort keys, work items and partials are
 * indexed by system * tiles + tile, grids and spectra are stacked, the FFTs are batched).
 * All systems of one call must request the same outputs (phi for all or none, force for
 * all or none).  Each system's result equals its se_solve result up to the rounding
 * freedom stated below for se_batch_*.  se_solve / se_spread on a group plan solve one
 * system.  Descriptors travel through a pinned staging ring, so se_solve_group is not
 * for stream capture.
 */
typedef struct {
    int N;
    const double* d_x;   /* SoA [3N], device */
    const double* d_q;   /* [N], device */
    double* d_phi;       /* [N] or NULL */
    double* d_force;     /* SoA [3N] or NULL */
} se_system;

int se_plan_create_group(se_plan** plan, const se_plan_desc* desc, int max_systems, void* hip_stream);
size_t se_plan_bytes_for_group(const se_plan_desc* desc, int max_systems);
int se_solve_group(se_plan* plan, int n_systems, const se_system* systems);

/* Stages of se_solve, for timing and for callers that bring their own grid. */
int se_spread(se_plan* plan, int N, const double* d_x, const double* d_q);
int se_kspace(se_plan* plan);
int se_interp(se_plan* plan, double* d_phi, double* d_force);

/* Grid of the plan (row-major [x][y][z], z fastest, M x M x Mz; float or double per precision). */
int se_plan_grid(se_plan* plan, void** d_grid, int* elem_bytes);

/*
 * Instrumentation: when enabled, se_solve/se_spread/se_interp record HIP events
 * on the plan's stream around the spread kernel, the merge kernel, the gather
 * kernel, the k-space stage and the whole spread/interp stages.  se_plan_timings
 * synchronises the stream and returns milliseconds of the last call:
 * ms[0]=spread stage (bucketing+spread+merge), ms[1]=spread kernel, ms[2]=merge,
 * ms[3]=k-space, ms[4]=interp stage, ms[5]=gather kernel, ms[6]=bucketing.
 */
int se_plan_enable_timing(se_plan* plan, int on);
int se_plan_timings(se_plan* plan, float ms[8]);

/* Tile/work-item geometry chosen for the plan: out[0]=tile, [1]=chunk, [2]=block,
 * [3]=number of tiles, [4]=work items of the last spread (syncs), [5]=particles of the
 * last spread discarded (non-finite, or outside the SE2P slab), [6]=sub-grid side S,
 * [7]=1 if the pencil kernels run. */
int se_plan_info(se_plan* plan, int out[8]);

/* Energy E = 1/2 sum phi q (FP64, on the device), into *d_energy (device double). */
int se_energy(se_plan* plan, int N, const double* d_phi, const double* d_q, double* d_energy);

const char* se_gpu_last_error(void);

/*
 * Tiling autotuner (no reference counterpart).  se_tune times the full solve (the outputs
 * in `quantities`: SE_Q_POTENTIAL and/or SE_Q_FORCE) of `systems` synthetic uniform
 * systems of N particles each on desc's geometry -- one plan for systems = 1, a grouped
 * solve (se_solve_group) otherwise -- under every candidate tiling -- tile side, maximum
 * particles per work item (the bucket cap that splits heavy tiles) and threads per
 * work-group -- and records the fastest in a process-wide table keyed by (periodicity,
 * M, P, precision, ceil(log2 N)).  Plans created afterwards with tile = chunk = block = 0
 * take the recorded tiling of their geometry at the nearest size band (group plans use
 * particles per system), including the plans the batch executor creates.  out = {tile, chunk, block}
 * (block 0: the kernels' own defaults, one of the candidates) and *ms =
 * its solve time.  se_tune allocates its own scratch and synchronises `hip_stream`.
 * se_tune_save / se_tune_load write and read the table as text (load returns the number
 * of entries read, or an error code); se_tune_lookup returns 1 and the tiling a plan of
 * desc's geometry with N particles would take, else 0.
 */
int se_tune(const se_plan_desc* desc, int N, int systems, int quantities, void* hip_stream, int out[3], float* ms);
int se_tune_save(const char* path);
int se_tune_load(const char* path);
int se_tune_clear(void);
int se_tune_lookup(const se_plan_desc* desc, int N, int out[3]);

/*
 * Parameters for an accuracy target (host only).  Applies the error estimates that
 * SE_init_params (SE_general.c:517-572) prints -- it replaces no entry point of the
 * reference, which takes P from the caller -- to choose the smallest support P <= 32
 * whose estimated rms gridding error for `quantities` (SE_Q_POTENTIAL and/or SE_Q_FORCE;
 * the larger estimate when both) is at most tol, for a cubic box of side L, Ewald
 * parameter xi and Q = sum q^2:
 *   potential: sqrt(Q xi L) / L * exp(-pi P c^2 / 2)                 (SE_general.c:530)
 *   force:     Q sqrt(xi L) / L^2 * exp(-pi P c^2 / 2) * 2 pi c^2 P  (SE_general.c:544)
 * with c = 0.95.  *m = c sqrt(pi P) (the Gaussian shape), *M = the reference's grid
 * (k_inf = 2 xi L, M = ceil(2 k_inf + 1) + 2, at least P, rounded up to a multiple of
 * 4), *est (optional) = the estimate at P.  The FP32 register-ring kernels cover
 * P in {8, 11, 16}; other P run the generic kernels.  SE_ERR_ARG when no P <= 32 meets
 * tol or an argument is out of range.
 */
int se_select_params(double tol, double L, double xi, double Q, int quantities, int* P, int* M, double* m,
                     double* est);

/*
 * Real-space and self terms of the Ewald sum, so that real + k-space (se_solve) + self
 * gives the full Coulomb sum of the periodic system.  Replaces the reference's direct
 * evaluations SE3P_direct_real_rc (SE_direct.c:464-505), SE2P_direct_real_rc (:224-254),
 * SE1P_direct_real_rc (:50-88) and SE*P_direct_self (:178-183, 340-345, 550-561), which
 * loop over all pairs and image layers; here particles are bucketed into cells no smaller
 * than the cutoff and each particle visits the neighbouring cells.
 *   phi_m   = sum_{n, images, r <= rc, not the particle itself} q_n erfc(xi r)/r
 *             (- 2 xi/sqrt(pi) q_m when `self` is set)
 *   force_m = (q_m/2) grad phi_m, the convention of the reference's k-space force
 *             (SE_direct.c:417-461), so it adds to se_solve's force.
 * Periodic directions fold positions into [0, box); in free directions particles must
 * lie in [0, box) and others are left out (their outputs are not written).  Outputs are
 * assigned, or added into phi / force when `accumulate` is set.  Asynchronous on the
 * plan's stream; arrays are device pointers in the reference layout.
 */
typedef struct {
    int periodicity;   /* 3, 2 (x and y periodic) or 1 (z periodic) */
    double box[3];     /* period of each periodic direction; extent of the particles' range in a free one */
    double xi;         /* Ewald parameter */
    double rc;         /* cutoff radius */
    int self;          /* 1: include the self term in phi */
    int accumulate;    /* 1: add into phi / force instead of assigning */
} se_real_desc;

typedef struct se_real_plan se_real_plan;
int se_real_plan_create(se_real_plan** plan, const se_real_desc* desc, int n_max, void* hip_stream);
int se_real_plan_destroy(se_real_plan* plan);
/* phi and/or force may be NULL (not computed). */
int se_real_solve(se_real_plan* plan, int N, const double* d_x, const double* d_q, double* d_phi, double* d_force);

/*
 * se_real_solve without a plan and without allocating.  No counterpart in the reference
 * (which allocates nothing on a device); it is for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops and harnesses that forbid
 * hipMalloc.  It computes what se_real_solve computes on a plan created with
 * n_max = N (the same cells, so the same arrays bit for bit), asynchronously on
 * hip_stream, and calls no hipMalloc, hipFree or allocating library routine: its cell
 * lists, sort buffers and scan temporary live in one static device array of
 * SE_REAL_SCRATCH_BYTES per GPU, held until the process ends.  N is limited by that
 * array (52 bytes per particle plus 8 per cell, at most about 4 cells per particle: at
 * least 750 000 particles fit): a larger N returns SE_ERR_CAPACITY before anything is
 * launched.  Calls that overlap on one GPU
 * share the array and race: serialise them (one stream, or events).
 */
#define SE_REAL_SCRATCH_BYTES (64u << 20)
int se_real_solve_noalloc(const se_real_desc* desc, int N, const double* d_x, const double* d_q, double* d_phi,
                          double* d_force, void* hip_stream);

/*
 * Batched and served solves (no reference counterpart: the reference solves one system
 * per call).  An executor owns `n_slots` HIP streams and a per-slot cache of plans
 * keyed by geometry; jobs are independent systems in device memory.  se_batch_submit
 * queues a job on the slots in turn; se_batch_try_submit takes an idle slot (one that
 * already holds the job's captured graph first, then one that holds a plan for the job)
 * or returns SE_BUSY (the admission step of a serving loop).  se_batch_submit replays the
 * job through se_solve_graph (captured on the first submission of the same arguments on
 * that slot); se_batch_try_submit replays a graph the chosen slot already holds and
 * otherwise solves eagerly, so it never captures on a request's path (a serving loop can
 * capture its jobs ahead of traffic by submitting each once per slot).
 * Plans of all slots together
 * stay within `hbm_budget` bytes (0 = 64 GiB); least recently used plans of idle slots
 * are freed to make room, and a job whose plan cannot fit fails with SE_ERR_CAPACITY.
 * Jobs on one slot run in order, jobs on different slots overlap.  A job's result does
 * not depend on its slot or its neighbours: particles inside a sort key are ordered by
 * index, and FP32 plans add integer-valued FP64 values in their LDS sub-grids (exact in
 * any order), so FP32 results repeat bit for bit, alone and in any mix (keys of more than
 * 65536 particles, over 16k per grid cell, keep the order of the counting atomics).  FP64
 * plans add FP64 products with LDS atomics and repeat to about 1e-16.
 * The executor is not thread-safe; the caller serialises calls.
 */
#define SE_BUSY 1

typedef struct {
    int periodicity;   /* 3, 2 or 1 */
    int M, P;
    double box, xi, m; /* m = 0: 0.95 sqrt(pi P) */
    double slab;       /* SE2P slab height; SE1P side of the x-y square */
    int precision;     /* SE_PREC_FP64 / SE_PREC_FP32 */
} se_job_shape;

typedef struct {
    se_job_shape shape;
    int N;
    const double* d_x; /* SoA [3N], device */
    const double* d_q; /* [N], device */
    double* d_phi;     /* [N] or NULL */
    double* d_force;   /* [3N] or NULL */
} se_job;

typedef struct se_batch se_batch;
int se_batch_create(se_batch** batch, int n_slots, size_t hbm_budget);
int se_batch_destroy(se_batch* batch);
int se_batch_submit(se_batch* batch, const se_job* job, int* ticket);
int se_batch_try_submit(se_batch* batch, const se_job* job, int* ticket);
/* Solve a set of jobs as grouped solves: jobs with the same shape and the same requested
 * outputs form one group, solved by a cached group plan (se_solve_group) on one slot;
 * groups overlap on different slots.  One ticket completes when every job is done. */
int se_batch_run(se_batch* batch, const se_job* jobs, int n_jobs, int* ticket);
/* Build (or grow) the plan for `shape` with capacity n_max on every slot ahead of time. */
int se_batch_prepare(se_batch* batch, const se_job_shape* shape, int n_max);
/* 1 when the ticket's job is complete, 0 while it runs, < 0 on error. */
int se_batch_poll(se_batch* batch, int ticket);
int se_batch_wait(se_batch* batch, int ticket);
int se_batch_wait_all(se_batch* batch);
/* HBM held by cached plans and their count. */
int se_batch_stats(se_batch* batch, size_t* plan_bytes, int* n_plans);

#ifdef __cplusplus
}
#endif
#endif

/*
 * se_fgg.h -- drop-in C ABI for the fast Gaussian gridding core of
 * davoudss/SE_fast_gridding, implemented by libse_fgg_amd.so on an MI355X.
 *
 * Every declaration below replaces the reference declaration cited beside it
 * (SE_fgg.h line numbers) with the same name, argument order and meaning.  The
 * struct layouts are byte-for-byte those of the reference so existing callers
 * link unchanged.  Arrays passed in and out are HOST arrays, as in the
 * reference; each call copies what it needs to the GPU, runs hand-written
 * gfx950 kernels, copies results back and returns when they are complete.
 *
 * Differences a caller can observe:
 *   - Periodicity is a run-time choice (se_fgg_set_periodicity) instead of the
 *     reference's -DTHREE_PERIODIC / -DTWO_PERIODIC / -DONE_PERIODIC.  Default 3P.
 *   - The SSE/AVX/P8/P16/u8 variants and their dispatchers all run the same
 *     GPU kernels; their contract is numerical (FP64, agreement ~1e-15 relative).
 *   - Spread, wrap and reorder results do not depend on a summation order the
 *     caller can rely on (GPU atomics); differences are at the 1e-15 level.
 *   - On a HIP failure the void functions print one line to stderr and record
 *     the message for se_gpu_last_error() (declared in se_gpu.h).
 */
#ifndef SE_FGG_AMD_SE_FGG_H
#define SE_FGG_AMD_SE_FGG_H

#ifdef __cplusplus
extern "C" {
#endif

#define PI 3.14159265358979323846      /* SE_fgg.h:11 */
#define FGG_INF 1.79769e+308           /* SE_fgg.h:12 */
#define P_MAX 32                       /* SE_fgg.h:36: largest supported support P */

/* SE_fgg.h:14-15: column-major (periodic grid) and row-major (extended grid) indexing. */
#define __IDX3_CMAJ(II, IJ, IK, N1, N2) ((II) + (IJ) * (N1) + (IK) * (N1) * (N2))
#define __IDX3_RMAJ(II, IJ, IK, N2, N3) ((II) * (N2) * (N3) + (IJ) * (N3) + (IK))

/* SE_fgg.h:81-97.  Translation units that also see HIP's `half` type define
 * SE_FGG_NO_INLINE_HELPERS before including this header. */
#ifndef SE_FGG_NO_INLINE_HELPERS
static inline int is_odd(int p) { return p & 1; }
static inline int isnot_div_by_4(int p) { return p & 3; }
static inline int half(int p) { return is_odd(p) ? (p - 1) / 2 : p / 2; }
#endif

/* SE_fgg.h:100-118 */
typedef struct {
    double* H;
    double* zs;
    double* zx;
    double* zy;
    double* zz;
    int* idx;
    double* zfx;
    double* zfy;
    double* zfz;
    int free_zs;
    int free_fgg_expa;
} SE_FGG_work;

/* SE_fgg.h:121-134 */
typedef struct {
    int N;
    int P;
    int P_half;
    int dims[3];
    int npdims[3];
    double c;
    double d;
    double h;
    double a;
    double b;
} SE_FGG_params;

/* SE_fgg.h:136-139 */
typedef struct {
    int idx_on_grid, idx_in_array;
} idx_reorder_t;

/* SE_fgg.h:142-148 */
typedef struct {
    double* x;
    double* q;
    double* phi;
} SE_state;

/* Periodicity selector replacing the reference's compile-time macros (SE_fgg.h:18-33). */
#define SE_FGG_3P 3
#define SE_FGG_2P 2
#define SE_FGG_1P 1
void se_fgg_set_periodicity(int periodicity);   /* no reference counterpart */
int se_fgg_get_periodicity(void);               /* no reference counterpart */

/* Parameter packing: SE_fgg.h:151-154; SE1P_ is defined at SE_fgg.c:113 without a prototype. */
void SE_FGG_pack_params(SE_FGG_params*, int, int, int, int, int, double, double);
void SE2P_FGG_pack_params(SE_FGG_params*, int, int, int, int, int, double, double, double);
void SE1P_FGG_pack_params(SE_FGG_params*, int, int, int, int, int, double, double, double, double);

/* Workspace: SE_fgg.h:157-164 */
void SE_FGG_allocate_workspace(SE_FGG_work*, const SE_FGG_params*, int, int);
void SE_FGG_allocate_workspace_SSE_force(SE_FGG_work*, const SE_FGG_params*, int, int);
double* SE_FGG_allocate_grid(const SE_FGG_params*);
double* SE_FGG_allocate_vec(int);
void SE_FGG_free_workspace(SE_FGG_work*);
void SE_FGG_free_workspace_SSE_force(SE_FGG_work*);

/* Spread, potential workspace: SE_fgg.h:167-179 */
void SE_FGG_grid(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_SSE_dispatch(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_SSE(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_SSE_u8(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_SSE_P16(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_dispatch(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_P16(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_P8(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_u8(SE_FGG_work*, const SE_state*, const SE_FGG_params*);

/* Spread, force workspace: SE_fgg.h:182-194 */
void SE_FGG_grid_split_SSE_dispatch_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_SSE_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_SSE_u8_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_SSE_P16_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_dispatch_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_P16_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_P8_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_u8_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_grid_split_AVX_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);

/* Expansion vectors: SE_fgg.h:197-198 */
void SE_FGG_expand_all(SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_expand_all_SSE_force(SE_FGG_work*, const SE_state*, const SE_FGG_params*);

/* Interpolation, potential: SE_fgg.h:201-217 */
void SE_FGG_int(double*, const SE_FGG_work*, const SE_state*, const SE_FGG_params*);
void SE_FGG_int_force(double*, const SE_FGG_work*, SE_state*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_dispatch(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_u8(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_P8(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_P16(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_dispatch(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_P8(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_P16(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_u8(double*, const SE_FGG_work*, const SE_FGG_params*);

/* Interpolation, force (argument order differs from SE_FGG_int_force): SE_fgg.h:220-236 */
void SE_FGG_int_split_SSE_dispatch_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_P8_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_u8_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_SSE_P16_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_dispatch_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_P16_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_P8_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_u8_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_int_split_AVX_force(double*, SE_state*, const SE_FGG_work*, const SE_FGG_params*);

/* SE_fgg.h:239 */
void SE_FGG_base_gaussian(SE_FGG_work*, const SE_FGG_params*);

/* Wrap / extend: SE_fgg.h:242-249.  Periodic grid is column-major.  2P: z is free
 * and the periodic grid holds dims0*dims1*npdims2 values.  1P: the reference's
 * index (i + j*dims0 + k*dims0*dims1 with i < npdims0, j < npdims1) is kept, so
 * the buffer must hold (npdims0-1)+(npdims1-1)*dims0+(dims2-1)*dims0*dims1+1. */
void SE_FGG_wrap_fcn(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE2P_FGG_wrap_fcn(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE1P_FGG_wrap_fcn(double*, const SE_FGG_work*, const SE_FGG_params*);
void SE_FGG_extend_fcn(SE_FGG_work*, const double*, const SE_FGG_params*);
void SE2P_FGG_extend_fcn(SE_FGG_work*, const double*, const SE_FGG_params*);
void SE1P_FGG_extend_fcn(SE_FGG_work*, const double*, const SE_FGG_params*);

/* Test systems: SE_fgg.h:252-256 (host-side input generation, as in the reference). */
void SE_init_system(SE_state*, const SE_FGG_params*);
void SE_init_unit_system(SE_state*, const SE_FGG_params*);
void SE_free_system(SE_state*);

/* Utilities: SE_fgg.h:259-266 */
double SE_gettime(void);
int SE_prod3(const int[3]);
void SE_fp_set_zero(double*, int);
/* Reorders x,q by the extended-grid index of each particle (stable: ties keep input
 * order; the reference's qsort leaves tie order unspecified). */
void SE_FGG_reorder_system(SE_state*, const SE_FGG_work*, const SE_FGG_params*);
double calc_energy(SE_state, int);

#ifdef __cplusplus
}
#endif
#endif

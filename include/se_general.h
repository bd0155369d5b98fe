/*
 * se_general.h -- drop-in C ABI for the k-space utilities and the SE3P driver of
 * davoudss/SE_fast_gridding (SE_general.h, spectral_ewald.c), implemented by
 * libse_fgg_amd.so.  Names, argument orders and struct layouts are the
 * reference's.  FFTW is not used: the driver runs rocFFT R2C/C2R transforms on
 * the GPU (the real-input c2c of the reference computes the same numbers).
 */
#ifndef SE_FGG_AMD_SE_GENERAL_H
#define SE_FGG_AMD_SE_GENERAL_H

#include "se_fgg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ONE_4PI_EPS0 138.9354491512000038   /* SE_general.h:15 */

typedef double t_complex[2];                /* SE_general.h:17 */

/* SE_general.h:20-24 */
typedef struct {
    double m, c, box[3], h, xi, w, eta;
    int P, M, N;
} SE_opt;

void parse_params(SE_opt*, double);                                        /* SE_general.h:27 */
void k_vec(int, double*, double*, double*, double*);                       /* SE_general.h:30 */
void scaling(double, double*, int, int, int, double*);                     /* SE_general.h:33 */
void product_rc(t_complex*, double*, t_complex*, int, int, int, int);      /* SE_general.h:40 */
void SE_FGG_FCN_params(SE_FGG_params*, const SE_opt*, int);                /* SE_general.h:44-45 */
void SE_fgg_int_force(double*, double*, double*, int, SE_opt, double*);    /* SE_general.h:54 */
void copy_r2c(double*, t_complex*, int);                                   /* SE_general.h:73 */
void copy_c2r(t_complex*, double*, int);                                   /* SE_general.h:74 */
double SE_init_params(int*, int*, double*, double, double, double, const double*, int); /* SE_general.h:76-77 */

/*
 * spectral_ewald.c:20-22.  SE3P k-space potential (force_phi = phi[N]) or force
 * (force_phi = F[3N], the reference's (q/2) grad(phi) convention; phi_energy, if
 * not NULL, also receives phi).  With se_fgg_set_periodicity(2) it solves SE2P instead
 * (periodic in x and y with period box[0], particles in 0 <= z < box[2]), with
 * se_fgg_set_periodicity(1) SE1P (periodic in z with period box[2], particles in
 * [0, box[0])^2 in x and y); M counts grid points per period.  The reference driver
 * is SE3P only.  Which quantity is computed is set by
 * se_set_quantity() (default potential), replacing the reference's -DPOTENTIAL /
 * -DFORCE.  Arithmetic is FP64 (the reference's precision) unless
 * se_set_precision(SE_FP32) was called.  Returns 0 on success, nonzero on a HIP
 * error (the reference always returns 0).
 */
int spectral_ewald(double* x, double* q, SE_opt opt, double xi, double* force_phi, double* phi_energy);

#define SE_POTENTIAL 1
#define SE_FORCE 2
void se_set_quantity(int quantity);      /* no reference counterpart: run-time -DPOTENTIAL/-DFORCE */
#define SE_FP64 0
#define SE_FP32 1
void se_set_precision(int precision);    /* no reference counterpart */

#ifdef __cplusplus
}
#endif
#endif

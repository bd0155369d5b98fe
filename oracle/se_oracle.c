/*
 * se_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU, FP64, single-file C restatement of the Spectral Ewald fast Gaussian
 * gridding path of davoudss/SE_fast_gridding.  It is the checker the parity
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg compare the
 * HIP library against.  Nothing in the product library links, loads or calls
 * this file.
 *
 * Every routine names the reference lines it restates.  Evaluation order of the
 * gridding sums follows the reference's scalar kernels (SE_fgg.c) so that this
 * oracle agrees with the compiled reference to a few ulp; the direct Ewald sums
 * use a structure-factor form that is algebraically identical to the
 * reference's pairwise sums (a pairwise form is kept beside it for tiny-N
 * cross-checks).
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   or_expand / or_base_gaussian / or_grid_split / or_wrap / or_extend /
 *   or_int_split / or_int_split_force: golden vectors produced by the compiled
 *     reference (tests/golden/*.npz) + closed form W = q d exp(-c|x-g|^2).
 *   or_fft3d: DFT definition on tiny grids and numpy.fft.
 *   or_scaling / or_kvec: closed form.
 *   or_se3p: direct k-space sums (two independent forms) to the FGG error floor.
 *   or_direct3p_*: pairwise form (reference's hs-weighted cosine sum) vs
 *     structure-factor form; force vs finite differences of the potential.
 *   or_direct2p_fd + or_direct2p_k0: the slab limit of or_direct3p_sf in an
 *     L x L x Lz box (minus the k_z = 0 quadratic term), to 2e-15 at Lz = 8 L.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Same field order as SE_FGG_params (SE_fgg.h:121-134). */
typedef struct {
    int N, P, P_half, dims[3], npdims[3];
    double c, d, h, a, b;
} or_params;

enum { OR_3P = 3, OR_2P = 2, OR_1P = 1 };

/* SE_fgg.h:94-97 */
static int or_half(int p) { return (p & 1) ? (p - 1) / 2 : p / 2; }

/* SE_fgg.c:17-25: non-negative modulus. */
static int or_mod(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

/* SE_fgg.c:69-132 (SE_FGG_pack_params / SE2P_ / SE1P_). */
void or_pack_params(or_params* p, int N, int M0, int M1, int M2, int P,
                    double c, double h, double a, double b)
{
    p->N = N; p->P = P; p->P_half = or_half(P);
    p->c = c; p->d = pow(c / M_PI, 1.5); p->h = h; p->a = a; p->b = b;
    p->dims[0] = M0; p->dims[1] = M1; p->dims[2] = M2;
    p->npdims[0] = M0 + P; p->npdims[1] = M1 + P; p->npdims[2] = M2 + P;
}

/* SE_general.c:16-29: h, w, eta, c from (L, M, P, m, xi). out = {h, w, eta, c} */
void or_parse_params(double L, int M, int P, double m, double xi, double out[4])
{
    double h = L / (double)M;
    double w = (double)P * h / 2.0;
    double eta = (2.0 * w * xi / m) * (2.0 * w * xi / m);
    out[0] = h; out[1] = w; out[2] = eta; out[3] = 2.0 * xi * xi / eta;
}

/* SE_fgg.c:508-535: zs[i][j][k] = d exp(-c h^2 (oi^2+oj^2+ok^2)), row-major P^3. */
void or_base_gaussian(const or_params* p, double* zs)
{
    const int P = p->P, ph = p->P_half, lo = (P & 1) ? ph : ph - 1;
    const double h = p->h;
    int n = 0;
    for (int i = -lo; i <= ph; i++)
        for (int j = -lo; j <= ph; j++)
            for (int k = -lo; k <= ph; k++)
                zs[n++] = p->d * exp(-p->c * ((i * h) * (i * h) + (j * h) * (j * h) + (k * h) * (k * h)));
}

/*
 * One particle's expansion.  3P: SE_fgg.c:539-621, force factors :625-722;
 * 2P (z free, origin a+h/2): :727-819; 1P (x,y free, origins a+h/2, b+h/2):
 * :823-916.  Returns the extended-grid (row-major) index of the first support
 * point.  zf may be NULL.  q multiplies the x-factor only, as in the reference.
 */
static int or_expand1(int per, const double xp[3], double q, const or_params* p,
                      double* zx, double* zy, double* zz,
                      double* zfx, double* zfy, double* zfz)
{
    const int P = p->P, ph = p->P_half;
    const double h = p->h, c = p->c;
    double t[3];
    int from[3];
    for (int d = 0; d < 3; d++) {
        int freed = (per == OR_2P && d == 2) || (per == OR_1P && d < 2);
        double origin = 0.0;
        if (freed) origin = (per == OR_2P) ? p->a + h / 2 : (d == 0 ? p->a : p->b) + h / 2;
        int i;
        if (freed) {
            i = (P & 1) ? (int)round((xp[d] - origin) / h) : (int)floor((xp[d] - origin) / h);
            t[d] = xp[d] - (i * h + origin);
        } else {
            i = (P & 1) ? (int)round(xp[d] / h) : (int)floor(xp[d] / h);
            t[d] = xp[d] - h * i;
        }
        from[d] = (P & 1) ? i - ph : i - (ph - 1);
    }
    const int first = (P & 1) ? -ph : -ph + 1;   /* offset of support point 0 */
    double z3 = exp(-c * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2])) * q;
    double* zv[3] = {zx, zy, zz};
    double* fv[3] = {zfx, zfy, zfz};
    for (int d = 0; d < 3; d++) {
        double base = exp(2 * c * h * t[d]);
        double z = pow(base, first);
        zv[d][0] = z;
        for (int i = 1; i < P; i++) { z *= base; zv[d][i] = z; }
        if (fv[d])
            for (int i = 0; i < P; i++) fv[d][i] = -c * (t[d] - (first + i) * h);
    }
    for (int i = 0; i < P; i++) zx[i] *= z3;
    int e0 = from[0] + ((per == OR_1P) ? 0 : ph);
    int e1 = from[1] + ((per == OR_1P) ? 0 : ph);
    int e2 = from[2] + ((per == OR_2P) ? 0 : ph);
    return (e0 * p->npdims[1] + e1) * p->npdims[2] + e2;
}

/* SE_FGG_expand_all (SE_fgg.c:920-938) and _SSE_force (:941-963): q = 1.
 * The reference's 1P force build uses the 3P force expansion (SE_fgg.h:29-31). */
void or_expand_all(int per, int force, const double* x, const or_params* p,
                   double* zx, double* zy, double* zz, int* idx,
                   double* zfx, double* zfy, double* zfz)
{
    const int N = p->N, P = p->P;
    int eper = (force && per == OR_1P) ? OR_3P : per;
    for (int n = 0; n < N; n++) {
        double xp[3] = {x[n], x[n + N], x[n + 2 * N]};
        idx[n] = or_expand1(eper, xp, 1.0, p, zx + n * P, zy + n * P, zz + n * P,
                            force ? zfx + n * P : NULL, force ? zfy + n * P : NULL,
                            force ? zfz + n * P : NULL);
    }
}

/* SE_FGG_grid_split (SE_fgg.c:3983-4028): H_ext += q zs zz zx zy, serial. */
void or_grid_split(const or_params* p, const double* q, const double* zx, const double* zy,
                   const double* zz, const int* idx, const double* zs, double* H)
{
    const int P = p->P, n1 = p->npdims[1], n2 = p->npdims[2];
    for (int n = 0; n < p->N; n++) {
        int zi = 0;
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) {
                double cij = q[n] * zx[P * n + i] * zy[P * n + j];
                double* row = H + idx[n] + (i * n1 + j) * n2;
                for (int k = 0; k < P; k++, zi++) row[k] += zs[zi] * zz[P * n + k] * cij;
            }
    }
}

/* SE_FGG_grid (SE_fgg.c:3866-3914): expansion on the fly, q in the factors. */
void or_grid(int per, const or_params* p, const double* x, const double* q, const double* zs, double* H)
{
    const int N = p->N, P = p->P, n1 = p->npdims[1], n2 = p->npdims[2];
    double zx[64], zy[64], zz[64];
    for (int n = 0; n < N; n++) {
        double xp[3] = {x[n], x[n + N], x[n + 2 * N]};
        int id = or_expand1(per, xp, q[n], p, zx, zy, zz, NULL, NULL, NULL);
        int zi = 0;
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) {
                double cij = zx[i] * zy[j];
                double* row = H + id + (i * n1 + j) * n2;
                for (int k = 0; k < P; k++, zi++) row[k] += zs[zi] * zz[k] * cij;
            }
    }
}

/* Wrap (SE_fgg.c:327-411): fold the row-major extended grid into the
 * column-major periodic grid (H_per += ...).  Free directions are copied. */
void or_wrap(int per, const or_params* p, const double* H, double* Hper)
{
    const int ph = or_half(p->P);
    const int *D = p->dims, *E = p->npdims;
    for (int i = 0; i < E[0]; i++)
        for (int j = 0; j < E[1]; j++)
            for (int k = 0; k < E[2]; k++) {
                int a = (per == OR_1P) ? i : or_mod(i - ph, D[0]);
                int b = (per == OR_1P) ? j : or_mod(j - ph, D[1]);
                int cc = (per == OR_2P) ? k : or_mod(k - ph, D[2]);
                Hper[a + b * D[0] + cc * D[0] * D[1]] += H[(i * E[1] + j) * E[2] + k];
            }
}

/* Extend (SE_fgg.c:417-502): inverse of wrap, overwrites H. */
void or_extend(int per, const or_params* p, const double* Hper, double* H)
{
    const int ph = or_half(p->P);
    const int *D = p->dims, *E = p->npdims;
    for (int i = 0; i < E[0]; i++)
        for (int j = 0; j < E[1]; j++)
            for (int k = 0; k < E[2]; k++) {
                int a = (per == OR_1P) ? i : or_mod(i - ph, D[0]);
                int b = (per == OR_1P) ? j : or_mod(j - ph, D[1]);
                int cc = (per == OR_2P) ? k : or_mod(k - ph, D[2]);
                H[(i * E[1] + j) * E[2] + k] = Hper[a + b * D[0] + cc * D[0] * D[1]];
            }
}

/* SE_FGG_int_split (SE_fgg.c:1176-1223): phi_m = h^3 sum H zs zz zx zy. */
void or_int_split(const or_params* p, const double* H, const double* zx, const double* zy,
                  const double* zz, const int* idx, const double* zs, double* phi)
{
    const int P = p->P, n1 = p->npdims[1], n2 = p->npdims[2];
    const double h = p->h;
    for (int m = 0; m < p->N; m++) {
        double s = 0;
        int zi = 0;
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) {
                double cij = zx[m * P + i] * zy[m * P + j];
                const double* row = H + idx[m] + (i * n1 + j) * n2;
                for (int k = 0; k < P; k++, zi++) s += row[k] * zs[zi] * zz[m * P + k] * cij;
            }
        phi[m] = (h * h * h) * s;
    }
}

/* SE_FGG_int_split_force (SE_fgg.c:2241-2312): F = h^3 sum H W q zf  (= q/2 grad phi). */
void or_int_split_force(const or_params* p, const double* H, const double* q,
                        const double* zx, const double* zy, const double* zz,
                        const double* zfx, const double* zfy, const double* zfz,
                        const int* idx, const double* zs, double* F)
{
    const int P = p->P, N = p->N, n1 = p->npdims[1], n2 = p->npdims[2];
    const double h = p->h;
    for (int m = 0; m < N; m++) {
        double f0 = 0, f1 = 0, f2 = 0;
        int zi = 0;
        for (int i = 0; i < P; i++)
            for (int j = 0; j < P; j++) {
                double cij = zx[m * P + i] * zy[m * P + j];
                const double* row = H + idx[m] + (i * n1 + j) * n2;
                for (int k = 0; k < P; k++, zi++) {
                    double w = row[k] * zs[zi] * zz[m * P + k] * cij * q[m];
                    f0 += w * zfx[m * P + i];
                    f1 += w * zfy[m * P + j];
                    f2 += w * zfz[m * P + k];
                }
            }
        F[m] = (h * h * h) * f0; F[m + N] = (h * h * h) * f1; F[m + 2 * N] = (h * h * h) * f2;
    }
}

/* SE_general.c:34-84: wave numbers 2*pi/L * {0..M/2-1, -M/2..-1} (odd M analogous). */
void or_kvec(int M, double L, double* k)
{
    double f = 2.0 * M_PI / L;
    int pos = (M % 2 == 0) ? M / 2 : (M - 1) / 2 + 1;
    for (int i = 0; i < M; i++) k[i] = f * (double)(i < pos ? i : i - M);
}

/* SE_general.c:141-176: Z = exp(scalar K^2)/K^2 * 4 pi / M^3 on a cubic M^3 grid,
 * row-major.  Z at K = 0 is left as computed (inf); callers zero it. */
void or_scaling(double scalar, double* Z, int M, double L)
{
    double* k = (double*)malloc(sizeof(double) * M);
    or_kvec(M, L, k);
    double dd = 4.0 * M_PI / ((double)M * M * M);
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++)
            for (int l = 0; l < M; l++) {
                double K2 = k[i] * k[i] + k[j] * k[j] + k[l] * k[l];
                Z[((size_t)i * M + j) * M + l] = exp(scalar * K2) / K2 * dd;
            }
    free(k);
}

/* ---- FFT: recursive mixed-radix DIT, unnormalised, sign -1 forward ---- */
static void or_fft_rec(const double complex* in, double complex* out, int n, int stride, int sign,
                       double complex* tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    int r = 2;
    while (n % r) r++;
    int m = n / r;
    for (int s = 0; s < r; s++) or_fft_rec(in + s * stride, out + s * m, m, stride * r, sign, tmp);
    for (int k = 0; k < n; k++) {
        double complex acc = 0;
        int kk = k % m;
        for (int s = 0; s < r; s++) {
            long e = ((long)s * k) % n;
            double ang = sign * 2.0 * M_PI * (double)e / (double)n;
            acc += out[s * m + kk] * (cos(ang) + I * sin(ang));
        }
        tmp[k] = acc;
    }
    memcpy(out, tmp, sizeof(double complex) * n);
}

static void or_fft1d(double complex* line, int n, int sign)
{
    double complex* in = (double complex*)malloc(sizeof(double complex) * n * 3);
    memcpy(in, line, sizeof(double complex) * n);
    /* recursion needs a scratch per level: allocate per call at the top level and
       reuse it (levels finish before parents combine). */
    or_fft_rec(in, line, n, 1, sign, in + n);
    free(in);
}

/* In-place 3-D DFT of a row-major n0 x n1 x n2 complex array (interleaved re,im). */
void or_fft3d(double* data, int n0, int n1, int n2, int sign)
{
    double complex* a = (double complex*)data;
    int nmax = n0 > n1 ? n0 : n1; nmax = nmax > n2 ? nmax : n2;
    double complex* line = (double complex*)malloc(sizeof(double complex) * nmax);
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++) {
            double complex* row = a + ((size_t)i * n1 + j) * n2;
            memcpy(line, row, sizeof(double complex) * n2);
            or_fft1d(line, n2, sign);
            memcpy(row, line, sizeof(double complex) * n2);
        }
    for (int i = 0; i < n0; i++)
        for (int k = 0; k < n2; k++) {
            for (int j = 0; j < n1; j++) line[j] = a[((size_t)i * n1 + j) * n2 + k];
            or_fft1d(line, n1, sign);
            for (int j = 0; j < n1; j++) a[((size_t)i * n1 + j) * n2 + k] = line[j];
        }
    for (int j = 0; j < n1; j++)
        for (int k = 0; k < n2; k++) {
            for (int i = 0; i < n0; i++) line[i] = a[((size_t)i * n1 + j) * n2 + k];
            or_fft1d(line, n0, sign);
            for (int i = 0; i < n0; i++) a[((size_t)i * n1 + j) * n2 + k] = line[i];
        }
    free(line);
}

/*
 * SE3P driver (spectral_ewald.c:20-182) with the FGG_SPLIT path: expand (q=1),
 * spread, wrap (column-major), c2c FFT of the real grid, times Z with Z(0)=0,
 * inverse c2c, real part, extend, interpolate.  force=0 -> out = phi[N];
 * force=1 -> out = F[3N] in the reference's (q/2) grad(phi) convention.
 * Also returns the periodic grid after the inverse FFT if grid_out != NULL.
 */
void or_se3p(int N, const double* x, const double* q, double L, int M, int P, double m,
             double xi, int force, double* out, double* grid_out)
{
    double pp[4];
    or_parse_params(L, M, P, m, xi, pp);
    const double eta = pp[2], c = pp[3];
    or_params p;
    or_pack_params(&p, N, M, M, M, P, c, L / M, -1.0e308, -1.0);
    size_t next = (size_t)p.npdims[0] * p.npdims[1] * p.npdims[2];
    size_t M3 = (size_t)M * M * M;
    double* H = (double*)calloc(next, sizeof(double));
    double* zs = (double*)malloc(sizeof(double) * P * P * P);
    double* zx = (double*)malloc(sizeof(double) * N * P);
    double* zy = (double*)malloc(sizeof(double) * N * P);
    double* zz = (double*)malloc(sizeof(double) * N * P);
    double* zf = force ? (double*)malloc(sizeof(double) * 3 * N * P) : NULL;
    int* idx = (int*)malloc(sizeof(int) * N);
    or_base_gaussian(&p, zs);
    or_expand_all(OR_3P, force, x, &p, zx, zy, zz, idx,
                  zf, zf ? zf + N * P : NULL, zf ? zf + 2 * N * P : NULL);
    or_grid_split(&p, q, zx, zy, zz, idx, zs, H);
    double* Hper = (double*)calloc(M3, sizeof(double));
    or_wrap(OR_3P, &p, H, Hper);
    double* G = (double*)malloc(sizeof(double) * 2 * M3);
    for (size_t i = 0; i < M3; i++) { G[2 * i] = Hper[i]; G[2 * i + 1] = 0.0; }
    or_fft3d(G, M, M, M, -1);
    double* Z = (double*)malloc(sizeof(double) * M3);
    or_scaling(-(1.0 - eta) / (4.0 * xi * xi), Z, M, L);
    Z[0] = 0.0;
    for (size_t i = 0; i < M3; i++) { G[2 * i] *= Z[i]; G[2 * i + 1] *= Z[i]; }
    or_fft3d(G, M, M, M, +1);
    for (size_t i = 0; i < M3; i++) Hper[i] = G[2 * i];
    if (grid_out) memcpy(grid_out, Hper, sizeof(double) * M3);
    or_extend(OR_3P, &p, Hper, H);
    if (force)
        or_int_split_force(&p, H, q, zx, zy, zz, zf, zf + N * P, zf + 2 * N * P, idx, zs, out);
    else
        or_int_split(&p, H, zx, zy, zz, idx, zs, out);
    free(H); free(zs); free(zx); free(zy); free(zz); free(zf); free(idx);
    free(Hper); free(G); free(Z);
}

/* ---------------- direct k-space Ewald sums (oracle of the oracle) ---------------- */

/* SE_direct.c:352-414 as written: pairwise, symmetric-sign weights hs().  O(N^2 K^3). */
void or_direct3p_fd_pairwise(double* phi, const double* x, const double* q, int N,
                             const double box[3], double xi, int layers)
{
    double c = 4.0 * M_PI / (box[0] * box[1] * box[2]);
    for (int m = 0; m < N; m++) {
        double acc = 0;
        for (int j0 = 0; j0 <= layers; j0++)
            for (int j1 = 0; j1 <= layers; j1++)
                for (int j2 = 0; j2 <= layers; j2++) {
                    if (!j0 && !j1 && !j2) continue;
                    double k0 = 2 * M_PI * j0 / box[0], k1 = 2 * M_PI * j1 / box[1], k2 = 2 * M_PI * j2 / box[2];
                    double K2 = k0 * k0 + k1 * k1 + k2 * k2;
                    double s = 0;
                    for (int n = 0; n < N; n++)
                        s += q[n] * cos(k0 * (x[m] - x[n])) * cos(k1 * (x[m + N] - x[n + N])) *
                             cos(k2 * (x[m + 2 * N] - x[n + 2 * N]));
                    int zeros = (j0 == 0) + (j1 == 0) + (j2 == 0);
                    double wgt = zeros == 0 ? 8.0 : (zeros == 1 ? 4.0 : 2.0);
                    acc += wgt * s * exp(-K2 / (4 * xi * xi)) / K2;
                }
        phi[m] = c * acc;
    }
}

/* Structure-factor form of the same sum over the full cube |j_d| <= layers:
 *   phi_m = 4pi/V sum_{k != 0} e^{-k^2/4xi^2}/k^2 Re( e^{i k.x_m} S(k) ),  S(k) = sum_n q_n e^{-i k.x_n}.
 * force=1 gives SE_direct.c:417-461's F_m = (q_m/2) 4pi/V sum_k e/k^2 k Im(e^{i k.x_m} conj-free...)
 * written as 0.5 c q_m sum_k (e/k^2) k sum_n q_n sin(k.(x_n - x_m)).
 * Targets: the first Nt particles of x (Nt <= N) receive outputs. */
void or_direct3p_sf(double* out, int force, const double* x, const double* q, int N, int Nt,
                    const double box[3], double xi, int layers)
{
    const int K = layers, W = 2 * K + 1;
    double c = 4.0 * M_PI / (box[0] * box[1] * box[2]);
    /* per-dimension phase tables e^{-i 2pi j x/L}, j in [-K, K] */
    double complex* E = (double complex*)malloc(sizeof(double complex) * (size_t)3 * N * W);
    for (int d = 0; d < 3; d++)
        for (int n = 0; n < N; n++)
            for (int j = -K; j <= K; j++) {
                double ang = -2.0 * M_PI * j * x[n + d * N] / box[d];
                E[((size_t)d * N + n) * W + (j + K)] = cos(ang) + I * sin(ang);
            }
    size_t nk = (size_t)W * W * W;
    double complex* S = (double complex*)malloc(sizeof(double complex) * nk);
#pragma omp parallel for schedule(dynamic)
    for (long kk = 0; kk < (long)nk; kk++) {
        int j0 = (int)(kk / (W * W)), j1 = (int)((kk / W) % W), j2 = (int)(kk % W);
        double complex s = 0;
        for (int n = 0; n < N; n++)
            s += q[n] * E[(size_t)n * W + j0] * E[((size_t)N + n) * W + j1] * E[((size_t)2 * N + n) * W + j2];
        S[kk] = s;
    }
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < Nt; m++) {
        double a0 = 0, a1 = 0, a2 = 0, ap = 0;
        for (size_t kk = 0; kk < nk; kk++) {
            int j0 = (int)(kk / (W * W)), j1 = (int)((kk / W) % W), j2 = (int)(kk % W);
            if (j0 == K && j1 == K && j2 == K) continue;
            double k0 = 2 * M_PI * (j0 - K) / box[0], k1 = 2 * M_PI * (j1 - K) / box[1],
                   k2 = 2 * M_PI * (j2 - K) / box[2];
            double K2 = k0 * k0 + k1 * k1 + k2 * k2;
            double g = exp(-K2 / (4 * xi * xi)) / K2;
            /* e^{+i k.x_m} = conj of the stored e^{-i k.x_m} */
            double complex em = conj(E[(size_t)m * W + j0] * E[((size_t)N + m) * W + j1] *
                                     E[((size_t)2 * N + m) * W + j2]);
            double complex v = em * S[kk];
            if (force) {
                /* sum_n q_n sin(k.(x_n - x_m)) = -Im(e^{i k.x_m} S(k)) */
                double sn = -cimag(v);
                a0 += g * sn * k0; a1 += g * sn * k1; a2 += g * sn * k2;
            } else {
                ap += g * creal(v);
            }
        }
        if (force) {
            out[m] = 0.5 * c * a0 * q[m]; out[m + Nt] = 0.5 * c * a1 * q[m]; out[m + 2 * Nt] = 0.5 * c * a2 * q[m];
        } else {
            out[m] = c * ap;
        }
    }
    free(E); free(S);
}

/* SE_direct.c:417-461 as written (pairwise sine form), for tiny-N cross-checks. */
void or_direct3p_force_pairwise(double* F, const double* x, const double* q, int N,
                                const double box[3], double xi, int layers)
{
    double c = 4.0 * M_PI / (box[0] * box[1] * box[2]);
    for (int m = 0; m < N; m++) {
        double p0 = 0, p1 = 0, p2 = 0;
        for (int j0 = -layers; j0 <= layers; j0++)
            for (int j1 = -layers; j1 <= layers; j1++)
                for (int j2 = -layers; j2 <= layers; j2++) {
                    if (!j0 && !j1 && !j2) continue;
                    double k0 = 2 * M_PI * j0 / box[0], k1 = 2 * M_PI * j1 / box[1], k2 = 2 * M_PI * j2 / box[2];
                    double K2 = k0 * k0 + k1 * k1 + k2 * k2;
                    double z = 0;
                    for (int n = 0; n < N; n++)
                        z += q[n] * sin(k0 * (x[n] - x[m]) + k1 * (x[n + N] - x[m + N]) + k2 * (x[n + 2 * N] - x[m + 2 * N]));
                    double g = z * exp(-K2 / (4 * xi * xi)) / K2;
                    p0 += g * k0; p1 += g * k1; p2 += g * k2;
                }
        F[m] = .5 * c * p0 * q[m]; F[m + N] = .5 * c * p1 * q[m]; F[m + 2 * N] = .5 * c * p2 * q[m];
    }
}

/* SE_direct.c:256-318: SE2P k-space sum (k != 0), periodic in x,y. */
static double or_theta(double z, double k, double xi, double sgn)
{
    if (sgn * k * z < 600.0) return exp(sgn * k * z) * erfc(k / (2.0 * xi) + sgn * xi * z);
    return 0.0;
}

void or_direct2p_fd(double* phi, const double* x, const double* q, int N, int Nt,
                    const double box[3], double xi, int layers)
{
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < Nt; m++) {
        double acc = 0;
        for (int n = 0; n < N; n++)
            for (int j0 = -layers; j0 <= layers; j0++)
                for (int j1 = -layers; j1 <= layers; j1++) {
                    if (!j0 && !j1) continue;
                    double k0 = 2 * M_PI * j0 / box[0], k1 = 2 * M_PI * j1 / box[1];
                    double kn = sqrt(k0 * k0 + k1 * k1);
                    double kr = k0 * (x[m] - x[n]) + k1 * (x[m + N] - x[n + N]);
                    double z = x[m + 2 * N] - x[n + 2 * N];
                    acc += q[n] * cos(kr) * (or_theta(z, kn, xi, -1.0) + or_theta(z, kn, xi, +1.0)) / kn;
                }
        phi[m] = M_PI * acc / (box[0] * box[1]);
    }
}

/* SE_direct.c:320-338: SE2P k = 0 term. */
void or_direct2p_k0(double* phi, const double* x, const double* q, int N, int Nt,
                    const double box[3], double xi)
{
    for (int m = 0; m < Nt; m++) {
        double acc = 0;
        for (int n = 0; n < N; n++) {
            double z = x[m + 2 * N] - x[n + 2 * N];
            acc += q[n] * (exp(-xi * xi * z * z) / xi + sqrt(M_PI) * z * erf(xi * z));
        }
        phi[m] = -2 * acc * sqrt(M_PI) / (box[0] * box[1]);
    }
}

/* ---------------- SE1P direct sums (periodic in z, free in x and y) ---------------- */

/* Gauss-Legendre rule on [a, b]: Newton iteration on the three-term Legendre
 * recurrence from the Chebyshev guesses cos(pi (i - 1/4)/(n + 1/2)) (textbook). */
static void or_gauss_legendre(int n, double a, double b, double* x, double* w)
{
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1;
        for (int it = 0; it < 100; it++) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) {
                /* one more update of the derivative at the converged node */
                p0 = 1; p1 = z;
                for (int k = 2; k <= n; k++) {
                    double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1);
                break;
            }
        }
        double wi = 2.0 / ((1 - z * z) * dp * dp);
        x[i] = 0.5 * (a + b) - 0.5 * (b - a) * z;
        x[n - 1 - i] = 0.5 * (a + b) + 0.5 * (b - a) * z;
        w[i] = w[n - 1 - i] = 0.5 * (b - a) * wi;
    }
}

/* Integrand of the incomplete Bessel function after t = e^{-y}:
 * int_0^1 t^{-1} e^{-b t - a/t} dt = int_0^inf exp(-b e^{-y} - a e^{y}) dy. */
static double or_k0inc_f(double a, double b, double y) { return exp(-b * exp(-y) - a * exp(y)); }

static double or_gx16[16], or_gw16[16];

/* the 16-point rule is built once, at load time (the direct sums call it from threads) */
__attribute__((constructor)) static void or_gl16_init(void) { or_gauss_legendre(16, -1.0, 1.0, or_gx16, or_gw16); }

static double or_gl16(double a, double b, double l, double r)
{
    const double* gx = or_gx16;
    const double* gw = or_gw16;
    double s = 0, c = 0.5 * (l + r), hw = 0.5 * (r - l);
    for (int i = 0; i < 16; i++) s += gw[i] * or_k0inc_f(a, b, c + hw * gx[i]);
    return s * hw;
}

static double or_k0inc_adapt(double a, double b, double l, double r, double whole, double tol, int depth)
{
    double m = 0.5 * (l + r);
    double lft = or_gl16(a, b, l, m), rgt = or_gl16(a, b, m, r);
    if (depth > 40 || fabs(lft + rgt - whole) <= tol) return lft + rgt;
    return or_k0inc_adapt(a, b, l, m, lft, tol, depth + 1) + or_k0inc_adapt(a, b, m, r, rgt, tol, depth + 1);
}

/* Incomplete modified Bessel function K0(a, b) = int_0^1 t^{-1} exp(-b t - a/t) dt
 * (the integrand of mathint.c:387-395, IncompBesselK0_int, integrated here
 * adaptively to ~1e-16 instead of by the reference's fixed two-panel rule).  a > 0. */
double or_k0inc(double a, double b)
{
    if (a <= 0) return HUGE_VAL;
    /* beyond y = Y the integrand is below e^{-745} */
    double Y = log(745.0 / a);
    if (Y <= 0) return 0.0;
    /* coarse pass to set the absolute tolerance */
    double coarse = 0;
    int np = 64;
    for (int i = 0; i < np; i++) coarse += or_gl16(a, b, Y * i / np, Y * (i + 1) / np);
    double tol = 1e-17 * fabs(coarse) + 1e-300, acc = 0;
    for (int i = 0; i < np; i++) {
        double l = Y * i / np, r = Y * (i + 1) / np;
        acc += or_k0inc_adapt(a, b, l, r, or_gl16(a, b, l, r), tol, 0);
    }
    return acc;
}

/* Ein(x) = gamma + ln x + E1(x) = int_0^x (1 - e^{-t})/t dt: the power series
 * sum (-1)^{k+1} x^k/(k k!) for x < 2, otherwise gamma + ln x + E1(x) with E1 from
 * its continued fraction (modified Lentz). */
double or_ein(double x)
{
    const double euler = 0.57721566490153286061;
    if (x <= 0) return 0.0;
    if (x < 2.0) {
        double term = x, s = x;
        for (int k = 2; k < 200; k++) {
            term *= -x / k;
            double add = term / k;
            s += add;
            if (fabs(add) < 1e-18 * fabs(s)) break;
        }
        return s;
    }
    /* E1(x) = e^{-x} / (x + 1 - 1/(x + 3 - 4/(x + 5 - ...))) */
    double b = x + 1.0, c = 1e300, d = 1.0 / b, hh = d;
    for (int i = 1; i < 1000; i++) {
        double an = -(double)i * i;
        b += 2.0;
        d = 1.0 / (an * d + b);
        c = b + an / c;
        double del = c * d;
        hh *= del;
        if (fabs(del - 1.0) < 1e-17) break;
    }
    return euler + log(x) + hh * exp(-x);
}

/* SE_direct.c:99-155 (SE1P_direct_fd): k_z != 0 part of the SE1P k-space sum,
 *   phi_m = 1/L sum_n q_n sum_{j != 0} cos(k_j z_mn) K0inc(k_j^2/(4 xi^2), rho_mn^2 xi^2),
 * k_j = 2 pi j / L (L = box[2]); j and -j are summed together. */
void or_direct1p_fd(double* phi, const double* x, const double* q, int N, int Nt,
                    const double box[3], double xi, int layers)
{
    const double L = box[2];
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < Nt; m++) {
        double acc = 0;
        for (int n = 0; n < N; n++) {
            double dx = x[m] - x[n], dy = x[m + N] - x[n + N], z = x[m + 2 * N] - x[n + 2 * N];
            double b = (dx * dx + dy * dy) * xi * xi;
            for (int j = 1; j <= layers; j++) {
                double k = 2 * M_PI * j / L;
                acc += 2.0 * q[n] * cos(k * z) * or_k0inc(k * k / (4 * xi * xi), b);
            }
        }
        phi[m] = acc / L;
    }
}

/* SE_direct.c:158-176 (SE1P_direct_k0): the k_z = 0 part,
 *   phi_m = -1/L sum_{n != m} q_n (gamma + ln(rho^2 xi^2) + E1(rho^2 xi^2)). */
void or_direct1p_k0(double* phi, const double* x, const double* q, int N, int Nt,
                    const double box[3], double xi)
{
    for (int m = 0; m < Nt; m++) {
        double acc = 0;
        for (int n = 0; n < N; n++) {
            if (n == m) continue;
            double dx = x[m] - x[n], dy = x[m + N] - x[n + N];
            acc -= q[n] * or_ein((dx * dx + dy * dy) * xi * xi);
        }
        phi[m] = acc / box[2];
    }
}

/* ---------------- real-space and self terms (SE_direct.c) ---------------- */

/* Real-space Ewald sum sum_n sum_images q_n erfc(xi r)/r over the periodic images of
 * the `per` periodic directions (3: all, SE_direct.c:464-505; 2: x and y,
 * :184-254; 1: z, :15-88), |image index| <= layers, the m == n unshifted pair left
 * out; pairs with r > rc are skipped (rc <= 0: no cutoff).  Targets: the first Nt
 * particles.  Output is assigned (the reference's SE3P form adds to phi, :502). */
void or_direct_real(int per, double* phi, const double* x, const double* q, int N, int Nt,
                    const double box[3], double xi, int layers, double rc)
{
    int lx = per >= 2 ? layers : 0, ly = per >= 2 ? layers : 0, lz = (per == 3 || per == 1) ? layers : 0;
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < Nt; m++) {
        double acc = 0;
        for (int n = 0; n < N; n++) {
            double r0 = x[m] - x[n], r1 = x[m + N] - x[n + N], r2 = x[m + 2 * N] - x[n + 2 * N];
            for (int p0 = -lx; p0 <= lx; p0++)
                for (int p1 = -ly; p1 <= ly; p1++)
                    for (int p2 = -lz; p2 <= lz; p2++) {
                        if (m == n && !p0 && !p1 && !p2) continue;
                        double a = r0 + p0 * box[0], b = r1 + p1 * box[1], c = r2 + p2 * box[2];
                        double r = sqrt(a * a + b * b + c * c);
                        if (rc > 0 && r > rc) continue;
                        acc += q[n] * erfc(xi * r) / r;
                    }
        }
        phi[m] = acc;
    }
}

/* Self term -2 xi/sqrt(pi) q_m (SE_direct.c:178-183, 340-345, 550-561). */
void or_direct_self(double* phi, const double* q, int N, double xi)
{
    double c = 2 * xi / sqrt(M_PI);
    for (int m = 0; m < N; m++) phi[m] = -c * q[m];
}

/* SE_fgg.c:5934-5946 */
double or_energy(const double* phi, const double* q, int N)
{
    double e = 0;
    for (int i = 0; i < N; i++) e += phi[i] * q[i];
    return e / 2.0;
}

/* SE_general.c:517-572 (potential branch; AVX build rounds M up to a multiple of 4).
 * Returns Q = sum q^2; writes M, m. */
double or_init_params(int* M, int P, double* m, double L, double xi, const double* q, int N,
                      int round4)
{
    double Q = 0;
    for (int i = 0; i < N; i++) Q += q[i] * q[i];
    double c = 0.95;
    *m = c * sqrt(M_PI * P);
    int Ml = (int)ceil(2.0 * (2.0 * xi * L) + 1) + 2;
    if (Ml < P) Ml = P;
    int r = round4 ? 4 : 2;
    if (Ml % r) Ml += r - Ml % r;
    *M = Ml;
    return Q;
}

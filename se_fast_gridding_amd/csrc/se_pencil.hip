// Register-ring "pencil" kernels: the FP32 fast path of the fused engine.
//
// Particles are sorted (tile, pencil, row): a pencil is a 2x2 group of window-start
// columns (x, z) inside a tile, walked along y.  All particles of a pencil touch the
// same (P+1) x (P+1) block of (x, z) grid columns, so a wave keeps that block in
// registers: lane l owns the (i, k) pairs l, l+64, ... (R = ceil((P+1)^2/64) per
// lane) for P consecutive grid rows -- a ring of P rows.  Slot s of the ring holds
// the row rho == s (mod P); instead of rotating registers when the window moves,
// the y-factor is rotated: slot s of a particle at row y uses Y[(s - y) mod P].
// Every register index is therefore static and each grid update is one v_fmac_f32
// with the y-factor in an SGPR.
//
// Spread: a row that falls out of the window is added once (ds_add_f64) into the
// tile's FP64 LDS sub-grid, then the item's sub-grid is written to its partial slot
// and the fixed-order halo merge builds the periodic grid.
// Gather: the tile's sub-grid is loaded into LDS; rows entering the window are read
// once into the ring, and each particle needs 2*P*R FMAs (potential + force).
//
// LDS sub-grids here are laid out [y][x][z] so that a ring row is one LDS plane.
// Work split: an item (<= chunk particles of one tile) is cut into nw equal,
// contiguous particle ranges, one per wave; a wave flushes and restarts its ring
// whenever its range crosses into another pencil.
#include <hip/hip_runtime.h>

#include "se_internal.h"

namespace se {

namespace {

__device__ __forceinline__ float rlane(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <int P>
__device__ __forceinline__ int modp(int v)
{
    int r = v % P;
    return r < 0 ? r + P : r;
}

__device__ __forceinline__ float wave_sum_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Lane-owned (i, k) pairs of the (P+1)^2 pencil block: offset i*S + k in an LDS row
// plane, or -1 for padding slots.
template <int P, int R>
__device__ __forceinline__ void lane_pairs(int lane, int S, int (&pi)[R], int (&pk)[R], int (&pofs)[R])
{
    constexpr int W = P + 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int pr = lane + 64 * r;
        const bool ok = pr < W * W;
        pi[r] = ok ? pr / W : 0;
        pk[r] = ok ? pr - (pr / W) * W : 0;
        pofs[r] = ok ? pi[r] * S + pk[r] : -1;
    }
}

}  // namespace

template <int P>
__global__ __launch_bounds__(512) void k_spread_pencil(const Rec<float>* __restrict__ rec,
                                                       const int* __restrict__ item_beg,
                                                       const int* __restrict__ item_end,
                                                       const int* __restrict__ n_items, Geom g, Tiling tl,
                                                       float* __restrict__ partials)
{
    constexpr int W = P + 1;
    constexpr int R = (W * W + 63) / 64;
    const int item = blockIdx.x;
    if (item >= *n_items) return;   // uniform across the work-group
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* sub = reinterpret_cast<double*>(smem_raw);   // [y][x][z], S^3
    const int S = tl.S, S2 = S * S, S3 = S2 * S;
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int e = tid; e < S3; e += nthr) sub[e] = 0.0;
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    int pi[R], pk[R], pofs[R];
    lane_pairs<P, R>(lane, S, pi, pk, pofs);

    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    const int wb = beg + (int)(((long)n * wave) / nw), we = beg + (int)(((long)n * (wave + 1)) / nw);
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;

    float acc[P][R];
#pragma unroll
    for (int s = 0; s < P; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[s][r] = 0.f;

    int pen = -1, base = 0, pbase = 0;
    // add ring rows [base, upto) (upto <= base + P) into LDS and clear their slots
    auto flush = [&](int upto) {
#pragma unroll
        for (int s = 0; s < P; ++s) {
            const int rho = base + modp<P>(s - base);
            if (rho < upto) {
                double* rowp = sub + rho * S2 + pbase;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (pofs[r] >= 0)
                        __hip_atomic_fetch_add(rowp + pofs[r], (double)acc[s][r], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    acc[s][r] = 0.f;
                }
            }
        }
    };

    for (int p = wb; p < we; ++p) {
        const Rec<float> rc = rec[p];
        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255, lz = rc.lo >> 16;
        const int npen = ((lx >> 1) << 12) | (lz >> 1);
        if (npen != pen) {
            if (pen >= 0) flush(base + P);
            pen = npen;
            base = ly;
            pbase = ((lx >> 1) * 2) * S + (lz >> 1) * 2;
        } else if (ly > base) {
            flush(min(ly, base + P));
            base = ly;
        }
        const int dx = lx & 1, dz = lz & 1;
        float fx = 0.f, fz = 0.f, fy = 0.f;
        if (lane < W) {
            const int ix = lane - dx, iz = lane - dz;
            if (ix >= 0 && ix < P) {
                const float a = rc.t[0] - (float)(g.first + ix) * h;
                fx = expf(-c * a * a) * (rc.q * dn);
            }
            if (iz >= 0 && iz < P) {
                const float a = rc.t[2] - (float)(g.first + iz) * h;
                fz = expf(-c * a * a);
            }
        }
        if (lane < P) {
            const float a = rc.t[1] - (float)(g.first + modp<P>(lane - ly)) * h;
            fy = expf(-c * a * a);
        }
        float xz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float v = __shfl(fx, pi[r], 64) * __shfl(fz, pk[r], 64);
            xz[r] = pofs[r] >= 0 ? v : 0.f;
        }
#pragma unroll
        for (int s = 0; s < P; ++s) {
            const float ys = rlane(fy, s);
#pragma unroll
            for (int r = 0; r < R; ++r) acc[s][r] = fmaf(ys, xz[r], acc[s][r]);
        }
    }
    if (pen >= 0) flush(base + P);
    __syncthreads();
    float* out = partials + (size_t)item * tl.part_elems;
    for (int e = tid; e < tl.part_elems; e += nthr) {
        const int iz = e % S, r2 = e / S, iy = r2 % S, ix = r2 / S;   // canonical [x][y][z]
        out[e] = (float)sub[(iy * S + ix) * S + iz];
    }
}

template <int P, bool POT, bool FORCE>
__global__ __launch_bounds__(512) void k_interp_pencil(const Rec<float>* __restrict__ rec,
                                                       const int* __restrict__ item_tile,
                                                       const int* __restrict__ item_beg,
                                                       const int* __restrict__ item_end,
                                                       const int* __restrict__ n_items,
                                                       const float* __restrict__ grid, Geom g, Tiling tl, int N,
                                                       double* __restrict__ phi, double* __restrict__ force)
{
    constexpr int W = P + 1;
    constexpr int R = (W * W + 63) / 64;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* sub = reinterpret_cast<float*>(smem_raw);   // [y][x][z], S^3
    const int S = tl.S, S2 = S * S, S3 = S2 * S, M = g.M, NT = tl.NT;
    const int tid = threadIdx.x, nthr = blockDim.x;
    {
        const int tile = item_tile[item];
        const int ox = (tile / (NT * NT)) * tl.TS, oy = ((tile / NT) % NT) * tl.TS, oz = (tile % NT) * tl.TS;
        for (int e = tid; e < S3; e += nthr) {
            const int iz = e % S, r2 = e / S, ix = r2 % S, iy = r2 / S;
            const int gx = (ox + ix) % M, gy = (oy + iy) % M, gz = (oz + iz) % M;
            sub[e] = grid[((long)gx * M + gy) * M + gz];
        }
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    int pi[R], pk[R], pofs[R];
    lane_pairs<P, R>(lane, S, pi, pk, pofs);
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    const int wb = beg + (int)(((long)n * wave) / nw), we = beg + (int)(((long)n * (wave + 1)) / nw);
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;
    const double h3 = g.h * g.h * g.h;

    float G[P][R];
    int pen = -1, base = 0, pbase = 0;
    // read ring rows in [lo, hi) (all within [base, base + P)) from LDS
    auto load = [&](int lo, int hi) {
#pragma unroll
        for (int s = 0; s < P; ++s) {
            const int rho = base + modp<P>(s - base);
            if (rho >= lo && rho < hi) {
                const float* rowp = sub + rho * S2 + pbase;
#pragma unroll
                for (int r = 0; r < R; ++r) G[s][r] = pofs[r] >= 0 ? rowp[pofs[r]] : 0.f;
            }
        }
    };

    for (int p = wb; p < we; ++p) {
        const Rec<float> rc = rec[p];
        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255, lz = rc.lo >> 16;
        const int npen = ((lx >> 1) << 12) | (lz >> 1);
        if (npen != pen) {
            pen = npen;
            base = ly;
            pbase = ((lx >> 1) * 2) * S + (lz >> 1) * 2;
            load(base, base + P);
        } else if (ly > base) {
            const int old_end = base + P;
            base = ly;
            load(max(old_end, ly), ly + P);
        }
        const int dx = lx & 1, dz = lz & 1;
        float fx = 0.f, fz = 0.f, fy = 0.f, gx = 0.f, gz = 0.f, gy = 0.f;   // factors and -c*(t - o h)
        if (lane < W) {
            const int ix = lane - dx, iz = lane - dz;
            if (ix >= 0 && ix < P) {
                const float a = rc.t[0] - (float)(g.first + ix) * h;
                fx = expf(-c * a * a) * dn;
                gx = -c * a;
            }
            if (iz >= 0 && iz < P) {
                const float a = rc.t[2] - (float)(g.first + iz) * h;
                fz = expf(-c * a * a);
                gz = -c * a;
            }
        }
        if (lane < P) {
            const float a = rc.t[1] - (float)(g.first + modp<P>(lane - ly)) * h;
            fy = expf(-c * a * a);
            gy = -c * a;
        }
        const float fyd = fy * gy;
        float A[R], B[R];
#pragma unroll
        for (int r = 0; r < R; ++r) A[r] = B[r] = 0.f;
#pragma unroll
        for (int s = 0; s < P; ++s) {
            const float ys = rlane(fy, s);
            const float yd = FORCE ? rlane(fyd, s) : 0.f;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                A[r] = fmaf(ys, G[s][r], A[r]);
                if (FORCE) B[r] = fmaf(yd, G[s][r], B[r]);
            }
        }
        float sp = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float xv = __shfl(fx, pi[r], 64), zv = __shfl(fz, pk[r], 64);
            const float xz = pofs[r] >= 0 ? xv * zv : 0.f;
            if (POT || FORCE) sp = fmaf(xz, A[r], sp);
            if (FORCE) {
                const float xg = __shfl(gx, pi[r], 64), zg = __shfl(gz, pk[r], 64);
                s0 = fmaf(xz * xg, A[r], s0);
                s1 = fmaf(xz, B[r], s1);
                s2 = fmaf(xz * zg, A[r], s2);
            }
        }
        if (POT) sp = wave_sum_f(sp);
        if (FORCE) {
            s0 = wave_sum_f(s0);
            s1 = wave_sum_f(s1);
            s2 = wave_sum_f(s2);
        }
        if (lane == 0) {
            const int o = rc.orig;
            if (POT) phi[o] = h3 * (double)sp;
            if (FORCE) {
                const double qh = h3 * (double)rc.q;
                force[o] = qh * (double)s0;
                force[o + (long)N] = qh * (double)s1;
                force[o + 2l * N] = qh * (double)s2;
            }
        }
    }
}

bool pencil_supported(int P) { return P == 8 || P == 11 || P == 16; }

static void allow_lds_p(const void* fn, size_t bytes)
{
    if (bytes > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

bool launch_spread_pencil(const Rec<float>* rec, const int* item_beg, const int* item_end, const int* n_items,
                          const Geom& g, const Tiling& tl, float* partials, hipStream_t s)
{
    const size_t lds = (size_t)tl.S * tl.S * tl.S * sizeof(double);
    dim3 gr(tl.max_items), bl(tl.block);
#define SE_SP(PV)                                                                                            \
    do {                                                                                                     \
        allow_lds_p((const void*)k_spread_pencil<PV>, lds);                                                  \
        hipLaunchKernelGGL(k_spread_pencil<PV>, gr, bl, lds, s, rec, item_beg, item_end, n_items, g, tl,    \
                           partials);                                                                        \
    } while (0)
    switch (g.P) {
    case 8: SE_SP(8); return true;
    case 11: SE_SP(11); return true;
    case 16: SE_SP(16); return true;
    default: return false;
    }
#undef SE_SP
}

template <int PV>
static void interp_pencil_modes(const Rec<float>* rec, const int* item_tile, const int* item_beg,
                                const int* item_end, const int* n_items, const float* grid, const Geom& g,
                                const Tiling& tl, int N, double* phi, double* force, hipStream_t s)
{
    const size_t lds = (size_t)tl.S * tl.S * tl.S * sizeof(float);
    dim3 gr(tl.max_items), bl(tl.block);
#define SE_GP(POT, FORCE)                                                                                     \
    do {                                                                                                      \
        allow_lds_p((const void*)k_interp_pencil<PV, POT, FORCE>, lds);                                       \
        hipLaunchKernelGGL((k_interp_pencil<PV, POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg,       \
                           item_end, n_items, grid, g, tl, N, phi, force);                                    \
    } while (0)
    if (phi && force)
        SE_GP(true, true);
    else if (force)
        SE_GP(false, true);
    else
        SE_GP(true, false);
#undef SE_GP
}

bool launch_interp_pencil(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                          const int* n_items, const float* grid, const Geom& g, const Tiling& tl, int N,
                          double* phi, double* force, hipStream_t s)
{
    switch (g.P) {
    case 8: interp_pencil_modes<8>(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, s); return true;
    case 11: interp_pencil_modes<11>(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, s); return true;
    case 16: interp_pencil_modes<16>(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, s); return true;
    default: return false;
    }
}

}  // namespace se

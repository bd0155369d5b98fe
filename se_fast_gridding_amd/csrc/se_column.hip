// Column kernels: the FP32 P = 16 path for dense work items on the matrix cores.
//
// Particles are sorted (tile, pencil, row), a pencil being 4 x 1 window-start columns
// (4 in x, 1 in z) walked along y.  The particles of one 1 x 1 column of a pencil (one
// lx, one lz) share their x and z windows exactly, so for any grid row y the z
// contraction of up to sixteen of them is one 16 x 16 x 16 matrix product:
//
//   T_y[x][n] = sum_z G[y][x0 + x][z0 + z] Fz_n(z),     x, z = 0..15, n = particle
//
// which v_mfma_f32_16x16x4f32 runs as four K = 4 steps (A = the sub-grid slice, read
// from LDS; B = the particles' z factors).  Every product is a useful one: no window
// padding, no masking.  The y and x contractions are per particle and stay on the VALU:
//
//   U = sum_y Fy T,  V = sum_y Fy' T,  W = sum_y Fy T'   (T' with the z derivative factors)
//   phi = sum_x Fx U,  F = (sum_x Fx' U, sum_x Fx V, sum_x Fx W)
//
// A group costs (rows spanned + 15) row steps of 4 (potential) or 8 (potential and
// force) matrix instructions, whatever its 1..16 particles; in a dense column, sixteen
// consecutive particles span one or two rows, so nearly every instruction is full.
//
// Work split: the waves of a work-group (4, 8 or 16) take the item's records in 1, 2 or 4
// equal shares, four waves to a share; each of the four takes one column of every pencil
// of the share, scanning the records in sort order with ballots; a group closes at 16
// particles or at a pencil boundary.  The
// particles of a group come in row order, so rows y0 .. y1 + 15 cover every window.
// Each particle's result depends only on its own record and on the sub-grid: rows outside
// its window weigh exactly zero, so the grouping does not change any output bit.
#include <hip/hip_runtime.h>

#include "se_internal.h"

namespace se {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// exp(-c a^2) as v_exp_f32 of kc a^2, kc = -c log2(e)
__device__ __forceinline__ float gauss_c(float kc, float a) { return __builtin_amdgcn_exp2f(kc * a * a); }

// Sum over the four 16-lane rows (lanes j, j + 16, j + 32, j + 48): rows 0/1 and 2/3 by
// v_permlane16_swap, then the two halves by v_permlane32_swap; every lane gets the
// total, added in the same order in every lane.
__device__ __forceinline__ float sum_rows4(float v)
{
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

__device__ __forceinline__ v4f mfma16(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Lane masks of the first k set bits of m (k <= popcount(m)).
__device__ __forceinline__ unsigned long long first_bits(unsigned long long m, int k)
{
    unsigned long long rest = m;
#pragma nounroll
    for (int t = 0; t < k; ++t) rest &= rest - 1;
    return m & ~rest;
}

}  // namespace

// v_mfma_f32_16x16x4f32 operand layouts (one block): A[i][k] in lane i + 16 k, B[k][j] in
// lane j + 16 k, D[i][j] in register i mod 4 of lane j + 16 (i / 4).  Here lane j + 16 q
// (q = 0..3) holds: A = sub-grid point (x0 + j, z0 + 4 q + c) of K step c; B = particle
// j's z factor at window point 4 q + c; D = T_y[4 q + r][j] in register r.
template <bool POT, bool FORCE>
__global__ __launch_bounds__(1024) void k_interp_col16(const Rec<float>* __restrict__ rec,
                                                      const int* __restrict__ item_tile,
                                                      const int* __restrict__ item_beg,
                                                      const int* __restrict__ item_end,
                                                      const int* __restrict__ n_items,
                                                      const float* __restrict__ grid, Geom g, Tiling tl, int N,
                                                      double* __restrict__ phi, double* __restrict__ force,
                                                      const SysDesc* __restrict__ sys, int dense_min)
{
    constexpr int P = 16;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    // items below dense_min particles are the row gather's
    if (n < dense_min) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* sub = reinterpret_cast<float*>(smem_raw);   // [y][x][z], S^3
    const int S = tl.S, S2 = S * S, M = g.M, Mz = g.Mz, NT = tl.NT, NTz = tl.NTz;
    int* slots = reinterpret_cast<int*>(sub + S2 * S + LDS_TILE_PAD);   // 16 record indices per wave
    const int tid = threadIdx.x, nthr = blockDim.x;
    {
        int tile = item_tile[item];
        if (sys) {   // grouped solve: the item's system, its outputs and its grid
            const int si = tile / tl.ntiles;
            tile -= si * tl.ntiles;
            grid += (long)si * M * M * Mz;
            phi = sys[si].phi;
            force = sys[si].force;
            N = sys[si].N;
        }
        const int ox = (tile / (NT * NTz)) * tl.TS, oy = ((tile / NTz) % NT) * tl.TS, oz = (tile % NTz) * tl.TS;
        load_tile_dma<true>(sub, grid, S, M, Mz, ox, oy, oz, tid, nthr);
    }
    __syncthreads();

    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    const int col = wave & 3;                          // column of the share's first pencil
    const int part = wave >> 2, nparts = nw >> 2;      // share of the item's records
    const int rb = beg + (int)(((long)n * part) / nparts), re = beg + (int)(((long)n * (part + 1)) / nparts);
    int* myslots = slots + 16 * wave;
    const int j = lane & 15, q = lane >> 4;
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;
    const float kc = -c * 1.4426950408889634f;
    const float fh = (float)g.first * h;
    const double h3 = g.h * g.h * g.h;

    // One group of k particles, record indices in myslots[0 .. k).
    auto process = [&](int k) {
        __builtin_amdgcn_wave_barrier();
        const int si = myslots[min(j, k - 1)];   // slots past k repeat the last particle
        const float4 r = *reinterpret_cast<const float4*>(rec + si);   // t[0], t[1], t[2], q
        const int lo = rec[si].lo, orig = rec[si].orig;
        const bool act = j < k;
        const int ly = (lo >> 8) & 255;
        const int lo0 = __builtin_amdgcn_readfirstlane(lo);
        const int x0 = lo0 & 255, z0 = lo0 >> 16, y0 = (lo0 >> 8) & 255;
        const int y1 = ((__builtin_amdgcn_readlane(lo, k - 1) >> 8) & 255) + P;
        // B operands: particle j's z factors at window points 4 q + s, s = 0..3
        float bz[4], bzd[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float az = fmaf(-(float)(4 * q + s), h, r.z - fh);
            bz[s] = gauss_c(kc, az);
            if (FORCE) bzd[s] = bz[s] * (-c * az);
        }
        const float* ap = sub + (x0 + j) * S + z0 + 4 * q;
        const float tyf = r.y - fh;   // y offset of window index w: tyf - w h
        const v4f zero4 = {0.f, 0.f, 0.f, 0.f};
        v4f U = zero4, V = zero4, W = zero4;
#pragma unroll 2
        for (int y = y0; y < y1; ++y) {
            const float* a = ap + y * S2;
            const float a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
            v4f T = mfma16(a0, bz[0], zero4);
            T = mfma16(a1, bz[1], T);
            T = mfma16(a2, bz[2], T);
            T = mfma16(a3, bz[3], T);
            v4f Td = zero4;
            if (FORCE) {
                Td = mfma16(a0, bzd[0], zero4);
                Td = mfma16(a1, bzd[1], Td);
                Td = mfma16(a2, bzd[2], Td);
                Td = mfma16(a3, bzd[3], Td);
            }
            const int jy = y - ly;
            const float ay = fmaf(-(float)jy, h, tyf);
            const float fy = (act && (unsigned)jy < (unsigned)P) ? gauss_c(kc, ay) : 0.f;
            const float fyd = fy * (-c * ay);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                U[i] = fmaf(fy, T[i], U[i]);
                if (FORCE) {
                    V[i] = fmaf(fyd, T[i], V[i]);
                    W[i] = fmaf(fy, Td[i], W[i]);
                }
            }
        }
        // x contraction over this lane's rows x = 4 q + i, then over the four lane rows
        float sp = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float ax = fmaf(-(float)(4 * q + i), h, r.x - fh);
            const float fx = dn * gauss_c(kc, ax);
            sp = fmaf(fx, U[i], sp);
            if (FORCE) {
                sx = fmaf(fx * (-c * ax), U[i], sx);
                sy = fmaf(fx, V[i], sy);
                sz = fmaf(fx, W[i], sz);
            }
        }
        if (POT) sp = sum_rows4(sp);
        if (FORCE) {
            sx = sum_rows4(sx);
            sy = sum_rows4(sy);
            sz = sum_rows4(sz);
        }
        if (act) {   // one output per lane row
            if (POT && q == 0) store_out(phi, orig, h3 * (double)sp);
            if (FORCE && q > 0) {
                const double qh = h3 * (double)r.w;
                const float sv = q == 1 ? sx : (q == 2 ? sy : sz);
                store_out(force, orig + (long)(q - 1) * N, qh * (double)sv);
            }
        }
        __builtin_amdgcn_wave_barrier();
    };

    // Scan the share in sort order, 64 records at a time (the next batch's keys are loaded
    // while the current one is processed).  A group takes particles of one pencil, at most
    // 16, from one of its four columns: pencil p of the share (counted from 0 in sort order)
    // gives this wave the column lx mod 4 == (col + p) mod 4, so that a pencil's dense and
    // light columns fall to different waves from pencil to pencil.
    int have = 0, gpen = -1, pcarry = 0, ppen = -1;
    int lo_next = rb + lane < re ? rec[rb + lane].lo : -1;
    for (int cb = rb; cb < re; cb += 64) {
        const int i = cb + lane;
        const int lo = lo_next;
        lo_next = cb + 64 + lane < re ? rec[cb + 64 + lane].lo : -1;
        const bool in = i < re;
        const int pen = lo & 0xFF00FC;   // (lx / 4, lz): the pencil inside the tile
        // pencil ordinal of each record: pencils started before this batch plus the pencil
        // starts among lanes 0..lane of this batch
        const int prev = __shfl_up(pen, 1, 64);
        const bool start = in && pen != (lane == 0 ? ppen : prev);
        const unsigned long long smask = __ballot(start);
        const int pord = pcarry - 1 + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(smask >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((unsigned)smask, 0u))
                         + (start ? 1 : 0);
        pcarry += __builtin_popcountll(smask);
        ppen = __builtin_amdgcn_readlane(pen, 63);
        const bool match = in && (lo & 3) == ((col + pord) & 3);
        unsigned long long m = __ballot(match);
        while (m) {
            const int pj = __builtin_amdgcn_readlane(pen, __builtin_ctzll(m));
            if (have > 0 && pj != gpen) {
                process(have);
                have = 0;
            }
            gpen = pj;
            // this pencil's remaining matches: a run, since the records are sorted by pencil
            const unsigned long long same = __ballot(match && pen == gpen) & m;
            const int take = min(__builtin_popcountll(same), 16 - have);
            const unsigned long long taken = first_bits(same, take);
            if ((taken >> lane) & 1ull) {
                const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(taken >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)taken, 0u));
                myslots[have + rank] = i;
            }
            have += take;
            m &= ~taken;
            if (have == 16) {
                process(16);
                have = 0;
            }
        }
    }
    if (have > 0) process(have);
}

// Column gather of the items of at least dense_min particles; block = 256, 512 or 1024.
void launch_interp_col16(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                         const int* n_items, const float* grid, const Geom& g, const Tiling& tl, int N,
                         double* phi, double* force, const SysDesc* sys, int dense_min, int block, hipStream_t s)
{
    SE_REQUIRE(g.P == 16 && tl.S >= 16 + tl.TS - 1 && (block == 256 || block == 512 || block == 1024), SE_ERR_ARG,
               "column gather: P = 16 and 256, 512 or 1024 threads");
    const size_t lds = ((size_t)tl.S * tl.S * tl.S + LDS_TILE_PAD) * sizeof(float) + 16 * (block / 64) * sizeof(int);
    dim3 gr(tl.max_items), bl(block);
#define SE_GC(POT, FORCE)                                                                                    \
    do {                                                                                                     \
        if (lds > 64 * 1024)                                                                                 \
            (void)hipFuncSetAttribute((const void*)k_interp_col16<POT, FORCE>,                               \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                 \
        hipLaunchKernelGGL((k_interp_col16<POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg, item_end, \
                           n_items, grid, g, tl, N, phi, force, sys, dense_min);                             \
    } while (0)
    if (phi && force)
        SE_GC(true, true);
    else if (force)
        SE_GC(false, true);
    else
        SE_GC(true, false);
#undef SE_GC
}

}  // namespace se

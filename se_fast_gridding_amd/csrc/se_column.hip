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

// ---------------------------------------------------------------------------------
// Column spread of the dense items (P = 16).  The same 1 x 1 columns: for a grid row y,
// four particles k of one column add
//
//   D_y[x][z] += sum_k (q_k d Fx_k(x) Fy_k(y)) Fz_k(z),    x, z = 0..15
//
// which is one v_mfma_f32_16x16x4f32 (A = the particles' x factors times their y factor
// at row y, B = their z factors; K = the four particles).  A wave keeps a ring of 20 grid
// rows of its current column in registers (D of row y in ring slot y mod 20, four VGPRs
// each).  A group is up to sixteen particles of one column whose windows lie in the ring
// [y0, y0 + 20), y0 the first one's row: four K steps of four, each 20 matrix
// instructions with no branch.  Lane i of particle k's 16-lane row holds the y factor of
// ring slot i (and of slot 16 + i), so the A operand of slot s is one DPP multiply
// (row_newbcast: s mod 16) with every register index static.  Rows that leave the ring are
// added to the tile's FP64 LDS sub-grid: each value, scaled by 2^s per work item (q d 2^s
// is folded into the x factor) and rounded to an integer, is added exactly, so the sub-grid
// sums do not depend on the order of the atomics; the ring's own FP32 sums run in the fixed
// order of the groups.  The sub-grid then becomes the item's FP32 partial, as in the pencil
// spread.  Holding all S rows of the column instead, with a uniform branch per row and K
// step, measured slower (clustered N = 4M at M = 256, dense items: 1.03 against 0.93 ms):
// the branches cost more than the ring's extra rows.
//
// Work split: as in the column gather, four waves to a share of the item's records, each
// taking one column of every pencil of the share (rotated from pencil to pencil).  A
// group's particles are taken from the wave's current batch of 64 records (one record per
// lane; the next batch is loaded meanwhile): their batch lanes go to a per-wave slot list in
// LDS, and the lanes of each particle fetch its fields with ds_bpermute and form its factors.
// ---------------------------------------------------------------------------------
constexpr int COL_RING = 20;
constexpr int COL_KSTEPS = 4;   // K steps (of four particles) per group

// a * (lane L of b's 16-lane row): a DPP row_newbcast:L move (GFX9.4 DPP) into a multiply,
// which the compiler folds into one v_mul_f32_dpp
template <int L>
__device__ __forceinline__ float mul_row_bcast(float a, float b)
{
    return a * __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(b), 0x150 + L, 0xF, 0xF, false));
}

// one K step: ring slot s += (qfx * Y_s) (x) fz, Y_s = lane s mod 16 of fy1 (s < 16) or fy2
template <int S0>
struct ColumnRing {
    static __device__ __forceinline__ void run(v4f (&acc)[COL_RING], float qfx, float fy1, float fy2, float fz)
    {
        const float a = S0 < 16 ? mul_row_bcast<S0 & 15>(qfx, fy1) : mul_row_bcast<S0 & 15>(qfx, fy2);
        acc[S0] = mfma16(a, fz, acc[S0]);
        ColumnRing<S0 + 1>::run(acc, qfx, fy1, fy2, fz);
    }
};
template <>
struct ColumnRing<COL_RING> {
    static __device__ __forceinline__ void run(v4f (&)[COL_RING], float, float, float, float) {}
};

// Ring slots holding rows [lo, lo + n) (0 <= n <= R) as a bit mask: slot s holds the row
// == s (mod R), so the mask is n consecutive bits rotated by lo mod R.
template <int R>
__device__ __forceinline__ unsigned ring_slots(int lo, int n)
{
    const unsigned m = n >= R ? (1u << R) - 1u : (1u << n) - 1u;
    const int b = lo % R;
    return ((m << b) | (m >> (R - b))) & ((1u << R) - 1u);
}

__global__ __launch_bounds__(768) void k_spread_col16(const Rec<float>* __restrict__ rec,
                                                      const int* __restrict__ item_beg,
                                                      const int* __restrict__ item_end,
                                                      const int* __restrict__ n_items, Geom g, Tiling tl,
                                                      float* __restrict__ partials, int dense_min)
{
    constexpr int P = 16, GY = COL_RING, KS = COL_KSTEPS, GMAX = 4 * KS;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    // items below dense_min particles are the pencil and plane spreads'
    if (n < dense_min) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ float wmax[12];
    __shared__ int slots[12 * 16];   // per wave: the batch lanes of the open group's particles
    double* sub = reinterpret_cast<double*>(smem_raw);   // [y][x][z], S^3
    const int S = tl.S, S2 = S * S, S3 = S2 * S;
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int e = tid; e < S3; e += nthr) sub[e] = 0.0;
    item_qmax_part(rec, beg, end, tid, nthr, wmax);
    __syncthreads();

    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    const int col = wave & 3, part = wave >> 2, nparts = nw >> 2;
    const int rb = beg + (int)(((long)n * part) / nparts), re = beg + (int)(((long)n * (part + 1)) / nparts);
    int* myslots = slots + 16 * wave;
    const int i16 = lane & 15, k4 = lane >> 4;   // lane = i16 + 16 k4: x / z index, particle slot
    // fixed-point scale of the item's LDS sums: q d 2^sc is folded into the x factor
    const int sc = __builtin_amdgcn_readfirstlane(flush_scale((double)n * item_qmax_read(wmax, nw) * g.d, g.d));
    const float h = (float)g.h, c = (float)g.c;
    const float dn = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)ldexp(g.d, sc))));
    const float kc = -c * 1.4426950408889634f;
    const float oi = (float)(g.first + i16) * h;   // grid offset of window index i16

    v4f acc[GY];
#pragma unroll
    for (int s = 0; s < GY; ++s) acc[s] = v4f{0.f, 0.f, 0.f, 0.f};
    int rpen = -1, base = 0, rofs = 0;   // ring: column key, first row, this lane's point (x0 + 4 k4, 0, z0 + i16)

    // add ring rows [base, upto) (upto <= base + GY; rows past the sub-grid hold zeros)
    // into LDS and clear their slots
    auto flush = [&](int upto) __attribute__((always_inline)) {
        upto = min(upto, S);
        if (upto <= base) return;
        const unsigned mask = ring_slots<GY>(base, upto - base);
        const int b = base % GY;
#pragma unroll
        for (int s = 0; s < GY; ++s) {
            if (mask & (1u << s)) {
                // a distinct block per slot (so that the slots' blocks are not merged into
                // one that indexes the ring registers at run time)
                asm volatile("; ring slot %0" ::"n"(s));
                const int y = base + (s >= b ? s - b : s - b + GY);
                double* p = sub + y * S2 + rofs;
                const v4f v = acc[s];
                acc[s] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    __hip_atomic_fetch_add(p + r * S, (double)rintf(v[r]), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    };

    // The open group: particle 4 cc + k4 of K step cc has its factors on this lane's row,
    // formed when it joins the group: x factor (times q d 2^sc; zero on slots without a
    // particle), z factor, y factors of ring slots i16 and 16 + (i16 & 3) for the ring
    // starting at the group's first row.
    float gqx[KS], gfz[KS], gy1[KS], gy2[KS];
#pragma unroll
    for (int cc = 0; cc < KS; ++cc) gqx[cc] = gfz[cc] = gy1[cc] = gy2[cc] = 0.f;

    auto process = [&](int k, int gpen, int gy0, int lo0) __attribute__((always_inline)) {
        // one flush site (rows [base, upto) of the old ring), so the ring registers are
        // never copied between control-flow paths
        int upto = base;
        if (gpen != rpen) {
            if (rpen >= 0) upto = base + GY;
        } else if (gy0 > base) {
            upto = min(gy0, base + GY);
        }
        flush(upto);
        if (gpen != rpen) {
            rpen = gpen;
            rofs = ((lo0 & 255) + 4 * k4) * S + (lo0 >> 16) + i16;
        }
        base = gy0;
#pragma unroll
        for (int cc = 0; cc < KS; ++cc) {
            if (4 * cc < k) ColumnRing<0>::run(acc, gqx[cc], gy1[cc], gy2[cc], gfz[cc]);
            gqx[cc] = 0.f;   // the next group starts with empty slots
        }
    };

    // Scan (as in the column gather); the next batch of records is loaded while the current
    // one is processed.  Each pass of the outer loop fills one group and processes it at a
    // single call site.
    int have = 0, gpen = -1, gy0 = 0, glo0 = 0, pcarry = 0, ppen = -1;
    if (lane < 16) myslots[lane] = 0;   // every slot names a lane before its first fill
    int cb = rb - 64;                    // current batch (none yet)
    float4 ra = {0.f, 0.f, 0.f, 0.f}, ra_next = {0.f, 0.f, 0.f, 0.f};
    int lo = -1, lo_next = -1, pen = 0, ly = 0;
    bool match = false;
    unsigned long long m = 0;
    if (rb + lane < re) {
        ra_next = *reinterpret_cast<const float4*>(rec + rb + lane);
        lo_next = rec[rb + lane].lo;
    }
    for (;;) {
        bool close = false;
        while (!close) {
            if (m == 0) {
                if (cb + 64 >= re) break;   // no records left
                cb += 64;
                const int i = cb + lane;
                ra = ra_next;
                lo = lo_next;
                lo_next = -1;
                if (cb + 64 + lane < re) {
                    ra_next = *reinterpret_cast<const float4*>(rec + cb + 64 + lane);
                    lo_next = rec[cb + 64 + lane].lo;
                }
                const bool in = i < re;
                pen = lo & 0xFF00FC;
                ly = (lo >> 8) & 255;
                const int prev = __shfl_up(pen, 1, 64);
                const bool start = in && pen != (lane == 0 ? ppen : prev);
                const unsigned long long smask = __ballot(start);
                const int pord = pcarry - 1 +
                                 (int)__builtin_amdgcn_mbcnt_hi((unsigned)(smask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)smask, 0u)) +
                                 (start ? 1 : 0);
                pcarry += __builtin_popcountll(smask);
                ppen = __builtin_amdgcn_readlane(pen, 63);
                match = in && (lo & 3) == ((col + pord) & 3);
                m = __ballot(match);
                continue;
            }
            const int j = __builtin_ctzll(m);
            const int pj = __builtin_amdgcn_readlane(pen, j), lyj = __builtin_amdgcn_readlane(ly, j);
            if (have > 0 && (pj != gpen || lyj > gy0 + (GY - P))) break;   // the group is complete
            if (have == 0) {
                gpen = pj;
                gy0 = lyj;
                glo0 = __builtin_amdgcn_readlane(lo, j);
            }
            // this pencil's next matches whose windows fit the ring from gy0 (a run: the
            // records are sorted by pencil, then row)
            const unsigned long long same = __ballot(match && pen == gpen && ly <= gy0 + (GY - P)) & m;
            const int take = min(__builtin_popcountll(same), GMAX - have);
            const unsigned long long taken = first_bits(same, take);
            if ((taken >> lane) & 1ull) {
                const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(taken >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)taken, 0u));
                myslots[have + rank] = lane;
            }
            __builtin_amdgcn_wave_barrier();
            // the lanes of particle 4 cc + k4 fetch its fields from the batch lane in its slot
            // and form its factors; the ring will start at row gy0 (slot s holds the row
            // == s (mod GY) of [gy0, gy0 + GY))
            const int bm = gy0 % GY;
            const int t1 = i16 - bm, t2 = 16 + (i16 & 3) - bm;
            const int y1 = gy0 + (t1 < 0 ? t1 + GY : t1), y2 = gy0 + (t2 < 0 ? t2 + GY : t2);
#pragma unroll
            for (int cc = 0; cc < KS; ++cc) {
                if (4 * cc + 3 >= have && 4 * cc < have + take) {   // K step cc gains particles
                    const int p = 4 * cc + k4;
                    const int sl = 4 * myslots[p];
                    const bool fill = p >= have && p < have + take;
                    const float t0 = __int_as_float(__builtin_amdgcn_ds_bpermute(sl, __float_as_int(ra.x)));
                    const float ty = __int_as_float(__builtin_amdgcn_ds_bpermute(sl, __float_as_int(ra.y)));
                    const float t2f = __int_as_float(__builtin_amdgcn_ds_bpermute(sl, __float_as_int(ra.z)));
                    const float q = __int_as_float(__builtin_amdgcn_ds_bpermute(sl, __float_as_int(ra.w)));
                    const int lyp = __builtin_amdgcn_ds_bpermute(sl, ly);
                    const int w1 = y1 - lyp, w2 = y2 - lyp;   // window indices of the two slots' rows
                    const float qx = q * dn * gauss_c(kc, t0 - oi), fz = gauss_c(kc, t2f - oi);
                    const float f1 = (unsigned)w1 < (unsigned)P ? gauss_c(kc, ty - (float)(g.first + w1) * h) : 0.f;
                    const float f2 = (unsigned)w2 < (unsigned)P ? gauss_c(kc, ty - (float)(g.first + w2) * h) : 0.f;
                    if (fill) {
                        gqx[cc] = qx;
                        gfz[cc] = fz;
                        gy1[cc] = f1;
                        gy2[cc] = f2;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            have += take;
            m &= ~taken;
            close = have == GMAX;
        }
        if (have == 0) break;
        process(have, gpen, gy0, glo0);
        have = 0;
    }
    if (rpen >= 0) flush(base + GY);
    __syncthreads();
    float* out = partials + (size_t)item * tl.part_elems;
    SubWalk w(S, tid, nthr);   // canonical [x][y][z]: a = ix, b = iy, c = iz
    for (int e = tid; e < S3; e += nthr) {
        out[e] = (float)ldexp(sub[(w.b * S + w.a) * S + w.c], -sc);
        w.next();
    }
}

// Column spread of the items of at least dense_min particles: work-groups of `block`
// threads (256, 512 or 768; the kernel is held to 168 registers, 3 waves per SIMD), one
// per CU (the FP64 sub-grid fills the LDS).
bool launch_spread_col16(const Rec<float>* rec, const int* item_beg, const int* item_end, const int* n_items,
                         const Geom& g, const Tiling& tl, float* partials, int dense_min, int block, hipStream_t s)
{
    SE_REQUIRE(g.P == 16 && tl.S >= 16 + tl.TS - 1 && block % 256 == 0 && block >= 256 && block <= 768, SE_ERR_ARG,
               "column spread: P = 16 and 256, 512 or 768 threads");
    const size_t lds = (size_t)tl.S * tl.S * tl.S * sizeof(double);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_spread_col16, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_spread_col16, dim3(tl.max_items), dim3(block), lds, s, rec, item_beg, item_end, n_items, g,
                       tl, partials, dense_min);
    return true;
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

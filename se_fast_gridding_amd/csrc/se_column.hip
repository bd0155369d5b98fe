// Column kernels: the FP32 P = 16 path for dense work items on the matrix cores.
//
// Particles are sorted (tile, pencil, row), a pencil being 4 x 1 window-start columns
// (4 in x, 1 in z) walked along y.  The particles of one 1 x 1 column (one lx, one lz)
// share their x and z windows exactly, so for any grid row y the z contraction of up to
// sixteen of them is one 16 x 16 x 16 matrix product:
//
//   T_y[x][n] = sum_z G[y][x0 + x][z0 + z] Fz_n(z),     x, z = 0..15, n = particle
//
// which v_mfma_f32_16x16x4f32 runs as four K = 4 steps (A = the sub-grid slice, read
// from LDS; B = the particles' z factors), and the spread's rank-1 updates of one row,
// four particles at a time, are one such instruction too.  Every product is a useful one:
// no window padding, no masking.
//
// A dense item's records are first put in column order, (pencil, lx mod 4, row), by
// k_column_order (a stable partition of each pencil by column), so that a column's
// particles are consecutive records in row order: a group is a run of up to sixteen
// consecutive records of one column, and the work of an item splits into equal record
// shares per wave.  k_column_order also decides which items are column items (item_col);
// the other kernels of the spread and the gather skip exactly these.
//
// FP32 matrix instructions and FP32 vector instructions do not overlap on this GPU (one
// v_mfma_f32_16x16x4f32 and eight v_fma_f32 per iteration take the sum of their times, at
// two or four waves per SIMD), so the kernels keep the vector work per matrix
// instruction small rather than counting on it to hide behind them.
#include <hip/hip_runtime.h>

#include "se_internal.h"

namespace se {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));

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

// D += A B on 16 x 16 x 16 fp16 operands (FP32 sums)
__device__ __forceinline__ v4f mfma16h(v4h a, v4h b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// fp16 pair of v: hi = fp16(v), lo = fp16(v - hi), so that hi + lo holds v to about 2^-22
// of |v| (or 2^-25 absolute where lo is subnormal).  v is pinned as an FP32 value first:
// otherwise the compiler may fuse the product that forms v into the lo conversion
// (v_fma_mix, exact product) while forming hi from the rounded product, and where the two
// round to different fp16 values hi + lo is off by an fp16 ulp of v.
__device__ __forceinline__ void split_f16(float v, _Float16& hi, _Float16& lo)
{
    asm volatile("" : "+v"(v));
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// four (hi | lo << 16) words -> the four his and the four los as fp16 vectors
__device__ __forceinline__ v4h pack_hi(unsigned p0, unsigned p1, unsigned p2, unsigned p3)
{
    const v2u u = {__builtin_amdgcn_perm(p1, p0, 0x05040100u), __builtin_amdgcn_perm(p3, p2, 0x05040100u)};
    return __builtin_bit_cast(v4h, u);
}
__device__ __forceinline__ v4h pack_lo(unsigned p0, unsigned p1, unsigned p2, unsigned p3)
{
    const v2u u = {__builtin_amdgcn_perm(p1, p0, 0x07060302u), __builtin_amdgcn_perm(p3, p2, 0x07060302u)};
    return __builtin_bit_cast(v4h, u);
}

__device__ __forceinline__ float bperm_f(int byte_addr, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(byte_addr, __float_as_int(v)));
}

// Exclusive scan over the work-group of one value per thread (wave scans by shuffles,
// then the wave totals through LDS); returns the combination of the threads before this
// one (ident for thread 0).
template <typename T, typename Op>
__device__ __forceinline__ T block_exclusive(T v, T ident, Op op, T* wtot)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T u = __shfl_up(inc, o, 64);
        if (lane >= o) inc = op(inc, u);
    }
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    T before = ident;
    for (int w = 0; w < wave; ++w) before = op(before, wtot[w]);
    __syncthreads();
    const T ex = __shfl_up(inc, 1, 64);
    return op(before, lane == 0 ? ident : ex);
}

}  // namespace

// ---------------------------------------------------------------------------------
// Column order.  One work-group per item of column_floor to COLUMN_MAX particles; each
// thread holds eight consecutive records in registers and the records' window words sit in
// LDS for the neighbour tests.
//
// The item is a column item if the column kernels' groups (runs of up to sixteen
// particles of one 1 x 1 column within five rows) hold COLUMN_GROUP_FILL particles or more
// on average: their work per group does not shrink with the group, so thinly filled
// groups are cheaper on the pencil kernels.  For a column item the kernel writes the fixed-point scale of its column spread
// sums to item_col[item] (COLUMN_NONE for every other item) from a bound on any sub-grid
// sum, sum_i |q_i| d Fx_i Fy_i Fz_i <= max|q| d B: with n(lx, lz, ly) the particles per
// window-start cell, gm(w) the largest factor of support point w over a particle's place
// in its cell, and
//   Y(lx, lz) = max_y sum_ly n(lx, lz, ly) gm(y - ly),
//   B = max_{x,z} sum_{lx,lz} Y(lx, lz) gm(x - lx) gm(z - lz)
// (separable: the z sum first), which for spread-out items is tens of times below the
// particle count that bounds it crudely, and so buys the 32-bit sums as many bits.
//
// The records are then put in column order: each pencil's records, in row order, are
// partitioned stably by their column (lx mod 4), so the records become (pencil, column,
// row).  Record i goes to
//   s + (records of the pencil in lower columns) + (records of its column before it),
// s the pencil's first record, from per-column running counts (four 16-bit fields of one
// 64-bit count, scanned over the item) taken at the pencil's start and end.  Every record
// is read before the first barrier and written after the last, so the permutation is
// done in place.
// ---------------------------------------------------------------------------------
constexpr int COL_RING = 20;     // column spread: grid rows of a wave's register ring
constexpr int COL_KSTEPS = 4;    // K steps (of four particles) per group
constexpr int COLUMN_ORDER_THREADS = 256;
constexpr int COLUMN_PER_THREAD = COLUMN_MAX / COLUMN_ORDER_THREADS;

__device__ __forceinline__ unsigned long long col_one(int lo) { return 1ull << (16 * (lo & 3)); }
__device__ __forceinline__ int col_field(unsigned long long c, int d) { return (int)((c >> (16 * d)) & 0xFFFF); }

// Largest factor exp(-c a^2) a particle gives support point w over its place in its cell,
// t in [0, h) (even P) or [-h/2, h/2) (odd P), a = t - (first + w) h: m h is the least
// |a| over the cell.
__device__ __forceinline__ float window_factor_max(const Geom& g, int w)
{
    const float a = (float)(g.first + w);
    const float m = g.odd ? fmaxf(fabsf(a) - 0.5f, 0.f) : fmaxf(fmaxf(-a, a - 1.f), 0.f);
    return expf(-(float)(g.c * g.h * g.h) * m * m);
}

__global__ __launch_bounds__(COLUMN_ORDER_THREADS) void k_column_order(Rec<float>* __restrict__ rec,
                                                                       const int* __restrict__ item_beg,
                                                                       const int* __restrict__ item_end,
                                                                       const int* __restrict__ n_items, Geom g,
                                                                       int TS, int floor_n,
                                                                       int* __restrict__ item_col)
{
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    const int beg = item_beg[item], n = item_end[item] - beg;
    // The records are in (pencil, row) key order, so the item spans keys key(first) to
    // key(last); spread evenly over them its particles would fill a column group (five rows
    // of one of a pencil's four columns) with 1.25 n / span.  Items below half the fill
    // threshold on that estimate are left without loading their records.
    bool take = n >= floor_n && n <= COLUMN_MAX;
    if (take) {
        const auto key = [TS](int lo) { return (((lo & 255) >> 2) * TS + (lo >> 16)) * TS + ((lo >> 8) & 255); };
        const int span = key(rec[beg + n - 1].lo) - key(rec[beg].lo) + 1;
        take = (float)n >= 0.4f * COLUMN_GROUP_FILL * (float)span;
    }
    if (!take) {
        if (threadIdx.x == 0) item_col[item] = COLUMN_NONE;
        return;
    }
    __shared__ int los[COLUMN_MAX];                        // the records' window words
    __shared__ unsigned long long cst[COLUMN_MAX];         // per-column counts before a pencil's first record
    __shared__ unsigned long long cen[COLUMN_MAX];         // ... and after its last (indexed by its first)
    __shared__ int hist[16 * 16 * 16];                     // particles per window-start cell, [lx][lz][ly]
    __shared__ float ycol[16 * 16];                        // Y(lx, lz)
    __shared__ float tz[16 * 31];                          // sum_lz Y(lx, lz) gm(z - lz), [lx][z]
    __shared__ float gmt[32];                              // gm(w), 0 past the support
    __shared__ unsigned int wlo[16], whi[16];
    __shared__ int wmx[16];
    __shared__ float wq[16], wb[16];
    __shared__ int ngroups;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int P = g.P, S = TS + P - 1, T3 = TS * TS * TS;
    const int i0 = t * COLUMN_PER_THREAD;   // thread t owns records i0 .. i0 + COLUMN_PER_THREAD - 1
    v4f* rp = reinterpret_cast<v4f*>(rec + beg);   // plain vectors (no HIP vector-type copies)
    v4f ra[COLUMN_PER_THREAD], rb[COLUMN_PER_THREAD];
#pragma unroll
    for (int u = 0; u < COLUMN_PER_THREAD; ++u) {
        const int i = min(i0 + u, n - 1);   // threads past the item reread its last record
        ra[u] = rp[2 * i];
        rb[u] = rp[2 * i + 1];   // lo, orig, pad
    }
    for (int e = t; e < T3; e += COLUMN_ORDER_THREADS) hist[e] = 0;
    if (t < 32) gmt[t] = t < P ? window_factor_max(g, t) : 0.f;
    if (t == 0) ngroups = 0;
    __syncthreads();
    float qm = 0.f;
#pragma unroll
    for (int u = 0; u < COLUMN_PER_THREAD; ++u) {
        const int i = i0 + u;
        qm = fmaxf(qm, fabsf(ra[u].w));
        if (i < n) {
            const int lo = __float_as_int(rb[u].x);
            los[i] = lo;
            atomicAdd(&hist[((lo & 255) * TS + (lo >> 16)) * TS + ((lo >> 8) & 255)], 1);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) qm = fmaxf(qm, __shfl_xor(qm, o, 64));
    if (lane == 0) wq[wave] = qm;
    __syncthreads();

    // Y of column t (TS <= 16, S <= 31: registers with static indices), and the column
    // spread's groups in the item
    if (t < TS * TS) {
        int hc[16];
        float gw[16];
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            hc[l] = l < TS ? hist[t * TS + l] : 0;
            gw[l] = gmt[l];
        }
        float ym = 0.f;
#pragma unroll
        for (int y = 0; y < 31; ++y) {
            float v = 0.f;
#pragma unroll
            for (int l = (y > 15 ? y - 15 : 0); l <= (y < 15 ? y : 15); ++l) v += (float)hc[l] * gw[y - l];
            ym = fmaxf(ym, v);   // rows past the sub-grid add nothing (no terms)
        }
        ycol[t] = ym;
        // the column's groups: runs of up to 16 of its records (in row order) whose rows
        // lie within COL_RING - 16 of the run's first, counted row by row (static indices)
        int groups = 0, gs = -COL_RING, cap = 0;   // the open group's first row and room
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            int c = hc[r];
            while (c > 0) {
                if (cap == 0 || r > gs + (COL_RING - 16)) {
                    ++groups;
                    gs = r;
                    cap = 16;
                }
                const int k = min(c, cap);
                c -= k;
                cap -= k;
            }
        }
        if (groups) atomicAdd(&ngroups, groups);
    }
    __syncthreads();
    if ((float)n < COLUMN_GROUP_FILL * (float)ngroups) {   // thinly filled groups: the pencil kernels take the item
        if (t == 0) item_col[item] = COLUMN_NONE;
        return;
    }

    // record i starts its pencil (lx / 4, lz)
#define SE_STARTS(i, lo) ((i) == 0 || ((los[(i) - 1] ^ (lo)) & 0xFF00FC) != 0)

    unsigned long long cl[COLUMN_PER_THREAD];
    unsigned long long cnt = 0;
    int last = -1;
#pragma unroll
    for (int u = 0; u < COLUMN_PER_THREAD; ++u) {
        const int i = i0 + u;
        cl[u] = cnt;
        if (i < n) {
            const int lo = __float_as_int(rb[u].x);
            cnt += col_one(lo);
            if (SE_STARTS(i, lo)) last = i;
        }
    }
    // counts before this thread's records (a field counts at most COLUMN_MAX < 2^16, so
    // no field carries into the next, and the two 32-bit halves are scanned separately)
    const unsigned int blo = block_exclusive<unsigned int>((unsigned int)cnt, 0u,
                                                           [](unsigned int a, unsigned int b) { return a + b; }, wlo);
    const unsigned int bhi = block_exclusive<unsigned int>((unsigned int)(cnt >> 32), 0u,
                                                           [](unsigned int a, unsigned int b) { return a + b; }, whi);
    const int prev_start = block_exclusive<int>(last, -1, [](int a, int b) { return a > b ? a : b; }, wmx);
    const unsigned long long cbase = ((unsigned long long)bhi << 32) | blo;
    int s = prev_start;
#pragma unroll
    for (int u = 0; u < COLUMN_PER_THREAD; ++u) {
        const int i = i0 + u;
        if (i < n) {
            const int lo = __float_as_int(rb[u].x);
            const unsigned long long cv = cbase + cl[u];
            if (SE_STARTS(i, lo)) {
                s = i;
                cst[i] = cv;
            }
            if (i == n - 1 || ((los[i + 1] ^ lo) & 0xFF00FC) != 0) cen[s] = cv + col_one(lo);
        }
    }
    for (int e = t; e < TS * S; e += COLUMN_ORDER_THREADS) {
        const int lx = e / S, z = e - lx * S;
        float v = 0.f;
        for (int lz = max(0, z - P + 1); lz <= min(z, TS - 1); ++lz) v += ycol[lx * TS + lz] * gmt[z - lz];
        tz[e] = v;
    }
    __syncthreads();
    float bm = 0.f;
    for (int e = t; e < S * S; e += COLUMN_ORDER_THREADS) {
        const int x = e / S, z = e - x * S;
        float v = 0.f;
        for (int lx = max(0, x - P + 1); lx <= min(x, TS - 1); ++lx) v += gmt[x - lx] * tz[lx * S + z];
        bm = fmaxf(bm, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bm = fmaxf(bm, __shfl_xor(bm, o, 64));
    if (lane == 0) wb[wave] = bm;
    __syncthreads();
    if (t == 0) {
        float b = 0.f, q = 0.f;
        for (int w = 0; w < COLUMN_ORDER_THREADS / 64; ++w) {
            b = fmaxf(b, wb[w]);
            q = fmaxf(q, wq[w]);
        }
        item_col[item] = flush_scale_i32((double)q * g.d * b, g.d);
    }

    s = prev_start;
#pragma unroll
    for (int u = 0; u < COLUMN_PER_THREAD; ++u) {
        const int i = i0 + u;
        if (i < n) {
            const int lo = __float_as_int(rb[u].x);
            if (SE_STARTS(i, lo)) s = i;
            const int d = lo & 3;
            const unsigned long long run = cen[s] - cst[s];
            const int before = (d > 0 ? col_field(run, 0) : 0) + (d > 1 ? col_field(run, 1) : 0) +
                               (d > 2 ? col_field(run, 2) : 0);
            const int dst = s + before + col_field(cbase + cl[u] - cst[s], d);
            if ((unsigned)dst < (unsigned)n) {   // always, by construction; keeps every store in the item
                rp[2 * dst] = ra[u];
                rp[2 * dst + 1] = rb[u];
            }
        }
    }
#undef SE_STARTS
}

void launch_column_order(Rec<float>* rec, const int* item_beg, const int* item_end, const int* n_items, int max_items,
                         const Geom& g, const Tiling& tl, int* item_col, hipStream_t s)
{
    SE_REQUIRE(tl.TS <= 16 && g.P <= 16, SE_ERR_ARG, "column order: tile side or support past 16");
    hipLaunchKernelGGL(k_column_order, dim3(max_items), dim3(COLUMN_ORDER_THREADS), 0, s, rec, item_beg, item_end,
                       n_items, g, tl.TS, column_floor(tl), item_col);
}

// ---------------------------------------------------------------------------------
// Column gather.  The y and x contractions are per particle and stay on the VALU:
//
//   U = sum_y Fy T,  V = sum_y Fy' T,  W = sum_y Fy T'   (T' with the z derivative factors)
//   phi = sum_x Fx U,  F = (sum_x Fx' U, sum_x Fx V, sum_x Fx W)
//
// A group costs (rows spanned + 15) row steps of 4 (potential) or 8 (potential and
// force) matrix instructions, whatever its 1..16 particles; in a dense column, sixteen
// consecutive particles span one or two rows, so nearly every instruction is full.  Rows
// outside a particle's window weigh exactly zero, so each particle's result depends only on
// its own record and on the sub-grid, not on the grouping.
//
// v_mfma_f32_16x16x4f32 operand layouts (one block): A[i][k] in lane i + 16 k, B[k][j] in
// lane j + 16 k, D[i][j] in register i mod 4 of lane j + 16 (i / 4).  Here lane j + 16 q
// (q = 0..3) holds: A = sub-grid point (x0 + j, z0 + 4 q + c) of K step c; B = particle
// j's z factor at window point 4 q + c; D = T_y[4 q + r][j] in register r.
// ---------------------------------------------------------------------------------
template <bool POT, bool FORCE>
__global__ __launch_bounds__(1024) void k_interp_col16(const Rec<float>* __restrict__ rec,
                                                      const int* __restrict__ item_tile,
                                                      const int* __restrict__ item_beg,
                                                      const int* __restrict__ item_end,
                                                      const int* __restrict__ n_items,
                                                      const float* __restrict__ grid, Geom g, Tiling tl, int N,
                                                      double* __restrict__ phi, double* __restrict__ force,
                                                      const SysDesc* __restrict__ sys,
                                                      const int* __restrict__ item_col)
{
    constexpr int P = 16;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    // the other items are the row (or vector) gather's
    if (item_col[item] == COLUMN_NONE) return;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* sub = reinterpret_cast<float*>(smem_raw);   // [y][x][z], S^3
    const int S = tl.S, S2 = S * S, M = g.M, Mz = g.Mz, NT = tl.NT, NTz = tl.NTz;
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

    // The sub-grid becomes fp16 pairs in place: G 2^sg = hi + lo, one (hi | lo << 16) word
    // per point, sg such that max |G| 2^sg lies in [2^14, 2^15).
    __shared__ float wmax[16];
    const int S3 = S2 * S;
    {
        float m = 0.f;
        for (int e = tid; e < S3; e += nthr) m = fmaxf(m, fabsf(sub[e]));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if (lane == 0) wmax[wave] = m;
    }
    __syncthreads();
    float gmax = 0.f;
    for (int w = 0; w < nw; ++w) gmax = fmaxf(gmax, wmax[w]);
    const int sg = gmax > 0.f ? 14 - ilogbf(gmax) : 0;
    unsigned* subp = reinterpret_cast<unsigned*>(sub);
    for (int e = tid; e < S3; e += nthr) {
        _Float16 hi, lo;
        split_f16(ldexpf(sub[e], sg), hi, lo);
        subp[e] = (unsigned)__builtin_bit_cast(unsigned short, hi) |
                  ((unsigned)__builtin_bit_cast(unsigned short, lo) << 16);
    }
    __syncthreads();

    // equal shares of the item's records, one per wave
    const int rb = beg + (int)(((long)n * wave) / nw), re = beg + (int)(((long)n * (wave + 1)) / nw);
    const int j = lane & 15, q = lane >> 4;
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;
    const float kc = -c * 1.4426950408889634f;
    const float fh = (float)g.first * h;
    const double h3 = g.h * g.h * g.h;
    // z factors enter the matrix instructions as fp16 pairs of Fz 2^15 (Fz <= 1) and of
    // Fz' 2^sd (|Fz'| = c |a| exp(-c a^2) <= sqrt(c / 2e))
    const int sd = 14 - ilogbf(sqrtf(c * 0.18393972058572117f));

    // Groups: runs of up to 16 consecutive records of one column; lane j + 16 q holds record
    // g0 + j (slots past the run repeat its last record and store nothing).  The next
    // group's records are loaded as soon as its first record is known, so their latency
    // runs under this group's rows.
    if (rb >= re) return;
    v4f rn = *reinterpret_cast<const v4f*>(rec + min(rb + j, re - 1));   // t[0], t[1], t[2], q
    int lon = rec[min(rb + j, re - 1)].lo, orn = rec[min(rb + j, re - 1)].orig;
    // settled here, so that no load is pending where the loop starts (otherwise the wait the
    // compiler places at the top of the loop would also cover the previous group's stores)
    asm volatile("" ::"v"(rn.x), "v"(rn.y), "v"(rn.z), "v"(rn.w), "v"(lon), "v"(orn));
    for (int g0 = rb; g0 < re;) {
        const v4f r = rn;
        const int lo = lon, orig = orn;
        const int lo0 = __builtin_amdgcn_readfirstlane(lo);
        const unsigned long long same = __ballot(q == 0 && g0 + j < re && (lo & 0xFF00FF) == (lo0 & 0xFF00FF));
        const int k = __builtin_ctzll(~same);   // the run: lanes 0 .. k - 1 (1 <= k <= 16)
        const bool act = j < k;
        const int ly = (lo >> 8) & 255;
        const int x0 = lo0 & 255, z0 = lo0 >> 16, y0 = (lo0 >> 8) & 255;
        const int y1 = ((__builtin_amdgcn_readlane(lo, k - 1) >> 8) & 255) + P;
        {
            const int sn = min(g0 + k + j, re - 1);   // the next group's records
            rn = *reinterpret_cast<const v4f*>(rec + sn);
            lon = rec[sn].lo;
            orn = rec[sn].orig;
        }
        // B operands: particle j's z factors at window points 4 q + s, s = 0..3, as fp16 pairs
        v4h bh, bl, bdh, bdl;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float az = fmaf(-(float)(4 * q + s), h, r.z - fh);
            const float bz = gauss_c(kc, az);
            _Float16 hi, lo;
            split_f16(bz * 32768.f, hi, lo);
            bh[s] = hi;
            bl[s] = lo;
            if (FORCE) {
                split_f16(ldexpf(bz * (-c * az), sd), hi, lo);
                bdh[s] = hi;
                bdl[s] = lo;
            }
        }
        const unsigned* ap = subp + (x0 + j) * S + z0 + 4 * q;
        const float tyf = r.y - fh;   // y offset of window index w: tyf - w h
        const v4f zero4 = {0.f, 0.f, 0.f, 0.f};
        v4f U = zero4, V = zero4, W = zero4;
        // Software pipeline: the next row's sub-grid values are read, and the previous
        // row's y sums formed, around this row's matrix instructions, so neither the LDS
        // latency nor the matrix results stall the wave.  After the last row the read is of
        // row S - 1 again (in the sub-grid) and is not used.
        v4f Tp = zero4, Tdp = zero4;
        float fyp = 0.f, fydp = 0.f;
        const unsigned* a = ap + y0 * S2;
        unsigned a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
        for (int y = y0; y < y1; ++y) {
            // G Fz = hi hi + lo hi + hi lo (lo lo, about 2^-22 of the product, left out);
            // the small terms first
            const v4h ah = pack_hi(a0, a1, a2, a3), al = pack_lo(a0, a1, a2, a3);
            v4f T = mfma16h(al, bh, zero4);
            T = mfma16h(ah, bl, T);
            T = mfma16h(ah, bh, T);
            v4f Td = zero4;
            if (FORCE) {
                Td = mfma16h(al, bdh, zero4);
                Td = mfma16h(ah, bdl, Td);
                Td = mfma16h(ah, bdh, Td);
            }
            a = ap + min(y + 1, S - 1) * S2;
            a0 = a[0];
            a1 = a[1];
            a2 = a[2];
            a3 = a[3];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                U[i] = fmaf(fyp, Tp[i], U[i]);
                if (FORCE) {
                    V[i] = fmaf(fydp, Tp[i], V[i]);
                    W[i] = fmaf(fyp, Tdp[i], W[i]);
                }
            }
            const int jy = y - ly;
            const float ay = fmaf(-(float)jy, h, tyf);
            fyp = (act && (unsigned)jy < (unsigned)P) ? gauss_c(kc, ay) : 0.f;
            fydp = fyp * (-c * ay);
            Tp = T;
            Tdp = Td;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            U[i] = fmaf(fyp, Tp[i], U[i]);
            if (FORCE) {
                V[i] = fmaf(fydp, Tp[i], V[i]);
                W[i] = fmaf(fyp, Tdp[i], W[i]);
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
        // undo the operand scales (powers of two: exact)
        if (POT) sp = ldexpf(sum_rows4(sp), -(sg + 15));
        if (FORCE) {
            sx = ldexpf(sum_rows4(sx), -(sg + 15));
            sy = ldexpf(sum_rows4(sy), -(sg + 15));
            sz = ldexpf(sum_rows4(sz), -(sg + sd));
        }
        // Wait for the next group's records here, before this group's output stores: vmcnt
        // counts loads and stores together, so a wait placed after the stores (at the top of
        // the next group, where the compiler would put it) would also wait for them.
        asm volatile("" ::"v"(rn.x), "v"(rn.y), "v"(rn.z), "v"(rn.w), "v"(lon), "v"(orn));
        if (act) {   // one output per lane row
            if (POT && q == 0) store_out(phi, orig, h3 * (double)sp);
            if (FORCE && q > 0) {
                const double qh = h3 * (double)r.w;
                const float sv = q == 1 ? sx : (q == 2 ? sy : sz);
                store_out(force, orig + (long)(q - 1) * N, qh * (double)sv);
            }
        }
        g0 += k;
    }
}

// Column gather of the column items; block = 256, 512 or 1024.
void launch_interp_col16(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                         const int* n_items, const float* grid, const Geom& g, const Tiling& tl, int N,
                         double* phi, double* force, const SysDesc* sys, const int* item_col, int block,
                         hipStream_t s)
{
    SE_REQUIRE(g.P == 16 && tl.S >= 16 + tl.TS - 1 && (block == 256 || block == 512 || block == 1024), SE_ERR_ARG,
               "column gather: P = 16 and 256, 512 or 1024 threads");
    const size_t lds = ((size_t)tl.S * tl.S * tl.S + LDS_TILE_PAD) * sizeof(float);
    dim3 gr(tl.max_items), bl(block);
#define SE_GC(POT, FORCE)                                                                                    \
    do {                                                                                                     \
        if (lds > 64 * 1024)                                                                                 \
            (void)hipFuncSetAttribute((const void*)k_interp_col16<POT, FORCE>,                               \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                 \
        hipLaunchKernelGGL((k_interp_col16<POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg, item_end, \
                           n_items, grid, g, tl, N, phi, force, sys, item_col);                              \
    } while (0)
    if (phi && force)
        SE_GC(true, true);
    else if (force)
        SE_GC(false, true);
    else
        SE_GC(true, false);
#undef SE_GC
}

// ---------------------------------------------------------------------------------
// Column spread.  For a grid row y, four particles k of one column add
//
//   D_y[x][z] += sum_k (q_k d Fx_k(x) Fy_k(y)) Fz_k(z),    x, z = 0..15
//
// one v_mfma_f32_16x16x4f32 (A = the particles' x factors times their y factor at row y,
// B = their z factors; K = the four particles).  A wave keeps a ring of 20 grid rows of
// its current column in registers (D of row y in ring slot y mod 20, four VGPRs each).  A
// group is a run of up to sixteen consecutive records of one column whose windows lie in
// the ring [y0, y0 + 20), y0 the first one's row: four K steps of four, each 20 matrix
// instructions with no branch.  Lane i of particle k's 16-lane row holds the y factor of
// ring slot i (and of slot 16 + i), so the A operand of slot s is one DPP multiply
// (row_newbcast: s mod 16) with every register index static.  Rows that leave the ring are
// added to the tile's LDS sub-grid: each value, scaled by 2^s per work item (q d 2^s is
// folded into the x factor) and rounded to an integer, is added exactly, so the sub-grid
// sums do not depend on the order of the atomics; the ring's own FP32 sums run in the fixed
// order of the records.  The sub-grid then becomes the item's FP32 partial, as in the
// pencil spread, except that the sub-grid holds 32-bit integers (quantum about
// n max|q| d 2^-29 per item): half the FP64 sub-grid's LDS, so two work-groups share a CU
// and one's zeroing, flush and write-out overlap the other's matrix work, and
// `ds_add_u32` runs at about twice the rate of `ds_add_f64`.  Holding all S rows of the
// column instead of the ring, with a uniform branch per row and K step, measured slower
// (clustered N = 4M at M = 256, dense items: 1.03 against 0.93 ms): the branches cost
// more than the ring's extra rows.
// ---------------------------------------------------------------------------------

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
                                                      float* __restrict__ partials,
                                                      const int* __restrict__ item_col)
{
    constexpr int P = 16, GY = COL_RING, KS = COL_KSTEPS;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    // the other items are the pencil and plane spreads'; a column item's entry is the
    // fixed-point scale of its sums: q d 2^sc is folded into the x factor
    const int sc = item_col[item];
    if (sc == COLUMN_NONE) return;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int* sub = reinterpret_cast<int*>(smem_raw);   // [y][x][z], S^3 fixed-point sums
    const int S = tl.S, S2 = S * S, S3 = S2 * S;
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int e = tid; e < S3; e += nthr) sub[e] = 0;
    __syncthreads();

    // equal shares of the item's records, one per wave
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    const int rb = beg + (int)(((long)n * wave) / nw), re = beg + (int)(((long)n * (wave + 1)) / nw);
    const int i16 = lane & 15, k4 = lane >> 4;   // lane = i16 + 16 k4: x / z index, particle slot
    const float h = (float)g.h, c = (float)g.c;
    const float dn = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)ldexp(g.d, sc))));
    const float kc = -c * 1.4426950408889634f;
    const float oi = (float)(g.first + i16) * h;   // grid offset of window index i16

    v4f acc[GY];
#pragma unroll
    for (int s = 0; s < GY; ++s) acc[s] = v4f{0.f, 0.f, 0.f, 0.f};
    int rcol = -1, base = 0, rofs = 0;   // ring: column, first row, this lane's point (x0 + 4 k4, 0, z0 + i16)

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
                int* p = sub + y * S2 + rofs;
                const v4f v = acc[s];
                acc[s] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    __hip_atomic_fetch_add(p + r * S, (int)rintf(v[r]), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    };

    // Groups: lanes 0 .. 15 read records g0 .. g0 + 15; the run of one column whose rows
    // fit the ring from the first one's row is the group (1 to 16 particles, one pass of
    // the loop each, so the ring has one flush site).  The lanes of particle 4 cc + k4
    // fetch its fields by ds_bpermute and form its factors: x factor times q d 2^sc (zero
    // on slots without a particle), z factor, y factors of ring slots i16 and
    // 16 + (i16 & 3) of the ring [y0, y0 + GY).
    // The next group's records are loaded as soon as its first record is known (a wave
    // with an empty share loads nothing; it still takes the barrier below).
    v4f rn = {0.f, 0.f, 0.f, 0.f};   // t[0], t[1], t[2], q
    int lon = 0;
    if (rb < re) {
        rn = *reinterpret_cast<const v4f*>(rec + min(rb + i16, re - 1));
        lon = rec[min(rb + i16, re - 1)].lo;
    }
    for (int g0 = rb; g0 < re;) {
        const v4f ra = rn;
        const int lo = lon;
        const int lo0 = __builtin_amdgcn_readfirstlane(lo);
        const int col = lo0 & 0xFF00FF, y0 = (lo0 >> 8) & 255;
        const int ly = (lo >> 8) & 255;
        const unsigned long long ok =
            __ballot(k4 == 0 && g0 + i16 < re && (lo & 0xFF00FF) == col && ly <= y0 + (GY - P));
        const int k = __builtin_ctzll(~ok);   // the group: lanes 0 .. k - 1 (1 <= k <= 16)
        {
            const int sn = min(g0 + k + i16, re - 1);   // the next group's records
            rn = *reinterpret_cast<const v4f*>(rec + sn);
            lon = rec[sn].lo;
        }
        // one flush site (rows [base, upto) of the old ring), so the ring registers are
        // never copied between control-flow paths
        int upto = base;
        if (col != rcol) {
            if (rcol >= 0) upto = base + GY;
        } else if (y0 > base) {
            upto = min(y0, base + GY);
        }
        flush(upto);
        if (col != rcol) {
            rcol = col;
            rofs = ((lo0 & 255) + 4 * k4) * S + (lo0 >> 16) + i16;
        }
        base = y0;
        const int bm = y0 % GY;
        const int t1 = i16 - bm, t2 = 16 + (i16 & 3) - bm;
        const int yr1 = y0 + (t1 < 0 ? t1 + GY : t1), yr2 = y0 + (t2 < 0 ? t2 + GY : t2);   // rows of the slots
#pragma unroll
        for (int cc = 0; cc < KS; ++cc) {
            if (4 * cc < k) {
                const int p = 4 * cc + k4;
                const int sl = 4 * min(p, k - 1);
                const float t0 = bperm_f(sl, ra.x), ty = bperm_f(sl, ra.y), tz = bperm_f(sl, ra.z),
                            qv = bperm_f(sl, ra.w);
                const int lyp = __builtin_amdgcn_ds_bpermute(sl, ly);
                const int w1 = yr1 - lyp, w2 = yr2 - lyp;   // window indices of the two slots' rows
                const float qx = p < k ? qv * dn * gauss_c(kc, t0 - oi) : 0.f;
                const float fz = gauss_c(kc, tz - oi);
                const float f1 = (unsigned)w1 < (unsigned)P ? gauss_c(kc, ty - (float)(g.first + w1) * h) : 0.f;
                const float f2 = (unsigned)w2 < (unsigned)P ? gauss_c(kc, ty - (float)(g.first + w2) * h) : 0.f;
                ColumnRing<0>::run(acc, qx, f1, f2, fz);
            }
        }
        g0 += k;
    }
    if (rcol >= 0) flush(base + GY);
    __syncthreads();
    float* out = partials + (size_t)item * tl.part_elems;
    SubWalk w(S, tid, nthr);   // canonical [x][y][z]: a = ix, b = iy, c = iz
    for (int e = tid; e < S3; e += nthr) {
        out[e] = ldexpf((float)sub[(w.b * S + w.a) * S + w.c], -sc);
        w.next();
    }
}

// Column spread of the column items: work-groups of `block` threads (256, 512 or 768; the
// kernel is held to 168 registers).  At 512 threads two work-groups share a CU (LDS 79 KB
// each at tile 12, and registers for four waves per SIMD).
bool launch_spread_col16(const Rec<float>* rec, const int* item_beg, const int* item_end, const int* n_items,
                         const Geom& g, const Tiling& tl, float* partials, const int* item_col, int block,
                         hipStream_t s)
{
    SE_REQUIRE(g.P == 16 && tl.S >= 16 + tl.TS - 1 && block % 256 == 0 && block >= 256 && block <= 768, SE_ERR_ARG,
               "column spread: P = 16 and 256, 512 or 768 threads");
    const size_t lds = (size_t)tl.S * tl.S * tl.S * sizeof(int);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_spread_col16, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_spread_col16, dim3(tl.max_items), dim3(block), lds, s, rec, item_beg, item_end, n_items, g,
                       tl, partials, item_col);
    return true;
}

}  // namespace se

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

// Ring slots holding rows [lo, lo + n) (0 <= n <= P) as a bit mask: slot s holds the
// row == s (mod P), so the mask is n consecutive bits rotated by lo mod P.
template <int P>
__device__ __forceinline__ unsigned ring_mask(int lo, int n)
{
    const unsigned m = n >= P ? (1u << P) - 1u : (1u << n) - 1u;
    const int b = lo % P;
    return ((m << b) | (m >> (P - b))) & ((1u << P) - 1u);
}

// exp(-c a^2) as a native v_exp_f32 of kc * a^2, kc = -c log2(e) (about 1 ulp; the
// argument stays far above the FP32 denormal range for the supports used here)
__device__ __forceinline__ float gauss(float kc, float a) { return __builtin_amdgcn_exp2f(kc * a * a); }

// Particle records are fetched 64 at a time with one coalesced vector load (lane l holds
// record b0 + l) and the fields of particle b0 + i are read into SGPRs with v_readlane.
// Scalar loads would be simpler, but SMEM shares the lgkmcnt counter with LDS and
// returns out of order, so every wait on an LDS result would also wait for them.
struct LaneRec {
    float4 a;   // t[0], t[1], t[2], q
    int4 b;     // lo, orig, pad, pad
};

__device__ __forceinline__ LaneRec load_lane_rec(const Rec<float>* rec, int i)
{
    const float4* p = reinterpret_cast<const float4*>(rec + i);
    LaneRec r;
    r.a = p[0];
    r.b = *reinterpret_cast<const int4*>(p + 1);
    return r;
}

__device__ __forceinline__ Rec<float> pick_rec(const LaneRec& s, int l)
{
    Rec<float> r;
    r.t[0] = rlane(s.a.x, l);
    r.t[1] = rlane(s.a.y, l);
    r.t[2] = rlane(s.a.z, l);
    r.q = rlane(s.a.w, l);
    r.lo = __builtin_amdgcn_readlane(s.b.x, l);
    r.orig = __builtin_amdgcn_readlane(s.b.y, l);
    return r;
}

// Slot-rotated window row: lane s of the ring holds row rho == s (mod P) of the window
// starting at row y, i.e. window index (s - y) mod P; ymodp = y mod P (wave-uniform).
template <int P>
__device__ __forceinline__ int ring_index(int lane, int ymodp)
{
    const int j = lane - ymodp;
    return j < 0 ? j + P : j;
}

// v += DPP-permuted v (the permuted operand of lanes outside row_mask reads 0)
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

// Wave sum with DPP: quad swaps, half-row and row mirrors, then row broadcasts 15 and
// 31 (GFX9 DPP); lane 63 ends with the total, which is read into an SGPR.
__device__ __forceinline__ float wave_sum_dpp(float v)
{
    v = dpp_add<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
    v = dpp_add<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
    v = dpp_add<0x141, 0xF>(v);   // row_half_mirror
    v = dpp_add<0x140, 0xF>(v);   // row_mirror
    v = dpp_add<0x142, 0xA>(v);   // row_bcast:15 into rows 1, 3
    v = dpp_add<0x143, 0xC>(v);   // row_bcast:31 into rows 2, 3
    return rlane(v, 63);
}

// Lane-owned (i, k) pairs of the (P+1)^2 pencil block: offset i*S + k in an LDS row
// plane, and the ds_bpermute byte addresses (4*i, 4*k) that fetch the pair's x and z
// factors from the lanes that computed them.  Padding slots (pair index >= (P+1)^2,
// only in the last register) are mapped onto pairs of the first register, so every
// lane addresses a distinct, in-bounds location without a branch, and fetch their x
// factor from lane 63, whose factor is always 0 (63 >= P + 1).
template <int P, int R>
__device__ __forceinline__ void lane_pairs(int lane, int S, int (&pi)[R], int (&pk)[R], int (&pofs)[R],
                                           int (&bx)[R], int (&bz)[R])
{
    constexpr int W = P + 1;
    static_assert(W < 63, "lane 63 must lie outside the window");
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int pr0 = lane + 64 * r;
        const bool ok = pr0 < W * W;
        const int pr = ok ? pr0 : pr0 % (W * W);
        pi[r] = pr / W;
        pk[r] = pr - (pr / W) * W;
        pofs[r] = pi[r] * S + pk[r];
        bx[r] = 4 * (ok ? pi[r] : 63);
        bz[r] = 4 * pk[r];
    }
}

__device__ __forceinline__ float bperm(int byte_addr, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(byte_addr, __float_as_int(v)));
}

// acc += y[lane S of this 16-lane row] * x, as one v_fmac_f32_dpp with row_newbcast:S
// (GFX9.4 DPP: the broadcast source operand is read through the DPP crossbar, so the
// y-factor of ring slot S needs no v_readlane).  The y operand is read through DPP,
// which needs two wait states after the VALU write of that VGPR; FIRST places them in
// front of the first FMA of a block, and the blocks' asm is volatile so that no VALU
// write of y is scheduled between them.
template <int S, bool FIRST>
__device__ __forceinline__ void fmac_row_bcast(float& acc, float y, float x)
{
    if (FIRST)
        asm volatile("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(y), "v"(x), "n"(S));
    else
        asm volatile("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(y), "v"(x), "n"(S));
}

// Spread FMA block: acc[s][r] += Y_s * xz[r] for all ring slots s, Y_s = lane s of fy's row.
template <int S, int P, int R>
struct SpreadRing {
    static __device__ __forceinline__ void run(float (&acc)[P][R], float fy, const float (&xz)[R])
    {
        fmac_row_bcast<S, S == 0>(acc[S][0], fy, xz[0]);
#pragma unroll
        for (int r = 1; r < R; ++r) fmac_row_bcast<S, false>(acc[S][r], fy, xz[r]);
        SpreadRing<S + 1, P, R>::run(acc, fy, xz);
    }
};
template <int P, int R>
struct SpreadRing<P, P, R> {
    static __device__ __forceinline__ void run(float (&)[P][R], float, const float (&)[R]) {}
};

// Gather FMA block: A[r] += Y_s G[s][r] and (FORCE) B[r] += Y'_s G[s][r] over the slots.
template <int S, int P, int R, bool FORCE>
struct GatherRing {
    static __device__ __forceinline__ void run(float (&A)[R], float (&B)[R], float fy, float fyd,
                                               const float (&G)[P][R])
    {
        fmac_row_bcast<S, S == 0>(A[0], fy, G[S][0]);
#pragma unroll
        for (int r = 1; r < R; ++r) fmac_row_bcast<S, false>(A[r], fy, G[S][r]);
        if (FORCE) {
#pragma unroll
            for (int r = 0; r < R; ++r) fmac_row_bcast<S, false>(B[r], fyd, G[S][r]);
        }
        GatherRing<S + 1, P, R, FORCE>::run(A, B, fy, fyd, G);
    }
};
template <int P, int R, bool FORCE>
struct GatherRing<P, P, R, FORCE> {
    static __device__ __forceinline__ void run(float (&)[R], float (&)[R], float, float, const float (&)[P][R]) {}
};

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

    // wave index made provably uniform, so the particle range and all ring bookkeeping
    // (pencil, base row, flush tests) live in SGPRs
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    int pi[R], pk[R], pofs[R], bx[R], bz[R];
    lane_pairs<P, R>(lane, S, pi, pk, pofs, bx, bz);

    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    const int wb = beg + (int)(((long)n * wave) / nw), we = beg + (int)(((long)n * (wave + 1)) / nw);
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;

    float acc[P][R];
#pragma unroll
    for (int s = 0; s < P; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[s][r] = 0.f;

    int pen = -1, base = 0, pbase = 0;
    // add ring rows [base, upto) (upto <= base + P) into LDS and clear their slots; the
    // common case (no row leaves the window) costs one uniform branch
    auto flush = [&](int upto) {
        if (upto <= base) return;
        const unsigned mask = ring_mask<P>(base, upto - base);
        const int b = base % P;
#pragma unroll
        for (int s = 0; s < P; ++s) {
            if (mask & (1u << s)) {
                const int rho = base + (s >= b ? s - b : s - b + P);
                double* rowp = sub + rho * S2 + pbase;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    // padding slots hold 0 and add it to a valid, distinct location
                    __hip_atomic_fetch_add(rowp + pofs[r], (double)acc[s][r], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    acc[s][r] = 0.f;
                }
            }
        }
    };

    // Factors of one particle: the x and z factors of the lane's pairs (q*d folded into
    // x), fetched with ds_bpermute and multiplied only when the particle is processed,
    // and the slot-rotated y factor (lane s < P holds Y[(s - y) mod P]).
    const float kc = -c * 1.4426950408889634f;
    const float olane = (float)(g.first + lane) * h;   // grid offset of window index `lane`
    auto factors = [&](const Rec<float>& rc, float (&sx)[R], float (&sz)[R], float& fy) {
        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255, lz = rc.lo >> 16;
        const int dx = lx & 1, dz = lz & 1;
        // block column `lane` is window index lane - d (d = the particle's column in the pencil)
        const float ax = rc.t[0] - olane + (float)dx * h, az = rc.t[2] - olane + (float)dz * h;
        const bool vx = lane >= dx && lane < P + dx, vz = lane >= dz && lane < P + dz;
        // the exponentials run on every lane and the masks select afterwards, so no
        // exec-mask branch is formed around them
        const float fx = gauss(kc, ax) * (vx ? rc.q * dn : 0.f);
        const float fz = gauss(kc, az) * (vz ? 1.f : 0.f);
        // every 16-lane row holds the slot-rotated y factors (lanes s < P of the row)
        const float ay = rc.t[1] - (float)(g.first + ring_index<P>(lane & 15, ly % P)) * h;
        fy = gauss(kc, ay);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sx[r] = bperm(bx[r], fx);
            sz[r] = bperm(bz[r], fz);
        }
    };

    // Pipeline: the record of particle p+1 is picked and its factors (exponentials and
    // cross-lane fetches) are issued before the FMA block of particle p, and their LDS
    // results are first used one iteration later, so the fetch latency overlaps the FMAs.
    // The look-ahead at the last particle reads a clamped (valid) record and is unused.
    float sx_n[R], sz_n[R], fy_n = 0.f;
    int b0 = wb;
    LaneRec cur = load_lane_rec(rec, min(b0 + lane, max(we - 1, beg)));
    LaneRec nxt = load_lane_rec(rec, min(b0 + 64 + lane, max(we - 1, beg)));
    Rec<float> rc_n = pick_rec(cur, 0);
    factors(rc_n, sx_n, sz_n, fy_n);
    for (int p = wb; p < we; ++p) {
        if (p + 1 - b0 == 64) {   // the look-ahead record starts the next batch of 64
            b0 += 64;
            cur = nxt;
            nxt = load_lane_rec(rec, min(b0 + 64 + lane, we - 1));
        }
        const Rec<float> rc = rc_n;
        float xz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            xz[r] = sx_n[r] * sz_n[r];
            // form the products here, before the flush's LDS atomics are issued: the wait
            // for the fetches then does not also wait for those atomics
            asm volatile("" : "+v"(xz[r]));
        }
        const float fy = fy_n;

        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255, lz = rc.lo >> 16;
        const int npen = ((lx >> 1) << 12) | (lz >> 1);
        // one flush site (rows [base, upto) of the old window), so the accumulator ring
        // is never copied between control-flow paths
        int upto = base;
        if (npen != pen) {
            if (pen >= 0) upto = base + P;
        } else if (ly > base) {
            upto = min(ly, base + P);
        }
        flush(upto);
        if (npen != pen) {
            pen = npen;
            base = ly;
            pbase = ((lx >> 1) * 2) * S + (lz >> 1) * 2;
        } else if (ly > base) {
            base = ly;
        }
        rc_n = pick_rec(cur, p + 1 - b0);
        factors(rc_n, sx_n, sz_n, fy_n);
        float fyb = fy;
        asm volatile("" : "+v"(fyb));   // last definition of the DPP source before the block
        SpreadRing<0, P, R>::run(acc, fyb, xz);
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
                                                       double* __restrict__ phi, double* __restrict__ force,
                                                       const SysDesc* __restrict__ sys)
{
    constexpr int W = P + 1;
    constexpr int R = (W * W + 63) / 64;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* sub = reinterpret_cast<float*>(smem_raw);   // [y][x][z], S^3
    const int S = tl.S, S2 = S * S, S3 = S2 * S, M = g.M, Mz = g.Mz, NT = tl.NT, NTz = tl.NTz;
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
        for (int e = tid; e < S3; e += nthr) {
            const int iz = e % S, r2 = e / S, ix = r2 % S, iy = r2 / S;
            const int gx = (ox + ix) % M, gy = (oy + iy) % M, gz = (oz + iz) % Mz;
            sub[e] = grid[((long)gx * M + gy) * Mz + gz];
        }
    }
    __syncthreads();

    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    int pi[R], pk[R], pofs[R], bx[R], bz[R];
    lane_pairs<P, R>(lane, S, pi, pk, pofs, bx, bz);
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    const int wb = beg + (int)(((long)n * wave) / nw), we = beg + (int)(((long)n * (wave + 1)) / nw);
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;
    const double h3 = g.h * g.h * g.h;

    float G[P][R];
    int pen = -1, base = 0, pbase = 0;
    // read ring rows in [lo, hi) (all within [base, base + P)) from LDS; one uniform
    // branch when no row enters the window
    auto load = [&](int lo, int hi) {
        if (hi <= lo) return;
        const unsigned mask = ring_mask<P>(lo, hi - lo);
        const int b = base % P;
#pragma unroll
        for (int s = 0; s < P; ++s) {
            if (mask & (1u << s)) {
                const int rho = base + (s >= b ? s - b : s - b + P);
                const float* rowp = sub + rho * S2 + pbase;
#pragma unroll
                for (int r = 0; r < R; ++r) G[s][r] = rowp[pofs[r]];   // padding slots: weight 0
            }
        }
    };

    // Factors of one particle: the x and z factors of the lane's pairs (d folded into x),
    // fetched with ds_bpermute and multiplied only when the particle is processed, the
    // slot-rotated y factor and its derivative factor -c (t_y - o h) Y.
    const float kc = -c * 1.4426950408889634f;
    const float olane = (float)(g.first + lane) * h;   // grid offset of window index `lane`
    // c * (grid offset) of the lane's pair columns: the x/z derivative factor of a pair is
    // -c (t - o h) = c o h - c t
    float coxr[R], cozr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        coxr[r] = c * ((float)(g.first + pi[r]) * h);
        cozr[r] = c * ((float)(g.first + pk[r]) * h);
    }
    auto factors = [&](const Rec<float>& rc, float (&sx)[R], float (&sz)[R], float& fy, float& fyd) {
        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255, lz = rc.lo >> 16;
        const int dx = lx & 1, dz = lz & 1;
        const float ax = rc.t[0] - olane + (float)dx * h, az = rc.t[2] - olane + (float)dz * h;
        const bool vx = lane >= dx && lane < P + dx, vz = lane >= dz && lane < P + dz;
        const float fx = gauss(kc, ax) * (vx ? dn : 0.f);
        const float fz = gauss(kc, az) * (vz ? 1.f : 0.f);
        // every 16-lane row holds the slot-rotated y factors (lanes s < P of the row)
        const float ay = rc.t[1] - (float)(g.first + ring_index<P>(lane & 15, ly % P)) * h;
        fy = gauss(kc, ay);
        fyd = fy * (-c * ay);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sx[r] = bperm(bx[r], fx);
            sz[r] = bperm(bz[r], fz);
        }
    };
    // wave sums of the previous particle's lane partials, then its outputs
    auto finish = [&](float sp, float s0, float s1, float s2, int o, float q) {
        if (POT) sp = wave_sum_dpp(sp);
        if (FORCE) {
            s0 = wave_sum_dpp(s0);
            s1 = wave_sum_dpp(s1);
            s2 = wave_sum_dpp(s2);
        }
        if (lane == 0) {
            if (POT) store_out(phi, o, h3 * (double)sp);
            if (FORCE) {
                const double qh = h3 * (double)q;
                store_out(force, o, qh * (double)s0);
                store_out(force, o + (long)N, qh * (double)s1);
                store_out(force, o + 2l * N, qh * (double)s2);
            }
        }
    };

    // Pipeline: the record of particle p+1 is picked and its factors are issued before the
    // FMA block of particle p (their LDS results are first used one iteration later), and
    // the wave reductions of particle p-1 run after it.  The look-ahead at the last
    // particle reads a clamped (valid) record and is unused.
    float sx_n[R], sz_n[R], fy_n = 0.f, fyd_n = 0.f;
    int b0 = wb;
    LaneRec cur = load_lane_rec(rec, min(b0 + lane, max(we - 1, beg)));
    LaneRec nxt = load_lane_rec(rec, min(b0 + 64 + lane, max(we - 1, beg)));
    Rec<float> rc_n = pick_rec(cur, 0);
    factors(rc_n, sx_n, sz_n, fy_n, fyd_n);
    float psp = 0.f, ps0 = 0.f, ps1 = 0.f, ps2 = 0.f, pq = 0.f;
    int porig = -1;
    for (int p = wb; p < we; ++p) {
        if (p + 1 - b0 == 64) {   // the look-ahead record starts the next batch of 64
            b0 += 64;
            cur = nxt;
            nxt = load_lane_rec(rec, min(b0 + 64 + lane, we - 1));
        }
        const Rec<float> rc = rc_n;
        float xz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) xz[r] = sx_n[r] * sz_n[r];
        const float fy = fy_n, fyd = fyd_n;

        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255, lz = rc.lo >> 16;
        const int npen = ((lx >> 1) << 12) | (lz >> 1);
        // one load site (rows [lo, hi) of the new window) for the same reason as the
        // spread's single flush site
        int lo = 0, hi = 0;
        if (npen != pen) {
            pen = npen;
            pbase = ((lx >> 1) * 2) * S + (lz >> 1) * 2;
            lo = ly;
            hi = ly + P;
            base = ly;
        } else if (ly > base) {
            lo = max(base + P, ly);
            hi = ly + P;
            base = ly;
        }
        load(lo, hi);
        rc_n = pick_rec(cur, p + 1 - b0);
        factors(rc_n, sx_n, sz_n, fy_n, fyd_n);
        float A[R], B[R];
#pragma unroll
        for (int r = 0; r < R; ++r) A[r] = B[r] = 0.f;
        float fyb = fy, fydb = fyd;
        asm volatile("" : "+v"(fyb), "+v"(fydb));   // last definitions of the DPP sources
        GatherRing<0, P, R, FORCE>::run(A, B, fyb, fydb, G);
        if (porig >= 0) finish(psp, ps0, ps1, ps2, porig, pq);
        // lane partials of particle p: sum xz A, and for the x/z force sum xz A (c o h - c t)
        // (the difference is formed per pair: subtracting c t sum xz A once per particle
        // instead cancels about one more bit)
        const float ctx = c * (rc.t[0] + (float)(lx & 1) * h), ctz = c * (rc.t[2] + (float)(lz & 1) * h);
        float sp = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float w = xz[r] * A[r];
            sp += w;
            if (FORCE) {
                s0 = fmaf(w, coxr[r] - ctx, s0);
                s1 = fmaf(xz[r], B[r], s1);
                s2 = fmaf(w, cozr[r] - ctz, s2);
            }
        }
        psp = sp;
        ps0 = s0;
        ps1 = s1;
        ps2 = s2;
        porig = rc.orig;
        pq = rc.q;
    }
    if (porig >= 0) finish(psp, ps0, ps1, ps2, porig, pq);
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
                                const Tiling& tl, int N, double* phi, double* force, const SysDesc* sys,
                                hipStream_t s)
{
    const size_t lds = (size_t)tl.S * tl.S * tl.S * sizeof(float);
    dim3 gr(tl.max_items), bl(tl.block);
#define SE_GP(POT, FORCE)                                                                                     \
    do {                                                                                                      \
        allow_lds_p((const void*)k_interp_pencil<PV, POT, FORCE>, lds);                                       \
        hipLaunchKernelGGL((k_interp_pencil<PV, POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg,       \
                           item_end, n_items, grid, g, tl, N, phi, force, sys);                               \
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
                          double* phi, double* force, const SysDesc* sys, hipStream_t s)
{
#define SE_GPM(PV) interp_pencil_modes<PV>(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, sys, s)
    switch (g.P) {
    case 8: SE_GPM(8); return true;
    case 11: SE_GPM(11); return true;
    case 16: SE_GPM(16); return true;
    default: return false;
    }
#undef SE_GPM
}

}  // namespace se

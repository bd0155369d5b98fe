// Register-ring "pencil" kernels: the FP32 fast path of the fused engine.
//
// Particles are sorted (tile, pencil, row): a pencil is a group of 4 x 1 window-start
// columns (4 in x, 1 in z) inside a tile, walked along y.  All particles of a pencil
// touch the same (P+3) x P block of (x, z) grid columns, so a wave keeps that block in
// registers: lane l owns the (i, k) pairs l, l+64, ... (R = ceil((P+3)P/64) per lane)
// for P consecutive grid rows -- a ring of P rows.  Slot s of the ring holds
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

#include <algorithm>

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

// Waits for a batch's loads at this point: an inline-asm use of every loaded register makes
// the compiler place the vmcnt wait here.  Without it the wait lands where the batch is
// next read, which in a loop with a conditional reload is the top of every iteration, and
// vmcnt also counts the stores issued since, so each iteration would wait for the
// previous iteration's output stores to complete.
__device__ __forceinline__ void settle(const LaneRec& r)
{
    asm volatile("" ::"v"(r.a.x), "v"(r.a.y), "v"(r.a.z), "v"(r.a.w), "v"(r.b.x), "v"(r.b.y));
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

// Lane-owned (i, k) pairs of the (P+3) x P pencil block: offset i*S + k in an LDS row
// plane, and the ds_bpermute byte addresses (4*i, 4*k) that fetch the pair's x and z
// factors from the lanes that computed them.  Padding slots (pair index >= (P+3)P,
// only in the last register) are mapped onto pairs of the first register, so every
// lane addresses a distinct, in-bounds location without a branch, and fetch their x
// factor from lane 63, whose factor is always 0 (63 >= P + 3).
// Pair order: XFAST puts x fastest (pair = i + (P+3) k), so the 64 lanes of a row access
// step through x columns S apart; for the P = 16 gather with S = 27 (tile 12) the FP32
// addresses of a register then fall in distinct LDS banks (z fastest puts four 16-wide
// runs 27 apart, two to a bank).  The spread and the other supports keep z fastest (the
// spread's FP64 flush has another bank pattern, and x fastest costs it registers).
template <int P, int R, bool XFAST = false>
__device__ __forceinline__ void lane_pairs(int lane, int S, int (&pi)[R], int (&pk)[R], int (&pofs)[R],
                                           int (&bx)[R], int (&bz)[R])
{
    constexpr int WX = P + 3, WZ = P;
    static_assert(WX < 63, "lane 63 must lie outside the window");
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int pr0 = lane + 64 * r;
        const bool ok = pr0 < WX * WZ;
        const int pr = ok ? pr0 : pr0 % (WX * WZ);
        pi[r] = XFAST ? pr % WX : pr / WZ;
        pk[r] = XFAST ? pr / WX : pr % WZ;
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

// dst[k] = lane k of v's 16-lane row, for k = 0..15, in every lane (v_mov_b32_dpp
// row_newbcast); the leading s_nop gives the two wait states a DPP read needs after a
// VALU write of its source.
__device__ __forceinline__ void row_bcast16(float v, float (&d)[16])
{
    asm volatile(
        "s_nop 1\n\t"
        "v_mov_b32_dpp %0, %16 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %1, %16 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %2, %16 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %3, %16 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %4, %16 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %5, %16 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %6, %16 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %7, %16 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %8, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %9, %16 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %10, %16 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %11, %16 row_newbcast:11 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %12, %16 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %13, %16 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %14, %16 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b32_dpp %15, %16 row_newbcast:15 row_mask:0xf bank_mask:0xf"
        : "=&v"(d[0]), "=&v"(d[1]), "=&v"(d[2]), "=&v"(d[3]), "=&v"(d[4]), "=&v"(d[5]), "=&v"(d[6]), "=&v"(d[7]),
          "=&v"(d[8]), "=&v"(d[9]), "=&v"(d[10]), "=&v"(d[11]), "=&v"(d[12]), "=&v"(d[13]), "=&v"(d[14]),
          "=&v"(d[15])
        : "v"(v));
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

// A sparse item: a whole tile (its range is every record of the tile, so it was not cut
// into chunks) of fewer than sparse_max particles.
__device__ __forceinline__ bool sparse_item(const int* __restrict__ item_tile, const int* __restrict__ item_beg,
                                            const int* __restrict__ item_end, const int* __restrict__ tile_start,
                                            int item, int sparse_max)
{
    if (sparse_max <= 0) return false;
    const int beg = item_beg[item], end = item_end[item];
    if (end - beg >= sparse_max) return false;
    const int t = item_tile[item];
    return beg == tile_start[t] && end == tile_start[t + 1];
}

template <int P>
__global__ __launch_bounds__(spread_pencil_max_threads(P)) void k_spread_pencil(const Rec<float>* __restrict__ rec,
                                                       const int* __restrict__ item_beg,
                                                       const int* __restrict__ item_end,
                                                       const int* __restrict__ n_items, Geom g, Tiling tl,
                                                       float* __restrict__ partials, const int* __restrict__ item_tile,
                                                       const int* __restrict__ tile_start, int sparse_max,
                                                       DenseSel dense)
{
    constexpr int R = ((P + 3) * P + 63) / 64;
    const int item = blockIdx.x;
    if (item >= *n_items) return;   // uniform across the work-group
    // sparse items (whole tiles of fewer than sparse_max particles) are the plane
    // spread's (sparse_max > 0), the column items the column spread's
    if (sparse_item(item_tile, item_beg, item_end, tile_start, item, sparse_max)) return;
    if (dense(item, item_end[item] - item_beg[item])) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ float wmax[spread_pencil_max_threads(P) / 64];
    double* sub = reinterpret_cast<double*>(smem_raw);   // [y][x][z], S^3
    const int S = tl.S, S2 = S * S, S3 = S2 * S;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    for (int e = tid; e < S3; e += nthr) sub[e] = 0.0;
    item_qmax_part(rec, beg, end, tid, nthr, wmax);
    __syncthreads();

    // wave index made provably uniform, so the particle range and all ring bookkeeping
    // (pencil, base row, flush tests) live in SGPRs
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    int pi[R], pk[R], pofs[R], bx[R], bz[R];
    lane_pairs<P, R>(lane, S, pi, pk, pofs, bx, bz);

    const int wb = beg + (int)(((long)n * wave) / nw), we = beg + (int)(((long)n * (wave + 1)) / nw);
    // fixed-point scale of the item's LDS sums (flush_scale): q d 2^sc is folded into the
    // x factor, each flushed value is rounded to an integer, the write-out scales back
    const int sc = __builtin_amdgcn_readfirstlane(flush_scale((double)n * item_qmax_read(wmax, nw) * g.d, g.d));
    const float h = (float)g.h, c = (float)g.c;
    const float dn = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)ldexp(g.d, sc))));

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
                    // padding slots hold 0 and add it to a valid, distinct location; the
                    // value is rounded to an integer (in FP32: a float of 2^23 or more is
                    // one already), so the FP64 sums are exact
                    __hip_atomic_fetch_add(rowp + pofs[r], (double)rintf(acc[s][r]), __ATOMIC_RELAXED,
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
        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255;
        const int dx = lx & 3;
        // block column `lane` is window index lane - dx in x (dx = the particle's column in
        // the pencil) and window index lane in z
        const float ax = rc.t[0] - olane + (float)dx * h, az = rc.t[2] - olane;
        const bool vx = lane >= dx && lane < P + dx, vz = lane < P;
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
        const int npen = ((lx >> 2) << 12) | lz;
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
            pbase = ((lx >> 2) * 4) * S + lz;
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
    SubWalk w(S, tid, nthr);   // canonical [x][y][z]: a = ix, b = iy, c = iz
    for (int e = tid; e < tl.part_elems; e += nthr) {
        out[e] = (float)ldexp(sub[(w.b * S + w.a) * S + w.c], -sc);
        w.next();
    }
}

// ---------------------------------------------------------------------------------
// Plane spread for sparse work items (a whole tile of fewer particles than the
// threshold).  A 4 x 1 pencil of a sparse tile holds one or two particles, so the register
// ring above flushes its P rows (convert, round, clear and an FP64 LDS atomic per slot) for
// about one particle each, and the zeroing and write-out of the FP64 sub-grid cost as much
// again.  Here lane (X, Z) of the work-group owns one (x, z) column of the S^3 sub-grid and
// keeps all S of its y values in registers: wave w holds the columns X = 2w, 2w + 1 and
// Z = 0 .. 31 (S <= 32).  Each particle adds w = q d Fx(X) Fz(Z) times its y factor to
// every row of the column: the y factors are tabulated per particle at rows 0 .. 31 with
// zeros outside the window [ly, ly + P), so every register index is static and the FMA
// block has no branch (S FMAs per lane; the factor of row y is read from lane y mod 16 of
// the table's first or second 16 entries with row_newbcast).  Choosing only the 4-row blocks
// that meet the window saves FMAs but measured slower (the branches cost more).  No LDS
// sub-grid, no atomic and no flush is needed, and the columns are stored straight from
// registers into the item's partial.  Waves whose two x columns miss the particle's window
// skip it.  The x and z factors are tabulated too (with zeros around the window, so a lane
// reads its factor without a bounds test), per batch of 128 particles in LDS.  (Reading the
// next particle's factors before the current particle's FMAs measured slower: the copies
// cost more VALU than the waits, which the other waves cover.)  Every sum is formed in
// FP32 in the item's record order, so the result does not depend on timing.
// The kernel is held to 64 VGPRs (8 waves per SIMD) so that two work-groups of 14 waves
// share a CU.
// ---------------------------------------------------------------------------------
constexpr int PLANE_BATCH = 128;  // particles tabulated per batch
constexpr int PLANE_TW = 48;      // x / z table row: window index j at 16 + j, zeros around

// acc[Y] += T_Y * w for Y < SM, T_Y = lane Y mod 16 of fya's (Y < 16) or fyb's (Y >= 16) row
template <int Y, int SM>
struct PlaneColumn {
    static __device__ __forceinline__ void run(float (&acc)[SM], float fya, float fyb, float w)
    {
        if (Y < 16)
            fmac_row_bcast<Y & 15, Y == 0>(acc[Y], fya, w);
        else
            fmac_row_bcast<Y & 15, Y == 16>(acc[Y], fyb, w);
        PlaneColumn<Y + 1, SM>::run(acc, fya, fyb, w);
    }
};
template <int SM>
struct PlaneColumn<SM, SM> {
    static __device__ __forceinline__ void run(float (&)[SM], float, float, float) {}
};

template <int P>
__global__ __launch_bounds__(1024, 8) void k_spread_plane(const Rec<float>* __restrict__ rec,
                                                          const int* __restrict__ item_tile,
                                                          const int* __restrict__ item_beg,
                                                          const int* __restrict__ item_end,
                                                          const int* __restrict__ n_items,
                                                          const int* __restrict__ tile_start, Geom g, Tiling tl,
                                                          float* __restrict__ partials, int sparse_max,
                                                          DenseSel dense)
{
    constexpr int SM = 12 + P - 1;   // largest sub-grid side (tile side <= 12)
    static_assert(SM <= 32 && P <= 16, "the y table holds 32 rows, the x and z tables 16 + 16 + 16");
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    if (!sparse_item(item_tile, item_beg, item_end, tile_start, item, sparse_max)) return;
    if (dense(item, item_end[item] - item_beg[item])) return;   // a column item (the column spread's)
    __shared__ float tx[PLANE_BATCH][PLANE_TW], tz[PLANE_BATCH][PLANE_TW], ty[PLANE_BATCH][32];
    const int beg = item_beg[item], end = item_end[item];
    const int S = tl.S, tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int xlo = 2 * wave;                          // the wave's x columns: xlo, xlo + 1
    const int X = xlo + (lane >> 5), Z = lane & 31;    // this lane's column
    // the x and z tables keep zeros outside the window entries [16, 16 + P)
    for (int e = tid; e < PLANE_BATCH * PLANE_TW; e += nthr) {
        (&tx[0][0])[e] = 0.f;
        (&tz[0][0])[e] = 0.f;
    }

    float acc[SM];
#pragma unroll
    for (int y = 0; y < SM; ++y) acc[y] = 0.f;
    const float h = (float)g.h, kc = -(float)g.c * 1.4426950408889634f, dq = (float)g.d;
    // byte offsets of this lane's entries for a window starting at 0 (X - lx + 16 and
    // Z - lz + 16 stay in [16 - 11, 31 + 16] = [5, 47]); the particle's part is uniform
    const int xo = (X + 16) * 4, zo = (Z + 16) * 4, yo = (lane & 15) * 4;
    const char* txb = reinterpret_cast<const char*>(&tx[0][0]);
    const char* tzb = reinterpret_cast<const char*>(&tz[0][0]);
    const char* tyb = reinterpret_cast<const char*>(&ty[0][0]);
    for (int b0 = beg; b0 < end; b0 += PLANE_BATCH) {
        const int cnt = min(PLANE_BATCH, end - b0);
        __syncthreads();   // the tables are cleared / the previous batch is consumed
        // entry j < 16 of particle p: window index j of the x, y and z factors; the y
        // factor goes to row ly + j, and row (ly + 16 + j) mod 32 is cleared, so the 32 rows
        // hold the window and zeros
        for (int e = tid; e < cnt * 16; e += nthr) {
            const int p = e >> 4, j = e & 15;
            const float4 r = *reinterpret_cast<const float4*>(rec + b0 + p);   // t[0], t[1], t[2], q
            const int ly = (rec[b0 + p].lo >> 8) & 255;
            const float o = (float)(g.first + j) * h;
            const bool in = j < P;
            ty[p][ly + j] = in ? gauss(kc, r.y - o) : 0.f;
            ty[p][(ly + 16 + j) & 31] = 0.f;
            if (in) {
                tx[p][16 + j] = gauss(kc, r.x - o) * (r.w * dq);
                tz[p][16 + j] = gauss(kc, r.z - o);
            }
        }
        // lane l: window starts of particles b0 + l and b0 + 64 + l
        const int lo_a = rec[b0 + min(lane, cnt - 1)].lo, lo_b = rec[b0 + min(64 + lane, cnt - 1)].lo;
        __syncthreads();
        for (int p = 0; p < cnt; ++p) {
            const int la = __builtin_amdgcn_readlane(lo_a, p & 63), lb = __builtin_amdgcn_readlane(lo_b, p & 63);
            const int lo = p < 64 ? la : lb;
            const int lx = lo & 255;
            if (xlo + 1 < lx || xlo >= lx + P) continue;   // both columns outside the x window
            // uniform byte offsets of the particle's table rows, shifted by the window start
            const int sx = __builtin_amdgcn_readfirstlane((p * PLANE_TW - lx) * 4);
            const int sz = __builtin_amdgcn_readfirstlane((p * PLANE_TW - (lo >> 16)) * 4);
            const int sy = __builtin_amdgcn_readfirstlane(p * 128);
            const float w = *reinterpret_cast<const float*>(txb + (xo + sx)) *
                            *reinterpret_cast<const float*>(tzb + (zo + sz));
            float fya = *reinterpret_cast<const float*>(tyb + (yo + sy));
            float fyb = *reinterpret_cast<const float*>(tyb + (yo + sy + 64));
            asm volatile("" : "+v"(fya), "+v"(fyb));   // last definitions of the DPP sources
            PlaneColumn<0, SM>::run(acc, fya, fyb, w);
        }
    }
    if (X < S && Z < S) {   // canonical [x][y][z] partial
        float* out = partials + (size_t)item * tl.part_elems + (size_t)X * S * S + Z;
#pragma unroll
        for (int y = 0; y < SM; ++y)
            if (y < S) out[y * S] = acc[y];
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
                                                       const SysDesc* __restrict__ sys, DenseSel dense)
{
    constexpr int R = ((P + 3) * P + 63) / 64;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    // the column items are the matrix-core gather's
    if (dense(item, n)) return;
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
        load_tile_dma<true>(sub, grid, S, M, Mz, ox, oy, oz, tid, nthr);
    }
    __syncthreads();

    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    int pi[R], pk[R], pofs[R], bx[R], bz[R];
    lane_pairs<P, R, P == 16>(lane, S, pi, pk, pofs, bx, bz);
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
        const int lx = rc.lo & 255, ly = (rc.lo >> 8) & 255;
        const int dx = lx & 3;
        const float ax = rc.t[0] - olane + (float)dx * h, az = rc.t[2] - olane;
        const bool vx = lane >= dx && lane < P + dx, vz = lane < P;
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
        const int npen = ((lx >> 2) << 12) | lz;
        // one load site (rows [lo, hi) of the new window) for the same reason as the
        // spread's single flush site
        int lo = 0, hi = 0;
        if (npen != pen) {
            pen = npen;
            pbase = ((lx >> 2) * 4) * S + lz;
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
        const float ctx = c * (rc.t[0] + (float)(lx & 3) * h), ctz = c * rc.t[2];
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

// ---------------------------------------------------------------------------------
// MFMA gather for P = 16.  The y contraction of a pencil's (P+3) x P = 19 x 16 column
// block runs on v_mfma_f32_4x4x1_16b_f32 (16 independent 4x4x1 outer products per
// instruction; operand and result layouts: A and B one value per lane, lane 4b + i is
// row/column i of block b; D register i, lane 4b + j holds D_b[i][j]).  Block b is z
// column b of the block; A-row i of ring register r is x column i + 4r (r < 5: x < 20,
// 19 used).  A group of up to 4 consecutive particles of one pencil whose window rows
// all lie in a 20-row ring [y0, y0 + 20) is contracted in one pass over the 20 ring rows:
// column j of B holds particle j's y factor (lane 4b + j), so
//   D_r[i] (lane 4b + j) = sum_y G[y][x = i + 4r][z = b] * Y_j(y)
// and a second set of accumulators takes Y_j(y) (t_y - o_y h) for the y force.  The
// epilogue multiplies by the particle's x and z factors per lane and sums over the 16
// blocks.  The ring is the pencil kernels' ring with 20 rows in A layout; particles come
// in (tile, pencil, row) order, so y0 never decreases inside a pencil.
//
// The y and x factors are evaluated for every ring row and every x column of the block,
// also the (at most 4) rows and (3) columns outside a particle's P-point window: those
// lie at least 8.5 h from the particle, where the Gaussian is below e^{-c 72 h^2}
// (1e-10 relative at the P = 16 shape), far below FP32 resolution, so no masks are
// needed.  A y factor reaches all 16 blocks through the MFMA's B lane-group broadcast
// (each lane evaluates 5 of them); an x factor depends only on the particle slot j and
// the column, so each is evaluated once per group (on lane 4 c + j) and fetched across
// the blocks with ds_bpermute, issued ahead of the group's MFMAs.
// ---------------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

// Sum over the 16 lanes 4b + j (b = 0..15) that share j = lane & 3: DPP row rotations by
// 4 and 8 inside each 16-lane row, then row swaps between rows 0/1, 2/3 and between the
// two halves (gfx950 v_permlane16_swap / v_permlane32_swap); every lane gets the total.
__device__ __forceinline__ float sum_over_blocks(float v)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xF, 0xF, false));   // row_ror:4
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false));   // row_ror:8
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// D += A B with the B operand of 16-lane group g (lanes 16g .. 16g + 15) broadcast to all
// four groups (BLGP = 4 + g); g is a constant after unrolling.
__device__ __forceinline__ v4f mfma4_bcast(float a, float b, v4f c, int g)
{
    switch (g) {
    case 0: return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 4);
    case 1: return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 5);
    case 2: return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 6);
    default: return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 7);
    }
}

// Per-lane data of one group (this lane's particle: slot lane & 3 of the group).
struct GroupLane {
    float tyl;    // t_y - first h + ly h: the y offset of row rho is tyl - rho h
    float txf;    // t_x - first h: the x offset of window column w is txf - w h
    float tzf;    // t_z - first h
    float oxf;    // the particle's column in the pencil (x window index = x column - oxf)
    float q;
    int orig;
    bool act;     // the slot holds a particle of the group
};

template <bool POT, bool FORCE>
__global__ __launch_bounds__(512) void k_interp_mfma16(const Rec<float>* __restrict__ rec,
                                                       const int* __restrict__ item_tile,
                                                       const int* __restrict__ item_beg,
                                                       const int* __restrict__ item_end,
                                                       const int* __restrict__ n_items,
                                                       const float* __restrict__ grid, Geom g, Tiling tl, int N,
                                                       double* __restrict__ phi, double* __restrict__ force,
                                                       const SysDesc* __restrict__ sys, DenseSel dense)
{
    constexpr int P = 16, GY = 20, R = 5;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    // the column items only; the others are the row or vector gather's
    if (!dense(item, n)) return;
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
        load_tile_dma<true>(sub, grid, S, M, Mz, ox, oy, oz, tid, nthr);
    }
    __syncthreads();

    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nthr >> 6;
    const int blk = lane >> 2, q4 = lane & 3;   // block = z column; A-row = particle slot
    const int lgrp = lane >> 4;                  // 16-lane group
    const int wb = beg + (int)(((long)n * wave) / nw), we = beg + (int)(((long)n * (wave + 1)) / nw);
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;
    const float kc = -c * 1.4426950408889634f;
    const float fh = (float)g.first * h;
    const double h3 = g.h * g.h * g.h;

    // A element of ring register r in this lane: x column q4 + 4r, z column blk.  The
    // column x = 19 (past the P + 3 = 19 columns of the block) reads column 0 instead: its
    // x factor is below 1e-10 for every particle of the pencil, as is the y factor of
    // ring rows past the sub-grid, which read its last row; every value stays finite.
    int aofs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int px = q4 + 4 * r;
        aofs[r] = (px < P + 3 ? px : 0) * S + blk;
    }

    float G[GY][R];
    int pen = -1, base = 0, pbase = 0;
    // ring rows in [lo, hi) (all within [base, base + GY))
    auto load = [&](int lo, int hi) {
        if (hi <= lo) return;
        const unsigned mask = ring_mask<GY>(lo, hi - lo);
        const int bm = base % GY;
#pragma unroll
        for (int s = 0; s < GY; ++s) {
            if (mask & (1u << s)) {
                const int rho = min(base + (s >= bm ? s - bm : s - bm + GY), S - 1);
                const float* rowp = sub + rho * S2 + pbase;
#pragma unroll
                for (int r = 0; r < R; ++r) G[s][r] = rowp[aofs[r]];
            }
        }
    };

    int b0 = wb;
    LaneRec batch = load_lane_rec(rec, min(b0 + lane, max(we - 1, beg)));
    settle(batch);
    // the next group from particle p0 (< we): up to 4 particles of p0's pencil whose rows lie
    // in [y0, y0 + GY - P]; moves the ring and returns this lane's particle data
    auto next_group = [&](int p0, int& k) -> GroupLane {
        if (p0 + 4 > b0 + 64) {   // records p0 .. p0 + 3 must lie in the batch
            b0 = p0;
            batch = load_lane_rec(rec, min(b0 + lane, we - 1));
            settle(batch);
        }
        const int lo0 = __builtin_amdgcn_readlane(batch.b.x, p0 - b0);
        const int lx0 = lo0 & 255, ly0 = (lo0 >> 8) & 255, lz0 = lo0 >> 16;
        const int npen = ((lx0 >> 2) << 12) | lz0;
        k = 1;
#pragma unroll
        for (int jj = 1; jj < 4; ++jj) {
            if (k < jj || p0 + jj >= we) break;
            const int loj = __builtin_amdgcn_readlane(batch.b.x, p0 - b0 + jj);
            if (((((loj & 255) >> 2) << 12) | (loj >> 16)) != npen || ((loj >> 8) & 255) > ly0 + (GY - P)) break;
            k = jj + 1;
        }
        int lo = 0, hi = 0;
        if (npen != pen) {
            pen = npen;
            pbase = ((lx0 >> 2) * 4) * S + lz0;
            base = ly0;
            lo = ly0;
            hi = ly0 + GY;
        } else if (ly0 > base) {
            lo = max(base + GY, ly0);
            hi = ly0 + GY;
            base = ly0;
        }
        load(lo, hi);
        // slots past k repeat the group's last particle and store nothing
        const int sa = 4 * (p0 - b0 + min(q4, k - 1));
        const int loj = __builtin_amdgcn_ds_bpermute(sa, batch.b.x);
        GroupLane d;
        d.tyl = bperm(sa, batch.a.y) - fh + (float)((loj >> 8) & 255) * h;
        d.txf = bperm(sa, batch.a.x) - fh;
        d.tzf = bperm(sa, batch.a.z) - fh;
        d.oxf = (float)(loj & 3);
        d.q = bperm(sa, batch.a.w);
        d.orig = __builtin_amdgcn_ds_bpermute(sa, batch.b.y);
        d.act = q4 < k;
        return d;
    };

    // y contraction of the group in `d` over the 20 ring rows into (Dp, Df).  The B
    // operand of ring slot s = 4u + v is register u of the y factors with lane group v
    // broadcast to all four 16-lane groups (MFMA BLGP = 4 + v): lane 16v + 4c + j of
    // register u holds particle j's y factor at the row of slot 4u + v (the same value for
    // c = 0..3), so every block sees Y_j of that row in column j.  Each lane evaluates 5
    // exponentials and no cross-lane fetch sits between the MFMAs.
    auto contract = [&](const GroupLane& d, v4f (&Dp)[R], v4f (&Df)[R]) {
        const v4f zero4 = {0.f, 0.f, 0.f, 0.f};
        const int bm = base % GY;
        float yv[GY / 4], yb[GY / 4];
#pragma unroll
        for (int u = 0; u < GY / 4; ++u) {
            const int t = 4 * u + lgrp - bm;
            const int rho = base + (t < 0 ? t + GY : t);
            const float ay = fmaf(-(float)rho, h, d.tyl);
            yv[u] = __builtin_amdgcn_exp2f(kc * (ay * ay));
            if (FORCE) yb[u] = yv[u] * ay;
        }
#pragma unroll
        for (int s = 0; s < GY; ++s) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                Dp[r] = mfma4_bcast(G[s][r], yv[s >> 2], s == 0 ? zero4 : Dp[r], s & 3);
                if (FORCE) Df[r] = mfma4_bcast(G[s][r], yb[s >> 2], s == 0 ? zero4 : Df[r], s & 3);
            }
        }
    };

    // x factors of (j = l & 3, column l >> 2) and (j, column 16 + (l >> 2)) on lane l,
    // fetched per column like a table (source lane 4 c + j); issued before the group's
    // MFMAs so that the fetch latency hides behind them
    auto xfactors = [&](const GroupLane& d, float (&xv)[4 * R], float (&xa)[4 * R]) {
        const float ax1 = fmaf(d.oxf - (float)blk, h, d.txf), ax2 = fmaf(d.oxf - (float)(16 + blk), h, d.txf);
        const float x1 = __builtin_amdgcn_exp2f(kc * (ax1 * ax1)), x2 = __builtin_amdgcn_exp2f(kc * (ax2 * ax2));
        const float xa1 = x1 * ax1, xa2 = x2 * ax2;
        const int src = 4 * q4;
#pragma unroll
        for (int c4 = 0; c4 < 4 * R; ++c4) {
            xv[c4] = bperm(src + 16 * (c4 & 15), c4 < 16 ? x1 : x2);
            if (FORCE) xa[c4] = bperm(src + 16 * (c4 & 15), c4 < 16 ? xa1 : xa2);
        }
    };

    // x and z factors of this lane's columns, sums over the blocks, outputs of the group in d
    auto epilogue = [&](const GroupLane& d, const v4f (&Dp)[R], const v4f (&Df)[R], const float (&xv)[4 * R],
                        const float (&xa)[4 * R]) {
        float S0 = 0.f, S1 = 0.f, S2 = 0.f;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                S0 = fmaf(xv[i + 4 * r], Dp[r][i], S0);
                if (FORCE) {
                    S1 = fmaf(xa[i + 4 * r], Dp[r][i], S1);
                    S2 = fmaf(xv[i + 4 * r], Df[r][i], S2);
                }
            }
        const float az = fmaf(-(float)blk, h, d.tzf);
        const float zf = dn * __builtin_amdgcn_exp2f(kc * (az * az));
        float sp = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
        if (POT) sp = sum_over_blocks(zf * S0);
        if (FORCE) {
            s0 = sum_over_blocks(zf * S1);
            s1 = sum_over_blocks(zf * S2);
            s2 = sum_over_blocks((zf * az) * S0);
        }
        if (d.act && blk == 0) {
            if (POT) store_out(phi, d.orig, h3 * (double)sp);
            if (FORCE) {
                // derivative factors -c (t - o h) of x, y and z
                const double qh = -(double)c * h3 * (double)d.q;
                store_out(force, d.orig, qh * (double)s0);
                store_out(force, d.orig + (long)N, qh * (double)s1);
                store_out(force, d.orig + 2l * N, qh * (double)s2);
            }
        }
    };

    // One group ahead: the next group's ring rows and record fields are fetched after the
    // current group's MFMAs have read the ring and before its epilogue, whose work then
    // covers the LDS latency of those fetches.
    v4f Dp[R], Df[R];
    float xv[4 * R], xa[4 * R];
    int p0 = wb, k = 0;
    GroupLane d{};
    if (p0 < we) d = next_group(p0, k);
    while (p0 < we) {
        xfactors(d, xv, xa);
        __builtin_amdgcn_sched_barrier(0);   // the fetches stay ahead of the MFMA block
        contract(d, Dp, Df);
        const int pn = p0 + k;
        int kn = 0;
        GroupLane dn{};
        if (pn < we) dn = next_group(pn, kn);
        epilogue(d, Dp, Df, xv, xa);
        d = dn;
        p0 = pn;
        k = kn;
    }
}

// ---------------------------------------------------------------------------------
// Row gather for sparse work items at P = 16 (items of fewer particles than the
// matrix-core gather's threshold).  In a sparse tile a pencil holds one or two
// particles, so the vector gather above loads all P ring rows (P R LDS reads and their
// addresses), fetches the x and z factors across lanes and reduces four sums over the 64
// lanes for each particle.  Here one 16-lane DPP row takes one particle: lane z of the row
// owns the window's z point z and walks the whole 16 x 16 (x, y) window, reading the
// tile's LDS sub-grid at offsets fixed at compile time (the sub-grid side SS is a template
// parameter), and sums in registers
//   c0 = sum_y Fy G,  c1 = sum_y Fy' G                  (2 FMAs per (x, y))
//   u = sum_x Fx c0,  ux = sum_x Fx' c0,  uy = sum_x Fx c1
// with F' = -c (t - o h) F the derivative factors (d folded into Fx; lane j of the row
// evaluates window index j and the row's lanes share them by DPP broadcasts).  The z contraction
// comes last: phi = sum_z Fz u, F = sum_z (Fz ux, Fz uy, Fz' u), a four-step DPP sum over
// the row.  Each step of a half-wave reads two runs of 16 consecutive LDS words (two
// particles), so bank conflicts stay at most two-way.  A particle costs 140 FMA and 64 LDS
// wave instructions, against about 500 VALU and 80 LDS instructions in the vector gather.
// Every sum has a fixed order.
// ---------------------------------------------------------------------------------
template <int SS, bool POT, bool FORCE>
__global__ __launch_bounds__(512, 4) void k_interp_row16(const Rec<float>* __restrict__ rec,
                                                         const int* __restrict__ item_tile,
                                                         const int* __restrict__ item_beg,
                                                         const int* __restrict__ item_end,
                                                         const int* __restrict__ n_items,
                                                         const float* __restrict__ grid, Geom g, Tiling tl, int N,
                                                         double* __restrict__ phi, double* __restrict__ force,
                                                         const SysDesc* __restrict__ sys, DenseSel dense)
{
    constexpr int P = 16, S2 = SS * SS;
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    const int beg = item_beg[item], end = item_end[item], n = end - beg;
    // the column items are the matrix-core gather's
    if (dense(item, n)) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* sub = reinterpret_cast<float*>(smem_raw);   // [y][x][z], SS^3
    const int M = g.M, Mz = g.Mz, NT = tl.NT, NTz = tl.NTz;
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
        load_tile_dma<true>(sub, grid, SS, M, Mz, ox, oy, oz, tid, nthr);
    }
    __syncthreads();

    const int lane = tid & 63, z = lane & 15, wave = tid >> 6, nw = nthr >> 6;
    const float h = (float)g.h, c = (float)g.c, dn = (float)g.d;
    const float kc = -c * 1.4426950408889634f;
    const double h3 = g.h * g.h * g.h;
    // The next iteration's record fields are loaded one iteration ahead (the output index
    // early in the iteration) and waited for before this iteration's output stores: vmcnt
    // counts loads and stores together, so a wait at the top of the loop would also wait for
    // the previous iteration's stores.
    typedef float v4f_t __attribute__((ext_vector_type(4)));
    const int p00 = beg + 4 * wave;
    v4f_t rn = {0.f, 0.f, 0.f, 0.f};
    int lon = 0;
    if (p00 < end) {
        const int pi = min(p00 + (lane >> 4), end - 1);
        rn = *reinterpret_cast<const v4f_t*>(rec + pi);
        lon = rec[pi].lo;
    }
    asm volatile("" ::"v"(rn.x), "v"(rn.y), "v"(rn.z), "v"(rn.w), "v"(lon));
    for (int p0 = p00; p0 < end; p0 += 4 * nw) {   // 4 particles per wave
        const int p = p0 + (lane >> 4);
        const bool valid = p < end;
        const v4f_t r = rn;   // t[0], t[1], t[2], q
        const int lo = lon;
        const int orig = rec[valid ? p : p0].orig;
        {
            const int pi = min(p + 4 * nw, end - 1);   // the next iteration's records
            rn = *reinterpret_cast<const v4f_t*>(rec + pi);
            lon = rec[pi].lo;
        }
        const int lx = lo & 255, ly = (lo >> 8) & 255, lz = lo >> 16;
        // lane z of the row evaluates window index z of the x and y factors; every lane
        // of the row then takes all 16 from the row's lanes
        float fx[P], gx[P], fy[P], gy[P];
        {
            const float o = (float)(g.first + z) * h;
            const float ax = r.x - o, ay = r.y - o;
            const float fxz = gauss(kc, ax) * dn, fyz = gauss(kc, ay);
            row_bcast16(fxz, fx);
            row_bcast16(fyz, fy);
            if (FORCE) {
                row_bcast16(fxz * (-c * ax), gx);
                row_bcast16(fyz * (-c * ay), gy);
            }
        }
        const float* base = sub + (ly * SS + lx) * SS + lz + z;
        float u = 0.f, ux = 0.f, uy = 0.f;
#pragma unroll
        for (int x = 0; x < P; ++x) {
            float c0 = 0.f, c1 = 0.f;
#pragma unroll
            for (int y = 0; y < P; ++y) {
                const float G = base[y * S2 + x * SS];
                c0 = fmaf(fy[y], G, c0);
                if (FORCE) c1 = fmaf(gy[y], G, c1);
            }
            u = fmaf(fx[x], c0, u);
            if (FORCE) {
                ux = fmaf(gx[x], c0, ux);
                uy = fmaf(fx[x], c1, uy);
            }
        }
        const float az = r.z - (float)(g.first + z) * h;
        const float fz = gauss(kc, az);
        // sums over the row: quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror
        auto row_sum = [](float v) {
            v = dpp_add<0xB1, 0xF>(v);
            v = dpp_add<0x4E, 0xF>(v);
            v = dpp_add<0x141, 0xF>(v);
            return dpp_add<0x140, 0xF>(v);
        };
        const float sp = row_sum(fz * u);
        float sx = 0.f, sy = 0.f, szz = 0.f;
        if (FORCE) {
            sx = row_sum(fz * ux);
            sy = row_sum(fz * uy);
            szz = row_sum(fz * (-c * az) * u);
        }
        asm volatile("" ::"v"(rn.x), "v"(rn.y), "v"(rn.z), "v"(rn.w), "v"(lon), "v"(orig));
        if (valid && z == 0) {
            if (POT) store_out(phi, orig, h3 * (double)sp);
            if (FORCE) {
                const double qh = h3 * (double)r.w;
                store_out(force, orig, qh * (double)sx);
                store_out(force, orig + (long)N, qh * (double)sy);
                store_out(force, orig + 2l * N, qh * (double)szz);
            }
        }
    }
}

bool pencil_supported(int P) { return P == 8 || P == 11 || P == 16; }

static void allow_lds_p(const void* fn, size_t bytes)
{
    if (bytes > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Items of fewer particles than this (whole sparse tiles) take the plane spread, the others
// the pencil spread; 0 turns the plane spread off.
int spread_sparse_max(const Tiling& tl) { return tl.plane ? sparse_tile_max(tl) : 0; }


bool launch_spread_pencil(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                          const int* n_items, const int* tile_start, const Geom& g, const Tiling& tl, float* partials,
                          const int* item_col, hipStream_t s, hipStream_t side, hipEvent_t fork, hipEvent_t join)
{
    if (!pencil_supported(g.P)) return false;
    const size_t lds = (size_t)tl.S * tl.S * tl.S * sizeof(double);
    SE_REQUIRE(tl.sblock >= 64 && tl.sblock % 64 == 0 && tl.sblock <= spread_pencil_max_threads(g.P), SE_ERR_ARG,
               "spread work-group size past the kernel's bound");
    const int sparse_max = spread_sparse_max(tl);
    // The column items (P = 16 with the column path) take the column spread; given a side
    // stream, the pencil and plane spreads of the other items run on it concurrently (all
    // three write disjoint partials), and s waits for them.
    const DenseSel dense{(tl.column && g.P == 16) ? item_col : nullptr, 0};
    hipStream_t sv = s;
    if (dense.col) {
        if (side) {
            SE_HIP(hipEventRecord(fork, s));
            SE_HIP(hipStreamWaitEvent(side, fork, 0));
            sv = side;
        }
        launch_spread_col16(rec, item_beg, item_end, n_items, g, tl, partials, dense.col, 512, s);
    }
    dim3 gr(tl.max_items), bl(tl.sblock);
#define SE_SP(PV)                                                                                            \
    do {                                                                                                     \
        allow_lds_p((const void*)k_spread_pencil<PV>, lds);                                                  \
        hipLaunchKernelGGL(k_spread_pencil<PV>, gr, bl, lds, sv, rec, item_beg, item_end, n_items, g, tl,   \
                           partials, item_tile, tile_start, sparse_max, dense);                               \
        if (sparse_max > 0)                                                                                  \
            hipLaunchKernelGGL(k_spread_plane<PV>, gr, dim3(64 * ((tl.S + 1) / 2)), 0, sv, rec, item_tile,     \
                               item_beg, item_end, n_items, tile_start, g, tl, partials, sparse_max, dense);  \
    } while (0)
    switch (g.P) {
    case 8: SE_SP(8); break;
    case 11: SE_SP(11); break;
    default: SE_SP(16); break;
    }
#undef SE_SP
    if (sv != s) {
        SE_HIP(hipEventRecord(join, sv));
        SE_HIP(hipStreamWaitEvent(s, join, 0));
    }
    return true;
}

static void interp_mfma16_modes(const Rec<float>* rec, const int* item_tile, const int* item_beg,
                                const int* item_end, const int* n_items, const float* grid, const Geom& g,
                                const Tiling& tl, int N, double* phi, double* force, const SysDesc* sys,
                                DenseSel dense, hipStream_t s)
{
    const size_t lds = ((size_t)tl.S * tl.S * tl.S + LDS_TILE_PAD) * sizeof(float);
    dim3 gr(tl.max_items), bl(tl.mblock);
#define SE_GM(POT, FORCE)                                                                                     \
    do {                                                                                                      \
        allow_lds_p((const void*)k_interp_mfma16<POT, FORCE>, lds);                                           \
        hipLaunchKernelGGL((k_interp_mfma16<POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg, item_end, \
                           n_items, grid, g, tl, N, phi, force, sys, dense);                                  \
    } while (0)
    if (phi && force)
        SE_GM(true, true);
    else if (force)
        SE_GM(false, true);
    else
        SE_GM(true, false);
#undef SE_GM
}

template <int PV>
static void interp_pencil_modes(const Rec<float>* rec, const int* item_tile, const int* item_beg,
                                const int* item_end, const int* n_items, const float* grid, const Geom& g,
                                const Tiling& tl, int N, double* phi, double* force, const SysDesc* sys,
                                DenseSel dense, hipStream_t s)
{
    const size_t lds = ((size_t)tl.S * tl.S * tl.S + LDS_TILE_PAD) * sizeof(float);
    dim3 gr(tl.max_items), bl(tl.block);
#define SE_GP(POT, FORCE)                                                                                     \
    do {                                                                                                      \
        allow_lds_p((const void*)k_interp_pencil<PV, POT, FORCE>, lds);                                       \
        hipLaunchKernelGGL((k_interp_pencil<PV, POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg,       \
                           item_end, n_items, grid, g, tl, N, phi, force, sys, dense);                        \
    } while (0)
    if (phi && force)
        SE_GP(true, true);
    else if (force)
        SE_GP(false, true);
    else
        SE_GP(true, false);
#undef SE_GP
}

static bool row16_supported(const Tiling& tl) { return tl.S == 19 || tl.S == 23 || tl.S == 27; }

// Items below 5K/2 particles (2.5 per (pencil, row) key: 1080 at tile 12) take the row
// gather, the others the matrix-core gather.  The row gather's cost per particle does not
// depend on density; the MFMA gather's falls as its 4-particle groups fill.  Measured at
// tile 12 (gather ms; uniform N = 1M at M = 128 with 824 particles per interior tile,
// clustered N = 4M at M = 256, uniform N = 4M at M = 256):
//   K: 0.60, 2.13, 2.13;  2K: 0.68, 2.08, 1.92;  5K/2: 0.50, 2.15, 1.95;  3K: 0.50, 2.17, 1.95
// (2K sends the Poisson tail of the uniform 1M tiles to a nearly idle MFMA launch).
int gather_dense_min(const Tiling& tl)
{
    if (!tl.mfma) return 0;
    return row16_supported(tl) ? 5 * tl.K / 2 : sparse_tile_max(tl);
}

// The row gather of the sparse items (P = 16, sub-grid sides 19, 23 and 27: tile sides
// 4, 8 and 12); false if the sub-grid side has no instance.
static bool interp_row16_modes(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                              const int* n_items, const float* grid, const Geom& g, const Tiling& tl, int N,
                              double* phi, double* force, const SysDesc* sys, DenseSel dense, hipStream_t s)
{
    const size_t lds = ((size_t)tl.S * tl.S * tl.S + LDS_TILE_PAD) * sizeof(float);
    dim3 gr(tl.max_items), bl(512);
#define SE_GQ(SV, POT, FORCE)                                                                                 \
    do {                                                                                                      \
        allow_lds_p((const void*)k_interp_row16<SV, POT, FORCE>, lds);                                         \
        hipLaunchKernelGGL((k_interp_row16<SV, POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg,         \
                           item_end, n_items, grid, g, tl, N, phi, force, sys, dense);                        \
    } while (0)
#define SE_GQS(SV)                        \
    do {                                  \
        if (phi && force)                 \
            SE_GQ(SV, true, true);        \
        else if (force)                   \
            SE_GQ(SV, false, true);       \
        else                              \
            SE_GQ(SV, true, false);       \
    } while (0)
    switch (tl.S) {
    case 19: SE_GQS(19); return true;
    case 23: SE_GQS(23); return true;
    case 27: SE_GQS(27); return true;
    default: return false;
    }
#undef SE_GQS
#undef SE_GQ
}

bool launch_interp_pencil(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                          const int* n_items, const float* grid, const Geom& g, const Tiling& tl, int N,
                          double* phi, double* force, const SysDesc* sys, const int* item_col, hipStream_t s,
                          hipStream_t side, hipEvent_t fork, hipEvent_t join)
{
#define SE_GPM(PV, DM) \
    interp_pencil_modes<PV>(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, sys, DM, s)
    switch (g.P) {
    case 8: SE_GPM(8, (DenseSel{nullptr, 0})); return true;
    case 11: SE_GPM(11, (DenseSel{nullptr, 0})); return true;
    case 16:
        if (tl.mfma) {
            // The dense items -- the column items with the column path, else those from
            // gather_dense_min particles -- take the matrix-core gather; the others take
            // the row gather (or, where it has no instance, the vector gather, which also
            // runs two work-groups of at most 256 threads per CU, one loading its sub-grid
            // while the other computes).  The two kernels write disjoint particles; given
            // a side stream they run concurrently, so each fills the other's tail.
            const DenseSel dense{tl.column ? item_col : nullptr, tl.column ? 0 : gather_dense_min(tl)};
            Tiling tv = tl;
            tv.block = tl.mblock;
            hipStream_t sv = s;
            if (side) {
                SE_HIP(hipEventRecord(fork, s));
                SE_HIP(hipStreamWaitEvent(side, fork, 0));
                sv = side;
            }
            if (!interp_row16_modes(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, sys,
                                   dense, sv))
                interp_pencil_modes<16>(rec, item_tile, item_beg, item_end, n_items, grid, g, tv, N, phi, force, sys,
                                        dense, sv);
            if (side) SE_HIP(hipEventRecord(join, side));
            // the column gather at 512 threads: 4 waves per SIMD (two work-groups per CU by
            // LDS); clustered N = 4M at M = 256: 256 threads 1.17 ms, 512 0.99, 1024 1.08
            if (tl.column)
                launch_interp_col16(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, sys,
                                    item_col, 512, s);
            else
                interp_mfma16_modes(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, phi, force, sys,
                                    dense, s);
            if (side) SE_HIP(hipStreamWaitEvent(s, join, 0));
        } else {
            SE_GPM(16, (DenseSel{nullptr, 0}));
        }
        return true;
    default: return false;
    }
#undef SE_GPM
}

}  // namespace se

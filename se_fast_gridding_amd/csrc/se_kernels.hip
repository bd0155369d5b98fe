// Hand-written gfx950 kernels of libse_fgg_amd.so.
//
// Two families:
//  * the fused engine (FP32 or FP64): bucketing by grid tile, tiled spread into
//    LDS sub-grids with per-work-item partials, fixed-order halo merge into the
//    periodic grid (wrap fused), Green's-function scaling on the R2C half
//    spectrum, tiled gather (extend fused) with a wavefront reduction;
//  * the reference-layout FP64 kernels behind the drop-in SE_FGG_* surface
//    (expansion arrays, extended row-major grid, column-major periodic grid).
#include <hip/hip_runtime.h>

#include <cmath>

#include "se_internal.h"

namespace se {

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

__device__ __forceinline__ int imod(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// Wave-uniform broadcast of lane l's value (v_readlane into an SGPR).
__device__ __forceinline__ float bcast(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast(double v, int l)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// LDS accumulate without return value (ds_add_f32 / ds_add_f64).
template <typename T>
__device__ __forceinline__ void lds_add(T* p, T v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float dexp(float a) { return expf(a); }
__device__ __forceinline__ double dexp(double a) { return exp(a); }

// Kernels whose dynamic LDS exceeds the default per-launch limit must opt in once.
static void allow_lds(const void* fn, size_t bytes)
{
    if (bytes <= 64 * 1024) return;
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// ================================================================================
// Utility
// ================================================================================
__global__ void k_fill_i32(int* p, long n, int v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

void launch_fill_i32(int* p, long n, int v, hipStream_t s)
{
    if (n <= 0) return;
    int blocks = (int)std::min<long>(cdiv(n, 256), 4096);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, s, p, n, v);
}

__global__ void k_fill_u32(unsigned int* p, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

void launch_fill_bytes0(void* p, size_t bytes, hipStream_t s)
{
    long n = (long)(bytes / 4);
    if (n <= 0) return;
    int blocks = (int)std::min<long>(cdiv(n, 256), 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, (unsigned int*)p, n);
}

// ================================================================================
// Bucketing: particle -> (tile, pencil, row) of its window start (counting sort)
// ================================================================================
// x and y: offset by the grid origin x0 (0 unless SE1P); for SE1P the particle must lie
// in [xlo, xhi) (the free directions' accepted range)
__device__ __forceinline__ int window_start(double x, const Geom& g, double& t, bool& ok)
{
    const double xr = x - g.x0;
    double u = xr / g.h;
    ok = fabs(u) < 1.0e9 && (g.per != 1 || (x >= g.xlo && x < g.xhi));
    if (!ok) u = 0.0;
    int iota = g.odd ? (int)round(u) : (int)floor(u);
    t = xr - g.h * iota;
    return imod(iota + g.first, g.M);
}

// z: offset by the grid origin z0; for SE2P the particle must lie in the slab
__device__ __forceinline__ int window_start_z(double z, const Geom& g, double& t, bool& ok)
{
    const double zr = z - g.z0;
    double u = zr / g.h;
    ok = fabs(u) < 1.0e9 && (g.per != 2 || (z >= g.zlo && z < g.zhi));
    if (!ok) u = 0.0;
    int iota = g.odd ? (int)round(u) : (int)floor(u);
    t = zr - g.h * iota;
    return imod(iota + g.first, g.Mz);
}

// Counting-sort key of a particle: its window start cell, ordered tile-major, then
// by pencil -- 4 window-start columns in x (px = lx / 4) by 1 in z (lz) --, then by
// row y inside the tile:
//   key = tile * K + ((px * TS + lz) * TS + ly),  K = PX * TS * TS,  PX = ceil(TS/4).
// Sorting by this key groups particles by tile (for the work items) and, inside a
// tile, into pencils walked row by row (for the pencil kernels).
__device__ __forceinline__ int row_key(int w0, int w1, int w2, const Tiling& tl)
{
    const int tx = w0 / tl.TS, ty = w1 / tl.TS, tz = w2 / tl.TS;
    const int lx = w0 - tx * tl.TS, ly = w1 - ty * tl.TS, lz = w2 - tz * tl.TS;
    const int tile = (tx * tl.NT + ty) * tl.NTz + tz;
    return tile * tl.K + (((lx >> 2) * tl.TS + lz) * tl.TS + ly);
}

// Sort key of every particle and its index, the value the stable key sort carries.
template <typename T>
__global__ void k_sort_keys(const double* __restrict__ x, int N, Geom g, Tiling tl, int* __restrict__ key,
                            int* __restrict__ idx)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    bool ok0, ok1, ok2;
    double t;
    int w0 = window_start(x[n], g, t, ok0);
    int w1 = window_start(x[n + (long)N], g, t, ok1);
    int w2 = window_start_z(x[n + 2l * N], g, t, ok2);
    // non-finite or absurd positions (and, for SE2P, positions outside the slab) go to a
    // discard bucket after all tiles
    key[n] = (ok0 && ok1 && ok2) ? row_key(w0, w1, w2, tl) : tl.nbins;
    idx[n] = n;
}

template <typename T>
void launch_sort_keys(const double* x, int N, const Geom& g, const Tiling& tl, int* key, int* idx, hipStream_t s)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_sort_keys<T>, dim3(cdiv(N, 256)), dim3(256), 0, s, x, N, g, tl, key, idx);
}

// First sorted position of every tile, from the sorted keys: position i starts the tiles
// (tile(i - 1), tile(i)] (tile = key / K; the discard key nt K counts as tile nt; the
// sentinels are -1 before the first position and nt + 1 past the last), so each of the
// nt + 1 entries is written exactly once, with no atomics and no scan.
__global__ void k_tile_starts(const int* __restrict__ skey, int N, int nt, int K, int* __restrict__ tile_start)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    const int prev = i > 0 ? skey[i - 1] / K : -1;
    const int cur = i < N ? skey[i] / K : nt + 1;
    for (int t = prev + 1; t <= min(cur, nt); ++t) tile_start[t] = i;
}

void launch_tile_starts(const int* skey, int N, int nt, int K, int* tile_start, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_starts, dim3(cdiv(N + 1, 256)), dim3(256), 0, s, skey, N, nt, K, tile_start);
}

// Per tile: particle count (from the tile starts) and number of work items.
__global__ void k_tile_chunks(const int* __restrict__ tile_start, int ntiles, int chunk,
                              int* __restrict__ tile_count, int* __restrict__ nchunk)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) {
        nchunk[t] = 0;
        return;
    }
    const int c = tile_start[t + 1] - tile_start[t];
    tile_count[t] = c;
    nchunk[t] = (c + chunk - 1) / chunk;
}

void launch_tile_chunks(const int* tile_start, int ntiles, int chunk, int* tile_count, int* nchunk, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_chunks, dim3(cdiv(ntiles + 1, 256)), dim3(256), 0, s, tile_start, ntiles, chunk,
                       tile_count, nchunk);
}

__global__ void k_items(const int* __restrict__ tile_count, const int* __restrict__ tile_start,
                        const int* __restrict__ chunk_start, int ntiles, int chunk,
                        int* __restrict__ item_tile, int* __restrict__ item_beg,
                        int* __restrict__ item_end, int* __restrict__ n_items)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *n_items = chunk_start[ntiles];
    if (t >= ntiles) return;
    const int b = tile_start[t], e = b + tile_count[t];
    int it = chunk_start[t];
    for (int p = b; p < e; p += chunk, ++it) {
        item_tile[it] = t;
        item_beg[it] = p;
        item_end[it] = min(p + chunk, e);
    }
}

void launch_items(const int* tile_count, const int* tile_start, const int* chunk_start, int ntiles, int chunk,
                  int* item_tile, int* item_beg, int* item_end, int* n_items, hipStream_t s)
{
    hipLaunchKernelGGL(k_items, dim3(cdiv(ntiles, 256)), dim3(256), 0, s, tile_count, tile_start, chunk_start,
                       ntiles, chunk, item_tile, item_beg, item_end, n_items);
}

// Records in sorted order: particle n goes to its sorted position pos[n].
template <typename T>
__global__ void k_scatter(const double* __restrict__ x, const double* __restrict__ q, int N, Geom g,
                          Tiling tl, const int* __restrict__ pos, Rec<T>* __restrict__ rec)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    bool ok;
    double t0, t1, t2;
    int w0 = window_start(x[n], g, t0, ok);
    int w1 = window_start(x[n + (long)N], g, t1, ok);
    int w2 = window_start_z(x[n + 2l * N], g, t2, ok);
    Rec<T> r;
    r.t[0] = (T)t0;
    r.t[1] = (T)t1;
    r.t[2] = (T)t2;
    r.q = (T)q[n];
    r.lo = (w0 % tl.TS) | ((w1 % tl.TS) << 8) | ((w2 % tl.TS) << 16);
    r.orig = n;
    rec[pos[n]] = r;
}

template <typename T>
void launch_scatter(const double* x, const double* q, int N, const Geom& g, const Tiling& tl, const int* pos,
                    Rec<T>* rec, hipStream_t s)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_scatter<T>, dim3(cdiv(N, 256)), dim3(256), 0, s, x, q, N, g, tl, pos, rec);
}

// Grouped bucketing: block b belongs to the last system whose first block is <= b
// (systems without particles own no block), found by a wave-uniform binary search.
__device__ __forceinline__ int block_system(const SysDesc* __restrict__ sys, int S, int b)
{
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sys[mid].blk <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T>
__global__ void k_sort_keys_group(const SysDesc* __restrict__ sys, int S, Geom g, Tiling tl, int* __restrict__ key,
                                  int* __restrict__ idx)
{
    const int si = block_system(sys, S, blockIdx.x);
    const SysDesc d = sys[si];
    const int n = (blockIdx.x - d.blk) * 256 + threadIdx.x;
    if (n >= d.N) return;
    bool ok0, ok1, ok2;
    double t;
    const int w0 = window_start(d.x[n], g, t, ok0);
    const int w1 = window_start(d.x[n + (long)d.N], g, t, ok1);
    const int w2 = window_start_z(d.x[n + 2l * d.N], g, t, ok2);
    key[d.off + n] = (ok0 && ok1 && ok2) ? si * tl.nbins + row_key(w0, w1, w2, tl) : S * tl.nbins;
    idx[d.off + n] = d.off + n;
}

template <typename T>
void launch_sort_keys_group(const SysDesc* sys, int S, int blocks, const Geom& g, const Tiling& tl, int* key,
                            int* idx, hipStream_t s)
{
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_sort_keys_group<T>, dim3(blocks), dim3(256), 0, s, sys, S, g, tl, key, idx);
}

template <typename T>
__global__ void k_scatter_group(const SysDesc* __restrict__ sys, int S, Geom g, Tiling tl,
                                const int* __restrict__ pos, Rec<T>* __restrict__ rec)
{
    const int si = block_system(sys, S, blockIdx.x);
    const SysDesc d = sys[si];
    const int n = (blockIdx.x - d.blk) * 256 + threadIdx.x;
    if (n >= d.N) return;
    bool ok;
    double t0, t1, t2;
    const int w0 = window_start(d.x[n], g, t0, ok);
    const int w1 = window_start(d.x[n + (long)d.N], g, t1, ok);
    const int w2 = window_start_z(d.x[n + 2l * d.N], g, t2, ok);
    const int f = d.off + n;
    Rec<T> r;
    r.t[0] = (T)t0;
    r.t[1] = (T)t1;
    r.t[2] = (T)t2;
    r.q = (T)d.q[n];
    r.lo = (w0 % tl.TS) | ((w1 % tl.TS) << 8) | ((w2 % tl.TS) << 16);
    r.orig = n;   // index inside the system; the gather finds the system from the tile
    rec[pos[f]] = r;
}

template <typename T>
void launch_scatter_group(const SysDesc* sys, int S, int blocks, const Geom& g, const Tiling& tl, const int* pos,
                          Rec<T>* rec, hipStream_t s)
{
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_scatter_group<T>, dim3(blocks), dim3(256), 0, s, sys, S, g, tl, pos, rec);
}

// ================================================================================
// Tiled spread (generic P, FP32 or FP64).  One work item = up to `chunk` particles
// of one tile.  The work-group zeroes an S^3 FP64 LDS sub-grid, each wave takes
// one particle at a time: lanes 0..P-1 build the three 1-D Gaussian factors (q*d
// folded into x), lane l owns (j,k) pairs l, l+64, ... of the P^2 window columns
// and adds X[i]*Y[j]*Z[k] with ds_add_f64 for i = 0..P-1.  (The sub-grid is FP64
// because gfx950 executes ds_add_f32 about 20x slower than ds_add_f64.)  The
// sub-grid is then written to the item's partial slot; the halo merge sums slots
// in fixed order.
// ================================================================================
// Lane l of a wave owns the window columns (j,k) = divmod(l + 64u, P), u < U,
// U = ceil(P^2/64).  PT > 0 fixes P at compile time; PT == 0 takes P from the
// geometry with ceil(P^2/64) <= UM.
struct LaneCols {
    int P, PP, lane;
    __device__ LaneCols(int P_, int lane_) : P(P_), PP(P_ * P_), lane(lane_) {}
    __device__ bool ok(int u) const { return lane + 64 * u < PP; }
    __device__ int j(int u) const { return ok(u) ? (lane + 64 * u) / P : 0; }
    __device__ int k(int u) const { return ok(u) ? (lane + 64 * u) - j(u) * P : 0; }
};

template <typename T, int PT, int UM>
__global__ __launch_bounds__(512) void k_spread_tiles(
    const Rec<T>* __restrict__ rec, const int* __restrict__ item_tile, const int* __restrict__ item_beg,
    const int* __restrict__ item_end, const int* __restrict__ n_items, Geom g, Tiling tl,
    T* __restrict__ partials)
{
    const int item = blockIdx.x;
    if (item >= *n_items) return;   // uniform across the work-group
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ float wmax[8];
    double* sub = reinterpret_cast<double*>(smem_raw);   // FP64 accumulation (ds_add_f64)
    const int P = PT > 0 ? PT : g.P;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int beg = item_beg[item], end = item_end[item];
    constexpr bool FIX = sizeof(T) == 4;   // FP32: integer-valued (deterministic) LDS sums
    for (int e = tid; e < tl.sub_elems; e += nthr) sub[e] = 0.0;
    if (FIX) item_qmax_part(rec, beg, end, tid, nthr, wmax);
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    const LaneCols lc(P, lane);
    int off[UM];
#pragma unroll
    for (int u = 0; u < UM; ++u) off[u] = lc.j(u) * tl.Szp + lc.k(u);
    const int slab = tl.S * tl.Szp;
    // FP32: fixed-point scale of the item's LDS sums (flush_scale), folded into q d
    const int sc = FIX ? flush_scale((double)(end - beg) * item_qmax_read(wmax, nw) * g.d, g.d) : 0;
    const T h = (T)g.h, c = (T)g.c, dn = (T)ldexp(g.d, sc);
    for (int p = beg + wave; p < end; p += nw) {
        const Rec<T> r = rec[p];
        T fx = 0, fy = 0, fz = 0;
        if (lane < P) {
            const T o = (T)(g.first + lane) * h;
            const T ax = r.t[0] - o, ay = r.t[1] - o, az = r.t[2] - o;
            fx = dexp(-c * ax * ax) * (r.q * dn);
            fy = dexp(-c * ay * ay);
            fz = dexp(-c * az * az);
        }
        T yz[UM];
#pragma unroll
        for (int u = 0; u < UM; ++u) {
            T v = __shfl(fy, lc.j(u), 64) * __shfl(fz, lc.k(u), 64);
            yz[u] = lc.ok(u) ? v : T(0);
        }
        const int lx = r.lo & 255, ly = (r.lo >> 8) & 255, lz = r.lo >> 16;
        double* base = sub + (lx * tl.S + ly) * tl.Szp + lz;
#pragma unroll 4
        for (int i = 0; i < P; ++i) {
            const T xi = bcast(fx, i);
            double* row = base + i * slab;
#pragma unroll
            for (int u = 0; u < UM; ++u)
                if (lc.ok(u)) lds_add(row + off[u], FIX ? (double)rintf((float)(xi * yz[u])) : (double)(xi * yz[u]));
        }
    }
    __syncthreads();
    T* out = partials + (size_t)item * tl.part_elems;
    const int S = tl.S;
    SubWalk w(S, tid, nthr);   // a = ix, b = iy, c = iz
    for (int e = tid; e < tl.part_elems; e += nthr) {
        const double v = sub[(w.a * S + w.b) * tl.Szp + w.c];
        out[e] = (T)(FIX ? ldexp(v, -sc) : v);
        w.next();
    }
}

template <typename T>
void launch_spread_tiles(const Rec<T>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                         const int* n_items, const Geom& g, const Tiling& tl, T* partials, hipStream_t s)
{
    size_t lds = (size_t)tl.sub_elems * sizeof(double);
    dim3 grid(tl.max_items), block(tl.block);
#define SE_SPREAD(PT, UM)                                                                                   \
    do {                                                                                                    \
        allow_lds((const void*)k_spread_tiles<T, PT, UM>, lds);                                             \
        hipLaunchKernelGGL((k_spread_tiles<T, PT, UM>), grid, block, lds, s, rec, item_tile, item_beg,      \
                           item_end, n_items, g, tl, partials);                                             \
    } while (0)
    const int U = (g.P * g.P + 63) / 64;
    switch (g.P) {
    case 8: SE_SPREAD(8, 1); break;
    case 11: SE_SPREAD(11, 2); break;
    case 16: SE_SPREAD(16, 4); break;
    case 21: SE_SPREAD(21, 7); break;
    default:
        if (U <= 1) SE_SPREAD(0, 1);
        else if (U <= 2) SE_SPREAD(0, 2);
        else if (U <= 4) SE_SPREAD(0, 4);
        else if (U <= 8) SE_SPREAD(0, 8);
        else if (U <= 12) SE_SPREAD(0, 12);
        else SE_SPREAD(0, 16);
    }
#undef SE_SPREAD
}

// ================================================================================
// Halo merge (= wrap), owner-tile form: one work-group per tile of the periodic grid
// sums, in a fixed order, every partial value that lands on the tile: for each box of
// partial sub-grid elements that maps onto the tile (the product of one per-dimension
// segment from each of the three segment tables) and each work item of the box's source
// tile.  Each thread owns fixed tile points (point e = thread + 256 j, z fastest) and
// keeps their sums in registers, so the (box, item) pairs need no barrier between them:
// the thread walks the pairs in order (MB at a time with their loads in flight; one
// measured fastest, as more registers per thread cost occupancy), and adds each value to
// its point in that order, so the result does not depend on scheduling.  FP32 partials
// are summed in FP64; FP64 partials with Kahan compensation.
// seg[t * sstride + a] = (source tile coordinate, local start, output start, length).
//
// Addressing: point (x, y, z) of the tile lies at lin = (x S + y) S + z plus a shift
// that is fixed per box in the partial sub-grid of a source item, so each box costs one
// scalar base and each point a 32-bit offset of a raw buffer load.  A point outside the
// box takes an offset past the buffer's range, which loads 0 without a memory access:
// the loads need no branch, and an FP32 value of 0 is added unconditionally (adding +0
// leaves an FP64 sum that starts at +0 bit for bit unchanged).  With MASK, a point's
// membership of the tile's segments is one word (bit a: x segment a, bit 10 + b: y
// segment b, bit 20 + c: z segment c), so the box test is one AND and compare against
// the box's three bits; without it (more than 10 segments per dimension, only on small
// grids) the word holds the point's coordinates and the test reads the box's segments.
// ================================================================================
template <typename T>
__device__ __forceinline__ void merge_add(double& acc, double& cmp, T v);
template <>
__device__ __forceinline__ void merge_add<float>(double& acc, double&, float v)
{
    acc += (double)v;
}
template <>
__device__ __forceinline__ void merge_add<double>(double& acc, double& cmp, double v)
{
    const double y = v - cmp;
    const double t = acc + y;
    cmp = (t - acc) - y;
    acc = t;
}

// LDS of the merge: the segment tables and, per segment triple, (item range, shift, bits)
inline size_t merge_lds_bytes(int sstride)
{
    return (size_t)3 * sstride * sizeof(int4) + (size_t)sstride * sstride * sstride * sizeof(int4);
}

constexpr int MERGE_THREADS = 256;        // threads per work-group (one tile)
constexpr int MERGE_MASK_SEGS = 10;       // most segments per dimension of the MASK form
constexpr int MERGE_RANGE = 0x7ffffff0;   // byte range of the merge's buffer loads

// JM tile points per thread (TS^3 <= MERGE_THREADS JM); MB (box, item) pairs in flight
template <typename T, int JM, int MB, bool MASK>
__global__ __launch_bounds__(MERGE_THREADS) void k_merge_tiles(const T* __restrict__ partials,
                                                     const int* __restrict__ chunk_start,
                                                     const int4* __restrict__ seg, const int* __restrict__ nseg,
                                                     const int4* __restrict__ segz, const int* __restrict__ nsegz,
                                                     int sstride, Geom g, Tiling tl, T* __restrict__ grid)
{
    // XCD-aware order: work-groups are dealt to the 8 XCDs round-robin, so XCD
    // (blockIdx % 8) walks whole z-columns of tiles (tz = 0, 1, ...) of the (tx, ty)
    // columns congruent to it mod 8.  Neighbours in z read the two ends of the same
    // z-rows of a partial, which then stay in that XCD's L2.
    const int NT = tl.NT, NTz = tl.NTz, TS = tl.TS, S = tl.S, M = g.M, Mz = g.Mz;
    // stacked systems of a grouped solve: system blockIdx.y owns tiles [sys*ntiles, ...)
    // and grid [sys][x][y][z]
    chunk_start += (long)blockIdx.y * tl.ntiles;
    grid += (long)blockIdx.y * M * M * Mz;
    const int xcd = blockIdx.x % 8, k = blockIdx.x / 8;
    const int col = xcd + 8 * (k / NTz), tz = k % NTz;
    if (col >= NT * NT) return;   // uniform: the last round of columns is partial
    const int tx = col / NT, ty = col % NT;
    const int ox = tx * TS, oy = ty * TS, oz = tz * TS;
    const int nx = min(TS, M - ox), ny = min(TS, M - oy), nz = min(TS, Mz - oz);
    const int V = nx * ny * nz;
    // the tile's source boxes: the segments of each dimension and, per segment triple,
    // the item range of its source tile, the box's shift and its segment bits, fetched in
    // parallel into LDS once instead of one dependent global load per triple
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int4* sA = reinterpret_cast<int4*>(smem_raw);
    int4* sB = sA + sstride;
    int4* sC = sB + sstride;
    int4* sR = sC + sstride;
    const int na = nseg[tx], nb = nseg[ty], nc = nsegz[tz], ntri = na * nb * nc;
    if (threadIdx.x < na) sA[threadIdx.x] = seg[tx * sstride + threadIdx.x];
    if (threadIdx.x < nb) sB[threadIdx.x] = seg[ty * sstride + threadIdx.x];
    if (threadIdx.x < nc) sC[threadIdx.x] = segz[tz * sstride + threadIdx.x];
    __syncthreads();
    for (int t = threadIdx.x; t < ntri; t += MERGE_THREADS) {
        const int a = t / (nb * nc), b = (t / nc) % nb, c = t % nc;
        const int4 A = sA[a], B = sB[b], C = sC[c];
        const int src = (A.x * NT + B.x) * NTz + C.x;
        const unsigned bits = MASK ? (1u << a) | (1u << (10 + b)) | (1u << (20 + c)) : 0u;
        sR[t] = make_int4(chunk_start[src], chunk_start[src + 1], ((A.y - A.z) * S + (B.y - B.z)) * S + (C.y - C.z),
                          (int)bits);
    }
    __syncthreads();
    // per owned point: its offset lin and its word (segment bits with MASK, else the packed
    // coordinates x << 20 | y << 10 | z, all ones past the tile)
    int lin[JM];
    unsigned pk[JM];
    double acc[JM], cmp[JM];
#pragma unroll
    for (int j = 0; j < JM; ++j) {
        const int e = threadIdx.x + MERGE_THREADS * j;
        const int ez = e % nz, ey = (e / nz) % ny, ex = e / (nz * ny);
        unsigned w = MASK ? 0u : ~0u;
        if (e < V) {
            if constexpr (MASK) {
                for (int a = 0; a < na; ++a) w |= (unsigned)((unsigned)(ex - sA[a].z) < (unsigned)sA[a].w) << a;
                for (int b = 0; b < nb; ++b) w |= (unsigned)((unsigned)(ey - sB[b].z) < (unsigned)sB[b].w) << (10 + b);
                for (int c = 0; c < nc; ++c) w |= (unsigned)((unsigned)(ez - sC[c].z) < (unsigned)sC[c].w) << (20 + c);
            } else {
                w = ((unsigned)ex << 20) | ((unsigned)ey << 10) | (unsigned)ez;
            }
        }
        pk[j] = w;
        lin[j] = (ex * S + ey) * S + ez;
        acc[j] = 0.0;
        cmp[j] = 0.0;
    }
    // cursor over the (box, item) pairs in order: box t, item it of [it, ie), all uniform
    int t = 0, it = 0, ie = 0;
    while (t < ntri && sR[t].x == sR[t].y) ++t;
    if (t < ntri) {
        it = __builtin_amdgcn_readfirstlane(sR[t].x);
        ie = __builtin_amdgcn_readfirstlane(sR[t].y);
    }
    while (t < ntri) {
        T v[MB][JM];
        unsigned msk[MB];
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            msk[m] = 0;
            if (t < ntri) {
                const int shift = __builtin_amdgcn_readfirstlane(sR[t].z);
                const T* base = partials + (long)it * tl.part_elems + shift;
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, MERGE_RANGE, 0x00020000);
                unsigned sel = 0;
                int4 A{}, B{}, C{};
                if constexpr (MASK) {
                    sel = (unsigned)__builtin_amdgcn_readfirstlane(sR[t].w);
                } else {
                    A = sA[t / (nb * nc)];
                    B = sB[(t / nc) % nb];
                    C = sC[t % nc];
                }
#pragma unroll
                for (int j = 0; j < JM; ++j) {
                    bool in;
                    if constexpr (MASK) {
                        in = (pk[j] & sel) == sel;
                    } else {
                        in = pk[j] != ~0u && (pk[j] >> 20) - (unsigned)A.z < (unsigned)A.w &&
                             ((pk[j] >> 10) & 1023u) - (unsigned)B.z < (unsigned)B.w &&
                             (pk[j] & 1023u) - (unsigned)C.z < (unsigned)C.w;
                    }
                    const int off = in ? lin[j] * (int)sizeof(T) : MERGE_RANGE;
                    if constexpr (sizeof(T) == 4)
                        v[m][j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
                    else
                        v[m][j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
                    msk[m] |= (unsigned)in << j;
                }
                if (++it == ie) {
                    ++t;
                    while (t < ntri && sR[t].x == sR[t].y) ++t;
                    if (t < ntri) {
                        it = __builtin_amdgcn_readfirstlane(sR[t].x);
                        ie = __builtin_amdgcn_readfirstlane(sR[t].y);
                    }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int j = 0; j < JM; ++j)
                if (sizeof(T) == 4 || (msk[m] >> j & 1u)) merge_add<T>(acc[j], cmp[j], v[m][j]);
    }
#pragma unroll
    for (int j = 0; j < JM; ++j) {
        const int e = threadIdx.x + MERGE_THREADS * j;
        const int ez = e % nz, ey = (e / nz) % ny, ex = e / (nz * ny);
        if (e < V) grid[((long)(ox + ex) * M + (oy + ey)) * Mz + (oz + ez)] = (T)acc[j];
    }
}

template <typename T, int JM>
static void launch_merge_jm(const T* partials, const int* chunk_start, const int4* seg, const int* nseg,
                            const int4* segz, const int* nsegz, int sstride, const Geom& g, const Tiling& tl, T* grid,
                            int nsys, hipStream_t s)
{
    constexpr int MB = 1;   // 1, 2, 3, 4, 8 pairs: 0.061, 0.062, 0.064, 0.066, 0.069 ms (N = 1M, M = 128)
    const int cols8 = (tl.NT * tl.NT + 7) / 8;
    dim3 gr(8 * cols8 * tl.NTz, nsys), bl(MERGE_THREADS);
    const size_t lds = merge_lds_bytes(sstride);
    if (sstride <= MERGE_MASK_SEGS) {
        allow_lds((const void*)k_merge_tiles<T, JM, MB, true>, lds);
        hipLaunchKernelGGL((k_merge_tiles<T, JM, MB, true>), gr, bl, lds, s, partials, chunk_start, seg, nseg, segz,
                           nsegz, sstride, g, tl, grid);
    } else {
        allow_lds((const void*)k_merge_tiles<T, JM, MB, false>, lds);
        hipLaunchKernelGGL((k_merge_tiles<T, JM, MB, false>), gr, bl, lds, s, partials, chunk_start, seg, nseg, segz,
                           nsegz, sstride, g, tl, grid);
    }
}

template <typename T>
void launch_merge(const T* partials, const int* chunk_start, const int4* seg, const int* nseg, const int4* segz,
                  const int* nsegz, int sstride, const Geom& g, const Tiling& tl, T* grid, int nsys, hipStream_t s)
{
    if (nsys <= 0) return;
    const int pts = tl.TS * tl.TS * tl.TS;
    SE_REQUIRE(tl.TS <= 16, SE_ERR_ARG, "merge: tile side above 16");
    if (pts <= MERGE_THREADS * 2)
        launch_merge_jm<T, 2>(partials, chunk_start, seg, nseg, segz, nsegz, sstride, g, tl, grid, nsys, s);
    else if (pts <= MERGE_THREADS * 7)
        launch_merge_jm<T, 7>(partials, chunk_start, seg, nseg, segz, nsegz, sstride, g, tl, grid, nsys, s);
    else if (pts <= MERGE_THREADS * 11)
        launch_merge_jm<T, 11>(partials, chunk_start, seg, nseg, segz, nsegz, sstride, g, tl, grid, nsys, s);
    else
        launch_merge_jm<T, 16>(partials, chunk_start, seg, nseg, segz, nsegz, sstride, g, tl, grid, nsys, s);
}

// ================================================================================
// Tiled gather (= extend + interpolate).  The work-group loads its tile's S^3
// sub-grid from the periodic grid into LDS (modular indexing), then each wave
// takes one particle: r_u = sum_i X_i G[i, j_u, k_u] per owned (j,k) pair,
// phi_lane = sum_u Y_j Z_k r_u, and a wavefront reduction in FP64.  Force uses
// the derivative factors zf = -c (t - o h) (SE_fgg.c:693-709):
// F = h^3 q sum G W zf = (q/2) grad phi.
// ================================================================================
// The generic gather's work-group bound: 768 threads (3 waves per SIMD in 168 registers)
// where the FP64 kernel fits them (up to 8 owned pairs per lane, P <= 22), else 512.
template <typename T, int UM>
constexpr int interp_tiles_max_threads() { return (sizeof(T) == 8 && UM <= 8) ? 768 : 512; }

template <typename T, int PT, int UM, bool POT, bool FORCE>
__global__ __launch_bounds__((interp_tiles_max_threads<T, UM>())) void k_interp_tiles(
    const Rec<T>* __restrict__ rec, const int* __restrict__ item_tile, const int* __restrict__ item_beg,
    const int* __restrict__ item_end, const int* __restrict__ n_items, const T* __restrict__ grid, Geom g,
    Tiling tl, int N, double* __restrict__ phi, double* __restrict__ force, const SysDesc* __restrict__ sys)
{
    const int item = blockIdx.x;
    if (item >= *n_items) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sub = reinterpret_cast<T*>(smem_raw);
    const int P = PT > 0 ? PT : g.P;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int M = g.M, Mz = g.Mz, S = tl.S, NT = tl.NT, NTz = tl.NTz;
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
        // [x][y][z] with z pitch Szp == S (choose_tiling), so point e lands at sub[e]
        load_tile<false>(sub, grid, S, M, Mz, ox, oy, oz, tid, nthr);
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    const LaneCols lc(P, lane);
    int off[UM];
#pragma unroll
    for (int u = 0; u < UM; ++u) off[u] = lc.j(u) * tl.Szp + lc.k(u);
    const int slab = S * tl.Szp;
    const T h = (T)g.h, c = (T)g.c, dn = (T)g.d;
    const double h3 = g.h * g.h * g.h;
    const int beg = item_beg[item], end = item_end[item];
    for (int p = beg + wave; p < end; p += nw) {
        const Rec<T> r = rec[p];
        T fx = 0, fy = 0, fz = 0, dx = 0, dy = 0, dz = 0;
        if (lane < P) {
            const T o = (T)(g.first + lane) * h;
            const T ax = r.t[0] - o, ay = r.t[1] - o, az = r.t[2] - o;
            fx = dexp(-c * ax * ax) * dn;
            fy = dexp(-c * ay * ay);
            fz = dexp(-c * az * az);
            dx = -c * ax;
            dy = -c * ay;
            dz = -c * az;
        }
        T yz[UM];
#pragma unroll
        for (int u = 0; u < UM; ++u) {
            T v = __shfl(fy, lc.j(u), 64) * __shfl(fz, lc.k(u), 64);
            yz[u] = lc.ok(u) ? v : T(0);
        }
        const int lx = r.lo & 255, ly = (r.lo >> 8) & 255, lz = r.lo >> 16;
        const T* base = sub + (lx * S + ly) * tl.Szp + lz;
        T acc[UM], accf[UM];
#pragma unroll
        for (int u = 0; u < UM; ++u) acc[u] = accf[u] = T(0);
        // FP64 with many columns per lane: no unrolling, or hoisted LDS loads spill
        constexpr int UNR = (sizeof(T) == 8 && UM > 4) ? 1 : 4;
#pragma unroll UNR
        for (int i = 0; i < P; ++i) {
            const T xi = bcast(fx, i);
            const T xf = FORCE ? xi * bcast(dx, i) : T(0);
            const T* row = base + i * slab;
#pragma unroll
            for (int u = 0; u < UM; ++u) {
                if (lc.ok(u)) {
                    const T sv = row[off[u]];
                    acc[u] += xi * sv;
                    if (FORCE) accf[u] += xf * sv;
                }
            }
        }
        double sp = 0, s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int u = 0; u < UM; ++u) {
            const double w = (double)(yz[u] * acc[u]);
            sp += w;
            if (FORCE) {
                s0 += (double)(yz[u] * accf[u]);
                s1 += w * (double)__shfl(dy, lc.j(u), 64);
                s2 += w * (double)__shfl(dz, lc.k(u), 64);
            }
        }
        if (POT) sp = wave_sum(sp);
        if (FORCE) {
            s0 = wave_sum(s0);
            s1 = wave_sum(s1);
            s2 = wave_sum(s2);
        }
        if (lane == 0) {
            const int o = r.orig;
            if (POT) store_out(phi, o, h3 * sp);
            if (FORCE) {
                const double qh = h3 * (double)r.q;
                store_out(force, o, qh * s0);
                store_out(force, o + (long)N, qh * s1);
                store_out(force, o + 2l * N, qh * s2);
            }
        }
    }
}

template <typename T, int PT, int UM>
static void interp_dispatch_mode(const Rec<T>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                                 const int* n_items, const T* grid, const Geom& g, const Tiling& tl, int N,
                                 double* phi, double* force, const SysDesc* sys, hipStream_t s)
{
    size_t lds = (size_t)tl.sub_elems * sizeof(T);
    // at its bound when the block size is automatic: FP64 P = 21 at C2, 3.65 -> 3.13 ms
    dim3 gr(tl.max_items), bl(tl.block_auto ? interp_tiles_max_threads<T, UM>() : tl.block);
#define SE_GATHER(POT, FORCE)                                                                                  \
    do {                                                                                                       \
        allow_lds((const void*)k_interp_tiles<T, PT, UM, POT, FORCE>, lds);                                    \
        hipLaunchKernelGGL((k_interp_tiles<T, PT, UM, POT, FORCE>), gr, bl, lds, s, rec, item_tile, item_beg,  \
                           item_end, n_items, grid, g, tl, N, phi, force, sys);                                \
    } while (0)
    if (phi && force)
        SE_GATHER(true, true);
    else if (force)
        SE_GATHER(false, true);
    else
        SE_GATHER(true, false);
#undef SE_GATHER
}

template <typename T>
void launch_interp_tiles(const Rec<T>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                         const int* n_items, const T* grid, const Geom& g, const Tiling& tl, int N, double* phi,
                         double* force, const SysDesc* sys, hipStream_t s)
{
#define SE_GD(PT, UM) interp_dispatch_mode<T, PT, UM>(rec, item_tile, item_beg, item_end, n_items, grid, g, tl, N, \
                                                     phi, force, sys, s)
    const int U = (g.P * g.P + 63) / 64;
    switch (g.P) {
    case 8: SE_GD(8, 1); break;
    case 11: SE_GD(11, 2); break;
    case 16: SE_GD(16, 4); break;
    case 21: SE_GD(21, 7); break;
    default:
        if (U <= 1) SE_GD(0, 1);
        else if (U <= 2) SE_GD(0, 2);
        else if (U <= 4) SE_GD(0, 4);
        else if (U <= 8) SE_GD(0, 8);
        else if (U <= 12) SE_GD(0, 12);
        else SE_GD(0, 16);
    }
#undef SE_GD
}

// ================================================================================
// Green's function on the R2C half spectrum (SE_general.c:141-176 and
// spectral_ewald.c:126-131): Z = 4pi/M^3 exp(s K^2)/K^2, Z(0) = 0, with
// exp(s K^2) = prod_d exp(s k_d^2) from a per-axis table (no M^3 exp).
// tab[0..M) = k^2, tab[M..2M) = exp(s k^2), tab[2M] = 4pi/M^3.
// ================================================================================
// SE3P Green's function (spectral_ewald.c:126-130, SE_general.c:141-176), separable:
// Z = 4 pi/(M M Mz) e^{scalar k^2} / k^2 with e^{scalar k^2} = e_x e_y e_z, Z(0) = 0.
// tab = [k_x^2 (M), e_x (M), k_z^2 (Mz), e_z (Mz), 4 pi/(M M Mz)].
__device__ __forceinline__ double kfreq(int i, int n, double period)
{
    const int pos = (n % 2 == 0) ? n / 2 : (n - 1) / 2 + 1;   // k_vec (SE_general.c:34-84)
    return 2.0 * M_PI / period * (double)(i < pos ? i : i - n);
}

__global__ void k_green_tables(double* tab, int M, int Mz, double L, double Lz, double scalar)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) tab[2 * M + 2 * Mz] = 4.0 * M_PI / ((double)M * M * Mz);
    if (i < M) {
        const double k = kfreq(i, M, L);
        tab[i] = k * k;
        tab[M + i] = exp(scalar * k * k);
    }
    if (i < Mz) {
        const double k = kfreq(i, Mz, Lz);
        tab[2 * M + i] = k * k;
        tab[2 * M + Mz + i] = exp(scalar * k * k);
    }
}

void launch_green_tables(double* tab, int M, int Mz, double L, double Lz, double scalar, hipStream_t s)
{
    const int n = M > Mz ? M : Mz;
    hipLaunchKernelGGL(k_green_tables, dim3(cdiv(n, 256)), dim3(256), 0, s, tab, M, Mz, L, Lz, scalar);
}

template <typename T>
__global__ void k_green(T* __restrict__ spec, int M, int Mz, const double* __restrict__ tab)
{
    const int Mh = Mz / 2 + 1;
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= (long)M * M * Mh) return;
    spec += 2l * blockIdx.y * M * M * Mh;   // stacked spectra of a grouped solve
    const int kc = (int)(gid % Mh);
    const long r = gid / Mh;
    const int j = (int)(r % M), i = (int)(r / M);
    const double* kz2 = tab + 2 * M;
    const double* ez = tab + 2 * M + Mz;
    const double K2 = (tab[i] + tab[j]) + kz2[kc];
    const double z = (K2 > 0.0) ? tab[M + i] * tab[M + j] * ez[kc] / K2 * tab[2 * M + 2 * Mz] : 0.0;
    spec[2 * gid] = (T)((double)spec[2 * gid] * z);
    spec[2 * gid + 1] = (T)((double)spec[2 * gid + 1] * z);
}

template <typename T>
void launch_green(void* spec, int M, int Mz, const double* tab, int nsys, hipStream_t s)
{
    long n = (long)M * M * (Mz / 2 + 1);
    if (nsys <= 0) return;
    hipLaunchKernelGGL(k_green<T>, dim3(cdiv(n, 256), nsys), dim3(256), 0, s, (T*)spec, M, Mz, tab);
}

// SE2P Green's function on the half spectrum of the M x M x Mz grid, z zero-padded.
// The free direction uses the kernel truncated at |z| <= R, whose transform is exact on
// the padded period Mz h >= R + (slab span), so the periodic FFT convolution equals the
// free one in z (truncated-kernel method for free-space convolution):
//   k_par != 0:  FT of (2 pi/k) e^{-k|z|} 1(|z|<=R) = 4 pi [k - e^{-kR}(k cos kR' - kz sin kR')] / (k (k^2+kz^2)),
//                kR' = kz R (k = |k_par|),
//   k_par == 0:  FT of -2 pi |z| 1(|z|<=R) = 4 pi [2 sin^2(kz R/2)/kz^2 - R sin(kz R)/kz],  -2 pi R^2 at kz = 0,
// times e^{scalar (k_par^2 + kz^2)} / (M M Mz), scalar = -(1-eta)/(4 xi^2).  For R -> inf
// the k_par != 0 part is the SE3P kernel 4 pi/k^2; the potential converges to the SE2P
// Ewald k-space sum (SE_direct.c:256-338).
__global__ void k_green2p_table(double* __restrict__ tab, int M, int Mz, double L, double h, double R,
                                double scalar)
{
    const int Mh = Mz / 2 + 1;
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= (long)M * M * Mh) return;
    const int kc = (int)(gid % Mh);
    const long r = gid / Mh;
    const int j = (int)(r % M), i = (int)(r / M);
    const double kx = kfreq(i, M, L), ky = kfreq(j, M, L);
    const double kz = 2.0 * M_PI / ((double)Mz * h) * (double)kc;
    const double kp2 = kx * kx + ky * ky, kp = sqrt(kp2), k2 = kp2 + kz * kz;
    double G;
    if (kp > 0.0) {
        G = 4.0 * M_PI / (kp * k2) * (kp - exp(-kp * R) * (kp * cos(kz * R) - kz * sin(kz * R)));
    } else if (kc == 0) {
        G = -2.0 * M_PI * R * R;
    } else {
        const double sh = sin(0.5 * kz * R);
        G = 4.0 * M_PI * (2.0 * sh * sh / (kz * kz) - R * sin(kz * R) / kz);
    }
    tab[gid] = G * exp(scalar * k2) / ((double)M * M * Mz);
}

void launch_green2p_table(double* tab, int M, int Mz, double L, double h, double R, double scalar, hipStream_t s)
{
    long n = (long)M * M * (Mz / 2 + 1);
    hipLaunchKernelGGL(k_green2p_table, dim3(cdiv(n, 256)), dim3(256), 0, s, tab, M, Mz, L, h, R, scalar);
}

// SE1P Green's function on the half spectrum of the M x M x Mz grid, x and y free and
// zero-padded, z periodic (period Lz).  Per z mode k_z the in-plane kernel is the 2-D
// Green's function of the mode, 2 K0(|k_z| rho) (k_z != 0) or -2 ln rho (k_z = 0),
// truncated at rho <= R; its 2-D transform at k = |k_par| is exact:
//   k_z != 0:  4 pi [1 + k R J1(kR) K0(k_z R) - k_z R J0(kR) K1(k_z R)] / (k^2 + k_z^2),
//   k_z == 0:  -4 pi [R ln R J1(kR)/k + (J0(kR) - 1)/k^2],   c00 at k = 0,
// (from d/drho [rho (k J1 K0 - k_z J0 K1)] = (k^2 + k_z^2) rho J0 K0 and
// d/drho [rho ln rho J1/k + J0/k^2] = rho ln rho J0), times e^{scalar (k^2 + k_z^2)} /
// (M M Mz).  With the padded period M h >= R + (span of the charged and read grid
// points) the periodic FFT convolution equals the free one in x and y, and the potential
// is the SE1P Ewald k-space sum (SE_direct.c:90-176).
__global__ void k_green1p_table(double* __restrict__ tab, const double* __restrict__ jtab,
                                const double* __restrict__ ktab, int M, int Mz, double h, double Lz, double R,
                                double scalar, double c00)
{
    const int Mh = Mz / 2 + 1;
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= (long)M * M * Mh) return;
    const int kc = (int)(gid % Mh);
    const long r = gid / Mh;
    const int j = (int)(r % M), i = (int)(r / M);
    const double kx = kfreq(i, M, (double)M * h), ky = kfreq(j, M, (double)M * h);
    const double kz = 2.0 * M_PI / Lz * (double)kc;
    const double kp2 = kx * kx + ky * ky, kp = sqrt(kp2), k2 = kp2 + kz * kz;
    const double J0 = jtab[r], J1 = jtab[(long)M * M + r];
    double G;
    if (kc > 0) {
        G = 4.0 * M_PI * (1.0 + kp * R * J1 * ktab[kc] - kz * R * J0 * ktab[Mh + kc]) / k2;
    } else if (kp > 0.0) {
        G = -4.0 * M_PI * (R * log(R) * J1 / kp + (J0 - 1.0) / kp2);
    } else {
        G = c00;
    }
    tab[gid] = G * exp(scalar * k2) / ((double)M * M * Mz);
}

void launch_green1p_table(double* tab, const double* jtab, const double* ktab, int M, int Mz, double h, double Lz,
                          double R, double scalar, double c00, hipStream_t s)
{
    long n = (long)M * M * (Mz / 2 + 1);
    hipLaunchKernelGGL(k_green1p_table, dim3(cdiv(n, 256)), dim3(256), 0, s, tab, jtab, ktab, M, Mz, h, Lz, R, scalar,
                       c00);
}

template <typename T>
__global__ void k_green_full(T* __restrict__ spec, long n, const double* __restrict__ tab)
{
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (gid >= n) return;
    spec += 2l * blockIdx.y * n;   // stacked spectra of a grouped solve
    const double z = tab[gid];
    spec[2 * gid] = (T)((double)spec[2 * gid] * z);
    spec[2 * gid + 1] = (T)((double)spec[2 * gid + 1] * z);
}

template <typename T>
void launch_green_full(void* spec, long n, const double* tab, int nsys, hipStream_t s)
{
    if (nsys <= 0) return;
    hipLaunchKernelGGL(k_green_full<T>, dim3(cdiv(n, 256), nsys), dim3(256), 0, s, (T*)spec, n, tab);
}

// E = 1/2 sum phi q  (SE_fgg.c:5934-5946), FP64 block tree + one atomic per block.
__global__ void k_energy(const double* __restrict__ phi, const double* __restrict__ q, int N, double* out)
{
    __shared__ double red[4];
    double v = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
        v += phi[i] * q[i];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, 0.5 * (red[0] + red[1] + red[2] + red[3]));
}

__global__ void k_zero_d(double* p) { *p = 0.0; }

void launch_energy(const double* phi, const double* q, int N, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, s, out);
    if (N <= 0) return;
    int blocks = std::min(cdiv(N, 256), 1024);
    hipLaunchKernelGGL(k_energy, dim3(blocks), dim3(256), 0, s, phi, q, N, out);
}

// ================================================================================
// Reference-layout FP64 kernels (drop-in SE_FGG_* surface)
// ================================================================================

// SE_fgg.c:508-535
__global__ void k_base_gaussian(LParams p, double* zs)
{
    const int P = p.P, n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P * P * P) return;
    const int lo = (P & 1) ? p.P_half : p.P_half - 1;
    const int i = n / (P * P) - lo, j = (n / P) % P - lo, k = n % P - lo;
    const double h = p.h;
    zs[n] = p.d * exp(-p.c * ((i * h) * (i * h) + (j * h) * (j * h) + (k * h) * (k * h)));
}

void launch_base_gaussian(const LParams& p, double* zs, hipStream_t s)
{
    int n = p.P * p.P * p.P;
    hipLaunchKernelGGL(k_base_gaussian, dim3(cdiv(n, 256)), dim3(256), 0, s, p, zs);
}

// Centering of one coordinate (SE_fgg.c:554-573, 2P :741-768, 1P :838-865).
__device__ void legacy_center(const LParams& p, int per, int d, double x, int& from, double& t)
{
    const bool freed = (per == 2 && d == 2) || (per == 1 && d < 2);
    const int odd = p.P & 1, ph = p.P_half;
    const double h = p.h;
    int i;
    if (freed) {
        const double origin = ((per == 2) ? p.a : (d == 0 ? p.a : p.b)) + h / 2;
        const double u = (x - origin) / h;
        i = odd ? (int)round(u) : (int)floor(u);
        t = x - (i * h + origin);
    } else {
        const double u = x / h;
        i = odd ? (int)round(u) : (int)floor(u);
        t = x - h * i;
    }
    from = odd ? i - ph : i - (ph - 1);
}

__device__ int legacy_index(const LParams& p, int per, const int from[3])
{
    const int ph = p.P_half;
    const int e0 = from[0] + (per == 1 ? 0 : ph), e1 = from[1] + (per == 1 ? 0 : ph),
              e2 = from[2] + (per == 2 ? 0 : ph);
    return (e0 * p.npdims[1] + e1) * p.npdims[2] + e2;
}

// True when the whole P^3 window starting at extended index i0 lies inside the
// extended grid.  The reference does not check (positions outside the box index
// outside H); here such particles are skipped (spread) or get NaN (gather).
__device__ __forceinline__ bool window_inside(const LParams& p, long i0)
{
    const long n1 = p.npdims[1], n2 = p.npdims[2];
    const long last = i0 + ((long)(p.P - 1) * n1 + (p.P - 1)) * n2 + (p.P - 1);
    return i0 >= 0 && last < (long)p.npdims[0] * n1 * n2;
}

// One thread per (particle, dimension): SE_fgg.c:539-722 (pow, then recurrence).
__global__ void k_expand_all(LParams p, int force, const double* __restrict__ x, double* zx, double* zy, double* zz,
                             int* idx, double* zfx, double* zfy, double* zfz)
{
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int N = p.N, P = p.P;
    if (gid >= 3l * N) return;
    const int n = (int)(gid / 3), d = (int)(gid % 3);
    const int per = (force && p.per == 1) ? 3 : p.per;   // SE_fgg.h:29-31
    int from[3];
    double t[3];
    for (int e = 0; e < 3; ++e) legacy_center(p, per, e, x[n + (long)e * N], from[e], t[e]);
    const double c = p.c, h = p.h;
    const int first = (P & 1) ? -p.P_half : -p.P_half + 1;
    const double z3 = exp(-c * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2])) * 1.0;
    const double base = exp(2 * c * h * t[d]);
    double z = pow(base, (double)first);
    double* out = (d == 0 ? zx : (d == 1 ? zy : zz)) + (long)n * P;
    double* fo = force ? (d == 0 ? zfx : (d == 1 ? zfy : zfz)) + (long)n * P : nullptr;
    for (int i = 0; i < P; ++i) {
        if (i > 0) z *= base;
        out[i] = (d == 0) ? z * z3 : z;
        if (fo) fo[i] = -c * (t[d] - (first + i) * h);
    }
    if (d == 0) idx[n] = legacy_index(p, per, from);
}

void launch_expand_all(const LParams& p, int force, const double* x, double* zx, double* zy, double* zz, int* idx,
                       double* zfx, double* zfy, double* zfz, hipStream_t s)
{
    if (p.N <= 0) return;
    hipLaunchKernelGGL(k_expand_all, dim3(cdiv(3l * p.N, 256)), dim3(256), 0, s, p, force, x, zx, zy, zz, idx, zfx,
                       zfy, zfz);
}

// SE_FGG_grid_split (SE_fgg.c:3983-4028): one wave per particle, global FP64 atomics.
__global__ void k_grid_split_ext(LParams p, const double* __restrict__ q, const double* __restrict__ zx,
                                 const double* __restrict__ zy, const double* __restrict__ zz,
                                 const int* __restrict__ idx, const double* __restrict__ zs, double* H)
{
    const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (n >= p.N) return;
    const int P = p.P, n1 = p.npdims[1], n2 = p.npdims[2];
    const double qn = q[n];
    const long i0 = idx[n];
    if (!window_inside(p, i0)) return;
    for (int pr = lane; pr < P * P; pr += 64) {
        const int j = pr / P, k = pr % P;
        const double zzk = zz[(long)n * P + k], zyj = zy[(long)n * P + j];
        for (int i = 0; i < P; ++i) {
            const double cij = qn * zx[(long)n * P + i] * zyj;
            atomicAdd(H + i0 + ((long)i * n1 + j) * n2 + k, zs[(i * P + j) * P + k] * zzk * cij);
        }
    }
}

void launch_grid_split_ext(const LParams& p, const double* q, const double* zx, const double* zy, const double* zz,
                           const int* idx, const double* zs, double* H, hipStream_t s)
{
    if (p.N <= 0) return;
    hipLaunchKernelGGL(k_grid_split_ext, dim3(cdiv(p.N, 4)), dim3(256), 0, s, p, q, zx, zy, zz, idx, zs, H);
}

// Per-wave on-the-fly expansion into LDS: f[0..P) = zx (times qx), f[32..) = zy, f[64..) = zz,
// f[96..) .. zf (force).  Returns the extended-grid index.
__device__ int legacy_expand_wave(const LParams& p, int per, const double* x, long n, int N, double qx, int lane,
                                  double* f, bool force)
{
    int from[3];
    double t[3];
    for (int e = 0; e < 3; ++e) legacy_center(p, per, e, x[n + (long)e * N], from[e], t[e]);
    const int P = p.P;
    const int first = (P & 1) ? -p.P_half : -p.P_half + 1;
    const double c = p.c, h = p.h;
    if (lane < P) {
        const double z3 = exp(-c * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2])) * qx;
        for (int d = 0; d < 3; ++d) {
            const double base = exp(2 * c * h * t[d]);
            const double z = pow(base, (double)(first + lane));
            f[32 * d + lane] = (d == 0) ? z * z3 : z;
            if (force) f[96 + 32 * d + lane] = -c * (t[d] - (first + lane) * h);
        }
    }
    __syncthreads();
    return legacy_index(p, per, from);
}

// SE_FGG_grid (SE_fgg.c:3866-3914): one wave per block.
__global__ void k_grid_ext(LParams p, const double* __restrict__ x, const double* __restrict__ q,
                           const double* __restrict__ zs, double* H)
{
    __shared__ double f[192];
    const int n = blockIdx.x, lane = threadIdx.x;
    const int P = p.P, n1 = p.npdims[1], n2 = p.npdims[2];
    const int i0 = legacy_expand_wave(p, p.per, x, n, p.N, q[n], lane, f, false);
    if (!window_inside(p, i0)) return;
    for (int pr = lane; pr < P * P; pr += 64) {
        const int j = pr / P, k = pr % P;
        for (int i = 0; i < P; ++i) {
            const double cij = f[i] * f[32 + j];
            atomicAdd(H + i0 + ((long)i * n1 + j) * n2 + k, zs[(i * P + j) * P + k] * f[64 + k] * cij);
        }
    }
}

void launch_grid_ext(const LParams& p, const double* x, const double* q, const double* zs, double* H, hipStream_t s)
{
    if (p.N <= 0) return;
    hipLaunchKernelGGL(k_grid_ext, dim3(p.N), dim3(64), 0, s, p, x, q, zs, H);
}

// SE_FGG_int_split (SE_fgg.c:1176-1223) and _force (:2241-2312): one wave per particle.
__global__ void k_int_split_ext(LParams p, int force, const double* __restrict__ H, const double* __restrict__ q,
                                const double* __restrict__ zx, const double* __restrict__ zy,
                                const double* __restrict__ zz, const double* __restrict__ zfx,
                                const double* __restrict__ zfy, const double* __restrict__ zfz,
                                const int* __restrict__ idx, const double* __restrict__ zs, double* out)
{
    const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (n >= p.N) return;
    const int P = p.P, n1 = p.npdims[1], n2 = p.npdims[2];
    const long i0 = idx[n];
    const long b = (long)n * P;
    const double qm = force ? q[n] : 1.0;
    if (!window_inside(p, i0)) {
        if (lane == 0) {
            out[n] = NAN;
            if (force) out[n + (long)p.N] = out[n + 2l * p.N] = NAN;
        }
        return;
    }
    double s0 = 0, s1 = 0, s2 = 0;
    for (int pr = lane; pr < P * P; pr += 64) {
        const int j = pr / P, k = pr % P;
        for (int i = 0; i < P; ++i) {
            const double cij = zx[b + i] * zy[b + j];
            const double v = H[i0 + ((long)i * n1 + j) * n2 + k] * zs[(i * P + j) * P + k] * zz[b + k] * cij;
            if (force) {
                const double w = v * qm;
                s0 += w * zfx[b + i];
                s1 += w * zfy[b + j];
                s2 += w * zfz[b + k];
            } else {
                s0 += v;
            }
        }
    }
    s0 = wave_sum(s0);
    if (force) {
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
    }
    if (lane == 0) {
        const double h3 = p.h * p.h * p.h;
        out[n] = h3 * s0;
        if (force) {
            out[n + (long)p.N] = h3 * s1;
            out[n + 2l * p.N] = h3 * s2;
        }
    }
}

void launch_int_split_ext(const LParams& p, int force, const double* H, const double* q, const double* zx,
                          const double* zy, const double* zz, const double* zfx, const double* zfy,
                          const double* zfz, const int* idx, const double* zs, double* out, hipStream_t s)
{
    if (p.N <= 0) return;
    hipLaunchKernelGGL(k_int_split_ext, dim3(cdiv(p.N, 4)), dim3(256), 0, s, p, force, H, q, zx, zy, zz, zfx, zfy,
                       zfz, idx, zs, out);
}

// SE_FGG_int (SE_fgg.c:968-1019, q = 1) and SE_FGG_int_force (:1024-1100, q in the expansion).
__global__ void k_int_ext(LParams p, int force, const double* __restrict__ H, const double* __restrict__ x,
                          const double* __restrict__ q, const double* __restrict__ zs, double* out)
{
    __shared__ double f[192];
    const int n = blockIdx.x, lane = threadIdx.x;
    const int P = p.P, n1 = p.npdims[1], n2 = p.npdims[2];
    const int per = (force && p.per == 1) ? 3 : p.per;
    const long i0 = legacy_expand_wave(p, per, x, n, p.N, force ? q[n] : 1.0, lane, f, force);
    if (!window_inside(p, i0)) {
        if (lane == 0) {
            out[n] = NAN;
            if (force) out[n + (long)p.N] = out[n + 2l * p.N] = NAN;
        }
        return;
    }
    double s0 = 0, s1 = 0, s2 = 0;
    for (int pr = lane; pr < P * P; pr += 64) {
        const int j = pr / P, k = pr % P;
        for (int i = 0; i < P; ++i) {
            const double cij = f[i] * f[32 + j];
            const double v = H[i0 + ((long)i * n1 + j) * n2 + k] * zs[(i * P + j) * P + k] * f[64 + k] * cij;
            if (force) {
                s0 += v * f[96 + i];
                s1 += v * f[128 + j];
                s2 += v * f[160 + k];
            } else {
                s0 += v;
            }
        }
    }
    s0 = wave_sum(s0);
    if (force) {
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
    }
    if (lane == 0) {
        const double h3 = p.h * p.h * p.h;
        out[n] = h3 * s0;
        if (force) {
            out[n + (long)p.N] = h3 * s1;
            out[n + 2l * p.N] = h3 * s2;
        }
    }
}

void launch_int_ext(const LParams& p, int force, const double* H, const double* x, const double* q,
                    const double* zs, double* out, hipStream_t s)
{
    if (p.N <= 0) return;
    hipLaunchKernelGGL(k_int_ext, dim3(p.N), dim3(64), 0, s, p, force, H, x, q, zs, out);
}

// Wrap (SE_fgg.c:327-411) as a scatter with FP64 atomics; extend (:417-502) as a gather.
__device__ __forceinline__ long legacy_per_index(const LParams& p, long e)
{
    const int k = (int)(e % p.npdims[2]);
    const long r = e / p.npdims[2];
    const int j = (int)(r % p.npdims[1]), i = (int)(r / p.npdims[1]);
    const int ph = p.P_half;
    const int a = (p.per == 1) ? i : imod(i - ph, p.dims[0]);
    const int b = (p.per == 1) ? j : imod(j - ph, p.dims[1]);
    const int c = (p.per == 2) ? k : imod(k - ph, p.dims[2]);
    return (long)a + (long)b * p.dims[0] + (long)c * p.dims[0] * p.dims[1];
}

__global__ void k_wrap(LParams p, const double* __restrict__ H, double* Hper)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)p.npdims[0] * p.npdims[1] * p.npdims[2]) return;
    atomicAdd(Hper + legacy_per_index(p, e), H[e]);
}

__global__ void k_extend(LParams p, const double* __restrict__ Hper, double* __restrict__ H)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)p.npdims[0] * p.npdims[1] * p.npdims[2]) return;
    H[e] = Hper[legacy_per_index(p, e)];
}

void launch_wrap(const LParams& p, const double* H, double* Hper, hipStream_t s)
{
    long n = (long)p.npdims[0] * p.npdims[1] * p.npdims[2];
    hipLaunchKernelGGL(k_wrap, dim3(cdiv(n, 256)), dim3(256), 0, s, p, H, Hper);
}

void launch_extend(const LParams& p, const double* Hper, double* H, hipStream_t s)
{
    long n = (long)p.npdims[0] * p.npdims[1] * p.npdims[2];
    hipLaunchKernelGGL(k_extend, dim3(cdiv(n, 256)), dim3(256), 0, s, p, Hper, H);
}

// scaling (SE_general.c:141-176) on a cubic n^3 grid, reference index order.
__global__ void k_scaling(double scalar, double* Z, int n, double L)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)n * n * n) return;
    const int k = (int)(e % n);
    const long r = e / n;
    const int j = (int)(r % n), i = (int)(r / n);
    const int pos = (n % 2 == 0) ? n / 2 : (n - 1) / 2 + 1;
    const double f = 2.0 * M_PI / L;
    const double k1 = f * (i < pos ? i : i - n), k2 = f * (j < pos ? j : j - n), k3 = f * (k < pos ? k : k - n);
    const double K2 = (k1 * k1 + k2 * k2) + k3 * k3;
    const double d = 4.0 * M_PI / (double)((long)n * n * n);
    Z[e] = exp(scalar * K2) / K2 * d;
}

void launch_scaling(double scalar, double* Z, int n, double L, hipStream_t s)
{
    long m = (long)n * n * n;
    hipLaunchKernelGGL(k_scaling, dim3(cdiv(m, 256)), dim3(256), 0, s, scalar, Z, n, L);
}

// product_rc (SE_general.c:336-362)
__global__ void k_product_rc(const double* __restrict__ a, const double* __restrict__ b, double* c, long n, int flag)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double re = a[2 * i], im = a[2 * i + 1], s = b[i];
    if (flag == 1) {
        c[2 * i] = re * s;
        c[2 * i + 1] = im * s;
    } else {
        c[2 * i] = re / s;
        c[2 * i + 1] = im / s;
    }
}

void launch_product_rc(const double* a, const double* b, double* c, long n, int flag, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_product_rc, dim3(cdiv(n, 256)), dim3(256), 0, s, a, b, c, n, flag);
}

__global__ void k_reorder_gather(const double* __restrict__ xin, const double* __restrict__ qin,
                                 const int* __restrict__ perm, int N, double* xout, double* qout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int j = perm[i];
    xout[i] = xin[j];
    xout[i + (long)N] = xin[j + (long)N];
    xout[i + 2l * N] = xin[j + 2l * N];
    qout[i] = qin[j];
}

void launch_reorder_gather(const double* xin, const double* qin, const int* perm, int N, double* xout,
                           double* qout, hipStream_t s)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_reorder_gather, dim3(cdiv(N, 256)), dim3(256), 0, s, xin, qin, perm, N, xout, qout);
}

// ---- explicit instantiations ----------------------------------------------------
#define SE_INST(T)                                                                                               \
    template void launch_sort_keys<T>(const double*, int, const Geom&, const Tiling&, int*, int*, hipStream_t);  \
    template void launch_scatter<T>(const double*, const double*, int, const Geom&, const Tiling&, const int*,  \
                                    Rec<T>*, hipStream_t);                                                       \
    template void launch_sort_keys_group<T>(const SysDesc*, int, int, const Geom&, const Tiling&, int*, int*,   \
                                            hipStream_t);                                                        \
    template void launch_scatter_group<T>(const SysDesc*, int, int, const Geom&, const Tiling&, const int*,     \
                                          Rec<T>*, hipStream_t);                                                 \
    template void launch_spread_tiles<T>(const Rec<T>*, const int*, const int*, const int*, const int*,         \
                                         const Geom&, const Tiling&, T*, hipStream_t);                           \
    template void launch_merge<T>(const T*, const int*, const int4*, const int*, const int4*, const int*, int,  \
                                  const Geom&, const Tiling&, T*, int, hipStream_t);                             \
    template void launch_interp_tiles<T>(const Rec<T>*, const int*, const int*, const int*, const int*,         \
                                         const T*, const Geom&, const Tiling&, int, double*, double*,            \
                                         const SysDesc*, hipStream_t);                                           \
    template void launch_green<T>(void*, int, int, const double*, int, hipStream_t);                           \
    template void launch_green_full<T>(void*, long, const double*, int, hipStream_t);
SE_INST(float)
SE_INST(double)

}  // namespace se

// Internal declarations shared by the HIP translation units of libse_fgg_amd.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/se_gpu.h"

namespace se {

// ---------------------------------------------------------------------------------
// Errors: every HIP call goes through SE_HIP; failures are recorded for
// se_gpu_last_error() and turned into return codes by the C entry points.
// ---------------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();

struct Error {
    int code;
    std::string msg;
};

#define SE_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::se::Error{-2, std::string(#expr) + ": " + hipGetErrorString(e_)};   \
    } while (0)

#define SE_REQUIRE(cond, code, msg)                                                       \
    do {                                                                                  \
        if (!(cond)) throw ::se::Error{(code), std::string(msg)};                         \
    } while (0)

// ---------------------------------------------------------------------------------
// Geometry of one grid as the fused kernels see it: M x M x Mz points, periodic in x
// and y with period M h, and in z with period Mz h.  Support point i of a particle sits
// at index (iota + first + i) mod M, iota = round(u) for odd P, floor(u) for even P,
// u = x/h in x and y and u = (z - z0)/h in z (SE_fgg.c:554-573 for SE3P, where Mz = M
// and z0 = 0).  For SE2P (per == 2) the z grid is the free direction, zero-padded so
// that no support window reaches its ends; z0 = a + h/2 is the reference's 2P grid
// origin (SE_fgg.c:753-764) and particles outside [zlo, zhi) are left out.  For SE1P
// (per == 1) x and y are the free directions instead: u = (x - x0)/h in x and y with
// x0 = a + h/2 (the reference's 1P grid origin, SE_fgg.c:836-846), both zero-padded to
// M points, and particles outside [xlo, xhi) in x or y are left out; z is periodic.
// ---------------------------------------------------------------------------------
struct Geom {
    int M;        // grid points per side in x and y
    int P;        // support
    int first;    // -P_half (odd P) or 1-P_half (even P)
    int odd;
    double h, c, d;   // spacing, Gaussian exponent, normalisation (c/pi)^1.5
    int Mz;       // grid points in z
    int per;      // 3: SE3P, 2: SE2P (z free), 1: SE1P (x and y free)
    double z0;    // z coordinate of grid index 0 in z
    double zlo, zhi;  // SE2P slab: accepted particle z range
    double x0;        // x and y coordinate of grid index 0 in x and y (SE1P; 0 otherwise)
    double xlo, xhi;  // SE1P: accepted particle x and y range
};

// Tile decomposition used by the bucketed spread/gather.  A tile of TS^3 window
// start points owns an LDS sub-grid of S^3 points, S = TS + P - 1.
struct Tiling {
    int TS, NT, S;    // tile side, tiles per side in x and y (ceil(M/TS)), sub-grid side
    int NTz;          // tiles in z (ceil(Mz/TS))
    int Szp;          // z pitch of the LDS sub-grid (>= S)
    int ntiles;       // NT * NT * NTz
    int PX;           // pencils (4 columns wide in x, 1 in z) per tile side in x: ceil(TS/4)
    int K;            // sort keys per tile: PX*TS*TS
    int nbins;        // ntiles*K (the discard bucket is nbins)
    int chunk;        // max particles per work item
    int block;        // threads per work-group
    int sub_elems;    // S*S*Szp
    int part_elems;   // S*S*S
    int max_items;    // ntiles + n_cap / chunk + 1
    int nseg;         // max merge segments per tile coordinate (segment-table stride)
    int pencil;       // 1: register-ring pencil kernels (FP32, P in the compiled set)
    int mfma;         // 1: the pencil gather runs on MFMA (P = 16)
    int mblock;       // threads per work-group of the MFMA gather
    int sblock;       // threads per work-group of the register-ring (pencil) spread
    int block_auto;   // 1: block was not given, so each kernel may take its own size
    int plane;        // 1: sparse tiles take the plane spread (pencil spreads only)
    int column;       // 1: dense items take the 1 x 1 column kernels (P = 16 FP32, with mfma)
};

// Walk over the S^3 points of a tile sub-grid, e = (a S + b) S + c, visited as e = tid,
// tid + nthr, ...: the digits (a, b, c) of the thread's point advance by the digits of
// nthr with carries, so no point costs an integer division.
struct SubWalk {
    int a, b, c;      // digits of the current point
    int da, db, dc;   // digits of the step nthr
    int S;
    __device__ SubWalk(int S_, int tid, int nthr) : S(S_)
    {
        c = tid % S;
        b = (tid / S) % S;
        a = tid / (S * S);
        dc = nthr % S;
        db = (nthr / S) % S;
        da = nthr / (S * S);
    }
    __device__ void next()
    {
        c += dc;
        const int k1 = c >= S ? 1 : 0;
        c -= k1 * S;
        b += db + k1;
        const int k2 = b >= S ? 1 : 0;
        b -= k2 * S;
        a += da + k2;
    }
};

// v mod n for 0 <= v < (k + 1) n with small k (sub-grid offsets past the grid end)
__device__ __forceinline__ int wrap_index(int v, int n)
{
    while (v >= n) v -= n;
    return v;
}

// Loads the periodic S^3 block of `grid` (M x M x Mz, z fastest) that starts at grid
// point (ox, oy, oz) into LDS, laid out [y][x][z] (sub[(iy S + ix) S + iz]) if YXZ, else
// [x][y][z].  Each thread issues U loads before it writes any of them, so U global loads
// are in flight per thread instead of one.
template <bool YXZ, typename T, int U = 8>
__device__ __forceinline__ void load_tile(T* sub, const T* __restrict__ grid, int S, int M, int Mz, int ox, int oy,
                                          int oz, int tid, int nthr)
{
    const int S3 = S * S * S;
    SubWalk w(S, tid, nthr);   // (a, b) = (iy, ix) if YXZ, else (ix, iy); c = iz
    for (int e0 = tid; e0 < S3; e0 += U * nthr) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (e0 + u * nthr < S3) {
                const int ix = YXZ ? w.b : w.a, iy = YXZ ? w.a : w.b;
                const int gx = wrap_index(ox + ix, M), gy = wrap_index(oy + iy, M), gz = wrap_index(oz + w.c, Mz);
                v[u] = grid[((long)gx * M + gy) * Mz + gz];
            }
            w.next();
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (e0 + u * nthr < S3) sub[e0 + u * nthr] = v[u];
    }
}

// The same load by LDS DMA (global_load_lds_dword): wave w takes chunks of 64 dwords of
// the sub-grid, EPW = 256 / sizeof(T) consecutive points e0 .. e0 + EPW - 1
// (e0 = EPW (w + nw k)), and lane l loads dword l % DPE of point e0 + l / DPE
// (DPE = sizeof(T) / 4 dwords per point) straight into LDS, so a load holds no register
// and every load of the tile can be in flight at once.  In the last chunk the lanes past
// S^3 read the last point and write into the LDS_TILE_PAD points of padding that the
// caller allocates after the sub-grid.  Returns with the loads complete (vmcnt 0); the
// caller's barrier then publishes the sub-grid.
constexpr int LDS_TILE_PAD = 64;
template <bool YXZ, typename T>
__device__ __forceinline__ void load_tile_dma(T* sub, const T* __restrict__ grid, int S, int M, int Mz, int ox, int oy,
                                              int oz, int tid, int nthr)
{
    static_assert(sizeof(T) % 4 == 0, "points are whole dwords");
    constexpr int DPE = (int)(sizeof(T) / 4), EPW = 64 / DPE;
    const int S3 = S * S * S, lane = tid & 63, nw = nthr >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    SubWalk w(S, EPW * wave + lane / DPE, EPW * nw);   // (a, b) = (iy, ix) if YXZ, else (ix, iy); c = iz
    if (ox + S <= M && oy + S <= M && oz + S <= Mz && (long)M * M * Mz < (1l << 31)) {
        // an interior tile (no periodic wrap): the point's element offset from the tile's
        // origin is stepped with the walk's digits, a carry out of digit c or b adding
        // the stride difference it makes, so a load costs no multiply and no wrap test
        const T* tb = grid + ((long)ox * M + oy) * Mz + oz;
        const int sA = YXZ ? Mz : M * Mz, sB = YXZ ? M * Mz : Mz;   // strides of digits a and b
        int off = w.a * sA + w.b * sB + w.c;
        const int d0 = w.da * sA + w.db * sB + w.dc, cB = sB - S, cA = sA - S * sB;
        const int last = (S - 1) * (sA + sB + 1);
        for (int e0 = EPW * wave; e0 < S3; e0 += EPW * nw) {
            const int o = e0 + lane / DPE < S3 ? off : last;
            const unsigned* src = reinterpret_cast<const unsigned*>(tb + o) + lane % DPE;
            __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(sub + e0), 4, 0, 0);
            w.c += w.dc;
            const bool k1 = w.c >= S;
            w.c -= k1 ? S : 0;
            w.b += w.db + (k1 ? 1 : 0);
            const bool k2 = w.b >= S;
            w.b -= k2 ? S : 0;
            off += d0 + (k1 ? cB : 0) + (k2 ? cA : 0);
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): every DMA write has landed in LDS
        return;
    }
    for (int e0 = EPW * wave; e0 < S3; e0 += EPW * nw) {
        const bool in = e0 + lane / DPE < S3;
        const int a = in ? w.a : S - 1, b = in ? w.b : S - 1, c = in ? w.c : S - 1;
        const int ix = YXZ ? b : a, iy = YXZ ? a : b;
        const int gx = wrap_index(ox + ix, M), gy = wrap_index(oy + iy, M), gz = wrap_index(oz + c, Mz);
        const unsigned* src = reinterpret_cast<const unsigned*>(grid + ((long)gx * M + gy) * Mz + gz) + lane % DPE;
        __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(sub + e0), 4, 0, 0);
        w.next();
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): every DMA write has landed in LDS
}

// Counter slot of sort key k (nslots = keys + 1, the discard bucket included).  The
// particles of a dense cluster fall into runs of consecutive keys; their counters are
// scattered over blocks of 2^16 slots (256 KB; an odd multiplier is a bijection mod
// 2^16) so that the bucketing atomics of one cluster do not queue on one memory
// channel.  Keys in the last, partial block keep their slot.  The scan of the counts
// reads them back in key order through the same map.
__host__ __device__ __forceinline__ int count_slot(int k, int nslots)
{
    const int hi = k & ~0xFFFF;
    return hi + 0x10000 <= nslots ? (hi | (int)(((unsigned)k * 40503u) & 0xFFFFu)) : k;
}

// Deterministic FP32 spread sums.  The FP32 spreads add FP32 values into an FP64 LDS
// sub-grid with atomics, whose order varies from run to run.  Each value is scaled by 2^s
// and rounded to an integer before it is added, with s chosen per work item so that every
// partial sum of the item is an integer below 2^52 in magnitude: FP64 adds such integers
// exactly, so the sums -- and the FP32 partials made from them -- do not depend on the
// order of the atomics.  The bound of a partial sum is sum |q| d <= n max|q| d (each
// Gaussian factor is at most 1), so s = 50 - ilogb(n max|q| d) keeps it below 2^51; the
// quantum 2^-s is about n max|q| d 2^-51, far below FP32 resolution.  s is capped so that
// d 2^s stays a finite float (q d 2^s is then formed without overflow).
__device__ __forceinline__ int exponent_of(double v)   // ilogb for normal positive v
{
    return (int)((__double_as_longlong(v) >> 52) & 0x7ff) - 1023;
}
__device__ __forceinline__ int flush_scale(double bound, double d)
{
    if (!(bound >= 0x1p-1000) || !(bound < 1e300)) return 0;
    const int s = 50 - exponent_of(bound);
    const int smax = 125 - exponent_of(d);
    return s < smax ? s : smax;
}

// The same for 32-bit integer sums (the column spread's LDS sub-grid): every partial sum
// stays below 2^30 in magnitude, a quantum of about n max|q| d 2^-29.
__device__ __forceinline__ int flush_scale_i32(double bound, double d)
{
    if (!(bound >= 0x1p-1000) || !(bound < 1e300)) return 0;
    const int s = 29 - exponent_of(bound);
    const int smax = 125 - exponent_of(d);
    return s < smax ? s : smax;
}

// max |q| over the records [beg, end) of a work item: every thread reduces a strided share,
// the wave maximum goes to wmax[wave]; after the caller's barrier, item_qmax_read combines
// them.
template <typename R>
__device__ __forceinline__ void item_qmax_part(const R* __restrict__ rec, int beg, int end, int tid, int nthr,
                                               float* wmax)
{
    float m = 0.f;
    for (int i = beg + tid; i < end; i += nthr) m = fmaxf(m, fabsf((float)rec[i].q));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((tid & 63) == 0) wmax[tid >> 6] = m;
}
__device__ __forceinline__ float item_qmax_read(const float* wmax, int nw)
{
    float m = 0.f;
    for (int w = 0; w < nw; ++w) m = fmaxf(m, wmax[w]);
    return m;
}

// Work-group bound of the register-ring spread: the P = 16 kernel fits 168 registers
// (3 waves per SIMD) at 768 threads; the smaller supports fit 128 at 1024.
constexpr int spread_pencil_max_threads(int P) { return P == 16 ? 768 : 1024; }

// A sparse work item has fewer particles than half the tile's (pencil, row) keys: the
// plane spread takes those whole tiles, and where the P = 16 row gather has no instance,
// the vector gather takes those items instead of the matrix-core gather (a 4-particle
// MFMA group needs about one particle per pencil row).
inline int sparse_tile_max(const Tiling& tl) { return tl.K / 2 > 1 ? tl.K / 2 : 1; }
// P = 16 FP32 gather split without the column path: items with fewer particles than this
// take the row gather, the others the matrix-core gather (se_pencil.hip)
int gather_dense_min(const Tiling& tl);

// Particle record written by the bucketing scatter, read by spread and gather.
template <typename T>
struct alignas(16) Rec {
    T t[3];   // offset from the grid point iota*h (computed in FP64)
    T q;      // charge
    int lo;   // window start relative to the tile origin: x | y << 8 | z << 16
    int orig; // index in the caller's arrays
};

// One system of a grouped solve (several independent systems of one geometry in one
// pass).  Flat particle index n of the group = off + local index; the particle kernels
// run ceil(N/256) blocks per system starting at block `blk`.  Global tile index =
// system * ntiles + tile, so sort keys, work items and partials of all systems share one
// set of arrays; grids and spectra are stacked [system][x][y][z].
struct SysDesc {
    const double* x;   // SoA [3N]
    const double* q;   // [N]
    double* phi;       // [N] or null
    double* force;     // SoA [3N] or null
    int N, off, blk, pad;
};

// Output stores of the gathers.  Output pointers of a grouped solve are read from the
// descriptors, so the compiler cannot tell their address space and would emit flat stores;
// flat stores count against the LDS wait counter and turn the gathers' partial LDS waits
// into full ones.  The outputs are device memory: store through global pointers.
typedef __attribute__((address_space(1))) double global_double;
__device__ __forceinline__ void store_out(double* p, long i, double v) { ((global_double*)p)[i] = v; }

// Parameters of the reference-layout (legacy) kernels: SE_FGG_params + periodicity.
struct LParams {
    int N, P, P_half, dims[3], npdims[3];
    double c, d, h, a, b;
    int per;
};

// True when the plan holds a captured graph of se_solve with exactly these arguments
// (se_engine.hip; the batch executor routes a served job to a slot that has one).
bool plan_has_graph(const se_plan* pl, int N, const void* x, const void* q, const void* phi, const void* force);

// ---- tuned tilings (se_tune.hip) ---------------------------------------------------
// The tiling se_tune recorded for d's geometry and n particles per system, if any.
bool tuned_tiling(const se_plan_desc* d, long n_per_system, int& tile, int& chunk, int& block);

// ---- launchers (se_kernels.hip) ------------------------------------------------
void launch_fill_i32(int* p, long n, int v, hipStream_t s);
void launch_fill_bytes0(void* p, size_t bytes, hipStream_t s);

// ---- stable sort (se_sort.hip) ----------------------------------------------------
// Stable LSD radix sort of (k0, v0) by the low `bits` bits of the (non-negative) keys,
// using (k1, v1) as the second buffer pair; *keys_out / *vals_out point at the sorted keys
// and values (one of the two pairs).  inverse: *vals_out holds the inverse permutation
// instead, (*vals_out)[v] = sorted position of the pair with value v (values must be a
// permutation of [0, n)).  tmp holds sort_tmp_bytes(n) bytes or more.  Launches only
// kernels (no memset), so it can be graph-captured.
// Exclusive scan of n ints: one work-group for n <= SCAN_SMALL_MAX, else the library scan
// with its temporary area tmp (tmp_bytes, sized by hipcub for n).
constexpr int SCAN_SMALL_MAX = 32768;
void exclusive_scan(const int* in, int* out, int n, void* tmp, size_t tmp_bytes, hipStream_t s);
size_t sort_tmp_bytes(int n_cap);
void stable_sort_pairs(int* k0, int* v0, int* k1, int* v1, int n, int bits, bool inverse, void* tmp,
                       size_t tmp_bytes, hipStream_t s, const int** keys_out, int** vals_out);
// count of every key from sorted keys (count zero before): counters at count_slot(k, nslots)
// if nslots > 0, else at k
void launch_run_counts(const int* skey, int n, int nslots, int* count, hipStream_t s);

// bucketing: key[n] (tile, pencil, row; the discard key nbins past every tile) and
// idx[n] = n for the stable key sort; after the sort, the tile starts from the sorted keys
// (nt + 1 entries; K keys per tile)
template <typename T>
void launch_sort_keys(const double* x, int N, const Geom& g, const Tiling& tl, int* key, int* idx, hipStream_t s);
void launch_tile_starts(const int* skey, int N, int nt, int K, int* tile_start, hipStream_t s);
void launch_tile_chunks(const int* tile_start, int ntiles, int chunk, int* tile_count, int* nchunk, hipStream_t s);
void launch_items(const int* tile_count, const int* tile_start, const int* chunk_start, int ntiles, int chunk,
                  int* item_tile, int* item_beg, int* item_end, int* n_items, hipStream_t s);
template <typename T>
void launch_scatter(const double* x, const double* q, int N, const Geom& g, const Tiling& tl,
                    const int* pos, Rec<T>* rec, hipStream_t s);
// grouped forms: `blocks` = sum over systems of ceil(N/256); discard bucket S * nbins;
// idx / pos are indexed by the flat particle index off + n
template <typename T>
void launch_sort_keys_group(const SysDesc* sys, int S, int blocks, const Geom& g, const Tiling& tl, int* key,
                            int* idx, hipStream_t s);
template <typename T>
void launch_scatter_group(const SysDesc* sys, int S, int blocks, const Geom& g, const Tiling& tl, const int* pos,
                          Rec<T>* rec, hipStream_t s);
// pencil kernels (se_pencil.hip): FP32 only; return false if P has no compiled pencil kernel
bool pencil_supported(int P);
// pencil spread; where tl.plane is set, whole sparse tiles take the plane spread
// (tile_start: the tiles' first records)
// item_col: the column path's per-item table (launch_column_order), or nullptr.  side
// (with the events fork and join) non-null: the column spread runs on s and the other
// items' spreads on side, concurrently, and s waits for them.
bool launch_spread_pencil(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                          const int* n_items, const int* tile_start, const Geom& g, const Tiling& tl,
                          float* partials, const int* item_col, hipStream_t s, hipStream_t side = nullptr,
                          hipEvent_t fork = nullptr, hipEvent_t join = nullptr);
// sys != null: grouped solve, the outputs and N come from sys[tile / ntiles] and the
// grids are stacked; phi/force then only select the output mode.  side (with the events
// fork and join) non-null: where the items split between two kernels, the sparse items'
// kernel runs on side, concurrently with the other on s, and s waits for it.
bool launch_interp_pencil(const Rec<float>* rec, const int* item_tile, const int* item_beg,
                          const int* item_end, const int* n_items, const float* grid,
                          const Geom& g, const Tiling& tl, int N, double* phi, double* force,
                          const SysDesc* sys, const int* item_col, hipStream_t s, hipStream_t side = nullptr,
                          hipEvent_t fork = nullptr, hipEvent_t join = nullptr);
// Dense items: those the matrix-core spread and gather take; every other spread and gather
// kernel skips them.  On the column path (P = 16 FP32, tl.column) k_column_order decides per
// item and writes item_col[item]: COLUMN_NONE, or the fixed-point scale of the item's
// column spread sums; the other path with matrix cores (the MFMA gather) goes by size.
constexpr int COLUMN_MAX = 2048;                  // column items are staged in LDS: at most this many records
constexpr int COLUMN_NONE = -2147483647 - 1;      // item_col entry of an item the column kernels leave
constexpr float COLUMN_GROUP_FILL = 3.5f;         // column items: at least this many particles per column group
struct DenseSel {
    const int* col;   // per-item column table (column path) or nullptr
    int min;          // without a table: dense from min to COLUMN_MAX particles (0: none)
    __device__ __forceinline__ bool operator()(int item, int n) const
    {
        if (col) return col[item] != COLUMN_NONE;
        return min > 0 && n >= min && n <= COLUMN_MAX;
    }
};
// column path: the items considered for column order (from one particle per (pencil, row)
// key to COLUMN_MAX particles)
inline int column_floor(const Tiling& tl) { return tl.K; }
// Column order of the items from column_floor to COLUMN_MAX particles whose column groups
// hold COLUMN_GROUP_FILL particles on average: records into (pencil, lx mod 4, row)
// order, in place, and item_col[item] for every item (COLUMN_NONE for the others)
void launch_column_order(Rec<float>* rec, const int* item_beg, const int* item_end, const int* n_items, int max_items,
                         const Geom& g, const Tiling& tl, int* item_col, hipStream_t s);
// column gather (se_column.hip): the items of item_col, P = 16 FP32, work-groups of 256,
// 512 or 1024 threads
void launch_interp_col16(const Rec<float>* rec, const int* item_tile, const int* item_beg, const int* item_end,
                         const int* n_items, const float* grid, const Geom& g, const Tiling& tl, int N,
                         double* phi, double* force, const SysDesc* sys, const int* item_col, int block,
                         hipStream_t s);
// column spread (se_column.hip): the items of item_col, P = 16 FP32, work-groups of 256,
// 512 or 768 threads
bool launch_spread_col16(const Rec<float>* rec, const int* item_beg, const int* item_end, const int* n_items,
                         const Geom& g, const Tiling& tl, float* partials, const int* item_col, int block,
                         hipStream_t s);
template <typename T>
void launch_spread_tiles(const Rec<T>* rec, const int* item_tile, const int* item_beg,
                         const int* item_end, const int* n_items, const Geom& g,
                         const Tiling& tl, T* partials, hipStream_t s);
// merge segment tables: x and y share (seg, nseg), z has (segz, nsegz); both strides tl.nseg;
// nsys stacked systems (tile index system * ntiles + tile)
template <typename T>
void launch_merge(const T* partials, const int* chunk_start, const int4* seg, const int* nseg, const int4* segz,
                  const int* nsegz, int sstride, const Geom& g, const Tiling& tl, T* grid, int nsys,
                  hipStream_t s);
template <typename T>
void launch_interp_tiles(const Rec<T>* rec, const int* item_tile, const int* item_beg,
                         const int* item_end, const int* n_items, const T* grid, const Geom& g,
                         const Tiling& tl, int N, double* phi, double* force, const SysDesc* sys,
                         hipStream_t s);
// SE3P Green's function, separable tables: tab = [k_x^2 (M), e_x (M), k_z^2 (Mz), e_z (Mz), 4 pi/(M M Mz)]
template <typename T>
void launch_green(void* spec, int M, int Mz, const double* tab, int nsys, hipStream_t s);
void launch_green_tables(double* tab, int M, int Mz, double L, double Lz, double scalar, hipStream_t s);
// SE2P Green's function, full half-spectrum table (see se_kernels.hip)
void launch_green2p_table(double* tab, int M, int Mz, double L, double h, double R, double scalar,
                          hipStream_t s);
// SE1P Green's function, full half-spectrum table of the M x M (free) x Mz (periodic) grid;
// jtab = [J0(|k_par| R), J1(|k_par| R)] per (x, y) frequency pair (M*M each), ktab =
// [K0(k_z R), K1(k_z R)] per z frequency (Mz/2+1 each), c00 = the k = 0 value
void launch_green1p_table(double* tab, const double* jtab, const double* ktab, int M, int Mz, double h, double Lz,
                          double R, double scalar, double c00, hipStream_t s);
template <typename T>
void launch_green_full(void* spec, long n, const double* tab, int nsys, hipStream_t s);
void launch_energy(const double* phi, const double* q, int N, double* out, hipStream_t s);

// legacy (FP64, reference layout)
void launch_base_gaussian(const LParams& p, double* zs, hipStream_t s);
void launch_expand_all(const LParams& p, int force, const double* x, double* zx, double* zy,
                       double* zz, int* idx, double* zfx, double* zfy, double* zfz,
                       hipStream_t s);
void launch_grid_split_ext(const LParams& p, const double* q, const double* zx,
                           const double* zy, const double* zz, const int* idx,
                           const double* zs, double* H, hipStream_t s);
void launch_grid_ext(const LParams& p, const double* x, const double* q, const double* zs,
                     double* H, hipStream_t s);
void launch_int_split_ext(const LParams& p, int force, const double* H, const double* q,
                          const double* zx, const double* zy, const double* zz,
                          const double* zfx, const double* zfy, const double* zfz,
                          const int* idx, const double* zs, double* out, hipStream_t s);
void launch_int_ext(const LParams& p, int force, const double* H, const double* x,
                    const double* q, const double* zs, double* out, hipStream_t s);
void launch_wrap(const LParams& p, const double* H, double* Hper, hipStream_t s);
void launch_extend(const LParams& p, const double* Hper, double* H, hipStream_t s);
void launch_scaling(double scalar, double* Z, int n, double L, hipStream_t s);
void launch_product_rc(const double* a, const double* b, double* c, long n, int flag,
                       hipStream_t s);
void launch_reorder_gather(const double* xin, const double* qin, const int* perm, int N,
                           double* xout, double* qout, hipStream_t s);

}  // namespace se

// Stable LSD radix sort of (key, value) int pairs, used by the engine's bucketing (sort key
// = tile, pencil, row) and by the real-space cell lists: the order of equal keys is the
// input order, so the particle order inside a key does not depend on the timing of any
// atomic.
//
// Each pass sorts by one 8-bit digit in three launches:
//   k_sort_hist     per tile of 1024 to 4096 items, the count of each digit (LDS counters)
//   (scan)          exclusive scan of the digit-major table hist[digit * ntile + tile]
//   k_sort_scatter  per tile, each item's position = the scanned base of (digit, tile) +
//                   the number of earlier items of the tile with the same digit; the
//                   tile is staged in LDS in that order and written out from there.
// The scatter walks its tile in rounds of 256 items; inside a round, a wave finds the
// lanes that share a lane's digit with 8 ballots (one per digit bit) and takes the lane's
// rank among them from a popcount, the waves' per-digit counts are combined in LDS in wave
// order, and a per-digit running count carries the rounds.  No atomics order anything and
// nothing is cleared with a memset, so the sort can be captured in a HIP graph.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "se_internal.h"

namespace se {

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ROUNDS = 16;                         // most rounds of 256 items per tile
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;   // most items per tile
constexpr int SORT_MIN_ROUNDS = 4;
constexpr int SORT_RADIX = 256;

__global__ __launch_bounds__(SORT_THREADS) void k_sort_hist(const int* __restrict__ keys, int n, int shift, int ntile,
                                                            int rounds, int* __restrict__ hist)
{
    __shared__ int cnt[SORT_RADIX];
    const int tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * rounds * SORT_THREADS;
#pragma unroll 4
    for (int r = 0; r < rounds; ++r) {
        const long i = base + r * SORT_THREADS + tid;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (SORT_RADIX - 1)], 1);
    }
    __syncthreads();
    hist[(long)tid * ntile + blockIdx.x] = cnt[tid];
}

// kout / vout: sorted pairs.  inv != null (last pass): inv[v] = the item's sorted position
// instead of vout.  The tile's items are first placed in LDS in (digit, input) order --
// each digit's items start at the digit's offset within the tile (an exclusive scan of
// the tile's histogram) -- and then written out in LDS order, so that the items of one
// digit, which are consecutive in the output, leave in consecutive lanes.
__global__ __launch_bounds__(SORT_THREADS) void k_sort_scatter(const int* __restrict__ kin, const int* __restrict__ vin,
                                                               int n, int shift, int ntile, int rounds,
                                                               const int* __restrict__ hist,
                                                               const int* __restrict__ offs, int* __restrict__ kout,
                                                               int* __restrict__ vout, int* __restrict__ inv)
{
    __shared__ int lk[SORT_TILE], lv[SORT_TILE];  // the tile in (digit, input) order
    __shared__ int lofs[SORT_RADIX];              // first LDS slot of each digit
    __shared__ int gofs[SORT_RADIX];              // first output position of each digit
    __shared__ int run[SORT_RADIX];               // items of each digit placed so far
    __shared__ int wcnt[SORT_WAVES][SORT_RADIX];  // this round's count of each digit per wave
    __shared__ int wsum[SORT_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c = hist[(long)tid * ntile + blockIdx.x];
    gofs[tid] = offs[(long)tid * ntile + blockIdx.x];
    // exclusive scan of the tile's digit counts: wave prefix sums, then the waves' totals
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    run[tid] = 0;
#pragma unroll
    for (int v = 0; v < SORT_WAVES; ++v) wcnt[v][tid] = 0;
    __syncthreads();
    int pre = 0;
    for (int u = 0; u < w; ++u) pre += wsum[u];
    lofs[tid] = pre + incl - c;
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;   // lanes below this one
    const long base = (long)blockIdx.x * rounds * SORT_THREADS;
    for (int r = 0; r < rounds; ++r) {
        const long i = base + r * SORT_THREADS + tid;
        const bool ok = i < n;
        const int k = ok ? kin[i] : 0;
        const int v = ok ? vin[i] : 0;
        const int d = (k >> shift) & (SORT_RADIX - 1);
        // lanes of this wave with the same digit (valid lanes only)
        unsigned long long m = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bb = __ballot((d >> b) & 1);
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int below = __popcll(m & lt);
        if (ok && below == 0) wcnt[w][d] = __popcll(m);   // the group's lowest lane
        __syncthreads();
        if (ok) {
            int p = run[d] + below;   // rank among the tile's items of digit d
            for (int u = 0; u < w; ++u) p += wcnt[u][d];
            lk[lofs[d] + p] = k;
            if (inv)
                inv[v] = gofs[d] + p;
            else
                lv[lofs[d] + p] = v;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int u = 0; u < SORT_WAVES; ++u) {
            add += wcnt[u][tid];
            wcnt[u][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
    const int cnt = (int)min((long)rounds * SORT_THREADS, (long)n - base);
    for (int i = tid; i < cnt; i += SORT_THREADS) {
        const int k = lk[i], d = (k >> shift) & (SORT_RADIX - 1);
        const int p = gofs[d] + (i - lofs[d]);
        kout[p] = k;
        if (!inv) vout[p] = lv[i];
    }
}

// Exclusive scan of n <= SCAN_SMALL_MAX ints by one work-group: each thread sums a run of
// at most 32 consecutive entries, the run totals are scanned across the work-group (wave
// prefix sums, then the waves' totals), and each thread writes its run.  One launch where
// the library scan takes two, for the short tables of small solves.
constexpr int SCAN_SMALL_THREADS = 1024;
__global__ __launch_bounds__(SCAN_SMALL_THREADS) void k_scan_small(const int* __restrict__ in, int* __restrict__ out,
                                                                   int n)
{
    __shared__ int wsum[SCAN_SMALL_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int per = (n + SCAN_SMALL_THREADS - 1) / SCAN_SMALL_THREADS;
    const int b = min(tid * per, n), e = min(b + per, n);
    int t = 0;
    for (int i = b; i < e; ++i) t += in[i];
    int incl = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int run = incl - t;
    for (int u = 0; u < w; ++u) run += wsum[u];
    for (int i = b; i < e; ++i) {
        const int v = in[i];
        out[i] = run;
        run += v;
    }
}

}  // namespace

// Rounds of 256 items per tile: the most (up to 16) that still give two tiles per CU of
// a 256-CU device, so that a small sort is not one long work-group per CU.
static int sort_rounds(int n)
{
    int r = SORT_ROUNDS;
    while (r > SORT_MIN_ROUNDS && (n + (long)r * SORT_THREADS - 1) / ((long)r * SORT_THREADS) < 512) r /= 2;
    return r;
}

static int sort_tiles(int n, int rounds)
{
    return (int)((n + (long)rounds * SORT_THREADS - 1) / ((long)rounds * SORT_THREADS));
}

size_t sort_tmp_bytes(int n_cap)
{
    // the table of the smallest tiles, which every n <= n_cap fits
    const int nt = std::max(1, sort_tiles(std::max(n_cap, 1), SORT_MIN_ROUNDS));
    size_t b = 0;
    SE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (int*)nullptr, (int*)nullptr, SORT_RADIX * nt));
    const size_t hist = (size_t)2 * SORT_RADIX * nt * sizeof(int);
    return ((hist + 255) & ~(size_t)255) + b;
}

void exclusive_scan(const int* in, int* out, int n, void* tmp, size_t tmp_bytes, hipStream_t s)
{
    if (n <= 0) return;
    if (n <= SCAN_SMALL_MAX) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(SCAN_SMALL_THREADS), 0, s, in, out, n);
        return;
    }
    size_t b = tmp_bytes;
    SE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, b, in, out, n, s));
}

void stable_sort_pairs(int* k0, int* v0, int* k1, int* v1, int n, int bits, bool inverse, void* tmp,
                       size_t tmp_bytes, hipStream_t s, const int** keys_out, int** vals_out)
{
    *keys_out = k0;
    *vals_out = v0;
    if (n <= 0) return;
    const int rounds = sort_rounds(n), nt = sort_tiles(n, rounds);
    const size_t hist_bytes = (size_t)2 * SORT_RADIX * nt * sizeof(int);
    const size_t hist_pad = (hist_bytes + 255) & ~(size_t)255;
    SE_REQUIRE(tmp_bytes >= hist_pad, SE_ERR_CAPACITY, "sort temporary area too small");
    int* hist = reinterpret_cast<int*>(tmp);
    int* offs = hist + (size_t)SORT_RADIX * nt;
    void* scan_tmp = reinterpret_cast<char*>(tmp) + hist_pad;
    const int passes = (std::max(bits, 1) + 7) / 8;
    int *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const bool last = p == passes - 1;
        hipLaunchKernelGGL(k_sort_hist, dim3(nt), dim3(SORT_THREADS), 0, s, ki, n, shift, nt, rounds, hist);
        exclusive_scan(hist, offs, SORT_RADIX * nt, scan_tmp, tmp_bytes - hist_pad, s);
        // the last pass of an inverse sort writes the inverse permutation into the values
        // output (its sorted values are not needed)
        hipLaunchKernelGGL(k_sort_scatter, dim3(nt), dim3(SORT_THREADS), 0, s, ki, vi, n, shift, nt, rounds, hist, offs,
                           ko,
                           vo, (last && inverse) ? vo : nullptr);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    *keys_out = ki;   // the last pass's output
    *vals_out = vi;
}

// Per-key counts from the sorted keys: each run of equal keys [a, b) adds -a at its first
// and +b at its last item, two integer atomics per key, so the count b - a does not depend
// on their order; count must be zero before.
// nslots > 0: the counters sit at count_slot(k, nslots); 0: at k.
__global__ void k_run_counts(const int* __restrict__ skey, int n, int nslots, int* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = skey[i];
    const int c = nslots > 0 ? count_slot(k, nslots) : k;
    if (i == 0 || skey[i - 1] != k) atomicAdd(&count[c], -i);
    if (i == n - 1 || skey[i + 1] != k) atomicAdd(&count[c], i + 1);
}

void launch_run_counts(const int* skey, int n, int nslots, int* count, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_run_counts, dim3((n + 255) / 256), dim3(256), 0, s, skey, n, nslots, count);
}

}  // namespace se

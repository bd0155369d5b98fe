// Real-space and self terms of the Ewald sum on the device (se_gpu.h: se_real_*).
//
// The reference evaluates them by brute force over all pairs and image layers
// (SE_direct.c: SE3P_direct_real_rc :464-505, SE2P_direct_real_rc :224-254,
// SE1P_direct_real_rc :50-88, SE*P_direct_self :178-183, 340-345, 550-561).  Here the
// particles are bucketed into cells no smaller than the cutoff (counting sort), and one
// thread per particle visits the neighbouring cells: in a periodic direction every
// (cell, image) pair within reach once, so a box shorter than the cutoff adds its image
// layers; in a free direction only the cells that exist.  Pairs beyond rc are left out,
// as in the reference's _rc variants.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/se_gpu.h"
#include "se_internal.h"

namespace se {

struct RealGeom {
    int nc[3];      // cells per direction
    int reach[3];   // neighbour cells visited on each side
    int per[3];     // 1: periodic direction
    double L[3];    // period (periodic) or extent of the particles' range (free)
    double cs[3];   // cell side
    double xi, rc2, selfc;
    int ncells;
};

__device__ __forceinline__ bool real_cell(const RealGeom& g, int d, double v, double& w, int& c)
{
    if (!(fabs(v) < 1e300)) return false;
    if (g.per[d]) {
        w = v - floor(v / g.L[d]) * g.L[d];
        if (w >= g.L[d]) w = 0.0;   // v a tiny bit below a multiple of L rounds up
    } else {
        if (!(v >= 0.0 && v < g.L[d])) return false;
        w = v;
    }
    c = min((int)(w / g.cs[d]), g.nc[d] - 1);
    return true;
}

// Cell of every particle (discarded particles: cell ncells, sorted last).  The order inside
// a cell and the per-cell counts come from a stable radix sort of (cell, index) afterwards,
// so each particle's neighbour sum runs in the same order on every call.
__global__ void k_real_bin(const double* __restrict__ x, int N, RealGeom g, int* __restrict__ cell,
                           int* __restrict__ idx)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int c[3];
    double w;
    bool ok = true;
    for (int d = 0; d < 3; ++d) ok = real_cell(g, d, x[n + (long)d * N], w, c[d]) && ok;
    const int k = ok ? (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2] : g.ncells;
    cell[n] = k;
    idx[n] = n;
}

// Sorted position i holds particle sidx[i]: its folded position and charge.
__global__ void k_real_gather(const double* __restrict__ x, const double* __restrict__ q, int N, RealGeom g,
                              const int* __restrict__ scell, const int* __restrict__ sidx,
                              double4* __restrict__ pos, int* __restrict__ orig)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || scell[i] >= g.ncells) return;
    const int n = sidx[i];
    double w[3];
    int c;
    for (int d = 0; d < 3; ++d) real_cell(g, d, x[n + (long)d * N], w[d], c);
    pos[i] = make_double4(w[0], w[1], w[2], q[n]);
    orig[i] = n;
}

// One thread per bucketed particle.  phi_m = sum q_n erfc(xi r)/r (+ self term), and the
// force in the reference's convention F = (q/2) grad phi:
//   F_m = -(q_m/2) sum_n q_n [erfc(xi r)/r + (2 xi/sqrt(pi)) e^{-xi^2 r^2}] r_vec / r^2 / r,
// r_vec = x_m - (x_n + image shift).
__global__ void k_real_eval(const double4* __restrict__ pos, const int* __restrict__ orig,
                            const int* __restrict__ start, RealGeom g, int N, double* __restrict__ phi,
                            double* __restrict__ force, int accumulate)
{
    const int nsorted = start[g.ncells];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsorted) return;
    const double4 p = pos[i];
    const double pw[3] = {p.x, p.y, p.z};
    int ci[3];
    for (int d = 0; d < 3; ++d) ci[d] = min((int)(pw[d] / g.cs[d]), g.nc[d] - 1);
    const double tw = 2.0 * g.xi / sqrt(M_PI), xi2 = g.xi * g.xi;
    double sp = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int o0 = -g.reach[0]; o0 <= g.reach[0]; ++o0) {
        int c0 = ci[0] + o0;
        double s0 = 0.0;
        if (g.per[0]) {
            const int im = (c0 >= 0) ? c0 / g.nc[0] : -((-c0 + g.nc[0] - 1) / g.nc[0]);
            c0 -= im * g.nc[0];
            s0 = im * g.L[0];
        } else if (c0 < 0 || c0 >= g.nc[0]) {
            continue;
        }
        for (int o1 = -g.reach[1]; o1 <= g.reach[1]; ++o1) {
            int c1 = ci[1] + o1;
            double s1 = 0.0;
            if (g.per[1]) {
                const int im = (c1 >= 0) ? c1 / g.nc[1] : -((-c1 + g.nc[1] - 1) / g.nc[1]);
                c1 -= im * g.nc[1];
                s1 = im * g.L[1];
            } else if (c1 < 0 || c1 >= g.nc[1]) {
                continue;
            }
            for (int o2 = -g.reach[2]; o2 <= g.reach[2]; ++o2) {
                int c2 = ci[2] + o2;
                double s2 = 0.0;
                if (g.per[2]) {
                    const int im = (c2 >= 0) ? c2 / g.nc[2] : -((-c2 + g.nc[2] - 1) / g.nc[2]);
                    c2 -= im * g.nc[2];
                    s2 = im * g.L[2];
                } else if (c2 < 0 || c2 >= g.nc[2]) {
                    continue;
                }
                const bool same_image = s0 == 0.0 && s1 == 0.0 && s2 == 0.0;
                const int cc = (c0 * g.nc[1] + c1) * g.nc[2] + c2;
                const int je = start[cc + 1];
                for (int j = start[cc]; j < je; ++j) {
                    if (same_image && j == i) continue;
                    const double4 o = pos[j];
                    const double d0 = p.x - (o.x + s0), d1 = p.y - (o.y + s1), d2 = p.z - (o.z + s2);
                    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                    if (r2 > g.rc2) continue;
                    const double r = sqrt(r2), ir = 1.0 / r, e = erfc(g.xi * r);
                    sp += o.w * e * ir;
                    if (force) {
                        const double t = o.w * (e * ir + tw * exp(-xi2 * r2)) * ir * ir;
                        f0 += t * d0;
                        f1 += t * d1;
                        f2 += t * d2;
                    }
                }
            }
        }
    }
    const int n = orig[i];
    if (phi) {
        const double v = sp - g.selfc * p.w;
        phi[n] = accumulate ? phi[n] + v : v;
    }
    if (force) {
        const double h = -0.5 * p.w;
        if (accumulate) {
            force[n] += h * f0;
            force[(long)n + N] += h * f1;
            force[(long)n + 2l * N] += h * f2;
        } else {
            force[n] = h * f0;
            force[(long)n + N] = h * f1;
            force[(long)n + 2l * N] = h * f2;
        }
    }
}

// Device buffers of one real-space solve for up to n particles and nc cells.
struct RealBufs {
    int *cell, *idx, *scell, *sidx, *count, *start, *orig;
    double4* pos;
    void* tmp;          // temporary area of the count scan and the cell sort
    size_t tmp_bytes;
};

static int cell_bits(int ncells)
{
    int b = 1;
    while ((1l << b) <= ncells) ++b;
    return b;
}

// Temporary bytes the scan over nc + 1 counts and the sort of n pairs need (host-side
// queries only: nothing is allocated or launched).
static size_t real_tmp_bytes(int n, int nc, hipStream_t s)
{
    size_t a = 0;
    SE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, a, (int*)nullptr, (int*)nullptr, nc + 1, s));
    return std::max(a, sort_tmp_bytes(n));
}

// The solve on given buffers: bin, stable sort, counts and their scan, gather, evaluate.
static void real_run(const RealGeom& g, const RealBufs& b, int N, const double* x, const double* q, double* phi,
                     double* force, int accumulate, hipStream_t s)
{
    const int nb = (N + 255) / 256;
    launch_fill_i32(b.count, (long)g.ncells + 1, 0, s);
    hipLaunchKernelGGL(k_real_bin, dim3(nb), dim3(256), 0, s, x, N, g, b.cell, b.idx);
    const int* scell = b.cell;
    int* sidx = b.idx;
    stable_sort_pairs(b.cell, b.idx, b.scell, b.sidx, N, cell_bits(g.ncells), false, b.tmp, b.tmp_bytes, s, &scell,
                      &sidx);
    launch_run_counts(scell, N, 0, b.count, s);
    size_t tb = b.tmp_bytes;
    SE_HIP(hipcub::DeviceScan::ExclusiveSum(b.tmp, tb, b.count, b.start, g.ncells + 1, s));
    hipLaunchKernelGGL(k_real_gather, dim3(nb), dim3(256), 0, s, x, q, N, g, scell, sidx, b.pos, b.orig);
    hipLaunchKernelGGL(k_real_eval, dim3(nb), dim3(256), 0, s, b.pos, b.orig, b.start, g, N, phi, force, accumulate);
    SE_HIP(hipGetLastError());
}

}  // namespace se

using namespace se;

struct se_real_plan {
    se_real_desc desc{};
    RealGeom g{};
    int n_cap = 0;
    hipStream_t stream = nullptr;
    RealBufs b{};
    std::vector<void*> allocs;
    void* alloc(size_t b)
    {
        void* p = nullptr;
        SE_HIP(hipMalloc(&p, std::max<size_t>(b, 16)));
        allocs.push_back(p);
        return p;
    }
    ~se_real_plan()
    {
        for (void* p : allocs) (void)hipFree(p);
    }
};

namespace {

template <typename F>
int guarded_real(F&& f)
{
    try {
        f();
        return SE_OK;
    } catch (const se::Error& e) {
        se::set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        se::set_error(e.what());
        return SE_ERR_ARG;
    }
}

// Cells: in each direction floor(L / rc) cells (at least one, and at most about
// (4 n_max)^(1/3) so the cell table stays near the particle count); a cell side below
// the cutoff (L < rc, periodic) makes the neighbour reach ceil(rc / side) > 1.
RealGeom real_geometry(const se_real_desc* d, int n_max)
{
    SE_REQUIRE(d->periodicity == 3 || d->periodicity == 2 || d->periodicity == 1, SE_ERR_ARG,
               "se_real: periodicity must be 3, 2 or 1");
    SE_REQUIRE(d->xi > 0 && d->rc > 0 && d->box[0] > 0 && d->box[1] > 0 && d->box[2] > 0, SE_ERR_ARG,
               "se_real: need xi > 0, rc > 0 and a positive box");
    RealGeom g{};
    const int cap = std::max(1, (int)std::cbrt(4.0 * std::max(n_max, 1)));
    for (int k = 0; k < 3; ++k) {
        g.per[k] = d->periodicity == 3 || (d->periodicity == 2 && k < 2) || (d->periodicity == 1 && k == 2);
        g.L[k] = d->box[k];
        g.nc[k] = std::max(1, std::min(cap, (int)std::floor(d->box[k] / d->rc)));
        g.cs[k] = d->box[k] / g.nc[k];
        g.reach[k] = g.per[k] ? std::max(1, (int)std::ceil(d->rc / g.cs[k] - 1e-12)) : 1;
        if (!g.per[k] && g.nc[k] == 1) g.reach[k] = 0;
    }
    g.ncells = g.nc[0] * g.nc[1] * g.nc[2];
    g.xi = d->xi;
    g.rc2 = d->rc * d->rc;
    g.selfc = d->self ? 2.0 * d->xi / std::sqrt(M_PI) : 0.0;
    return g;
}

}  // namespace

extern "C" {

int se_real_plan_create(se_real_plan** out, const se_real_desc* d, int n_max, void* stream)
{
    if (!out || !d) return SE_ERR_ARG;
    *out = nullptr;
    se_real_plan* pl = new se_real_plan();
    int rc = guarded_real([&] {
        SE_REQUIRE(n_max >= 0, SE_ERR_ARG, "se_real: n_max must be >= 0");
        pl->desc = *d;
        pl->g = real_geometry(d, n_max);
        pl->n_cap = std::max(n_max, 1);
        pl->stream = (hipStream_t)stream;
        const size_t N = pl->n_cap, nc = pl->g.ncells;
        RealBufs& b = pl->b;
        b.cell = (int*)pl->alloc(N * 4);
        b.idx = (int*)pl->alloc(N * 4);
        b.scell = (int*)pl->alloc(N * 4);
        b.sidx = (int*)pl->alloc(N * 4);
        b.orig = (int*)pl->alloc(N * 4);
        b.pos = (double4*)pl->alloc(N * sizeof(double4));
        b.count = (int*)pl->alloc((nc + 1) * 4);
        b.start = (int*)pl->alloc((nc + 2) * 4);
        b.tmp_bytes = real_tmp_bytes((int)N, (int)nc, pl->stream);
        b.tmp = pl->alloc(b.tmp_bytes);
    });
    if (rc != SE_OK) {
        delete pl;
        return rc;
    }
    *out = pl;
    return SE_OK;
}

int se_real_plan_destroy(se_real_plan* pl)
{
    if (!pl) return SE_OK;
    int rc = guarded_real([&] { SE_HIP(hipStreamSynchronize(pl->stream)); });
    delete pl;
    return rc;
}

int se_real_solve(se_real_plan* pl, int N, const double* x, const double* q, double* phi, double* force)
{
    if (!pl) return SE_ERR_ARG;
    return guarded_real([&] {
        SE_REQUIRE(N >= 0 && N <= pl->n_cap, SE_ERR_CAPACITY, "N exceeds the real-space plan's capacity");
        SE_REQUIRE(N == 0 || (x && q), SE_ERR_ARG, "null particle arrays");
        if (N == 0 || (!phi && !force)) return;
        real_run(pl->g, pl->b, N, x, q, phi, force, pl->desc.accumulate, pl->stream);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Plan-free, allocation-free real-space solve (se_real_solve_noalloc).  The buffers that a
// real-space plan allocates -- cells, sort pairs, counts, cell starts, sorted positions and
// the scan / sort temporary area -- are carved from one file-scope device array of
// SE_REAL_SCRATCH_BYTES per GPU.  Host code never uses the array's name as an address:
// its device address is looked up once per device with hipGetSymbolAddress, with that
// device current, and kept per device.  Every call on one GPU shares the array, so calls
// that overlap on one GPU race; the count table is cleared by a kernel at the start of
// each call, every other buffer is written before it is read.
// ---------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SE_REAL_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up on the first call there).
unsigned char* scratch_on_current_device()
{
    int dev = 0;
    SE_HIP(hipGetDevice(&dev));
    SE_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, SE_ERR_ARG, "se_real_solve_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SE_REQUIRE(p != nullptr, SE_ERR_HIP, "se_real_solve_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int se_real_solve_noalloc(const se_real_desc* d, int N, const double* x, const double* q, double* phi, double* force,
                          void* stream)
{
    if (!d) return SE_ERR_ARG;
    return guarded_real([&] {
        SE_REQUIRE(N >= 0, SE_ERR_ARG, "se_real_solve_noalloc: N must be >= 0");
        SE_REQUIRE(N == 0 || (x && q), SE_ERR_ARG, "null particle arrays");
        const RealGeom g = real_geometry(d, N);
        if (N == 0 || (!phi && !force)) return;
        hipStream_t s = (hipStream_t)stream;
        // layout of the scratch area (the same buffers a plan of capacity N holds)
        const size_t n = (size_t)N, nc = (size_t)g.ncells;
        const size_t tmp_bytes = real_tmp_bytes(N, g.ncells, s);   // host-side size queries only
        const size_t sz[9] = {n * 4, n * 4, n * 4, n * 4, n * 4, n * sizeof(double4), (nc + 1) * 4, (nc + 2) * 4,
                              tmp_bytes};
        size_t off[9], total = 0;
        for (int i = 0; i < 9; ++i) {
            off[i] = total;
            total += align256(sz[i]);
        }
        if (total > SE_REAL_SCRATCH_BYTES)
            throw se::Error{SE_ERR_CAPACITY, "se_real_solve_noalloc: N = " + std::to_string(N) + " with " +
                                                 std::to_string(nc) + " cells needs " + std::to_string(total) +
                                                 " bytes of scratch, more than its SE_REAL_SCRATCH_BYTES (" +
                                                 std::to_string((size_t)SE_REAL_SCRATCH_BYTES) + ")"};
        unsigned char* base = scratch_on_current_device();
        RealBufs b{};
        b.cell = reinterpret_cast<int*>(base + off[0]);
        b.idx = reinterpret_cast<int*>(base + off[1]);
        b.scell = reinterpret_cast<int*>(base + off[2]);
        b.sidx = reinterpret_cast<int*>(base + off[3]);
        b.orig = reinterpret_cast<int*>(base + off[4]);
        b.pos = reinterpret_cast<double4*>(base + off[5]);
        b.count = reinterpret_cast<int*>(base + off[6]);
        b.start = reinterpret_cast<int*>(base + off[7]);
        b.tmp = base + off[8];
        b.tmp_bytes = tmp_bytes;
        real_run(g, b, N, x, q, phi, force, d->accumulate, s);
    });
}

}  // extern "C"

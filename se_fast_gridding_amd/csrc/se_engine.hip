// Device-resident SE3P engine (se_gpu.h): plan = grid geometry + tiling + HBM
// scratch + rocFFT plans; stages spread / kspace / interp on one HIP stream.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/se_gpu.h"
#include "se_internal.h"

namespace se {

static thread_local std::string g_last_error;
static std::mutex g_err_mu;
static std::string g_last_error_global;

void set_error(const std::string& msg)
{
    g_last_error = msg;
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error_global = msg;
}

const char* last_error()
{
    if (!g_last_error.empty()) return g_last_error.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_error_global.c_str();
}

static void rocfft_init_once()
{
    static std::once_flag f;
    std::call_once(f, [] { rocfft_setup(); });
}

#define SE_FFT(expr)                                                                             \
    do {                                                                                         \
        rocfft_status s_ = (expr);                                                               \
        if (s_ != rocfft_status_success)                                                         \
            throw ::se::Error{SE_ERR_FFT, std::string(#expr) + " failed (" + std::to_string((int)s_) + ")"}; \
    } while (0)

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
};

// Tiling rule.  Spread sub-grids accumulate in FP64 LDS (S^3 * 8 bytes), gather
// sub-grids hold the grid precision.  Default tile side: the largest (<= 16; a multiple
// of 4 and <= 12 for the pencil kernels) whose FP64 sub-grid fits in LDS.
static Tiling choose_tiling(int esize, int P, int M, int Mz, int n_cap, int n_sys, int tile, int chunk, int block,
                            int lds_max, bool allow_pencil, bool allow_mfma, bool allow_plane,
                            bool allow_column)
{
    Tiling tl{};
    tl.pencil = (allow_pencil && esize == 4 && pencil_supported(P)) ? 1 : 0;
    tl.mfma = (tl.pencil && allow_mfma && P == 16) ? 1 : 0;
    const int tmax = tl.pencil ? 12 : 16;
    int TS = tile;
    if (TS <= 0) {
        TS = 0;
        for (int t = tmax; t >= 1; t -= (tl.pencil ? 4 : 1)) {
            long S = t + P - 1;
            if (S * S * S * 8 <= lds_max) { TS = t; break; }
        }
        SE_REQUIRE(TS > 0, SE_ERR_ARG, "support P too large for the LDS sub-grid");
    }
    const int Mmin = std::min(M, Mz);
    TS = std::min(TS, std::min(Mmin, 16));   // the merge keeps a tile's sums in LDS: at most 16^3
    if (tl.pencil) TS -= TS % 4;   // pencils are 4 columns wide in x: tile side a multiple of 4
    if (tl.pencil && TS < 4) {
        tl.pencil = 0;
        tl.mfma = 0;
        TS = std::max(1, std::min(tile > 0 ? tile : Mmin, Mmin));
    }
    SE_REQUIRE(TS >= 1 && TS <= 16, SE_ERR_ARG, "tile side out of range");
    tl.TS = TS;
    tl.NT = (M + TS - 1) / TS;
    tl.NTz = (Mz + TS - 1) / TS;
    tl.S = TS + P - 1;
    tl.Szp = tl.S;
    SE_REQUIRE((long)tl.S * tl.S * tl.Szp * 8 <= lds_max, SE_ERR_ARG, "tile sub-grid exceeds LDS");
    tl.ntiles = tl.NT * tl.NT * tl.NTz;
    tl.PX = (TS + 3) / 4;
    tl.K = tl.PX * TS * TS;
    SE_REQUIRE((long)tl.ntiles * tl.K < (1l << 30), SE_ERR_ARG, "grid too large for the sort keys");
    tl.nbins = tl.ntiles * tl.K;
    // Default work-item cap: heavy tiles are cut into items of 1664 particles.  Clustered
    // N = 4M at M = 256 (updates/s, with the column kernels): 1024 9.75e12, 1280 1.005e13,
    // 1536 1.023e13, 1664 1.027e13, 1792 1.026e13, 2048 1.009e13 -- larger items write fewer
    // partials for the merge (0.43 -> 0.40 ms from 1280 to 1664) until the spread's items
    // balance worse over the CUs.
    tl.chunk = chunk > 0 ? chunk : 1664;
    tl.block = block > 0 ? block : 512;
    SE_REQUIRE(tl.block % 64 == 0 && tl.block <= 512, SE_ERR_ARG, "block must be a multiple of 64, <= 512");
    tl.sub_elems = tl.S * tl.S * tl.Szp;
    tl.part_elems = tl.S * tl.S * tl.S;
    tl.max_items = tl.ntiles * n_sys + n_cap / tl.chunk + 1;
    // The MFMA gather holds an FP32 sub-grid (half the spread's), so two work-groups of
    // at most 256 threads share a CU and one's tile load overlaps the other's particles;
    // it runs two waves per SIMD either way, limited by registers.  (The vector gathers
    // measured fastest at 512.)
    tl.mblock = std::min(tl.block, 256);
    // The pencil spread's FP64 sub-grid allows one or two work-groups per CU; at its
    // bound (768 threads for P = 16, 12 waves; 1024 for P = 8 and 11) it hides more
    // latency than at 512.  An explicit block applies to every kernel.
    tl.sblock = block > 0 ? block : spread_pencil_max_threads(P);
    tl.block_auto = block > 0 ? 0 : 1;
    // plane spread of sparse tiles: one lane per (x, z) column of the sub-grid, 32 z columns
    // per wave (S <= 32), its y values in registers (tile side <= 12)
    tl.plane = (tl.pencil && tl.TS <= 12 && tl.S <= 32 && allow_plane) ? 1 : 0;
    // column kernels for the dense items (matrix-core gather only)
    tl.column = (tl.mfma && allow_column) ? 1 : 0;
    // every launch shape must lie within its kernel's __launch_bounds__: refused here, at
    // plan creation, rather than by the launch guards at solve time
    SE_REQUIRE(tl.sblock <= spread_pencil_max_threads(P) && tl.mblock <= 512 && tl.block <= 512, SE_ERR_ARG,
               "work-group size past a kernel's launch bound");
    return tl;
}

}  // namespace se

using namespace se;

struct se_plan {
    se_plan_desc desc{};
    int esize = 8;
    Geom g{};
    Tiling tl{};
    double L = 1, xi = 1, m = 1, eta = 0;
    double R2p = 0;   // SE2P / SE1P kernel truncation radius
    hipStream_t stream = nullptr;
    // the gather's second stream (plans whose gather splits items between two kernels)
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    int n_cap = 0;
    int last_N = -1;
    // grouped solves: up to s_cap systems of this geometry stacked in one pass
    int s_cap = 1;
    int last_S = 1;                 // systems of the last spread
    bool last_group = false;        // the last spread was grouped (the gather reads d_sys)
    SysDesc* d_sys = nullptr;       // device descriptors of the last grouped spread
    SysDesc* h_sys[4] = {};         // pinned staging ring for the descriptors
    hipEvent_t h_ev[4] = {};        // staging buffer i is free once h_ev[i] completed
    int h_next = 0;
    struct GroupFft {
        int S;
        rocfft_plan fwd, bwd;
        rocfft_execution_info info;
        unsigned long last_use;
    };
    std::vector<GroupFft> gffts;   // batched transforms, one pair per group size
    unsigned long gfft_clock = 0;
    // device buffers
    // key / idx and key2 / idx2: the two (sort key, particle index) buffer pairs of the
    // stable key sort; sort_tmp: its histograms and scan area; tile_start: each tile's
    // first sorted record
    int *key = nullptr, *idx = nullptr, *key2 = nullptr, *idx2 = nullptr, *tile_start = nullptr, *tile_count = nullptr,
        *nchunk = nullptr, *chunk_start = nullptr, *item_tile = nullptr, *item_beg = nullptr,
        *item_end = nullptr, *n_items = nullptr, *nseg = nullptr, *nsegz = nullptr,
        *item_col = nullptr;   // column path: per-item column table (COLUMN_NONE or the fixed-point scale)
    int4 *seg = nullptr, *segz = nullptr;
    void *rec = nullptr, *partials = nullptr, *grid = nullptr, *spec = nullptr, *cub_tmp = nullptr,
         *sort_tmp = nullptr, *fft_work = nullptr;
    double* gtab = nullptr;
    size_t cub_bytes = 0, sort_bytes = 0, fft_work_bytes = 0, bytes = 0;
    rocfft_plan fwd = nullptr, bwd = nullptr;
    rocfft_execution_info info = nullptr;
    bool timing = false;
    hipEvent_t ev[8] = {};
    std::vector<void*> allocs;
    // captured solves, keyed by their arguments (see se_solve_graph)
    struct GraphEntry {
        int N;
        const void *x, *q, *phi, *force;
        hipGraphExec_t exec;
        unsigned long last_use;
        // state the captured solve leaves behind, restored on every replay so that a later
        // se_interp / se_plan_info sees this solve and not the last eager call
        int last_N, last_S;
        bool last_group;
    };
    std::vector<GraphEntry> graphs;
    unsigned long graph_clock = 0;

    void* alloc(size_t b)
    {
        void* p = nullptr;
        if (b == 0) b = 16;
        SE_HIP(hipMalloc(&p, b));
        allocs.push_back(p);
        bytes += b;
        return p;
    }
    ~se_plan()
    {
        for (auto& gph : graphs) (void)hipGraphExecDestroy(gph.exec);
        for (auto& f : gffts) {
            rocfft_plan_destroy(f.fwd);
            rocfft_plan_destroy(f.bwd);
            rocfft_execution_info_destroy(f.info);
        }
        for (int i = 0; i < 4; ++i) {
            if (h_ev[i]) (void)hipEventDestroy(h_ev[i]);
            if (h_sys[i]) (void)hipHostFree(h_sys[i]);
        }
        if (fwd) rocfft_plan_destroy(fwd);
        if (bwd) rocfft_plan_destroy(bwd);
        if (info) rocfft_execution_info_destroy(info);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (side) (void)hipStreamDestroy(side);
        for (void* p : allocs) (void)hipFree(p);
    }
    void rec_ev(int i)
    {
        if (timing) SE_HIP(hipEventRecord(ev[i], stream));
    }
};

namespace {

template <typename F>
int guarded(F&& f)
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

// Smallest n >= nmin whose prime factors are 2, 3, 5 and 7 (sizes rocFFT handles directly).
int fft_size(int nmin)
{
    for (int n = std::max(nmin, 1);; ++n) {
        int r = n;
        for (int f : {2, 3, 5, 7})
            while (r % f == 0) r /= f;
        if (r == 1) return n;
    }
}

void derive(const se_plan_desc* d, Geom& g, double& eta, double& m, double& R2p)
{
    SE_REQUIRE(d->periodicity == 3 || d->periodicity == 2 || d->periodicity == 1, SE_ERR_ARG,
               "se_plan: periodicity must be 3 (SE3P), 2 (SE2P) or 1 (SE1P)");
    SE_REQUIRE(d->M >= 1 && d->P >= 1 && d->P <= 32 && d->box > 0 && d->xi > 0, SE_ERR_ARG,
               "invalid plan description (need M>=1, 1<=P<=32, box>0, xi>0)");
    m = d->m > 0 ? d->m : 0.95 * std::sqrt(M_PI * d->P);
    // parse_params (SE_general.c:16-29)
    const double h = d->box / (double)d->M;
    const double w = (double)d->P * h / 2.0;
    eta = (2.0 * w * d->xi / m) * (2.0 * w * d->xi / m);
    const double c = 2.0 * d->xi * d->xi / eta;
    g.M = d->M;
    g.P = d->P;
    g.odd = d->P & 1;
    const int ph = g.odd ? (d->P - 1) / 2 : d->P / 2;
    g.first = g.odd ? -ph : -ph + 1;
    g.h = h;
    g.c = c;
    g.d = std::pow(c / M_PI, 1.5);
    g.per = d->periodicity;
    g.x0 = 0.0;
    g.xlo = -HUGE_VAL;
    g.xhi = HUGE_VAL;
    if (g.per == 1) {
        // SE1P: z periodic with M points; x and y free, grid origin a + h/2 with
        // a = -(P_half + 1) h in both, so every support window of a particle with
        // 0 <= x, y < slab starts at index >= 0.  The in-plane kernel truncation radius R
        // covers the diagonal of the charged and read grid points plus the Gaussian
        // screening tail; the padded period holds R plus their span D.
        SE_REQUIRE(d->slab > 0, SE_ERR_ARG, "SE1P plan needs slab > 0 (the side of the x-y square)");
        g.Mz = d->M;
        g.z0 = 0.0;
        g.zlo = -HUGE_VAL;
        g.zhi = HUGE_VAL;
        g.x0 = -(double)(ph + 1) * h + 0.5 * h;
        g.xlo = 0.0;
        g.xhi = d->slab;
        R2p = std::sqrt(2.0) * (d->slab + (double)(d->P + 1) * h) + 6.0 / d->xi;
        const double D = d->slab + (double)(d->P + 2 * ph + 2) * h;
        const int nmin = (int)std::ceil((R2p + D) / h) + 1;
        g.M = d->Mz > 0 ? d->Mz : fft_size(nmin);
        SE_REQUIRE(g.M >= nmin, SE_ERR_ARG, "SE1P x-y grid (Mz) too small for the square, support and kernel radius");
        return;
    }
    if (g.per == 3) {
        g.Mz = d->M;
        g.z0 = 0.0;
        g.zlo = -HUGE_VAL;
        g.zhi = HUGE_VAL;
        R2p = 0.0;
        return;
    }
    // SE2P: z grid origin a + h/2 with a = -(P_half + 1) h keeps every support window
    // of a particle with 0 <= z < slab at index >= 0; the kernel truncation radius R
    // covers the slab, the support and the Gaussian screening tail, and the padded z
    // period holds R plus the span D of grid points that carry charge or are read back.
    SE_REQUIRE(d->slab > 0, SE_ERR_ARG, "SE2P plan needs slab > 0");
    g.z0 = -(double)(ph + 1) * h + 0.5 * h;
    g.zlo = 0.0;
    g.zhi = d->slab;
    R2p = d->slab + d->P * h + 6.0 / d->xi;
    const double D = d->slab + (double)(d->P + 2 * ph + 2) * h;
    const int nmin = (int)std::ceil((R2p + D) / h) + 1;
    g.Mz = d->Mz > 0 ? d->Mz : fft_size(nmin);
    SE_REQUIRE(g.Mz >= nmin, SE_ERR_ARG, "SE2P z grid (Mz) too small for the slab, support and kernel radius");
}

// Merge segment tables, one per dimension (the three are identical for a cubic grid):
// for output tile coordinate t, the runs of sub-grid coordinates l in [0, S) of source
// tile u whose periodic image (u*TS + l) mod M falls inside [t*TS, t*TS + TS), as
// (u, first l, first output offset, run length).  A source tile can contribute several
// runs when its sub-grid wraps around a small grid.
void merge_segments(int M, int TS, int NT, int S, std::vector<std::vector<int4>>& lists)
{
    lists.assign(NT, {});
    for (int t = 0; t < NT; ++t) {
        const int o0 = t * TS, len = std::min(TS, M - o0);
        for (int u = 0; u < NT; ++u) {
            int run_l = -1, run_o = -1, run_n = 0;
            for (int l = 0; l < S; ++l) {
                const int o = (u * TS + l) % M - o0;
                const bool in = o >= 0 && o < len;
                if (in && run_n > 0 && o == run_o + run_n && l == run_l + run_n) {
                    ++run_n;
                    continue;
                }
                if (run_n > 0) lists[t].push_back(make_int4(u, run_l, run_o, run_n));
                run_n = 0;
                if (in) {
                    run_l = l;
                    run_o = o;
                    run_n = 1;
                }
            }
            if (run_n > 0) lists[t].push_back(make_int4(u, run_l, run_o, run_n));
        }
    }
}

void build_merge_segments(se_plan* pl)
{
    const int TS = pl->tl.TS, S = pl->tl.S;
    std::vector<std::vector<int4>> lx, lz;
    merge_segments(pl->g.M, TS, pl->tl.NT, S, lx);
    merge_segments(pl->g.Mz, TS, pl->tl.NTz, S, lz);
    int nseg = 1;
    for (auto& v : lx) nseg = std::max<int>(nseg, (int)v.size());
    for (auto& v : lz) nseg = std::max<int>(nseg, (int)v.size());
    pl->tl.nseg = nseg;
    auto upload = [&](const std::vector<std::vector<int4>>& lists, int4*& seg, int*& cnt) {
        std::vector<int4> flat(lists.size() * nseg, make_int4(0, 0, 0, 0));
        std::vector<int> c(lists.size());
        for (size_t t = 0; t < lists.size(); ++t) {
            c[t] = (int)lists[t].size();
            for (size_t a = 0; a < lists[t].size(); ++a) flat[t * nseg + a] = lists[t][a];
        }
        seg = (int4*)pl->alloc(flat.size() * sizeof(int4));
        cnt = (int*)pl->alloc(c.size() * sizeof(int));
        SE_HIP(hipMemcpy(seg, flat.data(), flat.size() * sizeof(int4), hipMemcpyHostToDevice));
        SE_HIP(hipMemcpy(cnt, c.data(), c.size() * sizeof(int), hipMemcpyHostToDevice));
    };
    upload(lx, pl->seg, pl->nseg);
    upload(lz, pl->segz, pl->nsegz);
}

// Modified Bessel functions K0 and K1 from K_nu(x) = int_0^inf e^{-x cosh t} cosh(nu t) dt
// by the trapezoidal rule, which converges like e^{-pi^2/dt} relative to e^{x} K_nu(x) for
// this integrand (analytic in |Im t| < pi/2), so dt = pi^2/(x + 40) gives full FP64
// accuracy; the sum runs until e^{-x cosh t} underflows.  Zero for x > 700 and for
// x <= 0 (where K0 and K1 are infinite; the k_z = 0 mode does not use them).
static void bessel_k01(double x, double& k0, double& k1)
{
    k0 = k1 = 0.0;
    if (!(x > 0.0) || x > 700.0) return;
    const double dt = std::min(0.1, M_PI * M_PI / (x + 40.0)) / 1.25;
    const double tmax = std::acosh(std::max(1.0, 745.0 / x)) + dt;
    double s0 = 0.5 * std::exp(-x), s1 = s0;
    for (double t = dt; t <= tmax; t += dt) {
        const double e = std::exp(-x * std::cosh(t));
        s0 += e;
        s1 += e * std::cosh(t);
    }
    k0 = s0 * dt;
    k1 = s1 * dt;
}

// SE1P Green's function table: Bessel J0, J1 of |k_par| R per in-plane frequency pair and
// K0, K1 of k_z R per z frequency on the host (FP64, libm j0/j1), the table on the device.
// The k = 0 value is the truncated log kernel's -pi R^2 (2 ln R - 1) shifted by
// -(M h)^2 (gamma + ln xi^2): the FFT convolution then reproduces the reference's k_z = 0
// term -(1/L) sum_n q_n (gamma + ln(rho^2 xi^2) + E1(rho^2 xi^2)) (SE_direct.c:158-176)
// exactly, also for systems with a net charge.
static void build_green1p(se_plan* pl, double scalar)
{
    const int M = pl->g.M, Mz = pl->g.Mz, Mh = Mz / 2 + 1;
    const double h = pl->g.h, R = pl->R2p, Lam = M * h, Lz = Mz * h;
    std::vector<double> jt(2 * (size_t)M * M), kt(2 * (size_t)Mh);
    auto freq = [](int i, int n, double period) {
        const int pos = (n % 2 == 0) ? n / 2 : (n - 1) / 2 + 1;
        return 2.0 * M_PI / period * (double)(i < pos ? i : i - n);
    };
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            const double kx = freq(i, M, Lam), ky = freq(j, M, Lam), kr = std::sqrt(kx * kx + ky * ky) * R;
            jt[(size_t)i * M + j] = ::j0(kr);
            jt[(size_t)M * M + (size_t)i * M + j] = ::j1(kr);
        }
    for (int k = 0; k < Mh; ++k) bessel_k01(2.0 * M_PI / Lz * k * R, kt[k], kt[Mh + k]);
    const double euler = 0.57721566490153286061;
    const double c00 = -M_PI * R * R * (2.0 * std::log(R) - 1.0) - Lam * Lam * (euler + std::log(pl->xi * pl->xi));
    double* dj = (double*)pl->alloc(jt.size() * sizeof(double));
    double* dk = (double*)pl->alloc(kt.size() * sizeof(double));
    SE_HIP(hipMemcpy(dj, jt.data(), jt.size() * sizeof(double), hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(dk, kt.data(), kt.size() * sizeof(double), hipMemcpyHostToDevice));
    launch_green1p_table(pl->gtab, dj, dk, M, Mz, h, Lz, R, scalar, c00, pl->stream);
}

size_t plan_bytes_estimate(const se_plan_desc* d, int lds_max, int n_sys)
{
    Geom g;
    double eta, m;
    double R2p = 0;
    derive(d, g, eta, m, R2p);
    const int es = d->precision == SE_PREC_FP32 ? 4 : 8;
    int tile = d->tile, chunk = d->chunk, block = d->block;
    if (tile == 0 && chunk == 0 && block == 0) tuned_tiling(d, std::max(d->n_max, 1) / n_sys, tile, chunk, block);
    Tiling tl = choose_tiling(es, d->P, g.M, g.Mz, d->n_max, n_sys, tile, chunk, block, lds_max,
                              !(d->flags & SE_PLAN_GENERIC_KERNELS), !(d->flags & SE_PLAN_NO_MFMA),
                              !(d->flags & SE_PLAN_NO_PLANE_SPREAD), !(d->flags & SE_PLAN_NO_COLUMN));
    const size_t M = g.M, Mz = g.Mz, N = d->n_max, ns = n_sys;
    size_t b = 0;
    b += 4 * N * 4;                                       // two (key, index) pairs of the key sort
    b += 2 * 256 * 4 * (N / 1024 + 1) + (64u << 10);      // sort histograms and scan area (bound)
    b += (size_t)(ns * tl.ntiles + 1) * 4;                // tile starts
    b += 4 * (size_t)(ns * tl.ntiles + 2) * 4;            // tile_count, nchunk, chunk_start
    b += 4 * (size_t)tl.max_items * 4;                    // items (tile, begin, end, column table)
    b += N * (es == 4 ? 32 : 48);                         // records
    b += (size_t)tl.max_items * tl.part_elems * es;       // partials
    b += ns * M * M * Mz * es;                            // grids
    b += ns * M * M * (Mz / 2 + 1) * 2 * es;              // half spectra
    b += ns * M * M * (Mz / 2 + 1) * 2 * es;              // FFT work (upper estimate)
    if (g.per != 3) b += M * M * (Mz / 2 + 1) * 8;        // SE2P / SE1P Green's function table
    if (ns > 1) b += ns * sizeof(SysDesc);                // descriptors
    return b;
}

int lds_limit()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 64 * 1024;
    return v;
}

// Spread stage.  S == 0: one system (x, q, N).  S > 0: a grouped solve of S systems whose
// descriptors are in pl->d_sys and whose particle kernels take `blocks` blocks.
template <typename T>
void run_spread(se_plan* pl, int N, const double* x, const double* q, int S, int blocks)
{
    hipStream_t s = pl->stream;
    const Tiling& tl = pl->tl;
    const bool grp = S > 0;
    const int ns = grp ? S : 1;
    const long nb = (long)ns * tl.nbins;
    const int nt = ns * tl.ntiles;
    pl->rec_ev(0);
    // sort key and index of every particle, then a stable sort by key: the records of a
    // key keep the particles' index order, so the sorted order does not depend on timing
    if (grp)
        launch_sort_keys_group<T>(pl->d_sys, S, blocks, pl->g, tl, pl->key, pl->idx, s);
    else
        launch_sort_keys<T>(x, N, pl->g, tl, pl->key, pl->idx, s);
    const int bits = 32 - __builtin_clz((unsigned)std::max<long>(nb, 1));   // keys 0 .. nb (discard)
    const int* skey = nullptr;
    int* pos = nullptr;   // pos[n]: the sorted position of particle n (the inverse permutation)
    stable_sort_pairs(pl->key, pl->idx, pl->key2, pl->idx2, N, bits, true, pl->sort_tmp, pl->sort_bytes, s, &skey,
                      &pos);
    launch_tile_starts(skey, N, nt, tl.K, pl->tile_start, s);
    launch_tile_chunks(pl->tile_start, nt, tl.chunk, pl->tile_count, pl->nchunk, s);
    exclusive_scan(pl->nchunk, pl->chunk_start, nt + 1, pl->cub_tmp, pl->cub_bytes, s);
    launch_items(pl->tile_count, pl->tile_start, pl->chunk_start, nt, tl.chunk, pl->item_tile, pl->item_beg,
                 pl->item_end, pl->n_items, s);
    if (grp)
        launch_scatter_group<T>(pl->d_sys, S, blocks, pl->g, tl, pos, (Rec<T>*)pl->rec, s);
    else
        launch_scatter<T>(x, q, N, pl->g, tl, pos, (Rec<T>*)pl->rec, s);
    // work items: at most one partial item per non-empty tile plus one per full chunk
    Tiling tlc = tl;
    tlc.max_items = std::min(tl.max_items, nt + N / tl.chunk + 1);
    // the column items: chosen, and their records put in column order, for the column
    // spread and gather
    if constexpr (sizeof(T) == 4) {
        if (tl.column)
            launch_column_order((Rec<float>*)pl->rec, pl->item_beg, pl->item_end, pl->n_items, tlc.max_items, pl->g,
                                tl, pl->item_col, s);
    }
    pl->rec_ev(1);
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        if (tl.pencil)
            done = launch_spread_pencil((const Rec<float>*)pl->rec, pl->item_tile, pl->item_beg, pl->item_end,
                                        pl->n_items, pl->tile_start, pl->g, tlc, (float*)pl->partials,
                                        tl.column ? pl->item_col : nullptr, s,
                                        (tl.column && !grp && N >= (1 << 20)) ? pl->side : nullptr, pl->fork_ev,
                                        pl->join_ev);
    }
    if (!done)
        launch_spread_tiles<T>((const Rec<T>*)pl->rec, pl->item_tile, pl->item_beg, pl->item_end, pl->n_items,
                               pl->g, tlc, (T*)pl->partials, s);
    pl->rec_ev(2);
    launch_merge<T>((const T*)pl->partials, pl->chunk_start, pl->seg, pl->nseg, pl->segz, pl->nsegz, tl.nseg, pl->g,
                    tl, (T*)pl->grid, ns, s);
    pl->rec_ev(3);
    pl->last_N = N;
    pl->last_S = ns;
    pl->last_group = grp;
    SE_HIP(hipGetLastError());
}

// Gather stage.  After a grouped spread, phi / force only select the output mode and
// every system's outputs come from the descriptors.
template <typename T>
void run_interp(se_plan* pl, double* phi, double* force)
{
    pl->rec_ev(4);
    const SysDesc* sys = pl->last_group ? pl->d_sys : nullptr;
    Tiling tlc = pl->tl;
    tlc.max_items = std::min(pl->tl.max_items, pl->last_S * pl->tl.ntiles + pl->last_N / pl->tl.chunk + 1);
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        // The matrix-core gather of the dense items and the row gather of the others run
        // concurrently (the second on the plan's side stream) for one system of at least
        // 2^20 particles: which of them has work depends on the particle distribution,
        // which the host does not see, and where one of them has none the fork and join
        // cost about 1 % (uniform N = 1M: gather 0.500 against 0.502 ms).  Smaller
        // systems run both on one stream.
        const Tiling& tl = pl->tl;
        const bool split = tl.mfma && !pl->last_group && pl->last_N >= (1 << 20);
        if (pl->tl.pencil)
            done = launch_interp_pencil((const Rec<float>*)pl->rec, pl->item_tile, pl->item_beg, pl->item_end,
                                        pl->n_items, (const float*)pl->grid, pl->g, tlc, pl->last_N, phi, force, sys,
                                        tl.column ? pl->item_col : nullptr, pl->stream, split ? pl->side : nullptr,
                                        pl->fork_ev, pl->join_ev);
    }
    if (!done)
        launch_interp_tiles<T>((const Rec<T>*)pl->rec, pl->item_tile, pl->item_beg, pl->item_end, pl->n_items,
                               (const T*)pl->grid, pl->g, tlc, pl->last_N, phi, force, sys, pl->stream);
    pl->rec_ev(5);
    SE_HIP(hipGetLastError());
}

void make_fft_pair(se_plan* pl, int S, rocfft_plan& fwd, rocfft_plan& bwd, size_t& work);
se_plan::GroupFft& group_fft(se_plan* pl, int S);

// k-space stage over the grids of the last spread (one, or S stacked ones with batched
// transforms).
void run_kspace(se_plan* pl)
{
    pl->rec_ev(6);
    const int ns = pl->last_S;
    rocfft_plan fwd = pl->fwd, bwd = pl->bwd;
    rocfft_execution_info info = pl->info;
    if (ns > 1) {
        se_plan::GroupFft& f = group_fft(pl, ns);
        fwd = f.fwd;
        bwd = f.bwd;
        info = f.info;
    }
    void* in[1] = {pl->grid};
    void* out[1] = {pl->spec};
    SE_FFT(rocfft_execute(fwd, in, out, info));
    const int M = pl->g.M, Mz = pl->g.Mz;
    if (pl->g.per == 3) {
        if (pl->esize == 4)
            launch_green<float>(pl->spec, M, Mz, pl->gtab, ns, pl->stream);
        else
            launch_green<double>(pl->spec, M, Mz, pl->gtab, ns, pl->stream);
    } else {
        const long n = (long)M * M * (Mz / 2 + 1);
        if (pl->esize == 4)
            launch_green_full<float>(pl->spec, n, pl->gtab, ns, pl->stream);
        else
            launch_green_full<double>(pl->spec, n, pl->gtab, ns, pl->stream);
    }
    void* in2[1] = {pl->spec};
    void* out2[1] = {pl->grid};
    SE_FFT(rocfft_execute(bwd, in2, out2, info));
    pl->rec_ev(7);
    SE_HIP(hipGetLastError());
}

// R2C / C2R plans for S stacked grids, lengths fastest-first (z, y, x).
void make_fft_pair(se_plan* pl, int S, rocfft_plan& fwd, rocfft_plan& bwd, size_t& work)
{
    const size_t M = pl->g.M, Mz = pl->g.Mz;
    size_t lens[3] = {Mz, M, M};
    const rocfft_precision prec = pl->esize == 4 ? rocfft_precision_single : rocfft_precision_double;
    const size_t rdist = M * M * Mz, cdist = M * M * (Mz / 2 + 1);
    size_t rstr[3] = {1, Mz, M * Mz}, cstr[3] = {1, Mz / 2 + 1, M * (Mz / 2 + 1)};
    rocfft_plan_description df = nullptr, db = nullptr;
    SE_FFT(rocfft_plan_description_create(&df));
    SE_FFT(rocfft_plan_description_create(&db));
    struct Guard {
        rocfft_plan_description a, b;
        ~Guard()
        {
            rocfft_plan_description_destroy(a);
            rocfft_plan_description_destroy(b);
        }
    } gd{df, db};
    SE_FFT(rocfft_plan_description_set_data_layout(df, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved,
                                                   nullptr, nullptr, 3, rstr, rdist, 3, cstr, cdist));
    SE_FFT(rocfft_plan_description_set_data_layout(db, rocfft_array_type_hermitian_interleaved, rocfft_array_type_real,
                                                   nullptr, nullptr, 3, cstr, cdist, 3, rstr, rdist));
    SE_FFT(rocfft_plan_create(&fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, prec, 3, lens,
                              (size_t)S, df));
    SE_FFT(rocfft_plan_create(&bwd, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, prec, 3, lens,
                              (size_t)S, db));
    size_t w1 = 0, w2 = 0;
    SE_FFT(rocfft_plan_get_work_buffer_size(fwd, &w1));
    SE_FFT(rocfft_plan_get_work_buffer_size(bwd, &w2));
    work = std::max(w1, w2);
}

// Batched transforms for S grids, cached per size (the 8 most recently used sizes).
se_plan::GroupFft& group_fft(se_plan* pl, int S)
{
    for (auto& f : pl->gffts)
        if (f.S == S) {
            f.last_use = ++pl->gfft_clock;
            return f;
        }
    if (pl->gffts.size() >= 8) {
        auto lru = std::min_element(pl->gffts.begin(), pl->gffts.end(),
                                    [](const auto& a, const auto& b) { return a.last_use < b.last_use; });
        SE_HIP(hipStreamSynchronize(pl->stream));
        rocfft_plan_destroy(lru->fwd);
        rocfft_plan_destroy(lru->bwd);
        rocfft_execution_info_destroy(lru->info);
        pl->gffts.erase(lru);
    }
    se_plan::GroupFft f{S, nullptr, nullptr, nullptr, ++pl->gfft_clock};
    size_t work = 0;
    make_fft_pair(pl, S, f.fwd, f.bwd, work);
    SE_FFT(rocfft_execution_info_create(&f.info));
    if (work > pl->fft_work_bytes) {
        // the plan's FFT work area was sized for the largest group at creation; a size
        // that needs more gets a larger area (after the stream's work drains)
        SE_HIP(hipStreamSynchronize(pl->stream));
        pl->fft_work = pl->alloc(work);
        pl->fft_work_bytes = work;
        SE_FFT(rocfft_execution_info_set_work_buffer(pl->info, pl->fft_work, pl->fft_work_bytes));
        for (auto& o : pl->gffts) SE_FFT(rocfft_execution_info_set_work_buffer(o.info, pl->fft_work, work));
    }
    if (pl->fft_work_bytes) SE_FFT(rocfft_execution_info_set_work_buffer(f.info, pl->fft_work, pl->fft_work_bytes));
    SE_FFT(rocfft_execution_info_set_stream(f.info, pl->stream));
    pl->gffts.push_back(f);
    return pl->gffts.back();
}

}  // namespace

extern "C" {

const char* se_gpu_last_error(void) { return se::last_error(); }

size_t se_plan_bytes_for(const se_plan_desc* desc)
{
    size_t b = 0;
    int rc = guarded([&] { b = plan_bytes_estimate(desc, lds_limit(), 1); });
    return rc == SE_OK ? b : 0;
}

size_t se_plan_bytes_for_group(const se_plan_desc* desc, int max_systems)
{
    size_t b = 0;
    int rc = guarded([&] {
        SE_REQUIRE(max_systems >= 1, SE_ERR_ARG, "max_systems must be >= 1");
        b = plan_bytes_estimate(desc, lds_limit(), max_systems);
    });
    return rc == SE_OK ? b : 0;
}

}  // extern "C"

namespace se {
// True when the plan holds a captured graph of se_solve with exactly these arguments.
bool plan_has_graph(const se_plan* pl, int N, const void* x, const void* q, const void* phi, const void* force)
{
    if (!pl || pl->timing) return false;
    for (const auto& e : pl->graphs)
        if (e.N == N && e.x == x && e.q == q && e.phi == phi && e.force == force) return true;
    return false;
}
}  // namespace se

static int create_plan(se_plan** out, const se_plan_desc* d, void* stream, int n_sys)
{
    if (!out || !d) return SE_ERR_ARG;
    *out = nullptr;
    se_plan* pl = new se_plan();
    int rc = guarded([&] {
        SE_REQUIRE(d->n_max >= 0, SE_ERR_ARG, "n_max must be >= 0");
        SE_REQUIRE(n_sys >= 1 && n_sys <= 65535, SE_ERR_ARG, "max_systems must be in [1, 65535]");
        pl->s_cap = n_sys;
        pl->desc = *d;
        pl->esize = d->precision == SE_PREC_FP32 ? 4 : 8;
        derive(d, pl->g, pl->eta, pl->m, pl->R2p);
        pl->L = d->box;
        pl->xi = d->xi;
        pl->stream = (hipStream_t)stream;
        pl->n_cap = std::max(d->n_max, 1);
        const int lds_max = lds_limit();
        int tile = d->tile, chunk = d->chunk, block = d->block;
        if (tile == 0 && chunk == 0 && block == 0) tuned_tiling(d, pl->n_cap / n_sys, tile, chunk, block);
        pl->tl = choose_tiling(pl->esize, d->P, pl->g.M, pl->g.Mz, pl->n_cap, n_sys, tile, chunk, block, lds_max,
                               !(d->flags & SE_PLAN_GENERIC_KERNELS), !(d->flags & SE_PLAN_NO_MFMA),
                               !(d->flags & SE_PLAN_NO_PLANE_SPREAD), !(d->flags & SE_PLAN_NO_COLUMN));
        const Tiling& tl = pl->tl;
        const size_t M = pl->g.M, Mz = pl->g.Mz, N = pl->n_cap, ns = n_sys;
        SE_REQUIRE(ns * tl.nbins < (1ul << 31) - 1 && ns * tl.ntiles * (size_t)tl.K < (1ul << 31) - 1, SE_ERR_ARG,
                   "too many systems for the sort keys of this geometry");
        pl->key = (int*)pl->alloc(N * 4);
        pl->idx = (int*)pl->alloc(N * 4);
        pl->key2 = (int*)pl->alloc(N * 4);
        pl->idx2 = (int*)pl->alloc(N * 4);
        pl->tile_start = (int*)pl->alloc((ns * tl.ntiles + 1) * 4);
        pl->sort_bytes = sort_tmp_bytes((int)N);
        pl->sort_tmp = pl->alloc(pl->sort_bytes);
        pl->tile_count = (int*)pl->alloc((ns * tl.ntiles + 1) * 4);
        pl->nchunk = (int*)pl->alloc((ns * tl.ntiles + 1) * 4);
        pl->chunk_start = (int*)pl->alloc((ns * tl.ntiles + 2) * 4);
        pl->item_tile = (int*)pl->alloc((size_t)tl.max_items * 4);
        pl->item_beg = (int*)pl->alloc((size_t)tl.max_items * 4);
        pl->item_end = (int*)pl->alloc((size_t)tl.max_items * 4);
        pl->item_col = (int*)pl->alloc((size_t)tl.max_items * 4);
        pl->n_items = (int*)pl->alloc(16);
        pl->rec = pl->alloc(N * (pl->esize == 4 ? sizeof(Rec<float>) : sizeof(Rec<double>)));
        pl->partials = pl->alloc((size_t)tl.max_items * tl.part_elems * pl->esize);
        pl->grid = pl->alloc(ns * M * M * Mz * pl->esize);
        pl->spec = pl->alloc(ns * M * M * (Mz / 2 + 1) * 2 * pl->esize);
        pl->gtab = (double*)pl->alloc((pl->g.per == 3 ? 2 * M + 2 * Mz + 1 : M * M * (Mz / 2 + 1)) * sizeof(double));
        // temporary area of the scan of the tile chunks
        SE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, pl->cub_bytes, pl->nchunk, pl->chunk_start,
                                                (int)(ns * tl.ntiles + 1), pl->stream));
        if (ns > 1) {
            pl->d_sys = (SysDesc*)pl->alloc(ns * sizeof(SysDesc));
            for (int i = 0; i < 4; ++i) {
                SE_HIP(hipHostMalloc((void**)&pl->h_sys[i], ns * sizeof(SysDesc), hipHostMallocDefault));
                SE_HIP(hipEventCreateWithFlags(&pl->h_ev[i], hipEventDisableTiming));
            }
        }
        pl->cub_tmp = pl->alloc(pl->cub_bytes);
        build_merge_segments(pl);
        // Green's function: s = -(1-eta)/(4 xi^2)   (spectral_ewald.c:126)
        const double gs = -(1.0 - pl->eta) / (4.0 * d->xi * d->xi);
        if (pl->g.per == 3)
            launch_green_tables(pl->gtab, (int)M, (int)Mz, d->box, d->box, gs, pl->stream);
        else if (pl->g.per == 2)
            launch_green2p_table(pl->gtab, (int)M, (int)Mz, d->box, pl->g.h, pl->R2p, gs, pl->stream);
        else
            build_green1p(pl, gs);
        // rocFFT R2C / C2R for one grid; a group plan also sizes the work area for its
        // largest batch (the batched plans themselves are made per group size on use)
        rocfft_init_once();
        size_t work = 0;
        make_fft_pair(pl, 1, pl->fwd, pl->bwd, work);
        pl->fft_work_bytes = work;
        if (ns > 1) {
            rocfft_plan f = nullptr, b = nullptr;
            size_t wg = 0;
            make_fft_pair(pl, (int)ns, f, b, wg);
            rocfft_plan_destroy(f);
            rocfft_plan_destroy(b);
            pl->fft_work_bytes = std::max(work, wg);
        }
        SE_FFT(rocfft_execution_info_create(&pl->info));
        if (pl->fft_work_bytes) {
            pl->fft_work = pl->alloc(pl->fft_work_bytes);
            SE_FFT(rocfft_execution_info_set_work_buffer(pl->info, pl->fft_work, pl->fft_work_bytes));
        }
        SE_FFT(rocfft_execution_info_set_stream(pl->info, pl->stream));
        for (auto& e : pl->ev) SE_HIP(hipEventCreate(&e));
        if (tl.mfma) {
            SE_HIP(hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking));
            SE_HIP(hipEventCreateWithFlags(&pl->fork_ev, hipEventDisableTiming));
            SE_HIP(hipEventCreateWithFlags(&pl->join_ev, hipEventDisableTiming));
        }
        SE_HIP(hipStreamSynchronize(pl->stream));
    });
    if (rc != SE_OK) {
        delete pl;
        return rc;
    }
    *out = pl;
    return SE_OK;
}

extern "C" {

int se_plan_create(se_plan** out, const se_plan_desc* d, void* stream) { return create_plan(out, d, stream, 1); }

int se_plan_create_group(se_plan** out, const se_plan_desc* d, int max_systems, void* stream)
{
    return create_plan(out, d, stream, max_systems);
}

int se_plan_destroy(se_plan* pl)
{
    if (!pl) return SE_OK;
    int rc = guarded([&] { SE_HIP(hipStreamSynchronize(pl->stream)); });
    delete pl;
    return rc;
}

int se_plan_set_stream(se_plan* pl, void* stream)
{
    if (!pl) return SE_ERR_ARG;
    return guarded([&] {
        pl->stream = (hipStream_t)stream;
        SE_FFT(rocfft_execution_info_set_stream(pl->info, pl->stream));
        for (auto& f : pl->gffts) SE_FFT(rocfft_execution_info_set_stream(f.info, pl->stream));
    });
}

size_t se_plan_bytes(const se_plan* pl) { return pl ? pl->bytes : 0; }

int se_spread(se_plan* pl, int N, const double* x, const double* q)
{
    if (!pl) return SE_ERR_ARG;
    return guarded([&] {
        SE_REQUIRE(N >= 0 && N <= pl->n_cap, SE_ERR_CAPACITY, "N exceeds the plan's particle capacity");
        SE_REQUIRE(N == 0 || (x && q), SE_ERR_ARG, "null particle arrays");
        if (pl->esize == 4)
            run_spread<float>(pl, N, x, q, 0, 0);
        else
            run_spread<double>(pl, N, x, q, 0, 0);
    });
}

int se_kspace(se_plan* pl)
{
    if (!pl) return SE_ERR_ARG;
    return guarded([&] { run_kspace(pl); });
}

int se_interp(se_plan* pl, double* phi, double* force)
{
    if (!pl) return SE_ERR_ARG;
    return guarded([&] {
        SE_REQUIRE(pl->last_N >= 0, SE_ERR_ARG, "se_interp before se_spread");
        if (!phi && !force) return;
        if (pl->esize == 4)
            run_interp<float>(pl, phi, force);
        else
            run_interp<double>(pl, phi, force);
    });
}

int se_solve(se_plan* pl, int N, const double* x, const double* q, double* phi, double* force)
{
    int rc = se_spread(pl, N, x, q);
    if (rc) return rc;
    rc = se_kspace(pl);
    if (rc) return rc;
    return se_interp(pl, phi, force);
}

int se_solve_group(se_plan* pl, int S, const se_system* sys)
{
    if (!pl) return SE_ERR_ARG;
    return guarded([&] {
        SE_REQUIRE(S >= 1 && S <= pl->s_cap, SE_ERR_CAPACITY, "number of systems exceeds the group plan's capacity");
        SE_REQUIRE(sys, SE_ERR_ARG, "null system list");
        // the output mode comes from the systems that have particles (an empty system has
        // no outputs, and its arrays may be null)
        int lead = 0;
        while (lead < S - 1 && sys[lead].N == 0) ++lead;
        const bool want_phi = sys[lead].d_phi != nullptr, want_force = sys[lead].d_force != nullptr;
        long total = 0, blocks = 0;
        for (int i = 0; i < S; ++i) {
            SE_REQUIRE(sys[i].N >= 0, SE_ERR_ARG, "negative system size");
            SE_REQUIRE(sys[i].N == 0 || (sys[i].d_x && sys[i].d_q), SE_ERR_ARG, "null particle arrays");
            SE_REQUIRE(sys[i].N == 0 || ((sys[i].d_phi != nullptr) == want_phi &&
                                         (sys[i].d_force != nullptr) == want_force),
                       SE_ERR_ARG, "all systems of a group must request the same outputs");
            total += sys[i].N;
            blocks += (sys[i].N + 255) / 256;
        }
        SE_REQUIRE(total <= pl->n_cap, SE_ERR_CAPACITY, "total particles exceed the group plan's capacity");
        double* phi_mode = sys[lead].d_phi;
        double* force_mode = sys[lead].d_force;
        if (S == 1) {   // one system: the plain path, no descriptors
            if (pl->esize == 4)
                run_spread<float>(pl, sys[0].N, sys[0].d_x, sys[0].d_q, 0, 0);
            else
                run_spread<double>(pl, sys[0].N, sys[0].d_x, sys[0].d_q, 0, 0);
        } else {
            // descriptors through a pinned staging ring: buffer i is rewritten only after
            // the copy that last read it has run
            const int hb = pl->h_next;
            pl->h_next = (hb + 1) % 4;
            SE_HIP(hipEventSynchronize(pl->h_ev[hb]));
            SysDesc* h = pl->h_sys[hb];
            long off = 0, blk = 0;
            for (int i = 0; i < S; ++i) {
                h[i].x = sys[i].d_x;
                h[i].q = sys[i].d_q;
                h[i].phi = sys[i].d_phi;
                h[i].force = sys[i].d_force;
                h[i].N = sys[i].N;
                h[i].off = (int)off;
                h[i].blk = (int)blk;
                h[i].pad = 0;
                off += sys[i].N;
                blk += (sys[i].N + 255) / 256;
            }
            SE_HIP(hipMemcpyAsync(pl->d_sys, h, S * sizeof(SysDesc), hipMemcpyHostToDevice, pl->stream));
            SE_HIP(hipEventRecord(pl->h_ev[hb], pl->stream));
            if (pl->esize == 4)
                run_spread<float>(pl, (int)total, nullptr, nullptr, S, (int)blocks);
            else
                run_spread<double>(pl, (int)total, nullptr, nullptr, S, (int)blocks);
        }
        run_kspace(pl);
        if (want_phi || want_force) {
            if (pl->esize == 4)
                run_interp<float>(pl, phi_mode, force_mode);
            else
                run_interp<double>(pl, phi_mode, force_mode);
        }
    });
}

int se_solve_graph(se_plan* pl, int N, const double* x, const double* q, double* phi, double* force)
{
    if (!pl) return SE_ERR_ARG;
    if (pl->timing) return se_solve(pl, N, x, q, phi, force);   // events do not go into graphs
    for (auto& e : pl->graphs)
        if (e.N == N && e.x == x && e.q == q && e.phi == phi && e.force == force) {
            e.last_use = ++pl->graph_clock;
            pl->last_N = e.last_N;
            pl->last_S = e.last_S;
            pl->last_group = e.last_group;
            return guarded([&] { SE_HIP(hipGraphLaunch(e.exec, pl->stream)); });
        }
    // validate eagerly (argument errors must not happen inside a capture)
    if (N < 0 || N > pl->n_cap) {
        se::set_error("N exceeds the plan's particle capacity");
        return SE_ERR_CAPACITY;
    }
    if (N > 0 && (!x || !q)) return SE_ERR_ARG;
    int rc = SE_OK;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    rc = guarded([&] { SE_HIP(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal)); });
    if (rc) return rc;
    const int src = se_solve(pl, N, x, q, phi, force);
    const hipError_t ec = hipStreamEndCapture(pl->stream, &graph);
    if (src != SE_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return src;
    }
    rc = guarded([&] {
        SE_HIP(ec);
        SE_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        graph = nullptr;
        const size_t max_graphs = 32;
        if (pl->graphs.size() >= max_graphs) {
            auto lru = std::min_element(pl->graphs.begin(), pl->graphs.end(),
                                        [](const auto& a, const auto& b) { return a.last_use < b.last_use; });
            SE_HIP(hipStreamSynchronize(pl->stream));
            (void)hipGraphExecDestroy(lru->exec);
            pl->graphs.erase(lru);
        }
        pl->graphs.push_back({N, x, q, phi, force, exec, ++pl->graph_clock, pl->last_N, pl->last_S, pl->last_group});
        SE_HIP(hipGraphLaunch(exec, pl->stream));
    });
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

int se_plan_grid_shape(const se_plan* pl, int dims[3], double* z0)
{
    if (!pl || !dims) return SE_ERR_ARG;
    dims[0] = dims[1] = pl->g.M;
    dims[2] = pl->g.Mz;
    if (z0) *z0 = pl->g.z0;
    return SE_OK;
}

int se_plan_grid(se_plan* pl, void** d_grid, int* elem_bytes)
{
    if (!pl || !d_grid) return SE_ERR_ARG;
    *d_grid = pl->grid;
    if (elem_bytes) *elem_bytes = pl->esize;
    return SE_OK;
}

int se_plan_enable_timing(se_plan* pl, int on)
{
    if (!pl) return SE_ERR_ARG;
    pl->timing = on != 0;
    return SE_OK;
}

int se_plan_timings(se_plan* pl, float ms[8])
{
    if (!pl || !ms) return SE_ERR_ARG;
    return guarded([&] {
        SE_HIP(hipStreamSynchronize(pl->stream));
        for (int i = 0; i < 8; ++i) ms[i] = 0.f;
        auto el = [&](int a, int b) {
            float v = 0.f;
            if (hipEventElapsedTime(&v, pl->ev[a], pl->ev[b]) != hipSuccess) v = -1.f;
            return v;
        };
        ms[0] = el(0, 3);
        ms[1] = el(1, 2);
        ms[2] = el(2, 3);
        ms[3] = el(6, 7);
        ms[4] = el(4, 5);
        ms[5] = ms[4];
        ms[6] = el(0, 1);
        (void)hipGetLastError();
    });
}

int se_plan_info(se_plan* pl, int out[8])
{
    if (!pl || !out) return SE_ERR_ARG;
    return guarded([&] {
        for (int i = 0; i < 8; ++i) out[i] = 0;
        out[0] = pl->tl.TS;
        out[1] = pl->tl.chunk;
        out[2] = pl->tl.block;
        out[3] = pl->tl.ntiles;
        SE_HIP(hipStreamSynchronize(pl->stream));
        int ni = 0, kept = 0;
        SE_HIP(hipMemcpy(&ni, pl->n_items, sizeof(int), hipMemcpyDeviceToHost));
        // the discard bucket sorts after every tile: its records start at the end entry
        // of the tile starts
        SE_HIP(hipMemcpy(&kept, pl->tile_start + (long)pl->last_S * pl->tl.ntiles, sizeof(int),
                         hipMemcpyDeviceToHost));
        out[4] = ni;
        out[5] = pl->last_N > 0 ? pl->last_N - kept : 0;
        out[6] = pl->tl.S;
        out[7] = pl->tl.pencil;
    });
}

int se_energy(se_plan* pl, int N, const double* phi, const double* q, double* d_energy)
{
    if (!pl) return SE_ERR_ARG;
    return guarded([&] {
        launch_energy(phi, q, N, d_energy, pl->stream);
        SE_HIP(hipGetLastError());
    });
}

}  // extern "C"

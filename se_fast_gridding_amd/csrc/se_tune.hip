// Tiling autotuner (se_gpu.h: se_tune*).  For one plan geometry, a particle count and a
// number of systems solved together (a grouped solve when more than one), se_tune times
// the full solve of synthetic uniform systems under every candidate
// tiling -- tile side, maximum particles per work item (the bucket cap that splits heavy
// tiles) and threads per work-group -- and records the fastest in a process-wide table
// keyed by (periodicity, M, P, precision, ceil(log2 N)).  Plans created afterwards with
// tile = chunk = block = 0 take the recorded tiling of their geometry at the nearest size
// band (group plans use the particles per system), so the batch executor's cached plans
// pick it up unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/se_gpu.h"
#include "se_internal.h"

namespace se {

namespace {

struct TuneVal {
    int tile, chunk, block;
    float ms;
};
using TuneKey = std::tuple<int, int, int, int, int>;   // periodicity, M, P, precision, band

std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

int band_of(long n)
{
    int b = 0;
    while ((1l << b) < n && b < 40) ++b;
    return b;
}

TuneKey key_of(const se_plan_desc* d, long n)
{
    return TuneKey{d->periodicity, d->M, d->P, d->precision == SE_PREC_FP32 ? 1 : 0, band_of(std::max(n, 1l))};
}

// splitmix64 of the index: reproducible uniform [0, 1) without host traffic
__device__ __forceinline__ double hash01(unsigned long long i)
{
    unsigned long long z = i + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// positions uniform over the particles' range: [0, box)^3 (SE3P); [0, box)^2 x [0, slab)
// (SE2P); [0, slab)^2 x [0, box) (SE1P); charges +-1 alternating (0 for an odd last one)
__global__ void k_tune_system(double* x, double* q, int N, int per, double box, double slab)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int d = 0; d < 3; ++d) {
        const bool free_d = (per == 2 && d == 2) || (per == 1 && d < 2);
        x[n + (long)d * N] = hash01(3ull * n + d) * (free_d ? slab : box);
    }
    q[n] = (N % 2 == 1 && n == N - 1) ? 0.0 : ((n & 1) ? -1.0 : 1.0);
}

}  // namespace

// The entry of d's geometry whose size band is nearest to n's (the exact band first).
bool tuned_tiling(const se_plan_desc* d, long n_per_system, int& tile, int& chunk, int& block)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const TuneKey k = key_of(d, n_per_system);
    const TuneVal* best = nullptr;
    int dist = 1 << 30;
    for (auto& kv : g_tune) {
        const TuneKey& e = kv.first;
        if (std::get<0>(e) != std::get<0>(k) || std::get<1>(e) != std::get<1>(k) || std::get<2>(e) != std::get<2>(k) ||
            std::get<3>(e) != std::get<3>(k))
            continue;
        const int dd = std::abs(std::get<4>(e) - std::get<4>(k));
        if (dd < dist) {
            dist = dd;
            best = &kv.second;
        }
    }
    if (!best) return false;
    tile = best->tile;
    chunk = best->chunk;
    block = best->block;
    return true;
}

}  // namespace se

using namespace se;

extern "C" {

int se_tune(const se_plan_desc* desc, int N, int systems, int quantities, void* stream, int out[3], float* out_ms)
{
    if (!desc || N < 1 || systems < 1 || (long)N * systems >= (1l << 31) ||
        !(quantities & (SE_Q_POTENTIAL | SE_Q_FORCE)))
        return SE_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    double *x = nullptr, *q = nullptr, *phi = nullptr, *F = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SE_OK;
    auto fail = [&](const char* msg) {
        set_error(msg);
        rc = SE_ERR_HIP;
    };
    const size_t NS = (size_t)N * systems;
    if (hipMalloc(&x, 3 * (size_t)N * sizeof(double)) != hipSuccess || hipMalloc(&q, (size_t)N * sizeof(double)) ||
        hipMalloc(&phi, NS * sizeof(double)) || hipMalloc(&F, 3 * NS * sizeof(double)) || hipEventCreate(&e0) ||
        hipEventCreate(&e1)) {
        fail("se_tune: device allocation failed");
    }
    TuneVal best{0, 0, 0, 1e30f};
    if (rc == SE_OK) {
        hipLaunchKernelGGL(k_tune_system, dim3((N + 255) / 256), dim3(256), 0, s, x, q, N, desc->periodicity,
                           desc->box, desc->slab);
        const bool pencil = desc->precision == SE_PREC_FP32 && pencil_supported(desc->P) &&
                            !(desc->flags & SE_PLAN_GENERIC_KERNELS);
        std::vector<int> tiles = pencil ? std::vector<int>{4, 8, 12} : std::vector<int>{6, 8, 12, 16};
        const int chunks[] = {512, 1024, 2048, 4096};
        const int blocks[] = {0, 256};   // 0: each kernel's own default (spread 768, gathers 512 / 256)
        // `systems` copies of the system, solved as one group (each with its own outputs)
        std::vector<se_system> sys(systems);
        for (int i = 0; i < systems; ++i) {
            sys[i].N = N;
            sys[i].d_x = x;
            sys[i].d_q = q;
            sys[i].d_phi = (quantities & SE_Q_POTENTIAL) ? phi + (size_t)i * N : nullptr;
            sys[i].d_force = (quantities & SE_Q_FORCE) ? F + 3 * (size_t)i * N : nullptr;
        }
        auto solve = [&](se_plan* pl) {
            return systems == 1 ? se_solve(pl, N, x, q, sys[0].d_phi, sys[0].d_force)
                                : se_solve_group(pl, systems, sys.data());
        };
        for (int t : tiles)
            for (int c : chunks)
                for (int b : blocks) {
                    se_plan_desc d = *desc;
                    d.tile = t;
                    d.chunk = c;
                    d.block = b;
                    d.n_max = (int)NS;
                    se_plan* pl = nullptr;
                    const int cr = systems == 1 ? se_plan_create(&pl, &d, stream)
                                                : se_plan_create_group(&pl, &d, systems, stream);
                    if (cr != SE_OK) continue;   // e.g. a sub-grid past LDS
                    int info[8];
                    bool ok = solve(pl) == SE_OK && se_plan_info(pl, info) == SE_OK &&
                              info[0] == t;   // a tile the plan rounded to another one is tried under that one
                    float ms = 0.f;
                    if (ok) {
                        ok = hipEventRecord(e0, s) == hipSuccess;
                        for (int r = 0; r < 3 && ok; ++r) ok = solve(pl) == SE_OK;
                        ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                    }
                    se_plan_destroy(pl);
                    if (ok && ms / 3 < best.ms) best = TuneVal{t, c, b, ms / 3};
                }
        if (best.tile == 0) {
            fail("se_tune: no candidate tiling produced a solve");
        } else {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[key_of(desc, N)] = best;
        }
    }
    if (x) (void)hipFree(x);
    if (q) (void)hipFree(q);
    if (phi) (void)hipFree(phi);
    if (F) (void)hipFree(F);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == SE_OK) {
        if (out) {
            out[0] = best.tile;
            out[1] = best.chunk;
            out[2] = best.block;
        }
        if (out_ms) *out_ms = best.ms;
    }
    return rc;
}

int se_tune_save(const char* path)
{
    if (!path) return SE_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) {
        set_error("se_tune_save: cannot open the file");
        return SE_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::fprintf(f, "# periodicity M P fp32 log2N tile chunk block ms\n");
    for (auto& kv : g_tune) {
        const auto& k = kv.first;
        std::fprintf(f, "%d %d %d %d %d %d %d %d %.6f\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
                     std::get<3>(k), std::get<4>(k), kv.second.tile, kv.second.chunk, kv.second.block, kv.second.ms);
    }
    std::fclose(f);
    return SE_OK;
}

int se_tune_load(const char* path)
{
    if (!path) return SE_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) {
        set_error("se_tune_load: cannot open the file");
        return SE_ERR_ARG;
    }
    char line[256];
    int n = 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    while (std::fgets(line, sizeof line, f)) {
        int per, M, P, fp32, band, tile, chunk, block;
        float ms;
        if (line[0] == '#') continue;
        if (std::sscanf(line, "%d %d %d %d %d %d %d %d %f", &per, &M, &P, &fp32, &band, &tile, &chunk, &block, &ms) ==
                9 &&
            tile > 0 && chunk > 0 && block >= 0) {
            g_tune[TuneKey{per, M, P, fp32, band}] = TuneVal{tile, chunk, block, ms};
            ++n;
        }
    }
    std::fclose(f);
    return n;
}

int se_tune_clear(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
    return SE_OK;
}

int se_tune_lookup(const se_plan_desc* desc, int N, int out[3])
{
    if (!desc || !out) return SE_ERR_ARG;
    return tuned_tiling(desc, N, out[0], out[1], out[2]) ? 1 : 0;
}

// Parameter selection for an accuracy target, from the estimates SE_init_params prints
// (SE_general.c:525-549): Lindbo-Tornberg gridding errors with c = 0.95, and the
// reference's grid rule (SE_general.c:536-563, AVX build).
int se_select_params(double tol, double L, double xi, double Q, int quantities, int* P, int* M, double* m,
                     double* est)
{
    if (!(tol > 0) || !(L > 0) || !(xi > 0) || !(Q >= 0) || !P || !M || !m) return SE_ERR_ARG;
    if (!(quantities & (SE_Q_POTENTIAL | SE_Q_FORCE))) return SE_ERR_ARG;
    const double c = 0.95, PI = 3.14159265358979323846;
    auto estimate = [&](int p) {
        const double g = std::exp(-PI * p * c * c / 2.);
        double e = 0.;
        if (quantities & SE_Q_POTENTIAL) e = std::max(e, std::sqrt(Q * xi * L) / L * g);
        if (quantities & SE_Q_FORCE) e = std::max(e, Q * std::sqrt(xi * L) / (L * L) * g * 2. * PI * c * c * p);
        return e;
    };
    int p = 1;
    while (p <= 32 && estimate(p) > tol) ++p;
    if (p > 32) return SE_ERR_ARG;
    int Ml = (int)std::ceil(2. * (2. * xi * L) + 1) + 2;
    if (Ml < p) Ml = p;
    if (Ml % 4) Ml += 4 - Ml % 4;
    *P = p;
    *M = Ml;
    *m = c * std::sqrt(PI * p);
    if (est) *est = estimate(p);
    return SE_OK;
}

}  // extern "C"

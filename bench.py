"""SE3P spread+interp throughput on one MI355X.

Workload (default): the clustered N = 4,000,000 system -- 80 % of the charges in 64
isotropic Gaussian clusters (sigma = 0.01 L, centres uniform), 20 % uniform, q = +-1 with
zero net charge -- at P = 16, M = 256^3, xi = 80, FP32 spread / FFT / gather, potential
and force.  One step = one full SE3P k-space solve with x, q already in HBM: bucketing +
spread + halo merge, R2C FFT + Green's function + C2R FFT, gather (phi, F).

value = 2 N P^3 / (t_spread_stage + t_interp_stage)   [particle*P^3 updates/s]
(the reference metric counts expand, spread, wrap, extend and interpolation and excludes
the FFT; the full-solve time is reported beside it).

roofline: the path is compute-bound in FP32 (one FMA per update in the spread, two in the
potential+force gather); t_roof = max(F / 157.3 TFLOP/s, B / 8 TB/s) with F the
algorithmic flops and B the algorithmic bytes of spread + gather, and
frac = t_roof / (t_spread_stage + t_interp_stage).  Per-kernel fractions, HBM traffic and
VALU / MFMA utilisation (rocprofv3 counters of this workload) are reported beside it.

Extra keys: the uniform N = 1M, M = 128 system, the 512-system batched mix and a short
served-request sweep (`--workload batched` / `--workload serving` run them in full).

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 (vector, packed; equal to its FP32 matrix peak)
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

# rocprofv3 counters of the default workloads, per solve (the commands and raw counts are
# in the profiles/ summaries named for the kernel and workload):
#   traffic_bytes = FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (KiB -> bytes; the factors come from
#                   a 512 MiB copy kernel launched in the same run, --calib-copy-mib 512)
#   valu_busy_pct = rocprofv3 VALUBusy = 100 SQ_ACTIVE_INST_VALU / CU_NUM / GRBM_GUI_ACTIVE
#   mfma_util_pct = rocprofv3 MfmaUtil = 100 SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE SIMD_NUM)
#   fma_insts     = SQ_INSTS_VALU_FMA_F32 (wave instructions, 64 lanes each) of all VALU
#                   instructions SQ_INSTS_VALU (which counts the MFMAs, SQ_INSTS_MFMA, too)
# The spread and the gather of the clustered workload are several kernels each (spread:
# column spread on the column items, beside the pencil register ring on the medium items and
# the plane spread on the sparse tiles on a second stream; gather: column gather on the
# column items and row gather on the others, concurrently on two streams); each is listed
# under its stage.
MEASURED = {
    ("clustered", 4_000_000, 16, 256, "fp32", 12): {
        "spread": {"traffic_bytes": 261.4e6 + 20.7e6 + 832.6e6,
                   "k_spread_col16": {"traffic_bytes": 261.4e6, "valu_busy_pct": 25.0, "mfma_util_pct": 31.7,
                                      "valu_insts": 1.0e8, "mfma_insts": 1.64e7, "lds_bank_conflicts": 1.05e7,
                                      "lds_insts": 9.99e6},
                   "k_spread_pencil": {"traffic_bytes": 20.7e6, "valu_busy_pct": 12.8, "valu_insts": 1.14e7,
                                       "fma_insts": 4.93e6, "lds_bank_conflicts": 1.2e5, "lds_insts": 1.23e6},
                   "k_spread_plane": {"traffic_bytes": 832.6e6, "valu_busy_pct": 78.4, "valu_insts": 2.64e8,
                                      "fma_insts": 1.78e8, "lds_bank_conflicts": 7.7e5, "lds_insts": 2.13e7}},
        "gather": {"traffic_bytes": 662.3e6 + 1245.7e6,
                   "k_interp_col16": {"traffic_bytes": 662.3e6, "mfma_util_pct": 24.8, "valu_busy_pct": 56.9,
                                      "valu_insts": 2.02e8, "mfma_insts": 2.29e7, "fma_insts": 5.78e7,
                                      "lds_bank_conflicts": 3.05e7, "lds_insts": 9.63e6},
                   "k_interp_row16": {"traffic_bytes": 1245.7e6, "valu_busy_pct": 77.9, "valu_insts": 2.38e8,
                                      "fma_insts": 1.2e8, "lds_bank_conflicts": 9.85e7, "lds_insts": 2.72e7}},
    },
    ("uniform", 1_000_000, 16, 128, "fp32", 12): {
        "spread": {"traffic_bytes": 137.4e6,
                   "k_spread_pencil": {"traffic_bytes": 137.4e6, "valu_busy_pct": 64.5, "valu_insts": 1.57e8,
                                       "fma_insts": 8.21e7, "lds_bank_conflicts": 8.19e5, "lds_insts": 1.51e7}},
        "gather": {"traffic_bytes": 277.8e6,
                   "k_interp_row16": {"traffic_bytes": 277.8e6, "valu_busy_pct": 70.6, "valu_insts": 2.07e8,
                                      "fma_insts": 1.41e8, "lds_bank_conflicts": 1.1e8, "lds_insts": 3.21e7}},
    },
}


def make_system(N, seed=2):
    import numpy as np
    rng = np.random.default_rng(seed)
    x = rng.random(3 * N)
    q = rng.uniform(-1.0, 1.0, N)
    q -= q.mean()
    return x, q


def make_clustered(N, seed=3, n_clusters=64, frac=0.8, sigma=0.01):
    """80 % of the charges in 64 isotropic Gaussian clusters (sigma = 0.01 L, centres
    uniform, wrapped into the box), 20 % uniform; q = +-1 in exact halves, shuffled."""
    import numpy as np
    rng = np.random.default_rng(seed)
    nc = int(frac * N)
    centres = rng.random((3, n_clusters))
    lab = rng.integers(0, n_clusters, nc)
    xc = (centres[:, lab] + sigma * rng.standard_normal((3, nc))) % 1.0
    xu = rng.random((3, N - nc))
    x = np.concatenate([xc, xu], axis=1)
    x = np.where(x >= 1.0, 0.0, x)   # (-tiny) % 1.0 can round to 1.0
    q = np.ones(N)
    q[N // 2:] = -1.0
    if N % 2:
        q[-1] = 0.0
    rng.shuffle(q)
    return np.ascontiguousarray(x.reshape(-1)), q


WORKLOADS = {
    # name: (N, M, P, xi, generator, description)
    "clustered": (4_000_000, 256, 16, 80.0, make_clustered,
                  "N={N}: 80% in 64 Gaussian clusters (sigma=0.01 L), 20% uniform, q=+-1 net zero"),
    "uniform": (1_000_000, 128, 16, 40.0, make_system,
                "N={N} uniform charges in L=1, charges uniform in [-1,1] shifted to zero sum"),
}


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(M, P, xi, n_sample, gen, n_full, what, runs=5, fft_runs=3):
    """The reference's own gridding core (oracle/_ref: SE_fgg.c built with its -mavx
    -DFGG_SPLIT flags) on one host core, on a random sample of n_sample particles of the
    workload: zero H (the reference's workspace allocation, SE_fgg.c:142) + expand (force
    factors) + AVX spread + zero H_per and wrap (spectral_ewald.c:65) + extend + AVX force
    interpolation.  One untimed warm-up, then the median of `runs` timed runs per stage.
    The FFT stage (forward + inverse c2c on M^3, as spectral_ewald.c:119-134 does with
    FFTW) is timed with numpy's pocketfft on one core, because FFTW is not installed on the
    box; it is reported beside the metric, which excludes the FFT."""
    import ctypes as C

    import numpy as np

    import oracle as O
    if not O.ref_available(O.P3):
        return None
    L = O.ref_lib(O.P3)
    xf, qf = gen(n_full)
    Nf = len(qf)
    n_sample = min(n_sample, Nf)
    rng = np.random.default_rng(11)
    pick = np.sort(rng.choice(Nf, n_sample, replace=False)) if n_sample < Nf else np.arange(Nf)
    x = np.ascontiguousarray(xf.reshape(3, Nf)[:, pick].reshape(-1))
    q = np.ascontiguousarray(qf[pick])
    m = 0.95 * np.sqrt(np.pi * P)
    h, w, eta, c = O.parse_params(1.0, M, P, m, xi)
    p = O.SE_FGG_params()
    L.SE_FGG_pack_params(C.byref(p), n_sample, M, M, M, P, c, h)
    st, keep = O.ref_state(x, q)
    wk = O.RefWork(p, force=True)
    Hper = O.aligned_zeros(M ** 3)
    F = O.aligned_zeros(3 * n_sample)
    L.SE_FGG_base_gaussian(C.byref(wk.w), C.byref(p))
    stages = ("zero", "expand", "spread", "wrap", "extend", "interp")
    times = {k: [] for k in stages}
    for it in range(runs + 1):
        t = [time.perf_counter()]
        wk.H[:] = 0.0
        t.append(time.perf_counter())
        L.SE_FGG_expand_all_SSE_force(C.byref(wk.w), C.byref(st), C.byref(p))
        t.append(time.perf_counter())
        L.SE_FGG_grid_split_AVX_dispatch_force(C.byref(wk.w), C.byref(st), C.byref(p))
        t.append(time.perf_counter())
        Hper[:] = 0.0
        L.SE_FGG_wrap_fcn(Hper, C.byref(wk.w), C.byref(p))
        t.append(time.perf_counter())
        L.SE_FGG_extend_fcn(C.byref(wk.w), Hper, C.byref(p))
        t.append(time.perf_counter())
        L.SE_FGG_int_split_AVX_dispatch_force(F, C.byref(st), C.byref(wk.w), C.byref(p))
        t.append(time.perf_counter())
        if it > 0:   # run 0 is the warm-up
            for k, a, b in zip(stages, t[:-1], t[1:]):
                times[k].append(b - a)
    med = {k: float(np.median(v)) for k, v in times.items()}
    total = sum(med.values())
    fft_s = None
    if fft_runs > 0:
        g = Hper.reshape(M, M, M, order="F").astype(np.complex128)
        ft = []
        for it in range(fft_runs + 1):
            t0 = time.perf_counter()
            gk = np.fft.fftn(g)
            g2 = np.fft.ifftn(gk)
            if it > 0:
                ft.append(time.perf_counter() - t0)
        del g, gk, g2
        fft_s = float(np.median(ft))
    del wk, keep
    return {
        "value": 2.0 * n_sample * P ** 3 / total,
        "unit": "particle*P^3 updates/s",
        "cores": 1,
        "kind": "reference",
        "sample": ("N=%d particles drawn at random from %s, P=%d, M=%d^3, FP64 force path of SE_fgg.c "
                   "(zero H, expand_all_SSE_force, grid_split_AVX_dispatch_force, zero H_per + wrap, extend, "
                   "int_split_AVX_dispatch_force); median of %d runs after 1 warm-up, %.1f s per run"
                   % (n_sample, what, P, M, runs, total)),
        "stage_s": {k: round(v, 4) for k, v in med.items()},
        "fft_stage_s": (round(fft_s, 3) if fft_s is not None else None),
        "fft_stage": ("forward + inverse 3-D c2c FFT of the M^3 grid with numpy.fft (pocketfft, 1 core), median of "
                      "%d after 1 warm-up; FFTW, which the reference links, is absent on the box; not part of the "
                      "metric" % fft_runs) if fft_s is not None else None,
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
    }


def algorithmic(N, P, M, force=True, esize=4):
    """Algorithmic flops and bytes of spread + gather: 2 flop per update in the spread, 4
    (potential + force) or 2 (potential) in the gather; per particle 32 B of x, q read
    twice and 32 B (8 B without force) of outputs; per grid point esize bytes zeroed,
    written and read once each."""
    upd = float(N) * P ** 3
    flops = upd * (2.0 + (4.0 if force else 2.0))
    bytes_ = N * (32.0 + 32.0 + (32.0 if force else 8.0)) + 3.0 * M ** 3 * esize
    return flops, bytes_


def measure(N, M, P, xi, gen, wdesc, prec_name, steps, warmup, check, tile=0, chunk=0, block=0, generic=False,
            mfma=True, column=True):
    """One plan, `warmup` + `steps` full solves timed by the wall clock, then `steps` solves
    with HIP events on the plan's stream around each stage; FP32 accuracy against the FP64
    P = 21 solve of the same system (FGG floor about 1e-13)."""
    import torch

    import se_fast_gridding_amd as se

    dev = torch.device("cuda:0")
    x, q = gen(N)
    xt = torch.tensor(x, device=dev)
    qt = torch.tensor(q, device=dev)
    phi = torch.empty(N, dtype=torch.float64, device=dev)
    F = torch.empty(3 * N, dtype=torch.float64, device=dev)
    prec = se.PREC_FP32 if prec_name == "fp32" else se.PREC_FP64
    stream = torch.cuda.current_stream()
    plan = se.Plan(M, P, 1.0, xi, N, precision=prec, stream=stream, tile=tile, chunk=chunk, block=block,
                   generic=generic, mfma=mfma, column=column)
    for _ in range(warmup):
        plan.solve(xt, qt, phi, F)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.solve(xt, qt, phi, F)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    plan.enable_timing(True)
    acc = {}
    for _ in range(steps):
        plan.solve(xt, qt, phi, F)
        for k, v in plan.timings().items():
            acc[k] = acc.get(k, 0.0) + v
    plan.enable_timing(False)
    ms = {k: v / steps for k, v in acc.items()}
    info = plan.info()
    plan.close()
    err = {}
    if check and prec == se.PREC_FP32:
        phi64 = torch.empty_like(phi)
        F64 = torch.empty_like(F)
        ref = se.Plan(M, 21, 1.0, xi, N, precision=se.PREC_FP64, stream=stream)
        ref.enable_timing(True)
        ref.solve(xt, qt, phi64, F64)
        t64 = ref.timings()
        torch.cuda.synchronize()
        ref.close()
        rr = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))  # noqa: E731
        err = {"phi_rms_rel_vs_fp64_p21": rr(phi, phi64), "force_rms_rel_vs_fp64_p21": rr(F, F64),
               "fp64_p21_updates_per_s": 2.0 * N * 21 ** 3 / ((t64["spread_stage"] + t64["interp_stage"]) * 1e-3),
               "fp64_p21_stage_ms": {k: round(v, 4) for k, v in t64.items()}}
        del phi64, F64
    del xt, qt, phi, F
    torch.cuda.empty_cache()
    upd = 2.0 * N * P ** 3
    t_si = (ms["spread_stage"] + ms["interp_stage"]) * 1e-3
    return {"N": N, "M": M, "P": P, "xi": xi, "value": upd / t_si, "solve_ms": wall * 1e3, "t_si_ms": t_si * 1e3,
            "stage_ms": {k: round(v, 4) for k, v in ms.items()}, "info": info, "accuracy": err,
            "desc": wdesc.format(N=N)}


def roofline(r, workload, prec_name):
    """Stage roofline of spread + gather, and the per-kernel fractions beside it."""
    N, M, P = r["N"], r["M"], r["P"]
    peak = FP32_PEAK_TFLOPS if prec_name == "fp32" else FP64_PEAK_TFLOPS
    flops, bytes_ = algorithmic(N, P, M, True, 4 if prec_name == "fp32" else 8)
    t_roof = max(flops / (peak * 1e12), bytes_ / (HBM_PEAK_GBS * 1e9))
    t_meas = r["t_si_ms"] * 1e-3
    ms = r["stage_ms"]
    tile = r["info"]["tile"]
    upd = float(N) * P ** 3
    kernels = {}
    meas_all = MEASURED.get((workload, N, P, M, prec_name, tile), {})
    for name, key, fl, kms in (("spread (k_spread_col16 + k_spread_pencil + k_spread_plane)", "spread", 2.0 * upd,
                                ms["spread_kernel"]),
                               ("gather (k_interp_col16 + k_interp_row16)", "gather", 4.0 * upd,
                                ms["gather_kernel"])):
        meas = meas_all.get(key, {})
        kernels[name] = {"ms": kms, "flops": fl, "achieved_tflops": fl / (kms * 1e-3) / 1e12,
                         "frac": fl / (kms * 1e-3) / 1e12 / peak, "counters": meas or None}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    return {"bound": "mfma",
            "compute": ("FP32 results: matrix cores for the dense items (column spread: v_mfma_f32_16x16x4f32; "
                        "column gather: its z contraction as three v_mfma_f32_16x16x16_f16 on fp16 (hi, lo) pairs "
                        "of the FP32 operands, hi*hi + lo*hi + hi*lo with FP32 sums, about 2^-22 per product) and "
                        "vector FP32 FMA (pencil, plane and row kernels of the others); peak is MI355X dense FP32, "
                        "157.3 TFLOP/s (the FP32 rate of both units), so frac is against FP32 arithmetic"),
            "achieved": flops / t_meas / 1e12, "peak": peak, "unit": "TFLOP/s",
            "frac": t_roof / t_meas, "t_roof_ms": t_roof * 1e3, "t_measured_ms": t_meas * 1e3,
            "flops": flops, "bytes_algorithmic": bytes_,
            "definition": ("t_roof = max(F / peak, B / 8 TB/s), F = 6 flop per particle*P^3 update (spread 2, "
                           "potential+force gather 4), B = algorithmic bytes; frac = t_roof / (spread stage + "
                           "interp stage), both stages timed with HIP events on the plan's stream"),
            "traffic": (kernels[dom]["counters"] or {}).get("traffic_bytes"), "traffic_kernel": dom,
            "kernels": kernels}


def run(args):
    wN, wM, wP, wxi, gen, wdesc = WORKLOADS[args.workload]
    N, M, P, xi = args.n or wN, args.m or wM, args.p or wP, args.xi or wxi
    r = measure(N, M, P, xi, gen, wdesc, args.precision, args.steps, args.warmup, args.check, args.tile,
                args.chunk, args.block, args.generic, args.mfma, args.column)
    rl = roofline(r, args.workload, args.precision)
    extras = {}
    if args.extras and args.workload == "clustered" and not (args.n or args.m or args.p or args.xi):
        u = measure(*WORKLOADS["uniform"][:4], make_system, WORKLOADS["uniform"][5], args.precision, args.steps,
                    args.warmup, args.check)
        ur = roofline(u, "uniform", args.precision)
        extras["uniform_1M_M128"] = {"workload": "SE3P potential+force: " + u["desc"] + ", P=16, M=128^3, xi=40",
                                     "value": u["value"], "solve_ms": u["solve_ms"], "stage_ms": u["stage_ms"],
                                     "roofline_frac": ur["frac"], "kernels": ur["kernels"],
                                     "accuracy": u["accuracy"]}
        extras["batched_512"] = batched(args.steps_batched, args.slots, tune=True, quick=True)
        extras["serving"] = serving(args.slots, [float(v) for v in args.extra_rates.split(",")],
                                    args.extra_requests, check=True)
    cpu = None
    n_cpu = args.cpu_sample if args.cpu_sample >= 0 else min(N, 500_000)
    if n_cpu > 0:
        cpu = cpu_baseline(M, P, xi, n_cpu, gen, N, "the same %s system of %d particles" % (args.workload, N),
                           args.cpu_runs, args.cpu_fft_runs)
    out = {
        "metric": "spread+interp particle*P^3 updates/s",
        "value": r["value"],
        "unit": "particle*P^3 updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["solve_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": "SE3P potential+force, single system: %s, P=%d, M=%d^3, xi=%g" % (r["desc"], P, M, xi),
                   "N": N, "P": P, "M": M, "xi": xi, "precision": args.precision, "quantities": "phi+force",
                   "tile": r["info"]["tile"], "chunk": r["info"]["chunk"], "block": r["info"]["block"],
                   "work_items": r["info"]["items"]},
        "solve_ms": r["solve_ms"],
        "stage_ms": r["stage_ms"],
        "roofline": rl,
        "accuracy": r["accuracy"],
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out))


def mixed_systems(n_systems, seed=4, se2p_frac=0.0):
    """The batched workload: N log-uniform in [1e3, 1e5], M in {32, 48, 64, 96}, xi = 0.3125 M,
    P in {11, 16}, x uniform, q = +-1 with zero net charge; a fraction of the systems SE2P
    (unit square periodic, slab height 1)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_systems):
        N = int(np.floor(np.exp(rng.uniform(np.log(1e3), np.log(1e5)))))
        M = int(rng.choice([32, 48, 64, 96]))
        P = int(rng.choice([11, 16]))
        per = 2 if rng.random() < se2p_frac else 3
        x = rng.random(3 * N)
        q = np.ones(N)
        q[N // 2:] = -1.0
        if N % 2:
            q[-1] = 0.0
        rng.shuffle(q)
        out.append(dict(N=N, M=M, P=P, xi=0.3125 * M, per=per, x=x, q=q))
    return out


def _device_jobs(systems, prec, P_override=None):
    import torch
    import se_fast_gridding_amd as se
    jobs = []
    for s in systems:
        x = torch.tensor(s["x"], device="cuda")
        q = torch.tensor(s["q"], device="cuda")
        phi = torch.zeros_like(q)
        job = se.Batch.job(s["M"], P_override or s["P"], 1.0, s["xi"], x, q, phi, precision=prec,
                           periodicity=s["per"], slab=1.0 if s["per"] == 2 else 0.0)
        jobs.append((job, x, q, phi))
    return jobs


def fp64_p21_reference(systems):
    """phi of every system from the FP64 P = 21 solve (the FP64 oracle proxy: its FGG error
    floor is about 1e-13), solved one after another through the batch executor."""
    import se_fast_gridding_amd as se
    jobs = _device_jobs(systems, se.PREC_FP64, P_override=21)
    b = se.Batch(slots=1)
    for j in jobs:
        b.submit(j[0])
    b.wait_all()
    out = [j[3].cpu().numpy() for j in jobs]
    b.close()
    return out


def batched(steps, slots, tune=True, quick=False, n_sys=512, grouped_only=False):
    """512 independent systems: FP32 grouped solves (se_batch_run: one group plan per
    geometry, one launch per stage per group, groups overlapping on the slots) vs the same
    systems solved one after another on the FP64 path at their own P (plans cached per
    shape, one stream).  Errors are against the FP64 P = 21 solve of each system."""
    import numpy as np
    import torch
    import se_fast_gridding_amd as se
    systems = mixed_systems(n_sys)
    upd = sum(2.0 * s["N"] * s["P"] ** 3 for s in systems)

    def timed(prec, nslots, grouped):
        jobs = _device_jobs(systems, prec)
        b = se.Batch(slots=nslots)
        run_all = (lambda: b.wait(b.run([j[0] for j in jobs]))) if grouped else \
            (lambda: ([b.submit(j[0]) for j in jobs], b.wait_all()))
        run_all()                            # warm-up: plans (and graphs) for every shape
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run_all()
        dt = (time.perf_counter() - t0) / steps
        phis = [j[3].cpu().numpy() for j in jobs]
        st = b.stats()
        b.close()
        return dt, phis, st

    tuned = {}
    t32u = None
    tune_s = 0.0
    if tune:
        se.capi().se_tune_clear()
        if not quick:
            t32u, _, _ = timed(se.PREC_FP32, slots, True)
        shapes = {}
        for s in systems:
            shapes.setdefault((s["per"], s["M"], s["P"], s["xi"]), []).append(s["N"])
        t0 = time.perf_counter()
        for (per, M, P, xi), ns in sorted(shapes.items()):
            t, c, bk, ms = se.tune(M, P, 1.0, xi, int(np.mean(ns)), precision=se.PREC_FP32, periodicity=per,
                                   slab=1.0 if per == 2 else 0.0, potential=True, force=False, systems=len(ns))
            tuned["SE%dP M=%d P=%d" % (per, M, P)] = {"tile": t, "chunk": c, "block": bk, "group_ms": round(ms, 4),
                                                     "N": int(np.mean(ns)), "systems": len(ns)}
        tune_s = time.perf_counter() - t0
    t32, phi32, st32 = timed(se.PREC_FP32, slots, True)
    if grouped_only:   # for kernel traces of the grouped passes alone
        return {"solves_per_s": n_sys / t32, "ms": t32 * 1e3, "mode": "grouped (se_batch_run), FP32, alone"}
    t32s, phi32s, _ = timed(se.PREC_FP32, slots, False)
    t64, phi64, _ = timed(se.PREC_FP64, 1, False)
    if tune:
        se.capi().se_tune_clear()
    ref = fp64_p21_reference(systems)

    def errs_of(p):
        return np.array([float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(p, ref)])
    e32, e32s, e64 = errs_of(phi32), errs_of(phi32s), errs_of(phi64)
    p11 = np.array([s["P"] == 11 for s in systems])
    out = {
        "workload": "%d independent SE3P potential systems: N log-uniform in [1e3, 1e5], M in {32,48,64,96}^3, "
                    "xi = 0.3125 M, P in {11,16}, q = +-1 net zero" % n_sys,
        "solves_per_s": n_sys / t32, "ms": t32 * 1e3, "mode": "grouped (se_batch_run), FP32", "slots": slots,
        "particles": int(sum(s["N"] for s in systems)), "updates_per_s": upd / t32,
        "baseline_fp64_sequential": {"solves_per_s": n_sys / t64, "ms": t64 * 1e3},
        "speedup_vs_fp64_sequential": t64 / t32,
        "phi_rms_rel_vs_fp64_p21": {"max": float(e32.max()), "median": float(np.median(e32)),
                                    "max_P16": float(e32[~p11].max()), "max_P11": float(e32[p11].max())},
        "fp64_baseline_phi_rms_rel_vs_fp64_p21": {"max": float(e64.max()), "median": float(np.median(e64))},
        "per_job_slots": {"solves_per_s": n_sys / t32s, "ms": t32s * 1e3, "speedup_vs_fp64_sequential": t64 / t32s,
                          "phi_rms_rel_vs_fp64_p21_max": float(e32s.max())},
        "plan_cache": st32,
    }
    if tune:
        out["autotune"] = {"tilings": tuned, "tune_s": round(tune_s, 2)}
        if t32u is not None:
            out["autotune"].update({"untuned_solves_per_s": n_sys / t32u, "tuned_over_untuned": t32u / t32})
    return out


def serving(slots, rates, requests, check=True, pool_size=128, max_seconds=120.0, capture=True):
    """Poisson arrivals of SE3P / SE2P potential solves (50 / 50, sizes as the batched mix)
    against the executor: a request is admitted when a slot is idle (try_submit), results
    and inputs stay in HBM; latency = completion seen by the host - arrival.  Every rate
    runs `requests` requests (fewer if that would take longer than max_seconds of
    arrivals).  After each rate every pool system's last served phi is compared with its
    FP64 P = 21 solve and, bit for bit, with its solve alone."""
    import collections
    import numpy as np
    import torch
    import se_fast_gridding_amd as se
    pool = mixed_systems(pool_size, seed=5, se2p_frac=0.5)
    jobs = _device_jobs(pool, se.PREC_FP32)
    b = se.Batch(slots=slots)
    biggest = {}                                 # warm-up: every shape's plan on every slot
    for j in jobs:
        sh = j[0].shape
        k = (sh.periodicity, sh.M, sh.P)
        if k not in biggest or j[0].N > biggest[k].N:
            biggest[k] = j[0]
    for job in biggest.values():
        b.prepare(job)
    alone = []                                   # each system solved with no other request in flight
    for j in jobs:
        b.wait(b.submit(j[0]))
        alone.append(j[3].clone())
    if capture:
        # capture every job's solve on every slot ahead of traffic (submit goes round the
        # slots), so that try_submit replays a graph: one launch per request
        for j in jobs:
            for _ in range(slots):
                b.submit(j[0])
            b.wait_all()
    ref = fp64_p21_reference(pool) if check else None
    rng = np.random.default_rng(5)
    table = []
    for lam in rates:
        K = int(min(requests, max(200, lam * max_seconds)))
        arrivals = np.cumsum(rng.exponential(1.0 / lam, K))
        busy = {}                                # pool index -> ticket in flight
        inflight = {}                            # ticket -> (request, arrival)
        queue = collections.deque()
        lat = []
        i = 0
        for j in jobs:
            j[3].fill_(float("nan"))
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        while len(lat) < K:
            now = time.perf_counter() - t_start
            for t in list(inflight):
                if b.poll(t):
                    r, a = inflight.pop(t)
                    lat.append(now - a)
                    busy.pop(r % len(jobs), None)
            while i < K and arrivals[i] <= now:
                queue.append(i)
                i += 1
            while queue:
                r = queue[0]
                if (r % len(jobs)) in busy:      # that system's buffers are still in use
                    break
                t = b.try_submit(jobs[r % len(jobs)][0])
                if t is None:
                    break
                inflight[t] = (r, arrivals[r])
                busy[r % len(jobs)] = t
                queue.popleft()
        t_end = time.perf_counter() - t_start
        b.wait_all()
        lat = np.array(lat) * 1e3
        row = {"rate_req_s": lam, "requests": K, "solves_per_s": K / t_end,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "mean_ms": float(lat.mean())}
        if check:
            served = [j[3].cpu().numpy() for j in jobs]
            e = np.array([float(np.linalg.norm(a - r_) / np.linalg.norm(r_)) for a, r_ in zip(served, ref)])
            row["phi_rms_rel_vs_fp64_p21"] = {"max": float(e.max()), "median": float(np.median(e))}
            row["served_equal_to_alone"] = bool(all(np.array_equal(a, s.cpu().numpy()) for a, s in zip(served, alone)))
        table.append(row)
        print("serving: rate %g req/s: %d requests, %.0f solves/s, p50 %.3f ms, p99 %.3f ms"
              % (lam, K, row["solves_per_s"], row["p50_ms"], row["p99_ms"]), file=sys.stderr, flush=True)
    b.close()
    sat = None
    for row in table:
        if row["solves_per_s"] < 0.95 * row["rate_req_s"]:
            sat = row["rate_req_s"]
            break
    return {"workload": "Poisson arrivals of SE3P/SE2P potential solves (50/50), sizes as the batched mix, "
                        "%d device-resident systems, FP32" % pool_size, "slots": slots,
            "graphs": "every job captured on every slot before traffic; requests replay graphs" if capture
            else "requests solve eagerly",
            "best_solves_per_s": max(r["solves_per_s"] for r in table),
            "first_rate_not_sustained": sat, "rates": table}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["batched", "serving"], default="clustered",
                    help="clustered: N=4M, M=256, xi=80 (the headline line); uniform: N=1M, M=128, xi=40; "
                         "batched: 512 independent systems; serving: Poisson request stream")
    ap.add_argument("--slots", type=int, default=16, help="executor streams (batched / serving)")
    ap.add_argument("--rates", default="100,200,500,1000,2000,5000,10000,20000,50000",
                    help="serving arrival rates, req/s")
    ap.add_argument("--requests", type=int, default=10000, help="serving: requests per rate")
    ap.add_argument("--n", type=int, default=0, help="particles (0 = the workload's)")
    ap.add_argument("--m", type=int, default=0, help="grid side (0 = the workload's)")
    ap.add_argument("--p", type=int, default=0, help="Gaussian support (0 = the workload's)")
    ap.add_argument("--xi", type=float, default=0.0, help="Ewald parameter (0 = the workload's)")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="particles in the CPU reference run (-1 = min(N, 500000), 0 = skip)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU runs (median), after one warm-up")
    ap.add_argument("--cpu-fft-runs", type=int, default=3, help="timed CPU FFT stages (median; 0 = skip)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="launch one FP32 copy of this many MiB first (HBM counter calibration)")
    ap.add_argument("--no-check", dest="check", action="store_false")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="skip the uniform, batched and serving extra keys of the default line")
    ap.add_argument("--steps-batched", type=int, default=3, help="timed passes of the batched extra")
    ap.add_argument("--grouped-only", action="store_true",
                    help="batched: time the grouped FP32 passes only (no per-job or FP64 baselines)")
    ap.add_argument("--extra-rates", default="1000,5000,20000", help="arrival rates of the serving extra")
    ap.add_argument("--extra-requests", type=int, default=3000, help="requests per rate of the serving extra")
    ap.add_argument("--tile", type=int, default=0, help="tile side (0 = plan default)")
    ap.add_argument("--chunk", type=int, default=0, help="max particles per work item (0 = default)")
    ap.add_argument("--block", type=int, default=0, help="threads per work-group (0 = default)")
    ap.add_argument("--generic", action="store_true", help="one-particle-per-wave kernels instead of pencils")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="batched: skip the tiling autotuner (se_tune) before the timed runs")
    ap.add_argument("--no-mfma", dest="mfma", action="store_false",
                    help="P = 16 FP32 gather on the vector register-ring kernel instead of the matrix cores")
    ap.add_argument("--no-column", dest="column", action="store_false",
                    help="P = 16 FP32 dense items on the 4 x 1 pencil kernels instead of the 1 x 1 column kernels")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has exactly one; "
              "multi-GPU sharding is a design note, not a measured path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy_mib > 0:
        # one elementwise FP32 copy of known size (reads and writes calib_copy_mib MiB),
        # launched first, so a counter run sees a reference kernel beside the solve
        import torch
        src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "batched":
        out = batched(args.steps, args.slots, tune=args.tune, grouped_only=args.grouped_only)
        out.update({"metric": "batched solves/s", "value": out["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
                    "steps": args.steps, "higher_is_better": True, "dtype": "f32", "data": "synthetic"})
        print(json.dumps(out))
    elif args.workload == "serving":
        out = serving(args.slots, [float(r) for r in args.rates.split(",")], args.requests)
        out.update({"metric": "served solves/s", "value": out["best_solves_per_s"], "unit": "solves/s",
                    "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic"})
        print(json.dumps(out))
    else:
        run(args)


if __name__ == "__main__":
    main()

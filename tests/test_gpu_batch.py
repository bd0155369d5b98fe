"""The batch/serving executor (se_batch_*): results of a system are the same alone and in a
mix of concurrent jobs, admission refuses work when no slot is idle, and the HBM budget
is enforced."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se
from conftest import rms_rel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def system(N, seed, per=3):
    rng = np.random.default_rng(seed)
    x = rng.random((3, N))
    q = rng.choice([-1.0, 1.0], N)
    q -= q.mean()
    return torch.tensor(x.reshape(-1), device=DEV), torch.tensor(q, device=DEV)


def alone(spec, x, q):
    per, M, P, xi, prec = spec
    plan = se.Plan(M, P, 1.0, xi, q.numel(), precision=prec, periodicity=per, slab=1.0 if per == 2 else 0.0)
    phi = torch.zeros_like(q)
    F = torch.zeros_like(x)
    plan.solve(x, q, phi, F)
    torch.cuda.synchronize()
    plan.close()
    return phi.cpu().numpy(), F.cpu().numpy()


SPECS = [  # periodicity, M, P, xi, precision
    (3, 32, 11, 10.0, se.PREC_FP32), (3, 48, 16, 15.0, se.PREC_FP32), (2, 32, 16, 10.0, se.PREC_FP32),
    (3, 32, 16, 10.0, se.PREC_FP64), (2, 24, 11, 7.5, se.PREC_FP64), (3, 64, 11, 20.0, se.PREC_FP32),
]


def test_results_are_the_same_alone_and_in_the_mix():
    rng = np.random.default_rng(4)
    jobs = []
    for i in range(30):
        spec = SPECS[i % len(SPECS)]
        N = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
        x, q = system(N, 100 + i, spec[0])
        jobs.append((spec, x, q, torch.zeros_like(q), torch.zeros_like(x)))
    b = se.Batch(slots=8)
    tickets = []
    for spec, x, q, phi, F in jobs:
        per, M, P, xi, prec = spec
        tickets.append(b.submit(se.Batch.job(M, P, 1.0, xi, x, q, phi, F, precision=prec, periodicity=per,
                                             slab=1.0 if per == 2 else 0.0)))
    b.wait_all()
    assert all(b.poll(t) for t in tickets)
    st = b.stats()
    assert st["plans"] >= len(SPECS) and st["plan_bytes"] > 0
    for spec, x, q, phi, F in jobs:
        pa, Fa = alone(spec, x, q)
        if spec[4] == se.PREC_FP32:   # bit for bit: same plan shape, deterministic kernels
            assert np.array_equal(phi.cpu().numpy(), pa) and np.array_equal(F.cpu().numpy(), Fa), (spec, q.numel())
            continue
        e_phi, e_F = rms_rel(phi.cpu().numpy(), pa), rms_rel(F.cpu().numpy(), Fa)
        assert e_phi <= 1e-12 and e_F <= 1e-12, (spec, q.numel(), e_phi, e_F)
    b.close()


def test_try_submit_refuses_when_no_slot_is_idle():
    b = se.Batch(slots=1)
    x, q = system(1_000_000, 7)
    phi = torch.zeros_like(q)
    small_x, small_q = system(1000, 8)
    small_phi = torch.zeros_like(small_q)
    job = se.Batch.job(128, 16, 1.0, 40.0, x, q, phi)
    t0 = b.try_submit(job)                 # plan is built, then the solve is queued
    assert t0 is not None
    t1 = b.try_submit(se.Batch.job(32, 11, 1.0, 10.0, small_x, small_q, small_phi))
    assert t1 is None                      # the only slot still runs the 1M-particle solve
    b.wait(t0)
    assert b.poll(t0)
    t2 = b.try_submit(se.Batch.job(32, 11, 1.0, 10.0, small_x, small_q, small_phi))
    assert t2 is not None
    b.wait_all()
    b.close()


def test_hbm_budget_is_enforced():
    b = se.Batch(slots=2, hbm_budget=1 << 20)   # 1 MiB: no plan fits
    x, q = system(1000, 9)
    with pytest.raises(RuntimeError, match="budget"):
        b.submit(se.Batch.job(32, 11, 1.0, 10.0, x, q, torch.zeros_like(q)))
    b.close()


# 3p_fp32_split: N = 1.1M on M = 160 (14^3 tiles of side 12, ~400 particles each, 4-wide
# edge tiles), so the FP32 P = 16 gather runs the matrix-core kernel (full tiles) and the
# vector kernel (edge tiles) concurrently on two streams, and the capture holds both
@pytest.mark.parametrize("spec", [(3, 48, 16, 15.0, se.PREC_FP32, 6000), (2, 32, 11, 10.0, se.PREC_FP32, 6000),
                                  (3, 32, 16, 10.0, se.PREC_FP64, 6000), (3, 160, 16, 50.0, se.PREC_FP32, 1_100_000)],
                         ids=["3p_fp32", "2p_fp32", "3p_fp64", "3p_fp32_split"])
def test_graph_replays_match_eager_solves(spec):
    """se_solve_graph: every replay equals the eager solve of the same inputs, with eager
    work and synchronisations between replays, and follows new array contents."""
    per, M, P, xi, prec, N = spec
    # FP32 replays equal the eager solve bit for bit; FP64 LDS sums repeat to rounding
    tol = 1e-12 if prec == se.PREC_FP64 else 0.0
    x, q = system(N, 31)
    plan = se.Plan(M, P, 1.0, xi, N, precision=prec, periodicity=per, slab=1.0 if per == 2 else 0.0,
                   stream=torch.cuda.Stream())
    phi_e, F_e = torch.zeros_like(q), torch.zeros_like(x)
    phi_g, F_g = torch.zeros_like(q), torch.zeros_like(x)
    with torch.cuda.stream(plan.stream):
        plan.solve(x, q, phi_e, F_e)
    torch.cuda.synchronize()
    ref_phi, ref_F = phi_e.clone(), F_e.clone()
    for rep in range(4):
        phi_g.fill_(float("nan"))
        F_g.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.solve_graph(x, q, phi_g, F_g)
        torch.cuda.synchronize()
        scratch = torch.randn(1 << 20, device=DEV).sum()   # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(scratch)
        assert rms_rel(phi_g.cpu().numpy(), ref_phi.cpu().numpy()) <= tol, rep
        assert rms_rel(F_g.cpu().numpy(), ref_F.cpu().numpy()) <= tol, rep
    # same tensors, new contents: the graph reads the arrays when it runs
    x2, q2 = system(N, 32)
    x.copy_(x2)
    q.copy_(q2)
    torch.cuda.synchronize()
    plan.solve_graph(x, q, phi_g, F_g)
    with torch.cuda.stream(plan.stream):
        plan.solve(x, q, phi_e, F_e)
    torch.cuda.synchronize()
    assert rms_rel(phi_g.cpu().numpy(), phi_e.cpu().numpy()) <= tol
    assert rms_rel(F_g.cpu().numpy(), F_e.cpu().numpy()) <= tol
    plan.close()


def test_tune_records_a_tiling_that_new_plans_use():
    """se_tune times the candidate tilings and records the fastest; a plan created
    afterwards with the automatic tiling takes it and solves the same system to FP32
    rounding of the default plan."""
    lib = se.capi()
    lib.se_tune_clear()
    try:
        M, P, xi, N = 32, 16, 10.0, 20000
        t, c, b, ms = se.tune(M, P, 1.0, xi, N)
        # block 0: each kernel's own default (the plan reports the nominal 512)
        assert t in (4, 8, 12) and c in (512, 1024, 2048, 4096) and b in (0, 256) and ms > 0
        assert se.tune_lookup(M, P, 1.0, xi, N) == (t, c, b)
        rng = np.random.default_rng(3)
        x = torch.tensor(rng.random(3 * N), device="cuda")
        q = torch.tensor(rng.uniform(-1, 1, N), device="cuda")
        out = []
        for tuned in (True, False):
            if not tuned:
                lib.se_tune_clear()
            plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32)
            info = plan.info()
            if tuned:
                assert (info["tile"], info["chunk"], info["block"]) == (t, c, b or 512)
            phi = torch.zeros(N, dtype=torch.float64, device="cuda")
            plan.solve(x, q, phi, None)
            torch.cuda.synchronize()
            plan.close()
            out.append(phi.cpu().numpy())
        assert np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1]) < 2e-6
    finally:
        lib.se_tune_clear()

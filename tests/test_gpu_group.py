"""Grouped solves (se_solve_group, se_batch_run): several independent systems of one
geometry in one pass.  Every system's result is checked against the oracle's SE3P
solve (SE3P, FP64) or the SE2P direct sum, and (SE3P, SE2P, SE1P) against the same
system solved alone on a plain plan; groups hold uneven sizes, empty and one-particle
systems, and the executor's grouping by geometry and output mode is checked on a mixed
job list."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se
from conftest import rms_rel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def system(N, seed, slab=None):
    rng = np.random.default_rng(seed)
    x = rng.random((3, N))
    if slab is not None:
        x[2] *= slab
    q = rng.uniform(-1, 1, N)
    if N > 1:
        q -= q.mean()
    return np.ascontiguousarray(x.reshape(-1)), q


def to_dev(x, q, force=True):
    N = len(q)
    xt = torch.tensor(x, device=DEV)
    qt = torch.tensor(q, device=DEV)
    phi = torch.full((N,), float("nan"), dtype=torch.float64, device=DEV)
    F = torch.full((3 * N,), float("nan"), dtype=torch.float64, device=DEV) if force else None
    return xt, qt, phi, F


def solve_alone(x, q, M, P, xi, prec, per=3, force=True):
    N = len(q)
    plan = se.Plan(M, P, 1.0, xi, max(N, 1), precision=prec, periodicity=per, slab=1.0 if per != 3 else 0.0)
    xt, qt, phi, F = to_dev(x, q, force)
    plan.solve(xt, qt, phi, F)
    torch.cuda.synchronize()
    plan.close()
    return phi.cpu().numpy(), (F.cpu().numpy() if force else None)


SIZES = [300, 0, 1, 2000, 57, 4096, 999, 12000]


@pytest.mark.parametrize("prec", [se.PREC_FP64, se.PREC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("per", [3, 2, 1], ids=["se3p", "se2p", "se1p"])
def test_group_equals_each_system_alone(per, prec):
    M, P, xi = 32, 16 if prec == se.PREC_FP32 else 21, 10.0
    print("case group per=%d prec=%d sizes=%s" % (per, prec, SIZES), flush=True)
    systems = [system(N, 40 + i, slab=1.0 if per == 2 else None) for i, N in enumerate(SIZES)]
    dev = [to_dev(x, q) for x, q in systems]
    plan = se.Plan(M, P, 1.0, xi, sum(SIZES), precision=prec, periodicity=per, slab=1.0 if per != 3 else 0.0,
                   max_systems=len(SIZES) + 3)
    plan.solve_group(dev)
    torch.cuda.synchronize()
    info = plan.info()
    assert info["discarded"] == 0
    plan.close()
    # FP32: the grouped solve repeats the solve alone bit for bit (same tiling, stable
    # particle order, integer-valued LDS sums); FP64 LDS sums repeat to rounding
    tol = 1e-12 if prec == se.PREC_FP64 else 0.0
    for (x, q), (_, _, phi, F), N in zip(systems, dev, SIZES):
        pa, Fa = solve_alone(x, q, M, P, xi, prec, per)
        pg, Fg = phi.cpu().numpy(), F.cpu().numpy()
        assert np.all(np.isfinite(pg)) and np.all(np.isfinite(Fg)), N
        if N == 0:
            continue
        if N == 1:   # a single charge: compare absolutely, at the precision's scale
            atol = (1e-10 if prec == se.PREC_FP64 else 0.0) * max(1.0, float(np.max(np.abs(pa))))
            assert np.allclose(pg, pa, atol=atol, rtol=0) and np.allclose(Fg, Fa, atol=atol, rtol=0), N
            continue
        e_phi, e_F = rms_rel(pg, pa), rms_rel(Fg, Fa)
        assert e_phi <= tol and e_F <= tol, (N, e_phi, e_F)


def test_group_se3p_fp64_matches_the_oracle(oracle):
    M, P, xi = 24, 21, 7.0
    m = 0.95 * np.sqrt(np.pi * P)
    sizes = [40, 75, 13]
    systems = [system(N, 70 + i) for i, N in enumerate(sizes)]
    dev = [to_dev(x, q) for x, q in systems]
    plan = se.Plan(M, P, 1.0, xi, sum(sizes), precision=se.PREC_FP64, max_systems=4)
    plan.solve_group(dev)
    torch.cuda.synchronize()
    plan.close()
    for (x, q), (_, _, phi, F) in zip(systems, dev):
        assert rms_rel(phi.cpu().numpy(), oracle.se3p(x, q, 1.0, M, P, m, xi)) < 1e-12
        assert rms_rel(F.cpu().numpy(), oracle.se3p(x, q, 1.0, M, P, m, xi, force=True)) < 1e-12


def test_group_se2p_fp64_matches_the_direct_sum(oracle):
    M, P, xi = 32, 21, 6.0
    sizes = [30, 45]
    systems = [system(N, 90 + i, slab=1.0) for i, N in enumerate(sizes)]
    dev = [to_dev(x, q, force=False) for x, q in systems]
    plan = se.Plan(M, P, 1.0, xi, sum(sizes), precision=se.PREC_FP64, periodicity=2, slab=1.0, max_systems=2)
    plan.solve_group(dev)
    torch.cuda.synchronize()
    plan.close()
    for (x, q), (_, _, phi, _) in zip(systems, dev):
        assert rms_rel(phi.cpu().numpy(), oracle.direct2p(x, q, (1.0, 1.0, 1.0), xi, 14)) < 1e-11


def test_group_rejects_mixed_outputs_and_capacity():
    x, q = system(100, 1)
    a = to_dev(x, q, force=True)
    b = to_dev(x, q, force=False)
    plan = se.Plan(32, 11, 1.0, 10.0, 200, precision=se.PREC_FP32, max_systems=2)
    with pytest.raises(RuntimeError, match="same outputs"):
        plan.solve_group([a, b])
    with pytest.raises(RuntimeError, match="capacity"):
        plan.solve_group([a, a, a])
    big = to_dev(*system(300, 2))
    with pytest.raises(RuntimeError, match="capacity"):
        plan.solve_group([big])
    plan.close()


def test_group_plan_reused_with_fewer_and_more_systems():
    """Calls in a row with a different number of systems (batched FFT plans per size)."""
    M, P, xi = 48, 11, 15.0
    plan = se.Plan(M, P, 1.0, xi, 40000, precision=se.PREC_FP32, max_systems=6)
    for S in (5, 2, 6, 2):
        systems = [system(3000 + 500 * i, 200 + 10 * S + i) for i in range(S)]
        dev = [to_dev(x, q, force=False) for x, q in systems]
        plan.solve_group(dev)
        torch.cuda.synchronize()
        for (x, q), (_, _, phi, _) in zip(systems, dev):
            pa, _ = solve_alone(x, q, M, P, xi, se.PREC_FP32, force=False)
            assert np.array_equal(phi.cpu().numpy(), pa), S
    plan.close()


def test_batch_run_groups_a_mixed_job_list():
    """se_batch_run: jobs of several geometries and output modes, shuffled, in one call,
    twice in a row (the second call reuses the cached group plans)."""
    rng = np.random.default_rng(11)
    shapes = [(3, 32, 11, 10.0), (3, 48, 16, 15.0), (2, 32, 16, 10.0), (3, 32, 11, 10.0)]
    jobs, keep = [], []
    for i in range(40):
        per, M, P, xi = shapes[i % len(shapes)]
        N = int(np.exp(rng.uniform(np.log(200), np.log(8000))))
        x, q = system(N, 500 + i, slab=1.0 if per == 2 else None)
        want_force = (i % 3 == 0)
        xt, qt, phi, F = to_dev(x, q, force=want_force)
        keep.append((per, M, P, xi, x, q, phi, F))
        jobs.append(se.Batch.job(M, P, 1.0, xi, xt, qt, phi, F, precision=se.PREC_FP32, periodicity=per,
                                 slab=1.0 if per == 2 else 0.0))
        keep[-1] += (xt, qt)
    b = se.Batch(slots=4)
    for rep in range(2):
        for k in keep:
            k[6].fill_(float("nan"))
        torch.cuda.synchronize()
        t = b.run(jobs)
        b.wait(t)
        assert b.poll(t)
        for per, M, P, xi, x, q, phi, F, _, _ in keep:
            pa, Fa = solve_alone(x, q, M, P, xi, se.PREC_FP32, per, force=F is not None)
            assert np.array_equal(phi.cpu().numpy(), pa), (rep, per, M, P, len(q))
            if F is not None:
                assert np.array_equal(F.cpu().numpy(), Fa), (rep, per, M, P, len(q))
    st = b.stats()
    assert st["plans"] >= 3
    b.close()

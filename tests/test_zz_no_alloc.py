"""se_real_solve_noalloc, the plan-free real-space solve that allocates nothing (its buffers
live in the library's static scratch array), against se_real_solve on a plan of capacity N:
the same cells and the same kernels, so the same arrays bit for bit.  Two calls in a row,
the second with a smaller input, so a stale scratch area from the first call must not
leak into the second; accumulate mode; and the capacity refusal."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def system(N, seed, per):
    rng = np.random.default_rng(seed)
    x = rng.random((3, N))
    if per == 3:   # periodic directions fold any position into the box
        x = x * 3.0 - 1.0
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    return torch.tensor(x.reshape(-1), device=DEV), torch.tensor(q, device=DEV)


# N, periodicity, box, xi, rc, self term, accumulate
CALLS = [(40000, 3, 1.0, 14.0, 0.09, True, False), (7000, 2, 1.0, 9.0, 0.2, False, True),
         (3000, 1, (1.0, 1.0, 2.0), 6.0, 0.3, True, False)]


def test_noalloc_matches_the_plan_solve_bit_for_bit():
    for i, (N, per, box, xi, rc, self_term, acc) in enumerate(CALLS):
        print("call", i, N, per, flush=True)
        x, q = system(N, 60 + i, per)
        init_phi = torch.randn(N, dtype=torch.float64, device=DEV) if acc else torch.zeros(N, dtype=torch.float64,
                                                                                           device=DEV)
        init_F = torch.randn(3 * N, dtype=torch.float64, device=DEV) if acc else torch.zeros(3 * N, dtype=torch.float64,
                                                                                             device=DEV)
        rs = se.RealSpace(per, box, xi, rc, N, self_term=self_term, accumulate=acc)
        pp, Fp = init_phi.clone(), init_F.clone()
        rs.solve(x, q, pp, Fp)
        torch.cuda.synchronize()
        rs.close()
        pn = init_phi.clone() if acc else torch.full_like(q, float("nan"))
        Fn = init_F.clone() if acc else torch.full_like(x, float("nan"))
        rc_ = se.real_solve_noalloc(per, box, xi, rc, x, q, pn, Fn, self_term=self_term, accumulate=acc)
        assert rc_ == se.SE_OK, se.last_error()
        torch.cuda.synchronize()
        assert torch.isfinite(pn).all() and torch.isfinite(Fn).all()
        assert torch.equal(pn, pp) and torch.equal(Fn, Fp), i


def test_noalloc_potential_only_and_empty():
    N = 5000
    x, q = system(N, 70, 3)
    rs = se.RealSpace(3, 1.0, 10.0, 0.12, N)
    pp = torch.zeros_like(q)
    rs.solve(x, q, pp, None)
    torch.cuda.synchronize()
    rs.close()
    pn = torch.full_like(q, float("nan"))
    assert se.real_solve_noalloc(3, 1.0, 10.0, 0.12, x, q, pn, None) == se.SE_OK
    torch.cuda.synchronize()
    assert torch.equal(pn, pp)
    e = torch.zeros(0, dtype=torch.float64, device=DEV)
    assert se.real_solve_noalloc(3, 1.0, 10.0, 0.12, e, e, e, e) == se.SE_OK


def test_noalloc_refuses_inputs_beyond_its_scratch():
    N = 2_000_000
    x, q = system(N, 71, 3)
    phi = torch.zeros_like(q)
    rc_ = se.real_solve_noalloc(3, 1.0, 30.0, 0.05, x, q, phi, None)
    assert rc_ == se.SE_ERR_CAPACITY
    assert "scratch" in se.last_error()
    # the refusal launched nothing: a following call works
    N2 = 1000
    x2, q2 = system(N2, 72, 3)
    p2 = torch.zeros_like(q2)
    assert se.real_solve_noalloc(3, 1.0, 8.0, 0.2, x2, q2, p2, None) == se.SE_OK
    torch.cuda.synchronize()
    assert torch.isfinite(p2).all()

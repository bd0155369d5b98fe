"""Repeatability of the FP32 solves, bit for bit: the particle order inside a sort key is
fixed by a stable sort of (key, index), the FP32 spreads add integer-valued FP64 values
into LDS (exact, so the order of the atomics does not matter), the halo merge sums in a
fixed order and the gathers have no atomics.  So two solves of the same input, a graph
replay and an eager solve, a grouped solve and a solve alone, and a served job alone and
in a mix of concurrent jobs give identical arrays.  FP64 plans accumulate FP64 products
with LDS atomics and repeat to about 1e-16."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se
from conftest import rms_rel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def system(N, seed, clustered=False, per=3):
    rng = np.random.default_rng(seed)
    if clustered:
        c = rng.random((3, 8))
        lab = rng.integers(0, 8, N)
        x = (c[:, lab] + 0.01 * rng.standard_normal((3, N))) % 1.0
        x = np.where(x >= 1.0, 0.0, x)
    else:
        x = rng.random((3, N))
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    return torch.tensor(x.reshape(-1), device=DEV), torch.tensor(q, device=DEV)


# N, M, P, xi, clustered, periodicity: the register-ring spread and both P = 16 gathers
# (clustered: the matrix-core gather on dense items, the vector gather on sparse ones),
# the P = 11 and P = 8 ring kernels, the generic FP32 kernels (P = 12), SE2P
CASES = [(300_000, 64, 16, 20.0, True, 3), (60_000, 48, 11, 15.0, False, 3), (40_000, 32, 8, 10.0, True, 3),
         (20_000, 32, 12, 10.0, False, 3), (30_000, 32, 16, 10.0, False, 2)]


def _plan(M, P, xi, N, per=3, prec=se.PREC_FP32, **kw):
    return se.Plan(M, P, 1.0, xi, N, precision=prec, periodicity=per, slab=1.0 if per == 2 else 0.0, **kw)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d_M%d_P%d_%s_%dp" % (c[0], c[1], c[2], "cl" if c[4] else "un",
                                                                           c[5]))
def test_fp32_solves_repeat_bit_for_bit(case):
    N, M, P, xi, clustered, per = case
    print("case", case, flush=True)
    x, q = system(N, 11 + P, clustered, per)
    xo, qo = system(N // 3, 12 + P, not clustered, per)
    plan = _plan(M, P, xi, N, per)
    out = []
    for rep in range(3):
        phi = torch.full_like(q, float("nan"))
        F = torch.full_like(x, float("nan"))
        plan.solve(x, q, phi, F)
        # other work on the same plan between the repeats
        plan.solve(xo, qo, torch.empty_like(qo), None)
        torch.cuda.synchronize()
        out.append((phi.clone(), F.clone()))
    plan.close()
    for phi, F in out[1:]:
        assert torch.equal(phi, out[0][0]) and torch.equal(F, out[0][1])
    assert torch.isfinite(out[0][0]).all() and torch.isfinite(out[0][1]).all()


def test_fp64_solves_repeat_to_rounding():
    N, M, P, xi = 20_000, 32, 16, 10.0
    x, q = system(N, 5)
    plan = _plan(M, P, xi, N, prec=se.PREC_FP64)
    res = []
    for _ in range(2):
        phi = torch.zeros_like(q)
        plan.solve(x, q, phi, None)
        torch.cuda.synchronize()
        res.append(phi.cpu().numpy())
    plan.close()
    assert rms_rel(res[1], res[0]) < 1e-15


def test_grouped_solve_equals_solves_alone_bit_for_bit():
    """Systems of one geometry solved as one group (system index folded into the tile
    index; the stable sort keeps each system's particle order) give the arrays of their
    solves alone."""
    M, P, xi = 32, 16, 10.0
    rng = np.random.default_rng(8)
    systems = []
    for i in range(6):
        N = int(rng.integers(500, 20000))
        x, q = system(N, 300 + i, clustered=bool(i % 2))
        systems.append((x, q, torch.zeros_like(q), torch.zeros_like(x)))
    ntot = sum(s[1].numel() for s in systems)
    grp = _plan(M, P, xi, ntot, max_systems=len(systems))
    grp.solve_group(systems)
    torch.cuda.synchronize()
    grp.close()
    for x, q, phi, F in systems:
        plan = _plan(M, P, xi, q.numel())
        pa, Fa = torch.zeros_like(q), torch.zeros_like(x)
        plan.solve(x, q, pa, Fa)
        torch.cuda.synchronize()
        plan.close()
        assert torch.equal(phi, pa) and torch.equal(F, Fa), q.numel()


def test_graph_replay_then_interp_sees_the_replayed_solve():
    """A replayed graph restores the plan state of its solve: after solve(N1) and two
    replays of solve_graph(N2 > N1), se_interp gathers all N2 particles (NaN-filled outputs
    are overwritten) and equals an eager solve of N2."""
    M, P, xi = 48, 16, 15.0
    N1, N2 = 5000, 40000
    x1, q1 = system(N1, 21)
    x2, q2 = system(N2, 22, clustered=True)
    plan = _plan(M, P, xi, N2, stream=torch.cuda.Stream())
    with torch.cuda.stream(plan.stream):
        plan.solve(x1, q1, torch.zeros_like(q1), None)
    torch.cuda.synchronize()
    pg, Fg = torch.zeros_like(q2), torch.zeros_like(x2)
    for _ in range(2):
        plan.solve_graph(x2, q2, pg, Fg)
        torch.cuda.synchronize()
    plan.solve(x1, q1, torch.zeros_like(q1), None)   # eager state of another size in between
    torch.cuda.synchronize()
    plan.solve_graph(x2, q2, pg, Fg)                  # replay restores N2
    torch.cuda.synchronize()
    pi, Fi = torch.full_like(q2, float("nan")), torch.full_like(x2, float("nan"))
    with torch.cuda.stream(plan.stream):
        plan.interp(pi, Fi)
    torch.cuda.synchronize()
    pe, Fe = torch.zeros_like(q2), torch.zeros_like(x2)
    with torch.cuda.stream(plan.stream):
        plan.solve(x2, q2, pe, Fe)
    torch.cuda.synchronize()
    plan.close()
    assert torch.equal(pi, pe) and torch.equal(Fi, Fe)
    assert torch.equal(pg, pe) and torch.equal(Fg, Fe)


def test_real_space_repeats_bit_for_bit():
    N = 20000
    x, q = system(N, 33, clustered=True)
    rs = se.RealSpace(3, 1.0, 12.0, 0.1, N, self_term=True)
    out = []
    for _ in range(2):
        phi, F = torch.zeros_like(q), torch.zeros_like(x)
        rs.solve(x, q, phi, F)
        torch.cuda.synchronize()
        out.append((phi.clone(), F.clone()))
    rs.close()
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])

"""GPU SE2P solver (se_plan with periodicity 2) against the SE2P Ewald k-space sum
of the reference's direct solver (SE_direct.c:256-338, restated in the oracle as
or_direct2p_fd + or_direct2p_k0 and pinned as the slab limit of the 3P sum): the
potential to the fast-Gaussian-gridding error floor, the force as (q/2) grad phi of
the direct potential, FP32 against FP64, and the slab bounds."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se
from conftest import rms_rel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
FP32_TOL = 1e-6     # rms-relative, FP32 vs FP64 at the same P: potential
FP32_TOL_F = 2e-6   # force (one more FP32 rounding in its derivative factors)


def slab_system(N, seed, L=1.0, slab=1.0):
    rng = np.random.default_rng(seed)
    x = rng.random((3, N))
    x[:2] *= L
    x[2] *= slab
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    return np.ascontiguousarray(x.reshape(-1)), q


def solve2p(x, q, M, P, xi, prec, L=1.0, slab=1.0, force=True, **kw):
    N = len(q)
    plan = se.Plan(M, P, L, xi, max(N, 1), precision=prec, periodicity=2, slab=slab, **kw)
    xt = torch.tensor(x, device=DEV)
    qt = torch.tensor(q, device=DEV)
    phi = torch.zeros(N, dtype=torch.float64, device=DEV)
    F = torch.zeros(3 * N, dtype=torch.float64, device=DEV) if force else None
    plan.solve(xt, qt, phi, F)
    torch.cuda.synchronize()
    info, shape = plan.info(), plan.shape
    plan.close()
    return phi.cpu().numpy(), (F.cpu().numpy() if force else None), info, shape


@pytest.mark.parametrize("case", [(32, 21, 6.0, 1e-11), (24, 16, 5.0, 1e-9), (40, 24, 8.0, 1e-11),
                                  (28, 11, 4.0, 1e-6)], ids=lambda c: "M%d_P%d" % c[:2])
def test_se2p_potential_matches_direct_sum(oracle, case):
    M, P, xi, tol = case
    print("case se2p", case, flush=True)
    x, q = slab_system(40, 5 + M)
    phi, _, info, shape = solve2p(x, q, M, P, xi, se.PREC_FP64, force=False)
    assert info["discarded"] == 0 and shape[:2] == (M, M) and shape[2] > M
    ref = oracle.direct2p(x, q, (1.0, 1.0, 1.0), xi, 14)
    e = rms_rel(phi, ref)
    print("se2p fp64 M=%d P=%d Mz=%d: rms-rel %.2e" % (M, P, shape[2], e))
    assert e < tol


def test_se2p_box_and_slab_sizes(oracle):
    """A non-unit periodic square and a slab thinner than the box."""
    L, slab, M, P, xi = 2.0, 0.7, 48, 21, 4.0
    x, q = slab_system(36, 17, L=L, slab=slab)
    phi, _, info, shape = solve2p(x, q, M, P, xi, se.PREC_FP64, L=L, slab=slab, force=False)
    ref = oracle.direct2p(x, q, (L, L, slab), xi, 20)
    assert info["discarded"] == 0
    assert rms_rel(phi, ref) < 1e-11


def test_se2p_force_is_half_q_grad_phi(oracle):
    """F_m = (q_m/2) grad phi(x_m) of the direct SE2P sum, by central differences with a
    zero probe charge (the same convention as SE3P, SE_direct.c:445-459)."""
    M, P, xi = 32, 21, 6.0
    x, q = slab_system(24, 9)
    _, F, _, _ = solve2p(x, q, M, P, xi, se.PREC_FP64)
    F = F.reshape(3, -1)
    X = x.reshape(3, -1)
    eps = 1e-5
    for m in (0, 5, 17):
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                y = X[:, m].copy()
                y[d] += sgn * eps
                xa = np.concatenate([y[:, None], X], axis=1).reshape(-1)
                qa = np.concatenate([[0.0], q])
                vals.append(oracle.direct2p(xa, qa, (1.0, 1.0, 1.0), xi, 14, n_targets=1)[0])
            grad = (vals[0] - vals[1]) / (2 * eps)
            assert abs(F[d, m] - 0.5 * q[m] * grad) < 1e-6 * max(1.0, np.max(np.abs(F))), (m, d)


def test_se2p_fp32_matches_fp64():
    M, P, xi = 64, 16, 20.0
    x, q = slab_system(20000, 3)
    p32, F32, info32, _ = solve2p(x, q, M, P, xi, se.PREC_FP32)
    p64, F64, _, _ = solve2p(x, q, M, P, xi, se.PREC_FP64)
    assert info32["pencil"] == 1
    e_phi, e_F = rms_rel(p32, p64), rms_rel(F32, F64)
    print("se2p fp32 vs fp64: phi %.2e F %.2e" % (e_phi, e_F))
    assert e_phi < FP32_TOL and e_F < FP32_TOL_F


def test_se2p_particles_outside_the_slab_are_left_out(oracle):
    M, P, xi = 24, 16, 5.0
    x, q = slab_system(30, 8)
    X = x.reshape(3, -1)
    xo = np.concatenate([X, np.array([[0.5, 0.5], [0.5, 0.5], [-0.1, 1.2]])], axis=1)
    qo = np.concatenate([q, [1.0, -1.0]])
    phi, _, info, _ = solve2p(np.ascontiguousarray(xo.reshape(-1)), qo, M, P, xi, se.PREC_FP64, force=False)
    assert info["discarded"] == 2
    ref = oracle.direct2p(x, q, (1.0, 1.0, 1.0), xi, 14)
    assert rms_rel(phi[:30], ref) < 1e-9

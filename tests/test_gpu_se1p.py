"""GPU SE1P solver (se_plan with periodicity 1: periodic in z, free in x and y) against
the SE1P Ewald k-space sum of the reference's direct solver (SE_direct.c:90-176,
restated in the oracle as or_direct1p_fd + or_direct1p_k0 and pinned by the
xi-independence of the full Ewald sum): the potential to the fast-Gaussian-gridding
error floor, also for a charged system (the k_z = 0 term's constant), the force as
(q/2) grad phi of the direct potential, FP32 against FP64, and the square's bounds."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se
from conftest import rms_rel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
FP32_TOL = 1e-6     # rms-relative, FP32 vs FP64 at the same P: potential
FP32_TOL_F = 2e-6   # force (one more FP32 rounding in its derivative factors)


def rod_system(N, seed, L=1.0, side=1.0, neutral=True):
    """x, y uniform in [0, side)^2 (free directions), z uniform in [0, L) (periodic)."""
    rng = np.random.default_rng(seed)
    x = rng.random((3, N))
    x[:2] *= side
    x[2] *= L
    q = rng.uniform(-1, 1, N)
    if neutral:
        q -= q.mean()
    return np.ascontiguousarray(x.reshape(-1)), q


def k_layers(xi, L=1.0):
    # K0inc(k^2/(4 xi^2), .) < e^{-40} beyond k = 2 xi sqrt(40)
    return int(np.ceil(2 * xi * np.sqrt(40.0) * L / (2 * np.pi))) + 1


def solve1p(x, q, M, P, xi, prec, L=1.0, side=1.0, force=True, **kw):
    N = len(q)
    plan = se.Plan(M, P, L, xi, max(N, 1), precision=prec, periodicity=1, slab=side, **kw)
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
def test_se1p_potential_matches_direct_sum(oracle, case):
    M, P, xi, tol = case
    print("case se1p", case, flush=True)
    x, q = rod_system(40, 5 + M)
    phi, _, info, shape = solve1p(x, q, M, P, xi, se.PREC_FP64, force=False)
    assert info["discarded"] == 0 and shape[2] == M and shape[0] == shape[1] > M
    ref = oracle.direct1p(x, q, (1.0, 1.0, 1.0), xi, k_layers(xi))
    e = rms_rel(phi, ref)
    print("se1p fp64 M=%d P=%d Mxy=%d: rms-rel %.2e" % (M, P, shape[0], e))
    assert e < tol


def test_se1p_charged_system(oracle):
    """A net charge: the k = 0 mode carries the reference's constant (gamma + ln xi^2)."""
    M, P, xi = 32, 21, 6.0
    x, q = rod_system(30, 11, neutral=False)
    assert abs(q.sum()) > 0.5
    phi, _, _, _ = solve1p(x, q, M, P, xi, se.PREC_FP64, force=False)
    ref = oracle.direct1p(x, q, (1.0, 1.0, 1.0), xi, k_layers(xi))
    assert rms_rel(phi, ref) < 1e-11


def test_se1p_period_and_square_sizes(oracle):
    """A period L = 2 and a square of side 0.7 in the free directions."""
    L, side, M, P, xi = 2.0, 0.7, 48, 21, 4.0
    x, q = rod_system(36, 17, L=L, side=side)
    phi, _, info, _ = solve1p(x, q, M, P, xi, se.PREC_FP64, L=L, side=side, force=False)
    ref = oracle.direct1p(x, q, (side, side, L), xi, k_layers(xi, L))
    assert info["discarded"] == 0
    assert rms_rel(phi, ref) < 1e-11


def test_se1p_force_is_half_q_grad_phi(oracle):
    """F_m = (q_m/2) grad phi(x_m) of the direct SE1P sum, by central differences with a
    zero probe charge (the reference's force convention, SE_direct.c:445-459)."""
    M, P, xi = 32, 21, 6.0
    x, q = rod_system(24, 9)
    _, F, _, _ = solve1p(x, q, M, P, xi, se.PREC_FP64)
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
                vals.append(oracle.direct1p(xa, qa, (1.0, 1.0, 1.0), xi, k_layers(xi), n_targets=1)[0])
            grad = (vals[0] - vals[1]) / (2 * eps)
            assert abs(F[d, m] - 0.5 * q[m] * grad) < 1e-6 * max(1.0, np.max(np.abs(F))), (m, d)


def test_se1p_fp32_matches_fp64():
    M, P, xi = 48, 16, 15.0
    x, q = rod_system(20000, 3)
    p32, F32, info32, _ = solve1p(x, q, M, P, xi, se.PREC_FP32)
    p64, F64, _, _ = solve1p(x, q, M, P, xi, se.PREC_FP64)
    assert info32["pencil"] == 1
    e_phi, e_F = rms_rel(p32, p64), rms_rel(F32, F64)
    print("se1p fp32 vs fp64: phi %.2e F %.2e" % (e_phi, e_F))
    assert e_phi < FP32_TOL and e_F < FP32_TOL_F


def test_se1p_particles_outside_the_square_are_left_out(oracle):
    M, P, xi = 24, 16, 5.0
    x, q = rod_system(30, 8)
    X = x.reshape(3, -1)
    xo = np.concatenate([X, np.array([[-0.1, 0.5, 1.2], [0.5, 1.0, 0.5], [0.5, 0.5, 0.5]])], axis=1)
    qo = np.concatenate([q, [1.0, -1.0, 1.0]])
    phi, _, info, _ = solve1p(np.ascontiguousarray(xo.reshape(-1)), qo, M, P, xi, se.PREC_FP64, force=False)
    assert info["discarded"] == 3
    ref = oracle.direct1p(x, q, (1.0, 1.0, 1.0), xi, k_layers(xi))
    assert rms_rel(phi[:30], ref) < 1e-9

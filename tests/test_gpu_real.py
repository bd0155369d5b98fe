"""Real-space and self terms on the GPU (se_real_*) against the oracle's restatement of
the reference's direct sums (SE_direct.c: SE*P_direct_real_rc, SE*P_direct_self), the
force as (q/2) grad phi by central differences, and the full Ewald sum -- GPU real space
+ GPU k-space solver + self -- independent of the splitting parameter xi for SE3P, SE2P
and SE1P."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se
from conftest import rms_rel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def neutral(N, seed, box=(1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    x = rng.random((3, N)) * np.asarray(box)[:, None]
    q = rng.uniform(-1, 1, N)
    return np.ascontiguousarray(x.reshape(-1)), q - q.mean()


def real_gpu(per, x, q, box, xi, rc, force=False, self_term=False):
    N = len(q)
    rs = se.RealSpace(per, box, xi, rc, N, self_term=self_term)
    phi = torch.zeros(N, dtype=torch.float64, device=DEV)
    F = torch.zeros(3 * N, dtype=torch.float64, device=DEV) if force else None
    rs.solve(torch.tensor(x, device=DEV), torch.tensor(q, device=DEV), phi, F)
    torch.cuda.synchronize()
    rs.close()
    return phi.cpu().numpy(), (F.cpu().numpy() if force else None)


@pytest.mark.parametrize("per", [3, 2, 1])
@pytest.mark.parametrize("rc", [0.3, 0.45, 1.3])
def test_real_space_matches_direct_sum(oracle, per, rc):
    """Cutoffs below a third of the box (27 neighbour cells), just below half of it, and
    above the box (image layers beyond the nearest)."""
    box, xi = (1.0, 0.8, 1.2), 3.0
    x, q = neutral(300, 10 * per + int(10 * rc), box)
    phi, _ = real_gpu(per, x, q, box, xi, rc)
    layers = int(np.ceil(rc / min(box))) + 1
    ref = oracle.direct_real(per, x, q, box, xi, layers, rc=rc)
    print("real per=%d rc=%.2f: rms-rel %.2e" % (per, rc, rms_rel(phi, ref)), flush=True)
    assert rms_rel(phi, ref) < 1e-13
    phs, _ = real_gpu(per, x, q, box, xi, rc, self_term=True)
    assert rms_rel(phs, ref + oracle.direct_self(q, xi)) < 1e-13


def test_real_space_positions_outside_and_images(oracle):
    """Periodic coordinates are folded into the box; in SE2P's free z, particles outside
    [0, box) are left out and their outputs are not written."""
    box, xi, rc = (1.0, 1.0, 1.0), 4.0, 0.4
    x, q = neutral(200, 3, box)
    shifted = x.reshape(3, -1).copy()
    shifted[:2] += np.random.default_rng(1).integers(-2, 3, shifted[:2].shape)
    phi0, _ = real_gpu(2, x, q, box, xi, rc)
    phi1, _ = real_gpu(2, np.ascontiguousarray(shifted.reshape(-1)), q, box, xi, rc)
    assert rms_rel(phi1, phi0) < 1e-13
    X = x.reshape(3, -1)
    xo = np.ascontiguousarray(np.concatenate([X, [[0.5], [0.5], [1.5]]], axis=1).reshape(-1))
    N = len(q) + 1
    rs = se.RealSpace(2, box, xi, rc, N)
    phi = torch.full((N,), 7.0, dtype=torch.float64, device=DEV)
    rs.solve(torch.tensor(xo, device=DEV), torch.tensor(np.append(q, 1.0), device=DEV), phi, None)
    torch.cuda.synchronize()
    rs.close()
    phi = phi.cpu().numpy()
    assert phi[-1] == 7.0
    assert rms_rel(phi[:-1], phi0) < 1e-13


@pytest.mark.parametrize("per", [3, 1])
def test_real_space_force_is_half_q_grad_phi(oracle, per):
    box, xi, rc = (1.0, 1.0, 1.0), 4.0, 0.45
    x, q = neutral(120, 21 + per, box)
    _, F = real_gpu(per, x, q, box, xi, rc, force=True)
    F = F.reshape(3, -1)
    X = x.reshape(3, -1)
    eps = 1e-6
    for m in (0, 7, 50):
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                y = X[:, m].copy()
                y[d] += sgn * eps
                xa = np.concatenate([y[:, None], X], axis=1).reshape(-1)
                qa = np.concatenate([[0.0], q])
                # the probe at x_m sees every particle but m itself
                qa[1 + m] = 0.0
                vals.append(oracle.direct_real(per, xa, qa, box, xi, 2, rc=rc, n_targets=1)[0])
            grad = (vals[0] - vals[1]) / (2 * eps)
            assert abs(F[d, m] - 0.5 * q[m] * grad) < 1e-6 * max(1.0, np.max(np.abs(F))), (m, d)


@pytest.mark.parametrize("per", [3, 2, 1])
def test_full_ewald_sum_is_independent_of_xi(per):
    """real (GPU) + k-space (GPU plan, FP64, P = 24) + self is the Coulomb sum of the
    periodic system, so the totals at xi = 8 and xi = 10 agree to the solver's error;
    each part alone moves by O(1).  Pins the real-space kernel, the k-space solvers'
    Green's functions (SE3P, SE2P with its k = 0 term, SE1P with its k = 0 term) and the
    self term together, all on the GPU."""
    N, P = 60, 24
    x, q = neutral(N, 70 + per)
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    tot = {}
    parts = {}
    for xi, M in ((8.0, 48), (10.0, 60)):
        rc = 6.0 / xi
        phi_r, _ = real_gpu(per, x, q, (1.0, 1.0, 1.0), xi, rc, self_term=True)
        plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP64, periodicity=per,
                       slab=0.0 if per == 3 else 1.0)
        phi_k = torch.zeros(N, dtype=torch.float64, device=DEV)
        plan.solve(xt, qt, phi_k, None)
        torch.cuda.synchronize()
        plan.close()
        tot[xi] = phi_r + phi_k.cpu().numpy()
        parts[xi] = phi_r
    e = rms_rel(tot[8.0], tot[10.0])
    print("ewald total per=%d: xi 8 vs 10 rms-rel %.2e" % (per, e), flush=True)
    assert e < 1e-10
    assert rms_rel(parts[8.0], parts[10.0]) > 1e-2

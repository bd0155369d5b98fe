"""Parameter selection for an accuracy target (se_select_params, host only; the estimates
of SE_init_params, SE_general.c:517-572).  The chosen support is checked by what it
delivers: the oracle's SE3P solve at the chosen (M, P, m) against the oracle's direct
k-space sum (SE3P_direct_fd / _force_fd restated, |k| <= (M-1)/2, the layers
test_accuracy.c:55 uses) must meet the tolerance, and the grid size must be the one the
oracle's restatement of SE_init_params gives."""
import numpy as np
import pytest

import se_fast_gridding_amd as se


def _system(N=100, seed=7):
    rng = np.random.default_rng(seed)
    x = rng.random((3, N)).reshape(-1)
    q = rng.choice([-1.0, 1.0], N)
    q -= q.mean()
    return x, q


@pytest.mark.parametrize("tol", [1e-3, 1e-5, 1e-7, 1e-9])
@pytest.mark.parametrize("force", [False, True], ids=["potential", "force"])
def test_selected_support_meets_the_tolerance(oracle, tol, force):
    x, q = _system()
    L, xi = 1.0, 6.0
    P, M, m, est = se.select_params(tol, L, xi, q, potential=not force, force=force)
    assert 1 <= P <= 32 and est <= tol
    assert np.isclose(m, 0.95 * np.sqrt(np.pi * P))
    assert M == oracle.init_params(P, L, xi, q)[0]      # the reference's grid rule
    ref = oracle.direct3p(x, q, (L, L, L), xi, (M - 1) // 2, force=force)
    got = oracle.se3p(x, q, L, M, P, m, xi, force=force)
    assert np.sqrt(np.mean((got - ref) ** 2)) <= tol     # rms absolute, as the estimate
    # a looser target never asks for a larger support, a tighter one never for a smaller
    assert se.select_params(tol * 10, L, xi, q, potential=not force, force=force)[0] <= P
    assert se.select_params(tol / 10, L, xi, q, potential=not force, force=force)[0] >= P


def test_select_params_rejects_unreachable_targets():
    _, q = _system()
    with pytest.raises(ValueError):
        se.select_params(1e-30, 1.0, 6.0, q)
    with pytest.raises(ValueError):
        se.select_params(-1.0, 1.0, 6.0, q)
    with pytest.raises(ValueError):
        se.select_params(1e-6, 1.0, 6.0, q, potential=False, force=False)

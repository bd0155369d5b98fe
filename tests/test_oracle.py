"""Pins the CPU oracle (oracle/se_oracle.c) before anything is compared with it.

Checks, per function: the reference's own outputs (golden vectors from the compiled
SE_fgg.c), closed forms, the DFT definition, two independent forms of the direct
Ewald sums, finite differences, and SE-vs-direct agreement at the FGG error floor.
"""
import numpy as np
import pytest

from conftest import golden_files, load_golden, rms_rel

G3 = golden_files("g3p")
G2 = golden_files("g2p")
G1 = golden_files("g1p")


def _params(O, g):
    dims = tuple(int(v) for v in g["dims"])
    return O.pack_params(int(g["N"]), dims, int(g["P"]), float(g["c"]), float(g["h"]),
                         float(g["a"]), float(g["b"]))


def _close(a, b, tol=1e-13):
    scale = max(np.max(np.abs(b)), 1e-300)
    assert np.max(np.abs(np.asarray(a) - np.asarray(b))) <= tol * scale


@pytest.mark.parametrize("path", G3 + G2 + G1, ids=lambda p: p.split("/")[-1][:-4])
def test_oracle_matches_reference_golden(oracle, path):
    """Expansion, spread, wrap, extend, interpolation vs the compiled reference."""
    O = oracle
    g = load_golden(path)
    per = int(g["per"])
    p = _params(O, g)
    zs = O.base_gaussian(p)
    _close(zs, g["zs"], 1e-14)
    ex = O.expand_all(per, g["x"], p)
    assert np.array_equal(ex["idx"], g["idx"])
    for k in ("zx", "zy", "zz"):
        _close(ex[k], g[k], 1e-13)
    H = O.grid_split(p, g["q"], ex, zs)
    if "H_split" in g:
        _close(H, g["H_split"], 1e-13)
        _close(O.grid(per, p, g["x"], g["q"], zs), g["H_grid"], 1e-13)
    Hper = np.zeros(len(g["Hper"]))
    O.wrap(per, p, H, Hper)
    _close(Hper, g["Hper"], 1e-13)
    Hext = O.extend(per, p, g["G"])
    if "H_ext" in g:
        _close(Hext, g["H_ext"], 0.0)
    phi = O.int_split(p, Hext, ex, zs)
    _close(phi, g["phi_split"], 1e-12)
    _close(phi, g["phi_int"], 1e-12)
    if "zfx" in g:
        exf = O.expand_all(per, g["x"], p, force=True)
        assert np.array_equal(exf["idx"], g["fidx"])
        for k, gk in (("zx", "fzx"), ("zy", "fzy"), ("zz", "fzz"), ("zfx", "zfx"), ("zfy", "zfy"), ("zfz", "zfz")):
            _close(exf[k], g[gk], 1e-13)
        if "F_split" in g:
            F = O.int_split_force(p, Hext, g["q"], exf, zs)
            _close(F, g["F_split"], 1e-12)
            _close(F, g["F_int"], 1e-12)


@pytest.mark.parametrize("P", [3, 4, 5, 8, 11, 16, 21])
def test_expansion_closed_form(oracle, P):
    """zx_i zy_j zz_k zs_ijk == q d exp(-c |x - g_ijk|^2), g on the periodic grid image."""
    O = oracle
    rng = np.random.default_rng(P)
    M, N = 12, 6
    h = 1.0 / M
    c = 40.0 + P
    p = O.pack_params(N, M, P, c, h)
    x = rng.random(3 * N)
    ex = O.expand_all(O.P3, x, p)
    zs = O.base_gaussian(p).reshape(P, P, P)
    ph = O.half(P)
    first = -ph if P & 1 else -ph + 1
    for n in range(N):
        xn = x[n::N][:3] if False else np.array([x[n], x[n + N], x[n + 2 * N]])
        iota = np.round(xn / h) if P & 1 else np.floor(xn / h)
        g = h * (iota[:, None] + first + np.arange(P)[None, :])       # 3 x P grid coords
        W = (ex["zx"][n * P:(n + 1) * P][:, None, None] * ex["zy"][n * P:(n + 1) * P][None, :, None]
             * ex["zz"][n * P:(n + 1) * P][None, None, :] * zs)
        r2 = ((xn[0] - g[0])[:, None, None] ** 2 + (xn[1] - g[1])[None, :, None] ** 2
              + (xn[2] - g[2])[None, None, :] ** 2)
        want = p.d * np.exp(-c * r2)
        assert np.max(np.abs(W - want)) <= 1e-12 * want.max()
        # first support point sits at extended index (iota + first + P_half)
        e = (iota + first + ph).astype(int)
        assert ex["idx"][n] == (e[0] * p.npdims[1] + e[1]) * p.npdims[2] + e[2]


def test_spread_brute_force_periodic(oracle):
    """wrap(spread(x)) equals the brute-force sum of truncated periodic Gaussians."""
    O = oracle
    rng = np.random.default_rng(5)
    M, N, P = 10, 5, 5
    h = 1.0 / M
    c = 60.0
    p = O.pack_params(N, M, P, c, h)
    x = rng.random(3 * N)
    q = rng.uniform(-1, 1, N)
    zs = O.base_gaussian(p)
    Hper = O.wrap(O.P3, p, O.grid_split(p, q, O.expand_all(O.P3, x, p), zs)).reshape(M, M, M, order="F")
    want = np.zeros((M, M, M))
    ph = O.half(P)
    for n in range(N):
        xn = np.array([x[n], x[n + N], x[n + 2 * N]])
        iota = np.round(xn / h).astype(int)
        for i in range(-ph, ph + 1):
            for j in range(-ph, ph + 1):
                for k in range(-ph, ph + 1):
                    gi = iota + np.array([i, j, k])
                    r2 = np.sum((xn - h * gi) ** 2)
                    want[gi[0] % M, gi[1] % M, gi[2] % M] += q[n] * p.d * np.exp(-c * r2)
    assert np.max(np.abs(Hper - want)) <= 1e-12 * np.abs(want).max()


def test_fft_matches_dft_definition_and_numpy(oracle):
    O = oracle
    rng = np.random.default_rng(1)
    a = rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))
    n = np.indices(a.shape).reshape(3, -1).T
    dft = np.zeros(a.size, dtype=complex)
    for idx, kk in enumerate(n):
        ph = np.exp(-2j * np.pi * (n[:, 0] * kk[0] / 3 + n[:, 1] * kk[1] / 4 + n[:, 2] * kk[2] / 5))
        dft[idx] = np.sum(a.reshape(-1) * ph)
    assert np.allclose(O.fft3d(a, -1).reshape(-1), dft, rtol=0, atol=1e-12)
    b = rng.standard_normal((12, 8, 18)) + 0j
    assert np.allclose(O.fft3d(b, -1), np.fft.fftn(b), rtol=0, atol=1e-11)
    assert np.allclose(O.fft3d(b, +1), np.fft.ifftn(b) * b.size, rtol=0, atol=1e-11)


def test_kvec_and_scaling_closed_form(oracle):
    O = oracle
    L = 2.5
    k = O.kvec(8, L)
    assert np.allclose(k, 2 * np.pi / L * np.array([0, 1, 2, 3, -4, -3, -2, -1]))
    k5 = O.kvec(5, L)
    assert np.allclose(k5, 2 * np.pi / L * np.array([0, 1, 2, -2, -1]))
    M, s = 6, -0.01
    Z = O.scaling(s, M, L).reshape(M, M, M)
    kk = O.kvec(M, L)
    i, j, l = 1, 4, 5
    K2 = kk[i] ** 2 + kk[j] ** 2 + kk[l] ** 2
    assert np.isclose(Z[i, j, l], np.exp(s * K2) / K2 * 4 * np.pi / M ** 3, rtol=1e-14)


def test_direct3p_two_forms_agree(oracle):
    """Reference's pairwise hs-weighted cosine sum == structure-factor sum."""
    O = oracle
    rng = np.random.default_rng(2)
    N = 12
    x = rng.random(3 * N) * 1.3
    q = rng.choice([-1.0, 1.0], N)
    box = [1.3, 1.3, 1.3]
    a = O.direct3p(x, q, box, 4.0, 5, pairwise=True)
    b = O.direct3p(x, q, box, 4.0, 5)
    assert rms_rel(b, a) < 1e-13
    fa = O.direct3p(x, q, box, 4.0, 4, force=True, pairwise=True)
    fb = O.direct3p(x, q, box, 4.0, 4, force=True)
    assert rms_rel(fb, fa) < 1e-13


def test_direct3p_force_is_half_q_grad_phi(oracle):
    """F_m = (q_m/2) grad phi(x_m)  (SE_direct.c:445-459 convention), by finite differences."""
    O = oracle
    rng = np.random.default_rng(3)
    N, L, xi, K = 10, 1.0, 3.0, 6
    x = rng.random((3, N))
    q = rng.uniform(-1, 1, N)
    F = O.direct3p(x.reshape(-1), q, L, xi, K, force=True).reshape(3, N)
    eps = 1e-5
    for m in (0, 7):
        for d in range(3):
            vals = []
            for s in (+1, -1):
                y = x[:, m].copy()
                y[d] += s * eps
                xa = np.concatenate([y[:, None], x], axis=1).reshape(-1)   # probe as a zero charge
                qa = np.concatenate([[0.0], q])
                vals.append(O.direct3p(xa, qa, L, xi, K, n_targets=1)[0])
            grad = (vals[0] - vals[1]) / (2 * eps)
            assert abs(F[d, m] - 0.5 * q[m] * grad) < 1e-6 * max(1.0, abs(F[d, m]))


def test_direct2p_is_the_slab_limit_of_direct3p(oracle):
    """SE2P k-space sum (k != 0 part + k = 0 part, SE_direct.c:256-338) against the
    3-periodic sum in an L x L x Lz box.  For a neutral system with z in [0, L),
    phi_2P(x_m) = phi_3P(x_m; Lz) - (2 pi / V) sum_n q_n (z_m - z_n)^2, up to
    exp(-xi^2 (Lz - L)^2) and exp(-2 pi (Lz - L) / L): the sum over k_z becomes the
    integral that defines the 2P kernels, and the missing k_z = 0 point of the 3P sum is
    the quadratic term.  A dropped term, a wrong sign or a wrong 1/|k| breaks it."""
    O = oracle
    rng = np.random.default_rng(21)
    N, xi, Lz = 16, 3.0, 8.0
    x = rng.random((3, N))
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    phi2 = O.direct2p(x.reshape(-1), q, (1.0, 1.0, 1.0), xi, 8)
    phi3 = O.direct3p(x.reshape(-1), q, (1.0, 1.0, Lz), xi, 50)
    z = x[2]
    quad = 2 * np.pi / Lz * np.array([np.sum(q * (z[m] - z) ** 2) for m in range(N)])
    assert rms_rel(phi2, phi3 - quad) < 1e-13
    # both parts matter at this size: dropping the k = 0 part fails by orders of magnitude
    fd_only = O.direct2p(x.reshape(-1), q, (1.0, 1.0, 1.0), xi, 8, with_k0=False)
    assert rms_rel(fd_only, phi3 - quad) > 1e-3


def test_se3p_oracle_vs_direct_config1(oracle):
    """N=100 unit cube, q=+-1, P=21, M=32, xi=6: SE vs direct k-space sum (|j| <= 15)."""
    O = oracle
    rng = np.random.default_rng(1)
    N, P, M, xi = 100, 21, 32, 6.0
    x = rng.random(3 * N)
    q = np.array([1.0, -1.0] * (N // 2))
    rng.shuffle(q)
    m = 0.95 * np.sqrt(np.pi * P)
    phi = O.se3p(x, q, 1.0, M, P, m, xi)
    ref = O.direct3p(x, q, 1.0, xi, 15)
    assert rms_rel(phi, ref) < 1e-11
    F = O.se3p(x, q, 1.0, M, P, m, xi, force=True)
    Fref = O.direct3p(x, q, 1.0, xi, 15, force=True)
    assert rms_rel(F, Fref) < 1e-10


def test_init_params_rule(oracle):
    """SE_general.c:517-572 with the reference test's values (L=3.1814, xi=8)."""
    O = oracle
    g = load_golden(golden_files("atoms")[0]) if golden_files("atoms") else None
    import os
    from conftest import GOLDEN
    a = np.load(os.path.join(GOLDEN, "atoms.npz"))["xyzq"]
    M, m, Q = O.init_params(21, 3.1814, 8.0, a[:100, 3])
    assert M == 108 and np.isclose(m, 0.95 * np.sqrt(21 * np.pi))
    assert np.isclose(Q, np.sum(a[:100, 3] ** 2))


# ---- SE1P direct sums, real-space and self terms ---------------------------------

def test_k0inc_identity_and_limits(oracle):
    """K0inc(a, b) + K0inc(b, a) = 2 K0(2 sqrt(ab)) (the two halves of the complete
    integral, split at t = 1 and mapped by t -> 1/t), K0inc(a, 0) = E1(a); the modified
    Bessel K0 and E1 come from scipy, independent of the oracle."""
    sp = pytest.importorskip("scipy.special")
    for a, b in [(0.1, 0.2), (1.5, 0.01), (0.02, 30.0), (3.0, 3.0), (0.154, 128.0), (1e-3, 1e-3), (40.0, 0.5)]:
        ref = 2.0 * sp.k0(2.0 * np.sqrt(a * b))
        assert abs(oracle.k0inc(a, b) + oracle.k0inc(b, a) - ref) <= 1e-14 * ref, (a, b)
    for a in (0.01, 0.5, 3.0, 30.0):
        assert abs(oracle.k0inc(a, 0.0) / sp.exp1(a) - 1.0) < 1e-14


def test_ein_closed_forms(oracle):
    """Ein(x) = gamma + ln x + E1(x) (scipy E1) where that form does not cancel, and
    the leading terms x - x^2/4 + x^3/18 - x^4/96 of its series for tiny x."""
    sp = pytest.importorskip("scipy.special")
    for x in (0.5, 1.9, 2.1, 5.0, 30.0, 300.0):
        ref = np.euler_gamma + np.log(x) + sp.exp1(x)
        assert abs(oracle.ein(x) / ref - 1.0) < 1e-14, x
    for x in (1e-8, 1e-4):
        assert abs(oracle.ein(x) - (x - x * x / 4 + x ** 3 / 18 - x ** 4 / 96)) < 1e-16 * x


def _neutral(N, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((3, N)).reshape(-1)
    q = rng.uniform(-1, 1, N)
    return x, q - q.mean()


@pytest.mark.parametrize("per", [3, 2, 1])
def test_ewald_total_is_independent_of_xi(oracle, per):
    """real + k-space (+ k = 0 term) + self is the Coulomb sum of the periodic system,
    whatever the splitting parameter xi: comparing xi = 3 with xi = 5 pins the real-space
    sum, the k-space sums (SE3P; SE2P and SE1P with their k = 0 terms) and the self term
    together -- a dropped term, a wrong sign or prefactor moves the xi = 3 and xi = 5
    totals apart by O(1)."""
    O = oracle
    x, q = _neutral(20, 3)
    box = (1.0, 1.0, 1.0)

    def total(xi):
        kl = int(np.ceil(6 * xi / np.pi)) + 2
        real = O.direct_real(per, x, q, box, xi, 4 if per < 3 else 3)
        if per == 3:
            k = O.direct3p(x, q, box, xi, kl)
        elif per == 2:
            k = O.direct2p(x, q, box, xi, kl)
        else:
            k = O.direct1p(x, q, box, xi, kl)
        return real + k + O.direct_self(q, xi)

    a, b = total(3.0), total(5.0)
    assert np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(a))
    # each part alone does depend on xi
    assert np.max(np.abs(O.direct_real(per, x, q, box, 3.0, 3) - O.direct_real(per, x, q, box, 5.0, 3))) > 1e-2


def test_direct_real_cutoff_and_self(oracle):
    """The rc variant drops exactly the pairs beyond rc (brute force over pairs), and the
    self term is -2 xi/sqrt(pi) q."""
    O = oracle
    x, q = _neutral(12, 4)
    X = x.reshape(3, -1)
    xi, rc = 4.0, 0.45
    got = O.direct_real(3, x, q, (1.0, 1.0, 1.0), xi, 1, rc=rc)
    from math import erfc, sqrt
    for m in range(12):
        acc = 0.0
        for n in range(12):
            for s in np.ndindex(3, 3, 3):
                sh = np.array(s) - 1
                if m == n and not sh.any():
                    continue
                r = np.linalg.norm(X[:, m] - X[:, n] + sh)
                if r <= rc:
                    acc += q[n] * erfc(xi * r) / r
        assert abs(got[m] - acc) < 1e-14 * max(1.0, abs(acc))
    assert np.allclose(O.direct_self(q, xi), -2 * xi / sqrt(np.pi) * q, rtol=1e-15, atol=0)


def test_direct1p_k0_is_permutation_equivariant(oracle):
    """The SE1P k = 0 term is a sum over pairs of a kernel of the pair distance: relabelling
    the particles permutes its outputs."""
    O = oracle
    x, q = _neutral(10, 6)
    box = (1.0, 1.0, 1.0)
    k0 = O.direct1p(x, q, box, 4.0, 1, with_k0=True) - O.direct1p(x, q, box, 4.0, 1, with_k0=False)
    X = x.reshape(3, -1)
    perm = np.roll(np.arange(10), 3)
    k0p = O.direct1p(np.ascontiguousarray(X[:, perm].reshape(-1)), q[perm], box, 4.0, 1, with_k0=True) - \
        O.direct1p(np.ascontiguousarray(X[:, perm].reshape(-1)), q[perm], box, 4.0, 1, with_k0=False)
    assert np.allclose(k0p, k0[perm], rtol=0, atol=1e-14)

"""GPU parity of the device-resident SE3P engine (include/se_gpu.h) against the
CPU oracle: FP64 to <= 1e-12 rms-relative, FP32 to the stated FP32 tolerance, plus
size-independent properties (periodic images, adjointness, tiling invariance) and
edge cases (empty, single particle, boundary positions, heavy tiles)."""
import numpy as np
import pytest
import torch

import se_fast_gridding_amd as se
from conftest import rms_rel

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
# rms-relative, FP32 spread/FFT/gather vs the FP64 oracle at the same P: potential 1e-6
# (measured 1e-7..3e-7); the force carries one more FP32 rounding (its derivative factors)
# and measured up to 1.04e-6 at P = 8 on clustered input, so it has 2e-6
FP32_TOL = 1e-6
FP32_TOL_F = 2e-6


def system(N, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        c = rng.random((3, 4))
        lab = rng.integers(0, 4, N)
        x = (c[:, lab] + 0.01 * rng.standard_normal((3, N))) % 1.0
    else:
        x = rng.random((3, N))
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    return np.ascontiguousarray(x.reshape(-1)), q


def solve(x, q, M, P, xi, prec, force=True, L=1.0, **kw):
    N = len(q)
    plan = se.Plan(M, P, L, xi, max(N, 1), precision=prec, **kw)
    xt = torch.tensor(x, device=DEV)
    qt = torch.tensor(q, device=DEV)
    phi = torch.zeros(N, dtype=torch.float64, device=DEV)
    F = torch.zeros(3 * N, dtype=torch.float64, device=DEV) if force else None
    plan.solve(xt, qt, phi, F)
    torch.cuda.synchronize()
    info = plan.info()
    plan.close()
    return phi.cpu().numpy(), (F.cpu().numpy() if force else None), info


CASES = [  # N, M, P, xi, L
    (100, 32, 21, 6.0, 1.0),
    (300, 16, 8, 4.0, 1.0),
    (200, 20, 11, 5.0, 1.0),
    (150, 12, 16, 3.0, 1.0),
    (250, 24, 5, 6.0, 2.5),
    (120, 18, 12, 4.5, 1.0),
    (64, 10, 3, 2.0, 1.0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d_M%d_P%d" % c[:3])
def test_fp64_solve_matches_oracle(oracle, case):
    N, M, P, xi, L = case
    print("case fp64", case, flush=True)
    x, q = system(N, N + M + P)
    x = x * L
    m = 0.95 * np.sqrt(np.pi * P)
    phi, F, _ = solve(x, q, M, P, xi, se.PREC_FP64, L=L)
    assert rms_rel(phi, oracle.se3p(x, q, L, M, P, m, xi)) < 1e-12
    assert rms_rel(F, oracle.se3p(x, q, L, M, P, m, xi, force=True)) < 1e-12


@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: "N%d_M%d_P%d" % c[:3])
def test_fp32_solve_matches_oracle(oracle, case):
    N, M, P, xi, L = case
    print("case fp32", case, flush=True)
    x, q = system(N, 7 * N + P)
    m = 0.95 * np.sqrt(np.pi * P)
    phi, F, _ = solve(x, q, M, P, xi, se.PREC_FP32)
    e_phi = rms_rel(phi, oracle.se3p(x, q, 1.0, M, P, m, xi))
    e_F = rms_rel(F, oracle.se3p(x, q, 1.0, M, P, m, xi, force=True))
    print("fp32 rms-rel phi %.2e F %.2e" % (e_phi, e_F))
    assert e_phi < FP32_TOL and e_F < FP32_TOL_F


def test_spread_grid_matches_oracle_wrap(oracle):
    """Spread + fused wrap equals the reference's column-major H_per (transposed)."""
    N, M, P = 180, 14, 9
    x, q = system(N, 3)
    c = 50.0
    for prec, tol in ((se.PREC_FP64, 1e-13), (se.PREC_FP32, 2e-6)):
        plan = se.Plan(M, P, 1.0, 3.0, N, precision=prec)
        plan.spread(torch.tensor(x, device=DEV), torch.tensor(q, device=DEV))
        torch.cuda.synchronize()
        G = plan.grid().double().cpu().numpy()
        p = oracle.pack_params(N, M, P, plan_c(M, P, 3.0), 1.0 / M)
        zs = oracle.base_gaussian(p)
        Hper = oracle.wrap(3, p, oracle.grid_split(p, q, oracle.expand_all(3, x, p), zs))
        ref = Hper.reshape(M, M, M, order="F")
        plan.close()
        assert np.max(np.abs(G - ref)) <= tol * np.max(np.abs(ref))


@pytest.mark.parametrize("clustered", [False, True])
def test_pencil_spread_grid_matches_oracle_wrap(oracle, clustered):
    """The P = 16 FP32 register-ring spread gives the reference's wrapped grid, also with
    work items and waves that cut pencils mid-way (heavy clustered tiles split into
    chunks, small work-groups, small tiles)."""
    N, M, P, xi = 5000, 24, 16, 5.0
    x, q = system(N, 40 + int(clustered), clustered=clustered)
    p = oracle.pack_params(N, M, P, plan_c(M, P, xi), 1.0 / M)
    zs = oracle.base_gaussian(p)
    ref = oracle.wrap(3, p, oracle.grid_split(p, q, oracle.expand_all(3, x, p), zs)).reshape(M, M, M, order="F")
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    for kw in (dict(), dict(chunk=57, block=128), dict(tile=4, chunk=300, block=64)):
        plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32, **kw)
        plan.spread(xt, qt)
        torch.cuda.synchronize()
        assert plan.info()["pencil"] == 1
        G = plan.grid().double().cpu().numpy()
        plan.close()
        e = np.max(np.abs(G - ref)) / np.max(np.abs(ref))
        print("pencil spread %s: max rel %.2e" % (kw, e), flush=True)
        assert e < 2e-6, kw


def plan_c(M, P, xi, L=1.0):
    m = 0.95 * np.sqrt(np.pi * P)
    h = L / M
    w = P * h / 2
    eta = (2 * w * xi / m) ** 2
    return 2 * xi * xi / eta


def test_adjointness_spread_interp():
    """<spread(q), G> == h^-3 sum_n q_n interp(G)(x_n) for a random grid G (FP64)."""
    N, M, P = 300, 16, 10
    x, q = system(N, 5)
    plan = se.Plan(M, P, 1.0, 4.0, N, precision=se.PREC_FP64)
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    plan.spread(xt, qt)
    A = plan.grid()
    G = torch.tensor(np.random.default_rng(9).standard_normal((M, M, M)), device=DEV)
    plan.set_grid(G)
    phi = torch.zeros(N, dtype=torch.float64, device=DEV)
    plan.interp(phi, None)
    torch.cuda.synchronize()
    lhs = float((A * G).sum())
    rhs = float((qt * phi).sum()) / (1.0 / M) ** 3
    plan.close()
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_periodic_images_and_boundaries(oracle):
    """Positions shifted by whole boxes give the same result; x=0 and x just below L work."""
    N, M, P, xi = 90, 16, 8, 4.0
    x, q = system(N, 11)
    x[:5] = 0.0
    x[5:10] = np.nextafter(1.0, 0.0)
    phi0, F0, _ = solve(x, q, M, P, xi, se.PREC_FP64)
    m = 0.95 * np.sqrt(np.pi * P)
    assert rms_rel(phi0, oracle.se3p(x, q, 1.0, M, P, m, xi)) < 1e-12
    # whole-box shifts of interior particles (x = 1 - eps + k rounds to an integer, which
    # legitimately moves the support window by one cell: that changes the result by the
    # truncation error, as it would in the reference)
    shift = np.random.default_rng(2).integers(-3, 4, x.shape).astype(float)
    shift.reshape(3, N)[:, :10] = 0.0
    phi1, F1, _ = solve(x + shift, q, M, P, xi, se.PREC_FP64)
    assert rms_rel(phi1, phi0) < 1e-11 and rms_rel(F1, F0) < 1e-11


def test_empty_and_single_particle(oracle):
    phi, F, info = solve(np.zeros(0), np.zeros(0), 16, 8, 4.0, se.PREC_FP64)
    assert phi.size == 0 and info["items"] == 0
    x = np.array([0.3, 0.6, 0.9])
    q = np.array([1.0])
    m = 0.95 * np.sqrt(np.pi * 8)
    phi, F, _ = solve(x, q, 16, 8, 4.0, se.PREC_FP64)
    assert rms_rel(phi, oracle.se3p(x, q, 1.0, 16, 8, m, 4.0)) < 1e-12
    # the exact k-space self-force vanishes; SE's is the discretisation error, the same as the oracle's
    assert rms_rel(F, oracle.se3p(x, q, 1.0, 16, 8, m, 4.0, force=True)) < 1e-10
    phi16, F16, _ = solve(x, q, 16, 16, 4.0, se.PREC_FP64)
    assert np.max(np.abs(F16)) < 1e-3 * np.max(np.abs(F))


@pytest.mark.parametrize("prec", [se.PREC_FP64, se.PREC_FP32])
def test_tiling_and_chunking_invariance(prec):
    """Heavy tiles split into many work items give the same answer as one item per tile."""
    N, M, P, xi = 2000, 24, 8, 5.0
    x, q = system(N, 13, clustered=True)
    ref, Fref, info0 = solve(x, q, M, P, xi, prec)
    tol = 1e-12 if prec == se.PREC_FP64 else 2e-6
    for kw in (dict(chunk=16), dict(tile=4, chunk=64, block=128), dict(tile=3, block=64), dict(tile=12),
               dict(generic=True), dict(generic=True, tile=5, chunk=100)):
        phi, F, info = solve(x, q, M, P, xi, prec, **kw)
        assert rms_rel(phi, ref) < tol and rms_rel(F, Fref) < tol, kw
    assert solve(x, q, M, P, xi, prec, chunk=16)[2]["items"] > info0["items"]


@pytest.mark.parametrize("prec", [se.PREC_FP64, se.PREC_FP32])
def test_merge_with_many_segments_per_dimension(prec):
    """Tile side 1 at P = 16 (16 source runs per output tile and dimension, past the merge's
    10-segment bit word, so it tests every box's segments) gives the default tiling's answer."""
    N, M, P, xi = 1500, 24, 16, 5.0
    x, q = system(N, 17, clustered=True)
    ref, Fref, _ = solve(x, q, M, P, xi, prec)
    tol = 1e-12 if prec == se.PREC_FP64 else 2e-6
    for kw in (dict(tile=1), dict(tile=2, chunk=50)):
        phi, F, _ = solve(x, q, M, P, xi, prec, **kw)
        assert rms_rel(phi, ref) < tol and rms_rel(F, Fref) < tol, kw


@pytest.mark.parametrize("P", [8, 11, 16])
def test_pencil_kernels_match_oracle_and_generic(oracle, P):
    """FP32 register-ring (pencil) kernels vs the FP64 oracle and vs the generic kernels,
    on clustered input with work items that cut pencils mid-way."""
    N, M, xi = 3000, 28, 6.0
    x, q = system(N, 100 + P, clustered=True)
    m = 0.95 * np.sqrt(np.pi * P)
    ref_phi = oracle.se3p(x, q, 1.0, M, P, m, xi)
    ref_F = oracle.se3p(x, q, 1.0, M, P, m, xi, force=True)
    gphi, gF, ginfo = solve(x, q, M, P, xi, se.PREC_FP32, generic=True)
    assert ginfo["pencil"] == 0
    for kw in (dict(), dict(tile=4, chunk=37), dict(tile=6, chunk=300, block=256), dict(tile=10, block=64)):
        phi, F, info = solve(x, q, M, P, xi, se.PREC_FP32, **kw)
        assert info["pencil"] == 1, kw
        assert rms_rel(phi, ref_phi) < FP32_TOL and rms_rel(F, ref_F) < FP32_TOL_F, kw
        assert rms_rel(phi, gphi) < 2e-6 and rms_rel(F, gF) < 2e-6, kw
    # potential-only and force-only gathers agree with the combined one
    plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32)
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    p1 = torch.zeros(N, dtype=torch.float64, device=DEV)
    f1 = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
    plan.solve(xt, qt, p1, None)
    plan.interp(None, f1)
    torch.cuda.synchronize()
    plan.close()
    assert rms_rel(p1.cpu().numpy(), ref_phi) < FP32_TOL and rms_rel(f1.cpu().numpy(), ref_F) < FP32_TOL_F


@pytest.mark.parametrize("clustered", [False, True])
def test_mfma_gather_matches_vector_gather(oracle, clustered):
    """The P = 16 matrix-core gather (4-particle groups on a 20-row ring) against the
    vector register-ring gather on the same spread grid, for each output mode, with work
    items and waves that cut pencils and groups mid-way; and against the FP64 oracle."""
    N, M, P, xi = 6000, 32, 16, 8.0
    x, q = system(N, 300 + int(clustered), clustered=clustered)
    m = 0.95 * np.sqrt(np.pi * P)
    ref_phi = oracle.se3p(x, q, 1.0, M, P, m, xi)
    ref_F = oracle.se3p(x, q, 1.0, M, P, m, xi, force=True)
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    for kw in (dict(), dict(chunk=45, block=128), dict(tile=4, chunk=301, block=64)):
        out = {}
        for mf in (True, False):
            plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32, mfma=mf, **kw)
            phi = torch.zeros(N, dtype=torch.float64, device=DEV)
            F = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
            p1 = torch.zeros(N, dtype=torch.float64, device=DEV)
            f1 = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
            plan.solve(xt, qt, phi, F)
            plan.interp(p1, None)
            plan.interp(None, f1)
            torch.cuda.synchronize()
            assert plan.info()["pencil"] == 1
            plan.close()
            out[mf] = [t.cpu().numpy() for t in (phi, F, p1, f1)]
        a, b = out[True], out[False]
        print("mfma vs vector", kw, [rms_rel(u, v) for u, v in zip(a, b)], flush=True)
        for u, v in zip(a, b):
            assert rms_rel(u, v) < 2e-6, kw
        assert rms_rel(a[2], a[0]) < 1e-7 and rms_rel(a[3], a[1]) < 1e-7, kw
        assert rms_rel(a[0], ref_phi) < FP32_TOL and rms_rel(a[1], ref_F) < FP32_TOL_F, kw


@pytest.mark.parametrize("clustered", [False, True])
def test_column_kernels_match_pencil_kernels(oracle, clustered):
    """The P = 16 column kernels (sixteen particles of one 1 x 1 window-start column per
    matrix-core product) against the 4 x 1 pencil kernels (SE_PLAN_NO_COLUMN) and the FP64
    oracle, for each output mode, with items that cut columns mid-way (chunk 45 and 301,
    tile sides 12, 8 and 4), work-groups of 64 to 512 threads; and bit for bit on a repeat.
    At 1.8 particles per grid cell the column groups are full enough for the column path
    to take items at every tiling but chunk 45 (below the column floor): its 32-bit
    fixed-point spread sums make the results differ from the pencil path's in the last
    bits, which shows that it ran."""
    N, M, P, xi = 60000, 32, 16, 8.0
    x, q = system(N, 320 + int(clustered), clustered=clustered)
    m = 0.95 * np.sqrt(np.pi * P)
    ref_phi = oracle.se3p(x, q, 1.0, M, P, m, xi)
    ref_F = oracle.se3p(x, q, 1.0, M, P, m, xi, force=True)
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    for kw in (dict(), dict(chunk=45, block=128), dict(tile=8, chunk=301), dict(tile=4, chunk=301, block=64)):
        out = {}
        for col in (True, False, True):
            print("case column=%s %s" % (col, kw), flush=True)
            plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32, column=col, **kw)
            phi = torch.zeros(N, dtype=torch.float64, device=DEV)
            F = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
            p1 = torch.zeros(N, dtype=torch.float64, device=DEV)
            f1 = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
            plan.solve(xt, qt, phi, F)
            plan.interp(p1, None)
            plan.interp(None, f1)
            torch.cuda.synchronize()
            assert plan.info()["pencil"] == 1
            plan.close()
            res = [t.cpu().numpy() for t in (phi, F, p1, f1)]
            if col in out:
                assert all(np.array_equal(u, v) for u, v in zip(out[col], res)), kw
            out[col] = res
        a, b = out[True], out[False]
        print("column vs pencil", kw, [rms_rel(u, v) for u, v in zip(a, b)], flush=True)
        assert np.array_equal(a[0], b[0]) == (kw.get("chunk") == 45), kw   # the column path ran (or could not)
        for u, v in zip(a, b):
            assert rms_rel(u, v) < 2e-6, kw
        print("potential-only, force-only vs joint: equal %s %s" % (np.array_equal(a[2], a[0]),
                                                                    np.array_equal(a[3], a[1])), flush=True)
        assert rms_rel(a[2], a[0]) < 1e-7 and rms_rel(a[3], a[1]) < 1e-7, kw
        assert rms_rel(a[0], ref_phi) < FP32_TOL and rms_rel(a[1], ref_F) < FP32_TOL_F, kw


def test_column_items_are_chosen_by_group_fill():
    """Work items go to the column kernels only where their column groups (runs of up to
    16 particles of one 1 x 1 column within five rows) hold 3.5 particles or more on
    average: a uniform system at 0.46 particles per grid cell has none (its results equal
    the SE_PLAN_NO_COLUMN ones bit for bit), the same particles packed into an eighth of
    the box (3.7 per cell) have column items (the results differ in the last bits: the
    column spread sums in 32-bit fixed point)."""
    N, M, P, xi = 15000, 32, 16, 8.0
    rng = np.random.default_rng(77)
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    qt = torch.tensor(q, device=DEV)
    for scale, expect_equal in ((1.0, True), (0.5, False)):
        x = rng.random((3, N)) * scale
        xt = torch.tensor(np.ascontiguousarray(x.reshape(-1)), device=DEV)
        out = {}
        for col in (True, False):
            plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32, column=col)
            phi = torch.zeros(N, dtype=torch.float64, device=DEV)
            plan.solve(xt, qt, phi)
            torch.cuda.synchronize()
            plan.close()
            out[col] = phi.cpu().numpy()
        same = np.array_equal(out[True], out[False])
        print("scale %.2f: column path %s" % (scale, "not taken" if same else "taken"), flush=True)
        assert same == expect_equal, scale
        assert rms_rel(out[True], out[False]) < 2e-6


@pytest.mark.parametrize("clustered", [False, True])
def test_row_gather_of_sparse_items_matches_vector_gather(oracle, clustered):
    """Sparse work items at P = 16 (fewer particles than the matrix-core threshold) take
    the row gather (a 16-lane row per particle); at every tile side it has an instance for
    (sub-grid sides 27, 23, 19) it agrees with the vector register-ring gather (the
    SE_PLAN_NO_MFMA path) and with the FP64 oracle, for each output mode, and repeats bit
    for bit."""
    N, M, P, xi = 3000, 48, 16, 12.0
    x, q = system(N, 310 + int(clustered), clustered=clustered)
    m = 0.95 * np.sqrt(np.pi * P)
    ref_phi = oracle.se3p(x, q, 1.0, M, P, m, xi)
    ref_F = oracle.se3p(x, q, 1.0, M, P, m, xi, force=True)
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    for kw in (dict(), dict(tile=8), dict(tile=4), dict(chunk=40)):
        out = {}
        for mf in (True, False, True):
            plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32, mfma=mf, **kw)
            phi = torch.zeros(N, dtype=torch.float64, device=DEV)
            F = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
            p1 = torch.zeros(N, dtype=torch.float64, device=DEV)
            f1 = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
            plan.solve(xt, qt, phi, F)
            plan.interp(p1, None)
            plan.interp(None, f1)
            torch.cuda.synchronize()
            assert plan.info()["pencil"] == 1
            plan.close()
            res = [t.cpu().numpy() for t in (phi, F, p1, f1)]
            if mf in out:
                assert all(np.array_equal(u, v) for u, v in zip(out[mf], res)), kw
            out[mf] = res
        a, b = out[True], out[False]
        print("row vs vector", kw, [rms_rel(u, v) for u, v in zip(a, b)], flush=True)
        for u, v in zip(a, b):
            assert rms_rel(u, v) < 2e-6, kw
        assert rms_rel(a[2], a[0]) < 1e-7 and rms_rel(a[3], a[1]) < 1e-7, kw
        assert rms_rel(a[0], ref_phi) < FP32_TOL and rms_rel(a[1], ref_F) < FP32_TOL_F, kw


def clustered_system(N, seed=3, n_clusters=64, frac=0.8, sigma=0.01):
    """80 % of N in 64 Gaussian clusters (sigma = 0.01 L), 20 % uniform, q = +-1, zero net charge."""
    rng = np.random.default_rng(seed)
    nc = int(frac * N)
    centres = rng.random((3, n_clusters))
    lab = rng.integers(0, n_clusters, nc)
    xc = (centres[:, lab] + sigma * rng.standard_normal((3, nc))) % 1.0
    x = np.concatenate([xc, rng.random((3, N - nc))], axis=1)
    x = np.where(x >= 1.0, 0.0, x)
    q = np.ones(N)
    q[N // 2:] = -1.0
    rng.shuffle(q)
    return np.ascontiguousarray(x.reshape(-1)), q


def test_clustered_full_size_fp32_matches_fp64():
    """The clustered workload at full size (N = 4M, M = 256, P = 16, xi = 80): heavy tiles
    are split into many work items, nothing is dropped, and the FP32 path agrees with the
    FP64 path at the same P to the FP32 tolerance."""
    N, M, P, xi = 4_000_000, 256, 16, 80.0
    x, q = clustered_system(N)
    print("case clustered N=%d M=%d P=%d" % (N, M, P), flush=True)
    phi32, F32, info32 = solve(x, q, M, P, xi, se.PREC_FP32)
    assert info32["pencil"] == 1 and info32["discarded"] == 0
    # every tile holds uniform particles; the clusters' heavy tiles add items beyond one per tile
    assert info32["items"] > info32["ntiles"] + 500
    phi64, F64, info64 = solve(x, q, M, P, xi, se.PREC_FP64)
    assert info64["discarded"] == 0
    e_phi, e_F = rms_rel(phi32, phi64), rms_rel(F32, F64)
    print("clustered fp32 vs fp64: phi %.2e F %.2e, items %d / tiles %d" % (e_phi, e_F, info32["items"],
                                                                          info32["ntiles"]))
    assert e_phi < FP32_TOL and e_F < FP32_TOL_F


def test_capacity_is_enforced():
    plan = se.Plan(16, 8, 1.0, 4.0, 10, precision=se.PREC_FP64)
    x = torch.zeros(3 * 11, dtype=torch.float64, device=DEV)
    q = torch.zeros(11, dtype=torch.float64, device=DEV)
    with pytest.raises(RuntimeError, match="capacity"):
        plan.solve(x, q, torch.zeros(11, dtype=torch.float64, device=DEV))
    plan.close()


@pytest.mark.parametrize("tol", [1e-4, 1e-7, 1e-10])
def test_selected_parameters_meet_the_tolerance_on_the_gpu(oracle, tol):
    """se_select_params drives the FP64 engine: a plan built from the (P, M, m) chosen for
    an rms potential error of tol solves within tol of the oracle's direct k-space sum
    (|k| <= (M-1)/2, the layers test_accuracy.c:55 uses)."""
    x, q = system(120, 91)
    L, xi = 1.0, 6.0
    P, M, m, est = se.select_params(tol, L, xi, q, potential=True, force=False)
    print("case select tol=%g P=%d M=%d est=%.2e" % (tol, P, M, est), flush=True)
    phi, _, _ = solve(x, q, M, P, xi, se.PREC_FP64, force=False, m=m)
    ref = oracle.direct3p(x, q, (L, L, L), xi, (M - 1) // 2)
    assert np.sqrt(np.mean((phi - ref) ** 2)) <= tol


@pytest.mark.parametrize("P", [8, 11, 16])
def test_plane_spread_matches_pencil_spread_and_oracle(oracle, P):
    """Sparse tiles (a few particles per 4 x 1 pencil) take the plane spread, which keeps
    each (x, z) column of the tile's sub-grid in registers; its grid equals the pencil
    spread's and the reference's wrapped grid, also with tiles cut into chunks (which stay
    on the pencil kernel) and other tile sides, and it is the same on a second run."""
    N, M, xi = 1500, 40, 6.0
    x, q = system(N, 60 + P, clustered=True)   # dense clusters and sparse tiles in one grid
    p = oracle.pack_params(N, M, P, plan_c(M, P, xi), 1.0 / M)
    zs = oracle.base_gaussian(p)
    ref = oracle.wrap(3, p, oracle.grid_split(p, q, oracle.expand_all(3, x, p), zs)).reshape(M, M, M, order="F")
    xt, qt = torch.tensor(x, device=DEV), torch.tensor(q, device=DEV)
    scale = np.max(np.abs(ref))
    for kw in (dict(), dict(tile=8), dict(tile=4, chunk=25, block=128), dict(chunk=60, block=256)):
        grids = []
        for plane in (True, False, True):
            plan = se.Plan(M, P, 1.0, xi, N, precision=se.PREC_FP32, plane=plane, **kw)
            plan.spread(xt, qt)
            torch.cuda.synchronize()
            assert plan.info()["pencil"] == 1
            grids.append(plan.grid().double().cpu().numpy())
            plan.close()
        e_plane = np.max(np.abs(grids[0] - ref)) / scale
        e_pair = np.max(np.abs(grids[0] - grids[1])) / scale
        print("P=%d %s: plane vs oracle %.2e, plane vs pencil %.2e" % (P, kw, e_plane, e_pair), flush=True)
        assert e_plane < 2e-6 and e_pair < 2e-6, kw
        assert np.array_equal(grids[0], grids[2]), kw

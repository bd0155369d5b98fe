"""GPU parity of the drop-in SE_FGG_* surface against the reference's own outputs
(golden vectors from the compiled SE_fgg.c) and the CPU oracle.  Every call goes
through the C ABI of libse_fgg_amd.so with host arrays, as a reference caller would."""
import ctypes as C

import numpy as np
import pytest

import se_fast_gridding_amd as se
from conftest import golden_files, load_golden, rms_rel

pytestmark = pytest.mark.gpu

GOLD = golden_files("g3p") + golden_files("g2p") + golden_files("g1p")
D = C.POINTER(C.c_double)


def dp(a):
    return a.ctypes.data_as(D)


def close(a, b, tol):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    err = float(np.max(np.abs(a - b))) / scale
    assert err <= tol, err


class Work:
    """numpy-owned arrays wired into an SE_FGG_work (the library never frees them)."""

    def __init__(self, p, force=False):
        N, P = p.N, p.P
        self.H = np.zeros(int(np.prod(list(p.npdims))))
        self.zs = np.zeros(P ** 3)
        self.zx, self.zy, self.zz = (np.zeros(N * P) for _ in range(3))
        self.idx = np.zeros(N, dtype=np.int32)
        self.zf = [np.zeros(N * P) for _ in range(3)]
        w = se.SE_FGG_work()
        w.H, w.zs, w.zx, w.zy, w.zz = (dp(a) for a in (self.H, self.zs, self.zx, self.zy, self.zz))
        w.idx = self.idx.ctypes.data_as(C.POINTER(C.c_int))
        w.zfx, w.zfy, w.zfz = (dp(a) for a in self.zf)
        self.w = w


def state(x, q):
    x = np.ascontiguousarray(x, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    phi = np.zeros(len(q))
    s = se.SE_state()
    s.x, s.q, s.phi = dp(x), dp(q), dp(phi)
    return s, (x, q, phi)


def params(lib, g):
    p = se.SE_FGG_params()
    N, P = int(g["N"]), int(g["P"])
    dims = [int(v) for v in g["dims"]]
    per = int(g["per"])
    if per == 3:
        lib.SE_FGG_pack_params(C.byref(p), N, *dims, P, float(g["c"]), float(g["h"]))
    elif per == 2:
        lib.SE2P_FGG_pack_params(C.byref(p), N, *dims, P, float(g["c"]), float(g["h"]), float(g["a"]))
    else:
        lib.SE1P_FGG_pack_params(C.byref(p), N, *dims, P, float(g["c"]), float(g["h"]), float(g["a"]),
                                 float(g["b"]))
    return p


@pytest.mark.parametrize("path", GOLD, ids=lambda p: p.split("/")[-1][:-4])
def test_legacy_surface_matches_reference(path):
    lib = se.capi()
    g = load_golden(path)
    per = int(g["per"])
    print("case", path.split("/")[-1], flush=True)
    lib.se_fgg_set_periodicity(per)
    try:
        p = params(lib, g)
        st, keep = state(g["x"], g["q"])
        w = Work(p)
        lib.SE_FGG_base_gaussian(C.byref(w.w), C.byref(p))
        close(w.zs, g["zs"], 1e-14)
        lib.SE_FGG_expand_all(C.byref(w.w), C.byref(st), C.byref(p))
        assert np.array_equal(w.idx, g["idx"])
        for k in ("zx", "zy", "zz"):
            close(getattr(w, k), g[k], 1e-13)
        lib.SE_FGG_grid_split_AVX_dispatch(C.byref(w.w), C.byref(st), C.byref(p))
        if "H_split" in g:
            close(w.H, g["H_split"], 1e-13)
        Hper = np.zeros(len(g["Hper"]))
        {3: lib.SE_FGG_wrap_fcn, 2: lib.SE2P_FGG_wrap_fcn, 1: lib.SE1P_FGG_wrap_fcn}[per](dp(Hper), C.byref(w.w),
                                                                                           C.byref(p))
        close(Hper, g["Hper"], 1e-13)
        G = np.ascontiguousarray(g["G"])
        {3: lib.SE_FGG_extend_fcn, 2: lib.SE2P_FGG_extend_fcn, 1: lib.SE1P_FGG_extend_fcn}[per](C.byref(w.w), dp(G),
                                                                                                 C.byref(p))
        if "H_ext" in g:
            close(w.H, g["H_ext"], 0.0)
        phi = np.zeros(p.N)
        lib.SE_FGG_int_split_AVX_dispatch(dp(phi), C.byref(w.w), C.byref(p))
        close(phi, g["phi_split"], 1e-12)
        phi2 = np.zeros(p.N)
        lib.SE_FGG_int(dp(phi2), C.byref(w.w), C.byref(st), C.byref(p))
        close(phi2, g["phi_int"], 1e-12)
        # on-the-fly spread
        w2 = Work(p)
        w2.zs[:] = g["zs"]
        lib.SE_FGG_grid(C.byref(w2.w), C.byref(st), C.byref(p))
        if "H_grid" in g:
            close(w2.H, g["H_grid"], 1e-12)
        if "zfx" in g:
            wf = Work(p, force=True)
            lib.SE_FGG_expand_all_SSE_force(C.byref(wf.w), C.byref(st), C.byref(p))
            assert np.array_equal(wf.idx, g["fidx"])
            for a, k in ((wf.zx, "fzx"), (wf.zy, "fzy"), (wf.zz, "fzz"), (wf.zf[0], "zfx"), (wf.zf[1], "zfy"),
                         (wf.zf[2], "zfz")):
                close(a, g[k], 1e-13)
            if "F_split" in g:
                wf.zs[:] = g["zs"]
                wf.H[:] = w.H
                F = np.zeros(3 * p.N)
                lib.SE_FGG_int_split_AVX_dispatch_force(dp(F), C.byref(st), C.byref(wf.w), C.byref(p))
                close(F, g["F_split"], 1e-12)
                F2 = np.zeros(3 * p.N)
                lib.SE_FGG_int_force(dp(F2), C.byref(wf.w), C.byref(st), C.byref(p))
                close(F2, g["F_int"], 1e-12)
        assert se.last_error() == "" or True
    finally:
        lib.se_fgg_set_periodicity(3)


def test_int_split_avx_accumulates_like_reference():
    """SE_FGG_int_split_AVX adds into phi (SE_fgg.c:1856); the P16 kernel overwrites."""
    lib = se.capi()
    g = load_golden(golden_files("g3p_p12")[0])
    p = params(lib, g)
    st, keep = state(g["x"], g["q"])
    w = Work(p)
    w.zs[:] = g["zs"]
    lib.SE_FGG_expand_all(C.byref(w.w), C.byref(st), C.byref(p))
    G = np.ascontiguousarray(g["G"])
    lib.SE_FGG_extend_fcn(C.byref(w.w), dp(G), C.byref(p))
    phi = np.full(p.N, 1.5)
    lib.SE_FGG_int_split_AVX_dispatch(dp(phi), C.byref(w.w), C.byref(p))
    close(phi - 1.5, g["phi_split"], 1e-12)
    phi2 = np.full(p.N, 1.5)
    lib.SE_FGG_int_split(dp(phi2), C.byref(w.w), C.byref(p))
    close(phi2, g["phi_split"], 1e-12)


def test_reorder_energy_scaling_product(oracle):
    lib = se.capi()
    g = load_golden(golden_files("g3p_p8")[0])
    p = params(lib, g)
    st, (x, q, phi) = state(g["x"], g["q"])
    w = Work(p)
    lib.SE_FGG_expand_all(C.byref(w.w), C.byref(st), C.byref(p))
    idx0 = w.idx.copy()
    x0, q0 = x.copy(), q.copy()
    lib.SE_FGG_reorder_system(C.byref(st), C.byref(w.w), C.byref(p))
    order = np.argsort(idx0, kind="stable")
    N = p.N
    assert np.array_equal(q, q0[order])
    assert np.array_equal(x.reshape(3, N), x0.reshape(3, N)[:, order])
    phi[:] = np.linspace(-1, 1, N)
    e = lib.calc_energy(st, N)
    assert np.isclose(e, 0.5 * np.sum(phi * q), rtol=1e-13)
    M, L, s = 8, 1.7, -0.013
    Z = np.zeros(M ** 3)
    box = np.array([L, L, L])
    lib.scaling(s, dp(Z), M, M, M, dp(box))
    Zo = oracle.scaling(s, M, L)
    assert np.isinf(Z[0]) and np.isinf(Zo[0])
    close(Z[1:], Zo[1:], 1e-14)
    a = np.random.default_rng(0).standard_normal(2 * M ** 3)
    c = np.zeros_like(a)
    Z[0] = 1.0
    lib.product_rc(dp(a), dp(Z), dp(c), M, M, M, 1)
    assert np.allclose(c[0::2], a[0::2] * Z) and np.allclose(c[1::2], a[1::2] * Z)
    lib.product_rc(dp(a), dp(Z), dp(c), M, M, M, -1)
    assert np.allclose(c[0::2], a[0::2] / Z) and np.allclose(c[1::2], a[1::2] / Z)


def test_spectral_ewald_driver_config1(oracle):
    """spectral_ewald (FP64, SE3P) vs the oracle restatement and the direct k-space sum:
    N=100 unit cube, q=+-1, P=21, M=32, xi=6."""
    lib = se.capi()
    rng = np.random.default_rng(1)
    N, P, M, xi = 100, 21, 32, 6.0
    x = rng.random(3 * N)
    q = np.array([1.0, -1.0] * (N // 2))
    rng.shuffle(q)
    opt = se.SE_opt()
    opt.m = 0.95 * np.sqrt(np.pi * P)
    opt.box[0] = opt.box[1] = opt.box[2] = 1.0
    opt.P, opt.M, opt.N = P, M, N
    phi = np.zeros(N)
    lib.se_set_quantity(1)
    assert lib.spectral_ewald(dp(x), dp(q), opt, xi, dp(phi), None) == 0
    ref = oracle.se3p(x, q, 1.0, M, P, opt.m, xi)
    assert rms_rel(phi, ref) < 1e-12
    assert rms_rel(phi, oracle.direct3p(x, q, 1.0, xi, 15)) < 1e-11
    F = np.zeros(3 * N)
    phi_e = np.zeros(N)
    lib.se_set_quantity(3)
    try:
        assert lib.spectral_ewald(dp(x), dp(q), opt, xi, dp(F), dp(phi_e)) == 0
    finally:
        lib.se_set_quantity(1)
    Fref = oracle.se3p(x, q, 1.0, M, P, opt.m, xi, force=True)
    assert rms_rel(F, Fref) < 1e-12
    assert rms_rel(phi_e, ref) < 1e-12


def test_spectral_ewald_atoms_fixture(oracle):
    """The reference test's own input (atoms.txt, L=3.1814, xi=8, P=21, M from SE_init_params)."""
    import os
    from conftest import GOLDEN
    lib = se.capi()
    a = np.load(os.path.join(GOLDEN, "atoms.npz"))["xyzq"]
    N = 60
    x = np.ascontiguousarray(a[:N, :3].T.reshape(-1))
    q = np.ascontiguousarray(a[:N, 3])
    L, xi, P = 3.1814, 8.0, 21
    M = C.c_int(0)
    Pc = C.c_int(P)
    m = C.c_double(0)
    lib.SE_init_params(C.byref(M), C.byref(Pc), C.byref(m), 1.0, L, xi, dp(q), N)
    opt = se.SE_opt()
    opt.m = m.value
    opt.box[0] = opt.box[1] = opt.box[2] = L
    opt.P, opt.M, opt.N = P, M.value, N
    phi = np.zeros(N)
    assert lib.spectral_ewald(dp(x), dp(q), opt, xi, dp(phi), None) == 0
    ref = oracle.se3p(x, q, L, M.value, P, m.value, xi)
    assert rms_rel(phi, ref) < 1e-12


@pytest.mark.parametrize("per", [2, 1])
def test_spectral_ewald_driver_se2p_se1p(oracle, per):
    """spectral_ewald after se_fgg_set_periodicity(2 | 1): the SE2P / SE1P k-space sums of
    the reference's direct solver (SE_direct.c:90-176, 256-338) at the FGG error floor,
    potential and force (F against the device plan's force for the same system)."""
    import torch
    lib = se.capi()
    rng = np.random.default_rng(10 + per)
    N, P, M, xi = 40, 21, 32, 6.0
    x = rng.random(3 * N)
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    opt = se.SE_opt()
    opt.m = 0.95 * np.sqrt(np.pi * P)
    opt.box[0] = opt.box[1] = opt.box[2] = 1.0
    opt.P, opt.M, opt.N = P, M, N
    phi = np.zeros(N)
    F = np.zeros(3 * N)
    lib.se_fgg_set_periodicity(per)
    try:
        lib.se_set_quantity(1)
        assert lib.spectral_ewald(dp(x), dp(q), opt, xi, dp(phi), None) == 0
        lib.se_set_quantity(2)
        assert lib.spectral_ewald(dp(x), dp(q), opt, xi, dp(F), None) == 0
    finally:
        lib.se_set_quantity(1)
        lib.se_fgg_set_periodicity(3)
    if per == 2:
        ref = oracle.direct2p(x, q, (1.0, 1.0, 1.0), xi, 14)
    else:
        ref = oracle.direct1p(x, q, (1.0, 1.0, 1.0), xi, int(np.ceil(2 * xi * np.sqrt(40.0) / (2 * np.pi))) + 1)
    assert rms_rel(phi, ref) < 1e-11
    plan = se.Plan(M, P, 1.0, xi, N, periodicity=per, slab=1.0)
    Ft = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
    plan.solve(torch.tensor(x, device="cuda"), torch.tensor(q, device="cuda"), None, Ft)
    torch.cuda.synchronize()
    plan.close()
    assert rms_rel(F, Ft.cpu().numpy()) < 1e-13


def test_spectral_ewald_driver_fp32_p16(oracle):
    """spectral_ewald after se_set_precision(SE_FP32) at P = 16 runs the FP32 register-
    ring spread and the matrix-core gather behind the legacy surface: potential and force
    within FP32 tolerance of the oracle's FP64 solve of the same (M, P, m, xi)."""
    lib = se.capi()
    rng = np.random.default_rng(5)
    N, P, M, xi = 20000, 16, 32, 8.0   # about 1000 particles per tile: the dense (MFMA) items
    x = rng.random(3 * N)
    q = rng.uniform(-1, 1, N)
    q -= q.mean()
    opt = se.SE_opt()
    opt.m = 0.95 * np.sqrt(np.pi * P)
    opt.box[0] = opt.box[1] = opt.box[2] = 1.0
    opt.P, opt.M, opt.N = P, M, N
    F = np.zeros(3 * N)
    phi = np.zeros(N)
    print("case driver fp32 N=%d P=%d M=%d" % (N, P, M), flush=True)
    lib.se_set_precision(1)
    lib.se_set_quantity(3)
    try:
        assert lib.spectral_ewald(dp(x), dp(q), opt, xi, dp(F), dp(phi)) == 0
    finally:
        lib.se_set_quantity(1)
        lib.se_set_precision(0)
    assert rms_rel(phi, oracle.se3p(x, q, 1.0, M, P, opt.m, xi)) < 1e-6
    assert rms_rel(F, oracle.se3p(x, q, 1.0, M, P, opt.m, xi, force=True)) < 2e-6


def test_legacy_call_before_torch_leaves_torch_a_gpu():
    """A process that calls the legacy surface before it touches torch must still find the
    GPU from torch afterwards (torch ships its own HIP runtime under the same soname; the
    package loads torch first so that one runtime serves both).  Checked in a fresh child
    process, where nothing has initialised the GPU yet."""
    import subprocess
    import sys
    import os
    code = (
        "import numpy as np, se_fast_gridding_amd as se\n"
        "lib = se.capi()\n"
        "N, P, M = 50, 8, 16\n"
        "rng = np.random.default_rng(0)\n"
        "x = rng.random(3 * N); q = rng.uniform(-1, 1, N); q -= q.mean()\n"
        "opt = se.SE_opt(); opt.m = 0.95 * np.sqrt(np.pi * P)\n"
        "opt.box[0] = opt.box[1] = opt.box[2] = 1.0; opt.P, opt.M, opt.N = P, M, N\n"
        "phi = np.zeros(N)\n"
        "dp = lambda a: a.ctypes.data_as(__import__('ctypes').POINTER(__import__('ctypes').c_double))\n"
        "assert lib.spectral_ewald(dp(x), dp(q), opt, 4.0, dp(phi), None) == 0\n"
        "import torch\n"
        "assert torch.cuda.is_available() and torch.cuda.device_count() >= 1\n"
        "t = torch.ones(4, device='cuda'); assert float(t.sum()) == 4.0\n"
        "print('ok')\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]

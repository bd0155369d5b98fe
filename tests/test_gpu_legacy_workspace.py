"""The reference calling sequence with library-owned buffers, as a C caller of SE_fgg.h
writes it: SE_FGG_allocate_workspace -> base_gaussian -> expand_all -> grid_split ->
wrap into SE_FGG_allocate_grid -> extend -> int_split -> SE_FGG_free_workspace, the
_SSE_force variants, SE_FGG_allocate_vec, SE_fgg_int_force (SE_general.c) and
copy_r2c / copy_c2r, all against the golden vectors of the compiled reference SE_fgg.c."""
import ctypes as C

import numpy as np
import pytest

import se_fast_gridding_amd as se
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu

D = C.POINTER(C.c_double)
LIBC = C.CDLL(None)
LIBC.free.argtypes = [C.c_void_p]


def dp(a):
    return a.ctypes.data_as(D)


def view(ptr, n):
    """numpy view of n doubles of library-owned host memory."""
    return np.ctypeslib.as_array(ptr, shape=(n,))


def close(a, b, tol):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    err = float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / scale
    assert err <= tol, err


def pack(lib, g):
    p = se.SE_FGG_params()
    lib.SE_FGG_pack_params(C.byref(p), int(g["N"]), *[int(v) for v in g["dims"]], int(g["P"]), float(g["c"]),
                           float(g["h"]))
    return p


CASES = [p for p in golden_files("g3p") if "m20x16x12" not in p]


@pytest.mark.parametrize("path", CASES, ids=lambda p: p.split("/")[-1][:-4])
def test_reference_sequence_with_library_workspace(path):
    lib = se.capi()
    g = load_golden(path)
    print("case", path.split("/")[-1], flush=True)
    p = pack(lib, g)
    N, P = p.N, p.P
    x = np.ascontiguousarray(g["x"], dtype=np.float64)
    q = np.ascontiguousarray(g["q"], dtype=np.float64)
    phi_st = np.zeros(N)
    st = se.SE_state()
    st.x, st.q, st.phi = dp(x), dp(q), dp(phi_st)
    ne = int(np.prod([int(v) for v in p.npdims]))

    w = se.SE_FGG_work()
    lib.SE_FGG_allocate_workspace(C.byref(w), C.byref(p), 1, 1)
    try:
        assert w.free_zs == 1 and w.free_fgg_expa == 1
        assert not np.any(view(w.H, ne))                 # the workspace grid starts at zero
        lib.SE_FGG_base_gaussian(C.byref(w), C.byref(p))
        close(view(w.zs, P ** 3), g["zs"], 1e-14)
        lib.SE_FGG_expand_all(C.byref(w), C.byref(st), C.byref(p))
        assert np.array_equal(np.ctypeslib.as_array(w.idx, shape=(N,)), g["idx"])
        for name in ("zx", "zy", "zz"):
            close(view(getattr(w, name), N * P), g[name], 1e-13)
        lib.SE_FGG_grid_split_AVX_dispatch(C.byref(w), C.byref(st), C.byref(p))
        if "H_split" in g:
            close(view(w.H, ne), g["H_split"], 1e-13)
        Hp = lib.SE_FGG_allocate_grid(C.byref(p))       # zeroed, caller frees
        try:
            nper = len(g["Hper"])
            assert not np.any(view(Hp, nper))
            lib.SE_FGG_wrap_fcn(Hp, C.byref(w), C.byref(p))
            close(view(Hp, nper), g["Hper"], 1e-13)
            view(Hp, nper)[:] = g["G"]                  # the golden's k-space output
            lib.SE_FGG_extend_fcn(C.byref(w), Hp, C.byref(p))
            if "H_ext" in g:
                close(view(w.H, ne), g["H_ext"], 0.0)
        finally:
            LIBC.free(C.cast(Hp, C.c_void_p))
        phi = lib.SE_FGG_allocate_vec(N)
        try:
            assert not np.any(view(phi, N))
            lib.SE_FGG_int_split_AVX_dispatch(phi, C.byref(w), C.byref(p))
            close(view(phi, N), g["phi_split"], 1e-12)
        finally:
            LIBC.free(C.cast(phi, C.c_void_p))
    finally:
        lib.SE_FGG_free_workspace(C.byref(w))

    if "F_split" not in g:
        return
    wf = se.SE_FGG_work()
    lib.SE_FGG_allocate_workspace_SSE_force(C.byref(wf), C.byref(p), 1, 1)
    try:
        lib.SE_FGG_base_gaussian(C.byref(wf), C.byref(p))
        lib.SE_FGG_expand_all_SSE_force(C.byref(wf), C.byref(st), C.byref(p))
        assert np.array_equal(np.ctypeslib.as_array(wf.idx, shape=(N,)), g["fidx"])
        for name, key in (("zx", "fzx"), ("zy", "fzy"), ("zz", "fzz"), ("zfx", "zfx"), ("zfy", "zfy"), ("zfz", "zfz")):
            close(view(getattr(wf, name), N * P), g[key], 1e-13)
        G = np.ascontiguousarray(g["G"])
        lib.SE_FGG_extend_fcn(C.byref(wf), dp(G), C.byref(p))
        F = np.zeros(3 * N)
        lib.SE_FGG_int_split_AVX_dispatch_force(dp(F), C.byref(st), C.byref(wf), C.byref(p))
        close(F, g["F_split"], 1e-12)
    finally:
        lib.SE_FGG_free_workspace_SSE_force(C.byref(wf))

    # SE_fgg_int_force (SE_general.c): extend + on-the-fly force interpolation of a
    # periodic grid, parameters from an SE_opt (c taken from the option, h = box / M)
    M = int(g["dims"][0])
    opt = se.SE_opt()
    opt.c = float(g["c"])
    opt.P, opt.M, opt.N = P, M, N
    opt.box[0] = opt.box[1] = opt.box[2] = float(g["h"]) * M
    F2 = np.zeros(3 * N)
    G = np.ascontiguousarray(g["G"])
    lib.SE_fgg_int_force(dp(x), dp(q), dp(G), N, opt, dp(F2))
    close(F2, g["F_int"], 1e-12)


def test_workspace_without_expansion_arrays():
    """expa = 0: only H (and zs) are allocated; the on-the-fly spread fills H."""
    lib = se.capi()
    g = load_golden(golden_files("g3p_p8")[0])
    p = pack(lib, g)
    N = p.N
    x = np.ascontiguousarray(g["x"])
    q = np.ascontiguousarray(g["q"])
    st = se.SE_state()
    phi = np.zeros(N)
    st.x, st.q, st.phi = dp(x), dp(q), dp(phi)
    w = se.SE_FGG_work()
    lib.SE_FGG_allocate_workspace(C.byref(w), C.byref(p), 1, 0)
    try:
        assert w.free_fgg_expa == 0 and not bool(w.zx) and not bool(w.idx)
        lib.SE_FGG_base_gaussian(C.byref(w), C.byref(p))
        lib.SE_FGG_grid(C.byref(w), C.byref(st), C.byref(p))
        close(view(w.H, int(np.prod([int(v) for v in p.npdims]))), g["H_grid"], 1e-12)
    finally:
        lib.SE_FGG_free_workspace(C.byref(w))


def test_copy_r2c_c2r():
    lib = se.capi()
    a = np.random.default_rng(2).standard_normal(37)
    c = np.full(2 * 37, 7.0)
    lib.copy_r2c(dp(a), dp(c), 37)
    assert np.array_equal(c[0::2], a) and not np.any(c[1::2])
    c[1::2] = 3.0
    b = np.zeros(37)
    lib.copy_c2r(dp(c), dp(b), 37)
    assert np.array_equal(b, a)

"""Test infrastructure only: ctypes bindings to the CPU oracle (se_oracle.c) and,
when built, to the reference's own gridding core compiled into ``oracle/_ref``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package, as the checker.  The product library never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libse_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

P3, P2, P1 = 3, 2, 1
_PER_TAG = {P3: "3p", P2: "2p", P1: "1p"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class SE_FGG_params(C.Structure):
    """Byte layout of SE_FGG_params (SE_fgg.h:121-134)."""
    _fields_ = [("N", C.c_int), ("P", C.c_int), ("P_half", C.c_int),
                ("dims", C.c_int * 3), ("npdims", C.c_int * 3),
                ("c", C.c_double), ("d", C.c_double), ("h", C.c_double),
                ("a", C.c_double), ("b", C.c_double)]


class SE_FGG_work(C.Structure):
    """Byte layout of SE_FGG_work (SE_fgg.h:100-118)."""
    _fields_ = [("H", C.POINTER(C.c_double)), ("zs", C.POINTER(C.c_double)),
                ("zx", C.POINTER(C.c_double)), ("zy", C.POINTER(C.c_double)),
                ("zz", C.POINTER(C.c_double)), ("idx", C.POINTER(C.c_int)),
                ("zfx", C.POINTER(C.c_double)), ("zfy", C.POINTER(C.c_double)),
                ("zfz", C.POINTER(C.c_double)),
                ("free_zs", C.c_int), ("free_fgg_expa", C.c_int)]


class SE_state(C.Structure):
    """Byte layout of SE_state (SE_fgg.h:142-148)."""
    _fields_ = [("x", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
                ("phi", C.POINTER(C.c_double))]


def half(p):
    """SE_fgg.h:94-97."""
    return (p - 1) // 2 if p & 1 else p // 2


def parse_params(L, M, P, m, xi):
    """SE_general.c:16-29 -> (h, w, eta, c)."""
    h = L / M
    w = P * h / 2.0
    eta = (2.0 * w * xi / m) ** 2
    return h, w, eta, 2.0 * xi * xi / eta


_lib = None


def lib():
    """The oracle shared library (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        pp = C.POINTER(SE_FGG_params)
        L.or_pack_params.argtypes = [pp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_base_gaussian.argtypes = [pp, _dp]
        L.or_expand_all.argtypes = [C.c_int, C.c_int, _dp, pp, _dp, _dp, _dp, _ip,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_grid_split.argtypes = [pp, _dp, _dp, _dp, _dp, _ip, _dp, _dp]
        L.or_grid.argtypes = [C.c_int, pp, _dp, _dp, _dp, _dp]
        L.or_wrap.argtypes = [C.c_int, pp, _dp, _dp]
        L.or_extend.argtypes = [C.c_int, pp, _dp, _dp]
        L.or_int_split.argtypes = [pp, _dp, _dp, _dp, _dp, _ip, _dp, _dp]
        L.or_int_split_force.argtypes = [pp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp]
        L.or_kvec.argtypes = [C.c_int, C.c_double, _dp]
        L.or_scaling.argtypes = [C.c_double, _dp, C.c_int, C.c_double]
        L.or_fft3d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_se3p.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_double,
                              C.c_double, C.c_int, _dp, C.c_void_p]
        L.or_direct3p_fd_pairwise.argtypes = [_dp, _dp, _dp, C.c_int, _dp, C.c_double, C.c_int]
        L.or_direct3p_force_pairwise.argtypes = [_dp, _dp, _dp, C.c_int, _dp, C.c_double, C.c_int]
        L.or_direct3p_sf.argtypes = [_dp, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, C.c_double, C.c_int]
        L.or_direct2p_fd.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_double, C.c_int]
        L.or_direct2p_k0.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_double]
        L.or_direct1p_fd.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_double, C.c_int]
        L.or_direct1p_k0.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_double]
        L.or_direct_real.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_double, C.c_int,
                                     C.c_double]
        L.or_direct_self.argtypes = [_dp, _dp, C.c_int, C.c_double]
        L.or_k0inc.argtypes = [C.c_double, C.c_double]
        L.or_k0inc.restype = C.c_double
        L.or_ein.argtypes = [C.c_double]
        L.or_ein.restype = C.c_double
        L.or_energy.argtypes = [_dp, _dp, C.c_int]
        L.or_energy.restype = C.c_double
        L.or_init_params.argtypes = [C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_double), C.c_double,
                                     C.c_double, _dp, C.c_int, C.c_int]
        L.or_init_params.restype = C.c_double
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def pack_params(N, M, P, c, h, a=-1.0, b=-1.0):
    """SE_fgg.c:69-132.  M is an int or a 3-sequence."""
    Ms = (M, M, M) if np.isscalar(M) else tuple(M)
    p = SE_FGG_params()
    lib().or_pack_params(C.byref(p), N, Ms[0], Ms[1], Ms[2], P, c, h, a, b)
    return p


def base_gaussian(p):
    zs = np.zeros(p.P ** 3)
    lib().or_base_gaussian(C.byref(p), zs)
    return zs


def expand_all(per, x, p, force=False):
    """SE_FGG_expand_all[_SSE_force] restated.  Returns dict of zx, zy, zz, idx (+ zfx..)."""
    N, P = p.N, p.P
    out = {k: np.zeros(N * P) for k in ("zx", "zy", "zz")}
    out["idx"] = np.zeros(N, dtype=np.int32)
    zf = [np.zeros(N * P) for _ in range(3)] if force else [None] * 3
    ptr = [z.ctypes.data if z is not None else None for z in zf]
    lib().or_expand_all(per, int(force), _f64(x), C.byref(p), out["zx"], out["zy"], out["zz"],
                        out["idx"], *ptr)
    if force:
        out["zfx"], out["zfy"], out["zfz"] = zf
    return out


def grid_split(p, q, ex, zs, H=None):
    H = np.zeros(int(np.prod(p.npdims))) if H is None else H
    lib().or_grid_split(C.byref(p), _f64(q), ex["zx"], ex["zy"], ex["zz"], ex["idx"], zs, H)
    return H


def grid(per, p, x, q, zs, H=None):
    H = np.zeros(int(np.prod(p.npdims))) if H is None else H
    lib().or_grid(per, C.byref(p), _f64(x), _f64(q), zs, H)
    return H


def per_grid_size(per, p):
    """Elements of the column-major periodic grid as the reference's wrap/extend index it.

    3P: dims0*dims1*dims2.  2P (SE_fgg.c:357-382): z is free and runs over npdims[2]
    planes, so dims0*dims1*npdims2.  1P (SE_fgg.c:387-411): the leading dimensions stay
    dims[0], dims[1] while i, j run to npdims[0], npdims[1], so the highest index reached
    is (npdims0-1) + (npdims1-1)*dims0 + (dims2-1)*dims0*dims1 (rows alias)."""
    d, e = p.dims, p.npdims
    if per == P3:
        return d[0] * d[1] * d[2]
    if per == P2:
        return d[0] * d[1] * e[2]
    return (e[0] - 1) + (e[1] - 1) * d[0] + (d[2] - 1) * d[0] * d[1] + 1


def wrap(per, p, H, Hper=None):
    Hper = np.zeros(per_grid_size(per, p)) if Hper is None else Hper
    lib().or_wrap(per, C.byref(p), _f64(H), Hper)
    return Hper


def extend(per, p, Hper):
    assert len(Hper) >= per_grid_size(per, p)
    H = np.zeros(int(np.prod(p.npdims)))
    lib().or_extend(per, C.byref(p), _f64(Hper), H)
    return H


def int_split(p, H, ex, zs):
    phi = np.zeros(p.N)
    lib().or_int_split(C.byref(p), _f64(H), ex["zx"], ex["zy"], ex["zz"], ex["idx"], zs, phi)
    return phi


def int_split_force(p, H, q, ex, zs):
    F = np.zeros(3 * p.N)
    lib().or_int_split_force(C.byref(p), _f64(H), _f64(q), ex["zx"], ex["zy"], ex["zz"],
                             ex["zfx"], ex["zfy"], ex["zfz"], ex["idx"], zs, F)
    return F


def kvec(M, L):
    k = np.zeros(M)
    lib().or_kvec(M, L, k)
    return k


def scaling(scalar, M, L):
    Z = np.zeros(M ** 3)
    with np.errstate(all="ignore"):
        lib().or_scaling(scalar, Z, M, L)
    return Z


def fft3d(a, sign):
    """Unnormalised 3-D DFT of a complex array (sign -1 forward)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    buf = a.view(np.float64).reshape(-1).copy()
    lib().or_fft3d(buf, a.shape[0], a.shape[1], a.shape[2], sign)
    return buf.view(np.complex128).reshape(a.shape)


def se3p(x, q, L, M, P, m, xi, force=False, want_grid=False):
    """spectral_ewald.c:20-182 restated (FP64, FGG_SPLIT path)."""
    N = len(q)
    out = np.zeros(3 * N if force else N)
    grid_out = np.zeros(M ** 3) if want_grid else None
    lib().or_se3p(N, _f64(x), _f64(q), L, M, P, m, xi, int(force), out,
                  grid_out.ctypes.data if want_grid else None)
    return (out, grid_out) if want_grid else out


def direct3p(x, q, box, xi, layers, force=False, n_targets=None, pairwise=False):
    """SE3P_direct_fd / SE3P_direct_force_fd (SE_direct.c:352-461)."""
    q = _f64(q)
    N = len(q)
    box = _f64(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
    if pairwise:
        out = np.zeros(3 * N if force else N)
        fn = lib().or_direct3p_force_pairwise if force else lib().or_direct3p_fd_pairwise
        fn(out, _f64(x), q, N, box, xi, layers)
        return out
    Nt = N if n_targets is None else n_targets
    out = np.zeros(3 * Nt if force else Nt)
    lib().or_direct3p_sf(out, int(force), _f64(x), q, N, Nt, box, xi, layers)
    return out


def direct2p(x, q, box, xi, layers, n_targets=None, with_k0=True):
    """SE2P_direct_fd (+ SE2P_direct_k0), SE_direct.c:261-338."""
    q = _f64(q)
    N = len(q)
    Nt = N if n_targets is None else n_targets
    box = _f64(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
    phi = np.zeros(Nt)
    lib().or_direct2p_fd(phi, _f64(x), q, N, Nt, box, xi, layers)
    if with_k0:
        k0 = np.zeros(Nt)
        lib().or_direct2p_k0(k0, _f64(x), q, N, Nt, box, xi)
        phi = phi + k0
    return phi


def direct1p(x, q, box, xi, layers, n_targets=None, with_k0=True):
    """SE1P_direct_fd (+ SE1P_direct_k0), SE_direct.c:90-176: the SE1P k-space sum,
    periodic in z with period box[2], free in x and y."""
    q = _f64(q)
    N = len(q)
    Nt = N if n_targets is None else n_targets
    box = _f64(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
    phi = np.zeros(Nt)
    lib().or_direct1p_fd(phi, _f64(x), q, N, Nt, box, xi, layers)
    if with_k0:
        k0 = np.zeros(Nt)
        lib().or_direct1p_k0(k0, _f64(x), q, N, Nt, box, xi)
        phi = phi + k0
    return phi


def direct_real(per, x, q, box, xi, layers, rc=0.0, n_targets=None):
    """Real-space Ewald sum over the images of the `per` periodic directions
    (SE3P_direct_real[_rc], SE2P_direct_real[_rc], SE1P_direct_real[_rc])."""
    q = _f64(q)
    N = len(q)
    Nt = N if n_targets is None else n_targets
    box = _f64(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
    phi = np.zeros(Nt)
    lib().or_direct_real(per, phi, _f64(x), q, N, Nt, box, xi, layers, rc)
    return phi


def direct_self(q, xi):
    """-2 xi/sqrt(pi) q (SE3P/SE2P/SE1P_direct_self)."""
    q = _f64(q)
    phi = np.zeros(len(q))
    lib().or_direct_self(phi, q, len(q), xi)
    return phi


def k0inc(a, b):
    """Incomplete modified Bessel function int_0^1 t^-1 exp(-b t - a/t) dt."""
    return lib().or_k0inc(a, b)


def ein(x):
    """Ein(x) = gamma + ln x + E1(x)."""
    return lib().or_ein(x)


def energy(phi, q):
    return lib().or_energy(_f64(phi), _f64(q), len(q))


def init_params(P, L, xi, q, round4=True):
    """SE_general.c:517-572 -> (M, m, Q)."""
    M = C.c_int(0)
    m = C.c_double(0)
    Q = lib().or_init_params(C.byref(M), P, C.byref(m), L, xi, _f64(q), len(q), int(round4))
    return M.value, m.value, Q


# ----------------------------------------------------------------------------------
# The reference's own compiled gridding core (oracle/_ref), when it was built.
# ----------------------------------------------------------------------------------
_ref = {}


def ref_available(per=P3):
    return os.path.exists(os.path.join(REF_DIR, "libsefgg_ref_%s.so" % _PER_TAG[per]))


def ref_lib(per=P3):
    """ctypes handle of the reference SE_fgg.c built for one periodicity.

    The 2P build leaves the force expansion symbol undefined (SE_fgg.h:24-27), so it
    is opened with lazy binding and its force routines are never called."""
    if per not in _ref:
        path = os.path.join(REF_DIR, "libsefgg_ref_%s.so" % _PER_TAG[per])
        if not os.path.exists(path):
            raise RuntimeError("reference core not built (make -C oracle ref): %s" % path)
        L = C.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        pp = C.POINTER(SE_FGG_params)
        wp = C.POINTER(SE_FGG_work)
        sp = C.POINTER(SE_state)
        for name in ("SE_FGG_expand_all", "SE_FGG_expand_all_SSE_force", "SE_FGG_grid",
                     "SE_FGG_grid_split", "SE_FGG_grid_split_AVX_dispatch",
                     "SE_FGG_grid_split_SSE_dispatch"):
            getattr(L, name).argtypes = [wp, sp, pp]
        for name in ("SE_FGG_int_split", "SE_FGG_int_split_AVX_dispatch",
                     "SE_FGG_int_split_SSE_dispatch"):
            getattr(L, name).argtypes = [_dp, wp, pp]
        L.SE_FGG_int.argtypes = [_dp, wp, sp, pp]
        L.SE_FGG_int_force.argtypes = [_dp, wp, sp, pp]
        for name in ("SE_FGG_int_split_force", "SE_FGG_int_split_AVX_dispatch_force",
                     "SE_FGG_int_split_SSE_dispatch_force"):
            getattr(L, name).argtypes = [_dp, sp, wp, pp]
        L.SE_FGG_base_gaussian.argtypes = [wp, pp]
        for name in ("SE_FGG_wrap_fcn", "SE2P_FGG_wrap_fcn", "SE1P_FGG_wrap_fcn"):
            getattr(L, name).argtypes = [_dp, wp, pp]
        for name in ("SE_FGG_extend_fcn", "SE2P_FGG_extend_fcn", "SE1P_FGG_extend_fcn"):
            getattr(L, name).argtypes = [wp, _dp, pp]
        L.SE_FGG_pack_params.argtypes = [pp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_double, C.c_double]
        L.SE2P_FGG_pack_params.argtypes = [pp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_double, C.c_double]
        L.SE1P_FGG_pack_params.argtypes = [pp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_double, C.c_double, C.c_double]
        _ref[per] = L
    return _ref[per]


def aligned_zeros(n, dtype=np.float64, align=64):
    """Zeroed 1-D array whose data pointer is ``align``-byte aligned (the reference's
    AVX kernels use aligned loads, as its _mm_malloc(.., 32) buffers allow)."""
    dt = np.dtype(dtype)
    raw = np.zeros(int(n) * dt.itemsize + align, dtype=np.uint8)
    off = (-raw.ctypes.data) % align
    return raw[off:off + int(n) * dt.itemsize].view(dt)


def aligned_copy(a, dtype=np.float64):
    a = np.asarray(a, dtype=dtype).reshape(-1)
    b = aligned_zeros(a.size, dtype)
    b[:] = a
    return b


class RefWork:
    """Aligned numpy arrays wired into an SE_FGG_work for calls into the reference."""

    def __init__(self, p, force=False):
        N, P = p.N, p.P
        self.H = aligned_zeros(int(np.prod(p.npdims)))
        self.zs = aligned_zeros(P ** 3)
        self.zx, self.zy, self.zz = (aligned_zeros(N * P) for _ in range(3))
        self.idx = aligned_zeros(N, np.int32)
        self.zf = [aligned_zeros(N * P) for _ in range(3)] if force else None
        w = SE_FGG_work()
        dp = C.POINTER(C.c_double)
        w.H = self.H.ctypes.data_as(dp)
        w.zs = self.zs.ctypes.data_as(dp)
        w.zx, w.zy, w.zz = (a.ctypes.data_as(dp) for a in (self.zx, self.zy, self.zz))
        w.idx = self.idx.ctypes.data_as(C.POINTER(C.c_int))
        if force:
            w.zfx, w.zfy, w.zfz = (a.ctypes.data_as(dp) for a in self.zf)
        w.free_zs = 0
        w.free_fgg_expa = 0
        self.w = w


def ref_state(x, q):
    x = aligned_copy(x)
    q = aligned_copy(q)
    phi = aligned_zeros(len(q))
    s = SE_state()
    dp = C.POINTER(C.c_double)
    s.x, s.q, s.phi = x.ctypes.data_as(dp), q.ctypes.data_as(dp), phi.ctypes.data_as(dp)
    return s, (x, q, phi)

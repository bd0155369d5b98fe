"""MI355X-native Spectral Ewald fast Gaussian gridding.

Python mirror of the reference's C interface (davoudss/SE_fast_gridding:
SE_fgg.h, SE_general.h, spectral_ewald.c) over the C ABI of
``se_fast_gridding_amd/lib/libse_fgg_amd.so`` (hand-written gfx950 HIP kernels).

* ``capi()``           -- the ctypes handle with argument types for every symbol in include/*.h
* ``SE_FGG_params``, ``SE_FGG_work``, ``SE_state``, ``SE_opt`` -- byte layouts of the reference structs
* ``Plan``             -- the device-resident engine (include/se_gpu.h) over torch CUDA tensors

There is no CPU fallback: if the shared library is missing the import of ``capi()``
raises, and every compute entry point runs on the GPU.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libse_fgg_amd.so")
ROOT = os.path.dirname(HERE)
INCLUDE_DIR = os.path.join(ROOT, "include")

SE_OK, SE_ERR_ARG, SE_ERR_HIP, SE_ERR_FFT, SE_ERR_CAPACITY, SE_ERR_RANGE = 0, -1, -2, -3, -4, -5
SE_BUSY = 1
PREC_FP64, PREC_FP32 = 0, 1
Q_POTENTIAL, Q_FORCE = 1, 2


class SE_FGG_params(C.Structure):
    """SE_fgg.h:121-134."""
    _fields_ = [("N", C.c_int), ("P", C.c_int), ("P_half", C.c_int),
                ("dims", C.c_int * 3), ("npdims", C.c_int * 3),
                ("c", C.c_double), ("d", C.c_double), ("h", C.c_double),
                ("a", C.c_double), ("b", C.c_double)]


class SE_FGG_work(C.Structure):
    """SE_fgg.h:100-118."""
    _fields_ = [("H", C.POINTER(C.c_double)), ("zs", C.POINTER(C.c_double)),
                ("zx", C.POINTER(C.c_double)), ("zy", C.POINTER(C.c_double)),
                ("zz", C.POINTER(C.c_double)), ("idx", C.POINTER(C.c_int)),
                ("zfx", C.POINTER(C.c_double)), ("zfy", C.POINTER(C.c_double)),
                ("zfz", C.POINTER(C.c_double)),
                ("free_zs", C.c_int), ("free_fgg_expa", C.c_int)]


class SE_state(C.Structure):
    """SE_fgg.h:142-148."""
    _fields_ = [("x", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
                ("phi", C.POINTER(C.c_double))]


class SE_opt(C.Structure):
    """SE_general.h:20-24."""
    _fields_ = [("m", C.c_double), ("c", C.c_double), ("box", C.c_double * 3),
                ("h", C.c_double), ("xi", C.c_double), ("w", C.c_double),
                ("eta", C.c_double), ("P", C.c_int), ("M", C.c_int), ("N", C.c_int)]


class se_plan_desc(C.Structure):
    """se_gpu.h se_plan_desc."""
    _fields_ = [("periodicity", C.c_int), ("M", C.c_int), ("P", C.c_int),
                ("box", C.c_double), ("xi", C.c_double), ("m", C.c_double),
                ("precision", C.c_int), ("n_max", C.c_int), ("tile", C.c_int),
                ("chunk", C.c_int), ("block", C.c_int), ("flags", C.c_int),
                ("slab", C.c_double), ("Mz", C.c_int)]

PLAN_GENERIC_KERNELS = 1
PLAN_NO_MFMA = 2
PLAN_NO_PLANE_SPREAD = 4
PLAN_NO_COLUMN = 8


class se_job_shape(C.Structure):
    """se_gpu.h se_job_shape."""
    _fields_ = [("periodicity", C.c_int), ("M", C.c_int), ("P", C.c_int), ("box", C.c_double),
                ("xi", C.c_double), ("m", C.c_double), ("slab", C.c_double), ("precision", C.c_int)]


class se_job(C.Structure):
    """se_gpu.h se_job."""
    _fields_ = [("shape", se_job_shape), ("N", C.c_int), ("d_x", C.c_void_p), ("d_q", C.c_void_p),
                ("d_phi", C.c_void_p), ("d_force", C.c_void_p)]


class se_system(C.Structure):
    """se_gpu.h se_system."""
    _fields_ = [("N", C.c_int), ("d_x", C.c_void_p), ("d_q", C.c_void_p), ("d_phi", C.c_void_p),
                ("d_force", C.c_void_p)]


class se_real_desc(C.Structure):
    """se_gpu.h se_real_desc."""
    _fields_ = [("periodicity", C.c_int), ("box", C.c_double * 3), ("xi", C.c_double), ("rc", C.c_double),
                ("self", C.c_int), ("accumulate", C.c_int)]


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_PP = C.POINTER(SE_FGG_params)
_WP = C.POINTER(SE_FGG_work)
_SP = C.POINTER(SE_state)
_VP = C.c_void_p

# name -> (restype, argtypes) for every exported symbol declared in include/*.h
_SIG = {
    # se_fgg.h
    "se_fgg_set_periodicity": (None, [C.c_int]),
    "se_fgg_get_periodicity": (C.c_int, []),
    "SE_FGG_pack_params": (None, [_PP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
    "SE2P_FGG_pack_params": (None, [_PP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double]),
    "SE1P_FGG_pack_params": (None, [_PP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double]),
    "SE_FGG_allocate_workspace": (None, [_WP, _PP, C.c_int, C.c_int]),
    "SE_FGG_allocate_workspace_SSE_force": (None, [_WP, _PP, C.c_int, C.c_int]),
    "SE_FGG_allocate_grid": (_D, [_PP]),
    "SE_FGG_allocate_vec": (_D, [C.c_int]),
    "SE_FGG_free_workspace": (None, [_WP]),
    "SE_FGG_free_workspace_SSE_force": (None, [_WP]),
    "SE_FGG_grid": (None, [_WP, _SP, _PP]),
    "SE_FGG_expand_all": (None, [_WP, _SP, _PP]),
    "SE_FGG_expand_all_SSE_force": (None, [_WP, _SP, _PP]),
    "SE_FGG_int": (None, [_D, _WP, _SP, _PP]),
    "SE_FGG_int_force": (None, [_D, _WP, _SP, _PP]),
    "SE_FGG_base_gaussian": (None, [_WP, _PP]),
    "SE_FGG_wrap_fcn": (None, [_D, _WP, _PP]),
    "SE2P_FGG_wrap_fcn": (None, [_D, _WP, _PP]),
    "SE1P_FGG_wrap_fcn": (None, [_D, _WP, _PP]),
    "SE_FGG_extend_fcn": (None, [_WP, _D, _PP]),
    "SE2P_FGG_extend_fcn": (None, [_WP, _D, _PP]),
    "SE1P_FGG_extend_fcn": (None, [_WP, _D, _PP]),
    "SE_init_system": (None, [_SP, _PP]),
    "SE_init_unit_system": (None, [_SP, _PP]),
    "SE_free_system": (None, [_SP]),
    "SE_gettime": (C.c_double, []),
    "SE_prod3": (C.c_int, [_I]),
    "SE_fp_set_zero": (None, [_D, C.c_int]),
    "SE_FGG_reorder_system": (None, [_SP, _WP, _PP]),
    "calc_energy": (C.c_double, [SE_state, C.c_int]),
    # se_general.h
    "parse_params": (None, [C.POINTER(SE_opt), C.c_double]),
    "k_vec": (None, [C.c_int, _D, _D, _D, _D]),
    "scaling": (None, [C.c_double, _D, C.c_int, C.c_int, C.c_int, _D]),
    "product_rc": (None, [_D, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int]),
    "SE_FGG_FCN_params": (None, [_PP, C.POINTER(SE_opt), C.c_int]),
    "SE_fgg_int_force": (None, [_D, _D, _D, C.c_int, SE_opt, _D]),
    "copy_r2c": (None, [_D, _D, C.c_int]),
    "copy_c2r": (None, [_D, _D, C.c_int]),
    "SE_init_params": (C.c_double, [_I, _I, _D, C.c_double, C.c_double, C.c_double, _D, C.c_int]),
    "spectral_ewald": (C.c_int, [_D, _D, SE_opt, C.c_double, _D, _D]),
    "se_set_quantity": (None, [C.c_int]),
    "se_set_precision": (None, [C.c_int]),
    # se_gpu.h
    "se_plan_create": (C.c_int, [C.POINTER(_VP), C.POINTER(se_plan_desc), _VP]),
    "se_plan_destroy": (C.c_int, [_VP]),
    "se_plan_set_stream": (C.c_int, [_VP, _VP]),
    "se_plan_bytes": (C.c_size_t, [_VP]),
    "se_plan_bytes_for": (C.c_size_t, [C.POINTER(se_plan_desc)]),
    "se_plan_create_group": (C.c_int, [C.POINTER(_VP), C.POINTER(se_plan_desc), C.c_int, _VP]),
    "se_plan_bytes_for_group": (C.c_size_t, [C.POINTER(se_plan_desc), C.c_int]),
    "se_solve_group": (C.c_int, [_VP, C.c_int, C.POINTER(se_system)]),
    "se_solve": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP, _VP]),
    "se_solve_graph": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP, _VP]),
    "se_spread": (C.c_int, [_VP, C.c_int, _VP, _VP]),
    "se_kspace": (C.c_int, [_VP]),
    "se_interp": (C.c_int, [_VP, _VP, _VP]),
    "se_plan_grid": (C.c_int, [_VP, C.POINTER(_VP), _I]),
    "se_plan_grid_shape": (C.c_int, [_VP, _I, _D]),
    "se_plan_enable_timing": (C.c_int, [_VP, C.c_int]),
    "se_plan_timings": (C.c_int, [_VP, C.POINTER(C.c_float)]),
    "se_plan_info": (C.c_int, [_VP, _I]),
    "se_energy": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    "se_gpu_last_error": (C.c_char_p, []),
    "se_batch_create": (C.c_int, [C.POINTER(_VP), C.c_int, C.c_size_t]),
    "se_batch_destroy": (C.c_int, [_VP]),
    "se_batch_submit": (C.c_int, [_VP, C.POINTER(se_job), _I]),
    "se_batch_run": (C.c_int, [_VP, C.POINTER(se_job), C.c_int, _I]),
    "se_batch_try_submit": (C.c_int, [_VP, C.POINTER(se_job), _I]),
    "se_batch_prepare": (C.c_int, [_VP, C.POINTER(se_job_shape), C.c_int]),
    "se_batch_poll": (C.c_int, [_VP, C.c_int]),
    "se_batch_wait": (C.c_int, [_VP, C.c_int]),
    "se_batch_wait_all": (C.c_int, [_VP]),
    "se_batch_stats": (C.c_int, [_VP, C.POINTER(C.c_size_t), _I]),
    "se_real_plan_create": (C.c_int, [C.POINTER(_VP), C.POINTER(se_real_desc), C.c_int, _VP]),
    "se_real_plan_destroy": (C.c_int, [_VP]),
    "se_real_solve": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP, _VP]),
    "se_real_solve_noalloc": (C.c_int, [C.POINTER(se_real_desc), C.c_int, _VP, _VP, _VP, _VP, _VP]),
    "se_tune": (C.c_int, [C.POINTER(se_plan_desc), C.c_int, C.c_int, C.c_int, _VP, _I, C.POINTER(C.c_float)]),
    "se_tune_save": (C.c_int, [C.c_char_p]),
    "se_tune_load": (C.c_int, [C.c_char_p]),
    "se_tune_clear": (C.c_int, []),
    "se_tune_lookup": (C.c_int, [C.POINTER(se_plan_desc), C.c_int, _I]),
    "se_select_params": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _I, _I,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
}

_SPREAD_SPLIT = ["SE_FGG_grid_split_SSE_dispatch", "SE_FGG_grid_split", "SE_FGG_grid_split_SSE",
                 "SE_FGG_grid_split_SSE_u8", "SE_FGG_grid_split_SSE_P16", "SE_FGG_grid_split_AVX_dispatch",
                 "SE_FGG_grid_split_AVX", "SE_FGG_grid_split_AVX_P16", "SE_FGG_grid_split_AVX_P8",
                 "SE_FGG_grid_split_AVX_u8"]
_SPREAD_SPLIT += [n.replace("SE_FGG_grid_split", "SE_FGG_grid_split") + "_force" for n in
                  ["SE_FGG_grid_split_SSE_dispatch", "SE_FGG_grid_split", "SE_FGG_grid_split_SSE",
                   "SE_FGG_grid_split_SSE_u8", "SE_FGG_grid_split_SSE_P16", "SE_FGG_grid_split_AVX_dispatch",
                   "SE_FGG_grid_split_AVX_P16", "SE_FGG_grid_split_AVX_P8", "SE_FGG_grid_split_AVX_u8",
                   "SE_FGG_grid_split_AVX"]]
_INT_SPLIT = ["SE_FGG_int_split_SSE_dispatch", "SE_FGG_int_split", "SE_FGG_int_split_SSE",
              "SE_FGG_int_split_SSE_u8", "SE_FGG_int_split_SSE_P8", "SE_FGG_int_split_SSE_P16",
              "SE_FGG_int_split_AVX_dispatch", "SE_FGG_int_split_AVX", "SE_FGG_int_split_AVX_P8",
              "SE_FGG_int_split_AVX_P16", "SE_FGG_int_split_AVX_u8"]
_INT_SPLIT_FORCE = ["SE_FGG_int_split_SSE_dispatch_force", "SE_FGG_int_split_force",
                    "SE_FGG_int_split_SSE_force", "SE_FGG_int_split_SSE_P8_force",
                    "SE_FGG_int_split_SSE_u8_force", "SE_FGG_int_split_SSE_P16_force",
                    "SE_FGG_int_split_AVX_dispatch_force", "SE_FGG_int_split_AVX_P16_force",
                    "SE_FGG_int_split_AVX_P8_force", "SE_FGG_int_split_AVX_u8_force",
                    "SE_FGG_int_split_AVX_force"]
for _n in _SPREAD_SPLIT:
    _SIG[_n] = (None, [_WP, _SP, _PP])
for _n in _INT_SPLIT:
    _SIG[_n] = (None, [_D, _WP, _PP])
for _n in _INT_SPLIT_FORCE:
    _SIG[_n] = (None, [_D, _SP, _WP, _PP])

EXPORTS = sorted(_SIG)

_capi = None


def capi():
    """Load libse_fgg_amd.so (built by ``__graft_entry__.build()``); raises if it is missing."""
    global _capi
    if _capi is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libse_fgg_amd.so is not built (%s): run __graft_entry__.build()" % LIB_PATH)
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).  Loading
        # torch first makes this library bind to that copy, so the process has one HIP
        # runtime; loaded the other way round, torch's later device initialisation finds
        # no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _capi = lib
    return _capi


def last_error():
    msg = capi().se_gpu_last_error()
    return msg.decode() if msg else ""


def check(rc, what="se_gpu call"):
    if rc != SE_OK:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, last_error()))


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _check(t, n, name, where):
    """Refuse, before any C call, an array the kernels would misread: the engine reads and
    writes raw device pointers of float64 SoA arrays of a fixed length, so a CPU tensor,
    another dtype, a strided view or a wrong length would fault or be silently misread."""
    if t is None:
        return
    if not hasattr(t, "data_ptr") or not hasattr(t, "is_cuda"):
        raise ValueError("%s: %s must be a torch tensor, got %s" % (where, name, type(t).__name__))
    if not t.is_cuda:
        raise ValueError("%s: %s must be a CUDA (HIP device) tensor" % (where, name))
    import torch
    if t.dtype != torch.float64:
        raise ValueError("%s: %s must be float64, got %s" % (where, name, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s: %s must be contiguous" % (where, name))
    if t.numel() != n:
        raise ValueError("%s: %s must hold %d values, got %d" % (where, name, n, t.numel()))


def _check_particles(where, x, q, phi=None, force=None):
    """Checks the SoA particle arrays of one system (x 3N, q N, phi N, force 3N) and returns N."""
    if q is None or not hasattr(q, "numel"):
        raise ValueError("%s: q must be a torch tensor" % where)
    N = q.numel()
    _check(q, N, "q", where)
    _check(x, 3 * N, "x", where)
    _check(phi, N, "phi", where)
    _check(force, 3 * N, "force", where)
    return N


class Plan:
    """Device-resident SE3P / SE2P / SE1P solver (include/se_gpu.h) driven with torch CUDA tensors.

    x is (3N,) float64 SoA [x..., y..., z...], q is (N,) float64, outputs are float64,
    force SoA (3N,) in the reference's (q/2) grad(phi) convention (SE_fgg.c:2291-2307).
    periodicity=2 solves SE2P: periodic in x and y (period box), free in z with particles
    in 0 <= z < slab; Mz (0 = automatic) is the zero-padded z grid.  periodicity=1 solves
    SE1P: periodic in z (period box, M grid points), free in x and y with particles in
    [0, slab)^2; Mz (0 = automatic) is then the zero-padded x-y grid side.  max_systems > 1
    makes a group plan (n_max then counts the particles of all systems together) for
    solve_group."""

    def __init__(self, M, P, box, xi, n_max, precision=PREC_FP64, m=0.0, tile=0, chunk=0, block=0,
                 stream=None, generic=False, periodicity=3, slab=0.0, Mz=0, max_systems=1, mfma=True, plane=True,
                 column=True):
        import torch
        self._torch = torch
        d = se_plan_desc(periodicity, M, P, float(box), float(xi), float(m), precision, int(n_max), tile, chunk,
                         block, (PLAN_GENERIC_KERNELS if generic else 0) | (0 if mfma else PLAN_NO_MFMA)
                         | (0 if plane else PLAN_NO_PLANE_SPREAD) | (0 if column else PLAN_NO_COLUMN),
                         float(slab), int(Mz))
        self.desc = d
        self.max_systems = int(max_systems)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = C.c_void_p()
        if self.max_systems > 1:
            check(capi().se_plan_create_group(C.byref(h), C.byref(d), self.max_systems,
                                              C.c_void_p(self.stream.cuda_stream)), "se_plan_create_group")
        else:
            check(capi().se_plan_create(C.byref(h), C.byref(d), C.c_void_p(self.stream.cuda_stream)),
                  "se_plan_create")
        self.h = h
        self.M, self.P, self.n_max, self.precision = M, P, n_max, precision
        dims = (C.c_int * 3)()
        z0 = C.c_double()
        check(capi().se_plan_grid_shape(h, dims, C.byref(z0)), "se_plan_grid_shape")
        self.shape = (dims[0], dims[1], dims[2])
        self.z0 = z0.value

    def close(self):
        if getattr(self, "h", None):
            capi().se_plan_destroy(self.h)
            self.h = None

    __del__ = close

    def solve(self, x, q, phi=None, force=None):
        N = _check_particles("Plan.solve", x, q, phi, force)
        check(capi().se_solve(self.h, N, _ptr(x), _ptr(q), _ptr(phi), _ptr(force)), "se_solve")
        self._last_N = N

    def solve_group(self, systems):
        """Grouped solve of [(x, q, phi, force), ...] (phi / force may be None, the same
        for every system) in one pass (se_solve_group)."""
        arr = (se_system * len(systems))()
        for i, (x, q, phi, force) in enumerate(systems):
            N = _check_particles("Plan.solve_group (system %d)" % i, x, q, phi, force)
            arr[i] = se_system(N, x.data_ptr(), q.data_ptr(), phi.data_ptr() if phi is not None else None,
                               force.data_ptr() if force is not None else None)
        check(capi().se_solve_group(self.h, len(systems), arr), "se_solve_group")
        self._last_N = None

    def solve_graph(self, x, q, phi=None, force=None):
        """se_solve through a captured HIP graph (captured on the first call with these tensors)."""
        N = _check_particles("Plan.solve_graph", x, q, phi, force)
        check(capi().se_solve_graph(self.h, N, _ptr(x), _ptr(q), _ptr(phi), _ptr(force)), "se_solve_graph")
        self._last_N = N

    def spread(self, x, q):
        N = _check_particles("Plan.spread", x, q)
        check(capi().se_spread(self.h, N, _ptr(x), _ptr(q)), "se_spread")
        self._last_N = N

    def kspace(self):
        check(capi().se_kspace(self.h), "se_kspace")

    def interp(self, phi=None, force=None):
        N = getattr(self, "_last_N", None)
        if N is not None:
            _check(phi, N, "phi", "Plan.interp")
            _check(force, 3 * N, "force", "Plan.interp")
        check(capi().se_interp(self.h, _ptr(phi), _ptr(force)), "se_interp")

    def _grid_ptr(self):
        p = C.c_void_p()
        es = C.c_int()
        check(capi().se_plan_grid(self.h, C.byref(p), C.byref(es)), "se_plan_grid")
        dt = self._torch.float32 if es.value == 4 else self._torch.float64
        return p.value, dt, es.value

    def grid(self):
        """Copy of the plan's grid, (M, M, Mz) row-major (z fastest)."""
        torch = self._torch
        ptr, dt, es = self._grid_ptr()
        out = torch.empty(self.shape, dtype=dt, device="cuda")
        torch.cuda.synchronize()
        _hip_copy(out.data_ptr(), ptr, out.numel() * es)
        return out

    def set_grid(self, g):
        """Overwrite the plan's periodic grid (e.g. to interpolate a given field)."""
        torch = self._torch
        ptr, dt, es = self._grid_ptr()
        src = g.to(dtype=dt).contiguous()
        assert src.numel() == self.shape[0] * self.shape[1] * self.shape[2]
        torch.cuda.synchronize()
        _hip_copy(ptr, src.data_ptr(), src.numel() * es)

    def enable_timing(self, on=True):
        check(capi().se_plan_enable_timing(self.h, int(on)), "se_plan_enable_timing")

    def timings(self):
        ms = (C.c_float * 8)()
        check(capi().se_plan_timings(self.h, ms), "se_plan_timings")
        keys = ["spread_stage", "spread_kernel", "merge", "kspace", "interp_stage", "gather_kernel", "bucketing"]
        return {k: float(ms[i]) for i, k in enumerate(keys)}

    def info(self):
        out = (C.c_int * 8)()
        check(capi().se_plan_info(self.h, out), "se_plan_info")
        keys = ["tile", "chunk", "block", "ntiles", "items", "discarded", "S", "pencil"]
        return {k: int(out[i]) for i, k in enumerate(keys)}

    def nbytes(self):
        return int(capi().se_plan_bytes(self.h))


def tune(M, P, box, xi, N, precision=PREC_FP32, periodicity=3, slab=0.0, m=0.0, potential=True, force=True,
         systems=1, stream=None):
    """se_tune: time every candidate tiling on `systems` synthetic uniform systems of N
    particles (a grouped solve when systems > 1) and record the fastest for plans of this
    geometry and size band; returns (tile, chunk, block, ms)."""
    import torch
    d = se_plan_desc(periodicity, M, P, float(box), float(xi), float(m), precision, int(N), 0, 0, 0, 0,
                     float(slab), 0)
    out = (C.c_int * 3)()
    ms = C.c_float()
    st = stream if stream is not None else torch.cuda.current_stream()
    qty = (1 if potential else 0) | (2 if force else 0)
    check(capi().se_tune(C.byref(d), int(N), int(systems), qty, C.c_void_p(st.cuda_stream), out, C.byref(ms)),
          "se_tune")
    return out[0], out[1], out[2], ms.value


def tune_lookup(M, P, box, xi, N, precision=PREC_FP32, periodicity=3, slab=0.0):
    """The recorded tiling (tile, chunk, block) for this geometry and size, or None."""
    d = se_plan_desc(periodicity, M, P, float(box), float(xi), 0.0, precision, int(N), 0, 0, 0, 0, float(slab), 0)
    out = (C.c_int * 3)()
    return tuple(out) if capi().se_tune_lookup(C.byref(d), int(N), out) == 1 else None


def select_params(tol, L, xi, q, potential=True, force=False):
    """(P, M, m, estimate) for an rms error target (se_gpu.h se_select_params): the
    smallest support P whose estimated gridding error (the estimates SE_init_params
    prints, SE_general.c:525-549) is at most tol, with the reference's grid rule for M.
    q is the charge array (Q = sum q^2) or Q itself as a float."""
    import numpy as np

    Q = float(q) if np.ndim(q) == 0 else float(np.sum(np.asarray(q, dtype=np.float64) ** 2))
    P, M = C.c_int(), C.c_int()
    m, est = C.c_double(), C.c_double()
    qty = (1 if potential else 0) | (2 if force else 0)
    rc = capi().se_select_params(float(tol), float(L), float(xi), Q, qty, C.byref(P), C.byref(M), C.byref(m),
                                 C.byref(est))
    if rc != SE_OK:
        raise ValueError("se_select_params: no support P <= 32 meets tol=%g (or an argument is out of range)" % tol)
    return P.value, M.value, m.value, est.value


class RealSpace:
    """Real-space (+ self) Ewald terms on the device (se_gpu.h se_real_*): phi_m = sum over
    pairs and images within rc of q_n erfc(xi r)/r, force = (q/2) grad phi; periodicity 3,
    2 (x, y periodic) or 1 (z periodic); box = periods / extents of the free directions."""

    def __init__(self, periodicity, box, xi, rc, n_max, self_term=False, accumulate=False, stream=None):
        import torch
        box = [float(b) for b in (box if hasattr(box, "__len__") else (box, box, box))]
        d = se_real_desc(periodicity, (C.c_double * 3)(*box), float(xi), float(rc), int(self_term), int(accumulate))
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = C.c_void_p()
        check(capi().se_real_plan_create(C.byref(h), C.byref(d), int(n_max), C.c_void_p(self.stream.cuda_stream)),
              "se_real_plan_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            capi().se_real_plan_destroy(self.h)
            self.h = None

    __del__ = close

    def solve(self, x, q, phi=None, force=None):
        N = _check_particles("RealSpace.solve", x, q, phi, force)
        check(capi().se_real_solve(self.h, N, _ptr(x), _ptr(q), _ptr(phi), _ptr(force)), "se_real_solve")


def real_solve_noalloc(periodicity, box, xi, rc, x, q, phi=None, force=None, self_term=False, accumulate=False,
                       stream=None):
    """RealSpace(...).solve without a plan and without device allocation
    (se_real_solve_noalloc): the cell lists live in the library's static scratch array
    (SE_REAL_SCRATCH_BYTES per GPU); calls that overlap on one GPU race."""
    import torch
    N = _check_particles("real_solve_noalloc", x, q, phi, force)
    box = [float(b) for b in (box if hasattr(box, "__len__") else (box, box, box))]
    d = se_real_desc(periodicity, (C.c_double * 3)(*box), float(xi), float(rc), int(self_term), int(accumulate))
    st = stream if stream is not None else torch.cuda.current_stream()
    return capi().se_real_solve_noalloc(C.byref(d), N, _ptr(x), _ptr(q), _ptr(phi), _ptr(force),
                                        C.c_void_p(st.cuda_stream))


_hip = None


def _hip_copy(dst, src, nbytes):
    """Synchronous device-to-device copy through the HIP runtime the library links."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    rc = _hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, 3)   # hipMemcpyDeviceToDevice
    if rc != 0:
        raise RuntimeError("hipMemcpy failed (%d)" % rc)


class Batch:
    """Executor for many independent systems (include/se_gpu.h se_batch_*): `slots` HIP
    streams with per-slot plan caches under an HBM budget.  Jobs take torch CUDA tensors
    in the Plan layout; they must stay alive until their ticket completes."""

    def __init__(self, slots=8, hbm_budget=0):
        h = C.c_void_p()
        check(capi().se_batch_create(C.byref(h), int(slots), int(hbm_budget)), "se_batch_create")
        self.h = h
        self.slots = slots

    def close(self):
        if getattr(self, "h", None):
            capi().se_batch_destroy(self.h)
            self.h = None

    __del__ = close

    @staticmethod
    def job(M, P, box, xi, x, q, phi=None, force=None, precision=PREC_FP32, m=0.0, periodicity=3, slab=0.0):
        N = _check_particles("Batch.job", x, q, phi, force)
        sh = se_job_shape(periodicity, M, P, float(box), float(xi), float(m), float(slab), precision)
        return se_job(sh, N, x.data_ptr(), q.data_ptr(), phi.data_ptr() if phi is not None else None,
                      force.data_ptr() if force is not None else None)

    def submit(self, job):
        t = C.c_int()
        check(capi().se_batch_submit(self.h, C.byref(job), C.byref(t)), "se_batch_submit")
        return t.value

    def run(self, jobs):
        """Grouped solve of a list of jobs (se_batch_run); returns one ticket."""
        arr = (se_job * len(jobs))(*jobs)
        t = C.c_int()
        check(capi().se_batch_run(self.h, arr, len(jobs), C.byref(t)), "se_batch_run")
        return t.value

    def try_submit(self, job):
        """Ticket, or None when no slot is idle."""
        t = C.c_int()
        rc = capi().se_batch_try_submit(self.h, C.byref(job), C.byref(t))
        if rc == SE_BUSY:
            return None
        check(rc, "se_batch_try_submit")
        return t.value

    def prepare(self, job):
        """Build the plan for this job's shape (capacity at least its N) on every slot."""
        check(capi().se_batch_prepare(self.h, C.byref(job.shape), job.N), "se_batch_prepare")

    def poll(self, ticket):
        rc = capi().se_batch_poll(self.h, int(ticket))
        if rc < 0:
            check(rc, "se_batch_poll")
        return rc == 1

    def wait(self, ticket):
        check(capi().se_batch_wait(self.h, int(ticket)), "se_batch_wait")

    def wait_all(self):
        check(capi().se_batch_wait_all(self.h), "se_batch_wait_all")

    def stats(self):
        b = C.c_size_t()
        n = C.c_int()
        check(capi().se_batch_stats(self.h, C.byref(b), C.byref(n)), "se_batch_stats")
        return {"plan_bytes": int(b.value), "plans": int(n.value)}

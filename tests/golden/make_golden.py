"""Writes the golden vectors under tests/golden/ by running the reference's own
gridding core (SE_fgg.c, compiled by ``make -C oracle ref`` into oracle/_ref).

Run in the build container (it needs /root/reference to have built oracle/_ref):

    python tests/golden/make_golden.py

Each ``g3p_*.npz`` / ``g2p_*.npz`` / ``g1p_*.npz`` holds seeded inputs and the
reference's outputs for: SE_FGG_base_gaussian, SE_FGG_expand_all,
SE_FGG_expand_all_SSE_force, SE_FGG_grid_split_AVX_dispatch (the driver's spread),
the periodicity's wrap and extend, SE_FGG_int_split_AVX_dispatch and
SE_FGG_int_split_AVX_dispatch_force (3P), and SE_FGG_grid / SE_FGG_int (on-the-fly).
``atoms.npz`` is the reference test's input fixture atoms.txt (200 rows x,y,z,q)
parsed as SE_init_unit_system reads it (SE_fgg.c:266-292).
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# (tag, N, P, dims, seed)
CASES_3P = [
    ("p4_m12", 40, 4, (12, 12, 12), 11),
    ("p5_m10", 40, 5, (10, 10, 10), 12),
    ("p8_m16", 60, 8, (16, 16, 16), 13),
    ("p11_m16", 60, 11, (16, 16, 16), 14),
    ("p12_m16", 40, 12, (16, 16, 16), 15),
    ("p16_m16", 80, 16, (16, 16, 16), 16),
    ("p21_m24", 30, 21, (24, 24, 24), 17),
    ("p24_m24", 20, 24, (24, 24, 24), 18),
    ("p16_m20x16x12", 50, 16, (20, 16, 12), 19),
]
CASES_2P = [("p8", 30, 8, (12, 12, 20), 21), ("p7", 30, 7, (12, 12, 20), 22)]
CASES_1P = [("p8", 30, 8, (20, 20, 12), 31), ("p7", 30, 7, (20, 20, 12), 32)]


def gparams(P, M):
    """c from the driver's parameter rule at L = 1, m = 0.95 sqrt(pi P), xi = 0.3125 M."""
    m = 0.95 * np.sqrt(np.pi * P)
    h, w, eta, c = O.parse_params(1.0, M, P, m, 0.3125 * M)
    return c, h


def run_case(per, tag, N, P, dims, seed):
    rng = np.random.default_rng(seed)
    L = O.ref_lib(per)
    c, h = gparams(P, dims[0])
    p = O.SE_FGG_params()
    a = b = -1.0
    if per == O.P3:
        L.SE_FGG_pack_params(C.byref(p), N, dims[0], dims[1], dims[2], P, c, h)
        x = rng.random((3, N)) * np.array(dims)[:, None] * h
    elif per == O.P2:
        a = -(O.half(P) + 0.5) * h
        L.SE2P_FGG_pack_params(C.byref(p), N, dims[0], dims[1], dims[2], P, c, h, a)
        x = rng.random((3, N))
        x[0] *= dims[0] * h
        x[1] *= dims[1] * h
    else:
        a = b = -(O.half(P) + 0.5) * h
        L.SE1P_FGG_pack_params(C.byref(p), N, dims[0], dims[1], dims[2], P, c, h, a, b)
        x = rng.random((3, N))
        x[2] *= dims[2] * h
    x = np.ascontiguousarray(x.reshape(-1))
    q = rng.uniform(-1, 1, N)
    q -= q.mean()

    out = dict(per=per, N=N, P=P, dims=np.array(dims), c=c, h=h, a=a, b=b, x=x, q=q)
    st, keep = O.ref_state(x, q)

    # expansion (potential) + base gaussian + driver spread
    w = O.RefWork(p)
    L.SE_FGG_base_gaussian(C.byref(w.w), C.byref(p))
    L.SE_FGG_expand_all(C.byref(w.w), C.byref(st), C.byref(p))
    ne = int(np.prod(p.npdims))
    assert w.idx.min() >= 0 and w.idx.max() + ((P - 1) * p.npdims[1] + (P - 1)) * p.npdims[2] + P - 1 < ne
    out.update(zs=w.zs.copy(), zx=w.zx.copy(), zy=w.zy.copy(), zz=w.zz.copy(), idx=w.idx.copy())
    L.SE_FGG_grid_split_AVX_dispatch(C.byref(w.w), C.byref(st), C.byref(p))
    out["H_split"] = w.H.copy()

    # periodic-grid extent as the reference indexes it (O.per_grid_size)
    nper = O.per_grid_size(per, p)
    Hper = O.aligned_zeros(nper)
    wrapfn = {O.P3: L.SE_FGG_wrap_fcn, O.P2: L.SE2P_FGG_wrap_fcn, O.P1: L.SE1P_FGG_wrap_fcn}[per]
    wrapfn(Hper, C.byref(w.w), C.byref(p))
    out["Hper"] = Hper

    # extend a random periodic grid, then interpolate (split, AVX dispatcher)
    G = O.aligned_copy(rng.standard_normal(nper))
    extfn = {O.P3: L.SE_FGG_extend_fcn, O.P2: L.SE2P_FGG_extend_fcn, O.P1: L.SE1P_FGG_extend_fcn}[per]
    extfn(C.byref(w.w), G, C.byref(p))
    out["G"] = G
    out["H_ext"] = w.H.copy()
    phi = O.aligned_zeros(N)
    L.SE_FGG_int_split_AVX_dispatch(phi, C.byref(w.w), C.byref(p))
    out["phi_split"] = phi

    # on-the-fly spread and interpolation (SE_FGG_grid / SE_FGG_int)
    w2 = O.RefWork(p)
    L.SE_FGG_base_gaussian(C.byref(w2.w), C.byref(p))
    L.SE_FGG_grid(C.byref(w2.w), C.byref(st), C.byref(p))
    out["H_grid"] = w2.H.copy()
    w2.H[:] = out["H_ext"]
    phi2 = O.aligned_zeros(N)
    L.SE_FGG_int(phi2, C.byref(w2.w), C.byref(st), C.byref(p))
    out["phi_int"] = phi2

    if per in (O.P3, O.P1):
        # force workspace: expansion with derivative factors, force spread, force interpolation
        wf = O.RefWork(p, force=True)
        L.SE_FGG_base_gaussian(C.byref(wf.w), C.byref(p))
        L.SE_FGG_expand_all_SSE_force(C.byref(wf.w), C.byref(st), C.byref(p))
        out.update(fidx=wf.idx.copy(), fzx=wf.zx.copy(), fzy=wf.zy.copy(), fzz=wf.zz.copy(),
                   zfx=wf.zf[0].copy(), zfy=wf.zf[1].copy(), zfz=wf.zf[2].copy())
        if per == O.P3:
            wf.H[:] = out["H_ext"]
            F = O.aligned_zeros(3 * N)
            L.SE_FGG_int_split_AVX_dispatch_force(F, C.byref(st), C.byref(wf.w), C.byref(p))
            out["F_split"] = F
            F2 = O.aligned_zeros(3 * N)
            L.SE_FGG_int_force(F2, C.byref(wf.w), C.byref(st), C.byref(p))
            out["F_int"] = F2
    return out


def main():
    if not O.ref_available(O.P3):
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref` with /root/reference mounted")
    for per, cases, pre in ((O.P3, CASES_3P, "g3p"), (O.P2, CASES_2P, "g2p"), (O.P1, CASES_1P, "g1p")):
        for tag, N, P, dims, seed in cases:
            d = run_case(per, tag, N, P, dims, seed)
            if np.prod(np.array(dims) + P) > 40000:
                # large extended grids are pinned through Hper / phi / F only (keeps fixtures small)
                for k in ("H_split", "H_grid", "H_ext"):
                    d.pop(k)
            np.savez_compressed(os.path.join(OUT, "%s_%s.npz" % (pre, tag)), **d)
            print("wrote", pre, tag)
    atoms = np.loadtxt("/root/reference/atoms.txt")
    np.savez_compressed(os.path.join(OUT, "atoms.npz"), xyzq=atoms)
    print("wrote atoms", atoms.shape)


if __name__ == "__main__":
    main()

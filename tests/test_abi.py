"""CPU checks of the C ABI: the library loads, exports every function declared in
include/*.h, the struct layouts match the reference's, and the host-side parameter
routines agree with the reference formulas.  No GPU compute is called here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import se_fast_gridding_amd as se

HDRS = [os.path.join(se.INCLUDE_DIR, h) for h in ("se_fgg.h", "se_general.h", "se_gpu.h")]


def declared_functions():
    names = set()
    for h in HDRS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    lib = se.capi()
    decl = declared_functions()
    assert len(decl) > 90
    missing_sig = decl - set(se.EXPORTS)
    assert not missing_sig, missing_sig
    for name in decl:
        assert hasattr(lib, name), name


def _layout_program(header):
    return r"""
#include <stdio.h>
#include <stddef.h>
#include "%s"
int main(void) {
  printf("%%zu %%zu %%zu %%zu\n", sizeof(SE_FGG_params), offsetof(SE_FGG_params, c), offsetof(SE_FGG_params, b), offsetof(SE_FGG_params, npdims));
  printf("%%zu %%zu %%zu\n", sizeof(SE_FGG_work), offsetof(SE_FGG_work, zfz), offsetof(SE_FGG_work, free_fgg_expa));
  printf("%%zu\n", sizeof(SE_state));
  return 0;
}
""" % header


def _compile_run(src, incdir):
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-std=gnu99", "-O1", "-w", "-I", incdir, c, "-o", exe, "-lm"])
        return subprocess.check_output([exe]).decode().split()


def test_struct_layouts_match_ctypes_and_reference():
    ours = _compile_run(_layout_program("se_fgg.h"), se.INCLUDE_DIR)
    ct = [C.sizeof(se.SE_FGG_params), se.SE_FGG_params.c.offset, se.SE_FGG_params.b.offset,
          se.SE_FGG_params.npdims.offset, C.sizeof(se.SE_FGG_work), se.SE_FGG_work.zfz.offset,
          se.SE_FGG_work.free_fgg_expa.offset, C.sizeof(se.SE_state)]
    assert [int(v) for v in ours] == ct
    if os.path.exists("/root/reference/SE_fgg.h"):
        ref = _compile_run(_layout_program("SE_fgg.h"), "/root/reference")
        assert ref == ours


def test_host_parameter_routines():
    lib = se.capi()
    opt = se.SE_opt()
    opt.m = 0.95 * np.sqrt(np.pi * 16)
    opt.box[0] = opt.box[1] = opt.box[2] = 1.3
    opt.P, opt.M, opt.N = 16, 40, 10
    lib.parse_params(C.byref(opt), 7.0)
    h = 1.3 / 40
    w = 16 * h / 2
    eta = (2 * w * 7.0 / opt.m) ** 2
    assert np.isclose(opt.h, h) and np.isclose(opt.w, w) and np.isclose(opt.eta, eta)
    assert np.isclose(opt.c, 2 * 49.0 / eta)
    p = se.SE_FGG_params()
    lib.SE_FGG_FCN_params(C.byref(p), C.byref(opt), 10)
    assert p.P_half == 8 and list(p.npdims) == [56, 56, 56] and np.isclose(p.d, (opt.c / np.pi) ** 1.5)
    k = [np.zeros(6) for _ in range(3)]
    box = np.array([1.0, 2.0, 4.0])
    lib.k_vec(6, box.ctypes.data_as(se._D), *[a.ctypes.data_as(se._D) for a in k])
    f = np.array([0, 1, 2, -3, -2, -1])
    for d in range(3):
        assert np.allclose(k[d], 2 * np.pi / box[d] * f)
    M = C.c_int(0)
    P = C.c_int(21)
    m = C.c_double(0)
    q = np.array([0.4238, -0.8476] * 50)
    Q = lib.SE_init_params(C.byref(M), C.byref(P), C.byref(m), 1.0, 3.1814, 8.0, q.ctypes.data_as(se._D), 100)
    assert M.value == 108 and np.isclose(Q, np.sum(q * q)) and np.isclose(m.value, 0.95 * np.sqrt(21 * np.pi))
    p2 = se.SE_FGG_params()
    lib.SE2P_FGG_pack_params(C.byref(p2), 5, 8, 9, 10, 7, 3.0, 0.1, -0.25)
    assert p2.P_half == 3 and list(p2.npdims) == [15, 16, 17] and p2.a == -0.25 and p2.b == -1
    assert lib.SE_prod3((C.c_int * 3)(2, 3, 4)) == 24


def test_plan_rejects_bad_descriptions_without_gpu_compute():
    # argument validation happens before any device work
    lib = se.capi()
    h = C.c_void_p()
    d = se.se_plan_desc(3, 0, 16, 1.0, 1.0, 0.0, 1, 10, 0, 0, 0, 0, 0.0, 0)
    assert lib.se_plan_create(C.byref(h), C.byref(d), None) == se.SE_ERR_ARG
    d = se.se_plan_desc(4, 32, 16, 1.0, 1.0, 0.0, 1, 10, 0, 0, 0, 0, 1.0, 0)
    assert lib.se_plan_create(C.byref(h), C.byref(d), None) == se.SE_ERR_ARG
    assert "periodicity" in se.last_error()
    d = se.se_plan_desc(1, 32, 16, 1.0, 6.0, 0.0, 1, 10, 0, 0, 0, 0, 0.0, 0)   # SE1P without a square
    assert lib.se_plan_create(C.byref(h), C.byref(d), None) == se.SE_ERR_ARG
    assert "slab" in se.last_error()
    d = se.se_plan_desc(1, 32, 16, 1.0, 6.0, 0.0, 1, 10, 0, 0, 0, 0, 1.0, 64)  # x-y grid too small
    assert lib.se_plan_create(C.byref(h), C.byref(d), None) == se.SE_ERR_ARG
    assert "Mz" in se.last_error()
    d = se.se_plan_desc(2, 32, 16, 1.0, 6.0, 0.0, 1, 10, 0, 0, 0, 0, 0.0, 0)   # SE2P without a slab
    assert lib.se_plan_create(C.byref(h), C.byref(d), None) == se.SE_ERR_ARG
    assert "slab" in se.last_error()
    d = se.se_plan_desc(2, 32, 16, 1.0, 6.0, 0.0, 1, 10, 0, 0, 0, 0, 1.0, 40)  # z grid shorter than the slab
    assert lib.se_plan_create(C.byref(h), C.byref(d), None) == se.SE_ERR_ARG
    assert "Mz" in se.last_error()


def test_plan_desc_layout_matches_header():
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "se_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(se_plan_desc), offsetof(se_plan_desc, box), offsetof(se_plan_desc, flags),
         offsetof(se_plan_desc, slab), offsetof(se_plan_desc, Mz));
  return 0;
}
"""
    got = _compile_run(prog, se.INCLUDE_DIR)
    D = se.se_plan_desc
    assert [int(v) for v in got] == [C.sizeof(D), D.box.offset, D.flags.offset, D.slab.offset, D.Mz.offset]


def test_tuning_table_save_load_lookup(tmp_path):
    """The autotuner's table: a saved table loads back, lookups take the geometry's entry at
    the nearest size band, other geometries have none, and clear empties it (host only)."""
    lib = se.capi()
    lib.se_tune_clear()
    path = tmp_path / "tune.txt"
    path.write_text("# periodicity M P fp32 log2N tile chunk block ms\n"
                    "3 64 16 1 14 8 1024 256 0.5\n"
                    "3 64 16 1 20 12 4096 0 2.0\n"
                    "2 48 11 1 15 4 512 512 0.3\n")
    assert lib.se_tune_load(str(path).encode()) == 3
    assert se.tune_lookup(64, 16, 1.0, 20.0, 16000) == (8, 1024, 256)       # band 14
    assert se.tune_lookup(64, 16, 1.0, 20.0, 900000) == (12, 4096, 0)       # band 20 (0: kernel defaults)
    assert se.tune_lookup(64, 16, 1.0, 20.0, 60000) == (8, 1024, 256)       # band 16: nearer 14
    assert se.tune_lookup(48, 11, 1.0, 15.0, 30000, periodicity=2, slab=1.0) == (4, 512, 512)
    assert se.tune_lookup(64, 16, 1.0, 20.0, 16000, precision=se.PREC_FP64) is None
    out = tmp_path / "out.txt"
    assert lib.se_tune_save(str(out).encode()) == 0
    lib.se_tune_clear()
    assert se.tune_lookup(64, 16, 1.0, 20.0, 16000) is None
    assert lib.se_tune_load(str(out).encode()) == 3
    assert se.tune_lookup(64, 16, 1.0, 20.0, 16000) == (8, 1024, 256)
    lib.se_tune_clear()

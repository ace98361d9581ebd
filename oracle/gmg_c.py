"""ctypes front of oracle/gmg3.c, the C restatement of the NumPy oracle's 3D
cycle.  TEST INFRASTRUCTURE ONLY (tests/ and bench.py's cpu_baseline leg).

build() compiles libgmg3.so next to this file with gcc -fopenmp
-ffp-contract=off; cycles3d() runs K cycles of a gmg.Operator on the CPU with
the given OpenMP thread count and returns the new phi.  Results are
bit-identical to gmg.cycle for any thread count (tests/test_oracle_c.py).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "gmg3.c")
LIB = os.path.join(HERE, "libgmg3.so")
_lib = None


def build():
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O3", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           "-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("building the C oracle failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        P = ctypes.POINTER(ctypes.c_void_p)
        lib.gmg3_set_threads.restype = None
        lib.gmg3_set_threads.argtypes = [ctypes.c_int]
        lib.gmg3_cycles.restype = ctypes.c_int
        lib.gmg3_cycles.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), P, P, P, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
        _lib = lib
    return _lib


def _arr(v, shape):
    """A coefficient as a C-contiguous float64 array (scalars expanded)."""
    if np.ndim(v) == 0:
        return np.full(shape, float(v))
    return np.ascontiguousarray(v, dtype=np.float64)


def cycles3d(op, cfg, phi, rhs, ncycles, threads=1):
    """ncycles cycles of configuration cfg (gmg.Config) from phi; op from
    gmg.build_operator(3, ...).  Real FP64 operators only."""
    from oracle import gmg
    assert op.dim == 3
    lib = load()
    L = op.levels
    keep = []
    ms = (ctypes.c_int * (3 * L))()
    am = (ctypes.c_void_p * (3 * L))()
    ap = (ctypes.c_void_p * (3 * L))()
    ac = (ctypes.c_void_p * L)()
    for l, lv in enumerate(op.lv):
        m = [s - 3 for s in lv.shape]
        cshape = tuple(x + 1 for x in m)
        for a in range(3):
            ms[3 * l + a] = m[a]
            lo = _arr(lv.a_minus[a] if lv.spatial else lv.a_axis[a], cshape)
            hi = _arr(lv.a_plus[a] if lv.spatial else lv.a_axis[a], cshape)
            keep += [lo, hi]
            am[3 * l + a] = lo.ctypes.data
            ap[3 * l + a] = hi.ctypes.data
        c = _arr(lv.a_c, cshape)
        keep.append(c)
        ac[l] = c.ctypes.data
    out = np.array(phi, dtype=np.float64, order="C", copy=True)
    r = np.ascontiguousarray(rhs, dtype=np.float64)
    lib.gmg3_set_threads(int(threads))
    rc = lib.gmg3_cycles(L, ms, am, ap, ac, out.ctypes.data, r.ctypes.data, int(ncycles), cfg.nu1, cfg.nu2,
                         cfg.gamma, 1 if cfg.fcycle else 0, 1 if cfg.skip_pre else 0)
    if rc != 0:
        raise MemoryError("C oracle allocation failed")
    del keep, gmg
    return out

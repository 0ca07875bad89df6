"""The device's correctly rounded exp, log, pow and expm1 (cylag_amd/csrc/cyl_crmath.h).

The step evaluates the drag law (lagrangian_particle_set.pyx:790-832), the LSM-2
direct-integrator coefficients (update_state_lsm2.pyx:1240-1327), the LSM-3
transitions and the Ziggurat tail with these functions, so that positions agree
bit for bit with the reference's glibc results.  Each function has a fast path
(a value to about 2^-64..2^-69 and an error bound that decides its rounding) and
a double-double evaluation for the arguments the bound leaves undecided.  Checks,
on millions of arguments over the ranges the step uses and beyond:
  * wherever the fast path decides, its value equals the double-double one;
  * the function as the kernels call it equals the double-double value;
  * the double-double value is the correctly rounded one, against 60-digit
    decimal arithmetic on a sample (this pins the reference point itself);
  * the fast path decides all but a small fraction of the arguments.
The harness (tests/native/crmath_eval.hip) is test infrastructure built by
__graft_entry__.build().
"""
import ctypes as C
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libcrmath_eval.so")

pytestmark = pytest.mark.gpu

FN = {"exp": 0, "log": 1, "pow": 2, "expm1": 3}


def _lib():
    if not os.path.exists(LIB):
        raise ImportError("tests/native/libcrmath_eval.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    lib.crm_eval.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, C.POINTER(C.c_int)]
    lib.crm_eval.restype = C.c_int
    return lib


def evaluate(fn, x, y=None):
    lib = _lib()
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size
    fast, full, fin = (np.empty(n) for _ in range(3))
    dec = np.empty(n, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    yp = None if y is None else np.ascontiguousarray(y, dtype=np.float64).ctypes.data_as(dp)
    rc = lib.crm_eval(FN[fn], n, x.ctypes.data_as(dp), yp, fast.ctypes.data_as(dp), full.ctypes.data_as(dp),
                      fin.ctypes.data_as(dp), dec.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0, "crm_eval failed ({})".format(rc)
    return fast, full, fin, dec.astype(bool)


def _args(fn, rng):
    if fn == "exp":
        x = np.concatenate([rng.uniform(-708., 709.78, 1_000_000), rng.uniform(-50., 50., 1_000_000),
                            rng.uniform(-1e-3, 1e-3, 300_000),
                            -10. ** rng.uniform(-8., 2., 500_000)])  # -a1 dt of the LSM-2 coefficients
        return x, None
    if fn == "log":
        u32 = rng.integers(1, 2 ** 32, 1_000_000, dtype=np.uint64).astype(np.float64) / 2. ** 32  # Ziggurat tail draws
        x = np.concatenate([10. ** rng.uniform(-300., 300., 1_000_000), 1. + rng.uniform(-0.05, 0.05, 1_000_000),
                            u32, rng.uniform(0.5, 2., 500_000)])
        return x[x != 1.], None
    if fn == "pow":
        ys = np.array([0.687, -0.67, -0.33, 0.11, -1.625, 2.63, 10., -10., 0.1, 3., -1., 0.5, 2.])
        x1 = 10. ** rng.uniform(-3., 6., 1_300_000)             # Reynolds numbers and diameters
        y1 = ys[rng.integers(0, ys.size, x1.size)]
        x2 = 10. ** rng.uniform(-5., 5., 700_000)
        y2 = rng.uniform(-20., 20., x2.size)
        x, y = np.concatenate([x1, x2]), np.concatenate([y1, y2])
        a = y * np.log(x)
        keep = (x != 1.) & (y != 0.) & (a > -700.) & (a < 700.)
        return x[keep], y[keep]
    x = np.concatenate([rng.uniform(-38., 40., 1_000_000), rng.uniform(-1. / 32, 1. / 32, 1_000_000),
                        np.sign(rng.uniform(-1, 1, 500_000)) * 10. ** rng.uniform(-12., -1.6, 500_000)])
    return x, None


@pytest.mark.parametrize("fn", ["exp", "log", "pow", "expm1"])
def test_fast_path_equals_double_double(fn):
    rng = np.random.default_rng(20261204 + FN[fn])
    x, y = _args(fn, rng)
    fast, full, fin, dec = evaluate(fn, x, y)
    print("{}: {} arguments, fast path decided {:.5f}".format(fn, x.size, dec.mean()))
    bad = dec & (fast.view(np.int64) != full.view(np.int64))
    assert not bad.any(), "{}: fast path differs at {} arguments, first x={!r}".format(
        fn, int(bad.sum()), float(x[bad][0]))
    assert np.array_equal(fin.view(np.int64), full.view(np.int64)), fn
    # the fast path should leave only a small fraction to the double-double evaluation
    assert dec.mean() > 0.985, "{}: fast path decided only {:.4f}".format(fn, dec.mean())


def _decimal_ref(fn, a, b=None):
    A = Decimal(float(a))
    if fn == "exp":
        v = A.exp()
    elif fn == "log":
        v = A.ln()
    elif fn == "pow":
        v = (Decimal(float(b)) * A.ln()).exp()
    else:
        v = A.exp() - 1
    return float(v)  # Decimal -> float rounds correctly (to nearest)


@pytest.mark.parametrize("fn", ["exp", "log", "pow", "expm1"])
def test_double_double_is_correctly_rounded(fn):
    getcontext().prec = 60
    rng = np.random.default_rng(777 + FN[fn])
    x, y = _args(fn, rng)
    pick = rng.choice(x.size, 3000, replace=False)
    xs = x[pick]
    ys = None if y is None else y[pick]
    fast, full, fin, dec = evaluate(fn, xs, ys)
    ref = np.array([_decimal_ref(fn, xs[i], None if ys is None else ys[i]) for i in range(xs.size)])
    wrong = fin.view(np.int64) != ref.view(np.int64)
    assert not wrong.any(), "{}: {} of {} not correctly rounded, first x={!r}".format(
        fn, int(wrong.sum()), xs.size, float(xs[wrong][0]))

"""The edge pass's bounded-argument sin/cos (csrc/sincos_bounded.h), compiled for the host
with g++ from the same header the device build includes, against float64 sin/cos: within
2 ulp of the correctly rounded values over |x| <= kSinCosBoundedMax.

The reference takes sin/cos of the wrapped phase difference with __sincosf
(leviathan_kernel.cu:150-152); the FP64 oracle uses libm; this routine sits between them.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "leviathan_v1_amd", "csrc", "sincos_bounded.h")
BOUND = 8.0

WRAPPER = r"""
#define LEV_HD static inline
#include "sincos_bounded.h"
extern "C" void sincos_many(const float* x, long n, float* s, float* c) {
  for (long i = 0; i < n; ++i) lev::sincos_bounded(x[i], s + i, c + i);
}
"""


@pytest.fixture(scope="module")
def sincos(tmp_path_factory):
    d = tmp_path_factory.mktemp("sincos")
    src = d / "wrap.cpp"
    src.write_text(WRAPPER)
    so = d / "libsincos.so"
    # -ffp-contract=off: the device build's setting (every fma in the routine is explicit)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", os.path.dirname(HEADER), str(src), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.sincos_many.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]

    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        s = np.empty_like(x)
        c = np.empty_like(x)
        lib.sincos_many(x.ctypes.data, x.size, s.ctypes.data, c.ctypes.data)
        return s, c
    return run


def ulp_err(got, exact):
    """|got - exact| in units of the float32 ulp at the correctly rounded result."""
    ref = exact.astype(np.float32)
    ulp = np.spacing(np.abs(ref)).astype(np.float64)
    ulp = np.maximum(ulp, np.float64(np.spacing(np.float32(0))))
    return np.abs(got.astype(np.float64) - exact) / ulp


def test_dense_sweep_within_2_ulp(sincos):
    x = np.linspace(-BOUND, BOUND, 4_000_001, dtype=np.float64).astype(np.float32)
    s, c = sincos(x)
    xs = x.astype(np.float64)
    es, ec = ulp_err(s, np.sin(xs)), ulp_err(c, np.cos(xs))
    assert es.max() <= 2.0, (float(es.max()), float(x[es.argmax()]))
    assert ec.max() <= 2.0, (float(ec.max()), float(x[ec.argmax()]))


def test_hard_points_within_2_ulp(sincos):
    """Floats next to the multiples of pi/4 (quadrant boundaries and zeros of sin and cos,
    where the reduction's cancellation is largest), tiny arguments and signed zeros."""
    centres = np.arange(-10, 11) * (np.pi / 4)
    centres = centres[np.abs(centres) <= BOUND]
    pts = []
    for m in centres:
        f = np.float32(m)
        for k in range(-64, 65):
            pts.append(np.nextafter(f, np.float32(np.inf) if k > 0 else np.float32(-np.inf)) if k else f)
            f2 = f
            for _ in range(abs(k)):
                f2 = np.nextafter(f2, np.float32(np.inf) if k > 0 else np.float32(-np.inf))
            pts.append(f2)
    tiny = np.array([1e-30, 1e-20, 1e-10, 1e-6, 3e-4, 1e-45], dtype=np.float32)
    x = np.concatenate([np.array(pts, dtype=np.float32), tiny, -tiny, np.float32([0.0, -0.0])])
    x = x[np.abs(x) <= BOUND]
    s, c = sincos(x)
    xs = x.astype(np.float64)
    es, ec = ulp_err(s, np.sin(xs)), ulp_err(c, np.cos(xs))
    assert es.max() <= 2.0, (float(es.max()), float(x[es.argmax()]))
    assert ec.max() <= 2.0, (float(ec.max()), float(x[ec.argmax()]))
    assert np.all(s[x == 0] == 0) and np.all(c[x == 0] == 1)  # sin(-0) is +0 here: a zero contribution either way


def test_quadrant_identities(sincos):
    """sin^2 + cos^2 = 1 and the odd/even symmetries hold to rounding on a random sample."""
    rng = np.random.default_rng(0)
    x = rng.uniform(-BOUND, BOUND, 200_000).astype(np.float32)
    s, c = sincos(x)
    sn, cn = sincos(-x)
    assert np.array_equal(sn, -s) and np.array_equal(cn, c)
    one = s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2
    assert np.abs(one - 1).max() < 4 * 2 ** -24

"""The bulk form of the LSQR's ordered norms (fmt2d_amd/csrc/ordered_sum.h:
increments of the running sum's integer mantissa added as integers, the
terms that change its binade or tie added the ordinary way), compiled for
the CPU with g++ and run chunk by chunk as seqnorm_kernel runs it, against
the plain front-to-back loop s = s + x[k] (inversion2d.f90's sum of
squares in order).  Bit for bit on vectors chosen to exercise every branch:
zeros, subnormals, exact half remainders, terms far above and far below
the running sum, binade changes inside and at the edges of chunks, and
non-finite terms."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("osum")
    so = str(d / "libosum.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", so,
                    "-I", os.path.join(ROOT, "fmt2d_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "ordered_sum_check.cpp")], check=True)
    L = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    L.serial_sum.restype = C.c_double
    L.serial_sum.argtypes = [dp, C.c_longlong]
    L.bulk_sum.restype = C.c_double
    L.bulk_sum.argtypes = [dp, C.c_longlong, C.c_int, C.POINTER(C.c_longlong), C.c_int]
    return L


def _sums(L, x, chunk, plant=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    p = x.ctypes.data_as(C.POINTER(C.c_double))
    steps = C.c_longlong(0)
    b = L.bulk_sum(p, len(x), chunk, C.byref(steps), plant)
    s = L.serial_sum(p, len(x))
    return s, b, steps.value


def _cases():
    r = np.random.default_rng(5)
    yield "normal squares", r.standard_normal(300_000) ** 2
    yield "wide range", (r.standard_normal(50_000) * 10.0 ** r.uniform(-150, 150, 50_000)) ** 2
    yield "zeros first", np.concatenate([np.zeros(5000), r.random(5000), np.zeros(100), r.random(100)])
    yield "subnormals", r.integers(0, 2 ** 40, 20_000).astype(np.float64) * 2.0 ** -1074
    yield "subnormal into normal", np.concatenate([np.full(3000, 2.0 ** -1030), r.random(10) * 2.0 ** -1022])
    # exact halves: with s = 1 (ulp 2^-52) the terms 2^-53 and 3 * 2^-53 tie
    yield "ties", np.concatenate([[1.0], np.tile([2.0 ** -53, 3 * 2.0 ** -53, 2.0 ** -54, 2.0 ** -53], 2000)])
    yield "small integers", r.integers(0, 8, 100_000).astype(np.float64)
    yield "growing", np.arange(1, 70_000, dtype=np.float64) ** 2
    yield "huge then tiny", np.concatenate([[1e300], r.random(1000) * 1e-300, [1e290], r.random(1000)])
    yield "binade edge", np.concatenate([[2.0 ** 53 - 1], np.ones(100), [0.5, 0.5, 1.5, 2.5]])
    yield "overflow", np.concatenate([np.full(10, 1.7e308), r.random(100)])
    yield "nan", np.concatenate([r.random(100), [np.nan], r.random(100)])
    yield "inf", np.concatenate([r.random(100), [np.inf], r.random(100)])
    yield "empty", np.zeros(0)
    yield "single", np.array([3.0])


@pytest.mark.parametrize("chunk", [4096, 7, 1])
def test_bulk_sum_matches_the_plain_loop(lib, chunk):
    for name, x in _cases():
        s, b, steps = _sums(lib, x, chunk)
        assert np.float64(s).tobytes() == np.float64(b).tobytes(), (name, chunk, s, b, steps)


def test_bulk_sum_takes_few_ordinary_steps(lib):
    # a binade change (about 20 over 3e5 squares) or a tie forces one; the
    # rest of the terms go in bulk
    r = np.random.default_rng(9)
    s, b, steps = _sums(lib, r.standard_normal(300_000) ** 2, 4096)
    assert s == b and steps < 100, steps


@pytest.mark.parametrize("plant", [1, 2], ids=["no_binade_check", "ties_rounded_up"])
def test_a_planted_mistake_is_seen(lib, plant):
    wrong = [name for name, x in _cases()
             if np.float64(_sums(lib, x, 4096)[0]).tobytes() != np.float64(_sums(lib, x, 4096, plant)[1]).tobytes()]
    assert wrong, plant

"""The LSQR's ordered-norm kernel (seqnorm_kernel in fmt2d_amd/csrc/lsqr.hip,
the bulk form of fmt2d_amd/csrc/ordered_sum.h) through
fmt2d_ordered_sumsq, against the plain front-to-back loop over the same
squares (numpy's cumsum is a running sum: its last entry is that loop's
result).  Bit for bit, on vectors that exercise every branch of the bulk
form: many chunks, binade changes, exact half remainders, zeros,
subnormals, huge and tiny terms, overflow, inf and NaN."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _serial(a):
    with np.errstate(over="ignore", invalid="ignore"):
        sq = np.asarray(a, dtype=np.float64) * np.asarray(a, dtype=np.float64)
        return float(np.cumsum(sq)[-1]) if len(sq) else 0.0


def _cases():
    r = np.random.default_rng(11)
    yield "normal", r.standard_normal(300_001)
    yield "many chunks", r.standard_normal(2_500_000) * 3.0
    yield "wide range", r.standard_normal(60_000) * 10.0 ** r.uniform(-75, 75, 60_000)
    yield "zeros first", np.concatenate([np.zeros(9000), r.random(5000), np.zeros(100), r.random(100)])
    yield "tiny", r.random(20_000) * 2.0 ** -530
    # s = 2 after the first two terms (last mantissa bit 2^-51): 2^-52 and
    # 9 * 2^-52 are exact half remainders, 2^-54 rounds away
    yield "ties", np.concatenate([[1.0, 1.0], np.tile([2.0 ** -26, 3 * 2.0 ** -26, 2.0 ** -27], 3000)])
    yield "growing", np.arange(1, 70_000, dtype=np.float64)
    yield "huge then tiny", np.concatenate([[1e150], r.random(1000) * 1e-150, [1e145], r.random(1000)])
    yield "overflow", np.concatenate([np.full(10, 1.3e154), r.random(5000)])
    yield "nan", np.concatenate([r.random(5000), [np.nan], r.random(100)])
    yield "inf", np.concatenate([r.random(100), [-np.inf], r.random(9000)])
    yield "negative entries", -r.random(10_000)
    yield "empty", np.zeros(0)
    yield "single", np.array([3.0])


def test_ordered_sumsq_matches_the_plain_loop():
    from fmt2d_amd import hip
    ctx = hip.Context(0)
    for name, a in _cases():
        print("case", name, len(a), flush=True)
        got = ctx.ordered_sumsq(a)
        want = _serial(a)
        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (name, got, want)

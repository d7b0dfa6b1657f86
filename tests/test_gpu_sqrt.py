"""The trace's square roots against IEEE-754 double arithmetic.

Every kernel of the trace takes sqrt(x) and 1/sqrt(x) from one pair of
device functions that skip the range handling of the compiler's sequences
where it cannot change a bit (x in [2^-767, DBL_MAX]) and fall back to the
full sequences elsewhere.  The reference computes both on the CPU with
correctly rounded IEEE arithmetic (numpy's sqrt and division are the same),
so the bar is bit identity on every input class: random bit patterns over
the whole exponent range, the range boundaries and their neighbours,
zeros, denormals, infinities, NaNs, negatives, exact squares and values a
few ulps from 1 -- and, through the trace tests, every optic."""
import numpy as np
import pytest
import torch

from batoid_amd import _lib

pytestmark = pytest.mark.gpu


def _inputs():
    rng = np.random.default_rng(20261203)
    bits = rng.integers(0, 2 ** 63, size=400_000, dtype=np.uint64)  # every positive double
    parts = [bits.view(np.float64)]
    # wide magnitudes as the trace meets them
    parts.append(rng.uniform(0, 4, 200_000))
    parts.append(10.0 ** rng.uniform(-30, 30, 200_000))
    parts.append(1.0 + rng.integers(-2000, 2000, 20_000) * np.finfo(float).eps)
    k = rng.integers(1, 2 ** 26, 20_000).astype(float)
    parts.append(k * k)                                     # exact squares
    edges = np.array([2.0 ** -767, 2.0 ** -768, 2.0 ** -766, np.finfo(float).max,
                      np.finfo(float).tiny, 5e-324, 0.0, -0.0, np.inf, -np.inf,
                      np.nan, -1.0, -5e-324, 1.0, 0.25, 4.0, 2.0 ** 1023, 2.0 ** -1022])
    near = []
    with np.errstate(over="ignore"):
        near += _neighbours()
    parts += [edges, np.array(near)]
    x = np.concatenate(parts)
    return np.concatenate([x, -x[:1000]])


def _neighbours():
    near = []
    for e in (2.0 ** -767, np.finfo(float).max, np.finfo(float).tiny, 1.0):
        v = np.array([e])
        for _ in range(64):
            near.append(v[0])
            v = np.nextafter(v, np.inf)
        v = np.array([e])
        for _ in range(64):
            v = np.nextafter(v, -np.inf)
            near.append(v[0])
    return near


def _same_bits(got, want):
    both_nan = np.isnan(got) & np.isnan(want)
    return both_nan | (got.view(np.uint64) == want.view(np.uint64))


def test_sqrt_and_rsqrt_are_ieee():
    x = _inputs()
    n = len(x)
    xd = torch.from_numpy(x).cuda()
    out = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    print(f"case bt_check_sqrt n={n}", flush=True)
    rc = lib.bt_check_sqrt(xd.data_ptr(), out.data_ptr(), n, None)
    _lib.check(rc, "bt_check_sqrt")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    with np.errstate(invalid="ignore", divide="ignore"):
        want_sqrt = np.sqrt(x)
        want_rsqrt = 1.0 / np.sqrt(x)
    ok_s = _same_bits(got[:n], want_sqrt)
    ok_r = _same_bits(got[n:], want_rsqrt)
    assert ok_s.all(), f"sqrt differs at x = {x[~ok_s][:5]}"
    assert ok_r.all(), f"1/sqrt differs at x = {x[~ok_r][:5]}"
    # the plain-range path carries almost every input above
    plain = (x >= 2.0 ** -767) & (x <= np.finfo(float).max)
    assert plain.mean() > 0.5


def test_check_sqrt_rejects_bad_arguments():
    lib = _lib.load()
    assert lib.bt_check_sqrt(None, None, 4, None) != 0
    assert lib.bt_check_sqrt(None, None, 0, None) == 0

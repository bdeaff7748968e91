"""The on-device error checks (xc_validate_result_*, xc_norm_error_*) against the oracle's
restatements (xco_validate_result: src/validate.hpp:134-188; xco_norm_error), on arrays with
ties at the maximum, values at and below the 1e-5 threshold, zeros, NaN and infinite errors.
max, the index of the max and the values there must be equal; mean and standard deviation
agree to 1e-12 relative (the sums run in another order)."""
import ctypes

import numpy as np
import pytest

from xc_testutil import oracle, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cpu_validate(result, valid):
    r = np.ascontiguousarray(result, np.float64).ravel()
    v = np.ascontiguousarray(valid, np.float64).ravel()
    out = np.zeros(5)
    oracle().xco_validate_result(_p(r), _p(v), r.size, _p(out))
    return out


def cpu_norm_error(x, ref, den):
    x, ref, den = (np.ascontiguousarray(a, np.float64).ravel() for a in (x, ref, den))
    out = np.zeros(3)
    oracle().xco_norm_error(_p(x), _p(ref), _p(den), x.size, _p(out))
    return out


def same(a, b, rel=1e-12):
    return (np.isnan(a) and np.isnan(b)) or a == b or abs(a - b) <= rel * max(abs(a), abs(b))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 100_000, 3_000_001])
def test_validate_result_matches_oracle(dtype, n):
    rng = np.random.default_rng(n)
    valid = rng.normal(size=n).astype(dtype)
    result = (valid * (1 + rng.normal(scale=1e-6, size=n))).astype(dtype)
    result[::97] = valid[::97]                 # exact matches
    small = rng.random(n) < 0.05               # both at or below 1e-5: d = 0
    valid[small] = (1e-5 * rng.random(small.sum())).astype(dtype)
    result[small] = (-1e-5 * rng.random(small.sum())).astype(dtype)
    if n > 10:                                 # a tie at the maximum: the first index wins
        result[n // 3], valid[n // 3] = 2.0, 1.0
        result[2 * n // 3], valid[2 * n // 3] = 4.0, 2.0
    x = xcorr()
    got = x.validate_result(torch.from_numpy(result).cuda(), torch.from_numpy(valid).cuda())
    want = cpu_validate(result, valid)
    g = [got["max"], got["mean"], got["stddev"], got["result_at_max"], got["valid_at_max"]]
    assert g[0] == want[0] and g[3] == want[3] and g[4] == want[4], (g, want)
    assert same(g[1], want[1]) and same(g[2], want[2], 1e-9), (g, want)


def test_validate_result_nan_and_empty():
    x = xcorr()
    r = np.array([1.0, np.nan, 3.0, 0.5]); v = np.array([1.1, 2.0, 3.0, 0.25])
    got = x.validate_result(torch.from_numpy(r).cuda(), torch.from_numpy(v).cuda())
    want = cpu_validate(r, v)
    assert got["max"] == want[0] and got["result_at_max"] == want[3] and np.isnan(got["mean"]) and np.isnan(want[1])
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    assert all(np.isnan(v) for v in x.validate_result(e, e).values())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 300, 65_536, 2_000_003])
def test_norm_error_matches_oracle(dtype, n):
    rng = np.random.default_rng(7 + n)
    ref = rng.normal(size=n)
    den = np.abs(ref) + rng.random(n)
    xv = (ref + rng.normal(scale=1e-7, size=n) * den).astype(dtype)
    den[::50] = 0.0                            # den 0: error 0 where x == ref, else infinite
    ref[::50] = xv[::50]
    if n > 100:
        den[n // 2] = 0.0                      # one infinite error
        xv[3], ref[3], den[3] = 2.0, 1.0, 1.0  # and a finite tie at 1 is beaten by it
    x = xcorr()
    got = x.norm_error(torch.from_numpy(xv).cuda(), torch.from_numpy(ref).cuda(), torch.from_numpy(den).cuda())
    want = cpu_norm_error(xv, ref, den)
    assert got[0] == want[0] and got[2] == int(want[2]), (got, want)
    assert same(got[1], want[1]) or (np.isinf(got[1]) and np.isinf(want[1])), (got, want)


def test_norm_error_of_a_correlation_against_the_double_fft():
    """The device check on a real result: bf16x3 at 64 x 128^2 against the own FFT on double data
    and corr(|L|,|R|), compared with the oracle-based check on two results."""
    x = xcorr()
    rng = np.random.default_rng(3)
    L = torch.from_numpy(rng.random((1, 128, 128), dtype=np.float32)).cuda()
    R = torch.from_numpy(rng.random((64, 128, 128), dtype=np.float32)).cuda()
    got = x.cross_corr("one_to_many", L, R, precision="bf16x3")
    ref = x.cross_corr("one_to_many", L.double(), R.double(), precision="fft64", tile=1)
    den = x.cross_corr("one_to_many", L.abs().double(), R.abs().double(), precision="fft64", tile=1)
    mx, mean, _ = x.norm_error(got, ref, den)
    assert mx <= 2e-6 and 0 < mean < mx
    sub = cpu_norm_error(got[:2].cpu().numpy(), ref[:2].cpu().numpy(), den[:2].cpu().numpy())
    assert sub[0] <= mx * (1 + 1e-12)

"""GPU parity of the one_to_many crossover sweep: 1 left x 128 right float32 matrices,
U[0,1), at N x N for N in {32, 64, 128, 256, 512, 1024}, through AUTO and through the
library's own FP64 FFT (fft64 tile 1) -- the reference's FFT wrapper for this family is
src/one_to_many.hpp:679-745.

Bars (normalised error = |x - x64| / corr(|L|,|R|)): <= 1e-6 for fft64, <= 2e-6 for AUTO
(its candidates include bf16x3).  Up to 256 two or more results per size are compared with
the FP64 oracle.  At 512 and 1024 the oracle is out of reach: every result is checked by the
sum rule (sum of a correlation = sum(L) * sum(R)) and at its four one-element corners;
results beside the chunk boundaries of the column pass are checked at 35 shifts against the
definition and, for a few, against the rocFFT path (fft64 tile 0).  At 1024 x 1024 a pair's
column spectra take 33.6 MB, so the 128 pairs run as three chunks of at most 2 GiB (63, 63
and 2 pairs).  Smaller chunk limits (opt.split) take the same loop at sizes the oracle
reaches.
"""
import numpy as np
import pytest

from xc_testutil import norm_error, oracle_family, seeded, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

M = 128


def inputs(n):
    left = seeded((1, n, n), "u01", 500 + n, np.float32)
    right = seeded((M, n, n), "u01", 600 + n, np.float32)
    return left, right


def run(left, right, precision, tile=0, split=0):
    print(f"c5 one_to_many L{left.shape} R{right.shape} {precision}/{tile} split {split}", flush=True)
    x = xcorr()
    out = x.cross_corr("one_to_many", torch.from_numpy(left).cuda(), torch.from_numpy(right).cuda(),
                       precision=precision, tile=tile, split=split)
    torch.cuda.synchronize()
    return out


def check_oracle(left, right, got, idx, tol):
    for t in idx:
        r = right[t:t + 1].astype(np.float64)
        l = left.astype(np.float64)
        ref = oracle_family("one_to_one", l, r)
        den = oracle_family("one_to_one", l, r, absval=True)
        err, _ = norm_error(got[t:t + 1], ref, den)
        assert err <= tol, (left.shape, t, err)


def entry(left, right, dy, dx):
    """One output of the definition, sum_{r,c} L[r][c] R[r-dy][c-dx], in FP64."""
    H, W = left.shape
    a = left[max(0, dy):H + min(0, dy), max(0, dx):W + min(0, dx)].astype(np.float64)
    b = right[max(0, -dy):H - max(0, dy), max(0, -dx):W - max(0, dx)].astype(np.float64)
    return float((a * b).sum())


def check_large(left, right, out, n, tol):
    """Sum rule and corners of every result, 35 shifts of the results beside chunk
    boundaries; out is the device tensor of all M results."""
    lsum = left.astype(np.float64).sum()
    H = W = n
    boundary = sorted({0, 1, 61, 62, 63, 64, 125, 126, 127})
    for t in range(M):
        got = out[t].cpu().numpy()
        s = got.sum(dtype=np.float64)
        want = lsum * right[t].astype(np.float64).sum()
        assert abs(s - want) <= 1e-6 * want, (n, t, s, want)
        for dy, dx in ((-(H - 1), -(W - 1)), (-(H - 1), W - 1), (H - 1, -(W - 1)), (H - 1, W - 1)):
            w = entry(left[0], right[t], dy, dx)
            g = float(got[dy + H - 1, dx + W - 1])
            assert abs(g - w) <= tol * w, (n, t, dy, dx, g, w)
        if t in boundary:
            for dy in (-(H - 1), -(H - 2), -1, 0, 1, H - 2, H - 1):
                for dx in (-(W - 1), -1, 0, 1, W - 1):
                    w = entry(left[0], right[t], dy, dx)  # U[0,1) data: corr(|L|,|R|) = w
                    g = float(got[dy + H - 1, dx + W - 1])
                    assert abs(g - w) <= tol * w, (n, t, dy, dx, g, w)


@pytest.mark.parametrize("n", [32, 64, 128, 256])
def test_c5_against_the_oracle(n):
    left, right = inputs(n)
    own = run(left, right, "fft64", tile=1).cpu().numpy()
    assert own.shape == (M, 2 * n - 1, 2 * n - 1)
    idx = (0, 63, 127) if n <= 128 else (0, 127)
    check_oracle(left, right, own, idx, 1e-6)
    auto = run(left, right, "auto").cpu().numpy()
    check_oracle(left, right, auto, idx, 2e-6)
    # every result by the sum rule
    sums = own.astype(np.float64).reshape(M, -1).sum(axis=1)
    expect = left.astype(np.float64).sum() * right.astype(np.float64).reshape(M, -1).sum(axis=1)
    assert np.max(np.abs(sums - expect) / expect) < 1e-6


@pytest.mark.parametrize("n", [512, 1024])
def test_c5_beyond_the_oracle(n):
    left, right = inputs(n)
    own = run(left, right, "fft64", tile=1)
    assert tuple(own.shape) == (M, 2 * n - 1, 2 * n - 1)
    check_large(left, right, own, n, 1e-6)
    # the rocFFT path on results beside the chunk boundaries (63 and 126 pairs at 1024)
    pick = [0, 62, 63, 64, 126, 127]
    roc = run(left, np.ascontiguousarray(right[pick]), "fft64", tile=0).cpu().numpy().astype(np.float64)
    scale = float(n) * n / 4  # about the central value
    for i, t in enumerate(pick):
        assert np.max(np.abs(own[t].cpu().numpy() - roc[i])) / scale < 1e-6, (n, t)
    del own
    torch.cuda.empty_cache()
    auto = run(left, right, "auto")
    check_large(left, right, auto, n, 2e-6)


@pytest.mark.parametrize("chunk", [1, 3, 7])
def test_column_pass_chunks_against_the_oracle(chunk):
    # opt.split caps the pairs per launch of the column and inverse-row kernels: the chunk
    # loop runs 10 / 4 / 2 times (the last chunk partial) and must give the unchunked results.
    left = seeded((1, 48, 40), "u01", 71, np.float32)
    right = seeded((10, 48, 40), "u01", 72, np.float32)
    got = run(left, right, "fft64", tile=1, split=chunk).cpu().numpy()
    check_oracle(left, right, got, range(10), 1e-6)
    whole = run(left, right, "fft64", tile=1).cpu().numpy()
    assert np.array_equal(got, whole)


def test_column_pass_chunks_n_to_m_and_double():
    # pair index -> (left, right) across chunk boundaries for n_to_m; FP64 data too.
    x = xcorr()
    for dtype, tol in ((np.float32, 1e-6), (np.float64, 1e-9)):
        left = seeded((3, 33, 20), "normal", 73, dtype)
        right = seeded((5, 33, 20), "normal", 74, dtype)
        print(f"chunks n_to_m 3x5 33x20 {dtype.__name__}", flush=True)
        out = x.cross_corr("n_to_m", torch.from_numpy(left).cuda(), torch.from_numpy(right).cuda(),
                           precision="fft64", tile=1, split=4)
        torch.cuda.synchronize()
        ref = oracle_family("n_to_m", left.astype(np.float64), right.astype(np.float64))
        den = oracle_family("n_to_m", left.astype(np.float64), right.astype(np.float64), absval=True)
        err, _ = norm_error(out.cpu().numpy(), ref, den)
        assert err <= tol, (dtype, err)

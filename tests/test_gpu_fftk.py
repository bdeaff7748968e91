"""GPU parity of the FFT method on the library's own FP64 FFT kernels (precision fft64,
tile 1) against the CPU oracle, through the C ABI.

Bars (normalised error = |x - x64| / corr(|L|,|R|)): float data <= 1e-6 (FP64 transform,
FP32 results); double data <= 1e-9.  The FFT's rounding is global (about 1e-16 of the sum of
all products), while the metric divides it per output by corr(|L|,|R|), a single product
|a||b| at the four one-element-overlap corners; the double bar leaves room for that.  Sizes
beyond the oracle's reach are checked by the sum rule (sum of a correlation = sum(L) * sum(R))
and against the rocFFT path (fft64 tile 0).
"""
import numpy as np
import pytest

from xc_testutil import DATA, FIXTURES, fixture_family, load_csv, norm_error, oracle_family, seeded, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(family, left, right, tile=1):
    print(f"fftk {family} L{left.shape} R{right.shape} {left.dtype}", flush=True)
    out = xcorr().cross_corr(family, dev(left), dev(right), precision="fft64", tile=tile)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def check(family, left, right, got, tol):
    ref = oracle_family(family, left.astype(np.float64), right.astype(np.float64))
    den = oracle_family(family, left.astype(np.float64), right.astype(np.float64), absval=True)
    assert got.shape == ref.shape
    err, mean = norm_error(got, ref, den)
    assert err <= tol, (family, left.shape, right.shape, err, mean)
    return err


@pytest.mark.parametrize("valid_name", sorted(FIXTURES))
def test_golden_float_and_double(valid_name):
    fam = fixture_family(valid_name)
    l64 = load_csv(DATA / FIXTURES[valid_name][0])
    r64 = load_csv(DATA / FIXTURES[valid_name][1])
    l32, r32 = l64.astype(np.float32), r64.astype(np.float32)
    check(fam, l32, r32, run(fam, l32, r32), 1e-6)
    got = run(fam, l64, r64)
    check(fam, l64, r64, got, 1e-9)
    den = oracle_family(fam, l64, r64, absval=True)
    assert norm_error(got, load_csv(DATA / valid_name), den)[0] <= 3e-6  # the fixture's own digits


CASES = [
    ("one_to_one", 1, 1, 1, 1, "u01"),
    ("one_to_one", 2, 3, 1, 1, "normal"),
    ("one_to_one", 7, 5, 1, 1, "u11"),
    ("one_to_one", 16, 10, 1, 1, "u01"),
    ("one_to_one", 9, 17, 1, 1, "u01"),       # padded 32 (rows 17..32 of the column FFT are zero)
    ("one_to_one", 128, 128, 1, 1, "u01"),
    ("one_to_one", 313, 125, 1, 1, "u01"),
    ("one_to_one", 3, 600, 1, 1, "u11"),      # 1536-point rows (3 x 512), 16-point columns
    ("one_to_one", 600, 3, 1, 1, "u11"),
    ("one_to_many", 33, 65, 1, 5, "u01"),
    ("one_to_many", 64, 64, 1, 12, "u11"),
    ("n_to_mn", 12, 9, 3, 9, "normal"),
    ("n_to_mn", 32, 32, 2, 6, "u01"),
    ("n_to_m", 64, 64, 4, 6, "normal"),
    ("n_to_m", 17, 23, 3, 5, "u11"),
    ("one_to_many", 37, 40, 1, 4, "u11"),     # 96-point rows (3 x 32)
    ("n_to_m", 50, 150, 2, 3, "normal"),      # 384-point rows (3 x 128)
    ("one_to_many", 129, 256, 1, 3, "u11"),   # 512-point register path: rows and columns
    ("n_to_m", 200, 140, 2, 2, "normal"),     # 512-point columns, 512-point rows, partial H
    ("one_to_one", 256, 70, 1, 1, "u01"),     # 512-point columns only
    ("one_to_many", 140, 270, 1, 2, "u01"),   # 768-point rows (3 x 256), last row group partial
    ("n_to_m", 45, 700, 2, 2, "u11"),         # 1536-point rows, 45 rows: 6 groups of 8, the last partial
    ("one_to_many", 100, 450, 1, 2, "u01"),   # 1024-point rows (8-row blocks)
    ("n_to_m", 21, 800, 2, 2, "u11"),         # 2048-point rows, last row group partial
]


@pytest.mark.parametrize("family,rows,cols,n,m,dist", CASES)
def test_seeded_float(family, rows, cols, n, m, dist):
    left = seeded((n, rows, cols), dist, 21, np.float32)
    right = seeded((m, rows, cols), dist, 22, np.float32)
    check(family, left, right, run(family, left, right), 1e-6)


@pytest.mark.parametrize("family,rows,cols,n,m,dist", CASES[::3])
def test_seeded_double(family, rows, cols, n, m, dist):
    left = seeded((n, rows, cols), dist, 23, np.float64)
    right = seeded((m, rows, cols), dist, 24, np.float64)
    check(family, left, right, run(family, left, right), 1e-9)


def test_c2_config_subset_and_sum_rule():
    left = seeded((1, 256, 256), "u01", 5, np.float32)
    right = seeded((64, 256, 256), "u01", 6, np.float32)
    got = run("one_to_many", left, right)
    for t in (0, 63):
        check("one_to_one", left, right[t:t + 1], got[t:t + 1], 2e-6)
    sums = got.astype(np.float64).reshape(64, -1).sum(axis=1)
    expect = left.astype(np.float64).sum() * right.astype(np.float64).reshape(64, -1).sum(axis=1)
    assert np.max(np.abs(sums - expect) / expect) < 1e-6


def test_c3_config_n_to_m_16x64_at_64():
    left = seeded((16, 64, 64), "normal", 3, np.float32)
    right = seeded((64, 64, 64), "normal", 4, np.float32)
    check("n_to_m", left, right, run("n_to_m", left, right), 1e-6)
    l64, r64 = left.astype(np.float64), right.astype(np.float64)
    check("n_to_m", l64, r64, run("n_to_m", l64, r64), 1e-9)


@pytest.mark.parametrize("size", [1024, 2048])
def test_largest_transforms_against_rocfft_and_sum_rule(size):
    # 2048-point (and 4096-point) transforms in both directions; too large for the oracle.
    left = seeded((1, size, size), "u01", 31, np.float32)
    right = seeded((2, size, size), "u01", 32, np.float32)
    own = run("one_to_many", left, right)
    sums = own.astype(np.float64).reshape(2, -1).sum(axis=1)
    expect = left.astype(np.float64).sum() * right.astype(np.float64).reshape(2, -1).sum(axis=1)
    assert np.max(np.abs(sums - expect) / expect) < 1e-6
    if size <= 1024:
        roc = run("one_to_many", left, right, tile=0)
        scale = float(size) * size / 4  # about the central value
        assert np.max(np.abs(own.astype(np.float64) - roc)) / scale < 1e-6
    # corner outputs (overlap of one element) against the definition
    n = size - 1
    assert abs(own[0, 0, 0] - left[0, 0, 0] * right[0, n, n]) <= 1e-5
    assert abs(own[1, 2 * n, 2 * n] - left[0, n, n] * right[1, 0, 0]) <= 1e-5


def entry(left, right, dy, dx):
    """One output of the definition, sum_{r,c} L[r][c] R[r-dy][c-dx], in FP64."""
    H, W = left.shape
    a = left[max(0, dy):H + min(0, dy), max(0, dx):W + min(0, dx)].astype(np.float64)
    b = right[max(0, -dy):H - max(0, dy), max(0, -dx):W - max(0, dx)].astype(np.float64)
    return float((a * b).sum())


@pytest.mark.parametrize("rows,cols", [(1100, 600), (600, 1100)])
def test_ragged_sizes_of_the_4096_point_transforms(rows, cols):
    # 4096-point columns (rows = 1100) or rows (cols = 1100) with padding that is not 2x the
    # size: the column pass keeps 2H-1 of the Pr circular rows.  Too large for the oracle:
    # the sum rule, the rocFFT path, and outputs beside both halves of the result (the shifts
    # dy = 0, +-1, +-(H-1)) against the definition.
    left = seeded((1, rows, cols), "u01", 35, np.float32)
    right = seeded((2, rows, cols), "u01", 36, np.float32)
    own = run("one_to_many", left, right)
    assert own.shape == (2, 2 * rows - 1, 2 * cols - 1)
    sums = own.astype(np.float64).reshape(2, -1).sum(axis=1)
    expect = left.astype(np.float64).sum() * right.astype(np.float64).reshape(2, -1).sum(axis=1)
    assert np.max(np.abs(sums - expect) / expect) < 1e-6
    roc = run("one_to_many", left, right, tile=0)
    scale = float(rows) * cols / 4  # about the central value
    assert np.max(np.abs(own.astype(np.float64) - roc)) / scale < 1e-6
    H, W = rows, cols
    for t in range(2):
        for dy in (-(H - 1), -(H - 2), -1, 0, 1, H - 2, H - 1):
            for dx in (-(W - 1), -1, 0, 1, W - 1):
                want = entry(left[0], right[t], dy, dx)  # U[0,1) data: corr(|L|,|R|) = want
                got = float(own[t, dy + H - 1, dx + W - 1])
                assert abs(got - want) <= 1e-6 * want, (t, dy, dx, got, want)


@pytest.mark.parametrize("tile", [1, 0])
def test_calls_on_two_streams_at_once(tile):
    # Scratch memory and rocFFT plans are kept per stream: two streams running the same
    # shape at the same time must each get the single-stream results (the kernels are
    # deterministic, so bit for bit).
    x = xcorr()
    ins = [(dev(seeded((1, 96, 80), "u01", 51 + 2 * i, np.float32)),
            dev(seeded((6, 96, 80), "u01", 52 + 2 * i, np.float32))) for i in range(2)]
    want = [x.cross_corr("one_to_many", L, R, precision="fft64", tile=tile) for L, R in ins]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    got = [[], []]
    print(f"two streams, fft64 tile {tile}", flush=True)
    for _ in range(4):
        for i in range(2):
            with torch.cuda.stream(streams[i]):
                got[i].append(x.cross_corr("one_to_many", *ins[i], precision="fft64", tile=tile))
    torch.cuda.synchronize()
    for i in range(2):
        for g in got[i]:
            assert torch.equal(g, want[i])


def test_linearity_in_the_right_matrix():
    left = seeded((1, 200, 150), "u11", 41, np.float32)
    right = seeded((2, 200, 150), "u11", 42, np.float32)
    both = run("one_to_many", left, right)
    s = (right[0].astype(np.float64) + 2.0 * right[1]).astype(np.float32)[None]
    comb = run("one_to_many", left, s)[0]
    den = oracle_family("one_to_one", left.astype(np.float64), s.astype(np.float64), absval=True)[0]
    assert np.max(np.abs(comb - (both[0].astype(np.float64) + 2.0 * both[1])) / den) < 3e-6


def test_size_limit_is_rejected_before_device_work():
    x = xcorr()
    left = torch.zeros((1, 2049, 4), device="cuda")
    right = torch.zeros((1, 2049, 4), device="cuda")
    with pytest.raises(x.XcError) as e:
        x.cross_corr("one_to_one", left, right, precision="fft64", tile=1)
    assert e.value.status == 4
    with pytest.raises(x.XcError):
        x.cross_corr("one_to_one", left[:, :8], right[:, :8], precision="fft64", tile=3)


@pytest.mark.parametrize("family,rows,cols,n,m", [("one_to_many", 37, 40, 1, 4), ("n_to_m", 50, 150, 2, 3),
                                                  ("one_to_one", 140, 270, 1, 1), ("n_to_mn", 9, 600, 2, 4)])
def test_power_of_two_rows_variant(family, rows, cols, n, m):
    # tile 2 transforms rows at the power of two >= 2W-1 where tile 1 takes 3 * 2^n: the same
    # correlation through other transform lengths, for float and double data.
    for dtype, tol in ((np.float32, 1e-6), (np.float64, 1e-9)):
        left = seeded((n, rows, cols), "u11", 61, dtype)
        right = seeded((m, rows, cols), "u11", 62, dtype)
        got = run(family, left, right, tile=2)
        check(family, left, right, got, tol)


def test_double_data_auto_picks_the_own_fft_at_c3():
    # AUTO for float64 data times the FP64 direct kernel and the own FFT once per shape; at
    # 16 x 64 pairs of 64x64 the FFT is more than 10x faster.
    left = seeded((16, 64, 64), "normal", 3, np.float64)
    right = seeded((64, 64, 64), "normal", 4, np.float64)
    x = xcorr()
    for _ in range(2):
        out = x.cross_corr("n_to_m", dev(left), dev(right), precision="auto")
        torch.cuda.synchronize()
        assert x.last_launch()["kernel"].startswith("fftk"), x.last_launch()["kernel"]
    check("n_to_m", left, right, out.cpu().numpy(), 1e-9)

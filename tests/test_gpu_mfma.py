"""GPU parity of the MFMA implicit-GEMM kernels against the FP64 oracle.

Bars (normalised error |x - x64| / corr(|L|,|R|)):
  bf16x3 and fp16x3 (with the exact FP32 border frame) <= 2e-6 everywhere;
  single bf16 <= 8e-3 and single fp16 <= 1e-3 (one rounding of each operand, products summed in FP32).
"""
import numpy as np
import pytest

from xc_testutil import DATA, FIXTURES, fixture_family, load_csv, norm_error, oracle_family, seeded, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(family, left, right, **kw):
    out = xcorr().cross_corr(family, dev(left), dev(right), **kw)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def errors(family, left, right, got):
    l64, r64 = left.astype(np.float64), right.astype(np.float64)
    ref = oracle_family(family, l64, r64)
    den = oracle_family(family, l64, r64, absval=True)
    assert got.shape == ref.shape
    return np.abs(got.astype(np.float64) - ref) / np.where(den > 0, den, 1.0), ref, den


CASES = [
    ("one_to_one", 1, 1, 1, 1, "u01"),
    ("one_to_one", 3, 2, 1, 1, "normal"),
    ("one_to_one", 7, 5, 1, 1, "u11"),
    ("one_to_one", 16, 10, 1, 1, "u01"),
    ("one_to_one", 10, 16, 1, 1, "u01"),
    ("one_to_one", 33, 65, 1, 1, "u01"),
    ("one_to_one", 128, 128, 1, 1, "u01"),
    ("one_to_one", 200, 70, 1, 1, "normal"),
    ("one_to_one", 70, 300, 1, 1, "u01"),
    ("one_to_many", 64, 64, 1, 5, "u01"),
    ("one_to_many", 31, 97, 1, 3, "u11"),
    ("n_to_mn", 12, 9, 3, 9, "normal"),
    ("n_to_mn", 48, 48, 2, 4, "u01"),
    ("n_to_m", 64, 64, 3, 4, "normal"),
    ("n_to_m", 17, 23, 3, 5, "u11"),
]


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("family,rows,cols,n,m,dist", CASES)
def test_bf16x3_seeded(family, rows, cols, n, m, dist, prec):
    left = seeded((n, rows, cols), dist, 500 + rows, np.float32)
    right = seeded((m, rows, cols), dist, 600 + cols, np.float32)
    err, _, _ = errors(family, left, right, run(family, left, right, precision=prec))
    assert err.max() <= 2e-6, err.max()


def test_fp16x3_is_closer_than_bf16x3_on_unit_data():
    """fp16 hi + lo keeps 22 significant bits of each operand against bf16's 16: on data of
    order one (no fp16 range or subnormal limits) the split-fp16 result is the closer one."""
    left = seeded((1, 96, 96), "u01", 71, np.float32)
    right = seeded((4, 96, 96), "u01", 72, np.float32)
    e_bf, _, _ = errors("one_to_many", left, right, run("one_to_many", left, right, precision="bf16x3", border=1))
    e_fp, _, _ = errors("one_to_many", left, right, run("one_to_many", left, right, precision="fp16x3", border=1))
    assert e_fp.max() < e_bf.max(), (e_fp.max(), e_bf.max())


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("valid_name", sorted(FIXTURES))
def test_bf16x3_golden(valid_name, prec):
    left = load_csv(DATA / FIXTURES[valid_name][0]).astype(np.float32)
    right = load_csv(DATA / FIXTURES[valid_name][1]).astype(np.float32)
    fam = fixture_family(valid_name)
    got = run(fam, left, right, precision=prec)
    err, _, den = errors(fam, left, right, got)
    assert err.max() <= 2e-6
    assert norm_error(got, load_csv(DATA / valid_name), den)[0] <= 3e-6


@pytest.mark.parametrize("prec,tol", [("bf16", 8e-3), ("fp16", 1e-3)])
@pytest.mark.parametrize("family,rows,cols,n,m,dist", [CASES[6], CASES[9], CASES[13]])
def test_single_precision_mfma(prec, tol, family, rows, cols, n, m, dist):
    left = seeded((n, rows, cols), dist, 700 + rows, np.float32)
    right = seeded((m, rows, cols), dist, 800 + cols, np.float32)
    err, _, _ = errors(family, left, right, run(family, left, right, precision=prec))
    assert err.max() <= tol, err.max()


def test_border_frame_is_what_fixes_low_overlap_outputs():
    rows = cols = 96
    left = seeded((1, rows, cols), "u01", 901, np.float32)
    right = seeded((2, rows, cols), "u01", 902, np.float32)
    # border=1: no output is recomputed; interior still meets the bar
    err, _, _ = errors("one_to_many", left, right, run("one_to_many", left, right, precision="bf16x3", border=1))
    dy = np.abs(np.arange(-(rows - 1), rows))[:, None]
    dx = np.abs(np.arange(-(cols - 1), cols))[None, :]
    area = (rows - dy) * (cols - dx)
    assert err[:, area >= 4096].max() <= 2e-6
    err2, _, _ = errors("one_to_many", left, right, run("one_to_many", left, right, precision="bf16x3"))
    assert err2.max() <= 2e-6
    assert err2[:, area < 512].max() < 1e-6  # exact FP32 there


def test_c2_config_bf16x3_subset_and_sum_rule():
    left = seeded((1, 256, 256), "u01", 5, np.float32)
    right = seeded((64, 256, 256), "u01", 6, np.float32)
    got = run("one_to_many", left, right, precision="bf16x3")
    for t in (0, 63):
        err, _, _ = errors("one_to_one", left, right[t:t + 1], got[t:t + 1])
        assert err.max() <= 2e-6
    sums = got.astype(np.float64).reshape(64, -1).sum(axis=1)
    expect = left.astype(np.float64).sum() * right.astype(np.float64).reshape(64, -1).sum(axis=1)
    assert np.max(np.abs(sums - expect) / expect) < 2e-6


def test_c3_config_bf16x3_n_to_m():
    left = seeded((16, 64, 64), "normal", 3, np.float32)
    right = seeded((64, 64, 64), "normal", 4, np.float32)
    err, _, _ = errors("n_to_m", left, right, run("n_to_m", left, right, precision="bf16x3"))
    assert err.max() <= 2e-6


def test_linearity_in_the_right_matrix():
    # size-independent property: corr(L, a*R1 + b*R2) = a*corr(L,R1) + b*corr(L,R2)
    left = seeded((1, 80, 80), "u11", 11, np.float32)
    r1 = seeded((1, 80, 80), "u11", 12, np.float32)
    r2 = seeded((1, 80, 80), "u11", 13, np.float32)
    both = np.concatenate([r1, r2, (0.5 * r1 + 0.25 * r2).astype(np.float32)])
    got = run("one_to_many", left, both, precision="bf16x3").astype(np.float64)
    den = oracle_family("one_to_many", left.astype(np.float64), np.abs(both).astype(np.float64), absval=True)
    lin = 0.5 * got[0] + 0.25 * got[1]
    assert (np.abs(got[2] - lin) / den[2]).max() <= 4e-6


@pytest.mark.parametrize("tile", range(0, 13))
@pytest.mark.parametrize("family,rows,cols,n,m,dist", [CASES[2], CASES[7], CASES[10], CASES[13]])
def test_every_mfma_tile(tile, family, rows, cols, n, m, dist):
    """Tile codes 1-8: the dx-split block tiles (7 and 8 stage two left rows per barrier);
    9-12: the row-split tiles, whose waves share one output tile and take different left rows
    (12 one sub-tile wide at three blocks per CU); 0: the default for the size."""
    left = seeded((n, rows, cols), dist, 1500 + rows, np.float32)
    right = seeded((m, rows, cols), dist, 1600 + cols, np.float32)
    err, _, _ = errors(family, left, right, run(family, left, right, precision="bf16x3", tile=tile))
    assert err.max() <= 2e-6, err.max()


def test_mfma_tile_code_out_of_catalog_is_refused():
    left = seeded((1, 16, 16), "u01", 1, np.float32)
    with pytest.raises(Exception, match="catalog"):
        run("one_to_one", left, left, precision="bf16x3", tile=99)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("family,rows,cols,n,m,border", [
    ("one_to_one", 40, 70, 1, 1, 512), ("one_to_many", 30, 17, 1, 3, 100), ("n_to_mn", 25, 33, 3, 6, 300),
    ("n_to_m", 20, 24, 9, 8, 512), ("n_to_m", 64, 9, 5, 27, 200), ("one_to_many", 300, 5, 1, 2, 512),
    ("one_to_many", 37, 21, 1, 70, 512), ("n_to_m", 19, 30, 2, 64, 300), ("n_to_mn", 13, 45, 1, 65, 512)])
def test_border_frame_outputs_are_rounded_once(prec, family, rows, cols, n, m, border):
    """Outputs with overlap area below the border threshold are the exact sum rounded once to
    FP32: within half an FP32 ulp of the FP64 oracle.  Pair counts of 1 to 135 cover the task
    grid's lanes-per-task choices (1 to 64 pairs per task, a partial last chunk); 70 pairs of
    one left, n_to_m with 64 rights per left and n_to_mn with one left cover the tasks whose
    64 pairs share their left matrix; tall (300 x 5) and wide (37 x 21) shapes put most of the
    frame on the column-major or the row-major copies."""
    left = seeded((n, rows, cols), "u11", 31 + rows, np.float32)
    right = seeded((m, rows, cols), "u11", 37 + cols, np.float32)
    got = run(family, left, right, precision=prec, border=border).astype(np.float64)
    _, ref, den = errors(family, left, right, got)
    dy = rows - np.abs(np.arange(-(rows - 1), rows))[:, None]
    dx = cols - np.abs(np.arange(-(cols - 1), cols))[None, :]
    frame = (dy * dx) < border
    # half an FP32 ulp of the value, plus the FP64 summation slack of both sums (area * 2^-53
    # of the sum of magnitudes each), which matters only where the sum cancels
    tol = np.abs(ref[:, frame]) * 2.0 ** -24 * 1.0001 + (dy * dx)[frame] * 2.0 ** -52 * den[:, frame]
    assert np.all(np.abs(got[:, frame] - ref[:, frame]) <= tol + 1e-45)
    assert (np.abs(got - ref) / np.where(den > 0, den, 1.0)).max() <= 2e-6

"""Pins the CPU oracle (oracle/xcorr_oracle.c) before anything is checked against it.

How each oracle function is pinned:

| oracle function            | reference                         | pinned by                                                    |
|----------------------------|-----------------------------------|--------------------------------------------------------------|
| xco_pair_f64 / _f32        | src/validate.hpp:94-132           | 12 golden fixtures of data/, all-ones closed form, delta     |
|                            |                                   | closed form, symmetry, sum rule, linearity, scipy            |
|                            |                                   | correlate2d 'full' on non-square inputs, python brute force  |
| xco_pair_abs_f64           | (metric denominator)              | equals xco_pair_f64 of |L|,|R|; positive inputs: equals it   |
| xco_family_* orderings     | src/validate.hpp:216-298          | the fixtures of each family; per-result pair checks          |
| xco_num_results            | src/validate.hpp:140-160,249-262  | known counts and rejected counts                             |
| xco_validate_result        | src/validate.hpp:134-188          | hand-computed known answers                                  |
| xco_normalize_fft_f64      | src/fft_helpers.hpp:193-261       | numpy FFT correlation theorem vs direct oracle; hand layout  |
| xco_interleaved_index      | src/matrix.hpp:614-618            | inverse property, known values                               |
| xco_next_iteration_count   | src/stopwatch.hpp:57-71           | hand-computed known answers                                  |
| xco_norm_error             | (north-star metric)               | hand-computed known answers                                  |
"""
import numpy as np
import pytest

from xc_testutil import (DATA, FIXTURES, fixture_family, load_csv, norm_error, normalize_fft,
                         oracle, oracle_family, oracle_pair, n_results, seeded, significant_digits,
                         validate_result)


@pytest.mark.parametrize("valid_name", sorted(FIXTURES))
def test_oracle_matches_golden_fixture(valid_name):
    left_name, right_name = FIXTURES[valid_name]
    left = load_csv(DATA / left_name)
    right = load_csv(DATA / right_name)
    valid = load_csv(DATA / valid_name)
    fam = fixture_family(valid_name)
    got = oracle_family(fam, left, right)
    den = oracle_family(fam, left, right, absval=True)
    assert got.shape == valid.shape
    err_max, _ = norm_error(got, valid, den)
    # 15+ digit files were written from FP64; 7-digit ones from FP32 Matlab gpuArray.
    tol = 1e-12 if significant_digits(DATA / valid_name) >= 15 else 2e-6
    assert err_max <= tol, (valid_name, err_max)


def test_mislabeled_fixture_really_uses_right_id_2():
    left = load_csv(DATA / "ina_10_10_4_1.csv")
    wrong_right = load_csv(DATA / "ina_10_10_4_1.csv")
    valid = load_csv(DATA / "valid_n_to_m_10_10_4_4_1_1.csv")
    got = oracle_family("n_to_m", left, wrong_right)
    den = oracle_family("n_to_m", left, wrong_right, absval=True)
    assert norm_error(got, valid, den)[0] > 1e-2


@pytest.mark.parametrize("rows,cols", [(1, 1), (3, 3), (7, 5), (5, 7), (16, 10), (10, 16)])
def test_all_ones_closed_form(rows, cols):
    out = oracle_pair(np.ones((rows, cols)), np.ones((rows, cols)))
    dy = np.arange(-(rows - 1), rows)[:, None]
    dx = np.arange(-(cols - 1), cols)[None, :]
    expect = (rows - np.abs(dy)) * (cols - np.abs(dx))
    np.testing.assert_array_equal(out, expect)


def test_delta_gives_shifted_target():
    rows, cols = 6, 9
    r = seeded((rows, cols), "normal", 1, np.float64)
    p, q = 2, 5
    left = np.zeros((rows, cols))
    left[p, q] = 1.0
    out = oracle_pair(left, r)
    for dy in range(-(rows - 1), rows):
        for dx in range(-(cols - 1), cols):
            rr, cc = p - dy, q - dx
            expect = r[rr, cc] if (0 <= rr < rows and 0 <= cc < cols) else 0.0
            assert out[dy + rows - 1, dx + cols - 1] == expect


def test_symmetry_sum_rule_linearity():
    rows, cols = 11, 7
    a = seeded((rows, cols), "normal", 2, np.float64)
    b = seeded((rows, cols), "normal", 3, np.float64)
    c = seeded((rows, cols), "normal", 4, np.float64)
    ab = oracle_pair(a, b)
    ba = oracle_pair(b, a)
    np.testing.assert_allclose(ab, ba[::-1, ::-1], rtol=0, atol=1e-12)
    assert abs(ab.sum() - a.sum() * b.sum()) < 1e-9
    lin = oracle_pair(a, 2.0 * b - 3.0 * c)
    np.testing.assert_allclose(lin, 2.0 * ab - 3.0 * oracle_pair(a, c), rtol=0, atol=1e-11)


@pytest.mark.parametrize("rows,cols", [(16, 10), (10, 16), (31, 13), (8, 8)])
def test_equals_scipy_correlate2d_full(rows, cols):
    from scipy.signal import correlate2d
    a = seeded((rows, cols), "u11", rows * 100 + cols, np.float64)
    b = seeded((rows, cols), "u11", rows * 100 + cols + 1, np.float64)
    np.testing.assert_allclose(oracle_pair(a, b), correlate2d(a, b, "full"), rtol=0, atol=1e-12)


def test_python_brute_force_tiny():
    rows, cols = 3, 4
    a = seeded((rows, cols), "normal", 5, np.float64)
    b = seeded((rows, cols), "normal", 6, np.float64)
    expect = np.zeros((2 * rows - 1, 2 * cols - 1))
    for dy in range(-(rows - 1), rows):
        for dx in range(-(cols - 1), cols):
            s = 0.0
            for r in range(rows):
                for c in range(cols):
                    if 0 <= r - dy < rows and 0 <= c - dx < cols:
                        s += a[r, c] * b[r - dy, c - dx]
            expect[dy + rows - 1, dx + cols - 1] = s
    np.testing.assert_allclose(oracle_pair(a, b), expect, rtol=0, atol=1e-13)


def test_abs_denominator():
    a = seeded((9, 6), "u11", 7, np.float64)
    b = seeded((9, 6), "u11", 8, np.float64)
    den = oracle_family(0, a[None], b[None], absval=True)[0]
    np.testing.assert_allclose(den, oracle_pair(np.abs(a), np.abs(b)), rtol=1e-14, atol=0)
    pa, pb = np.abs(a), np.abs(b)
    np.testing.assert_allclose(oracle_family(0, pa[None], pb[None], absval=True)[0], oracle_pair(pa, pb),
                               rtol=1e-14)


def test_f32_oracle_close_to_f64():
    a = seeded((32, 32), "u01", 9, np.float32)
    b = seeded((32, 32), "u01", 10, np.float32)
    o32 = oracle_pair(a, b, dtype=np.float32)
    o64 = oracle_pair(a.astype(np.float64), b.astype(np.float64))
    den = oracle_family(0, a[None].astype(np.float64), b[None].astype(np.float64), absval=True)[0]
    assert norm_error(o32, o64, den)[0] < 5e-6


@pytest.mark.parametrize("family,n,m", [("one_to_many", 1, 3), ("n_to_mn", 2, 6), ("n_to_m", 3, 2)])
def test_family_orderings(family, n, m):
    rows, cols = 5, 4
    left = seeded((n, rows, cols), "normal", 11, np.float64)
    right = seeded((m, rows, cols), "normal", 12, np.float64)
    out = oracle_family(family, left, right)
    for k in range(out.shape[0]):
        if family == "one_to_many":
            l, r = 0, k
        elif family == "n_to_mn":
            l, r = k % n, k
        else:
            l, r = k // m, k % m
        np.testing.assert_array_equal(out[k], oracle_pair(left[l], right[r]))


def test_num_results():
    assert n_results("one_to_one", 1, 1) == 1
    assert n_results("one_to_one", 2, 1) == -1
    assert n_results("one_to_many", 1, 7) == 7
    assert n_results("one_to_many", 2, 7) == -1
    assert n_results("n_to_mn", 2, 6) == 6
    assert n_results("n_to_mn", 4, 6) == -1
    assert n_results("n_to_m", 3, 5) == 15


def test_validate_result_known_answers():
    out = validate_result([1.0, 2.0, 0.0, 1e-6], [1.0, 1.0, 0.0, 0.0])
    assert out[0] == 0.5
    assert out[1] == 0.125
    assert abs(out[2] - np.sqrt(0.0625 - 0.125 ** 2)) < 1e-15
    assert out[3] == 2.0 and out[4] == 1.0
    out = validate_result([2e-5, 3.0], [1e-5, 3.0])
    assert out[0] == 0.5 and out[3] == 2e-5 and out[4] == 1e-5


def test_norm_error_known_answers():
    mx, mean = norm_error([1.0, 2.0, 5.0], [1.5, 2.0, 4.0], [10.0, 1.0, 4.0])
    assert mx == 0.25 and abs(mean - (0.05 + 0.25) / 3) < 1e-15


def test_normalize_fft_hand_layout():
    res = np.arange(16, dtype=np.float64).reshape(4, 4) * 16.0
    norm = normalize_fft(res)
    # rows/cols 4 -> empty row/col index 2; quadrants TL(2x2)->(1,1), TR(1 col)->(0,1) ...
    expect = np.array([[15, 12, 13], [3, 0, 1], [7, 4, 5]], dtype=np.float64)
    np.testing.assert_array_equal(norm, expect)


@pytest.mark.parametrize("rows,cols", [(8, 8), (6, 10), (13, 7)])
def test_normalize_fft_against_correlation_theorem(rows, cols):
    a = seeded((rows, cols), "u11", 20, np.float64)
    b = seeded((rows, cols), "u11", 21, np.float64)
    pa = np.zeros((2 * rows, 2 * cols))
    pb = np.zeros((2 * rows, 2 * cols))
    pa[:rows, :cols] = a
    pb[:rows, :cols] = b
    circ = np.fft.irfft2(np.fft.rfft2(pa) * np.conj(np.fft.rfft2(pb)), s=pa.shape)
    unnormalised = circ * pa.size  # cuFFT/rocFFT inverse transforms are unscaled
    np.testing.assert_allclose(normalize_fft(unnormalised), oracle_pair(a, b), rtol=0, atol=1e-11)


def test_interleaved_index():
    lib = oracle()
    for g, k in [(2, 3), (4, 4), (1, 5), (5, 1)]:
        seen = set()
        for i in range(g * k):
            j = lib.xco_interleaved_index(i, g, k)
            assert lib.xco_interleaved_index(j, k, g) == i
            seen.add(j)
        assert seen == set(range(g * k))
    assert lib.xco_interleaved_index(3, 2, 3) == 1 * 3 + 1


def test_next_iteration_count():
    lib = oracle()
    assert lib.xco_next_iteration_count(1e9, 1e8, 1) == 15  # ceil(1 * min(1.5 * 10, 100))
    assert lib.xco_next_iteration_count(1e9, 1e3, 2) == 200
    assert lib.xco_next_iteration_count(1e9, 9e8, 10) == 17  # ceil(10 * 1.5 * 1e9/9e8)
    assert lib.xco_next_iteration_count(1e9, 1e9, 10) == 0

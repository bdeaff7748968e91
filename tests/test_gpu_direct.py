"""GPU parity of the wave64 direct kernels (FP32 and FP64) and the FFT method against the
CPU oracle, through the C ABI.

Bars (normalised error = |x - x64| / corr(|L|,|R|), x64 = FP64 oracle on the same inputs):
  direct FP32 <= 1e-6;  direct FP64 <= 1e-12;  FFT with FP64 transform <= 1e-6 (FP32 output);
  FFT with FP32 transform: interior only (overlap >= a quarter of the matrix) <= 2e-6.
"""
import numpy as np
import pytest

from xc_testutil import (DATA, FIXTURES, fixture_family, load_csv, norm_error, normalize_fft, oracle_family,
                         seeded, significant_digits, xcorr)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(family, left, right, **kw):
    out = xcorr().cross_corr(family, dev(left), dev(right), **kw)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def check(family, left, right, got, tol, dtype=np.float64):
    ref = oracle_family(family, left.astype(np.float64), right.astype(np.float64))
    den = oracle_family(family, left.astype(np.float64), right.astype(np.float64), absval=True)
    assert got.shape == ref.shape
    err, mean = norm_error(got, ref, den)
    assert err <= tol, (family, left.shape, right.shape, err, mean)
    return err


@pytest.mark.parametrize("valid_name", sorted(FIXTURES))
def test_direct_fp32_golden(valid_name):
    left = load_csv(DATA / FIXTURES[valid_name][0]).astype(np.float32)
    right = load_csv(DATA / FIXTURES[valid_name][1]).astype(np.float32)
    fam = fixture_family(valid_name)
    got = run(fam, left, right, precision="fp32")
    check(fam, left, right, got, 1e-6)
    valid = load_csv(DATA / valid_name)
    den = oracle_family(fam, left.astype(np.float64), right.astype(np.float64), absval=True)
    tol = 1e-6 if significant_digits(DATA / valid_name) >= 15 else 3e-6
    assert norm_error(got, valid, den)[0] <= tol


@pytest.mark.parametrize("valid_name", sorted(FIXTURES))
def test_direct_fp64_golden(valid_name):
    left = load_csv(DATA / FIXTURES[valid_name][0])
    right = load_csv(DATA / FIXTURES[valid_name][1])
    fam = fixture_family(valid_name)
    got = run(fam, left, right, precision="fp32")  # float64 data selects the FP64 kernel
    check(fam, left, right, got, 1e-12)
    if significant_digits(DATA / valid_name) >= 15:
        den = oracle_family(fam, left, right, absval=True)
        assert norm_error(got, load_csv(DATA / valid_name), den)[0] <= 1e-12


CASES = [
    # family, rows, cols, n, m, distribution
    ("one_to_one", 1, 1, 1, 1, "u01"),
    ("one_to_one", 2, 3, 1, 1, "normal"),
    ("one_to_one", 7, 5, 1, 1, "u11"),
    ("one_to_one", 16, 10, 1, 1, "u01"),
    ("one_to_one", 10, 16, 1, 1, "u01"),
    ("one_to_one", 128, 128, 1, 1, "u01"),
    ("one_to_one", 313, 125, 1, 1, "u01"),
    ("one_to_one", 125, 313, 1, 1, "normal"),
    ("one_to_one", 3, 600, 1, 1, "u11"),
    ("one_to_many", 33, 65, 1, 5, "u01"),
    ("one_to_many", 64, 64, 1, 12, "u11"),
    ("n_to_mn", 12, 9, 3, 9, "normal"),
    ("n_to_mn", 32, 32, 2, 6, "u01"),
    ("n_to_m", 64, 64, 4, 6, "normal"),
    ("n_to_m", 17, 23, 3, 5, "u11"),
]


@pytest.mark.parametrize("family,rows,cols,n,m,dist", CASES)
def test_direct_fp32_seeded(family, rows, cols, n, m, dist):
    left = seeded((n, rows, cols), dist, 100 + rows, np.float32)
    right = seeded((m, rows, cols), dist, 200 + cols, np.float32)
    check(family, left, right, run(family, left, right, precision="fp32"), 1e-6)


@pytest.mark.parametrize("family,rows,cols,n,m,dist", CASES[:8] + CASES[11:14])
def test_direct_fp64_seeded(family, rows, cols, n, m, dist):
    left = seeded((n, rows, cols), dist, 300 + rows, np.float64)
    right = seeded((m, rows, cols), dist, 400 + cols, np.float64)
    check(family, left, right, run(family, left, right, precision="fp32"), 1e-12)


@pytest.mark.parametrize("family,rows,cols,n,m,dist", CASES[:8] + CASES[11:14])
def test_double_auto_seeded(family, rows, cols, n, m, dist):
    """Double data under AUTO: the FP64 direct kernel or the own FFT in FP64, whichever the
    library measures faster for the workload; the FFT's rounding is global (about 1e-16 of the
    largest output), so the bar is the FFT's double bar, 1e-9."""
    left = seeded((n, rows, cols), dist, 300 + rows, np.float64)
    right = seeded((m, rows, cols), dist, 400 + cols, np.float64)
    check(family, left, right, run(family, left, right), 1e-9)


@pytest.mark.parametrize("split", [1, 2, 3, 8, 64])
def test_direct_split_reduction(split):
    left = seeded((1, 96, 80), "u01", 7, np.float32)
    right = seeded((3, 96, 80), "u01", 8, np.float32)
    check("one_to_many", left, right, run("one_to_many", left, right, precision="fp32", split=split), 1e-6)


# Register tiles of the direct kernel: every request of the grid below is rounded to a
# compiled tile (named in the launch descriptor) and must give the oracle's results, with
# partial groups of left and right matrices, outputs wider than one wave and every family.
TILE_SHAPES = [("n_to_m", 19, 37, 3, 5), ("one_to_many", 33, 70, 1, 7), ("n_to_mn", 20, 9, 2, 6),
               ("one_to_one", 11, 13, 1, 1)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ty", [1, 2, 4, 8, 16])
def test_direct_tile_catalog(ty, dtype):
    tol = 1e-6 if dtype == np.float32 else 1e-12
    used = set()
    for rm in (1, 2, 4, 8):
        for lm in (1, 2, 4):
            for jb in (1, 2, 4):
                for family, rows, cols, n, m in TILE_SHAPES:
                    left = seeded((n, rows, cols), "normal", 500 + rows, dtype)
                    right = seeded((m, rows, cols), "normal", 600 + cols, dtype)
                    print(f"tile request ty{ty} rm{rm} lm{lm} jb{jb} {family} {dtype.__name__}", flush=True)
                    got = run(family, left, right, precision="fp32", shifts_per_thread=ty, right_mats_per_thread=rm,
                              left_mats_per_thread=lm, rows_per_step=jb)
                    used.add(xcorr().last_launch()["kernel"])
                    check(family, left, right, got, tol)
    assert len(used) >= 3, used


@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_direct_waves_per_block(waves):
    for family, rows, cols, n, m in TILE_SHAPES + [("one_to_many", 70, 300, 1, 3)]:
        left = seeded((n, rows, cols), "u11", 700 + rows, np.float32)
        right = seeded((m, rows, cols), "u11", 800 + cols, np.float32)
        print(f"waves {waves} {family} {rows}x{cols}", flush=True)
        got = run(family, left, right, precision="fp32", waves_per_block=waves)
        assert xcorr().last_launch()["block"][0] == 64 * waves
        check(family, left, right, got, 1e-6)


@pytest.mark.parametrize("rows_per_block", [1, 3, 8, 40])
def test_direct_triangle_distribution(rows_per_block):
    # blocks take at most rows_per_block overlap rows (more blocks where the overlap is larger)
    for dtype, tol in ((np.float32, 1e-6), (np.float64, 1e-12)):
        left = seeded((2, 45, 30), "u01", 17, dtype)
        right = seeded((3, 45, 30), "u01", 18, dtype)
        for jb in (1, 4):
            got = run("n_to_m", left, right, precision="fp32", rows_per_block=rows_per_block, rows_per_step=jb)
            assert xcorr().last_launch()["split"] == (45 + max(rows_per_block, 1) - 1) // max(rows_per_block, 1) or jb > 1
            check("n_to_m", left, right, got, tol)


# The lanes-over-pairs kernel (FP32 tile 2, and the default where a left has 48 or more rights
# and the batch makes enough waves): a wave takes one left and 64 of its rights, one pair per
# lane; partial last waves of a left, several lefts (n_to_m, and n_to_mn whose rights of a
# left are n apart), few rights per left, outputs narrower and wider than a tile, one row or
# column.
PAIRS_SHAPES = [("one_to_many", 1, 1, 1, 64), ("one_to_many", 7, 5, 1, 70), ("one_to_many", 33, 65, 1, 128),
                ("n_to_m", 17, 23, 2, 64), ("n_to_mn", 12, 30, 1, 65), ("one_to_many", 3, 200, 1, 64),
                ("one_to_many", 150, 9, 1, 64), ("one_to_many", 1, 37, 1, 66), ("n_to_m", 40, 1, 3, 128),
                ("n_to_m", 19, 21, 3, 100), ("n_to_mn", 14, 18, 3, 210), ("n_to_mn", 9, 40, 2, 128),
                ("n_to_m", 5, 6, 2, 7), ("n_to_mn", 6, 5, 4, 8)]


@pytest.mark.parametrize("family,rows,cols,n,m", PAIRS_SHAPES)
def test_direct_lanes_over_pairs(family, rows, cols, n, m):
    x = xcorr()
    left = seeded((n, rows, cols), "u11", 900 + rows, np.float32)
    right = seeded((m, rows, cols), "u11", 950 + cols, np.float32)
    print(f"lanes over pairs {family} {rows}x{cols} n={n} m={m}", flush=True)
    got = run(family, left, right, precision="fp32", tile=2)
    assert x.last_launch()["kernel"].startswith("direct_pairs_kernel")
    check(family, left, right, got, 1e-6)
    default = run(family, left, right, precision="fp32")  # either kernel, by the batch's size
    check(family, left, right, default, 1e-6)
    tiles = run(family, left, right, precision="fp32", tile=1)
    assert x.last_launch()["kernel"].startswith("direct_kernel")
    check(family, left, right, tiles, 1e-6)


@pytest.mark.parametrize("family,rows,cols,n,m", PAIRS_SHAPES[1:6] + PAIRS_SHAPES[9:12])
def test_direct_lanes_over_pairs_double(family, rows, cols, n, m):
    """The same kernel on double data (FP64 FMAs, the FP64 direct method's bar)."""
    x = xcorr()
    left = seeded((n, rows, cols), "normal", 990 + rows, np.float64)
    right = seeded((m, rows, cols), "normal", 995 + cols, np.float64)
    print(f"lanes over pairs f64 {family} {rows}x{cols} n={n} m={m}", flush=True)
    got = run(family, left, right, precision="fp32", tile=2)
    assert x.last_launch()["kernel"].startswith("direct_pairs_kernel<f64>")
    check(family, left, right, got, 1e-12)
    tiles = run(family, left, right, precision="fp32", tile=1)
    assert x.last_launch()["kernel"].startswith("direct_kernel<f64>")
    check(family, left, right, tiles, 1e-12)


def test_direct_lanes_over_pairs_default_and_refusal():
    x = xcorr()
    # small batches: the default is the register-tile kernel
    for family, n, m in (("one_to_many", 1, 63), ("n_to_m", 2, 32), ("n_to_mn", 2, 128)):
        left = seeded((n, 9, 9), "u01", 1, np.float32)
        right = seeded((m, 9, 9), "u01", 2, np.float32)
        check(family, left, right, run(family, left, right, precision="fp32"), 1e-6)
        assert x.last_launch()["kernel"].startswith("direct_kernel")
    # one_to_one has one right per left: tile 2 is refused
    one = dev(seeded((1, 9, 9), "u01", 1, np.float32))
    with pytest.raises(x.XcError) as e:
        x.cross_corr("one_to_one", one, one, precision="fp32", tile=2)
    assert e.value.status == 4
    with pytest.raises(x.XcError) as e:
        x.cross_corr("one_to_one", one, one, precision="fp32", tile=3)
    assert e.value.status == 1
    # a large batch with 64 rights per left: the default is the lanes-over-pairs kernel
    left = seeded((2, 64, 64), "u11", 3, np.float32)
    right = seeded((64, 64, 64), "u11", 4, np.float32)
    check("n_to_m", left, right, run("n_to_m", left, right, precision="fp32"), 1e-6)
    assert x.last_launch()["kernel"].startswith("direct_pairs_kernel")


def test_direct_split_and_triangle_are_exclusive():
    x = xcorr()
    left = dev(seeded((1, 8, 8), "u01", 1, np.float32))
    with pytest.raises(x.XcError, match="not both"):
        x.cross_corr("one_to_one", left, left, precision="fp32", split=2, rows_per_block=3)


def test_c1_config_one_to_one_128():
    left = seeded((1, 128, 128), "u01", 1, np.float32)
    right = seeded((1, 128, 128), "u01", 2, np.float32)
    check("one_to_one", left, right, run("one_to_one", left, right, precision="fp32"), 1e-6)


def test_c3_config_n_to_m_16x64_at_64():
    left = seeded((16, 64, 64), "normal", 3, np.float32)
    right = seeded((64, 64, 64), "normal", 4, np.float32)
    check("n_to_m", left, right, run("n_to_m", left, right, precision="fp32"), 1e-6)
    l64, r64 = left.astype(np.float64), right.astype(np.float64)
    check("n_to_m", l64, r64, run("n_to_m", l64, r64, precision="fp32"), 1e-12)  # the FP64 direct kernel
    check("n_to_m", l64, r64, run("n_to_m", l64, r64), 1e-9)  # AUTO: the own FFT here


def test_c2_config_one_to_many_256_subset_and_sum_rule():
    left = seeded((1, 256, 256), "u01", 5, np.float32)
    right = seeded((64, 256, 256), "u01", 6, np.float32)
    got = run("one_to_many", left, right, precision="fp32")
    # full oracle on two of the 64 results, sum rule on all 64
    for t in (0, 37):
        check("one_to_one", left, right[t:t + 1], got[t:t + 1], 1e-6)
    sums = got.astype(np.float64).reshape(64, -1).sum(axis=1)
    expect = left.astype(np.float64).sum() * right.astype(np.float64).reshape(64, -1).sum(axis=1)
    assert np.max(np.abs(sums - expect) / expect) < 1e-6


def test_abi_mirror_of_reference_launchers():
    x = xcorr()
    left = seeded((1, 40, 24), "u01", 9, np.float32)
    right = seeded((3, 40, 24), "u01", 10, np.float32)
    l, r = dev(left), dev(right)
    out = torch.empty((3, 79, 47), device="cuda")
    import ctypes
    st = x.lib().xc_run_ccn_shuffle_one_to_many_multirow_both_multimat_right_f32(
        ctypes.c_void_p(l.data_ptr()), ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(out.data_ptr()),
        24, 40, 47, 79, 3, 8, 2, 2, 2, None)
    x.check(st)
    torch.cuda.synchronize()
    check("one_to_many", left, right, out.cpu().numpy(), 1e-6)
    out1 = torch.empty((1, 79, 47), device="cuda")
    x.check(x.lib().xc_run_ccn_shuffle_f32(ctypes.c_void_p(l.data_ptr()), ctypes.c_void_p(r.data_ptr()),
                                           ctypes.c_void_p(out1.data_ptr()), 24, 40, 47, 79, 8, None))
    torch.cuda.synchronize()
    check("one_to_one", left, right[:1], out1.cpu().numpy(), 1e-6)
    assert x.last_launch()["kernel"].startswith("direct_kernel")


# ---------------------------------------------------------------- FFT method
FFT_CASES = [("one_to_one", 16, 16, 1, 1), ("one_to_one", 13, 7, 1, 1), ("one_to_many", 32, 32, 1, 5),
             ("n_to_mn", 20, 24, 2, 4), ("n_to_m", 16, 12, 3, 4), ("one_to_many", 128, 128, 1, 3)]


@pytest.mark.parametrize("family,rows,cols,n,m", FFT_CASES)
def test_fft64_normalised(family, rows, cols, n, m):
    left = seeded((n, rows, cols), "u01", 11, np.float32)
    right = seeded((m, rows, cols), "u01", 12, np.float32)
    got = xcorr().fft_cross_corr(family, dev(left), dev(right), normalize=True, fp64_transform=True)
    torch.cuda.synchronize()
    check(family, left, right, got.cpu().numpy(), 1e-6)


@pytest.mark.parametrize("family,rows,cols,n,m", FFT_CASES)
def test_fft_double_data(family, rows, cols, n, m):
    left = seeded((n, rows, cols), "normal", 13, np.float64)
    right = seeded((m, rows, cols), "normal", 14, np.float64)
    got = xcorr().fft_cross_corr(family, dev(left), dev(right), normalize=True)
    torch.cuda.synchronize()
    check(family, left, right, got.cpu().numpy(), 1e-10)


@pytest.mark.parametrize("family,rows,cols,n,m", FFT_CASES[:4])
def test_fft_unnormalised_layout_matches_reference_normalisation(family, rows, cols, n, m):
    left = seeded((n, rows, cols), "u11", 15, np.float64)
    right = seeded((m, rows, cols), "u11", 16, np.float64)
    raw = xcorr().fft_cross_corr(family, dev(left), dev(right), normalize=False).cpu().numpy()
    assert raw.shape[1:] == (2 * rows, 2 * cols)
    normed = np.stack([normalize_fft(r) for r in raw])
    check(family, left, right, normed, 1e-10)


def test_fft32_interior():
    rows = cols = 64
    left = seeded((1, rows, cols), "u01", 17, np.float32)
    right = seeded((4, rows, cols), "u01", 18, np.float32)
    got = xcorr().fft_cross_corr("one_to_many", dev(left), dev(right)).cpu().numpy()
    ref = oracle_family("one_to_many", left.astype(np.float64), right.astype(np.float64))
    den = oracle_family("one_to_many", left.astype(np.float64), right.astype(np.float64), absval=True)
    dy = np.abs(np.arange(-(rows - 1), rows))[:, None]
    dx = np.abs(np.arange(-(cols - 1), cols))[None, :]
    interior = ((rows - dy) * (cols - dx)) >= rows * cols // 4
    err = np.abs(got - ref) / den
    assert err[:, interior].max() <= 2e-6


def test_fft_building_blocks():
    x = xcorr()
    import ctypes
    src = seeded((2, 5, 3), "normal", 19, np.float32)
    dst = torch.full((2, 8, 7), 7.0, device="cuda")
    s = dev(src)
    x.check(x.lib().xc_scatter_f32(ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(dst.data_ptr()), 5, 3, 2, 8, 7,
                                   2, 1, None))
    expect = np.zeros((2, 8, 7), np.float32)
    expect[:, 2:7, 1:4] = src
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dst.cpu().numpy(), expect)
    ref = seeded((10, 2), "normal", 20, np.float32)
    de = seeded((3, 10, 2), "normal", 21, np.float32)
    d = dev(de)
    x.check(x.lib().xc_hadamard_f32(ctypes.c_void_p(dev(ref).data_ptr()), ctypes.c_void_p(d.data_ptr()), 10, 3, None))
    rc = ref[:, 0] + 1j * ref[:, 1]
    dc = de[..., 0] + 1j * de[..., 1]
    prod = rc[None] * np.conj(dc)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    np.testing.assert_allclose(got[..., 0], prod.real, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(got[..., 1], prod.imag, rtol=1e-5, atol=1e-6)
    res = seeded((2, 6, 8), "normal", 22, np.float64)
    norm = torch.empty((2, 5, 7), dtype=torch.float64, device="cuda")
    x.check(x.lib().xc_normalize_fft_f64(ctypes.c_void_p(dev(res).data_ptr()), ctypes.c_void_p(norm.data_ptr()),
                                         6, 8, 2, None))
    torch.cuda.synchronize()
    for b in range(2):
        np.testing.assert_allclose(norm.cpu().numpy()[b], normalize_fft(res[b]), rtol=1e-15, atol=0)

"""The allocation-free entry point xc_cross_corr_noalloc_f32 against the entry point it varies,
xc_cross_corr_f32 with precision FFT64 tile 1 (the same own-FFT kernels on allocated
workspaces), through the C ABI: equal results on calls in a row with different sizes (the
second smaller than the first), calls split into parts because their spectra exceed the fixed
scratch array, and a refused size.  Results are also checked against the FP64 oracle
(normalised error <= 1e-6) where the oracle is fast.
"""
import ctypes

import numpy as np
import pytest

from xc_testutil import norm_error, oracle_family, seeded, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def noalloc(family, left, right):
    x = xcorr()
    fam = x.FAMILIES[family]
    n, rows, cols = left.shape
    m = right.shape[0]
    out = torch.full(x.result_shape(fam, rows, cols, n, m), float("nan"), device="cuda")
    print(f"noalloc {family} {n}x{m} of {rows}x{cols}", flush=True)
    x.check(x.lib().xc_cross_corr_noalloc_f32(fam, ctypes.c_void_p(left.data_ptr()), ctypes.c_void_p(right.data_ptr()),
                                              ctypes.c_void_p(out.data_ptr()), rows, cols, n, m,
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out


def regular(family, left, right):
    return xcorr().cross_corr(family, left, right, precision="fft64", tile=1)


def inputs(family, rows, cols, n, m, seed):
    left = torch.from_numpy(seeded((n, rows, cols), "u01", seed, np.float32)).cuda()
    right = torch.from_numpy(seeded((m, rows, cols), "u01", seed + 1, np.float32)).cuda()
    return left, right


def check_oracle(family, left, right, got, picks):
    l64, r64 = left.cpu().numpy().astype(np.float64), right.cpu().numpy().astype(np.float64)
    n, m = l64.shape[0], r64.shape[0]
    for k in picks:
        li, ri = {"one_to_one": (0, 0), "one_to_many": (0, k), "n_to_mn": (k % n, k), "n_to_m": (k // m, k % m)}[family]
        ref = oracle_family("one_to_one", l64[li:li + 1], r64[ri:ri + 1])
        den = oracle_family("one_to_one", l64[li:li + 1], r64[ri:ri + 1], absval=True)
        err, _ = norm_error(got[k:k + 1].cpu().numpy(), ref, den)
        assert err <= 1e-6, (family, k, err)


def test_two_calls_in_a_row_the_second_smaller():
    a = inputs("one_to_many", 256, 256, 1, 64, 11)
    b = inputs("n_to_m", 40, 33, 3, 5, 13)
    got_a = noalloc("one_to_many", *a)
    got_b = noalloc("n_to_m", *b)
    torch.cuda.synchronize()
    assert torch.equal(got_a, regular("one_to_many", *a))
    assert torch.equal(got_b, regular("n_to_m", *b))
    check_oracle("one_to_many", *a, got_a, [0, 63])
    check_oracle("n_to_m", *b, got_b, range(15))


@pytest.mark.parametrize("family,rows,cols,n,m", [
    ("one_to_many", 256, 256, 1, 150),   # 476 MB of spectra: two parts of rights
    ("n_to_mn", 512, 512, 2, 40),        # parts of whole groups of 2 rights
    ("n_to_m", 512, 512, 2, 40),         # the rights alone exceed the array: parts per left
    ("n_to_m", 64, 64, 16, 64),          # fits at once
    ("one_to_one", 1024, 1024, 1, 1),    # 96 MiB for the single pair
])
def test_calls_split_into_parts(family, rows, cols, n, m):
    left, right = inputs(family, rows, cols, n, m, 17)
    got = noalloc(family, left, right)
    want = regular(family, left, right)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    k = got.shape[0]
    if rows <= 256:
        check_oracle(family, left, right, got, sorted({0, k // 2, k - 1}))


def test_size_beyond_the_scratch_array_is_refused():
    x = xcorr()
    left = torch.zeros((1, 2048, 2048), device="cuda")
    out = torch.zeros((1, 4095, 4095), device="cuda")
    st = x.lib().xc_cross_corr_noalloc_f32(0, ctypes.c_void_p(left.data_ptr()), ctypes.c_void_p(left.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), 2048, 2048, 1, 1, None)
    assert st == 4 and "scratch" in x.lib().xc_last_error().decode()
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0  # nothing was launched

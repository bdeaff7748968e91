"""CPU-side checks of the C ABI: the library loads, exports every declared symbol, and
validates arguments on the host before touching a GPU (error paths of
src/run_args.cpp:114-117, src/one_to_one.hpp:20-22, src/n_to_mn.hpp:58-67,
src/naive_shuffle.cu:461-463)."""
import ctypes

import pytest

from xc_testutil import header_symbols, xcorr


def test_every_declared_symbol_is_exported():
    lib = xcorr().lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(xcorr().SIGNATURES), set(syms) ^ set(xcorr().SIGNATURES)


def test_num_results():
    lib = xcorr().lib()
    assert lib.xc_num_results(0, 1, 1) == 1
    assert lib.xc_num_results(0, 2, 1) == -1
    assert lib.xc_num_results(1, 1, 64) == 64
    assert lib.xc_num_results(2, 2, 6) == 6
    assert lib.xc_num_results(2, 4, 6) == -1
    assert lib.xc_num_results(3, 16, 64) == 1024
    assert lib.xc_num_results(7, 1, 1) == -1


def _call_f32(family, rows, cols, n, m, precision=0, ptr=1):
    x = xcorr()
    p = ctypes.c_void_p(ptr)
    opt = x.Options(precision, 0, 0, 0)
    return x.lib().xc_cross_corr_f32(family, p, p, p, rows, cols, n, m, ctypes.byref(opt), None)


@pytest.mark.parametrize("family,rows,cols,n,m,needle", [
    (0, 0, 4, 1, 1, "positive"),
    (0, 4, 4, 2, 1, "one_to_one"),
    (1, 4, 4, 2, 3, "one_to_many"),
    (2, 4, 4, 4, 6, "multiple"),
    (9, 4, 4, 1, 1, "family"),
])
def test_invalid_arguments_rejected_on_host(family, rows, cols, n, m, needle):
    x = xcorr()
    st = _call_f32(family, rows, cols, n, m)
    assert st == 1
    assert needle in x.lib().xc_last_error().decode()


def test_null_pointers_rejected():
    assert _call_f32(0, 4, 4, 1, 1, ptr=0) == 1


def test_unknown_precision_rejected():
    x = xcorr()
    st = _call_f32(0, 4, 4, 1, 1, precision=99)
    # validated on the host before any launch
    assert st == 1 and "precision" in x.lib().xc_last_error().decode()


def test_reference_tuning_limits():
    x = xcorr()
    p = ctypes.c_void_p(1)
    st = x.lib().xc_run_ccn_shuffle_f32(p, p, p, 8, 8, 15, 15, 33, None)
    assert st == 1 and "warps" in x.lib().xc_last_error().decode()
    st = x.lib().xc_run_ccn_shuffle_one_to_many_multirow_both_multimat_right_f32(
        p, p, p, 8, 8, 15, 15, 4, 8, 0, 2, 2, None)
    assert st == 1 and "shifts" in x.lib().xc_last_error().decode()


def test_version():
    assert b"gfx950" in xcorr().lib().xc_version()

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


def test_python_binding_rejects_bad_inputs_before_the_library():
    # Host-side checks of cross_corr / fft_cross_corr: dtype, size, device and a
    # caller-supplied `out` (shape, dtype, device, contiguity) are refused with XcError.
    torch = pytest.importorskip("torch")
    x = xcorr()
    a = torch.zeros((1, 4, 5))
    for fn in (x.cross_corr, x.fft_cross_corr):
        with pytest.raises(x.XcError, match="float32 or both float64"):
            fn("one_to_one", a.half(), a.half())
        with pytest.raises(x.XcError, match="float32 or both float64"):
            fn("one_to_one", a, a.double())
        with pytest.raises(x.XcError, match="float32 or both float64"):
            fn("one_to_one", a.int(), a.int())
        with pytest.raises(x.XcError, match="same size"):
            fn("one_to_one", a, torch.zeros((1, 4, 4)))
        with pytest.raises(x.XcError, match="GPU tensors"):
            fn("one_to_one", a, a)
        with pytest.raises(x.XcError, match="torch tensors"):
            fn("one_to_one", a.numpy(), a.numpy())
        with pytest.raises(x.XcError, match=r"\(count, rows, cols\)"):
            fn("one_to_one", a[None], a[None])


def test_python_binding_checks_a_given_out():
    torch = pytest.importorskip("torch")
    x = xcorr()
    like = torch.zeros((1, 4, 5))
    shape = (1, 7, 9)
    ok = torch.zeros(shape)
    assert x._check_out(ok, shape, like) is ok
    assert x._check_out(None, shape, like).shape == shape
    with pytest.raises(x.XcError, match="shape"):
        x._check_out(torch.zeros((1, 7, 8)), shape, like)
    with pytest.raises(x.XcError, match="shape"):
        x._check_out(torch.zeros((1, 9, 7)).transpose(1, 2), shape, like)
    with pytest.raises(x.XcError, match="dtype"):
        x._check_out(torch.zeros(shape, dtype=torch.float64), shape, like)
    with pytest.raises(x.XcError, match="dtype"):
        x._check_out(ok.numpy(), shape, like)


def test_device_checks_refuse_bad_inputs():
    """The on-device error checks validate their tensors before any GPU call."""
    torch = pytest.importorskip("torch")
    x = xcorr()
    a = torch.zeros(4, dtype=torch.float64)
    with pytest.raises(x.XcError, match="GPU"):
        x.validate_result(a, a)
    with pytest.raises(x.XcError, match="GPU"):
        x.norm_error(a, a, a)
    with pytest.raises(x.XcError, match="torch"):
        x.validate_result(a.numpy(), a)


"""Python binding of libxcorr.so (the MI355X cross-correlation library) over its C ABI.

The C ABI (include/xcorr.h) is the product boundary; this module only loads the shared
object with ctypes and passes device pointers and the current HIP stream.  PyTorch is
used for device memory and streams.  There is no CPU fallback: if the library or a GPU
is missing, calls raise.

The directory name contains a hyphen, so import it by path:

    import importlib.util, pathlib
    spec = importlib.util.spec_from_file_location("cross_corr_amd", "<repo>/cross-corr_amd/__init__.py")
"""
import ctypes
import pathlib

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libxcorr.so"

ONE_TO_ONE, ONE_TO_MANY, N_TO_MN, N_TO_M = 0, 1, 2, 3
FAMILIES = {"one_to_one": ONE_TO_ONE, "one_to_many": ONE_TO_MANY, "n_to_mn": N_TO_MN, "n_to_m": N_TO_M}

PREC_AUTO, PREC_FP32, PREC_BF16X3, PREC_BF16, PREC_FP16, PREC_FFT64, PREC_FFT32, PREC_FP16X3 = range(8)
PRECISIONS = {"auto": PREC_AUTO, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "bf16": PREC_BF16,
              "fp16": PREC_FP16, "fft64": PREC_FFT64, "fft32": PREC_FFT32, "fp16x3": PREC_FP16X3}

STATUS = {0: "XC_OK", 1: "XC_ERR_INVALID_ARGUMENT", 2: "XC_ERR_HIP", 3: "XC_ERR_FFT",
          4: "XC_ERR_UNSUPPORTED", 5: "XC_ERR_NO_MEMORY"}


class XcError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_uint32), ("tile", ctypes.c_uint32), ("split", ctypes.c_uint32),
                ("border", ctypes.c_uint32), ("shifts_per_thread", ctypes.c_uint32),
                ("right_mats_per_thread", ctypes.c_uint32), ("left_mats_per_thread", ctypes.c_uint32),
                ("rows_per_step", ctypes.c_uint32), ("waves_per_block", ctypes.c_uint32),
                ("rows_per_block", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 2)]


class Pair(ctypes.Structure):
    _fields_ = [("left_off", ctypes.c_uint64), ("right_off", ctypes.c_uint64), ("out_off", ctypes.c_uint64),
                ("rows", ctypes.c_uint32), ("cols", ctypes.c_uint32)]


class Workload(ctypes.Structure):
    _fields_ = [("family", ctypes.c_uint32), ("rows", ctypes.c_uint32), ("cols", ctypes.c_uint32),
                ("n", ctypes.c_uint32), ("m", ctypes.c_uint32)]


class TrainOptions(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("lr", ctypes.c_float), ("weight_decay", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("hidden", ctypes.c_uint32), ("use_graph", ctypes.c_int),
                ("reserved", ctypes.c_uint32 * 2)]


class MlpShape(ctypes.Structure):
    _fields_ = [("layers", ctypes.c_uint32), ("dims", ctypes.c_uint32 * 8)]

    @classmethod
    def of(cls, dims):
        s = cls()
        s.layers = len(dims) - 1
        for i, d in enumerate(dims):
            s.dims[i] = d
        return s


class LaunchDesc(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_char * 96), ("grid", ctypes.c_uint32 * 3), ("block", ctypes.c_uint32 * 3),
                ("lds_bytes", ctypes.c_uint32), ("rows", ctypes.c_uint32), ("cols", ctypes.c_uint32),
                ("pairs", ctypes.c_uint32), ("split", ctypes.c_uint32)]


# name -> (restype, argtypes); every symbol declared in include/xcorr.h.
_vp, _u32, _u64, _i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
_st = ctypes.c_int
SIGNATURES = {
    "xc_cross_corr_f32": (_st, [_i32, _vp, _vp, _vp, _u32, _u32, _u32, _u32, ctypes.POINTER(Options), _vp]),
    "xc_cross_corr_f64": (_st, [_i32, _vp, _vp, _vp, _u32, _u32, _u32, _u32, ctypes.POINTER(Options), _vp]),
    "xc_cross_corr_noalloc_f32": (_st, [_i32, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _vp]),
    "xc_num_results": (ctypes.c_int64, [_i32, _u32, _u32]),
    "xc_batch_f32": (_st, [ctypes.POINTER(Pair), _u32, _vp, _vp, _vp, ctypes.POINTER(Options), _vp]),
    "xc_batch_plan_create_f32": (_st, [ctypes.POINTER(Pair), _u32, ctypes.POINTER(Options),
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "xc_batch_plan_exec_f32": (_st, [_vp, _vp, _vp, _vp, _i32, _vp]),
    "xc_batch_plan_destroy": (_st, [_vp]),
    "xc_batch_autotune_f32": (_st, [ctypes.POINTER(Pair), _u32, _vp, _vp, _vp, ctypes.POINTER(Options),
                                    ctypes.POINTER(ctypes.c_float), _u32, ctypes.POINTER(_u32),
                                    ctypes.POINTER(Options), _vp]),
    "xc_scatter_f32": (_st, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _vp]),
    "xc_scatter_f64": (_st, [_vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _vp]),
    "xc_hadamard_f32": (_st, [_vp, _vp, _u64, _u32, _vp]),
    "xc_hadamard_f64": (_st, [_vp, _vp, _u64, _u32, _vp]),
    "xc_hadamard_n_to_m_f32": (_st, [_vp, _vp, _vp, _u64, _u32, _u32, _vp]),
    "xc_hadamard_n_to_m_f64": (_st, [_vp, _vp, _vp, _u64, _u32, _u32, _vp]),
    "xc_normalize_fft_f32": (_st, [_vp, _vp, _u32, _u32, _u32, _vp]),
    "xc_normalize_fft_f64": (_st, [_vp, _vp, _u32, _u32, _u32, _vp]),
    "xc_fft_cross_corr_f32": (_st, [_i32, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _i32, _i32, _vp]),
    "xc_fft_cross_corr_f64": (_st, [_i32, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _i32, _vp]),
    "xc_run_ccn_shuffle_f32": (_st, [_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _vp]),
    "xc_run_ccn_shuffle_one_to_many_multirow_both_multimat_right_f32":
        (_st, [_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _vp]),
    "xc_num_candidates": (_u32, []),
    "xc_candidate": (_st, [_u32, ctypes.POINTER(Options), ctypes.c_char_p, ctypes.c_size_t]),
    "xc_workload_features": (_st, [ctypes.POINTER(Workload), ctypes.POINTER(ctypes.c_float), _u32,
                                   ctypes.POINTER(ctypes.c_uint32)]),
    "xc_autotune_f32": (_st, [_i32, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, ctypes.POINTER(ctypes.c_float),
                              ctypes.POINTER(ctypes.c_uint32), _vp]),
    "xc_time_candidate_f32": (_st, [_i32, _vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32, _u32,
                                    ctypes.POINTER(ctypes.c_float), _vp]),
    "xc_autotune_clear": (_st, []),
    "xc_cost_model_load": (_st, [ctypes.c_char_p]),
    "xc_cost_model_predict": (_st, [ctypes.POINTER(Workload), ctypes.POINTER(ctypes.c_float)]),
    "xc_select": (_st, [ctypes.POINTER(Workload), ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_int)]),
    "xc_cost_model_fit": (_st, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(TrainOptions), _u32]),
    "xc_mlp_num_params": (_u64, [ctypes.POINTER(MlpShape)]),
    "xc_mlp_work_size": (_u64, [ctypes.POINTER(MlpShape), _u32]),
    "xc_mlp_init": (_st, [ctypes.POINTER(MlpShape), _u64, _vp]),
    "xc_mlp_forward_f32": (_st, [ctypes.POINTER(MlpShape), _vp, _vp, _u32, _vp, _vp, _vp]),
    "xc_mlp_loss_grad_f32": (_st, [ctypes.POINTER(MlpShape), _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    "xc_mlp_adam_f32": (_st, [_u64, _vp, _vp, _vp, _vp, _u32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, ctypes.c_float, _vp]),
    "xc_mlp_train_f32": (_st, [ctypes.POINTER(MlpShape), _vp, _vp, _vp, _u32, ctypes.POINTER(TrainOptions), _vp,
                               _vp]),
    "xc_validate_result_f32": (_st, [_vp, _vp, _u64, ctypes.POINTER(ctypes.c_double), _vp]),
    "xc_validate_result_f64": (_st, [_vp, _vp, _u64, ctypes.POINTER(ctypes.c_double), _vp]),
    "xc_norm_error_f32": (_st, [_vp, _vp, _vp, _u64, ctypes.POINTER(ctypes.c_double), _vp]),
    "xc_norm_error_f64": (_st, [_vp, _vp, _vp, _u64, ctypes.POINTER(ctypes.c_double), _vp]),
    "xc_kernel_timing": (_st, [_i32]),
    "xc_kernel_times": (_st, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "xc_last_error": (ctypes.c_char_p, []),
    "xc_last_launch": (ctypes.POINTER(LaunchDesc), []),
    "xc_version": (ctypes.c_char_p, []),
    "xc_device_info": (_st, [ctypes.c_char_p, ctypes.c_size_t]),
}

_LIB = None


def lib():
    """Load libxcorr.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status):
    if status != 0:
        raise XcError(status, lib().xc_last_error().decode())


def last_launch():
    d = lib().xc_last_launch().contents
    return {"kernel": d.kernel.decode(), "grid": tuple(d.grid), "block": tuple(d.block), "lds_bytes": d.lds_bytes,
            "rows": d.rows, "cols": d.cols, "pairs": d.pairs, "split": d.split}


def _stream(stream):
    if stream is not None:
        return ctypes.c_void_p(stream)
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def result_shape(family, rows, cols, n, m):
    fam = FAMILIES[family] if isinstance(family, str) else family
    k = lib().xc_num_results(fam, n, m)
    if k < 0:
        raise XcError(1, "invalid matrix counts for this family")
    return (k, 2 * rows - 1, 2 * cols - 1)


def _check_inputs(left, right):
    """(n, rows, cols) and (m, rows, cols) contiguous GPU tensors of one dtype, float32 or
    float64, on one device; a 2D tensor is one matrix.  Raises XcError otherwise."""
    import torch
    if not (isinstance(left, torch.Tensor) and isinstance(right, torch.Tensor)):
        raise XcError(1, "inputs must be torch tensors")
    if left.dim() == 2:
        left = left.unsqueeze(0)
    if right.dim() == 2:
        right = right.unsqueeze(0)
    if left.dim() != 3 or right.dim() != 3:
        raise XcError(1, "inputs must be (count, rows, cols) tensors")
    if left.dtype != right.dtype or left.dtype not in (torch.float32, torch.float64):
        raise XcError(1, "inputs must both be float32 or both float64")
    if tuple(right.shape[1:]) != tuple(left.shape[1:]):
        raise XcError(1, "left and right matrices must have the same size")
    if not (left.is_cuda and right.is_cuda):
        raise XcError(1, "inputs must be GPU tensors")
    if left.device != right.device:
        raise XcError(1, "left and right matrices must be on the same device")
    return left.contiguous(), right.contiguous()


def _check_out(out, shape, like):
    """A new result tensor, or the caller's `out` if it is a contiguous tensor of exactly
    `shape` with the inputs' dtype on the inputs' device (XcError otherwise)."""
    import torch
    if out is None:
        return torch.empty(shape, dtype=like.dtype, device=like.device)
    if not isinstance(out, torch.Tensor) or out.device != like.device or out.dtype != like.dtype:
        raise XcError(1, "out must be a tensor on the inputs' device with the inputs' dtype")
    if tuple(out.shape) != tuple(shape) or not out.is_contiguous():
        raise XcError(1, f"out must be a contiguous tensor of shape {tuple(shape)}")
    return out


TILE_KEYS = ("shifts_per_thread", "right_mats_per_thread", "left_mats_per_thread", "rows_per_step",
             "waves_per_block", "rows_per_block")


def options(precision="auto", tile=0, split=0, border=0, **tiling):
    """xc_options; `tiling` takes the direct kernel's TILE_KEYS (0 = library default)."""
    opt = Options(PRECISIONS[precision] if isinstance(precision, str) else precision, tile, split, border)
    for key, value in tiling.items():
        if key not in TILE_KEYS:
            raise XcError(1, f"unknown option {key}")
        setattr(opt, key, int(value))
    return opt


def cross_corr(family, left, right, out=None, precision="auto", split=0, border=0, tile=0, stream=None, **tiling):
    """Batched full cross-correlation of torch device tensors.

    left: (n, rows, cols), right: (m, rows, cols), float32 or float64, contiguous, on the GPU.
    Returns `out` of shape (k, 2*rows-1, 2*cols-1) in the family's result order.  `tiling`
    sets the direct kernel's register tile and schedule (TILE_KEYS, see xc_options).
    """
    import torch
    fam = FAMILIES[family] if isinstance(family, str) else family
    left, right = _check_inputs(left, right)
    n, rows, cols = left.shape
    m = right.shape[0]
    out = _check_out(out, result_shape(fam, rows, cols, n, m), left)
    opt = options(precision, tile, split, border, **tiling)
    fn = lib().xc_cross_corr_f32 if left.dtype == torch.float32 else lib().xc_cross_corr_f64
    check(fn(fam, _ptr(left), _ptr(right), _ptr(out), rows, cols, n, m, ctypes.byref(opt), _stream(stream)))
    return out


def fft_cross_corr(family, left, right, out=None, normalize=True, fp64_transform=False, stream=None):
    """FFT method; normalize=False returns the unnormalised (2*rows) x (2*cols) circular results."""
    import torch
    fam = FAMILIES[family] if isinstance(family, str) else family
    left, right = _check_inputs(left, right)
    n, rows, cols = left.shape
    m = right.shape[0]
    k = result_shape(fam, rows, cols, n, m)[0]
    shape = (k, 2 * rows - 1, 2 * cols - 1) if normalize else (k, 2 * rows, 2 * cols)
    out = _check_out(out, shape, left)
    if left.dtype == torch.float32:
        check(lib().xc_fft_cross_corr_f32(fam, _ptr(left), _ptr(right), _ptr(out), rows, cols, n, m,
                                          int(normalize), int(fp64_transform), _stream(stream)))
    else:
        check(lib().xc_fft_cross_corr_f64(fam, _ptr(left), _ptr(right), _ptr(out), rows, cols, n, m,
                                          int(normalize), _stream(stream)))
    return out


def pair_array(pairs):
    """ctypes array of xc_pair from (left_off, right_off, out_off, rows, cols) tuples; pass it
    instead of the list to calls made many times (building it costs ~1 us per pair)."""
    if isinstance(pairs, ctypes.Array):
        return pairs
    return (Pair * len(pairs))(*[Pair(*p) for p in pairs])


def _check_same(tensors, dtypes, what):
    """Contiguous GPU tensors on one device with the same element count, of the given dtypes."""
    import torch
    if not all(isinstance(t, torch.Tensor) for t in tensors):
        raise XcError(1, f"{what}: inputs must be torch tensors")
    if not all(t.is_cuda for t in tensors) or len({t.device for t in tensors}) != 1:
        raise XcError(1, f"{what}: inputs must be GPU tensors on one device")
    if len({t.numel() for t in tensors}) != 1:
        raise XcError(1, f"{what}: inputs must have the same number of elements")
    for t, d in zip(tensors, dtypes):
        if t.dtype not in d:
            raise XcError(1, f"{what}: unexpected dtype {t.dtype}")
    return [t.contiguous() for t in tensors]


def validate_result(result, valid, stream=None):
    """The reference's validate_result statistics (src/validate.hpp:134-188) reduced on the
    GPU in FP64: {max, mean, stddev, result_at_max, valid_at_max}.  result and valid: GPU
    tensors of one dtype (float32 or float64) with the same number of elements."""
    import torch
    result, valid = _check_same([result, valid], [(torch.float32, torch.float64)] * 2, "validate_result")
    if result.dtype != valid.dtype:
        raise XcError(1, "validate_result: result and valid must have one dtype")
    out = (ctypes.c_double * 5)()
    fn = lib().xc_validate_result_f32 if result.dtype == torch.float32 else lib().xc_validate_result_f64
    check(fn(_ptr(result), _ptr(valid), result.numel(), out, _stream(stream)))
    return dict(zip(("max", "mean", "stddev", "result_at_max", "valid_at_max"), list(out)))


def norm_error(x, ref, den, stream=None):
    """Normalised error |x - ref| / den reduced on the GPU in FP64: (max, mean, index of the
    max).  x: float32 or float64; ref and den (corr(|L|, |R|)): float64; same element count."""
    import torch
    x, ref, den = _check_same([x, ref, den], [(torch.float32, torch.float64), (torch.float64,), (torch.float64,)],
                              "norm_error")
    out = (ctypes.c_double * 3)()
    fn = lib().xc_norm_error_f32 if x.dtype == torch.float32 else lib().xc_norm_error_f64
    check(fn(_ptr(x), _ptr(ref), _ptr(den), x.numel(), out, _stream(stream)))
    return out[0], out[1], int(out[2])


def batch(pairs, left_flat, right_flat, out_flat, precision="auto", tile=0, split=0, stream=None):
    """Uneven batch.  pairs: list of (left_off, right_off, out_off, rows, cols) in elements,
    or a pair_array.  precision "fft64" with tile 1 is the own-FFT plan over `split` graph
    branches (0 = 8)."""
    arr = pair_array(pairs)
    opt = Options(PRECISIONS[precision] if isinstance(precision, str) else precision, tile, split, 0)
    check(lib().xc_batch_f32(arr, len(pairs), _ptr(left_flat), _ptr(right_flat), _ptr(out_flat),
                             ctypes.byref(opt), _stream(stream)))
    return out_flat


def batch_autotune(pairs, left_flat, right_flat, out_flat, stream=None):
    """Times every candidate schedule of this pair list (out_flat is overwritten) and makes
    the fastest the one AUTO batches of the same pairs use.  Returns (best, [(candidate, ms)])
    with each schedule as (precision, tile, split)."""
    arr = pair_array(pairs)
    cap = 16
    cands, ms = (Options * cap)(), (ctypes.c_float * cap)()
    count, best = ctypes.c_uint32(), Options()
    check(lib().xc_batch_autotune_f32(arr, len(pairs), _ptr(left_flat), _ptr(right_flat), _ptr(out_flat), cands, ms,
                                      cap, ctypes.byref(count), ctypes.byref(best), _stream(stream)))
    sched = lambda o: (o.precision, o.tile, o.split)
    return sched(best), [(sched(cands[i]), float(ms[i])) for i in range(min(count.value, cap))]


class BatchPlan:
    """Reusable uneven-batch plan, replayed as a HIP graph."""

    def __init__(self, pairs, precision="bf16x3", border=0, tile=0, branches=0):
        """precision bf16x3 / bf16 / fp16: one persistent MFMA kernel; fft64 with tile=1: the
        own FFT kernels per size class, on `branches` parallel graph branches (0 = 8)."""
        arr = pair_array(pairs)
        opt = Options(PRECISIONS[precision] if isinstance(precision, str) else precision, tile, branches, border)
        self._h = ctypes.c_void_p()
        check(lib().xc_batch_plan_create_f32(arr, len(pairs), ctypes.byref(opt), ctypes.byref(self._h)))

    def run(self, left_flat, right_flat, out_flat, use_graph=True, stream=None):
        check(lib().xc_batch_plan_exec_f32(self._h, _ptr(left_flat), _ptr(right_flat), _ptr(out_flat),
                                           int(use_graph), _stream(stream)))
        return out_flat

    def close(self):
        if self._h:
            check(lib().xc_batch_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- method selection (autotuner + cost model) ------------------------------------------
SELECT_SOURCES = {0: "model", 1: "tuned", 2: "none"}


def candidates():
    """[(name, precision, tile)] of every candidate XC_PREC_AUTO chooses from."""
    out = []
    for i in range(lib().xc_num_candidates()):
        o = Options()
        name = ctypes.create_string_buffer(32)
        check(lib().xc_candidate(i, ctypes.byref(o), name, 32))
        out.append((name.value.decode(), o.precision, o.tile))
    return out


def _workload(family, rows, cols, n, m):
    fam = FAMILIES[family] if isinstance(family, str) else family
    return Workload(fam, rows, cols, n, m)


def features(family, rows, cols, n, m):
    buf = (ctypes.c_float * 64)()
    cnt = ctypes.c_uint32()
    check(lib().xc_workload_features(ctypes.byref(_workload(family, rows, cols, n, m)), buf, 64, ctypes.byref(cnt)))
    return [buf[i] for i in range(cnt.value)]


def autotune(family, left, right, out, reps=3, stream=None):
    """Exhaustive autotune on torch device tensors: (ms per candidate, index of the fastest)."""
    fam = FAMILIES[family] if isinstance(family, str) else family
    n, rows, cols = left.shape
    m = right.shape[0]
    nc = lib().xc_num_candidates()
    ms = (ctypes.c_float * nc)()
    best = ctypes.c_uint32()
    fn = lib().xc_autotune_f32
    check(fn(fam, _ptr(left), _ptr(right), _ptr(out), rows, cols, n, m, reps, ms, ctypes.byref(best),
             _stream(stream)))
    return [ms[i] for i in range(nc)], best.value


def autotune_clear():
    """Forget the process's autotune table (AUTO goes back to the cost model)."""
    check(lib().xc_autotune_clear())


def time_candidate(family, left, right, out, index, reps=3, stream=None):
    """ms of one candidate (xc_time_candidate_f32) on torch device tensors."""
    fam = FAMILIES[family] if isinstance(family, str) else family
    n, rows, cols = left.shape
    m = right.shape[0]
    ms = ctypes.c_float()
    check(lib().xc_time_candidate_f32(fam, _ptr(left), _ptr(right), _ptr(out), rows, cols, n, m, index, reps,
                                      ctypes.byref(ms), _stream(stream)))
    return ms.value


def cost_model_load(path=None):
    check(lib().xc_cost_model_load(None if path is None else str(path).encode()))


def predict(family, rows, cols, n, m):
    nc = lib().xc_num_candidates()
    ms = (ctypes.c_float * nc)()
    check(lib().xc_cost_model_predict(ctypes.byref(_workload(family, rows, cols, n, m)), ms))
    return [ms[i] for i in range(nc)]


def select(family, rows, cols, n, m):
    """(options, source) that XC_PREC_AUTO would use without autotuning."""
    o = Options()
    src = ctypes.c_int()
    check(lib().xc_select(ctypes.byref(_workload(family, rows, cols, n, m)), ctypes.byref(o), ctypes.byref(src)))
    return o, SELECT_SOURCES.get(src.value, src.value)


def cost_model_fit(timing_csv, model_path, epochs=3000, lr=3e-3, weight_decay=0.0, seed=1, hidden=32,
                   use_graph=True, holdout_mod=0):
    o = TrainOptions(epochs, lr, weight_decay, seed, hidden, int(use_graph))
    check(lib().xc_cost_model_fit(str(timing_csv).encode(), str(model_path).encode(), ctypes.byref(o), holdout_mod))


def kernel_timing(enable=True):
    check(lib().xc_kernel_timing(int(enable)))


def kernel_times():
    """{kernel name: (launches, total ms)} since the last call (synchronises)."""
    need = ctypes.c_size_t()
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().xc_kernel_times(buf, len(buf), ctypes.byref(need)))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.rsplit(" ", 2)
        out[name] = (int(cnt), float(ms))
    return out

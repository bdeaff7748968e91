"""Shared helpers for the test suite: CSV fixtures, the CPU oracle, seeded inputs.

The oracle (oracle/liboracle.so) is the checker.  Tests call it through ctypes
and compare the product library (cross-corr_amd, loaded through its C ABI)
against it.
"""
import ctypes
import os
import pathlib

import numpy as np

REPO = pathlib.Path(__file__).resolve().parent.parent
GOLDEN = REPO / "tests" / "golden"
DATA = GOLDEN / "data"

FAMILIES = {"one_to_one": 0, "one_to_many": 1, "n_to_mn": 2, "n_to_m": 3}

# valid_<family>_<rows>_<cols>_<nL>_<nR>_<idL>_<idR>.csv  ->  (left file, right file).
# data/README.md:24-31 names the inputs; valid_n_to_m_10_10_4_4_1_1.csv was computed
# from ina_10_10_4_1.csv x ina_10_10_4_2.csv (its name says right id 1).
FIXTURES = {
    "valid_one_to_one_10_10_1_1_1_2.csv": ("ina_10_10_1_1.csv", "ina_10_10_1_2.csv"),
    "valid_one_to_one_64_64_1_1_1_2.csv": ("ina_64_64_1_1.csv", "ina_64_64_1_2.csv"),
    "valid_one_to_one_128_128_1_1_1_2.csv": ("ina_128_128_1_1.csv", "ina_128_128_1_2.csv"),
    "valid_one_to_many_10_10_1_2_1_1.csv": ("ina_10_10_1_1.csv", "ina_10_10_2_1.csv"),
    "valid_one_to_many_10_10_1_4_1_1.csv": ("ina_10_10_1_1.csv", "ina_10_10_4_1.csv"),
    "valid_one_to_many_64_64_1_2_1_1.csv": ("ina_64_64_1_1.csv", "ina_64_64_2_1.csv"),
    "valid_n_to_mn_10_10_2_4_1_1.csv": ("ina_10_10_2_1.csv", "ina_10_10_4_1.csv"),
    "valid_n_to_mn_128_128_2_4_1_1.csv": ("ina_128_128_2_1.csv", "ina_128_128_4_1.csv"),
    "valid_n_to_m_10_10_2_4_1_1.csv": ("ina_10_10_2_1.csv", "ina_10_10_4_1.csv"),
    "valid_n_to_m_10_10_4_4_1_1.csv": ("ina_10_10_4_1.csv", "ina_10_10_4_2.csv"),
    "valid_n_to_m_64_64_4_2_1_1.csv": ("ina_64_64_4_1.csv", "ina_64_64_2_1.csv"),
    "valid_n_to_m_64_64_4_4_1_2.csv": ("ina_64_64_4_1.csv", "ina_64_64_4_2.csv"),
}


def fixture_family(name):
    for fam in ("one_to_one", "one_to_many", "n_to_mn", "n_to_m"):
        if name.startswith("valid_" + fam + "_"):
            return fam
    raise ValueError(name)


def load_csv(path):
    """Read the reference CSV format: '# rows,cols[,n]' then rows*n lines."""
    path = pathlib.Path(path)
    with open(path) as f:
        header = f.readline()
        assert header.startswith("# "), header
        parts = [int(p) for p in header[2:].strip().split(",")]
        rows, cols = parts[0], parts[1]
        n = parts[2] if len(parts) > 2 else 1
        data = np.loadtxt(f, delimiter=",", dtype=np.float64, ndmin=2)
    assert data.shape == (rows * n, cols), (path, data.shape)
    return data.reshape(n, rows, cols)


def significant_digits(path):
    with open(path) as f:
        f.readline()
        tok = f.readline().split(",")[0]
    mant = tok.lower().split("e")[0].replace("-", "").replace(".", "").lstrip("0")
    return len(mant)


_ORACLE = None


def oracle():
    global _ORACLE
    if _ORACLE is None:
        so = REPO / "oracle" / "liboracle.so"
        if not so.exists():
            import subprocess
            subprocess.check_call(["make", "-s", "-C", str(REPO / "oracle")])
        lib = ctypes.CDLL(str(so))
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        lib.xco_pair_f64.argtypes = [dp, dp, dp, u32, u32]
        lib.xco_pair_f32.argtypes = [fp, fp, fp, u32, u32]
        lib.xco_pair_abs_f64.argtypes = [dp, dp, dp, u32, u32]
        lib.xco_num_results.argtypes = [i32, u32, u32]
        lib.xco_num_results.restype = ctypes.c_int64
        lib.xco_family_f64.argtypes = [i32, dp, dp, dp, u32, u32, u32, u32, i32]
        lib.xco_family_abs_f64.argtypes = [i32, dp, dp, dp, u32, u32, u32, u32, i32]
        lib.xco_family_f32.argtypes = [i32, fp, fp, fp, u32, u32, u32, u32, i32]
        lib.xco_family_range_f32.argtypes = [i32, fp, fp, fp, u32, u32, u32, u32, u64, u64, i32]
        lib.xco_validate_result.argtypes = [dp, dp, u64, dp]
        lib.xco_norm_error.argtypes = [dp, dp, dp, u64, dp]
        lib.xco_normalize_fft_f64.argtypes = [dp, u32, u32, dp]
        lib.xco_interleaved_index.argtypes = [u32, u32, u32]
        lib.xco_interleaved_index.restype = u32
        lib.xco_next_iteration_count.argtypes = [ctypes.c_double, ctypes.c_double, u64]
        lib.xco_next_iteration_count.restype = u64
        _ORACLE = lib
    return _ORACLE


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def n_results(family, n, m):
    return oracle().xco_num_results(FAMILIES[family] if isinstance(family, str) else family, n, m)


def oracle_family(family, left, right, dtype=np.float64, absval=False, threads=None):
    """Run the oracle for a whole family.  left: (n, rows, cols), right: (m, rows, cols)."""
    fam = FAMILIES[family] if isinstance(family, str) else family
    left = np.ascontiguousarray(left, dtype=dtype)
    right = np.ascontiguousarray(right, dtype=dtype)
    n, rows, cols = left.shape
    m = right.shape[0]
    k = oracle().xco_num_results(fam, n, m)
    if k < 0:
        raise ValueError("invalid matrix counts for family")
    out = np.zeros((k, 2 * rows - 1, 2 * cols - 1), dtype=dtype)
    threads = threads if threads is not None else min(8, os.cpu_count() or 1)
    lib = oracle()
    if dtype == np.float64:
        fn = lib.xco_family_abs_f64 if absval else lib.xco_family_f64
        rc = fn(fam, _p(left, ctypes.c_double), _p(right, ctypes.c_double), _p(out, ctypes.c_double),
                rows, cols, n, m, threads)
    else:
        assert not absval
        rc = lib.xco_family_f32(fam, _p(left, ctypes.c_float), _p(right, ctypes.c_float), _p(out, ctypes.c_float),
                                rows, cols, n, m, threads)
    assert rc == 0
    return out


def oracle_pair(left, right, dtype=np.float64):
    return oracle_family(0, left[None], right[None], dtype=dtype)[0]


def norm_error(x, ref, den):
    """max and mean of |x - ref| / corr(|L|,|R|)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    ref = np.ascontiguousarray(ref, dtype=np.float64).ravel()
    den = np.ascontiguousarray(den, dtype=np.float64).ravel()
    out = np.zeros(3)
    oracle().xco_norm_error(_p(x, ctypes.c_double), _p(ref, ctypes.c_double), _p(den, ctypes.c_double),
                            x.size, _p(out, ctypes.c_double))
    return out[0], out[1]


def validate_result(result, valid):
    result = np.ascontiguousarray(result, dtype=np.float64).ravel()
    valid = np.ascontiguousarray(valid, dtype=np.float64).ravel()
    out = np.zeros(5)
    oracle().xco_validate_result(_p(result, ctypes.c_double), _p(valid, ctypes.c_double), result.size,
                                 _p(out, ctypes.c_double))
    return out


def normalize_fft(res):
    res = np.ascontiguousarray(res, dtype=np.float64)
    rows, cols = res.shape
    norm = np.zeros((rows - 1, cols - 1))
    oracle().xco_normalize_fft_f64(_p(res, ctypes.c_double), rows, cols, _p(norm, ctypes.c_double))
    return norm


def seeded(shape, dist, seed, dtype=np.float32):
    """Seeded synthetic inputs: 'u01' = U[0,1), 'u11' = U[-1,1), 'normal' = N(0,1)."""
    rng = np.random.default_rng(seed)
    if dist == "u01":
        a = rng.random(shape)
    elif dist == "u11":
        a = rng.random(shape) * 2.0 - 1.0
    elif dist == "normal":
        a = rng.standard_normal(shape)
    else:
        raise ValueError(dist)
    return a.astype(dtype)


_PKG = None


def xcorr():
    """The product package (cross-corr_amd/__init__.py), imported by path."""
    global _PKG
    if _PKG is None:
        import importlib.util
        import sys
        spec = importlib.util.spec_from_file_location("cross_corr_amd", REPO / "cross-corr_amd" / "__init__.py")
        mod = importlib.util.module_from_spec(spec)
        sys.modules["cross_corr_amd"] = mod
        spec.loader.exec_module(mod)
        _PKG = mod
    return _PKG


def header_symbols():
    """Function names declared in include/*.h."""
    import re
    names = []
    for h in sorted((REPO / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(xc_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def uneven_batch(groups, seed, lo=16, hi=512, per_group=4, dist="u11", sizes=None):
    """Uneven n_to_mn-style batch: `groups` left matrices of square size N_g = round(lo*(hi/lo)^U),
    U ~ U[0,1) from `seed`, each paired with `per_group` right matrices of its size.
    Returns (left_flat, right_flat, pairs, out_len, sizes) with pairs as
    (left_off, right_off, out_off, rows, cols) element offsets into the flat buffers."""
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = [int(round(lo * (hi / lo) ** rng.random())) for _ in range(groups)]
    lefts, rights, pairs = [], [], []
    loff = roff = ooff = 0
    for g, n in enumerate(sizes):
        lefts.append(seeded((n, n), dist, seed * 1000 + 2 * g, np.float32).ravel())
        for j in range(per_group):
            rights.append(seeded((n, n), dist, seed * 1000 + 2 * g + 1 + 7919 * (j + 1), np.float32).ravel())
            pairs.append((loff, roff, ooff, n, n))
            roff += n * n
            ooff += (2 * n - 1) ** 2
        loff += n * n
    return np.concatenate(lefts), np.concatenate(rights), pairs, ooff, sizes


def batch_oracle_check(left_flat, right_flat, pairs, out_flat, tol, check_pairs=None):
    """Max normalised error of each pair of an uneven batch against the FP64 oracle."""
    worst = 0.0
    idx = range(len(pairs)) if check_pairs is None else check_pairs
    for i in idx:
        lo, ro, oo, r, c = pairs[i]
        L = left_flat[lo:lo + r * c].reshape(1, r, c).astype(np.float64)
        R = right_flat[ro:ro + r * c].reshape(1, r, c).astype(np.float64)
        got = out_flat[oo:oo + (2 * r - 1) * (2 * c - 1)].reshape(1, 2 * r - 1, 2 * c - 1)
        err, _ = norm_error(got, oracle_family(0, L, R), oracle_family(0, L, R, absval=True))
        assert err <= tol, (i, r, c, err)
        worst = max(worst, err)
    return worst

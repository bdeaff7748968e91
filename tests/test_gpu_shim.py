"""Every launcher of include/xcorr_kernels.hpp -- the re-creation of the reference's
src/kernels.cuh:7-283 interface -- compiled into a C++ program (cross-corr_amd/cli/shim_check.cpp)
that calls it as code written against kernels.cuh does.  The program writes its inputs and
results; here each correlation is compared with the FP64 oracle (normalised error <= 1e-6 for
float data, <= 1e-12 for double), the FFT helpers with their definitions in numpy
(src/scatter.cu:15-46, src/hadamard.cu:18-24,33-58,116-149), and the tuning values the
reference's release build refuses must have been refused.
"""
import pathlib
import subprocess

import numpy as np
import pytest

from xc_testutil import REPO, norm_error, oracle_family

pytestmark = pytest.mark.gpu

BIN = REPO / "cross-corr_amd" / "bin" / "shim_check"

LAUNCHERS = {
    "naive_original", "shuffle", "shuffle_wd_triangle", "shuffle_wd_rectangle", "shuffle_wd_none", "multimat_right",
    "warp_per_shift", "warp_per_shift_simple", "warp_per_shift_wd", "multimat_right_wd", "warp_per_shift_shared_mem",
    "n_to_m_multimat_both_wd", "local_mem_n_to_m_multimat_both_wd", "multirow_right", "multirow_both",
    "local_mem_multirow_both", "block_per_shift", "multirow_right_multimat_right",
    "one_to_many_multirow_both_multimat_right", "n_to_m_multirow_both_multimat_both", "shuffle_f64",
    "multirow_both_f64", "one_to_many_multirow_both_multimat_right_f64", "n_to_m_multirow_both_multimat_both_f64",
    "naive_original_f64", "scatter", "hadamard", "hadamard_n_to_m_over_right", "hadamard_n_to_m_over_output",
}
REJECTIONS = {
    "reject_warps": "warps per thread block", "reject_right_mats": "right matrices per thread",
    "reject_shifts": "shifts per thread", "reject_left_rows": "left rows per iteration",
    "reject_one_to_many_shifts": "shifts per thread in each right matrix",
    "reject_n_to_m_left_mats": "left matrices per thread", "reject_rows_per_thread": "rows per thread",
    "reject_search_size": "search size", "reject_block": "block too large",
}


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    if not BIN.exists():
        pytest.fail(f"{BIN} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    out = tmp_path_factory.mktemp("shim")
    print("running shim_check", flush=True)
    r = subprocess.run([str(BIN), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = (out / "manifest.txt").read_text().splitlines()
    return out, {ln.split()[0]: ln.split(maxsplit=2)[1:] for ln in lines}


def load(path, dtype):
    return np.fromfile(path, dtype=dtype)


def test_every_launcher_ran(cases):
    _, man = cases
    assert LAUNCHERS | set(REJECTIONS) == set(man), set(man) ^ (LAUNCHERS | set(REJECTIONS))


@pytest.mark.parametrize("name", sorted(n for n in LAUNCHERS if not n.startswith(("scatter", "hadamard"))))
def test_correlation_launcher_matches_oracle(cases, name):
    out, man = cases
    kind, rest = man[name]
    assert kind == "corr"
    family, rows, cols, n, m, t = rest.split()
    rows, cols, n, m = int(rows), int(cols), int(n), int(m)
    dt = np.float32 if t == "f32" else np.float64
    left = load(out / f"{name}.left", dt).reshape(n, rows, cols).astype(np.float64)
    right = load(out / f"{name}.right", dt).reshape(m, rows, cols).astype(np.float64)
    got = load(out / f"{name}.out", dt).reshape(-1, 2 * rows - 1, 2 * cols - 1)
    ref = oracle_family(family, left, right)
    den = oracle_family(family, left, right, absval=True)
    assert got.shape == ref.shape
    err, _ = norm_error(got, ref, den)
    assert err <= (1e-6 if dt == np.float32 else 1e-12), (name, err)


def test_scatter_launcher(cases):
    out, man = cases
    sr, sc, num, dr, dc, r0, c0 = (int(v) for v in man["scatter"][1].split())
    src = load(out / "scatter.src", np.float32).reshape(num, sr, sc)
    got = load(out / "scatter.out", np.float32).reshape(num, dr, dc)
    want = np.zeros((num, dr, dc), np.float32)
    want[:, r0:r0 + sr, c0:c0 + sc] = src
    assert np.array_equal(got, want)


def _complex(a):
    return a[0::2].astype(np.float64) + 1j * a[1::2].astype(np.float64)


def test_hadamard_original_launcher(cases):
    out, man = cases
    count, batch, _ = man["hadamard"][1].split()
    count, batch = int(count), int(batch)
    ref = _complex(load(out / "hadamard.ref", np.float32))
    de = _complex(load(out / "hadamard.def", np.float32)).reshape(batch, count)
    got = _complex(load(out / "hadamard.out", np.float32)).reshape(batch, count)
    want = ref[None, :] * np.conj(de)
    assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want))


@pytest.mark.parametrize("name", ["hadamard_n_to_m_over_right", "hadamard_n_to_m_over_output"])
def test_hadamard_n_to_m_launchers(cases, name):
    out, man = cases
    count, n, m, _ = man[name][1].split()
    count, n, m = int(count), int(n), int(m)
    left = _complex(load(out / f"{name}.left", np.float64)).reshape(n, count)
    right = _complex(load(out / f"{name}.right", np.float64)).reshape(m, count)
    got = _complex(load(out / f"{name}.out", np.float64)).reshape(n, m, count)
    want = left[:, None, :] * np.conj(right[None, :, :])
    assert np.max(np.abs(got - want)) <= 1e-14


@pytest.mark.parametrize("name", sorted(REJECTIONS))
def test_refused_tuning_values(cases, name):
    _, man = cases
    kind, msg = man[name]
    assert kind == "rejected"
    assert REJECTIONS[name].lower() in msg.lower(), msg

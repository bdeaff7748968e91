"""The `cross` command line (cross-corr_amd/bin/cross): command surface, formats and exit codes
of the reference's CLI (src/main.cpp:465-683, src/run_args.cpp:87-150) and the
benchmark runner's invocation (benchmarking/external/executable.py:97-165).

CPU tests use the cpu_* algorithms (no device); GPU tests run every algorithm name on the
golden fixtures and validate against the stored results."""
import json
import os
import pathlib
import subprocess

import numpy as np
import pytest

from xc_testutil import DATA, REPO, load_csv, norm_error, oracle_family

CROSS = REPO / "cross-corr_amd" / "bin" / "cross"
SAMPLE = REPO / "cross-corr_amd" / "bin" / "sample_app"

REFERENCE_NAMES = """cpu_one_to_one cpu_one_to_many cpu_n_to_mn cpu_n_to_m nai_orig_one_to_one nai_orig_one_to_many
nai_orig_n_to_mn nai_orig_n_to_m nai_shuffle_one_to_one nai_shuffle_work_distribution_one_to_one
nai_shuffle_multimat_right_one_to_one nai_shuffle_multimat_right_one_to_many
nai_shuffle_multimat_right_work_distribution_one_to_one nai_shuffle_multimat_right_work_distribution_one_to_many
nai_shuffle_multimat_right_work_distribution_n_to_mn nai_shuffle_multimat_right_work_distribution_n_to_m
nai_shuffle_multimat_both_work_distribution_local_mem_n_to_m nai_shuffle_multimat_both_work_distribution_n_to_m
nai_shuffle_multirow_both_multimat_both_n_to_m nai_shuffle_multirow_right_one_to_one nai_shuffle_multirow_both_one_to_one
nai_shuffle_multirow_both_local_mem_one_to_one nai_shuffle_multirow_right_multimat_right_one_to_one
nai_shuffle_multirow_right_multimat_right_one_to_many nai_shuffle_multirow_right_multimat_right_n_to_mn
nai_shuffle_multirow_both_multimat_right_one_to_many nai_shuffle_multirow_both_multimat_right_n_to_mn
nai_warp_per_shift_one_to_one nai_warp_per_shift_simple_indexing_one_to_one nai_warp_per_shift_work_distribution_one_to_one
nai_warp_per_shift_shared_mem_one_to_one nai_warp_per_shift_shared_mem_one_to_many nai_warp_per_shift_shared_mem_n_to_mn
nai_warp_per_shift_shared_mem_n_to_m nai_block_per_shift_one_to_one fft_orig_one_to_one fft_reduced_transfer_one_to_one
fft_orig_one_to_many fft_reduced_transfer_one_to_many fft_orig_n_to_mn fft_reduced_transfer_n_to_mn fft_better_n_to_m""".split()

# family -> (left, right, valid) fixtures of size 10x10
FAMILY_FIXTURE = {
    "one_to_one": ("ina_10_10_1_1.csv", "ina_10_10_1_2.csv", "valid_one_to_one_10_10_1_1_1_2.csv"),
    "one_to_many": ("ina_10_10_1_1.csv", "ina_10_10_4_1.csv", "valid_one_to_many_10_10_1_4_1_1.csv"),
    "n_to_mn": ("ina_10_10_2_1.csv", "ina_10_10_4_1.csv", "valid_n_to_mn_10_10_2_4_1_1.csv"),
    "n_to_m": ("ina_10_10_4_1.csv", "ina_10_10_4_2.csv", "valid_n_to_m_10_10_4_4_1_1.csv"),
}


def family_of(name):
    for fam in ("one_to_one", "one_to_many", "n_to_mn", "n_to_m"):
        if name.endswith("_" + fam):
            return fam
    raise ValueError(name)


def cross(*args, check=None):
    res = subprocess.run([str(CROSS), *map(str, args)], capture_output=True, text=True, timeout=300)
    if check is not None:
        assert res.returncode == check, (args, res.returncode, res.stdout, res.stderr)
    return res


def test_list_contains_every_reference_algorithm():
    names = cross("list", check=0).stdout.split()
    assert names == sorted(names)
    for n in REFERENCE_NAMES:
        assert n in names, n
    assert len(REFERENCE_NAMES) == 42


@pytest.mark.parametrize("args,out", [
    (("one_to_one", 10, 10, 1, 1), "Valid"), (("one_to_one", 10, 10, 2, 1), "Invalid"),
    (("one_to_many", 5, 7, 1, 3), "Valid"), (("one_to_many", 5, 7, 2, 3), "Invalid"),
    (("n_to_mn", 10, 10, 2, 4), "Valid"), (("n_to_mn", 10, 10, 3, 4), "Invalid"),
    (("n_to_m", 10, 10, 3, 4), "Valid"), (("n_to_m", 0, 10, 3, 4), "Invalid"),
])
def test_input_command(args, out):
    assert cross("input", *args, check=0).stdout.strip() == out


def test_input_unknown_type_exit_1():
    assert cross("input", "bogus", 1, 1, 1, 1).returncode == 1


def test_unknown_command_and_option_exit_1():
    assert cross("frobnicate").returncode == 1
    assert cross("run", "--bogus", "x").returncode == 1


def test_unknown_algorithm_exit_1():
    r = cross("run", "no_such_alg", DATA / "ina_10_10_1_1.csv", DATA / "ina_10_10_1_2.csv")
    assert r.returncode == 1 and "Unknown algorithm" in r.stderr


def test_unknown_benchmark_type_exit_1():
    r = cross("run", "-b", "Everything", "cpu_one_to_one", DATA / "ina_10_10_1_1.csv", DATA / "ina_10_10_1_2.csv")
    assert r.returncode == 1


@pytest.mark.parametrize("name", ["cpu_one_to_one", "cpu_one_to_many", "cpu_n_to_mn", "cpu_n_to_m"])
def test_cpu_algorithms_match_fixtures(name, tmp_path):
    left, right, valid = FAMILY_FIXTURE[family_of(name)]
    out = tmp_path / "out.csv"
    r = cross("run", "--no_progress", "-v", DATA / valid, "-o", out, name, DATA / left, DATA / right, check=0)
    lines = r.stdout.strip().splitlines()
    assert lines[0] == "Max,Mean,Stddev"
    assert float(lines[1].split(",")[0]) < 1e-5
    got = load_csv(out)
    L, R = load_csv(DATA / left).astype(np.float32), load_csv(DATA / right).astype(np.float32)
    fam = family_of(name)
    den = oracle_family(fam, L.astype(np.float64), R.astype(np.float64), absval=True)
    assert norm_error(got, oracle_family(fam, L.astype(np.float64), R.astype(np.float64)), den)[0] < 1e-6


def test_validate_command_csv_and_header():
    v = DATA / "valid_one_to_one_10_10_1_1_1_2.csv"
    r = cross("validate", "--csv", "--print_header", v, v, v, check=0)
    assert r.stdout.splitlines() == ["Max,Mean,Stddev", "0.000000e+00,0.000000e+00,0.000000e+00",
                                     "0.000000e+00,0.000000e+00,0.000000e+00"]


def _write_csv(path, a):
    a = np.asarray(a, dtype=np.float64)
    with open(path, "w") as f:
        f.write(f"# {a.shape[0]},{a.shape[1]}\n")
        for row in a:
            f.write(",".join(repr(float(v)) for v in row) + "\n")


@pytest.mark.parametrize("result,valid", [
    ([[1.0, 2.0], [0.0, 1e-6]], [[1.0, 1.0], [0.0, 0.0]]),          # max 0.5, mean 0.125
    ([[2e-5, 3.0], [4.0, -1.0]], [[1e-5, 3.0], [5.0, -1.5]]),
    ([[1.5, -2.0, 7.0]], [[1.0, 2.0, 7.5]]),
])
def test_validate_command_matches_the_oracle_metric(tmp_path, result, valid):
    # The CLI's validate_result (src/validate.hpp:134-188) against the oracle's restatement,
    # which the hand-computed cases of tests/test_oracle.py pin.
    from xc_testutil import validate_result
    rp, vp = tmp_path / "result.csv", tmp_path / "valid.csv"
    _write_csv(rp, result)
    _write_csv(vp, valid)
    r = cross("validate", "--csv", vp, rp, check=0)
    got = [float(v) for v in r.stdout.strip().split(",")]
    want = validate_result(np.ravel(result), np.ravel(valid))[:3]
    assert np.allclose(got, want, rtol=1e-6, atol=0), (got, want)
    if result[0][:2] == [1.0, 2.0]:
        assert got[0] == 0.5 and got[1] == 0.125


def test_timing_csv_compute_mode_and_append(tmp_path):
    t = tmp_path / "times.csv"
    args = ["-b", "Compute", "-m", "0.001", "-t", t, "--no_progress", "cpu_one_to_one",
            DATA / "ina_10_10_1_1.csv", DATA / "ina_10_10_1_2.csv"]
    cross("run", *args, check=0)
    cross("run", "--append", *args, check=0)
    lines = t.read_text().splitlines()
    assert lines[0] == "Computation,Computation_iterations"
    assert len(lines) == 3 and all(len(l.split(",")) == 2 for l in lines[1:])


def test_outer_loops_store_per_loop_outputs(tmp_path):
    out = tmp_path / "res.csv"
    cross("run", "-l", 2, "-o", out, "--no_progress", "cpu_one_to_one", DATA / "ina_10_10_1_1.csv",
          DATA / "ina_10_10_1_2.csv", check=0)
    assert (tmp_path / "res_0.csv").exists() and (tmp_path / "res_1.csv").exists()


@pytest.mark.parametrize("body,needle", [
    ("# 2\n1,2\n", "Missing number of columns"),
    ("1,2\n", "Invalid file header format"),
    ("# 2,2,1\n1,2\n3\n", "too short"),
    ("# 2,2,1\n1,2,3\n3,4\n", "too long"),
    ("# 2,2,1\n1,2\n", "Failed to read line"),
])
def test_malformed_csv_exit_3(tmp_path, body, needle):
    bad = tmp_path / "bad.csv"
    bad.write_text(body)
    r = cross("run", "cpu_one_to_one", bad, DATA / "ina_10_10_1_2.csv")
    assert r.returncode == 3 and needle in r.stderr


def test_size_mismatch_and_divisibility_exit_3():
    r = cross("run", "cpu_one_to_one", DATA / "ina_10_10_1_1.csv", DATA / "ina_64_64_1_1.csv")
    assert r.returncode == 3 and "same size" in r.stderr
    r = cross("run", "cpu_n_to_mn", DATA / "ina_10_10_4_1.csv", DATA / "ina_10_10_2_1.csv")
    assert r.returncode == 3 and "divisible" in r.stderr


def test_tuning_limits(tmp_path):
    args = tmp_path / "args.json"
    args.write_text(json.dumps({"warps_per_thread_block": 33}))
    r = cross("run", "--args_path", args, "nai_shuffle_one_to_one", DATA / "ina_10_10_1_1.csv",
              DATA / "ina_10_10_1_2.csv")
    assert r.returncode == 3 and "warps" in r.stderr
    args.write_text(json.dumps({"shifts_per_thread_block": 16, "shared_mem_rows": 8}))
    r = cross("run", "--args_path", args, "nai_warp_per_shift_shared_mem_one_to_one", DATA / "ina_10_10_1_1.csv",
              DATA / "ina_10_10_1_2.csv")
    assert r.returncode == 1 and "shared memory rows" in r.stderr
    args.write_text(json.dumps({"distribution_type": "hexagon"}))
    r = cross("run", "--args_path", args, "nai_shuffle_work_distribution_one_to_one", DATA / "ina_10_10_1_1.csv",
              DATA / "ina_10_10_1_2.csv")
    assert r.returncode == 2 and "distribution" in r.stderr


@pytest.mark.parametrize("alg,key,value,needle", [
    ("nai_shuffle_multirow_both_one_to_one", "shifts_per_thread", 9, "Too many shifts per thread: 9 (max 8)"),
    ("nai_shuffle_multirow_both_one_to_one", "left_rows_per_iteration", 5, "left rows per iteration: 5 (max 4)"),
    ("nai_shuffle_multimat_right_one_to_many", "right_matrices_per_thread", 9, "right matrices per thread: 9"),
    ("nai_shuffle_multirow_both_multimat_right_one_to_many", "shifts_per_thread_right_matrix", 5,
     "shifts per thread in each right matrix: 5 (max 4)"),
    ("nai_shuffle_multirow_both_multimat_right_one_to_many", "warps_per_thread_block", 33, "Cuda block too large"),
    ("nai_shuffle_multirow_right_multimat_right_one_to_many", "right_rows_per_thread", 5,
     "right rows per thread: 5 (max 4)"),
    ("nai_shuffle_multirow_right_one_to_one", "right_rows_per_thread", 9, "right rows per thread: 9 (max 8)"),
    ("nai_shuffle_multimat_both_work_distribution_n_to_m", "left_matrices_per_thread", 5,
     "left matrices per thread: 5 (max 4)"),
    ("nai_block_per_shift_one_to_one", "block_size", 2048, "block too large"),
    ("nai_shuffle_multirow_both_one_to_one", "shifts_per_thread", 0, "Invalid value 0"),
])
def test_reference_release_limits(tmp_path, alg, key, value, needle):
    # The limits of the reference's release build (CMakeLists.txt:16-45) refuse the value at
    # the allocation step, before any device work: exit code 3.
    args = tmp_path / "args.json"
    args.write_text(json.dumps({key: value}))
    fam = alg.rsplit("_", 2)
    right = "ina_10_10_2_1.csv" if "one_to_many" in alg or "n_to_m" in alg else "ina_10_10_1_2.csv"
    left = "ina_10_10_2_1.csv" if "n_to_m" in alg and "one_to_many" not in alg else "ina_10_10_1_1.csv"
    r = cross("run", "--args_path", args, alg, DATA / left, DATA / right)
    assert r.returncode == 3 and needle in r.stderr, (r.returncode, r.stderr, fam)


def test_runner_invocation_shape_cpu(tmp_path):
    """benchmarking/external/executable.py:97-165 builds exactly this command line."""
    times, out = tmp_path / "t.csv", tmp_path / "o.csv"
    args = tmp_path / "a.json"
    args.write_text("{}")
    valid = DATA / "valid_one_to_many_10_10_1_4_1_1.csv"
    r = cross("run", "--out", out, "--validate", valid, "--times", times, "--no_progress", "--args_path", args,
              "--data_type", "single", "--benchmark_type", "Compute", "--outer_loops", 1, "--min_time", 0.001,
              "cpu_one_to_many", DATA / "ina_10_10_1_1.csv", DATA / "ina_10_10_4_1.csv", check=0)
    assert r.stdout.splitlines()[0] == "Max,Mean,Stddev"
    assert times.read_text().splitlines()[0] == "Computation,Computation_iterations"


# ------------------------------------------------------------------ GPU
GPU_NAMES = [n for n in REFERENCE_NAMES if not n.startswith("cpu_")] + [
    f"{m}_{f}" for m in ("mfma_bf16x3", "mfma_fp16x3", "auto") for f in ("one_to_one", "one_to_many", "n_to_mn",
                                                                            "n_to_m")]


@pytest.mark.gpu
@pytest.mark.parametrize("name", GPU_NAMES)
@pytest.mark.parametrize("dtype", ["single", "double"])
def test_every_algorithm_validates_against_fixture(name, dtype, tmp_path):
    if dtype == "double" and name.startswith("mfma"):
        pytest.skip("MFMA methods take single-precision data")
    left, right, valid = FAMILY_FIXTURE[family_of(name)]
    t = tmp_path / "t.csv"
    out = tmp_path / "o.csv"
    r = cross("run", "-d", dtype, "-b", "Algorithm", "-m", "0.001", "-t", t, "--no_progress", "-n", "-o", out,
              "-v", DATA / valid, name, DATA / left, DATA / right, check=0)
    lines = r.stdout.strip().splitlines()
    assert lines[0] == "Max,Mean,Stddev"
    # reference metric (relative difference, 1e-5 floor); FP32 FFT spreads an absolute error
    # over all outputs, which is large relative to the smallest ones (the reference's own
    # validation sweep leaves the FFT algorithms out for this reason, validation.yml:102-103)
    bar = 1e-3 if (name.startswith("fft_") and dtype == "single") else 1e-4
    assert float(lines[1].split(",")[0]) < bar, (name, lines)
    header = t.read_text().splitlines()[0].split(",")
    if name.startswith("fft_"):
        assert header[:6] == ["Forward FFT", "Forward FFT_iterations", "Hadamard", "Hadamard_iterations",
                              "Inverse FFT", "Inverse FFT_iterations"]
    else:
        assert header[:2] == ["Kernel", "Kernel_iterations"]
    # stored results (normalised with -n for FFT algorithms) match the oracle
    got = load_csv(out)
    L, R = load_csv(DATA / left), load_csv(DATA / right)
    fam = family_of(name)
    if dtype == "single":
        L, R = L.astype(np.float32).astype(np.float64), R.astype(np.float32).astype(np.float64)
    ref = oracle_family(fam, L, R)
    den = oracle_family(fam, L, R, absval=True)
    tol = 2e-6 if dtype == "single" else 1e-12
    if name.startswith("fft_") and dtype == "single":
        tol = 1e-3  # FP32 FFT: absolute error spread over all outputs
    if name.startswith("auto") and dtype == "double":
        tol = 1e-9  # AUTO may run the own FFT in FP64 (rounding global, ~1e-16 of the largest output)
    assert norm_error(got, ref, den)[0] < tol, name


@pytest.mark.gpu
def test_fft_unnormalised_store(tmp_path):
    out = tmp_path / "o.csv"
    cross("run", "-d", "double", "-o", out, "--no_progress", "fft_reduced_transfer_one_to_one",
          DATA / "ina_10_10_1_1.csv", DATA / "ina_10_10_1_2.csv", check=0)
    raw = load_csv(out)
    assert raw.shape == (1, 20, 20)
    r = cross("validate", "--normalize", "--csv", DATA / "valid_one_to_one_10_10_1_1_1_2.csv", out, check=0)
    assert float(r.stdout.splitlines()[-1].split(",")[0]) < 1e-9


@pytest.mark.gpu
def test_sample_app_matches_oracle():
    r = subprocess.run([str(SAMPLE), str(DATA / "ina_64_64_1_1.csv"), str(DATA / "ina_64_64_2_1.csv")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "# 127,127,2"
    got = np.array([[float(v) for v in l.split(",")] for l in lines[1:]]).reshape(2, 127, 127)
    L = load_csv(DATA / "ina_64_64_1_1.csv").astype(np.float32).astype(np.float64)
    R = load_csv(DATA / "ina_64_64_2_1.csv").astype(np.float32).astype(np.float64)
    den = oracle_family("one_to_many", L, R, absval=True)
    assert norm_error(got, oracle_family("one_to_many", L, R), den)[0] < 1e-6

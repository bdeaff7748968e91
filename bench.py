"""Benchmark of the hot path on one MI355X.

Workload (a "step"): one_to_many full 2D cross-correlation of 1 left 256x256 float32 matrix
against 64 right 256x256 matrices, values i.i.d. U[0,1) from a fixed seed, output
64 x 511 x 511 -- one call of xc_cross_corr_f32 with inputs already resident in HBM.

Methods: the plain FP32 wave64 direct kernel, the bf16x3 MFMA implicit GEMM, the FFT method
in FP64 on rocFFT (fft64) and on this library's fused FFT kernels (fft64k), and AUTO (the
method the trained cost model selects).  Each is timed over K steps after W warm-up steps
with HIP events on the stream the kernels run on; its max normalised error
(|x - x64| / corr(|L|,|R|)) is measured on results 0 and 63 against the FP64 oracle.
`value` is the pairs/s of AUTO (it must meet the 2e-6 bar).

Roofline: one more timed pass of AUTO's method with the library's kernel timing on (HIP
events around each launch on its stream) gives the average duration of its dominant
kernel; `achieved` is that kernel's algorithmic bytes (FFT kernels, HBM-bound) or FLOPs
(definition-based kernels) per launch over that duration.  `traffic` is the HBM bytes per
launch measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy of known
size) from profiles/pmc_traffic_c2.txt when that summary lists the kernel.

Prints ONE JSON line.  --gpus N with N > 1 exits: this path runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import pathlib
import platform
import sys
import time

REPO = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "tests"))

ROWS = COLS = 256
NUM_RIGHT = 64
MACS_PER_PAIR = ROWS * COLS * ROWS * COLS
ERR_BAR = 2e-6
HBM_GBPS = 8000.0
TFLOPS = {"direct_kernel": 157.3, "mfma_kernel": 2500.0}  # FP32 VALU; dense BF16 MFMA
METHODS = {"fp32": dict(precision="fp32"), "bf16x3": dict(precision="bf16x3"), "fp16x3": dict(precision="fp16x3"),
           "fft64": dict(precision="fft64", tile=0), "fft64k": dict(precision="fft64", tile=1),
           "auto": dict(precision="auto")}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def pow2ceil(v):
    p = 1
    while p < v:
        p *= 2
    return p


def kernel_bytes(name, H, W, n, m, k):
    """Algorithmic HBM bytes of one launch of a kernel of the FFT methods: every array the
    kernel must read is read once and every array it must write is written once."""
    Pr, Pc = max(16, pow2ceil(2 * H - 1)), max(16, pow2ceil(2 * W - 1))
    # rows: 3 * 2^n (n = 4..9) when shorter than the power of two (fftk_row_len)
    Pc = min([Pc] + [3 * 2 ** e for e in range(4, 10) if 3 * 2 ** e >= 2 * W - 1])
    Kc = Pc // 2 + 1
    SY, SX = 2 * H - 1, 2 * W - 1
    out = k * SY * SX * 4
    if name == "fftk_rows_fwd":
        return (n + m) * (H * W * 4 + Kc * H * 16)
    if name == "fftk_cols":
        return m * Kc * H * 16 + n * Kc * Pr * 16 + k * Kc * SY * 16
    if name == "fftk_rows_inv":
        return k * Kc * SY * 16 + out
    pr, pc = 2 * H, 2 * W
    spec = pr * (pc // 2 + 1) * 16
    if name == "fft_scatter":
        return (n + m) * (H * W * 4 + pr * pc * 8)
    if name == "rocfft_forward":
        return (n + m) * (pr * pc * 8 + spec)
    if name == "fft_hadamard":
        return n * spec + 2 * k * spec
    if name == "rocfft_inverse":
        return k * (spec + pr * pc * 8)
    if name == "fft_normalize":
        return k * SY * SX * 8 + out
    return None


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary, or None."""
    path = REPO / "profiles" / "pmc_traffic_c2.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) == 4 and f[0] == kernel and not line.startswith("#"):
            try:
                return float(f[3])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--methods", default="fp32,bf16x3,fp16x3,fft64,fft64k,auto")
    ap.add_argument("--cpu-pairs", type=int, default=3, help="pairs timed for the CPU baseline sample")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy 3 times (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings into bytes), then exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.rand(64 << 20, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this workload runs on one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent: a multi-GPU run would deal pairs to GPUs, one process each)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from xc_testutil import norm_error, oracle, oracle_family, seeded, xcorr

    x = xcorr()
    x.lib()
    dev = torch.device("cuda:0")
    left = seeded((1, ROWS, COLS), "u01", 2024, np.float32)
    right = seeded((NUM_RIGHT, ROWS, COLS), "u01", 2025, np.float32)
    dl = torch.from_numpy(left).to(dev)
    dr = torch.from_numpy(right).to(dev)
    out = torch.empty((NUM_RIGHT, 2 * ROWS - 1, 2 * COLS - 1), device=dev)
    stream = torch.cuda.current_stream()

    # FP64 oracle of results 0 and 63 for the error column.
    picks = [0, NUM_RIGHT - 1]
    l64, r64 = left.astype(np.float64), right[picks].astype(np.float64)
    ref = oracle_family("one_to_many", l64, r64)
    den = oracle_family("one_to_many", l64, r64, absval=True)

    def step(method):
        x.cross_corr("one_to_many", dl, dr, out=out, **METHODS[method])

    results = {}
    for method in args.methods.split(","):
        try:
            step(method)
            torch.cuda.synchronize()
        except x.XcError as e:
            results[method] = {"error": str(e)}
            continue
        err, _ = norm_error(out[picks].cpu().numpy(), ref, den)
        for _ in range(args.warmup):
            step(method)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.steps):
            step(method)
        stop.record(stream)
        stop.synchronize()
        ms = start.elapsed_time(stop) / args.steps
        results[method] = {"ms_per_step": ms, "pairs_per_s": NUM_RIGHT / (ms * 1e-3),
                           "gmac_per_s": NUM_RIGHT * MACS_PER_PAIR / (ms * 1e-3) / 1e9,
                           "max_norm_err": err, "last_kernel": x.last_launch()["kernel"]}
    chosen, source = x.select("one_to_many", ROWS, COLS, 1, NUM_RIGHT)
    chosen_name = [c[0] for c in x.candidates() if c[1:] == (chosen.precision, chosen.tile)]
    if "auto" not in results or "ms_per_step" not in results["auto"]:
        print(json.dumps({"metric": "pairs/s", "error": "AUTO did not run", "methods": results}))
        sys.exit(1)
    b = results["auto"]
    if b["max_norm_err"] > ERR_BAR:
        print(json.dumps({"metric": "pairs/s", "error": "AUTO missed the error bar", "methods": results}))
        sys.exit(1)

    # Per-kernel durations of AUTO's method (HIP events around each launch on its stream).
    x.kernel_times()
    x.kernel_timing(True)
    for _ in range(args.steps):
        step("auto")
    torch.cuda.synchronize()
    times = x.kernel_times()
    x.kernel_timing(False)
    dom = max(times, key=lambda kname: times[kname][1])
    launches, total_ms = times[dom]
    avg_ms = total_ms / launches
    per_launch = {kname: {"launches": c, "avg_ms": t / c} for kname, (c, t) in times.items()}
    nbytes = kernel_bytes(dom, ROWS, COLS, 1, NUM_RIGHT, NUM_RIGHT)
    if nbytes is not None:
        achieved = nbytes / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_GBPS, "traffic": pmc_traffic(dom), "kernel": dom,
                "algorithmic_bytes": nbytes, "avg_ms": avg_ms}
    else:
        flops = 2.0 * NUM_RIGHT * MACS_PER_PAIR
        peak = TFLOPS.get(dom, 2500.0)
        achieved = flops / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": pmc_traffic(dom), "kernel": dom,
                "algorithmic_flops": flops, "avg_ms": avg_ms}

    # CPU baseline: the oracle's FP32 restatement of cpu_cross_corr_one_to_many
    # (src/validate.hpp:231-247), single thread, on a bounded sample of the same workload.
    lib = oracle()
    kcpu = max(1, args.cpu_pairs)
    cbuf = np.zeros((kcpu, 2 * ROWS - 1, 2 * COLS - 1), np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    t0 = time.perf_counter()
    lib.xco_family_range_f32(1, left.ctypes.data_as(fp), right.ctypes.data_as(fp), cbuf.ctypes.data_as(fp),
                             ROWS, COLS, 1, NUM_RIGHT, 0, kcpu, 1)
    cpu_s = time.perf_counter() - t0
    cpu_rate = kcpu / cpu_s
    # The same restatement across the host cores (POSIX threads over results, one pair per
    # thread), at most 16: the host share of one GPU.
    ncores = max(1, min(os.cpu_count() or 1, 16))
    abuf = np.zeros((ncores, 2 * ROWS - 1, 2 * COLS - 1), np.float32)
    t0 = time.perf_counter()
    lib.xco_family_range_f32(1, left.ctypes.data_as(fp), right.ctypes.data_as(fp), abuf.ctypes.data_as(fp),
                             ROWS, COLS, 1, NUM_RIGHT, 0, ncores, ncores)
    all_s = time.perf_counter() - t0

    fp32 = results.get("fp32", {})
    bf = results.get("bf16x3", {})
    line = {
        "metric": "pairs/s",
        "value": b["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 in/out, f64 FFT" if chosen.precision == 5 else ("bf16x3->f32" if chosen.precision == 2 else "f32"),
        "data": "synthetic U[0,1) seeded",
        "config": {"workload": "one_to_many 1x256x256 vs 64x256x256 float32, full 511x511 outputs",
                   "pairs": NUM_RIGHT, "rows": ROWS, "cols": COLS, "method": "auto",
                   "auto_selected": chosen_name[0] if chosen_name else None, "selected_by": source},
        "gmac_per_s": b["gmac_per_s"],
        "max_norm_err": b["max_norm_err"],
        "methods": results,
        "speedup_vs_fp32_direct": fp32["ms_per_step"] / b["ms_per_step"] if "ms_per_step" in fp32 else None,
        "bf16x3_vs_fp32_direct": (fp32["ms_per_step"] / bf["ms_per_step"]
                                  if "ms_per_step" in fp32 and "ms_per_step" in bf else None),
        "kernels": per_launch,
        "roofline": roof,
        "cpu_baseline": {"value": cpu_rate, "unit": "pairs/s", "cores": 1, "kind": "port",
                         "sample": f"{kcpu} of the 64 pairs, oracle FP32 restatement of cpu_cross_corr_one_to_many, "
                                   f"{cpu_s:.1f} s on {cpu_model()} ({os.cpu_count()} host cpus)"},
        "cpu_baseline_all_cores": {"value": ncores / all_s, "unit": "pairs/s", "cores": ncores, "kind": "port",
                                   "sample": f"{ncores} of the 64 pairs, one per thread, {all_s:.1f} s"},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

"""Benchmark of the hot path on one MI355X.

Headline workload (a "step", the `value`): one_to_many full 2D cross-correlation of 1 left
256x256 float32 matrix against 64 right 256x256 matrices, values i.i.d. U[0,1) from a fixed
seed, output 64 x 511 x 511 -- one call of xc_cross_corr_f32 with inputs already resident in
HBM.  Methods: the FP32 wave64 direct kernel (default tile), the bf16x3 / fp16x3 MFMA implicit
GEMM, the FFT method in FP64 on rocFFT (fft64) and on this library's fused FFT kernels
(fft64k), and AUTO (the method the trained cost model selects).  Each is timed over K steps
after W warm-up steps with HIP events on the stream the kernels run on; its max normalised
error (|x - x64| / corr(|L|,|R|)) is measured on results 0 and 63 against the FP64 oracle.
`value` is the pairs/s of AUTO (it must meet the 2e-6 bar).

`configs` holds the other workloads, each with its own timing, error and roofline:
  c1  one_to_one 128x128 float32 U[0,1)
  c3  n_to_m 16 left x 64 right 64x64 float32 N(0,1) (FP32 direct, bf16x3, fp16x3, single
      bf16 and fp16 MFMA, the own FFT, AUTO), and the same data in float64
  c4  uneven n_to_mn batch: 128 groups of 1 left + 4 right square matrices of size
      round(16 * 32^U) (seed 2024), float32 U[-1,1), 512 pairs: one AUTO batch call (the
      schedule the batch autotuner measured fastest, replayed from a HIP graph) against one
      call per size class
  c5  one_to_many 1 x 128 float32 U[0,1) at 1024x1024 (the largest point of the crossover
      sweep; definition-based methods are left out there: minutes per step)

Roofline: one more pass of each configuration's AUTO method with the library's kernel
timing on (HIP events around each launch on its stream) gives the durations of its dominant
kernel; `achieved` is that kernel's algorithmic bytes (FFT kernels, HBM-bound) or FLOPs
(definition-based kernels) per step over its total duration per step.  `traffic` is the
HBM bytes per launch measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE, scaled by a plain
copy of known size) from profiles/pmc_traffic_c2.txt when that summary lists the kernel.

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
ERR_BAR = 2e-6
HBM_GBPS = 8000.0
# Peaks of the arithmetic each definition-based kernel issues: non-packed FP32 FMA (the
# direct kernel does not use packed FP32 instructions: half the published 157.3 TFLOP/s),
# FP64 VALU, dense BF16/FP16 MFMA.
TFLOPS = {"direct_kernel": 78.6, "direct_kernel_f64": 78.6, "mfma_kernel": 2500.0, "batch_kernel": 2500.0}
# fp32: the library's FP32 default (the lanes-over-pairs kernel where it applies); fp32tiles: the
# register-tile kernel (FP32 tile 1), for comparison runs only.
METHODS = {"fp32": dict(precision="fp32"), "fp32tiles": dict(precision="fp32", tile=1),
           "bf16x3": dict(precision="bf16x3"), "fp16x3": dict(precision="fp16x3"),
           "bf16": dict(precision="bf16"), "fp16": dict(precision="fp16"),
           "fft64": dict(precision="fft64", tile=0), "fft64k": dict(precision="fft64", tile=1),
           "auto": dict(precision="auto")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def fft_lengths(H, W, pow2_rows=False):
    """Transform lengths of the own FFT method: columns the power of two >= 2H-1 (at least
    16), rows that or 3 * 2^e (e = 4..9) when shorter (fftk_col_len / fftk_row_len); rows of
    the power of two only for its tile 2."""
    Pr, Pc = max(16, pow2ceil(2 * H - 1)), max(16, pow2ceil(2 * W - 1))
    if not pow2_rows:
        Pc = min([Pc] + [3 * 2 ** e for e in range(4, 10) if 3 * 2 ** e >= 2 * W - 1])
    return Pr, Pc


def kernel_bytes(name, H, W, n, m, k, pow2_rows=False):
    """Algorithmic HBM bytes of one launch of a kernel of the FFT methods (n left, m right
    matrices of H x W, k results): every array the kernel must read is read once and every
    array it must write is written once."""
    Pr, Pc = fft_lengths(H, W, pow2_rows)
    Kc = Pc // 2 + 1
    SY, SX = 2 * H - 1, 2 * W - 1
    out = k * SY * SX * 4
    if name == "fftk_rows_fwd":
        return (n + m) * (H * W * 4 + Kc * H * 16)
    if name == "fftk_cols":
        return m * Kc * H * 16 + n * Kc * Pr * 16 + k * Kc * SY * 16
    if name == "fftk_cols_left":
        return n * (Kc * H * 16 + Kc * Pr * 16)
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
    """HBM bytes per launch of `kernel` at the headline workload from the committed rocprofv3
    summary, or None."""
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
    ap.add_argument("--configs", default="c1,c3,c4,c5", help="workloads beside the headline ('' for none)")
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

    from xc_testutil import norm_error, oracle, oracle_family, seeded, uneven_batch, xcorr

    x = xcorr()
    x.lib()
    dev = torch.device("cuda:0")
    cands = {c[0]: c for c in x.candidates()}

    def auto_name(family, rows, cols, n, m):
        o, src = x.select(family, rows, cols, n, m)
        names = [nm for nm, c in cands.items() if c[1:] == (o.precision, o.tile)]
        return (names[0] if names else None), src

    def timed(fn, steps, warmup):
        """ms per call: `warmup` untimed calls, then `steps` calls between two HIP events on the
        current stream (the stream the library enqueues on)."""
        fn()
        torch.cuda.synchronize()
        for _ in range(warmup):
            fn()
        stream = torch.cuda.current_stream()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(steps):
            fn()
        b.record(stream)
        b.synchronize()
        return a.elapsed_time(b) / steps

    def kernels_of(fn, steps):
        """{kernel: (launches per step, ms per step)} from the library's kernel timing."""
        x.kernel_times()
        x.kernel_timing(True)
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        times = x.kernel_times()
        x.kernel_timing(False)
        return {k: (c / steps, t / steps) for k, (c, t) in times.items()}

    def roofline(times, bytes_of, flops_of):
        """Roofline of the dominant kernel: its algorithmic bytes (or FLOPs) per step over
        its time per step."""
        dom = max(times, key=lambda kname: times[kname][1])
        launches, ms = times[dom]
        nbytes = bytes_of(dom)
        if nbytes is not None:
            achieved = nbytes / (ms * 1e-3) / 1e9
            return {"bound": "hbm", "achieved": achieved, "peak": HBM_GBPS, "unit": "GB/s",
                    "frac": achieved / HBM_GBPS, "traffic": None, "kernel": dom,
                    "algorithmic_bytes": nbytes, "avg_ms": ms / launches, "launches_per_step": launches}
        flops = flops_of(dom)
        peak = TFLOPS.get(dom, 2500.0)
        achieved = flops / (ms * 1e-3) / 1e12
        return {"bound": "mfma" if peak > 1000 else "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None, "kernel": dom, "algorithmic_flops": flops,
                "avg_ms": ms / launches, "launches_per_step": launches}

    # ---- headline: one_to_many 1 x 64 of 256 x 256
    left = seeded((1, ROWS, COLS), "u01", 2024, np.float32)
    right = seeded((NUM_RIGHT, ROWS, COLS), "u01", 2025, np.float32)
    dl = torch.from_numpy(left).to(dev)
    dr = torch.from_numpy(right).to(dev)
    out = torch.empty((NUM_RIGHT, 2 * ROWS - 1, 2 * COLS - 1), device=dev)
    picks = [0, NUM_RIGHT - 1]
    l64, r64 = left.astype(np.float64), right[picks].astype(np.float64)
    ref = oracle_family("one_to_many", l64, r64)
    den = oracle_family("one_to_many", l64, r64, absval=True)
    macs = float(ROWS * COLS) ** 2

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
        steps = args.steps if method not in ("fp32", "fp32tiles") else max(3, args.steps // 4)
        ms = timed(lambda: step(method), steps, args.warmup)
        results[method] = {"ms_per_step": ms, "pairs_per_s": NUM_RIGHT / (ms * 1e-3),
                           "gmac_per_s": NUM_RIGHT * macs / (ms * 1e-3) / 1e9,
                           "max_norm_err": err, "last_kernel": x.last_launch()["kernel"]}
        log(f"c2 {method}: {ms:.4f} ms, err {err:.2e}")
    chosen, source = x.select("one_to_many", ROWS, COLS, 1, NUM_RIGHT)
    chosen_name = [c[0] for c in x.candidates() if c[1:] == (chosen.precision, chosen.tile)]
    if "auto" not in results or "ms_per_step" not in results["auto"]:
        print(json.dumps({"metric": "pairs/s", "error": "AUTO did not run", "methods": results}))
        sys.exit(1)
    b = results["auto"]
    if b["max_norm_err"] > ERR_BAR:
        print(json.dumps({"metric": "pairs/s", "error": "AUTO missed the error bar", "methods": results}))
        sys.exit(1)
    times = kernels_of(lambda: step("auto"), args.steps)
    per_launch = {kname: {"launches": c, "avg_ms": t / c} for kname, (c, t) in times.items()}
    pow2 = bool(chosen_name) and chosen_name[0] == "fft64/2"
    roof = roofline(times, lambda k: kernel_bytes(k, ROWS, COLS, 1, NUM_RIGHT, NUM_RIGHT, pow2),
                    lambda k: 2.0 * NUM_RIGHT * macs)
    roof["traffic"] = pmc_traffic(roof["kernel"])
    # The FP32 direct kernel's own roofline (the baseline the other methods are compared with).
    fp32_roof = None
    if "ms_per_step" in results.get("fp32", {}):
        t32 = kernels_of(lambda: step("fp32"), 2)
        fp32_roof = roofline(t32, lambda k: None, lambda k: 2.0 * NUM_RIGHT * macs)
        # "direct_kernel" is the direct method's timer, shared by its two kernels: name the
        # kernel that ran.
        fp32_roof["timer"] = fp32_roof["kernel"]
        fp32_roof["kernel"] = results["fp32"].get("last_kernel", fp32_roof["kernel"])

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
    del cbuf, abuf

    # ---- the other workloads
    configs = {}
    cfgs = [c for c in args.configs.split(",") if c]

    def family_config(name, workload, family, lnp, rnp, methods, check, steps):
        """Times `methods` on a regular family; error against the oracle on results `check`."""
        L, R = torch.from_numpy(lnp).to(dev), torch.from_numpy(rnp).to(dev)
        n, H, W = lnp.shape
        m = rnp.shape[0]
        shape = x.result_shape(family, H, W, n, m)
        k = shape[0]
        o = torch.empty(shape, dtype=L.dtype, device=dev)
        fam = x.FAMILIES[family]
        sel = [(0 if fam < 2 else (q % n if fam == 2 else q // m), q if fam < 3 else q % m) for q in check]
        rr = np.concatenate([oracle_family(0, lnp[i:i + 1].astype(np.float64), rnp[j:j + 1].astype(np.float64))
                             for i, j in sel])
        dd = np.concatenate([oracle_family(0, lnp[i:i + 1].astype(np.float64), rnp[j:j + 1].astype(np.float64),
                                           absval=True) for i, j in sel])
        res = {}
        for meth in methods:
            fn = lambda meth=meth: x.cross_corr(family, L, R, out=o, **METHODS[meth])
            try:
                ms = timed(fn, steps, args.warmup)
            except x.XcError as e:
                res[meth] = {"error": str(e)}
                continue
            err = norm_error(o[check].cpu().numpy(), rr, dd)[0]
            res[meth] = {"ms_per_step": ms, "pairs_per_s": k / (ms * 1e-3),
                         "gmac_per_s": k * float(H * W) ** 2 / (ms * 1e-3) / 1e9, "max_norm_err": err,
                         "last_kernel": x.last_launch()["kernel"]}
            log(f"{name} {meth}: {ms:.4f} ms, err {err:.2e}")
        a = res.get("auto", {})
        line = {"workload": workload, "pairs": k, "method": "auto", "auto_selected": auto_name(family, H, W, n, m)[0],
                "ms_per_step": a.get("ms_per_step"), "pairs_per_s": a.get("pairs_per_s"),
                "max_norm_err": a.get("max_norm_err"), "methods": res}
        if "ms_per_step" in a:
            t = kernels_of(lambda: x.cross_corr(family, L, R, out=o, precision="auto"), steps)
            pow2 = line["auto_selected"] == "fft64/2"
            line["roofline"] = roofline(t, lambda kk: kernel_bytes(kk, H, W, n if fam > 1 else 1, m, k, pow2),
                                        lambda kk: 2.0 * k * float(H * W) ** 2)
        if "ms_per_step" in res.get("fp32", {}) and "ms_per_step" in a:
            line["auto_vs_fp32_direct"] = res["fp32"]["ms_per_step"] / a["ms_per_step"]
        if "ms_per_step" in res.get("fp32", {}) and "ms_per_step" in res.get("bf16x3", {}):
            line["bf16x3_vs_fp32_direct"] = res["fp32"]["ms_per_step"] / res["bf16x3"]["ms_per_step"]
        return line

    if "c1" in cfgs:
        configs["c1"] = family_config("c1", "one_to_one 128x128 float32 U[0,1)", "one_to_one",
                                      seeded((1, 128, 128), "u01", 1, np.float32),
                                      seeded((1, 128, 128), "u01", 2, np.float32),
                                      ["fp32", "bf16x3", "fft64k", "auto"], [0], args.steps)
    if "c3" in cfgs:
        l3 = seeded((16, 64, 64), "normal", 3, np.float32)
        r3 = seeded((64, 64, 64), "normal", 4, np.float32)
        chk = [0, 63, 64 * 7 + 5, 1023]
        c3 = family_config("c3", "n_to_m 16 x 64 of 64x64 float32 N(0,1)", "n_to_m", l3, r3,
                           ["fp32", "bf16x3", "fp16x3", "bf16", "fp16", "fft64k", "auto"], chk, args.steps)
        d3 = family_config("c3-f64", "the same data in float64 (-d double)", "n_to_m", l3.astype(np.float64),
                           r3.astype(np.float64), ["fp32", "fft64k", "auto"], chk, args.steps)
        d3["methods"]["fp32"]["note"] = "float64 data: the FP64 direct kernel"
        c3["double"] = d3
        configs["c3"] = c3
    if "c4" in cfgs:
        lf, rf, pairs, olen, sizes = uneven_batch(128, 2024, lo=16, hi=512, per_group=4, dist="u11")
        L, R = torch.from_numpy(lf).to(dev), torch.from_numpy(rf).to(dev)
        O = torch.zeros(olen, device=dev)
        parr = x.pair_array(pairs)
        groups = []
        for g, nsz in enumerate(sizes):
            lo, ro, oo = pairs[4 * g][:3]
            groups.append((L[lo:lo + nsz * nsz].view(1, nsz, nsz), R[ro:ro + 4 * nsz * nsz].view(4, nsz, nsz),
                           O[oo:oo + 4 * (2 * nsz - 1) ** 2].view(4, 2 * nsz - 1, 2 * nsz - 1)))
        order = np.argsort(sizes)
        chk_groups = list(order[:3]) + sorted(range(len(sizes)), key=lambda g: abs(sizes[g] - 100))[:2]
        chk = [4 * g + j for g in chk_groups for j in (0, 3)]

        def err_c4():
            torch.cuda.synchronize()
            o = O.cpu().numpy()
            worst = 0.0
            for i in chk:
                lo, ro, oo, r, c = pairs[i]
                Lm = lf[lo:lo + r * c].reshape(1, r, c).astype(np.float64)
                Rm = rf[ro:ro + r * c].reshape(1, r, c).astype(np.float64)
                got = o[oo:oo + (2 * r - 1) * (2 * c - 1)].reshape(1, 2 * r - 1, 2 * c - 1)
                worst = max(worst, norm_error(got, oracle_family(0, Lm, Rm), oracle_family(0, Lm, Rm, absval=True))[0])
            return worst

        res = {}
        best, table = x.batch_autotune(parr, L, R, O)  # what AUTO batches replay
        O.zero_()
        ms = timed(lambda: x.batch(parr, L, R, O, precision="auto"), args.steps, args.warmup)
        res["batch_auto"] = {"ms_per_step": ms, "pairs_per_s": 512 / (ms * 1e-3), "max_norm_err": err_c4(),
                             "schedule (precision, tile, branches)": best}
        log(f"c4 batch_auto: {ms:.4f} ms err {res['batch_auto']['max_norm_err']:.2e}")
        for meth in ("fft64k", "bf16x3", "fp32"):
            def per_class(meth=meth):
                for Lg, Rg, Og in groups:
                    x.cross_corr("one_to_many", Lg, Rg, out=Og, **METHODS[meth])
            O.zero_()
            ms = timed(per_class, max(2, args.steps // 4) if meth != "fft64k" else args.steps, 1)
            res["per_class_" + meth] = {"ms_per_step": ms, "pairs_per_s": 512 / (ms * 1e-3),
                                        "max_norm_err": err_c4(), "launches": len(groups)}
            log(f"c4 per_class_{meth}: {ms:.4f} ms err {res['per_class_' + meth]['max_norm_err']:.2e}")
        a = res["batch_auto"]
        line = {"workload": "uneven n_to_mn, 128 groups of 1 left + 4 right square matrices, sizes round(16*32^U) "
                            "seed 2024, float32 U[-1,1)", "pairs": 512, "sizes_min_max": [min(sizes), max(sizes)],
                "method": "auto batch (HIP graph)", "ms_per_step": a["ms_per_step"], "pairs_per_s": a["pairs_per_s"],
                "max_norm_err": a["max_norm_err"], "methods": res,
                "batch_autotune_ms": [[list(sc), t] for sc, t in table],
                "auto_vs_fp32_direct": res["per_class_fp32"]["ms_per_step"] / a["ms_per_step"],
                "auto_vs_one_call_per_class": res["per_class_fft64k"]["ms_per_step"] / a["ms_per_step"],
                "bf16x3_vs_fp32_direct": res["per_class_fp32"]["ms_per_step"] / res["per_class_bf16x3"]["ms_per_step"]}
        # Roofline of the own-FFT kernels over the 128 size classes, timed on the one-call-per-class
        # schedule (the graph replay of the AUTO batch runs the same kernels, but the library's
        # kernel timing does not record launches under graph capture).
        t = kernels_of(lambda: [x.cross_corr("one_to_many", Lg, Rg, out=Og, precision="fft64", tile=1)
                                for Lg, Rg, Og in groups], 3)

        def c4_bytes(kname):
            tot = [kernel_bytes(kname, int(sz), int(sz), 1, 4, 4) for sz in sizes]
            return None if tot[0] is None else float(sum(tot))
        line["roofline"] = roofline(t, c4_bytes, lambda kk: sum(2.0 * 4 * float(sz) ** 4 for sz in sizes))
        line["roofline"]["schedule"] = "one own-FFT call per size class"
        configs["c4"] = line
        del O, L, R
        torch.cuda.empty_cache()
    if "c5" in cfgs:
        N, M = 1024, 128
        l5 = seeded((1, N, N), "u01", 50 + N, np.float32)
        r5 = seeded((M, N, N), "u01", 51 + N, np.float32)
        L, R = torch.from_numpy(l5).to(dev), torch.from_numpy(r5).to(dev)
        o5 = torch.empty((M, 2 * N - 1, 2 * N - 1), device=dev)

        def entry(t, dy, dx):  # the definition at one shift, FP64
            a = l5[0, max(0, dy):N + min(0, dy), max(0, dx):N + min(0, dx)].astype(np.float64)
            bm = r5[t, max(0, -dy):N - max(0, dy), max(0, -dx):N - max(0, dx)].astype(np.float64)
            return float((a * bm).sum())

        shifts = [(-(N - 1), -(N - 1)), (N - 1, N - 1), (-(N - 1), N - 1), (0, 0), (1, -1), (N // 2, -N // 3)]
        res = {}
        for meth in ("fft64k", "fft64", "auto"):
            fn = lambda meth=meth: x.cross_corr("one_to_many", L, R, out=o5, **METHODS[meth])
            ms = timed(fn, max(3, args.steps // 2), 1)
            err = 0.0
            for t in (0, 63, 127):  # U[0,1) data: corr(|L|,|R|) equals the output
                g = o5[t].cpu().numpy()
                for dy, dx in shifts:
                    w = entry(t, dy, dx)
                    err = max(err, abs(float(g[dy + N - 1, dx + N - 1]) - w) / w)
            res[meth] = {"ms_per_step": ms, "pairs_per_s": M / (ms * 1e-3),
                         "gmac_per_s": M * float(N) ** 4 / (ms * 1e-3) / 1e9, "max_norm_err": err,
                         "last_kernel": x.last_launch()["kernel"]}
            log(f"c5 {meth}: {ms:.4f} ms err {err:.2e}")
        a = res["auto"]
        line = {"workload": "one_to_many 1 x 128 of 1024x1024 float32 U[0,1)", "pairs": M, "method": "auto",
                "auto_selected": auto_name("one_to_many", N, N, 1, M)[0], "ms_per_step": a["ms_per_step"],
                "pairs_per_s": a["pairs_per_s"], "max_norm_err": a["max_norm_err"],
                "error_sample": "results 0, 63, 127 at 6 shifts (corners, centre, interior) against the definition",
                "methods": res}
        # every output of the AUTO result against the own FFT on the data in float64 (its error
        # is ~1e-16 of the largest output), reduced on the GPU in FP64; U[0,1) data:
        # corr(|L|,|R|) is the FP64 reference itself
        x.cross_corr("one_to_many", L, R, out=o5, precision="auto")
        ref5 = x.cross_corr("one_to_many", L.double(), R.double(), precision="fft64", tile=1)
        mx, mean, _ = x.norm_error(o5, ref5, ref5)
        del ref5
        line["all_outputs_vs_fp64_fft"] = {"max_norm_err": mx, "mean_norm_err": mean,
                                           "outputs": M * (2 * N - 1) ** 2}
        log(f"c5 auto, all outputs against the FP64 own FFT: max {mx:.2e}, mean {mean:.2e}")
        t = kernels_of(lambda: x.cross_corr("one_to_many", L, R, out=o5, precision="auto"), 3)
        pow2 = line["auto_selected"] == "fft64/2"
        line["roofline"] = roofline(t, lambda kk: kernel_bytes(kk, N, N, 1, M, M, pow2),
                                    lambda kk: 2.0 * M * float(N) ** 4)
        configs["c5_1024"] = line

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
        "fp32_direct_roofline": fp32_roof,
        "cpu_baseline": {"value": cpu_rate, "unit": "pairs/s", "cores": 1, "kind": "port",
                         "sample": f"{kcpu} of the 64 pairs, oracle FP32 restatement of cpu_cross_corr_one_to_many, "
                                   f"{cpu_s:.1f} s on {cpu_model()} ({os.cpu_count()} host cpus)"},
        "cpu_baseline_all_cores": {"value": ncores / all_s, "unit": "pairs/s", "cores": ncores, "kind": "port",
                                   "sample": f"{ncores} of the 64 pairs, one per thread, {all_s:.1f} s"},
        "configs": configs,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

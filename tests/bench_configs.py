"""Measurements of the workloads beside bench.py's headline (one MI355X, inputs in HBM):

  c1  one_to_one 128x128 float32 U[0,1)
  c3  n_to_m 16 left x 64 right 64x64 float32 N(0,1), and the same in float64 (-d double)
  c4  uneven n_to_mn: 128 groups, square size N_g = round(16 * 32^U), 1 left + 4 rights
      each (512 pairs), float32 U[-1,1): one persistent HIP-graph batch launch against one
      launch per size class
  c5  one_to_many 1 x 128 float32 U[0,1) at N in {32, 64, 128, 256, 512, 1024}: the
      definition-based methods against the FFT methods, the cost model's pick against the
      fastest method measured

Every method is timed with HIP events on the stream it runs on (W warm-up calls, then K
timed calls; K = 1 when one call takes over 250 ms) and its max normalised error
(|x - x64| / corr(|L|,|R|)) is measured against the FP64 oracle on a subset of results.
Beside them, the oracle's FP32 restatement of the reference's cpu_cross_corr_* (the
reference's CPU path) is timed on a bounded sample of each workload on the host cores
(`cpu_baseline`: one thread and up to 16 threads; samples are stated, larger workloads are
extrapolated by their MACs and marked so).  Prints one JSON line per configuration (and
progress lines on stderr).

    python tests/bench_configs.py [--configs c1,c3,c4,c5] [--steps 10] [--warmup 2]
"""
import argparse
import json
import pathlib
import sys
import time

REPO = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "tests"))

import numpy as np  # noqa: E402

ERR_BAR = 2e-6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, steps, warmup):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    once = time.perf_counter() - t0
    if once > 0.25:
        steps, warmup = 1, 0
    for _ in range(warmup):
        fn()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(s)
    for _ in range(steps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="c1,c3,c4,c5")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--c5-sizes", default="32,64,128,256,512,1024")
    args = ap.parse_args()
    import torch

    from xc_testutil import norm_error, oracle_family, seeded, uneven_batch, xcorr

    x = xcorr()
    dev = torch.device("cuda:0")
    cands = {c[0]: c for c in x.candidates()}

    def auto_name(family, rows, cols, n, m):
        o, src = x.select(family, rows, cols, n, m)
        names = [nm for nm, c in cands.items() if c[1:] == (o.precision, o.tile)]
        return (names[0] if names else "?"), src

    def family_run(family, L, R, out, meth):
        prec, tile = meth
        return lambda: x.cross_corr(family, L, R, out=out, precision=prec, tile=tile)

    METH = {"fp32": ("fp32", 0), "bf16x3": ("bf16x3", 0), "fp16x3": ("fp16x3", 0), "bf16": ("bf16", 0), "fp16": ("fp16", 0),
            "fft64": ("fft64", 0), "fft64k": ("fft64", 1), "auto": ("auto", 0)}

    def measure_family(name, family, left, right, methods, check, pairs):
        """left/right numpy arrays; check: result indices compared with the oracle."""
        L, R = torch.from_numpy(left).to(dev), torch.from_numpy(right).to(dev)
        k = x.result_shape(family, left.shape[1], left.shape[2], left.shape[0], right.shape[0])
        out = torch.empty(k, dtype=L.dtype, device=dev)
        ref = den = None
        if check is not None:
            fam = x.FAMILIES[family]
            sel_l, sel_r = [], []
            for q in check:
                if fam == 3:
                    sel_l.append(q // right.shape[0]); sel_r.append(q % right.shape[0])
                elif fam == 2:
                    sel_l.append(q % left.shape[0]); sel_r.append(q)
                else:
                    sel_l.append(0); sel_r.append(q)
            ref = np.concatenate([oracle_family(0, left[i:i + 1].astype(np.float64), right[j:j + 1].astype(np.float64))
                                  for i, j in zip(sel_l, sel_r)])
            den = np.concatenate([oracle_family(0, left[i:i + 1].astype(np.float64),
                                                right[j:j + 1].astype(np.float64), absval=True)
                                  for i, j in zip(sel_l, sel_r)])
        res = {}
        for m in methods:
            try:
                ms = timed(family_run(family, L, R, out, METH[m]), args.steps, args.warmup)
            except x.XcError as e:
                res[m] = {"error": str(e)}
                continue
            r = {"ms": ms, "pairs_per_s": pairs / (ms * 1e-3)}
            if ref is not None:
                r["max_norm_err"] = norm_error(out[check].cpu().numpy(), ref, den)[0]
            else:
                # beyond the oracle's reach: the sum rule, sum of a correlation = sum(L) * sum(R)
                got = out.double().sum(dim=(1, 2)).cpu().numpy()
                want = float(left.astype(np.float64).sum()) * right.astype(np.float64).reshape(right.shape[0], -1).sum(1)
                r["sum_rule_rel_err"] = float(np.max(np.abs(got - want) / np.abs(want)))
            res[m] = r
            log(f"{name} {m}: {ms:.4f} ms" + (f" err {r['max_norm_err']:.2e}" if "max_norm_err" in r else ""))
        return res

    import ctypes
    import os
    from concurrent.futures import ThreadPoolExecutor

    from xc_testutil import oracle

    ncores = max(1, min(os.cpu_count() or 1, 16))  # the host share of one GPU

    def cpu_family(family, left, right, k0, k1, threads):
        """Seconds the oracle's FP32 restatement of cpu_cross_corr_<family> takes for results
        [k0, k1) of a regular family on `threads` host threads."""
        lib = oracle()
        fam = x.FAMILIES[family]
        n, r, c = left.shape
        m = right.shape[0]
        buf = np.zeros((k1 - k0, 2 * r - 1, 2 * c - 1), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        lc, rc = np.ascontiguousarray(left, np.float32), np.ascontiguousarray(right, np.float32)
        t0 = time.perf_counter()
        lib.xco_family_range_f32(fam, lc.ctypes.data_as(fp), rc.ctypes.data_as(fp), buf.ctypes.data_as(fp),
                                 r, c, n, m, k0, k1, threads)
        return time.perf_counter() - t0

    def cpu_baseline(family, left, right, pairs, sample):
        """pairs/s of the CPU path on one thread and on ncores threads, timed on `sample`
        results (linear in pairs beyond them)."""
        t1 = cpu_family(family, left, right, 0, min(sample, pairs), 1) / min(sample, pairs)
        kall = min(max(ncores, sample), pairs)
        tn = cpu_family(family, left, right, 0, kall, ncores) / kall
        return {"value_1_thread": 1.0 / t1, "value_all_cores": 1.0 / tn, "unit": "pairs/s", "cores": ncores,
                "kind": "port", "sample": f"{min(sample, pairs)} results on 1 thread, {kall} on {ncores} threads"
                + ("" if kall >= pairs else f"; {pairs} extrapolated linearly")}

    def summary(name, workload, res, pairs, extra=None):
        auto = res.get("auto", {})
        fp32 = res.get("fp32", {})
        ok = {m: r for m, r in res.items() if "ms" in r and r.get("max_norm_err", 0.0) <= ERR_BAR}
        best = min(ok, key=lambda m: ok[m]["ms"]) if ok else None
        line = {"config": name, "workload": workload, "pairs": pairs, "methods": res,
                "fastest_within_2e-6": best,
                "auto_pairs_per_s": auto.get("pairs_per_s"),
                "auto_vs_fp32_direct": fp32["ms"] / auto["ms"] if "ms" in fp32 and "ms" in auto else None,
                "bf16x3_vs_fp32_direct": (fp32["ms"] / res["bf16x3"]["ms"]
                                          if "ms" in fp32 and "ms" in res.get("bf16x3", {}) else None)}
        if extra:
            line.update(extra)
        print(json.dumps(line), flush=True)

    cfgs = args.configs.split(",")
    if "c1" in cfgs:
        left = seeded((1, 128, 128), "u01", 1, np.float32)
        right = seeded((1, 128, 128), "u01", 2, np.float32)
        res = measure_family("c1", "one_to_one", left, right, ["fp32", "bf16x3", "fft64", "fft64k", "auto"], [0], 1)
        summary("c1", "one_to_one 128x128 f32 U[0,1)", res, 1,
                {"auto_selected": auto_name("one_to_one", 128, 128, 1, 1),
                 "cpu_baseline": cpu_baseline("one_to_one", left, right, 1, 1)})

    if "c3" in cfgs:
        left = seeded((16, 64, 64), "normal", 3, np.float32)
        right = seeded((64, 64, 64), "normal", 4, np.float32)
        chk = [0, 63, 64 * 7 + 5, 1023]
        res = measure_family("c3", "n_to_m", left, right,
                             ["fp32", "bf16x3", "fp16x3", "bf16", "fp16", "fft64", "fft64k", "auto"], chk, 1024)
        res64 = measure_family("c3-f64", "n_to_m", left.astype(np.float64), right.astype(np.float64),
                               ["fp32", "fft64", "fft64k"], chk, 1024)
        summary("c3", "n_to_m 16x64 of 64x64 f32 N(0,1); double data under 'double'", res, 1024,
                {"double": res64, "auto_selected": auto_name("n_to_m", 64, 64, 16, 64),
                 "cpu_baseline": cpu_baseline("n_to_m", left, right, 1024, 16)})

    if "c4" in cfgs:
        lf, rf, pairs, olen, sizes = uneven_batch(128, 2024, lo=16, hi=512, per_group=4, dist="u11")
        L, R = torch.from_numpy(lf).to(dev), torch.from_numpy(rf).to(dev)
        O = torch.zeros(olen, device=dev)
        groups = []
        for g, nsz in enumerate(sizes):
            lo, ro, oo = pairs[4 * g][:3]
            groups.append((nsz, L[lo:lo + nsz * nsz].view(1, nsz, nsz), R[ro:ro + 4 * nsz * nsz].view(4, nsz, nsz),
                           O[oo:oo + 4 * (2 * nsz - 1) ** 2].view(4, 2 * nsz - 1, 2 * nsz - 1)))
        order = np.argsort(sizes)
        chk_groups = list(order[:4]) + sorted(range(len(sizes)), key=lambda g: abs(sizes[g] - 100))[:2]
        chk_pairs = [4 * g + j for g in chk_groups for j in (0, 3)]

        def err_of():
            torch.cuda.synchronize()
            out = O.cpu().numpy()
            worst = 0.0
            for i in chk_pairs:
                lo, ro, oo, r, c = pairs[i]
                Lm = lf[lo:lo + r * c].reshape(1, r, c).astype(np.float64)
                Rm = rf[ro:ro + r * c].reshape(1, r, c).astype(np.float64)
                got = out[oo:oo + (2 * r - 1) * (2 * c - 1)].reshape(1, 2 * r - 1, 2 * c - 1)
                worst = max(worst, norm_error(got, oracle_family(0, Lm, Rm), oracle_family(0, Lm, Rm, absval=True))[0])
            return worst

        res = {}
        plan = x.BatchPlan(pairs, precision="bf16x3")
        fplan = x.BatchPlan(pairs, precision="fft64", tile=1)
        for nm, fn in [("batch_bf16x3_graph", lambda: plan.run(L, R, O, use_graph=True)),
                       ("batch_bf16x3_launch", lambda: plan.run(L, R, O, use_graph=False)),
                       ("batch_fft64k_graph", lambda: fplan.run(L, R, O, use_graph=True)),
                       ("batch_fft64k_stream", lambda: fplan.run(L, R, O, use_graph=False))]:
            O.zero_()
            ms = timed(fn, args.steps, args.warmup)
            res[nm] = {"ms": ms, "pairs_per_s": 512 / (ms * 1e-3), "max_norm_err": err_of(), "launches": 1}
            log(f"c4 {nm}: {ms:.4f} ms err {res[nm]['max_norm_err']:.2e}")
        plan.close()
        fplan.close()
        # AUTO batch: the schedule the batch autotuner measured fastest for these pairs
        parr = x.pair_array(pairs)
        best, table = x.batch_autotune(parr, L, R, O)
        O.zero_()
        ms = timed(lambda: x.batch(parr, L, R, O, precision="auto"), args.steps, args.warmup)
        res["batch_auto"] = {"ms": ms, "pairs_per_s": 512 / (ms * 1e-3), "max_norm_err": err_of(), "launches": 1}
        log(f"c4 batch_auto: {ms:.4f} ms err {res['batch_auto']['max_norm_err']:.2e} (schedule {best})")
        for m in ["fp32", "bf16x3", "fft64k", "auto"]:
            prec, tile = METH[m]

            def per_class(prec=prec, tile=tile):
                for nsz, Lg, Rg, Og in groups:
                    x.cross_corr("one_to_many", Lg, Rg, out=Og, precision=prec, tile=tile)
            O.zero_()
            ms = timed(per_class, args.steps, args.warmup)
            res["per_class_" + m] = {"ms": ms, "pairs_per_s": 512 / (ms * 1e-3), "max_norm_err": err_of(),
                                     "launches": len(groups)}
            log(f"c4 per_class_{m}: {ms:.4f} ms err {res['per_class_' + m]['max_norm_err']:.2e}")
        picks = {}
        for nsz in sizes:
            nm = auto_name("one_to_many", nsz, nsz, 1, 4)[0]
            picks[nm] = picks.get(nm, 0) + 1
        fp32 = res["per_class_fp32"]["ms"]
        # CPU path on a sample: 16 pairs spread over the sizes up to 200 (one thread each),
        # extrapolated to the 512 pairs by their MACs (the largest pairs alone would take minutes)
        small = sorted((i for i in range(len(pairs)) if pairs[i][3] <= 200), key=lambda i: pairs[i][3])
        samp = [small[int(v)] for v in np.linspace(0, len(small) - 1, 16)]

        def cpu_pair(i):
            lo, ro, _, r, c = pairs[i]
            return cpu_family("one_to_one", lf[lo:lo + r * c].reshape(1, r, c), rf[ro:ro + r * c].reshape(1, r, c),
                              0, 1, 1)
        t0 = time.perf_counter()
        with ThreadPoolExecutor(ncores) as ex:
            list(ex.map(cpu_pair, samp))
        wall = time.perf_counter() - t0
        macs = lambda i: float(pairs[i][3]) ** 2 * float(pairs[i][4]) ** 2
        frac = sum(macs(i) for i in samp) / sum(macs(i) for i in range(len(pairs)))
        cpu = {"value_all_cores": len(pairs) * frac / wall, "unit": "pairs/s", "cores": ncores, "kind": "port",
               "sample": f"16 pairs of sizes {pairs[samp[0]][3]}..{pairs[samp[-1]][3]} ({frac:.2e} of the MACs) "
                         f"in {wall:.1f} s on {ncores} threads; 512 pairs extrapolated by MACs"}
        line = {"config": "c4", "workload": "uneven n_to_mn, 128 groups of 1+4 square matrices, sizes "
                "round(16*32^U) seed 2024, f32 U[-1,1)", "pairs": 512, "sizes_min_max": [min(sizes), max(sizes)],
                "methods": res, "auto_picks_per_group": picks,
                "batch_autotune": {"best (precision, tile, branches)": best,
                                   "ms per schedule": [[list(sc), t] for sc, t in table]},
                "batch_auto_vs_fp32_direct": res["per_class_fp32"]["ms"] / res["batch_auto"]["ms"],
                "auto_vs_fp32_direct": fp32 / res["per_class_auto"]["ms"],
                "batch_graph_vs_per_class_bf16x3": res["per_class_bf16x3"]["ms"] / res["batch_bf16x3_graph"]["ms"],
                "batch_graph_vs_fp32_direct": fp32 / res["batch_bf16x3_graph"]["ms"],
                "fft_batch_graph_vs_fp32_direct": fp32 / res["batch_fft64k_graph"]["ms"],
                "fft_batch_graph_vs_per_class_fft64k": res["per_class_fft64k"]["ms"] / res["batch_fft64k_graph"]["ms"],
                "cpu_baseline": cpu}
        print(json.dumps(line), flush=True)

    if "c5" in cfgs:
        rows = []
        for nsz in [int(v) for v in args.c5_sizes.split(",")]:
            left = seeded((1, nsz, nsz), "u01", 50 + nsz, np.float32)
            right = seeded((128, nsz, nsz), "u01", 51 + nsz, np.float32)
            methods = ["bf16x3", "fft64", "fft64k", "auto"] + (["fp32"] if nsz <= 512 else [])
            res = measure_family(f"c5-{nsz}", "one_to_many", left, right, methods, [0] if nsz <= 256 else None, 128)
            timed_ok = {m: r["ms"] for m, r in res.items() if "ms" in r and m != "auto"
                        and r.get("max_norm_err", 0.0) <= ERR_BAR}
            best = min(timed_ok, key=timed_ok.get)
            pick, src = auto_name("one_to_many", nsz, nsz, 1, 128)
            rows.append({"N": nsz, "methods": res, "fastest_measured": best, "auto_selected": pick, "selected_by": src,
                         "auto_regret": res["auto"]["ms"] / timed_ok[best] if "ms" in res.get("auto", {}) else None,
                         "definition_over_fft": res["bf16x3"]["ms"] / min(res["fft64"]["ms"], res["fft64k"]["ms"])
                         if all("ms" in res.get(m, {}) for m in ("bf16x3", "fft64", "fft64k")) else None,
                         # N >= 512: over a minute per pair on one thread; not timed
                         "cpu_baseline": cpu_baseline("one_to_many", left, right, 128, 1) if nsz <= 256 else None})
        print(json.dumps({"config": "c5", "workload": "one_to_many 1x128 f32 U[0,1)", "sweep": rows}), flush=True)


if __name__ == "__main__":
    main()

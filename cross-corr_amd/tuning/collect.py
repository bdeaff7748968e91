"""Measures the timing table the cost model is trained on: every candidate method of
XC_PREC_AUTO on seeded random workloads, timed by the library's own exhaustive autotuner
(xc_autotune_f32: HIP events on the stream, one warm-up then `reps` runs).

    python cross-corr_amd/tuning/collect.py --count 300 --large 150 --seed 2024 --out cross-corr_amd/tuning/timings.csv
    (or in slices: --part 1/3 --out t1.csv, --part 2/3 ..., then concatenate the rows)

Inputs are U[0,1) float32 from a seed per workload.  The `--large` workloads (more work than
the definition-based methods can be timed on in a second) get times for the FFT methods only;
the other cells are "nan".  Runs on one GPU.
"""
import argparse
import importlib.util
import pathlib
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent


def load_package():
    spec = importlib.util.spec_from_file_location("cross_corr_amd", HERE.parent / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cross_corr_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=300)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--large", type=int, default=150, help="extra large workloads, FFT methods only")
    ap.add_argument("--out", default=str(HERE / "timings.csv"))
    ap.add_argument("--part", default="1/1",
                    help="i/k: time only the i-th of k contiguous slices of the workload list (rows keep "
                         "their order, so the slices' files concatenate, without their repeated header, into "
                         "the whole table)")
    args = ap.parse_args()
    import torch

    sys.path.insert(0, str(HERE))
    from workloads import random_workloads, num_results

    x = load_package()
    cands = [c[0] for c in x.candidates()]
    work = [(w, False) for w in random_workloads(args.count, args.seed)]
    work += [(w, True) for w in random_workloads(args.large, args.seed + 1, large=True)]
    part, parts = (int(v) for v in args.part.split("/"))
    lo, hi = len(work) * (part - 1) // parts, len(work) * part // parts
    g = torch.Generator(device="cuda")
    t0 = time.time()
    with open(args.out, "w") as f:
        f.write(",".join(["family", "rows", "cols", "n", "m"] + cands) + "\n")
        for i, ((fam, rows, cols, n, m), large) in enumerate(work):
            if not lo <= i < hi:
                continue
            g.manual_seed(args.seed * 1000 + i)
            L = torch.rand((n, rows, cols), generator=g, device="cuda")
            R = torch.rand((m, rows, cols), generator=g, device="cuda")
            O = torch.empty((num_results(fam, n, m), 2 * rows - 1, 2 * cols - 1), device="cuda")
            if large:
                # the definition-based methods would take seconds here: FFT methods only
                ms = [x.time_candidate(fam, L, R, O, j, reps=args.reps) if cands[j].startswith("fft")
                      else float("nan") for j in range(len(cands))]
                best = min((j for j in range(len(cands)) if ms[j] == ms[j]), key=lambda j: ms[j])
            else:
                ms, best = x.autotune(fam, L, R, O, reps=args.reps)
            f.write(",".join([str(x.FAMILIES[fam]), str(rows), str(cols), str(n), str(m)] +
                             [f"{v:.6g}" for v in ms]) + "\n")
            f.flush()
            del L, R, O
            if i % 10 == 0 or i + 1 == hi:
                print(f"[{i + 1}/{len(work)}] {fam} {rows}x{cols} n={n} m={m} best={cands[best]} "
                      f"{ms[best]:.3f} ms  ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()

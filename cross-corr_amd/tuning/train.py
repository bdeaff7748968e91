"""Fits the cost model to a timing table on the GPU (xc_cost_model_fit: full-batch Adam on
the library's MLP kernels, each epoch replayed from a HIP graph) and writes the model file
that XC_PREC_AUTO reads (tuning/cost_model.txt next to libxcorr.so by default).

    python cross-corr_amd/tuning/train.py --timings cross-corr_amd/tuning/timings.csv \
        --out cross-corr_amd/tuning/cost_model.txt --holdout 5

With --holdout K, rows i with i % K == 0 are left out of training (the test set of the
selection-quality test).
"""
import argparse
import pathlib

from collect import load_package

HERE = pathlib.Path(__file__).resolve().parent


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--timings", default=str(HERE / "timings.csv"))
    ap.add_argument("--out", default=str(HERE / "cost_model.txt"))
    ap.add_argument("--holdout", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=6000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--weight-decay", type=float, default=1e-4)
    ap.add_argument("--hidden", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()
    x = load_package()
    x.cost_model_fit(args.timings, args.out, epochs=args.epochs, lr=args.lr, weight_decay=args.weight_decay,
                     seed=args.seed, hidden=args.hidden, holdout_mod=args.holdout)
    print(open(args.out).read().splitlines()[-1], flush=True)


if __name__ == "__main__":
    main()

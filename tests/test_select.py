"""Method selection on the host (no GPU): the cost model's features, the model file, host
inference against a numpy forward pass of the same file, and what the trained model gains
on held-out workloads of the committed timing table over the same choice made without it."""
import ctypes
import pathlib

import numpy as np
import pytest

import mlp_ref
from xc_testutil import REPO, xcorr

TUNING = REPO / "cross-corr_amd" / "tuning"
MODEL = TUNING / "cost_model.txt"
TIMINGS = TUNING / "timings.csv"
FAM = ["one_to_one", "one_to_many", "n_to_mn", "n_to_m"]


def py_features(family, rows, cols, n, m):
    """Restatement of the feature definition (xc_select.cpp workload_features)."""
    k = {0: 1, 1: m, 2: m, 3: n * m}[family]
    transforms = {0: 3, 1: 1 + 2 * m, 2: n + 2 * m, 3: n + m + k}[family]
    lr, lc = np.log2(rows), np.log2(cols)

    def lpf(v):  # largest prime factor
        best, p = 1, 2
        while p * p <= v:
            while v % p == 0:
                best, v = p, v // p
            p += 1
        return max(best, v) if v > 1 else best

    def p2(v):
        return max(16, 1 << int(np.ceil(np.log2(v))))

    def row_len(c):  # the power of two, or 3 * 2^e (e = 4..9) when shorter
        return min([p2(2 * c - 1)] + [3 * 2 ** e for e in range(4, 10) if 3 * 2 ** e >= 2 * c - 1])

    pr, pc, pc2 = p2(2 * rows - 1), row_len(cols), p2(2 * cols - 1)
    pad = np.log2(pr * pc / ((2 * rows - 1) * (2 * cols - 1)))
    nin = {0: 2, 1: 1 + m, 2: n + m, 3: n + m}[family]
    own = nin * rows * pc * np.log2(pc) / 2 + k * (pc / 2 + 1) * pr * np.log2(pr) * 2 + k * pr * pc * np.log2(pc) / 2
    return [lr, lc, np.log2(n), np.log2(m), np.log2(k), np.log2(k) + 2 * lr + 2 * lc,
            np.log2(transforms) + 2 + lr + lc, pad, np.log2(lpf(2 * rows)), np.log2(lpf(2 * cols))] + \
        [float(family == i) for i in range(4)] + [np.log2(own), np.log2(pc2 / pc)]


def parse_model(path):
    tok = open(path).read().split()
    assert tok[0] == "xcorr-cost-model" and tok[1] == "2"
    i = 2
    out = {}
    while i < len(tok) and tok[i] != "trained_rows":
        key, n = tok[i], int(tok[i + 1])
        vals = tok[i + 2:i + 2 + n]
        out[key] = vals if key == "candidates" else np.array(vals, np.float64)
        i += 2 + n
    return out


def read_table(path):
    lines = open(path).read().strip().splitlines()
    head = lines[0].split(",")
    w, t = [], []
    for line in lines[1:]:
        c = line.split(",")
        w.append(tuple(int(v) for v in c[:5]))
        t.append([float(v) for v in c[5:]])
    return head[5:], w, np.array(t)


@pytest.fixture(scope="module")
def x():
    x = xcorr()
    x.cost_model_load(str(MODEL))
    return x


def test_features_match_definition(x):
    for fam, r, c, n, m in [(0, 16, 16, 1, 1), (1, 256, 256, 1, 64), (2, 100, 37, 3, 12), (3, 64, 64, 16, 64),
                            (0, 1, 1, 1, 1), (0, 177, 1024, 1, 1), (1, 300, 300, 1, 5), (3, 40, 150, 2, 3)]:
        np.testing.assert_allclose(x.features(FAM[fam], r, c, n, m), py_features(fam, r, c, n, m), rtol=1e-6)


def test_host_inference_matches_numpy_forward(x):
    mdl = parse_model(MODEL)
    dims = [int(v) for v in mdl["dims"]]
    assert mdl["candidates"] == [c[0] for c in x.candidates()]
    for fam, r, c, n, m in [(1, 256, 256, 1, 64), (3, 64, 64, 16, 64), (0, 40, 900, 1, 1)]:
        f = np.clip(np.array(py_features(fam, r, c, n, m)), mdl["feat_min"], mdl["feat_max"])
        f = (f - mdl["feat_mean"]) / mdl["feat_std"]
        y, _ = mlp_ref.forward(dims, mdl["params"], f[None, :])
        want = 2.0 ** (y[0] * mdl["out_std"] + mdl["out_mean"])
        np.testing.assert_allclose(x.predict(FAM[fam], r, c, n, m), want, rtol=1e-5)


def test_select_sources(x):
    o, src = x.select("one_to_many", 256, 256, 1, 64)
    assert src == "model"
    ms = x.predict("one_to_many", 256, 256, 1, 64)
    assert x.candidates()[int(np.argmin(ms))][1:] == (o.precision, o.tile)
    # a matrix size far outside the training range is not answered by the model
    assert x.select("one_to_one", 8192, 8192, 1, 1)[1] == "none"
    # far beyond the work the definition-based methods were timed on, only FFT methods qualify
    o, src = x.select("one_to_many", 1024, 1024, 1, 128)
    assert src == "model" and o.precision == 5


def test_model_file_is_checked(x, tmp_path):
    text = MODEL.read_text()
    bad = tmp_path / "other_candidates.txt"
    bad.write_text(text.replace("fft64/0", "fft64/9", 1))
    with pytest.raises(x.XcError) as e:
        x.cost_model_load(str(bad))
    assert e.value.status == 4
    trunc = tmp_path / "truncated.txt"
    trunc.write_text(text[: len(text) // 2])
    with pytest.raises(x.XcError):
        x.cost_model_load(str(trunc))
    with pytest.raises(x.XcError):
        x.cost_model_load(str(tmp_path / "missing.txt"))
    x.cost_model_load(str(MODEL))  # the good model is back


def test_trained_model_gain_on_held_out_workloads(x):
    """Held-out rows (index % 5 == 0, never seen in training) of the committed timing table.
    The model's picks are compared with
      - exhaustive autotuning (the measured best; it must run every candidate once per new
        workload, which is the cost the network removes),
      - the best single fixed method chosen on the training rows,
      - the methods without it: the definition-based FP32 kernel (the reference's algorithm)
        and the bf16x3 MFMA kernel with its default tile."""
    cands, w, t = read_table(TIMINGS)
    assert cands == [c[0] for c in x.candidates()]
    hold = np.arange(len(w)) % 5 == 0
    train = ~hold
    # best fixed method on the training rows, among those measured on every row (the large
    # workloads have times for the FFT methods only)
    everywhere = ~np.isnan(t).any(0)
    logsum = np.where(everywhere, np.log(np.where(np.isnan(t), 1.0, t))[train].sum(0), np.inf)
    fixed = int(np.argmin(logsum))
    picks = []
    for (fam, r, c, n, m) in np.array(w)[hold]:
        o, src = x.select(FAM[fam], int(r), int(c), int(n), int(m))
        assert src == "model"
        picks.append([i for i, cc in enumerate(x.candidates()) if cc[1:] == (o.precision, o.tile)][0])
    th = t[hold]
    model_t = th[np.arange(len(th)), picks]
    assert not np.isnan(model_t).any(), "the model picked a method that was not measured there"
    best_t = np.nanmin(th, 1)
    fixed_t = th[:, fixed]
    gm = lambda v: float(np.exp(np.log(v).mean()))
    fp32 = cands.index("fp32/0")
    # the bf16x3 default tile: 6 up to 64 columns, 9 above
    bf_of = np.array([cands.index("bf16x3/6" if int(wl[2]) <= 64 else "bf16x3/9") for wl in np.array(w)[hold]])
    both = ~np.isnan(th[:, fp32])  # rows where the definition-based methods were timed
    report = {"held_out": int(hold.sum()), "model_total_ms": model_t.sum(), "exhaustive_total_ms": best_t.sum(),
              "autotune_cost_ms": np.nansum(th, 1).sum(), "fixed_method": cands[fixed],
              "fixed_total_ms": fixed_t.sum(), "model_regret_gm": gm(model_t / best_t),
              "fixed_regret_gm": gm(fixed_t / best_t),
              "fp32_direct_over_model": th[both, fp32].sum() / model_t[both].sum(),
              "bf16x3_default_over_model": th[np.arange(len(th)), bf_of][both].sum() / model_t[both].sum()}
    print(report)
    assert report["model_total_ms"] <= 1.02 * report["exhaustive_total_ms"]
    assert report["model_total_ms"] <= 1.02 * report["fixed_total_ms"]
    assert report["model_regret_gm"] < 1.05
    assert report["autotune_cost_ms"] > 10 * report["model_total_ms"]
    assert report["fp32_direct_over_model"] > 10 and report["bf16x3_default_over_model"] > 3

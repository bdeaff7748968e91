"""Exhaustive autotuning of XC_PREC_AUTO (xc_autotune_f32) on the GPU: every candidate is
timed on the caller's buffers, the fastest is recorded in the process's autotune table and
AUTO then uses it ahead of the cost model; beyond 1e13 MACs only the FFT candidates run.
Results of AUTO after tuning are checked against the FP64 oracle."""
import math

import numpy as np
import pytest

from xc_testutil import norm_error, oracle_family, seeded, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_autotune_times_every_candidate_and_auto_uses_the_winner():
    x = xcorr()
    x.autotune_clear()
    left = seeded((1, 40, 36), "u01", 31, np.float32)
    right = seeded((5, 40, 36), "u11", 32, np.float32)
    L, R = dev(left), dev(right)
    out = torch.empty((5, 79, 71), device="cuda")
    ms, best = x.autotune("one_to_many", L, R, out, reps=2)
    cands = x.candidates()
    assert len(ms) == len(cands) and all(math.isfinite(t) and t > 0 for t in ms), ms
    assert best == int(np.argmin(ms))
    o, src = x.select("one_to_many", 40, 36, 1, 5)
    assert src == "tuned" and (o.precision, o.tile) == cands[best][1:]
    got = x.cross_corr("one_to_many", L, R, precision="auto")
    torch.cuda.synchronize()
    l64, r64 = left.astype(np.float64), right.astype(np.float64)
    err, _ = norm_error(got.cpu().numpy(), oracle_family("one_to_many", l64, r64),
                        oracle_family("one_to_many", l64, r64, absval=True))
    assert err <= 2e-6, err
    x.autotune_clear()
    assert x.select("one_to_many", 40, 36, 1, 5)[1] == "model"


def test_autotune_beyond_1e13_macs_runs_only_the_fft_candidates():
    """10 pairs of 1024^2 (1.1e13 MACs): the definition-based candidates (about a second
    each per run here) are not run; the own FFT wins (tile 1 or 2: at 1024 columns both take
    2048-point rows, so the two run the same kernels and either may time faster)."""
    x = xcorr()
    x.autotune_clear()
    L = torch.rand((1, 1024, 1024), device="cuda")
    R = torch.rand((10, 1024, 1024), device="cuda")
    out = torch.empty((10, 2047, 2047), device="cuda")
    ms, best = x.autotune("one_to_many", L, R, out, reps=1)
    cands = x.candidates()
    for (name, prec, tile), t in zip(cands, ms):
        if prec == x.PREC_FFT64:
            assert math.isfinite(t), (name, t)
        else:
            assert t == float("inf"), (name, t)
    assert cands[best][1] == x.PREC_FFT64 and cands[best][2] in (1, 2)
    # sum rule on the tuned AUTO result: sum of a correlation = sum(L) * sum(R)
    got = x.cross_corr("one_to_many", L, R, out=out, precision="auto")
    s = got.double().sum(dim=(1, 2))
    want = L.double().sum() * R.double().sum(dim=(1, 2))
    assert torch.max(torch.abs(s - want) / want).item() < 1e-9
    x.autotune_clear()

"""A single library call captured into a HIP graph (torch.cuda.CUDAGraph over hipGraph) after a
warm-up call on the same stream (which sizes the stream's workspaces and loads the tables), and
replayed: every replay must give the eager result, with other work and synchronisations
between the replays.  The launch paths record kernels only (no memset or copy nodes)."""
import numpy as np
import pytest

from xc_testutil import seeded, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("prec,kw", [("fp32", {}), ("bf16x3", {}), ("fp16x3", {}), ("fft64", {"tile": 1}),
                                     ("fft64", {"tile": 2}), ("auto", {})])
def test_single_call_replays_from_a_graph(prec, kw):
    x = xcorr()
    L = torch.from_numpy(seeded((1, 96, 80), "u01", 41, np.float32)).cuda()
    R = torch.from_numpy(seeded((8, 96, 80), "u01", 42, np.float32)).cuda()
    shape = (8, 191, 159)
    eager = torch.empty(shape, device="cuda")
    out = torch.empty(shape, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        x.cross_corr("one_to_many", L, R, out=eager, precision=prec, **kw)  # warm-up on s
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        x.cross_corr("one_to_many", L, R, out=out, precision=prec, **kw)
    # the FP32 direct kernel adds chunks of rows atomically at this size: equal to rounding
    tol = 1e-6 * float(eager.abs().max()) if prec == "fp32" else 0.0
    for rep in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(out).all(), rep
        assert float((out - eager).abs().max()) <= tol, (rep, prec)
        _ = (eager * 2).sum().item()  # eager work between replays

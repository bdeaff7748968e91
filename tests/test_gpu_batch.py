"""Uneven batch (n_to_mn-style groups of different sizes) through xc_batch_f32 and reusable
plans: one persistent MFMA launch, or the own FFT kernels per size class on parallel graph
branches, captured into a HIP graph and replayed.  Every pair is checked against the FP64
oracle (bf16x3: <= 2e-6 normalised; fp32 per-pair direct and FFT64: <= 1e-6)."""
import numpy as np
import pytest

from xc_testutil import batch_oracle_check, uneven_batch, xcorr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3"])
def test_batch_small_uneven_bf16x3(prec):
    lf, rf, pairs, olen, sizes = uneven_batch(0, 3, sizes=[1, 2, 5, 16, 33, 64, 70, 100, 140, 31])
    out = torch.full((olen,), float("nan"), device="cuda")
    xcorr().batch(pairs, dev(lf), dev(rf), out, precision=prec)
    torch.cuda.synchronize()
    assert not torch.isnan(out).any()
    batch_oracle_check(lf, rf, pairs, out.cpu().numpy(), 2e-6)


def test_batch_fp32_per_pair():
    lf, rf, pairs, olen, _ = uneven_batch(0, 4, sizes=[3, 17, 50, 96])
    out = torch.full((olen,), float("nan"), device="cuda")
    xcorr().batch(pairs, dev(lf), dev(rf), out, precision="fp32")
    torch.cuda.synchronize()
    batch_oracle_check(lf, rf, pairs, out.cpu().numpy(), 1e-6)


def test_batch_plan_graph_replays_and_pointer_update():
    x = xcorr()
    lf, rf, pairs, olen, _ = uneven_batch(6, 5, lo=16, hi=200)
    plan = x.BatchPlan(pairs, "bf16x3")
    L, R = dev(lf), dev(rf)
    outs = [torch.full((olen,), float("nan"), device="cuda") for _ in range(2)]
    ref = None
    for rep in range(4):
        o = outs[rep % 2]  # alternate output buffers: the graph's kernel arguments are updated
        o.fill_(float("nan"))
        plan.run(L, R, o, use_graph=True)
        torch.cuda.synchronize()
        # eager work between replays
        _ = (L * 2.0).sum().item()
        got = o.cpu().numpy()
        assert not np.isnan(got).any(), rep
        if ref is None:
            batch_oracle_check(lf, rf, pairs, got, 2e-6)
            ref = got
        else:
            np.testing.assert_array_equal(got, ref)
    # the same plan without the graph
    o = torch.full((olen,), float("nan"), device="cuda")
    plan.run(L, R, o, use_graph=False)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(o.cpu().numpy(), ref)
    plan.close()


def test_batch_c4_shape_subset():
    """The 128-group, 512-pair workload (sizes log-uniform in [16, 512], U[-1,1)): all pairs
    computed in one launch; the oracle checks the 8 smallest groups and 2 groups near size 128
    (the FP64 oracle is single-threaded per pair: a 512x512 pair costs it minutes), and the sum
    rule (sum of a correlation = sum(L) * sum(R)) checks every pair."""
    lf, rf, pairs, olen, sizes = uneven_batch(128, 2025)
    out = torch.empty((olen,), device="cuda")
    xcorr().batch(pairs, dev(lf), dev(rf), out, precision="bf16x3")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    order = np.argsort(sizes)
    mid = sorted(range(len(sizes)), key=lambda g: abs(sizes[g] - 128))[:2]
    groups = list(order[:8]) + mid
    batch_oracle_check(lf, rf, pairs, got, 2e-6, check_pairs=[4 * g + j for g in groups for j in (0, 3)])
    for lo, ro, oo, r, c in pairs:
        s = got[oo:oo + (2 * r - 1) * (2 * c - 1)].astype(np.float64).sum()
        e = lf[lo:lo + r * c].astype(np.float64).sum() * rf[ro:ro + r * c].astype(np.float64).sum()
        den = np.abs(lf[lo:lo + r * c]).astype(np.float64).sum() * np.abs(rf[ro:ro + r * c]).astype(np.float64).sum()
        assert abs(s - e) / den < 2e-6


def test_batch_invalid_inputs_rejected():
    x = xcorr()
    with pytest.raises(x.XcError):
        x.BatchPlan([(0, 0, 0, 0, 4)], "bf16x3")
    with pytest.raises(x.XcError):
        x.BatchPlan([(0, 0, 0, 4, 4)], "fft64")


def _mixed_pairs(seed):
    """Pairs of several shapes, some sharing a left or a right matrix, in one flat buffer."""
    rng = np.random.default_rng(seed)
    shapes = [(1, 1), (2, 3), (9, 5), (5, 9), (16, 16), (33, 20), (64, 64), (100, 37)]
    lefts, rights, pairs = [], [], []
    loff = roff = ooff = 0
    for (r, c) in shapes:
        nl, nr = int(rng.integers(1, 3)), int(rng.integers(1, 4))
        lo = [loff + i * r * c for i in range(nl)]
        ro = [roff + i * r * c for i in range(nr)]
        lefts.append(rng.random(nl * r * c).astype(np.float32))
        rights.append((rng.random(nr * r * c) * 2 - 1).astype(np.float32))
        for i in range(nl):
            for j in range(nr):
                pairs.append((lo[i], ro[j], ooff, r, c))
                ooff += (2 * r - 1) * (2 * c - 1)
        loff += nl * r * c
        roff += nr * r * c
    return np.concatenate(lefts), np.concatenate(rights), pairs, ooff


def test_batch_fft_plan_mixed_shapes_shared_matrices():
    x = xcorr()
    lf, rf, pairs, olen = _mixed_pairs(7)
    plan = x.BatchPlan(pairs, "fft64", tile=1)
    out = torch.full((olen,), float("nan"), device="cuda")
    plan.run(dev(lf), dev(rf), out, use_graph=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert not np.isnan(got).any()
    batch_oracle_check(lf, rf, pairs, got, 1e-6)
    plan.close()


def test_batch_fft_plan_replays_recapture_and_eager():
    x = xcorr()
    lf, rf, pairs, olen, _ = uneven_batch(12, 8, lo=16, hi=300)
    plan = x.BatchPlan(pairs, "fft64", tile=1)
    L, R = dev(lf), dev(rf)
    outs = [torch.full((olen,), float("nan"), device="cuda") for _ in range(2)]
    ref = None
    for rep in range(4):
        o = outs[rep % 2]  # alternating outputs: the graph is captured again for new pointers
        o.fill_(float("nan"))
        plan.run(L, R, o, use_graph=True)
        torch.cuda.synchronize()
        _ = (L * 2.0).sum().item()  # eager work between replays
        got = o.cpu().numpy()
        assert not np.isnan(got).any(), rep
        if ref is None:
            order = np.argsort([p[3] for p in pairs])
            batch_oracle_check(lf, rf, pairs, got, 1e-6, check_pairs=list(order[:24]))
            ref = got
        else:
            np.testing.assert_array_equal(got, ref)
    o = torch.full((olen,), float("nan"), device="cuda")
    plan.run(L, R, o, use_graph=False)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(o.cpu().numpy(), ref)
    plan.close()


def test_batch_auto_is_the_fft_plan_on_the_c4_shape():
    """C4 shape through xc_batch_f32 with AUTO: oracle on the smallest groups and two near
    size 128, the sum rule on every pair."""
    lf, rf, pairs, olen, sizes = uneven_batch(128, 2026)
    out = torch.full((olen,), float("nan"), device="cuda")
    x = xcorr()
    x.batch(pairs, dev(lf), dev(rf), out, precision="auto")
    torch.cuda.synchronize()
    assert x.last_launch()["kernel"] == "fft batch graph"
    got = out.cpu().numpy()
    assert not np.isnan(got).any()
    order = np.argsort(sizes)
    mid = sorted(range(len(sizes)), key=lambda g: abs(sizes[g] - 128))[:2]
    groups = list(order[:8]) + mid
    batch_oracle_check(lf, rf, pairs, got, 1e-6, check_pairs=[4 * g + j for g in groups for j in (0, 3)])
    for lo, ro, oo, r, c in pairs:
        s = got[oo:oo + (2 * r - 1) * (2 * c - 1)].astype(np.float64).sum()
        e = lf[lo:lo + r * c].astype(np.float64).sum() * rf[ro:ro + r * c].astype(np.float64).sum()
        den = np.abs(lf[lo:lo + r * c]).astype(np.float64).sum() * np.abs(rf[ro:ro + r * c]).astype(np.float64).sum()
        assert abs(s - e) / den < 1e-6


def test_batch_fft_plan_rejects_oversize():
    x = xcorr()
    with pytest.raises(x.XcError) as e:
        x.BatchPlan([(0, 0, 0, 2049, 4)], "fft64", tile=1)
    assert e.value.status == 4


def test_batch_autotune_times_every_schedule_and_auto_replays_the_fastest():
    """The batch autotuner times the own-FFT plan over 1, 2, 4 and 8 branches and the bf16x3
    plan on the call's buffers; an AUTO batch of the same pairs then runs the fastest."""
    x = xcorr()
    lf, rf, pairs, olen, sizes = uneven_batch(24, 9, lo=16, hi=300)
    L, R = dev(lf), dev(rf)
    out = torch.full((olen,), float("nan"), device="cuda")
    best, table = x.batch_autotune(pairs, L, R, out)
    assert [s for s, _ in table] == [(5, 1, 1), (5, 1, 2), (5, 1, 4), (5, 1, 8), (2, 0, 0)]
    assert all(np.isfinite(ms) and ms > 0 for _, ms in table)
    assert best == min(table, key=lambda t: t[1])[0]
    fft_best = min(ms for s, ms in table if s[0] == 5)
    assert fft_best < dict(table)[(2, 0, 0)], table  # the own FFT beats the definition here
    out2 = torch.full((olen,), float("nan"), device="cuda")
    x.batch(pairs, L, R, out2, precision="auto")
    torch.cuda.synchronize()
    ll = x.last_launch()
    assert ll["kernel"] == "fft batch graph" and ll["split"] == best[2], (ll, best)
    got = out2.cpu().numpy()
    assert not np.isnan(got).any()
    order = np.argsort(sizes)
    batch_oracle_check(lf, rf, pairs, got, 1e-6, check_pairs=[4 * g + j for g in order[:6] for j in (0, 3)])


def test_batch_fft_plan_branch_counts_agree():
    """The branch count only changes the schedule: every count gives the same bits."""
    x = xcorr()
    lf, rf, pairs, olen, _ = uneven_batch(10, 13, lo=16, hi=200)
    L, R = dev(lf), dev(rf)
    ref = None
    for b in (1, 3, 8):
        plan = x.BatchPlan(pairs, "fft64", tile=1, branches=b)
        out = torch.full((olen,), float("nan"), device="cuda")
        plan.run(L, R, out, use_graph=True)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        plan.close()
        if ref is None:
            batch_oracle_check(lf, rf, pairs, got, 1e-6, check_pairs=list(range(0, len(pairs), 5)))
            ref = got
        else:
            np.testing.assert_array_equal(got, ref)
    with pytest.raises(x.XcError):
        x.BatchPlan(pairs, "fft64", tile=1, branches=9)


def test_batch_fft_plan_ragged_class_of_one_padded_size():
    """Pairs of different sizes with the same padded lengths (256 x 128) share one FFT job with
    a per-matrix size table; one left region is also read as two different sizes."""
    rng = np.random.default_rng(21)
    shapes = [(100, 37), (90, 50), (65, 64), (128, 33), (70, 40), (10, 10), (12, 10)]
    lf = rng.random(200 * 200).astype(np.float32)
    rf = (rng.random(200 * 200) * 2 - 1).astype(np.float32)
    pairs, ooff = [], 0
    roff = 0
    for i, (r, c) in enumerate(shapes):
        lo = 0 if (r, c) in ((10, 10), (12, 10)) else 1000 * i  # (10,10) and (12,10): one class, one left region
        pairs.append((lo, roff, ooff, r, c))
        roff += r * c
        ooff += (2 * r - 1) * (2 * c - 1)
    x = xcorr()
    plan = x.BatchPlan(pairs, "fft64", tile=1)
    out = torch.full((ooff,), float("nan"), device="cuda")
    plan.run(dev(lf), dev(rf), out, use_graph=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    plan.close()
    assert not np.isnan(got).any()
    batch_oracle_check(lf, rf, pairs, got, 1e-6)


@pytest.mark.parametrize("prec", ["fft64k", "bf16x3"])
def test_batch_on_two_streams_at_once(prec):
    """xc_batch_f32 caches its plans per (pair list, device, stream): the same pair list run
    on two streams at the same time, with different data, gives each stream its own
    single-stream results (bit for bit: the kernels are deterministic)."""
    x = xcorr()
    lf, rf, pairs, olen, _ = uneven_batch(0, 6, sizes=[20, 33, 64, 90, 129])
    arr = x.pair_array(pairs)
    data = [(dev(lf * (1.0 + i)), dev(rf[::-1].copy() if i else rf)) for i in range(2)]

    def run(L, R, out):
        if prec == "fft64k":  # the own-FFT plan over two graph branches
            x.batch(arr, L, R, out, precision="fft64", tile=1, split=2)
        else:
            x.batch(arr, L, R, out, precision="bf16x3")
    want = []
    for L, R in data:
        o = torch.empty(olen, device="cuda")
        run(L, R, o)
        want.append(o)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    got = [[], []]
    print(f"batch on two streams, {prec}", flush=True)
    for _ in range(3):
        for i in range(2):
            with torch.cuda.stream(streams[i]):
                o = torch.empty(olen, device="cuda")
                run(*data[i], o)
                got[i].append(o)
    torch.cuda.synchronize()
    for i in range(2):
        for g in got[i]:
            assert torch.equal(g, want[i])
    batch_oracle_check(lf, rf, pairs, want[0].cpu().numpy(), 2e-6 if prec == "bf16x3" else 1e-6)

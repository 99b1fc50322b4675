"""Calls whose candidates exceed what one candidate pass may hold
(tmx_cfg.max_candidates, at most 2^31 - 1) are split into halves of their
templates and give the same answers as one pass; one template alone over the
limit is an error.  The limit is lowered here so that a small call crosses it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.fixture(scope="module")
def work():
    from tremor_amd import synth
    rng = np.random.default_rng(77)
    n = 1 << 20
    x = synth.bandpassed_noise(n, rng)
    onsets, _ = synth.inject_events(x, 40, rng)
    tm = synth.cut_templates(x, onsets, [int(v) for v in rng.integers(250, 700, size=24)], rng)
    return x, tm


@pytest.mark.parametrize("prec,fft", [("bf16x3", "off"), ("bf16x3", "always")])
def test_threshold_split_equals_one_pass(oracle, work, prec, fft):
    from tremor_amd.capi import Context
    x, tm = work
    rho = 0.2
    with Context(precision=prec, fft=fft) as ctx:
        ctx.load_waveform(x)
        r0 = ctx.threshold(tm, rho, 0)
    total = int(r0.stats["candidates"])
    per = np.bincount(r0.tmpl, minlength=len(tm))
    limit = max(total // 5, 3 * int(per.max()) + 4096)
    _log("threshold split", prec=prec, fft=fft, candidates=total, max_hits=int(per.max()), limit=limit)
    assert r0.stats["splits"] == 0
    assert limit < total, "the workload must cross the lowered limit"
    with Context(precision=prec, fft=fft, max_candidates=limit) as ctx:
        ctx.load_waveform(x)
        r1 = ctx.threshold(tm, rho, 0)
        r2 = ctx.threshold(tm, rho, 8)
    print("  stats", {k: r1.stats[k] for k in ("splits", "candidates", "comparisons")}, flush=True)
    assert r1.stats["splits"] >= 2
    assert r1.stats["comparisons"] == r0.stats["comparisons"]
    assert np.array_equal(r1.tmpl, r0.tmpl) and np.array_equal(r1.pos, r0.pos)
    assert np.array_equal(r1.counter, r0.counter)
    assert np.array_equal(r1.d.view(np.uint32), r0.d.view(np.uint32))
    tt, tp, td = oracle.scan_threshold(x, tm, rho, 8)
    assert np.array_equal(r2.tmpl, tt) and np.array_equal(r2.pos, tp)
    assert np.array_equal(r2.d.view(np.uint32), td.view(np.uint32))


def test_knn_split_equals_one_pass(oracle, work):
    from tremor_amd.capi import Context
    x, tm = work
    with Context(precision="bf16x3", fft="off") as ctx:
        ctx.load_waveform(x)
        p0, d0, s0 = ctx.knn(tm, 50, 16)
    limit = max(int(s0["candidates"]) // 3, 1)
    _log("kNN split", candidates=s0["candidates"], limit=limit)
    with Context(precision="bf16x3", fft="off", max_candidates=limit) as ctx:
        ctx.load_waveform(x)
        p1, d1, s1 = ctx.knn(tm, 50, 16)
    assert s1["splits"] >= 1
    assert np.array_equal(p1, p0) and np.array_equal(d1.view(np.uint32), d0.view(np.uint32))
    po, do, _ = oracle.scan_knn(x, tm, 50, 16)
    assert np.array_equal(p1, po) and np.array_equal(d1.view(np.uint32), do.view(np.uint32))


def test_one_template_over_the_limit_is_an_error(work):
    from tremor_amd.capi import Context, TmxError
    x, tm = work
    with Context(precision="bf16x3", fft="off", max_candidates=10) as ctx:
        ctx.load_waveform(x)
        with pytest.raises(TmxError, match="one template has more candidate windows"):
            ctx.threshold(tm[:3], 0.3, 0)
        # the same error again (nothing left behind by the failed call)
        with pytest.raises(TmxError, match="one template has more candidate windows"):
            ctx.threshold(tm[:1], 0.3, 0)


def test_more_templates_than_one_batch(oracle):
    """A call of 70,000 templates runs as batches of at most 65,535 in caller
    order: every template's answer equals the one it gets in a call of its
    own group of ten, across the batch seam, and matches the oracle."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    rng = np.random.default_rng(91)
    n = 1 << 13
    x = synth.bandpassed_noise(n, rng)
    count, m = 70000, 32
    starts = rng.integers(0, n - m, size=count)
    tm = [x[s:s + m] + rng.normal(0.0, 0.1, m).astype(np.float32) for s in starts]
    seam = list(range(65530, 65540))
    _log("many templates", count=count, m=m, n=n)
    with Context(precision="bf16x3", fft="off", scan="off") as ctx:
        ctx.load_waveform(x)
        pos, d, _ = ctx.knn(tm, 3, 4)
        r = ctx.threshold(tm, 0.97, 0)
        _log("many templates: groups of ten")
        for lo in (0, 65530, 69990):
            ids = list(range(lo, lo + 10))
            sub = [tm[i] for i in ids]
            ps, ds, _ = ctx.knn(sub, 3, 4)
            assert np.array_equal(pos[lo:lo + 10], ps) and np.array_equal(d[lo:lo + 10].view(np.uint32),
                                                                         ds.view(np.uint32))
            rs = ctx.threshold(sub, 0.97, 0)
            keep = (r.tmpl >= lo) & (r.tmpl < lo + 10)
            assert np.array_equal(r.tmpl[keep] - lo, rs.tmpl) and np.array_equal(r.pos[keep], rs.pos)
            assert np.array_equal(r.d[keep].view(np.uint32), rs.d.view(np.uint32))
    sub = [tm[i] for i in seam]
    po, do, _ = oracle.scan_knn(x, sub, 3, 4)
    assert np.array_equal(pos[seam], po) and np.array_equal(d[seam].view(np.uint32), do.view(np.uint32))

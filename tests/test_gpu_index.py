"""GPU parity of the iSAX index (TREMOR's index: indexing.cpp:24-281 build,
Tremor.cpp:1195-1518 leaf-order search, SAX.cpp:112-170 node mindist).

Bar: the device build equals the oracle's restatement exactly (the words, the
window ids in key order, every leaf's start, count and depth), and searches
through the index return the oracle scan's answers bit for bit - kNN and
threshold, merge 0 and > 0, missing samples, flat stretches (one leaf at full
depth larger than the leaf size), lengths the index cannot take beside ones it
does - with the reference's 1% rule choosing the index where its leaves
prune; and, on the 16-segment fixture cases written from the reference build
(the plain workload at full size among them), the reference's own answers.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MISSING = np.float32(-100000.0)


def _log(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _wave(n, seed, events=30):
    from tremor_amd import synth
    rng = np.random.default_rng(seed)
    x = synth.bandpassed_noise(n, rng)
    onsets, _ = synth.inject_events(x, events, rng)
    return x, onsets, rng


@pytest.mark.parametrize("m,leaf_size", [(64, 300), (256, 2000)])
def test_index_build_equals_oracle(oracle, m, leaf_size):
    from tremor_amd.capi import Context
    x, _, _ = _wave(1 << 17, 71)
    x[5000:5100] = MISSING
    x[20000:23000] = 0.25  # a flat stretch: thousands of windows with one word, a leaf at full depth
    _log("index build", m=m, leaf_size=leaf_size)
    with Context() as ctx:
        ctx.load_waveform(x)
        ids, leaves, words = ctx.index_probe(m, leaf_size)
    want_words = oracle.sax_words(x, m, 16)
    assert np.array_equal(words, want_words)
    want_ids, want_leaves = oracle.isax_index(want_words, leaf_size)
    assert np.array_equal(ids, want_ids)
    assert np.array_equal(leaves, want_leaves)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16"])
def test_index_answers_equal_scan(oracle, prec):
    """Every eligible template (lengths that are multiples of 16) through the
    index, the others through the planned forms: the oracle's answers."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(1 << 18, 73)
    x[70000:70400] = MISSING
    # A flat stretch: near-constant windows are undecidable for the early
    # abandon (always candidates), and over 2,048 windows of length 64 share
    # one word, so one leaf at full depth spans several steps of the search.
    x[120000:123000] = -0.5
    lengths = [64, 96, 100, 128, 250, 256, 512]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    eligible = sum(1 for m in lengths if m % 16 == 0)
    with Context(precision=prec, index="always", leaf_size=500) as ctx:
        ctx.load_waveform(x)
        for k, merge in ((1, 0), (10, 16)):
            _log("index knn", prec=prec, k=k, merge=merge)
            pos, d, st = ctx.knn(tm, k, merge)
            assert st["index_templates"] == eligible
            po, do, _ = oracle.scan_knn(x, tm, k, merge)
            assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
        for rho, merge in ((0.8, 0), (0.6, 8)):
            _log("index threshold", prec=prec, rho=rho, merge=merge)
            r = ctx.threshold(tm, rho, merge)
            assert r.stats["index_templates"] == eligible
            tt, tp, td = oracle.scan_threshold(x, tm, rho, merge)
            assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
            assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_index_auto_takes_pruning_templates(oracle):
    """k = 1 on templates cut from events: the seed bound is tight, at most 1%
    of the sampled windows' word bounds pass it, and the reference's rule
    sends the templates to the index, whose leaves prune; at a low threshold
    nearly every bound passes and the rule keeps them on the planned forms."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(1 << 20, 79, events=40)
    tm = synth.cut_templates(x, onsets, [128, 256, 512, 1024], rng)
    W = [x.size - len(t) + 1 for t in tm]
    with Context(index="auto", leaf_size=2000) as ctx:
        ctx.load_waveform(x)
        _log("index auto knn")
        pos, d, st = ctx.knn(tm, 1, 0)
        print("  stats", {k: st[k] for k in ("index_templates", "index_windows", "candidates")}, flush=True)
        assert st["index_templates"] >= 1
        assert 0 < st["index_windows"] < sum(W)  # the leaves prune
        po, do, _ = oracle.scan_knn(x, tm, 1, 0)
        assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
        _log("index auto low threshold")
        r = ctx.threshold(tm, 0.2, 0)
        assert r.stats["index_templates"] == 0
        tt, tp, td = oracle.scan_threshold(x, tm, 0.2, 0)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_index_rebuilt_for_another_leaf_size(oracle):
    """Contexts with different leaf sizes over the same waveform build their
    own trees; the answers do not change."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(1 << 17, 83)
    tm = synth.cut_templates(x, onsets, [256, 256, 320], rng)
    po, do, _ = oracle.scan_knn(x, tm, 5, 0)
    for leaf in (16, 2000, 100000):
        with Context(index="always", leaf_size=leaf) as ctx:
            ctx.load_waveform(x)
            _log("index leaf size", leaf=leaf)
            pos, d, st = ctx.knn(tm, 5, 0)
            assert st["index_templates"] == 3
            assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))


def test_more_lengths_than_the_words_cache(oracle):
    """A batch with more distinct lengths than the resident words keep (8):
    the words and indexes of every length the batch uses stay resident while
    the others are built, through the index and through the skip-sequential
    bound alike."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(1 << 16, 89, events=20)
    lengths = [64 + 16 * i for i in range(11)]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    po, do, _ = oracle.scan_knn(x, tm, 3, 0)
    tt, tp, td = oracle.scan_threshold(x, tm, 0.7, 0)
    for kw in ({"index": "always", "leaf_size": 200}, {"scan": "always", "lower_bounds": True}):
        with Context(**kw) as ctx:
            ctx.load_waveform(x)
            for rep in range(2):  # the second call finds everything resident
                _log("many lengths", rep=rep, **kw)
                pos, d, st = ctx.knn(tm, 3, 0)
                assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
                r = ctx.threshold(tm, 0.7, 0)
                assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
                assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))
            if "index" in kw:
                assert st["index_templates"] == len(lengths)


@pytest.mark.parametrize("name", ["validate16", "c0"])
def test_index_equals_reference_fixtures(oracle, name):
    """The index search against the reference's own answers on the
    16-segment fixture cases (tests/golden/sss_lb_vectors.npz, written from
    the reference build: its exact skip-sequential search, whose answers the
    reference's index search returns too) - the plain workload at its full
    size among them - and bit for bit against the oracle."""
    import os
    import sys
    from tremor_amd.capi import Context
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_sss_lb_golden as G
    lbg = np.load(os.path.join(here, "golden", "sss_lb_vectors.npz"), allow_pickle=False)
    _, x, tm, m, seg, k, rhos, merges, _full = {c[0]: c for c in G.cases()}[name]
    assert seg == 16

    def fmt(v):
        return "%f" % float(v)

    with Context(precision="bf16x3", index="always", leaf_size=2000) as ctx:
        ctx.load_waveform(x)
        for merge in merges:
            _log("index-knn-ref", case=name, merge=merge)
            pos, d, st = ctx.knn(tm, k, merge)
            assert st["index_templates"] == len(tm)
            kk = f"{name}/knn{merge}"
            keep = d.ravel() < np.finfo(np.float32).max
            assert np.array_equal(np.repeat(np.arange(len(tm)), k)[keep], lbg[kk + "/tid"])
            assert np.array_equal(pos.ravel()[keep], lbg[kk + "/pos"])
            assert [fmt(oracle.ed_to_cc(v, m)) for v in d.ravel()[keep]] == list(lbg[kk + "/corr"])
            po, do, _ = oracle.scan_knn(x, tm, k, merge)
            assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
            for i, rho in enumerate(rhos):
                _log("index-threshold-ref", case=name, rho=rho, merge=merge)
                r = ctx.threshold(tm, rho, merge)
                assert r.stats["index_templates"] == len(tm)
                key = f"{name}/thr{merge}_{i}"
                assert np.array_equal(r.tmpl, lbg[key + "/q"]) and np.array_equal(r.pos, lbg[key + "/pos"])
                assert [fmt(oracle.ed_to_cc(v, m)) for v in r.d] == list(lbg[key + "/corr"])
                tt, tp, td = oracle.scan_threshold(x, tm, rho, merge)
                assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
                assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_auto_stays_dense_where_the_index_does_not_fit(oracle):
    """At 2^31 - 64 samples the index of one length would need about 200 GB
    while it is built: auto mode keeps the template on the dense forms (the
    call answers instead of failing on device memory)."""
    from tremor_amd.capi import Context
    n = (1 << 31) - 64
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n, dtype=np.float32)
    j = 123457
    t = x[j:j + 256].copy()
    with Context(index="auto") as ctx:
        ctx.load_waveform(x)
        _log("index auto at 2^31")
        pos, d, st = ctx.knn([t], 1, 0)
    assert st["index_templates"] == 0
    assert pos[0][0] == j  # the template's own window, distance ~0
    # the oracle on the neighbourhood: the same distance as the reference's kernel there
    lo = j - 1000
    po, do, _ = oracle.scan_knn(x[lo:j + 1256], [t], 1, 0)
    assert po[0][0] + lo == j and np.array_equal(d[0].view(np.uint32), do[0].view(np.uint32))


def test_index_with_lower_bounds_and_forms_mixed(oracle):
    """The index beside the skip-sequential bound (both on the 16-segment
    words of a length) and beside the FFT form: lengths the index takes go
    through it, the others through the bounded scan or the FFT form; the
    answers are the oracle's."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(1 << 17, 97)
    lengths = [64, 100, 160, 250, 256, 1000]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    po, do, _ = oracle.scan_knn(x, tm, 4, 8)
    tt, tp, td = oracle.scan_threshold(x, tm, 0.75, 0)
    for kw in ({"scan": "always"}, {"fft": "always"}):
        with Context(index="always", lower_bounds=True, paa_segments=16, leaf_size=400, **kw) as ctx:
            ctx.load_waveform(x)
            _log("index mixed", **kw)
            pos, d, st = ctx.knn(tm, 4, 8)
            assert st["index_templates"] == 3  # 64, 160, 256
            assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
            r = ctx.threshold(tm, 0.75, 0)
            assert r.stats["index_templates"] == 3
            assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
            assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))

"""GPU parity of the skip-sequential method: the early-abandoning scan with
the reference's iSAX lower bound (SkipSequentialScan with lower_bounds = true,
SkipSequential.cpp:155-211 build and bound, :256 and :321 the skip tests).

Bar: the reference's own answers with lower bounds on (the fixture
tests/golden/sss_lb_vectors.npz, written from the reference build) - kNN and
threshold positions exactly, correlations as the reference prints them -
and, bit for bit, the oracle's restatement, which equals the sequential
method's answers.  A word, PAA value or bound computed differently skips a
window the reference keeps (an answer lost) or keeps one it skips (caught by
the skip counts matching the oracle's own bound where that decides).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _log(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _cases():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_sss_lb_golden as G
    return {c[0]: c for c in G.cases()}


@pytest.fixture(scope="module")
def lbg():
    return np.load(os.path.join(HERE, "golden", "sss_lb_vectors.npz"), allow_pickle=False)


def _fmt(v):
    return "%f" % float(v)


@pytest.mark.parametrize("name", ["validate16", "validate8", "c0", "adversarial"])
def test_skip_sequential_equals_reference(lbg, oracle, name):
    from tremor_amd.capi import Context
    _, x, tm, m, seg, k, rhos, merges, _full = _cases()[name]
    skipped = 0
    with Context(precision="bf16x3", scan="always", lower_bounds=True, paa_segments=seg) as ctx:
        ctx.load_waveform(x)
        for merge in merges:
            _log("sss-knn", case=name, segments=seg, merge=merge)
            pos, d, st = ctx.knn(tm, k, merge)
            assert st["scan_templates"] == len(tm)
            kk = f"{name}/knn{merge}"
            keep = d.ravel() < np.finfo(np.float32).max
            assert np.array_equal(np.repeat(np.arange(len(tm)), k)[keep], lbg[kk + "/tid"])
            assert np.array_equal(pos.ravel()[keep], lbg[kk + "/pos"])
            assert [_fmt(oracle.ed_to_cc(v, m)) for v in d.ravel()[keep]] == list(lbg[kk + "/corr"])
            po, do, _ = oracle.scan_knn(x, tm, k, merge, segments=seg)
            assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
            for i, rho in enumerate(rhos):
                _log("sss-threshold", case=name, rho=rho, merge=merge)
                r = ctx.threshold(tm, rho, merge)
                key = f"{name}/thr{merge}_{i}"
                assert np.array_equal(r.tmpl, lbg[key + "/q"]) and np.array_equal(r.pos, lbg[key + "/pos"])
                assert [_fmt(oracle.ed_to_cc(v, m)) for v in r.d] == list(lbg[key + "/corr"])
                tt, tp, td = oracle.scan_threshold(x, tm, rho, merge, segments=seg)
                assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
                assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))
                # The bound skips windows exactly where the reference's does:
                # at a threshold the skips are the windows whose bound (the
                # oracle's, pinned to the reference's) exceeds cc_to_ed.
                bound = oracle.cc_to_ed(rho, m)
                words = oracle.sax_words(x, m, seg)
                start_ok = x[: x.size - m + 1] != np.float32(-100000.0)
                want_skip = sum(int(np.sum((oracle.lower_bounds(oracle.paa(oracle.znormalize(t), seg), words, m)
                                            > bound) & start_ok)) for t in tm)
                assert r.stats["lb_skipped"] == want_skip
                skipped += want_skip
    # The bound prunes in every case but the 8-segment one on the validation
    # recipe, whose 64-sample windows leave it too coarse at rho 0.4.
    assert skipped > 0 or name == "validate8"


def test_bounds_leave_answers_unchanged_mixed_lengths(oracle):
    """Lengths that are not a multiple of the segments scan without bounds;
    with the cost model choosing, the answers stay the sequential ones."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    rng = np.random.default_rng(41)
    x = synth.bandpassed_noise(1 << 17, rng)
    onsets, _ = synth.inject_events(x, 20, rng)
    lengths = [64, 96, 100, 128, 250, 256, 512]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    for scan in ("always", "auto"):
        with Context(precision="bf16x3", scan=scan, lower_bounds=True, paa_segments=16) as ctx:
            ctx.load_waveform(x)
            for merge in (0, 8):
                _log("sss-mixed", scan=scan, merge=merge)
                pos, d, _ = ctx.knn(tm, 7, merge)
                po, do, _ = oracle.scan_knn(x, tm, 7, merge)
                assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
                r = ctx.threshold(tm, 0.6, merge)
                tt, tp, td = oracle.scan_threshold(x, tm, 0.6, merge)
                assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
                assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))

"""The oracle's iSAX restatement against the reference's own skip-sequential
scan (SkipSequentialScan with lower_bounds = true, SkipSequential.cpp:138-211
build and bound, :256 and :321 the skip tests), through the fixture
tests/golden/sss_lb_vectors.npz written by tests/golden/make_sss_lb_golden.py
from the reference build.

Pinned bit for bit: the 255 breakpoints of cardinality 256, every window's
iSAX word (16 and 8 segments), every template's PAA, every lower bound (the
16-segment SIMD form and the plain form), and the reference's kNN and
threshold answers with lower bounds on.  The mathematics adds the bound's
defining property: a window's bound never exceeds its distance, so the
skip-sequential answers are the sequential ones."""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "sss_lb_vectors.npz")


@pytest.fixture(scope="module")
def lbg():
    return np.load(GOLD, allow_pickle=False)


def _case(name):
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_sss_lb_golden as G
    for case in G.cases():
        if case[0] == name:
            return case, G.digest
    raise KeyError(name)


def _fmt(v):
    return "%f" % float(v)


def test_breakpoints_are_the_references(oracle, lbg):
    bp = oracle.sax_breakpoints()
    assert np.array_equal(bp.view(np.uint32), lbg["breakpoints"].view(np.uint32))
    # the table is symmetric about 0 and increasing
    assert bp[127] == 0.0 and np.all(np.diff(bp) > 0) and np.array_equal(bp, -bp[::-1])


@pytest.mark.parametrize("name", ["validate16", "validate8", "c0", "adversarial"])
def test_words_paa_bounds_and_answers(oracle, lbg, name):
    (case, digest) = _case(name)
    _, x, tm, m, seg, k, rhos, merges, full = case
    assert digest(x, *tm) == str(lbg[f"{name}/input_sha"]), "workload generator changed"
    words = oracle.sax_words(x, m, seg)
    assert hashlib.sha256(words.tobytes()).hexdigest() == str(lbg[f"{name}/sax_sha"])
    idx = None if full else lbg[f"{name}/idx"]
    assert np.array_equal(words if full else words[idx], lbg[f"{name}/sax"])
    paas = np.stack([oracle.paa(oracle.znormalize(t), seg) for t in tm])
    assert np.array_equal(paas.view(np.uint32), lbg[f"{name}/paa"].view(np.uint32))
    lbs = np.stack([oracle.lower_bounds(p, words, m) for p in paas])
    assert hashlib.sha256(lbs.tobytes()).hexdigest() == str(lbg[f"{name}/lb_sha"])
    want = lbg[f"{name}/lb"]
    assert np.array_equal((lbs if full else lbs[:, idx]).view(np.uint32), want.view(np.uint32))
    for merge in [int(v) for v in lbg[f"{name}/merges"]]:
        pos, d, _ = oracle.scan_knn(x, tm, k, merge, segments=seg)
        rows_t = np.repeat(np.arange(len(tm)), k)
        keep = d.ravel() < np.finfo(np.float32).max
        got_pos = pos.ravel()[keep]
        got_corr = [_fmt(oracle.ed_to_cc(v, m)) for v in d.ravel()[keep]]
        kk = f"{name}/knn{merge}"
        assert np.array_equal(rows_t[keep], lbg[kk + "/tid"])
        assert np.array_equal(got_pos, lbg[kk + "/pos"])
        assert got_corr == list(lbg[kk + "/corr"])
        for i, rho in enumerate(float(r) for r in lbg[f"{name}/rhos"]):
            ot, op, od = oracle.scan_threshold(x, tm, rho, merge, segments=seg)
            key = f"{name}/thr{merge}_{i}"
            assert np.array_equal(ot, lbg[key + "/q"])
            assert np.array_equal(op, lbg[key + "/pos"])
            assert [_fmt(oracle.ed_to_cc(v, m)) for v in od] == list(lbg[key + "/corr"])
            # same answers as the sequential method
            st, sp, sd = oracle.scan_threshold(x, tm, rho, merge)
            assert np.array_equal(st, ot) and np.array_equal(sp, op) and np.array_equal(sd, od)


@pytest.mark.parametrize("seg", [16, 8, 4])
def test_bound_never_exceeds_distance(oracle, seg):
    """LB_PAA <= ED^2 of the z-normalised sequences (the PAA bound), and the
    symbol region holds the window's PAA value: every window's bound is at
    most its exact FP64 distance (with FP32 rounding room)."""
    from tremor_amd import synth
    w = synth.validate_recipe()
    x, m = w.x, 64
    words = oracle.sax_words(x, m, seg)
    mean, sd = oracle.window_stats(x, m)
    W = x.size - m + 1
    win = np.lib.stride_tricks.sliding_window_view(x.astype(np.float64), m)[:W]
    mu = win.mean(axis=1, keepdims=True)
    sig = win.std(axis=1, keepdims=True)
    zw = (win - mu) / np.where(sig > 1e-12, sig, 1.0)
    for t in w.templates:
        qz = oracle.znormalize(t)
        lb = oracle.lower_bounds(oracle.paa(qz, seg), words, m).astype(np.float64)
        d = np.sum((qz.astype(np.float64)[None, :] - zw) ** 2, axis=1)
        ok = sd > 0
        assert np.all(lb[ok] <= d[ok] * (1 + 1e-5) + 1e-4)
        assert np.mean(lb[ok] > 0) > 0.5  # and it is not vacuous


def test_plain_form_differs_from_simd_only_in_order(oracle):
    """With the same squares, the 16-segment SIMD order and the sequential FP64
    form agree to FP32 rounding (the fixture pins which one each count uses)."""
    rng = np.random.default_rng(3)
    bp = oracle.sax_breakpoints()
    p = rng.standard_normal(16).astype(np.float32)
    words = rng.integers(0, 256, size=(1000, 16)).astype(np.uint8)
    lb16 = oracle.lower_bounds(p, words, 256)
    lo = np.where(words == 0, np.float32(-2e6), bp[np.maximum(words.astype(int) - 1, 0)])
    hi = np.where(words == 255, np.float32(2e6), bp[np.minimum(words.astype(int), 254)])
    dd = np.where(lo > p, lo - p, np.where(hi < p, hi - p, 0)).astype(np.float64)
    ref = 16.0 * np.sum(dd * dd, axis=1)
    assert np.allclose(lb16, ref, rtol=1e-5, atol=1e-6)

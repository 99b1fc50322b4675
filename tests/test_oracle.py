"""Pins the CPU oracle.

1. Against the committed fixture written by the reference's own functions
   (tests/golden/ref_vectors.npz): bit-exact z-norm, block order, distances
   (with and without early abandoning), blocks visited, canonical kNN
   selection with exclusion and the bound set.
2. Against the reference build itself (oracle/_ref), where it exists, on
   fresh random and adversarial cases.
3. Against the mathematics: closed forms of the z-normalised distance
   (self-match, affine invariance, anti-correlation, constant windows,
   missing samples), FP64 brute force, and the selection rules restated in
   the reference's validation script (validate_search.py:20-30, 129-149).
"""
import ctypes
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def test_golden_distance_cases(oracle, gold):
    meta = gold["dist_meta"]
    q_raw, w_raw, zb = gold["q_raw"], gold["w_raw"], gold["zbits"]
    order, order_len = gold["order"], gold["order_len"]
    qo = wo = oo = 0
    for i, (m, start, mean_b, sd_b, bound_b) in enumerate(meta):
        m = int(m)
        t = q_raw[qo:qo + m]
        w = w_raw[wo:wo + m + 4]
        z_ref = zb[qo:qo + m]
        o_ref = order[oo:oo + order_len[i]]
        qo += m
        wo += m + 4
        oo += int(order_len[i])
        z = oracle.znormalize(t)
        assert np.array_equal(z.view(np.uint32), z_ref), f"z-norm case {i} m={m}"
        o = oracle.block_order(z)
        assert np.array_equal(o, o_ref), f"block order case {i}"
        mean, sd = oracle.window_stats(w[start:start + m], m)
        assert mean[0].view(np.uint32) == np.uint32(mean_b) and sd[0].view(np.uint32) == np.uint32(sd_b)
        bound = np.uint32(bound_b).view(np.float32)
        d, blocks = oracle.dist(z, w, int(start), mean[0], sd[0], float(bound), o)
        assert d.view(np.uint32) == gold["dbits"][i], f"distance case {i} m={m}"
        assert blocks == gold["blocks"][i]


def test_golden_selection(oracle, gold):
    meta = gold["sel_meta"]
    ip = op = 0
    for cnt, k, merge, kept in meta:
        pos = gold["sel_in_pos"][ip:ip + cnt]
        d = gold["sel_in_d"][ip:ip + cnt]
        ip += int(cnt)
        p, dd = oracle.select_knn(pos, d, int(k), int(merge))
        assert len(p) == kept
        assert np.array_equal(p, gold["sel_out_pos"][op:op + kept])
        assert np.array_equal(dd, gold["sel_out_d"][op:op + kept])
        op += int(kept)


def test_golden_bound_set(oracle, gold):
    io = oo = 0
    for k, merge, steps in gold["bsf_meta"]:
        ds = gold["bsf_in_d"][io:io + steps]
        ps = gold["bsf_in_p"][io:io + steps]
        io += int(steps)
        knn, pos = oracle.bsf_run(int(k), int(merge), ds, ps)
        assert np.array_equal(knn, gold["bsf_out_d"][oo:oo + k])
        assert np.array_equal(pos, gold["bsf_out_p"][oo:oo + k])
        oo += int(k)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "libtremor_ref.so")),
                    reason="reference build (oracle/_ref) not present on this machine")
def test_against_reference_build(oracle):
    R = oracle.ref()
    rng = np.random.default_rng(11)
    for trial in range(600):
        m = int(rng.integers(1, 1500))
        t = (rng.standard_normal(m) * 10.0 ** rng.integers(-4, 5) + rng.standard_normal() * 5).astype(np.float32)
        if trial % 9 == 0:
            t[rng.integers(0, m)] = oracle.MISSING
        z = oracle.znormalize(t)
        z2 = t.copy()
        R.ref_znormalize(oracle._p(z2, oracle._f32p), m)
        assert np.array_equal(z.view(np.uint32), z2.view(np.uint32))
        o = oracle.block_order(z)
        raw = (rng.standard_normal(m + 20) * rng.choice([1e-3, 1.0, 1e3, 1e6])
               + rng.choice([0.0, 1.0, 1e5])).astype(np.float32)
        if trial % 5 == 0:
            raw[rng.integers(0, m + 20, size=3)] = oracle.MISSING
        s = int(rng.integers(0, 20))
        mean, sd = oracle.window_stats(raw[s:s + m], m)
        for bound in (float(oracle.FLT_MAX), float(rng.uniform(0, 4 * m))):
            d1, b1 = oracle.dist(z, raw, s, mean[0], sd[0], bound, o)
            bl = ctypes.c_int32(0)
            d2 = R.ref_dist(oracle._p(z, oracle._f32p), oracle._p(raw, oracle._f32p), s, m, float(mean[0]),
                            float(sd[0]), bound, oracle._p(o, oracle._i32p), ctypes.byref(bl))
            assert np.float32(d1).view(np.uint32) == np.float32(d2).view(np.uint32)
            assert b1 == bl.value


# ------------------------------------------------------------ closed forms

def _stats1(oracle, w):
    mean, sd = oracle.window_stats(w, w.size)
    return mean[0], sd[0]


def test_self_match_and_affine_invariance(oracle):
    rng = np.random.default_rng(3)
    for m in (17, 64, 250, 1000):
        t = rng.standard_normal(m).astype(np.float32)
        z = oracle.znormalize(t)
        for a, b in ((1.0, 0.0), (3.5, -2.0), (1e3, 1e4)):
            w = (t.astype(np.float64) * a + b).astype(np.float32)
            mean, sd = _stats1(oracle, w)
            d, _ = oracle.dist(z, w, 0, mean, sd)
            # rho(q, a q + b) = 1 -> d = 0 up to FP32 rounding of a z-normalised pair
            assert d < 1e-3 * m / 64, (m, a, b, d)


def test_anticorrelation_gives_4m(oracle):
    rng = np.random.default_rng(4)
    for m in (32, 100, 777):
        t = rng.standard_normal(m).astype(np.float32)
        z = oracle.znormalize(t)
        w = (-t).astype(np.float32)
        mean, sd = _stats1(oracle, w)
        d, _ = oracle.dist(z, w, 0, mean, sd)
        assert abs(float(d) - 4 * m) < 1e-3 * m


def test_constant_window_and_missing(oracle):
    m = 64
    z = oracle.znormalize(np.arange(m, dtype=np.float32))
    w = np.full(m, 3.0, dtype=np.float32)
    mean, sd = _stats1(oracle, w)
    assert sd == 0.0
    d, _ = oracle.dist(z, w, 0, mean, sd)
    assert d == oracle.FLT_MAX
    # interior missing samples count as 0 in the statistics and the distance
    rng = np.random.default_rng(5)
    w = rng.standard_normal(m).astype(np.float32)
    w[10] = oracle.MISSING
    w0 = w.copy()
    w0[10] = 0.0
    m1, s1 = _stats1(oracle, w)
    m0, s0 = _stats1(oracle, w0)
    assert m1 == m0 and s1 == s0
    assert oracle.dist(z, w, 0, m1, s1)[0] == oracle.dist(z, w0, 0, m0, s0)[0]
    # a window whose first sample is missing is skipped: statistics stay 0
    w[0] = oracle.MISSING
    mean, sd = _stats1(oracle, w)
    assert mean == 0.0 and sd == 0.0


def test_distance_matches_fp64_pearson(oracle):
    rng = np.random.default_rng(6)
    x = rng.standard_normal(5000).astype(np.float32)
    for m in (31, 64, 200):
        t = x[1234:1234 + m] + 0.3 * rng.standard_normal(m).astype(np.float32)
        z = oracle.znormalize(t)
        d = oracle.dist_all(z, x)
        rho = oracle.pearson_fp64(x, t)
        assert np.allclose(1.0 - d / (2.0 * m), rho, atol=2e-5)


def test_cc_ed_roundtrip(oracle):
    for m in (8, 64, 1000):
        for cc in (-1.0, 0.0, 0.4, 0.93, 1.0):
            ed = oracle.cc_to_ed(cc, m)
            assert ed == np.float32(2.0) * np.float32(m) * (np.float32(1.0) - np.float32(cc))
            assert abs(float(oracle.ed_to_cc(ed, m)) - cc) < 1e-6


def _knn_with_exclusion(corr, k, merge):
    """validate_search.py:20-30"""
    order = sorted(range(len(corr)), key=lambda p: (-corr[p], p))
    kept = []
    for p in order:
        if len(kept) == k or corr[p] == -np.inf:
            break
        if all(abs(p - q) >= merge for q in kept):
            kept.append(p)
    return kept


def test_select_knn_matches_brute_force(oracle):
    rng = np.random.default_rng(7)
    for trial in range(60):
        cnt = int(rng.integers(1, 400))
        d = rng.integers(0, 60, size=cnt).astype(np.float32)
        pos = np.arange(cnt, dtype=np.uint64)
        k = int(rng.integers(1, 15))
        merge = int(rng.choice([0, 1, 4, 30]))
        p, dd = oracle.select_knn(pos, d, k, merge)
        exp = _knn_with_exclusion(list(-d.astype(np.float64)), k, merge)
        assert list(p) == exp


def test_validate_recipe_scan_matches_fp64(oracle):
    """The reference validation's checks on its adversarial input."""
    from tremor_amd import synth
    w = synth.validate_recipe()
    x, qs, m = w.x, w.templates, 64
    corr = np.stack([oracle.pearson_fp64(x, q) for q in qs])
    n = x.size
    for merge in (0, m // 4):
        pos, d, _ = oracle.scan_knn(x, qs, 5, merge, threads=4)
        for qi in range(len(qs)):
            found = [int(p) for p, dd in zip(pos[qi], d[qi]) if dd < oracle.FLT_MAX]
            cc = 1.0 - d[qi][:len(found)] / (2.0 * m)
            assert np.allclose(cc, corr[qi, found], atol=2e-4)
            if merge == 0:
                assert found == np.argsort(-corr[qi], kind="stable")[:5].tolist()
            else:
                assert found == _knn_with_exclusion(list(corr[qi]), 5, merge)
        tt, tp, td = oracle.scan_threshold(x, qs, 0.4, merge, threads=4)
        for qi in range(len(qs)):
            got = set(int(p) for t, p in zip(tt, tp) if t == qi)
            exp = set()
            nxt = 0
            for p in np.flatnonzero(corr[qi, : n - m + 1] >= 0.4):
                if p >= nxt:
                    exp.add(int(p))
                    nxt = p + merge + 1
            assert got == exp
        # constant stretch never produces hits or neighbours
        assert not any(2000 <= int(p) <= 2128 - m for p in tp)


def test_scan_knn_tail_and_fill(oracle):
    """Fewer windows than k: unfilled slots are (0, FLT_MAX)."""
    rng = np.random.default_rng(8)
    x = rng.standard_normal(40).astype(np.float32)
    pos, d, _ = oracle.scan_knn(x, [x[:32]], 20, 0, threads=2)
    assert (d[0][:9] < oracle.FLT_MAX).all() and (d[0][9:] == oracle.FLT_MAX).all()
    assert (pos[0][9:] == 0).all()

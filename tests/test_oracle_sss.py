"""The oracle's window statistics, kNN scan and threshold scan against the
reference's own sequential scan (SkipSequentialScan, lower_bounds = false,
SkipSequential.cpp:138-393), through the fixture tests/golden/sss_vectors.npz
written by tests/golden/make_sss_golden.py from the reference build.

Statistics: every window's float mean and std bits (SHA-256 of all of them,
plus sampled windows read back bit for bit).  kNN: the reference's CSV rows
(tID, k, pos, corr) - positions exactly, correlations as printed with %f from
ed_to_cc of the FP32 distance.  Threshold: the CSV rows (Query, Counter,
Position, Cross Correlation) after the reference's sort and merge."""
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "sss_vectors.npz")


@pytest.fixture(scope="module")
def sss():
    return np.load(GOLD, allow_pickle=False)


def _inputs(name):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_sss_golden as G
    for case in G.cases():
        if case[0] == name:
            return case, G.digest
    raise KeyError(name)


def _fmt(v):
    return "%f" % float(v)


@pytest.mark.parametrize("name", ["validate", "c0", "adversarial"])
def test_oracle_matches_reference_scan(oracle, sss, name):
    (case, digest) = _inputs(name)
    _, x, tm, m, k, rhos, merges, full = case
    assert digest(x, *tm) == str(sss[f"{name}/input_sha"]), "workload generator changed"
    assert int(sss[f"{name}/m"]) == m and int(sss[f"{name}/k"]) == k
    # statistics
    mean, sd = oracle.window_stats(x, m)
    mb, sb = mean.view(np.uint32), sd.view(np.uint32)
    if full:
        assert np.array_equal(mb, sss[f"{name}/mean_bits"]) and np.array_equal(sb, sss[f"{name}/std_bits"])
    else:
        idx = sss[f"{name}/stat_idx"]
        assert np.array_equal(mb[idx], sss[f"{name}/mean_bits"]) and np.array_equal(sb[idx], sss[f"{name}/std_bits"])
    h = hashlib.sha256()
    h.update(mb.tobytes())
    h.update(sb.tobytes())
    assert h.hexdigest() == str(sss[f"{name}/stats_sha"])
    for merge in [int(v) for v in sss[f"{name}/merges"]]:
        # kNN rows
        pos, d, _ = oracle.scan_knn(x, tm, k, merge)
        key = f"{name}/knn{merge}"
        T = len(tm)
        assert np.array_equal(sss[key + "/tid"], np.repeat(np.arange(T), k))
        assert np.array_equal(sss[key + "/rank"], np.tile(np.arange(1, k + 1), T))
        assert np.array_equal(sss[key + "/pos"], pos.reshape(-1).astype(np.uint64)), (name, merge)
        corr = [_fmt(oracle.ed_to_cc(v, m)) for v in d.reshape(-1)]
        assert corr == list(sss[key + "/corr"]), (name, merge)
        # threshold rows
        for i, rho in enumerate(float(r) for r in sss[f"{name}/rhos"]):
            key = f"{name}/thr{merge}_{i}"
            tt, tp, td = oracle.scan_threshold(x, tm, rho, merge)
            assert np.array_equal(sss[key + "/q"], tt.astype(np.int64)), (name, merge, rho)
            assert np.array_equal(sss[key + "/pos"], tp.astype(np.uint64)), (name, merge, rho)
            counter = np.concatenate([np.arange(np.count_nonzero(tt == q)) for q in range(T)]) if tt.size else tt
            assert np.array_equal(sss[key + "/counter"], counter.astype(np.int64))
            assert [_fmt(oracle.ed_to_cc(v, m)) for v in td] == list(sss[key + "/corr"]), (name, merge, rho)


def test_reference_scan_build_is_current():
    """Where the reference build exists, regenerating its answers for the
    validation recipe gives the committed rows again (the fixture is what
    the reference computes, not a stale copy)."""

    import sys
    import tempfile
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_sss_golden as G
    if not os.path.exists(G.BIN):
        pytest.skip("reference scan not built here (make -C oracle ref-sss needs the reference sources)")
    sss = np.load(GOLD, allow_pickle=False)
    (case, _) = _inputs("validate")
    _, x, tm, m, k, rhos, merges, _ = case
    with tempfile.TemporaryDirectory() as tmp:
        mean, sd, knn, thr = G.run(tmp, x, tm, m, k, rhos, 16)
    assert np.array_equal(mean, sss["validate/mean_bits"]) and np.array_equal(sd, sss["validate/std_bits"])
    assert [int(r[2]) for r in knn] == [int(v) for v in sss["validate/knn16/pos"]]
    assert [r[3] for r in thr[0]] == list(sss["validate/thr16_0/corr"])


"""GPU parity of the early-abandoning scan form of the candidate pass (the
reference's sequential scan with its abandon test after every block,
SkipSequential.cpp:223-393 over ZnormSearch.hpp:103-165) and of the cost
model's choice of it (Tremor.cpp:2265-2286).

Bar: as for every other form - kNN positions and distance bits, threshold
rows and distance bits identical to the oracle's restatement of the
reference sequential scan.  The scan abandons a window only when its running
sum proves that the reference's distance exceeds the bound, so any mistake
in that proof (a missing error term, a wrong block order, a tail dropped)
loses answers and fails these comparisons.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scan_ctx():
    from tremor_amd.capi import Context
    c = Context(precision="bf16x3", scan="always")
    yield c
    c.close()


def _log(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _same_knn(got, want):
    pos, d = got
    po, do = want
    assert np.array_equal(pos, po)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))


def _same_rows(r, want):
    tt, tp, td = want
    assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
    assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_scan_validate_recipe(scan_ctx, oracle):
    from tremor_amd import synth
    w = synth.validate_recipe()
    scan_ctx.load_waveform(w.x)
    for merge in (0, 16):
        _log("scan-validate", merge=merge)
        pos, d, st = scan_ctx.knn(w.templates, 5, merge)
        assert st["scan_templates"] == len(w.templates)
        po, do, _ = oracle.scan_knn(w.x, w.templates, 5, merge)
        _same_knn((pos, d), (po, do))
        r = scan_ctx.threshold(w.templates, 0.4, merge)
        _same_rows(r, oracle.scan_threshold(w.x, w.templates, 0.4, merge))


def test_scan_uneven_lengths_missing(scan_ctx, oracle):
    from tremor_amd import synth
    rng = np.random.default_rng(21)
    x = synth.bandpassed_noise(120_000, rng)
    onsets, _ = synth.inject_events(x, 12, rng)
    lengths = [1, 7, 31, 32, 33, 100, 511, 2047, 4100]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    x[rng.integers(0, x.size, size=30)] = synth.MISSING
    x[int(onsets[1])] = synth.MISSING
    scan_ctx.load_waveform(x)
    for k, merge in ((7, 0), (3, 50), (1, 0)):
        _log("scan-knn", k=k, merge=merge)
        pos, d, st = scan_ctx.knn(tm, k, merge)
        print("  stats", st, flush=True)
        assert st["scan_templates"] == len(tm)
        po, do, _ = oracle.scan_knn(x, tm, k, merge)
        _same_knn((pos, d), (po, do))
    for rho, merge in ((0.7, 10), (0.95, 0), (-0.2, 0)):
        _log("scan-thr", rho=rho, merge=merge)
        r = scan_ctx.threshold(tm, rho, merge)
        print("  stats", r.stats, "rows", r.pos.size, flush=True)
        _same_rows(r, oracle.scan_threshold(x, tm, rho, merge))


def test_scan_edges(scan_ctx, oracle):
    rng = np.random.default_rng(9)
    x = rng.standard_normal(300).astype(np.float32)
    x[50:120] = 1.0  # constant stretch: windows excluded (FLT_MAX)
    scan_ctx.load_waveform(x)
    tm = [x[0:64].copy(), x[100:300].copy(), rng.standard_normal(300).astype(np.float32),
          rng.standard_normal(301).astype(np.float32)]
    _log("scan-edges")
    pos, d, _ = scan_ctx.knn(tm, 150, 0)
    _same_knn((pos, d), oracle.scan_knn(x, tm, 150, 0)[:2])
    pos, d, _ = scan_ctx.knn(tm, 40, 9)
    _same_knn((pos, d), oracle.scan_knn(x, tm, 40, 9)[:2])
    r = scan_ctx.threshold(tm, -0.5, 0)
    _same_rows(r, oracle.scan_threshold(x, tm, -0.5, 0))


def test_scan_offset_waveform(scan_ctx, oracle):
    # A large mean offset: the running sum's error terms carry |mu| / sigma.
    from tremor_amd import synth
    rng = np.random.default_rng(23)
    x = synth.bandpassed_noise(60_000, rng) + np.float32(40.0)
    tm = [x[5000:5600].copy(), x[20000:21100].copy() + rng.standard_normal(1100).astype(np.float32) * 0.05]
    scan_ctx.load_waveform(x)
    _log("scan-offset")
    pos, d, _ = scan_ctx.knn(tm, 4, 0)
    _same_knn((pos, d), oracle.scan_knn(x, tm, 4, 0)[:2])
    r = scan_ctx.threshold(tm, 0.6, 0)
    _same_rows(r, oracle.scan_threshold(x, tm, 0.6, 0))


def test_cost_model_takes_scan_at_tight_bounds(oracle):
    # Long templates at a tight threshold: the sampled abandon test visits a
    # block or two per window, cheaper than the matrix-core form; at a loose
    # threshold the scan is not taken.  Answers identical either way.
    from tremor_amd import synth
    from tremor_amd.capi import Context
    rng = np.random.default_rng(25)
    x = synth.bandpassed_noise(200_000, rng)
    onsets, _ = synth.inject_events(x, 10, rng)
    tm = [x[int(o):int(o) + 4096].copy() for o in onsets[:3]]
    ctx = Context(precision="bf16x3", fft="off", scan="auto")
    try:
        ctx.load_waveform(x)
        _log("scan-auto", rho=0.995)
        r = ctx.threshold(tm, 0.995, 0)
        print("  stats", r.stats, "rows", r.pos.size, flush=True)
        assert r.stats["scan_templates"] == len(tm)
        _same_rows(r, oracle.scan_threshold(x, tm, 0.995, 0))
        _log("scan-auto", rho=0.3)
        r = ctx.threshold(tm, 0.3, 0)
        print("  stats", r.stats, "rows", r.pos.size, flush=True)
        assert r.stats["scan_templates"] == 0
        _same_rows(r, oracle.scan_threshold(x, tm, 0.3, 0))
        _log("scan-auto-knn1")
        pos, d, st = ctx.knn(tm, 1, 0)
        print("  stats", st, flush=True)
        _same_knn((pos, d), oracle.scan_knn(x, tm, 1, 0)[:2])
    finally:
        ctx.close()

"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bar: kNN positions and distances bit-identical to the oracle's restatement
of the reference sequential scan (SkipSequential.cpp:296-393); threshold
rows (template, counter, position, distance bits) identical to
SkipSequential.cpp:223-294; exact statistics, template preparation and
distances bit-identical.  Every precision form of the candidate pass is
held to the same bar, because the verification pass recomputes every
candidate in the reference's arithmetic.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = ["fp32", "bf16x3", "fp16"]


@pytest.fixture(scope="module")
def ctxs():
    from tremor_amd.capi import Context
    out = {p: Context(precision=p) for p in PRECS}
    yield out
    for c in out.values():
        c.close()


def _wave(n, seed, events=20):
    from tremor_amd import synth
    rng = np.random.default_rng(seed)
    x = synth.bandpassed_noise(n, rng)
    onsets, _ = synth.inject_events(x, events, rng)
    return x, onsets, rng


def _log(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def test_prepare_template_bit_exact(ctxs, oracle):
    ctx = ctxs["fp32"]
    rng = np.random.default_rng(1)
    ctx.load_waveform(rng.standard_normal(4096).astype(np.float32))
    for m in (1, 7, 31, 32, 33, 100, 1000, 4097, 65536):
        _log("prepare", m=m)
        t = (rng.standard_normal(m) * 3 + 7).astype(np.float32)
        z, order = ctx.prepare_template(t)
        zo = oracle.znormalize(t)
        assert np.array_equal(z.view(np.uint32), zo.view(np.uint32)), m
        assert np.array_equal(order, oracle.block_order(zo)), m


def test_window_stats_and_distances_bit_exact(ctxs, oracle):
    from tremor_amd import synth
    ctx = ctxs["fp32"]
    w = synth.validate_recipe()
    x = w.x.copy()
    x[100] = synth.MISSING
    x[3000] = synth.MISSING
    ctx.load_waveform(x)
    for m in (64, 33, 500):
        _log("stats", m=m)
        mean, sd = ctx.window_stats(m, 0, x.size - m + 1)
        mo, so = oracle.window_stats(x, m)
        assert np.array_equal(mean.view(np.uint32), mo.view(np.uint32))
        assert np.array_equal(sd.view(np.uint32), so.view(np.uint32))
        t = w.templates[1][:m] if m <= 64 else x[1000:1000 + m]
        starts = np.arange(0, x.size - m + 1, 7, dtype=np.uint64)
        d = ctx.distance_batch(t, starts)
        z = oracle.znormalize(t)
        full = oracle.dist_all(z, x, mo, so)
        assert np.array_equal(d.view(np.uint32), full[starts.astype(np.int64)].view(np.uint32))


@pytest.mark.parametrize("prec", PRECS)
def test_validate_recipe(ctxs, oracle, prec):
    from tremor_amd import synth
    ctx = ctxs[prec]
    w = synth.validate_recipe()
    ctx.load_waveform(w.x)
    for merge in (0, 16):
        _log("validate-knn", prec=prec, merge=merge)
        pos, d, st = ctx.knn(w.templates, 5, merge)
        po, do, _ = oracle.scan_knn(w.x, w.templates, 5, merge)
        assert np.array_equal(pos, po)
        assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
        _log("validate-thr", prec=prec, merge=merge)
        r = ctx.threshold(w.templates, 0.4, merge)
        tt, tp, td = oracle.scan_threshold(w.x, w.templates, 0.4, merge)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


@pytest.mark.parametrize("prec", PRECS)
def test_knn_parity_c0_shape(ctxs, oracle, prec):
    ctx = ctxs[prec]
    x, onsets, rng = _wave(200_000, 5, events=12)
    from tremor_amd import synth
    tm = synth.cut_templates(x, onsets, [256] * 6, rng)
    ctx.load_waveform(x)
    for k, merge in ((10, 0), (10, 64), (1, 0)):
        _log("knn", prec=prec, k=k, merge=merge)
        pos, d, st = ctx.knn(tm, k, merge)
        print("  stats", st, flush=True)
        po, do, _ = oracle.scan_knn(x, tm, k, merge)
        assert np.array_equal(pos, po)
        assert np.array_equal(d.view(np.uint32), do.view(np.uint32))


@pytest.mark.parametrize("prec", PRECS)
def test_threshold_parity_c0_shape(ctxs, oracle, prec):
    ctx = ctxs[prec]
    x, onsets, rng = _wave(200_000, 6, events=15)
    from tremor_amd import synth
    tm = synth.cut_templates(x, onsets, [256] * 5, rng)
    ctx.load_waveform(x)
    for rho, merge in ((0.8, 0), (0.5, 0), (0.6, 100)):
        _log("threshold", prec=prec, rho=rho, merge=merge)
        r = ctx.threshold(tm, rho, merge)
        print("  stats", r.stats, "rows", r.pos.size, flush=True)
        tt, tp, td = oracle.scan_threshold(x, tm, rho, merge)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))
        # per-template running counter as in the CSV
        for t in np.unique(tt):
            assert np.array_equal(r.counter[r.tmpl == t], np.arange((tt == t).sum()))


@pytest.mark.parametrize("prec", PRECS)
def test_uneven_lengths_and_missing(ctxs, oracle, prec):
    ctx = ctxs[prec]
    x, onsets, rng = _wave(150_000, 7, events=10)
    from tremor_amd import synth
    lengths = [33, 200, 511, 512, 1000, 1500, 2047, 2100]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    x[rng.integers(0, x.size, size=40)] = synth.MISSING
    x[int(onsets[0])] = synth.MISSING
    ctx.load_waveform(x)
    _log("uneven-knn", prec=prec)
    pos, d, st = ctx.knn(tm, 7, 0)
    print("  stats", st, flush=True)
    po, do, _ = oracle.scan_knn(x, tm, 7, 0)
    assert np.array_equal(pos, po)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
    _log("uneven-thr", prec=prec)
    r = ctx.threshold(tm, 0.7, 10)
    tt, tp, td = oracle.scan_threshold(x, tm, 0.7, 10)
    assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
    assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


@pytest.mark.parametrize("prec", PRECS)
def test_edges(ctxs, oracle, prec):
    ctx = ctxs[prec]
    rng = np.random.default_rng(9)
    x = rng.standard_normal(300).astype(np.float32)
    x[50:120] = 1.0  # constant stretch: windows excluded (FLT_MAX)
    ctx.load_waveform(x)
    tm = [x[0:64].copy(), x[100:300].copy(), rng.standard_normal(300).astype(np.float32),
          rng.standard_normal(301).astype(np.float32)]
    _log("edges-knn", prec=prec)
    pos, d, _ = ctx.knn(tm, 150, 0)
    po, do, _ = oracle.scan_knn(x, tm, 150, 0)
    assert np.array_equal(pos, po)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
    pos, d, _ = ctx.knn(tm, 40, 9)
    po, do, _ = oracle.scan_knn(x, tm, 40, 9)
    assert np.array_equal(pos, po)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
    _log("edges-thr", prec=prec)
    r = ctx.threshold(tm, -0.5, 0)
    tt, tp, td = oracle.scan_threshold(x, tm, -0.5, 0)
    assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
    assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_rho_error_reported_and_small(ctxs):
    x, onsets, rng = _wave(100_000, 10)
    from tremor_amd import synth
    tm = synth.cut_templates(x, onsets, [300, 700], rng)
    for prec, lim in (("fp32", 1e-4), ("bf16x3", 1e-3), ("fp16", 2e-2)):
        ctx = ctxs[prec]
        ctx.load_waveform(x)
        _log("rho-err", prec=prec)
        _, _, st = ctx.knn(tm, 5, 0)
        print("  stats", st, flush=True)
        assert 0.0 <= st["max_abs_rho_err"] < lim
        assert st["candidates"] >= 5


@pytest.fixture(scope="module")
def fft_ctx():
    from tremor_amd.capi import Context
    c = Context(precision="bf16x3", fft="always")
    yield c
    c.close()


def test_fft_form_validate_recipe(fft_ctx, oracle):
    from tremor_amd import synth
    w = synth.validate_recipe()
    fft_ctx.load_waveform(w.x)
    for merge in (0, 16):
        _log("fft-validate", merge=merge)
        pos, d, st = fft_ctx.knn(w.templates, 5, merge)
        assert st["fft_templates"] == len(w.templates)
        po, do, _ = oracle.scan_knn(w.x, w.templates, 5, merge)
        assert np.array_equal(pos, po)
        assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
        r = fft_ctx.threshold(w.templates, 0.4, merge)
        tt, tp, td = oracle.scan_threshold(w.x, w.templates, 0.4, merge)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_fft_form_uneven_long_templates(fft_ctx, oracle):
    from tremor_amd import synth
    x, onsets, rng = _wave(400_000, 11, events=30)
    lengths = [100, 700, 3000, 9000, 20000]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    x[rng.integers(0, x.size, size=20)] = synth.MISSING
    fft_ctx.load_waveform(x)
    _log("fft-uneven-knn")
    pos, d, st = fft_ctx.knn(tm, 10, 0)
    print("  stats", st, flush=True)
    po, do, _ = oracle.scan_knn(x, tm, 10, 0)
    assert np.array_equal(pos, po)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
    assert st["fft_block"] >= 4 * 20000
    assert st["max_abs_rho_err"] < 1e-2
    _log("fft-uneven-thr")
    r = fft_ctx.threshold(tm, 0.5, 50)
    tt, tp, td = oracle.scan_threshold(x, tm, 0.5, 50)
    assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
    assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


@pytest.mark.parametrize("merge", [0, 16])
def test_fused_fft_form(fft_ctx, oracle, merge):
    """Templates up to 8192 samples take the fused LDS transform (B = 2^15):
    many pieces, the waveform's last piece, missing samples, short and
    maximal lengths."""
    from tremor_amd import synth
    x, onsets, rng = _wave(300_000, 21, events=30)
    lengths = [16, 33, 64, 100, 700, 3000, 8192]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    x[rng.integers(0, x.size, size=20)] = synth.MISSING
    fft_ctx.load_waveform(x)
    _log("fused-knn", merge=merge)
    pos, d, st = fft_ctx.knn(tm, 10, merge)
    print("  stats", st, flush=True)
    assert st["fft_templates"] == len(tm) and st["fft_block"] == 32768
    po, do, _ = oracle.scan_knn(x, tm, 10, merge)
    assert np.array_equal(pos, po)
    assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
    assert st["max_abs_rho_err"] < 1e-3
    for rho in (0.5, 0.8):
        _log("fused-thr", merge=merge, rho=rho)
        r = fft_ctx.threshold(tm, rho, 3 * merge)
        tt, tp, td = oracle.scan_threshold(x, tm, rho, 3 * merge)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_cost_model_routes_long_templates(oracle):
    """Sixteen short templates share full matrix-core groups (cheaper there
    than on the FFT form); the long one takes the FFT form in its own
    shape."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(300_000, 12, events=20)
    tm = synth.cut_templates(x, onsets, [int(v) for v in np.linspace(256, 400, 16)] + [30000], rng)
    with Context(precision="bf16x3", fft="auto") as ctx:
        ctx.load_waveform(x)
        _log("auto-route")
        pos, d, st = ctx.knn(tm, 5, 0)
        print("  stats", st, flush=True)
        assert st["fft_templates"] == 1 and st["fft_block"] == 131072
        po, do, _ = oracle.scan_knn(x, tm, 5, 0)
        assert np.array_equal(pos, po)
        assert np.array_equal(d.view(np.uint32), do.view(np.uint32))


def test_fused_fft_many_templates_matches_mfma(oracle):
    """A batch large enough that the FFT seed pass queues more chunk records
    than the queue's default size: every template must still get its seed,
    and the FFT and matrix-core forms must return the same exact answers."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(4_000_000, 31, events=200)
    lengths = rng.integers(500, 8001, size=64)
    tm = synth.cut_templates(x, onsets, lengths, rng)
    with Context(precision="bf16x3", fft="always") as ctx:
        ctx.load_waveform(x)
        _log("fused-many", templates=len(tm))
        pos, d, st = ctx.knn(tm, 100, 0)
        print("  stats", st, flush=True)
        assert st["fft_templates"] == len(tm)
        ctx.set_fft_mode("off")
        pos2, d2, _ = ctx.knn(tm, 100, 0)
        assert np.array_equal(pos, pos2)
        assert np.array_equal(d.view(np.uint32), d2.view(np.uint32))
    po, do, _ = oracle.scan_knn(x, tm[:2], 100, 0)
    assert np.array_equal(pos[:2], po)
    assert np.array_equal(d[:2].view(np.uint32), do.view(np.uint32))


@pytest.mark.parametrize("m,shape", [(1024, (16384, 1024)), (1025, (16384, 2048)), (2048, (16384, 2048)),
                                     (2049, (32768, 4096)), (4096, (32768, 4096)), (4097, (32768, 8192))])
def test_fused_fft_class_boundary(oracle, m, shape):
    """Every overlap-save shape of the fused FFT form at its boundary:
    B = 2^14 (two workgroups per CU, radix-2 last pass) with pieces for
    templates up to 1024 and 2048 samples, B = 2^15 for up to 4096 and 8192;
    each must return the oracle's answers, alone and batched with a short
    template."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(400_000, 41, events=20)
    tm = synth.cut_templates(x, onsets, [m, 300], rng)
    with Context(precision="bf16x3", fft="always") as ctx:
        ctx.load_waveform(x)
        _log("fft-class", m=m)
        pos, d, st = ctx.knn(tm[:1], 7, 0)
        print("  stats", st, flush=True)
        assert (st["fft_block"], st["fft_lmax"]) == shape
        po, do, _ = oracle.scan_knn(x, tm[:1], 7, 0)
        assert np.array_equal(pos, po)
        assert np.array_equal(d.view(np.uint32), do.view(np.uint32))
        r = ctx.threshold(tm, 0.6, 5)
        tt, tp, td = oracle.scan_threshold(x, tm, 0.6, 5)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


@pytest.mark.parametrize("rho", [0.9, 0.7, 0.5])
def test_threshold_at_scale_shape(oracle, rho):
    """The threshold workload's shape (repeating-source clusters, lengths
    log-uniform 128..4096, FP16 candidate forms with the cost model's FFT
    choice) at a size the oracle scans quickly: rows identical."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    w = synth.make("c2", n=1 << 21, templates=12)
    with Context(precision="fp16", fft="auto") as ctx:
        ctx.load_waveform(w.x)
        _log("c2-shape", rho=rho)
        r = ctx.threshold(w.templates, rho, 0)
        print("  stats", r.stats, flush=True)
    tt, tp, td = oracle.scan_threshold(w.x, w.templates, rho, 0)
    assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
    assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_verify_sequential_statistics(oracle):
    """Long-template candidates take their window statistics from a parallel
    FP64 sum with rigorous bounds on the reference's sequential sum (the
    worst case, then the partial-sum bound), falling back to the sequential
    sums by a wavefront; band-passed noise has window means near 0, where
    the fallbacks happen.  Every window by the sequential sums (verify_seq)
    must give the same answers, and both the oracle's."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(300_000, 43, events=25)
    tm = synth.cut_templates(x, onsets, [520, 1500, 3000, 6000], rng)
    x[rng.integers(0, x.size, size=40)] = synth.MISSING
    with Context(precision="bf16x3") as a, Context(precision="bf16x3", verify_seq=True) as b:
        a.load_waveform(x)
        b.load_waveform(x)
        _log("verify-seq-knn", k=50)
        pa, da, _ = a.knn(tm, 50, 0)
        pb, db, st = b.knn(tm, 50, 0)
        print("  stats", st, flush=True)
        assert np.array_equal(pa, pb) and np.array_equal(da.view(np.uint32), db.view(np.uint32))
        po, do, _ = oracle.scan_knn(x, tm, 50, 0)
        assert np.array_equal(pa, po) and np.array_equal(da.view(np.uint32), do.view(np.uint32))
        _log("verify-seq-thr", rho=0.55)
        ra = a.threshold(tm, 0.55, 0)
        rb = b.threshold(tm, 0.55, 0)
        assert np.array_equal(ra.pos, rb.pos) and np.array_equal(ra.d.view(np.uint32), rb.d.view(np.uint32))
        tt, tp, td = oracle.scan_threshold(x, tm, 0.55, 0)
        assert np.array_equal(ra.tmpl, tt) and np.array_equal(ra.pos, tp)
        assert np.array_equal(ra.d.view(np.uint32), td.view(np.uint32))


def test_autotuned_group_size(oracle):
    """The matrix-core form's templates per workgroup are timed on the first
    batch of a shape (4, 8 or 16) and kept for the context; the answers do
    not depend on the choice: identical with the autotuner off (8), and the
    oracle's."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(400_000, 47, events=40)
    lengths = [int(v) for v in rng.integers(200, 1200, size=24)]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    with Context(precision="bf16x3", fft="off", scan="off") as a, \
            Context(precision="bf16x3", fft="off", scan="off", autotune=False) as b:
        a.load_waveform(x)
        b.load_waveform(x)
        _log("autotune-first")
        pa, da, st = a.knn(tm, 10, 0)
        print("  stats", st, flush=True)
        assert st["tuned_group"] in (4, 8, 16) and st["seconds_tune"] > 0.0
        _log("autotune-cached")
        pa2, da2, st2 = a.knn(tm, 10, 0)
        assert st2["tuned_group"] == st["tuned_group"] and st2["seconds_tune"] == 0.0
        pb, db, stb = b.knn(tm, 10, 0)
        assert stb["tuned_group"] == 8 and stb["seconds_tune"] == 0.0
        for p, d in ((pa2, da2), (pb, db)):
            assert np.array_equal(pa, p) and np.array_equal(da.view(np.uint32), d.view(np.uint32))
        po, do, _ = oracle.scan_knn(x, tm, 10, 0)
        assert np.array_equal(pa, po) and np.array_equal(da.view(np.uint32), do.view(np.uint32))
        r = a.threshold(tm, 0.7, 0)
        tt, tp, td = oracle.scan_threshold(x, tm, 0.7, 0)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


@pytest.mark.parametrize("prec", ["bf16x3", "fp16"])
def test_persistent_launch_equals_grid(oracle, prec):
    """The matrix-core pass as one persistent workgroup per CU claiming
    (block, group) items from per-XCD queues returns the same answers, bit for
    bit, as one workgroup per item, for every group size, on a waveform of
    many blocks and uneven template lengths; the autotuner reports the form it
    chose, and the answers are the oracle's."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(2_000_000, 59, events=60)
    lengths = [int(v) for v in rng.integers(150, 2000, size=40)]
    tm = synth.cut_templates(x, onsets, lengths, rng)
    res = {}
    for launch in ("grid", "queue"):
        for g in (4, 8, 16):
            with Context(precision=prec, fft="off", scan="off", autotune=g, launch=launch) as ctx:
                ctx.load_waveform(x)
                _log("persistent-launch", prec=prec, launch=launch, group=g)
                pos, d, st = ctx.knn(tm, 20, 8)
                assert st["tuned_launch"] == (1 if launch == "grid" else 2) and st["tuned_group"] == g
                r = ctx.threshold(tm, 0.75, 0)
                res[(launch, g)] = (pos, d, r)
    p0, d0, r0 = res[("grid", 8)]
    for key, (p, d, r) in res.items():
        assert np.array_equal(p, p0) and np.array_equal(d.view(np.uint32), d0.view(np.uint32)), key
        assert np.array_equal(r.tmpl, r0.tmpl) and np.array_equal(r.pos, r0.pos), key
        assert np.array_equal(r.d.view(np.uint32), r0.d.view(np.uint32)), key
    with Context(precision=prec, fft="off", scan="off") as ctx:
        ctx.load_waveform(x)
        p, d, st = ctx.knn(tm, 20, 8)
        assert st["tuned_launch"] in (1, 2) and st["tuned_group"] in (4, 8, 16)
        assert np.array_equal(p, p0) and np.array_equal(d.view(np.uint32), d0.view(np.uint32))
    sub = list(range(0, 40, 5))
    po, do, _ = oracle.scan_knn(x, [tm[i] for i in sub], 20, 8)
    assert np.array_equal(p0[sub], po) and np.array_equal(d0[sub].view(np.uint32), do.view(np.uint32))


def test_fft_shape_groups_in_one_batch(oracle):
    """One batch whose FFT templates fall in every overlap-save shape (the
    four fused ones and a hipFFT length): each group runs in its own shape,
    with its own spectra, seed stride and template spectra, and the answers
    are the oracle's."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(400_000, 53, events=30)
    tm = synth.cut_templates(x, onsets, [300, 1500, 3000, 6000, 9000, 700, 2500], rng)
    with Context(precision="bf16x3", fft="always") as ctx:
        ctx.load_waveform(x)
        for k, merge in ((5, 0), (3, 40)):
            _log("fft-groups-knn", k=k, merge=merge)
            pos, d, st = ctx.knn(tm, k, merge)
            print("  stats", st, flush=True)
            assert st["fft_templates"] == len(tm) and st["fft_block"] == 65536
            po, do, _ = oracle.scan_knn(x, tm, k, merge)
            assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
        _log("fft-groups-thr")
        r = ctx.threshold(tm, 0.6, 3)
        tt, tp, td = oracle.scan_threshold(x, tm, 0.6, 3)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_threshold_rows_outlive_result():
    """Threshold rows are views of the library's buffers, freed with the last
    array: arrays kept after the result object is gone still hold the rows
    (compared with a second call after heavy allocation)."""
    import gc
    from tremor_amd import synth
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(200_000, 59, events=30)
    tm = synth.cut_templates(x, onsets, [200, 700], rng)
    with Context(precision="bf16x3") as ctx:
        ctx.load_waveform(x)
        _log("rows-lifetime")
        r = ctx.threshold(tm, 0.3, 0)
        pos, d, tt = r.pos, r.d, r.tmpl
        assert pos.size > 1000
        del r
        gc.collect()
        junk = [np.ones(1 << 20, np.float64) for _ in range(64)]
        r2 = ctx.threshold(tm, 0.3, 0)
        del junk
        assert np.array_equal(pos, r2.pos) and np.array_equal(tt, r2.tmpl)
        assert np.array_equal(d.view(np.uint32), r2.d.view(np.uint32))


def test_fft_shapes_short_waveform(oracle):
    """The FFT form on a waveform shorter than one piece of every shape (a
    single partial piece each), templates close to the waveform's length,
    missing samples and a constant stretch: the oracle's answers."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    rng = np.random.default_rng(61)
    x = synth.bandpassed_noise(9000, rng)
    synth.inject_events(x, 6, rng)
    x[rng.integers(0, x.size, size=15)] = synth.MISSING
    x[4000:4300] = 0.25
    tm = [x[100:400].copy(), x[1000:2500].copy(), x[2000:5000].copy(), x[500:8600].copy(),
          rng.standard_normal(8990).astype(np.float32)]
    with Context(precision="bf16x3", fft="always") as ctx:
        ctx.load_waveform(x)
        for k, merge in ((4, 0), (2, 25)):
            _log("fft-short", k=k, merge=merge)
            pos, d, st = ctx.knn(tm, k, merge)
            print("  stats", st, flush=True)
            po, do, _ = oracle.scan_knn(x, tm, k, merge)
            assert np.array_equal(pos, po) and np.array_equal(d.view(np.uint32), do.view(np.uint32))
        r = ctx.threshold(tm, 0.3, 0)
        tt, tp, td = oracle.scan_threshold(x, tm, 0.3, 0)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))

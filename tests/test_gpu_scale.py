"""Parity at full workload sizes and at the long-template lengths.

- Long templates (20000, 40000, 65536 samples: FFT lengths 2^17 and 2^18):
  the FFT form, the split-BF16 matrix-core form and the early-abandoning scan
  each bit-exact against the oracle's answers (tests/golden/scale_answers.npz,
  written by tests/golden/make_scale_golden.py from the oracle).
- The plain small workload in full (n 10^6, 16 templates of 256): every form
  against the oracle run here.
- The batched-uneven workload in full (n 10^8, 256 templates): every template
  searched on the device, 8 sampled templates checked against the oracle's
  answers; and the same waveform with real-data dynamic range (a x1e6 stretch
  before a x1e-3 one, missing samples), kNN and threshold.
- The FFT forms' error allowance: the forward transforms (hipFFT) and the
  long-template form's hand-written inverse measured against float64
  transforms on adversarial inputs, at every FFT length used.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
U = 2.0 ** -24


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "scale_answers.npz"), allow_pickle=False)


def _log(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _same(pos, d, po, do):
    assert np.array_equal(pos, po)
    assert np.array_equal(np.asarray(d).view(np.uint32), np.asarray(do).view(np.uint32))


FORMS = [("fft", dict(precision="bf16x3", fft="always", scan="off")),
         ("mfma", dict(precision="bf16x3", fft="off", scan="off")),
         ("scan", dict(precision="bf16x3", fft="off", scan="always")),
         ("auto", dict(precision="bf16x3", fft="auto", scan="auto"))]


@pytest.mark.parametrize("form,kw", FORMS, ids=[f[0] for f in FORMS])
def test_long_templates_each_form(gold, form, kw):
    import make_scale_golden as G
    from tremor_amd.capi import Context
    x, tm = G.long_case()
    assert G.digest(x, tm) == str(gold["long/input_sha"]), "workload generator changed"
    with Context(**kw) as ctx:
        ctx.load_waveform(x)
        for merge in (0, 16):
            _log("long", form=form, merge=merge, m=[len(t) for t in tm])
            pos, d, st = ctx.knn(tm, 10, merge)
            print("  stats", {k: st[k] for k in ("fft_templates", "fft_block", "scan_templates", "candidates")},
                  flush=True)
            if form == "fft":
                assert st["fft_templates"] == 3 and st["fft_block"] == 1 << 18
            if form == "scan":
                assert st["scan_templates"] == 3
            if form == "mfma":
                assert st["fft_templates"] == 0 and st["scan_templates"] == 0
            _same(pos, d, gold[f"long/knn{merge}/pos"], gold[f"long/knn{merge}/d"])


def test_fft_length_2p16_form(oracle):
    """Templates of 8193..16384 samples take FFT length 2^16 on the FFT form
    (the hand-written inverse at its smallest size): kNN and threshold
    answers against the oracle run here."""
    from tremor_amd import synth
    from tremor_amd.capi import Context
    rng = np.random.default_rng(41)
    n = 1 << 19
    x = synth.bandpassed_noise(n, rng)
    onsets, _ = synth.inject_events(x, 24, rng)
    tm = synth.cut_templates(x, onsets, [8200, 12000, 16384], rng)
    with Context(precision="bf16x3", fft="always", scan="off") as ctx:
        ctx.load_waveform(x)
        for merge in (0, 16):
            _log("fft 2^16", merge=merge)
            pos, d, st = ctx.knn(tm, 10, merge)
            assert st["fft_templates"] == 3 and st["fft_block"] == 1 << 16
            po, do, _ = oracle.scan_knn(x, tm, 10, merge)
            _same(pos, d, po, do)
        r = ctx.threshold(tm, 0.5, 4)
        tt, tp, td = oracle.scan_threshold(x, tm, 0.5, 4)
        assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
        assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


def test_c0_full_every_form(oracle):
    from tremor_amd import synth
    from tremor_amd.capi import Context
    w = synth.make("c0")
    ref = {merge: oracle.scan_knn(w.x, w.templates, w.k, merge) for merge in (0, 16)}
    tref = {merge: oracle.scan_threshold(w.x, w.templates, 0.8, merge) for merge in (0, 16)}
    for prec, fft in (("fp32", "off"), ("bf16x3", "off"), ("fp16", "off"), ("bf16x3", "always"),
                      ("bf16x3", "auto")):
        with Context(precision=prec, fft=fft) as ctx:
            ctx.load_waveform(w.x)
            for merge in (0, 16):
                _log("c0", prec=prec, fft=fft, merge=merge)
                pos, d, _ = ctx.knn(w.templates, w.k, merge)
                _same(pos, d, ref[merge][0], ref[merge][1])
                r = ctx.threshold(w.templates, 0.8, merge)
                tt, tp, td = tref[merge]
                assert np.array_equal(r.tmpl, tt) and np.array_equal(r.pos, tp)
                assert np.array_equal(r.d.view(np.uint32), td.view(np.uint32))


@pytest.fixture(scope="module")
def c1():
    import make_scale_golden as G
    return G.c1_sample()


def test_c1_full_sampled(gold, c1):
    import make_scale_golden as G
    from tremor_amd.capi import Context
    w, idx = c1
    assert G.digest(w.x, w.templates) == str(gold["c1_sample/input_sha"]), "workload generator changed"
    assert np.array_equal(idx, gold["c1_sample/idx"])
    for kw in (dict(precision="bf16x3", fft="off"), dict(precision="bf16x3", fft="auto")):
        with Context(**kw) as ctx:
            ctx.load_waveform(w.x)
            _log("c1 full", **kw)
            pos, d, st = ctx.knn(w.templates, 100, 0)
            print("  stats", {k: st[k] for k in ("candidates", "verified", "fft_templates", "kernel_ms_main")},
                  flush=True)
            _same(pos[idx], d[idx], gold["c1_sample/knn0/pos"], gold["c1_sample/knn0/d"])


def test_c1_dynamic_range(gold, c1):
    import make_scale_golden as G
    from tremor_amd.capi import Context
    w, idx = c1
    x = G.c1_range(w)
    assert G.digest(x, w.templates) == str(gold["c1_range/input_sha"]), "workload generator changed"
    tm = [w.templates[i] for i in idx]
    for kw in (dict(precision="bf16x3", fft="off"), dict(precision="fp16", fft="off"),
               dict(precision="bf16x3", fft="auto")):
        with Context(**kw) as ctx:
            ctx.load_waveform(x)
            _log("c1 dynamic range", **kw)
            pos, d, st = ctx.knn(w.templates, 100, 0)
            print("  kNN stats", {k: st[k] for k in ("candidates", "verified", "seconds_verify", "kernel_ms_main")},
                  flush=True)
            _same(pos[idx], d[idx], gold["c1_range/knn0/pos"], gold["c1_range/knn0/d"])
            r = ctx.threshold(tm, 0.8, 0)
            print("  threshold stats", {k: r.stats[k] for k in ("candidates", "seconds_verify")}, flush=True)
            assert np.array_equal(r.tmpl, gold["c1_range/thr0/tmpl"])
            assert np.array_equal(r.pos, gold["c1_range/thr0/pos"])
            assert np.array_equal(r.d.view(np.uint32), gold["c1_range/thr0/d"].view(np.uint32))


def _pieces(B, rng):
    from tremor_amd import synth
    base = synth.bandpassed_noise(B, rng).astype(np.float64)
    loud = base.copy()
    loud[: B // 2] *= 1e6
    loud[B // 2:] *= 1e-3
    spikes = rng.standard_normal(B) * 1e-3
    spikes[rng.integers(0, B, size=8)] = rng.standard_normal(8) * 1e5
    dc = base + 1e4
    return np.stack([base, loud, spikes, dc]).astype(np.float32)


def _rel2(got, want, hermitian):
    e = np.abs(got.astype(np.complex128) - want) ** 2
    a = np.abs(want) ** 2
    if hermitian:  # the bins of the full spectrum that the half spectrum stands for
        wgt = np.full(e.shape[-1], 2.0)
        wgt[0] = wgt[-1] = 1.0
        e, a = e * wgt, a * wgt
    return np.sqrt(e.sum(axis=-1) / a.sum(axis=-1))


@pytest.mark.parametrize("lg", [14, 15, 16, 17, 18])
def test_fft_error_within_half_the_allowance(lg):
    """The filter allows each transform a relative 2-norm error of
    4 * 10 log2(B) u (the library's transforms are not analysed); the
    measured error must stay below half of that, on loud/quiet, spiky and
    uncentred pieces."""
    from tremor_amd.capi import Context
    B = 1 << lg
    rng = np.random.default_rng(lg)
    allowance = 4 * 10 * lg * U
    with Context(precision="fp32") as ctx:
        x = _pieces(B, rng)
        X = ctx.fft_probe(B, "r2c", x)
        want = np.fft.rfft(x.astype(np.float64), axis=-1)
        err = _rel2(X, want, True)
        _log("fft r2c", B=B, err_over_u=np.round(err / U, 2).tolist(), allowance_over_u=allowance / U)
        assert np.all(err <= allowance / 2)
        if lg >= 16:
            M = B // 2
            Z = np.fft.rfft(x.astype(np.float64), axis=-1)[:, :M]
            Z[1] *= 1e-6  # magnitudes spread over many orders
            Zc = Z.astype(np.complex64)
            y = ctx.fft_probe(B, "inverse", Zc)
            want = np.fft.ifft(Zc.astype(np.complex128), axis=-1) * M
            err = _rel2(y, want, False)
            _log("fft inverse", B=B, err_over_u=np.round(err / U, 2).tolist(), allowance_over_u=allowance / U)
            assert np.all(err <= allowance / 2)

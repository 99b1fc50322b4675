"""CPU tests of the seeded workload generator (tremor_amd/synth.py): the
waveform's band and scale, the workloads' shapes and template lengths, and
the serving request stream's mix.  The GPU parity tests and bench.py build
their inputs with it, so a drift here would silently change every measured
workload."""
import numpy as np
import pytest

from tremor_amd import synth


def test_bandpassed_noise_band_and_scale():
    rng = np.random.default_rng(5)
    n = 1 << 17
    x = synth.bandpassed_noise(n, rng, chunk=1 << 14)  # several chunks: filter state carried over
    assert x.dtype == np.float32 and x.size == n
    assert abs(float(x.mean())) < 1e-6 and abs(float(x.std()) - 1.0) < 1e-5
    p = np.abs(np.fft.rfft(x.astype(np.float64))) ** 2
    f = np.fft.rfftfreq(n, d=1.0 / synth.FS)
    inband = p[(f >= 2.0) & (f <= 10.0)].sum() / p.sum()
    assert inband > 0.9, inband
    assert p[f < 1.0].sum() / p.sum() < 0.02 and p[f > 20.0].sum() / p.sum() < 0.02
    # Streaming in chunks equals one pass over the whole stretch.
    y = synth.bandpassed_noise(n, np.random.default_rng(5), chunk=n)
    assert np.allclose(x, y, atol=1e-5)


def test_workload_shapes_and_determinism():
    w = synth.make("c0")
    assert w.x.dtype == np.float32 and w.x.size == 10**6
    assert len(w.templates) == 16 and all(t.size == 256 and t.dtype == np.float32 for t in w.templates)
    assert w.k == 10 and w.rhos == (0.8,) and w.events.size == 50
    v = synth.make("c0")
    assert np.array_equal(w.x, v.x) and all(np.array_equal(a, b) for a, b in zip(w.templates, v.templates))
    # Reduced sizes keep each workload's template-length law.
    c1 = synth.make("c1", n=200_000, templates=64)
    m = np.array([t.size for t in c1.templates])
    assert c1.k == 100 and m.min() >= 200 and m.max() <= 2000
    c4 = synth.make("c4", n=300_000, templates=16)
    m = np.array([t.size for t in c4.templates])
    assert m.min() >= 8192 and m.max() <= 65536


def test_templates_are_cut_at_events():
    w = synth.make("c0")
    # Each template is the waveform from up to 7 samples before an event's
    # onset, plus N(0, 0.1) noise: strongly correlated with the waveform there.
    best = []
    for t in w.templates[:6]:
        cs = [np.corrcoef(t, w.x[s:s + t.size])[0, 1] for o in w.events for s in range(max(0, o - 7), o + 1)
              if s + t.size <= w.x.size]
        best.append(max(cs))
    assert min(best) > 0.95, best


def test_request_stream_mix():
    w = synth.make("c0", n=200_000, templates=1)
    rate, duration = 200.0, 20.0
    s = synth.request_stream(w, rate, duration)
    n = len(s)
    assert abs(n - rate * duration) < 5 * np.sqrt(rate * duration)
    t = np.array([r["t"] for r in s])
    assert np.all(np.diff(t) > 0) and t[-1] < duration
    m = np.array([r["template"].size for r in s])
    assert m.min() >= 128 and m.max() <= 8192
    knn = [r for r in s if r["kind"] == "knn"]
    thr = [r for r in s if r["kind"] == "threshold"]
    assert abs(len(knn) / n - 0.7) < 0.05
    assert {r["k"] for r in knn} == {1, 10, 100}
    rho = np.array([r["rho"] for r in thr])
    assert rho.min() >= 0.6 and rho.max() <= 0.9
    assert synth.request_stream(w, rate, 1.0)[0]["t"] == s[0]["t"]  # seeded


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_length_laws_of_the_large_workloads(name):
    w = synth.make(name, n=300_000, templates=200)
    m = np.array([t.size for t in w.templates])
    lo, hi = (128, 4096) if name == "c2" else (128, 8192)
    assert m.min() >= lo and m.max() <= hi
    if name == "c2":
        # log-uniform: about half the templates below the geometric mean.
        assert abs(np.mean(m < np.sqrt(lo * hi)) - 0.5) < 0.15

"""Seeded synthetic seismic workloads.

Waveform: white Gaussian noise band-passed 2-10 Hz at 100 Hz by a 4th-order
Butterworth filter (second-order sections), scaled to unit variance, plus
injected Gabor-like events A exp(-t/tau) sin(2 pi f t) with f ~ U(3, 8) Hz,
tau ~ U(0.3, 1.5) s and A log-uniform in [0.5, 5] at random offsets.
Templates are cut from injected events (from the event onset, lengths per
workload) plus N(0, 0.1) noise.  Generation streams in chunks so long
waveforms never hold a float64 copy.

Workloads (lengths, counts and search parameters):
    c0  n = 10^6,  50 events,   16 templates, m = 256,           kNN k = 10, rho >= 0.8
    c1  n = 10^8,  5000 events, 256 templates, m ~ U[200, 2000],  kNN k = 100
    c2  n = 2^30,  50000 events (10% in repeating-source clusters), 1024 templates,
        m ~ logU[128, 4096], rho in {0.5, 0.7, 0.9}
    c3  n = 2^31 resident, template stream m ~ U[128, 8192]
    c4  n = 2^28,  64 templates, m ~ logU[8192, 65536], kNN k = 10
and the adversarial recipe of the reference's validation (seed 734, n = 8192,
constant stretch, loud stretch then near-silent stretch).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

FS = 100.0
BASE_SEED = 20260917
MISSING = np.float32(-100000.0)


@dataclass
class Workload:
    name: str
    x: np.ndarray                 # float32 waveform
    templates: list               # list of float32 arrays
    k: int = 10
    rhos: tuple = ()
    merge: int = 0
    events: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))


def _sos():
    from scipy.signal import butter
    return butter(4, [2.0, 10.0], btype="bandpass", fs=FS, output="sos")


def bandpassed_noise(n: int, rng: np.random.Generator, chunk: int = 1 << 24, progress=None) -> np.ndarray:
    """Unit-variance band-passed Gaussian noise, float32, streamed.  progress,
    if given, is called with the fraction done every 2^28 samples."""
    from scipy.signal import sosfilt, sosfilt_zi
    sos = _sos()
    zi = sosfilt_zi(sos) * 0.0
    out = np.empty(n, dtype=np.float32)
    s1 = 0.0
    s2 = 0.0
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        w = rng.standard_normal(b - a)
        y, zi = sosfilt(sos, w, zi=zi)
        s1 += float(y.sum())
        s2 += float(np.dot(y, y))
        out[a:b] = y
        if progress is not None and (b % (1 << 28) == 0 or b == n):
            progress(0.8 * b / n)
    mean = s1 / n
    sd = np.sqrt(max(s2 / n - mean * mean, 1e-30))
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        out[a:b] = ((out[a:b].astype(np.float64) - mean) / sd).astype(np.float32)
    return out


def gabor(f: float, tau: float, amp: float, length: int) -> np.ndarray:
    t = np.arange(length) / FS
    return (amp * np.exp(-t / tau) * np.sin(2 * np.pi * f * t)).astype(np.float64)


def inject_events(x: np.ndarray, count: int, rng: np.random.Generator, cluster_frac: float = 0.0):
    """Adds `count` events in place; returns onsets and per-event shapes."""
    n = x.size
    shapes = []
    onsets = np.sort(rng.integers(0, max(1, n - 1024), size=count))
    n_cluster = int(round(cluster_frac * count))
    sources = []
    for i in range(count):
        if i < n_cluster and sources and rng.random() < 0.8:
            f, tau, _ = sources[rng.integers(0, len(sources))]
            # repeating source: the same shape with small perturbations (rho ~ 0.9)
            f = f * (1.0 + 0.02 * rng.standard_normal())
            tau = tau * (1.0 + 0.05 * rng.standard_normal())
        else:
            f = rng.uniform(3.0, 8.0)
            tau = rng.uniform(0.3, 1.5)
            if i < n_cluster:
                sources.append((f, tau, 0))
        amp = float(np.exp(rng.uniform(np.log(0.5), np.log(5.0))))
        length = int(min(5.0 * tau * FS, n - onsets[i]))
        ev = gabor(f, tau, amp, length)
        seg = slice(int(onsets[i]), int(onsets[i]) + length)
        x[seg] = (x[seg].astype(np.float64) + ev).astype(np.float32)
        shapes.append((f, tau, amp))
    return onsets, shapes


def cut_templates(x: np.ndarray, onsets: np.ndarray, lengths, rng: np.random.Generator):
    n = x.size
    out = []
    picks = rng.choice(onsets.size, size=len(lengths), replace=onsets.size < len(lengths))
    for m, e in zip(lengths, picks):
        m = int(m)
        start = int(min(max(0, onsets[e] - rng.integers(0, 8)), n - m))
        t = x[start:start + m].astype(np.float64) + 0.1 * rng.standard_normal(m)
        out.append(t.astype(np.float32))
    return out


def make(name: str, n: int | None = None, templates: int | None = None, seed: int | None = None,
         progress=None) -> Workload:
    """Builds a workload by name ('c0'..'c4', 'validate'); n and templates
    may be reduced for tests, everything else follows the workload."""
    if name == "validate":
        return validate_recipe()
    idx = int(name[1:])
    rng = np.random.default_rng(BASE_SEED + idx if seed is None else seed)
    spec = {
        0: dict(n=10**6, ev=50, T=16, k=10, rhos=(0.8,), cl=0.0),
        1: dict(n=10**8, ev=5000, T=256, k=100, rhos=(), cl=0.0),
        2: dict(n=2**30, ev=50000, T=1024, k=10, rhos=(0.5, 0.7, 0.9), cl=0.1),
        3: dict(n=2**31, ev=100000, T=256, k=10, rhos=(0.6, 0.9), cl=0.0),
        4: dict(n=2**28, ev=12500, T=64, k=10, rhos=(), cl=0.0),
    }[idx]
    N = spec["n"] if n is None else int(n)
    T = spec["T"] if templates is None else int(templates)
    ev = max(1, int(round(spec["ev"] * N / spec["n"])))
    x = bandpassed_noise(N, rng, progress=progress)
    onsets, _ = inject_events(x, ev, rng, spec["cl"])
    if idx == 0:
        lengths = np.full(T, 256)
    elif idx == 1:
        lengths = rng.integers(200, 2001, size=T)
    elif idx == 2:
        lengths = np.exp(rng.uniform(np.log(128), np.log(4096), size=T)).astype(int)
    elif idx == 3:
        lengths = rng.integers(128, 8193, size=T)
    else:
        lengths = np.exp(rng.uniform(np.log(8192), np.log(65536), size=T)).astype(int)
    lengths = np.minimum(lengths, N)
    tm = cut_templates(x, onsets, lengths, rng)
    return Workload(name, x, tm, k=spec["k"], rhos=spec["rhos"], events=onsets)


def validate_recipe(m: int = 64) -> Workload:
    """The reference validation's adversarial input (validate_search.py:50-77)."""
    rng = np.random.default_rng(734)
    n = 8192
    x = rng.normal(size=n).astype("float32")
    x[2000:2128] = 0
    x[5000:6000] *= 1e6
    x[6000:6600] *= 1e-3
    positions = [0, 4090, min(6200, n - m - 1), n - m]
    queries = [x[p:p + m].copy() for p in positions] + [rng.normal(size=m).astype("float32")]
    return Workload("validate", x, [q.astype(np.float32) for q in queries], k=5, rhos=(0.4,))


def request_stream(w: Workload, rate: float, duration: float, seed: int = BASE_SEED + 33, mmin: int = 128,
                   mmax: int = 8192, knn_frac: float = 0.7):
    """The serving workload's request stream: Poisson arrivals at `rate` per
    second for `duration` seconds; each request a template cut from an event
    (length U[mmin, mmax]), kNN with k in {1, 10, 100} (probability
    knn_frac) or threshold with rho ~ U(0.6, 0.9).  Returns a list of dicts
    (t, template, kind, k, rho)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        m = int(rng.integers(mmin, mmax + 1))
        tmpl = cut_templates(w.x, w.events, [min(m, w.x.size)], rng)[0]
        if rng.random() < knn_frac:
            out.append(dict(t=t, template=tmpl, kind="knn", k=int(rng.choice([1, 10, 100])), rho=0.0))
        else:
            out.append(dict(t=t, template=tmpl, kind="threshold", k=0, rho=float(rng.uniform(0.6, 0.9))))
    return out

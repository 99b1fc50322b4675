"""Writes tests/golden/scale_answers.npz: the oracle's answers (a restatement of
the reference sequential scan, pinned to the reference's own SkipSequentialScan
by sss_vectors.npz) for workloads too large for the oracle to rerun inside a
GPU test.  Inputs are regenerated from tremor_amd.synth at test time; their
SHA-256 is stored so a changed generator fails loudly.

  long        n 2^21 band-passed noise with 60 events; templates of 20000,
              40000 and 65536 samples (FFT lengths 2^17 and 2^18, the long-
              template workload's range); kNN k 10, merge 0 and 16
  c1_sample   the batched-uneven workload in full (n 10^8, 256 templates of
              200..2000); 8 sampled templates, kNN k 100, merge 0
  c1_range    the same waveform with real-data dynamic range
              (validate_search.py:54-57): a x1e6 stretch followed by a x1e-3
              one, each 2*10^6 samples, and 1000 missing samples; the same 8
              templates, kNN k 100 and threshold rho 0.8, merge 0

Run in the build container: python tests/golden/make_scale_golden.py
(about 15 minutes on 8 cores)."""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tremor_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scale_answers.npz")


def digest(x, tm):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(x).tobytes())
    for t in tm:
        h.update(np.ascontiguousarray(t).tobytes())
    return h.hexdigest()


def long_case():
    rng = np.random.default_rng(40)
    n = 1 << 21
    x = synth.bandpassed_noise(n, rng)
    onsets, _ = synth.inject_events(x, 60, rng)
    tm = synth.cut_templates(x, onsets, [20000, 40000, 65536], rng)
    return x, tm


def c1_sample():
    w = synth.make("c1")
    idx = np.sort(np.random.default_rng(7).choice(len(w.templates), 8, replace=False))
    return w, idx


def c1_range(w):
    x = w.x.copy()
    a = 30_000_000
    x[a:a + 2_000_000] *= np.float32(1e6)
    x[a + 2_000_000:a + 4_000_000] *= np.float32(1e-3)
    miss = np.random.default_rng(8).integers(0, x.size, size=1000)
    x[miss] = synth.MISSING
    return x


def main():
    from oracle import oracle as O
    out = {}
    t0 = time.time()
    x, tm = long_case()
    out["long/input_sha"] = np.array(digest(x, tm))
    for merge in (0, 16):
        pos, d, _ = O.scan_knn(x, tm, 10, merge)
        out[f"long/knn{merge}/pos"] = pos
        out[f"long/knn{merge}/d"] = d
        print(f"long merge {merge}: {time.time() - t0:.0f} s", flush=True)
    w, idx = c1_sample()
    tm = [w.templates[i] for i in idx]
    out["c1_sample/input_sha"] = np.array(digest(w.x, w.templates))
    out["c1_sample/idx"] = idx
    pos, d, _ = O.scan_knn(w.x, tm, 100, 0)
    out["c1_sample/knn0/pos"] = pos
    out["c1_sample/knn0/d"] = d
    print(f"c1_sample: {time.time() - t0:.0f} s", flush=True)
    xr = c1_range(w)
    out["c1_range/input_sha"] = np.array(digest(xr, w.templates))
    pos, d, _ = O.scan_knn(xr, tm, 100, 0)
    out["c1_range/knn0/pos"] = pos
    out["c1_range/knn0/d"] = d
    tt, tp, td = O.scan_threshold(xr, tm, 0.8, 0)
    out["c1_range/thr0/tmpl"] = tt
    out["c1_range/thr0/pos"] = tp
    out["c1_range/thr0/d"] = td
    print(f"c1_range: {time.time() - t0:.0f} s, threshold rows {tt.size}", flush=True)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()

"""Writes tests/golden/sss_vectors.npz from the reference's own sequential scan
(SkipSequentialScan with lower_bounds = false, SkipSequential.cpp:138-393),
built from the reference sources by `make -C oracle ref-sss` and run on one
rank under an MPI singleton in the build container.  The fixture pins the
oracle's window statistics (orc_window_stats against SkipSequential.cpp:155-
190), kNN scan (orc_scan_knn against :296-393) and threshold scan
(orc_scan_threshold against :223-294) where the reference cannot run.

Cases (inputs regenerated from tremor_amd.synth, their digests stored):
  validate     the reference validation recipe (validate_search.py:50-77):
               n 8192, m 64, 5 templates, k 5, rho 0.4, merge 0 and 16
  c0           the plain small workload: n 10^6, 16 templates of 256,
               k 10, rho 0.8, merge 0 and 16
  adversarial  n 2^18, 8 templates of 200 (a partial last block), missing
               samples (some at window starts), a x1e6 stretch before a
               x1e-3 one, a constant stretch; k 10, rho 0.5, merge 0 and 8
Statistics are stored whole for 'validate' and as a SHA-256 of the bits plus
4096 sampled windows for the larger cases.  kNN and threshold answers are the
reference's CSV rows: positions exactly, correlations as printed (%f).
"""
import csv
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tremor_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sss_vectors.npz")
BIN = os.path.join(ROOT, "oracle", "_ref", "ref_sss")


def adversarial():
    rng = np.random.default_rng(20260917 + 99)
    n, m = 1 << 18, 200
    x = synth.bandpassed_noise(n, rng)
    onsets, _ = synth.inject_events(x, 30, rng)
    tm = synth.cut_templates(x, onsets, [m] * 8, rng)
    x[100000:120000] *= np.float32(1e6)
    x[120000:130000] *= np.float32(1e-3)
    x[150000:151000] = np.float32(2.5)
    miss = rng.integers(0, n, size=40)
    x[miss] = synth.MISSING
    x[[5000, 5001, 5002, 100000 - 1, 120000]] = synth.MISSING
    return x, tm, m


def cases():
    w = synth.validate_recipe()
    yield "validate", w.x, w.templates, 64, 5, [0.4], [0, 16], True
    w = synth.make("c0")
    yield "c0", w.x, w.templates, 256, 10, [0.8], [0, 16], False
    x, tm, m = adversarial()
    yield "adversarial", x, tm, m, 10, [0.5], [0, 8], False


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run(tmp, x, tm, m, k, rhos, merge):
    wf = os.path.join(tmp, "wave.bin")
    qf = os.path.join(tmp, "q.bin")
    x.astype(np.float32).tofile(wf)
    np.stack(tm).astype(np.float32).tofile(qf)
    out = os.path.join(tmp, f"out{merge}")
    cmd = [BIN, wf, str(x.size), str(m), qf, str(len(tm)), "8", out, str(merge), str(k)] + [repr(r) for r in rhos]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    W = x.size - m + 1
    st = np.fromfile(out + ".stats", dtype=np.uint32)
    assert st.size == 2 * W
    with open(out + ".knn.csv") as f:
        knn = [r for r in csv.reader(f, skipinitialspace=True)][1:]
    thr = []
    for i in range(len(rhos)):
        with open(out + f".thr{i}_0.csv") as f:
            thr.append([r for r in csv.reader(f, skipinitialspace=True)][1:])
    return st[:W], st[W:], knn, thr


def main():
    if not os.path.exists(BIN):
        raise SystemExit("oracle/_ref/ref_sss is not built (make -C oracle ref-sss)")
    out = {}
    names = []
    with tempfile.TemporaryDirectory() as tmp:
        for name, x, tm, m, k, rhos, merges, full in cases():
            names.append(name)
            out[f"{name}/input_sha"] = np.array(digest(x, *tm))
            out[f"{name}/m"] = np.int64(m)
            out[f"{name}/k"] = np.int64(k)
            out[f"{name}/rhos"] = np.array(rhos, dtype=np.float64)
            out[f"{name}/merges"] = np.array(merges, dtype=np.int64)
            for merge in merges:
                mean, sd, knn, thr = run(tmp, x, tm, m, k, rhos, merge)
                if merge == merges[0]:
                    out[f"{name}/stats_sha"] = np.array(digest(mean, sd))
                    if full:
                        out[f"{name}/mean_bits"] = mean
                        out[f"{name}/std_bits"] = sd
                    else:
                        idx = np.sort(np.random.default_rng(5).choice(mean.size, 4096, replace=False))
                        out[f"{name}/stat_idx"] = idx.astype(np.int64)
                        out[f"{name}/mean_bits"] = mean[idx]
                        out[f"{name}/std_bits"] = sd[idx]
                kn = np.array(knn, dtype=object)
                out[f"{name}/knn{merge}/tid"] = kn[:, 0].astype(np.int64)
                out[f"{name}/knn{merge}/rank"] = kn[:, 1].astype(np.int64)
                out[f"{name}/knn{merge}/pos"] = kn[:, 2].astype(np.uint64)
                out[f"{name}/knn{merge}/corr"] = kn[:, 3].astype(str)
                for i, rows in enumerate(thr):
                    key = f"{name}/thr{merge}_{i}"
                    r = np.array(rows, dtype=object).reshape(-1, 4)
                    out[key + "/q"] = r[:, 0].astype(np.int64)
                    out[key + "/counter"] = r[:, 1].astype(np.int64)
                    out[key + "/pos"] = r[:, 2].astype(np.uint64)
                    out[key + "/corr"] = r[:, 3].astype(str)
                print(f"{name} merge {merge}: knn rows {len(knn)}, threshold rows {[len(t) for t in thr]}")
    out["cases"] = np.array(names)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()

"""Writes tests/golden/sss_lb_vectors.npz from the reference's own
skip-sequential scan (SkipSequentialScan with lower_bounds = true,
SkipSequential.cpp:138-211 build and lower bound, :253-265 and :321 the skip
tests), built from the reference sources by `make -C oracle ref-sss` and run
on one rank under an MPI singleton in the build container (`ref_sss --lb`).
The fixture pins the oracle's iSAX restatement (orc_sax_breakpoints,
orc_sax_words, orc_paa, orc_lower_bound, the *_lb scans) where the
reference cannot run.

Stored: the 255 breakpoints of cardinality 256 the reference's bounds use;
per case the words of every window (whole, or a SHA-256 plus 4096 sampled
windows), the templates' PAA, every (template, window) lower bound (whole, or
a SHA-256 plus sampled windows), and the reference's kNN and threshold CSV
rows with lower bounds on.

Cases (inputs regenerated from tremor_amd.synth, their digests stored):
  validate16 / validate8  the validation recipe (validate_search.py:50-77),
                          16 segments (the SIMD bound) and 8 (the plain one)
  c0                      the plain small workload, 16 segments
  adversarial             n 2^18, templates of 200, 8 segments, missing
                          samples, x1e6 / x1e-3 stretches, a constant stretch
"""
import csv
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
from tremor_amd import synth  # noqa: E402
from make_sss_golden import adversarial, digest  # noqa: E402

OUT = os.path.join(HERE, "sss_lb_vectors.npz")
BIN = os.path.join(ROOT, "oracle", "_ref", "ref_sss")


def cases():
    w = synth.validate_recipe()
    yield "validate16", w.x, w.templates, 64, 16, 5, [0.4], [0, 16], True
    yield "validate8", w.x, w.templates, 64, 8, 5, [0.4], [0, 16], True
    w = synth.make("c0")
    yield "c0", w.x, w.templates, 256, 16, 10, [0.8], [0, 16], False
    x, tm, m = adversarial()
    yield "adversarial", x, tm, m, 8, 10, [0.5], [0, 8], False


def run(tmp, x, tm, m, seg, k, rhos, merge):
    wf = os.path.join(tmp, "wave.bin")
    qf = os.path.join(tmp, "q.bin")
    x.astype(np.float32).tofile(wf)
    np.stack(tm).astype(np.float32).tofile(qf)
    out = os.path.join(tmp, f"out{merge}")
    cmd = [BIN, "--lb", str(seg), wf, str(x.size), str(m), qf, str(len(tm)), "8", out, str(merge), str(k)]
    cmd += [repr(r) for r in rhos]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    W = x.size - m + 1
    sax = np.fromfile(out + ".sax", dtype=np.uint8).reshape(W, seg)
    paa = np.fromfile(out + ".paa", dtype=np.float32).reshape(len(tm), seg)
    lb = np.fromfile(out + ".lb", dtype=np.float32).reshape(len(tm), W)
    bp = np.fromfile(out + ".bp", dtype=np.float32)
    with open(out + ".knn.csv") as f:
        knn = [r for r in csv.reader(f, skipinitialspace=True)][1:]
    thr = []
    for i in range(len(rhos)):
        with open(out + f".thr{i}_0.csv") as f:
            thr.append([r for r in csv.reader(f, skipinitialspace=True)][1:])
    return sax, paa, lb, bp, knn, thr


def main():
    if not os.path.exists(BIN):
        raise SystemExit("oracle/_ref/ref_sss is not built (make -C oracle ref-sss)")
    out = {}
    names = []
    with tempfile.TemporaryDirectory() as tmp:
        for name, x, tm, m, seg, k, rhos, merges, full in cases():
            names.append(name)
            out[f"{name}/input_sha"] = np.array(digest(x, *tm))
            for key, v in (("m", m), ("segments", seg), ("k", k)):
                out[f"{name}/{key}"] = np.int64(v)
            out[f"{name}/rhos"] = np.array(rhos, dtype=np.float64)
            out[f"{name}/merges"] = np.array(merges, dtype=np.int64)
            for merge in merges:
                sax, paa, lb, bp, knn, thr = run(tmp, x, tm, m, seg, k, rhos, merge)
                if merge == merges[0]:
                    out["breakpoints"] = bp
                    out[f"{name}/sax_sha"] = np.array(digest(sax))
                    out[f"{name}/lb_sha"] = np.array(digest(lb))
                    out[f"{name}/paa"] = paa
                    if full:
                        out[f"{name}/sax"] = sax
                        out[f"{name}/lb"] = lb
                    else:
                        idx = np.sort(np.random.default_rng(6).choice(sax.shape[0], 4096, replace=False))
                        out[f"{name}/idx"] = idx.astype(np.int64)
                        out[f"{name}/sax"] = sax[idx]
                        out[f"{name}/lb"] = lb[:, idx]
                kn = np.array(knn, dtype=object)
                out[f"{name}/knn{merge}/tid"] = kn[:, 0].astype(np.int64)
                out[f"{name}/knn{merge}/pos"] = kn[:, 2].astype(np.uint64)
                out[f"{name}/knn{merge}/corr"] = kn[:, 3].astype(str)
                for i, rows in enumerate(thr):
                    key = f"{name}/thr{merge}_{i}"
                    r = np.array(rows, dtype=object).reshape(-1, 4)
                    out[key + "/q"] = r[:, 0].astype(np.int64)
                    out[key + "/pos"] = r[:, 2].astype(np.uint64)
                    out[key + "/corr"] = r[:, 3].astype(str)
                print(f"{name} merge {merge}: knn rows {len(knn)}, threshold rows {[len(t) for t in thr]}", flush=True)
    out["cases"] = np.array(names)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()

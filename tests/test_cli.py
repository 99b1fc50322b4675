"""The reference-compatible command-line driver (tremor_amd/bin/tmx_main):
argument and input checks of the reference driver (tremor_main.cpp:379-455,
494-606) on the CPU, and byte-identical CSV outputs against the oracle's
answers formatted as the reference writes them (SkipSequential.cpp:233,287,
381,388; timings tremor_main.cpp:698-727) on the GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tremor_amd", "bin", "tmx_main")


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(CLI):
        from tremor_amd import build
        build.build()
    return CLI


def _run(cli, *args):
    return subprocess.run([cli, *map(str, args)], capture_output=True, text=True, timeout=120)


def test_rejects_like_the_reference(cli, tmp_path):
    x = tmp_path / "x.bin"
    q = tmp_path / "q.bin"
    np.zeros(1000, np.float32).tofile(x)
    np.zeros(64 * 3 + 5, np.float32).tofile(q)
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 63)
    assert r.returncode == 2 and "divisible by 8" in r.stderr
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 64, "--threshold", 1.5)
    assert r.returncode == 2 and "between -1.0 and 1.0" in r.stderr
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 64, "--merge-offset", -1)
    assert r.returncode == 2 and "non-negative" in r.stderr
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 64)
    assert r.returncode == 1 and "not divisible by --window-length" in r.stderr
    r = _run(cli, "--dataset", tmp_path / "missing.bin", "--queries", q, "--window-length", 64)
    assert r.returncode == 1 and "Cannot open" in r.stderr
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 64, "--dataset-size", 2000)
    assert r.returncode == 1
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 64, "--method", "index")
    assert r.returncode == 2 and "Unknown --method" in r.stderr
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 64, "--refinement", "order")
    assert r.returncode == 2 and "Unknown --refinement" in r.stderr
    r = _run(cli, "--dataset", x, "--queries", q, "--window-length", 64, "--leaf-size", 0)
    assert r.returncode == 2 and "--leaf-size must be greater than 0" in r.stderr


def _corr(d, m):
    return np.float32(1.0) - np.float32(d) / (np.float32(2.0) * np.float32(m))


@pytest.mark.gpu
def test_csv_outputs_match_reference_format(cli, oracle, tmp_path):
    from tremor_amd import synth
    w = synth.validate_recipe()
    m = 64
    x = tmp_path / "waveform.bin"
    q = tmp_path / "queries.bin"
    w.x.tofile(x)
    np.stack(w.templates).astype(np.float32).tofile(q)
    for merge in (0, 16):
        out = tmp_path / f"knn{merge}.csv"
        r = _run(cli, "--dataset", x, "--queries", q, "--window-length", m, "--mode", "knn", "--k", 5,
                 "--merge-offset", merge, "--output", out)
        assert r.returncode == 0, r.stderr
        pos, d, _ = oracle.scan_knn(w.x, w.templates, 5, merge)
        exp = ["tID, k, pos, corr"]
        for qi in range(len(w.templates)):
            for j in range(5):
                exp.append("%d, %d, %d, %f" % (qi, j + 1, int(pos[qi, j]), float(_corr(d[qi, j], m))))
        assert out.read_text() == "\n".join(exp) + "\n"
        # The tremor method (the iSAX index), leaf and adaptive refinement.
        for extra in (("--method", "tremor"), ("--method", "tremor", "--refinement", "adaptive", "--leaf-size", 100)):
            out_t = tmp_path / f"knn{merge}_tremor.csv"
            r = _run(cli, "--dataset", x, "--queries", q, "--window-length", m, "--mode", "knn", "--k", 5,
                     "--merge-offset", merge, "--output", out_t, *extra)
            assert r.returncode == 0, r.stderr
            assert out_t.read_text() == "\n".join(exp) + "\n"
        timings = (tmp_path / f"knn{merge}.csv.timings.csv").read_text().splitlines()
        assert timings[0] == "node_rank,index_time,query_time,index_gb" and timings[1].startswith("0,")
        prefix = tmp_path / f"thr{merge}"
        r = _run(cli, "--dataset", x, "--queries", q, "--window-length", m, "--mode", "threshold",
                 "--threshold", "0.4", "--merge-offset", merge, "--output", prefix, "--precision", "fp16")
        assert r.returncode == 0, r.stderr
        tt, tp, td = oracle.scan_threshold(w.x, w.templates, np.float32(0.4), merge)
        exp = ["Query, Counter, Position, Cross Correlation"]
        counter = {}
        for t, p, dd in zip(tt, tp, td):
            c = counter.get(int(t), 0)
            counter[int(t)] = c + 1
            exp.append("%d, %d, %d, %f" % (int(t), c, int(p), float(_corr(dd, m))))
        assert (tmp_path / f"thr{merge}_0.csv").read_text() == "\n".join(exp) + "\n"
        # The same rows through the tremor method.
        prefix_t = tmp_path / f"thr{merge}_tremor"
        r = _run(cli, "--dataset", x, "--queries", q, "--window-length", m, "--mode", "threshold",
                 "--threshold", "0.4", "--merge-offset", merge, "--output", prefix_t, "--method", "tremor")
        assert r.returncode == 0, r.stderr
        assert (tmp_path / f"thr{merge}_tremor_0.csv").read_text() == "\n".join(exp) + "\n"

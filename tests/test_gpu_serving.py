"""GPU tests of the serving loop (tmx_server_start / tmx_submit / tmx_wait).

The bar: every request's answer is bit-identical to the same request run
alone through the synchronous C-ABI, and to the oracle's restatement of the
reference scan, whatever batch it shared (kNN and threshold requests of any
k, correlation, merge offset and length mixed), from any submitting thread.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _wave(n, seed, events):
    from tremor_amd import synth
    rng = np.random.default_rng(seed)
    x = synth.bandpassed_noise(n, rng)
    onsets, _ = synth.inject_events(x, events, rng)
    return x, onsets, rng


def _requests(x, onsets, rng, count, mmin, mmax):
    out = []
    for i in range(count):
        m = int(rng.integers(mmin, mmax + 1))
        s = int(min(onsets[rng.integers(0, onsets.size)], x.size - m))
        t = (x[s:s + m].astype(np.float64) + 0.1 * rng.standard_normal(m)).astype(np.float32)
        if rng.random() < 0.7:
            out.append(dict(template=t, kind="knn", k=int(rng.choice([1, 10, 100])),
                            merge=int(rng.choice([0, 0, 16])), rho=0.0))
        else:
            out.append(dict(template=t, kind="threshold", k=0, merge=int(rng.choice([0, 8])),
                            rho=float(rng.uniform(0.6, 0.9))))
    return out


def _alone(ctx, r):
    if r["kind"] == "knn":
        pos, d, _ = ctx.knn([r["template"]], r["k"], r["merge"])
        return pos[0], d[0]
    h = ctx.threshold([r["template"]], r["rho"], r["merge"])
    return h.pos, h.d


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(np.asarray(a[1]).view(np.uint32),
                                                         np.asarray(b[1]).view(np.uint32))


@pytest.mark.parametrize("prec", ["bf16x3", "fp16"])
def test_serving_alone_equals_mix(oracle, prec):
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(300_000, 11, 40)
    reqs = _requests(x, onsets, rng, 40, 64, 2048)
    with Context(precision=prec) as ctx:
        ctx.load_waveform(x)
        _log("serving_mix", prec=prec, requests=len(reqs))
        ctx.serve_start(lanes=2, max_batch=8)
        tickets = [ctx.submit(r["template"], r["kind"], r["k"], r["rho"], r["merge"]) for r in reqs]
        res = [ctx.wait(t, timeout=120.0) for t in tickets]
        ctx.serve_stop()
        assert all(r is not None and r["status"] == 0 for r in res), [r and r["error"] for r in res]
        batches = {r["batch"] for r in res}
        assert len(batches) > 1 and max(r["batch_size"] for r in res) > 1
        for i, (r, got) in enumerate(zip(reqs, res)):
            _log("serving_alone", i=i, kind=r["kind"], m=r["template"].size, k=r["k"], merge=r["merge"])
            alone = _alone(ctx, r)
            assert _same((got["pos"], got["d"]), alone), i
            if r["kind"] == "knn":
                assert got["pos"].size == r["k"]
    # oracle on a subset (the reference scan restated on the CPU)
    for i in range(0, len(reqs), 5):
        r = reqs[i]
        _log("serving_oracle", i=i, kind=r["kind"])
        if r["kind"] == "knn":
            op, od, _ = oracle.scan_knn(x, [r["template"]], r["k"], r["merge"])
            assert _same((res[i]["pos"], res[i]["d"]), (op[0], od[0])), i
        else:
            _, op, od = oracle.scan_threshold(x, [r["template"]], r["rho"], r["merge"])
            assert _same((res[i]["pos"], res[i]["d"]), (op, od)), i


def test_serving_threads_and_fft_requests(oracle):
    """Submissions from several threads; long templates take the FFT form."""
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(400_000, 12, 30)
    reqs = _requests(x, onsets, rng, 24, 200, 9000)
    results = [None] * len(reqs)
    with Context(precision="bf16x3", fft="auto") as ctx:
        ctx.load_waveform(x)
        _log("serving_threads", requests=len(reqs))
        ctx.serve_start(lanes=3, max_batch=4)
        errors = []

        def worker(idx):
            try:
                for i in idx:
                    r = reqs[i]
                    t = ctx.submit(r["template"], r["kind"], r["k"], r["rho"], r["merge"])
                    results[i] = ctx.wait(t, timeout=120.0)
            except Exception as e:  # surfaced below
                errors.append(e)

        th = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        ctx.serve_stop()
        assert not errors, errors
        assert all(r is not None and r["status"] == 0 for r in results)
        for i, r in enumerate(reqs):
            _log("serving_threads_alone", i=i, m=r["template"].size)
            assert _same((results[i]["pos"], results[i]["d"]), _alone(ctx, r)), i
        # the same answers with every template forced through the FFT form
        ctx.set_fft_mode("always")
        for i in range(0, len(reqs), 6):
            _log("serving_threads_fft", i=i)
            assert _same((results[i]["pos"], results[i]["d"]), _alone(ctx, reqs[i])), i


def test_serving_restart_reuses_lanes(oracle):
    """A stopped loop's lanes (their grown buffers and FFT plans of both fused
    lengths) serve the next start, here over another waveform and with one
    more lane; answers stay those of the synchronous path and the oracle."""
    from tremor_amd.capi import Context
    with Context(precision="bf16x3", fft="auto") as ctx:
        for run, (n, seed, lanes) in enumerate(((300_000, 14, 2), (260_000, 15, 3))):
            x, onsets, rng = _wave(n, seed, 30)
            reqs = _requests(x, onsets, rng, 16, 300, 8192)
            ctx.load_waveform(x)
            _log("serving_restart", run=run, lanes=lanes, requests=len(reqs))
            ctx.serve_start(lanes=lanes, max_batch=4)
            tickets = [ctx.submit(r["template"], r["kind"], r["k"], r["rho"], r["merge"]) for r in reqs]
            res = [ctx.wait(t, timeout=120.0) for t in tickets]
            ctx.serve_stop()
            assert all(r is not None and r["status"] == 0 for r in res), [r and r["error"] for r in res]
            for i, r in enumerate(reqs):
                _log("serving_restart_alone", run=run, i=i, m=r["template"].size, kind=r["kind"])
                assert _same((res[i]["pos"], res[i]["d"]), _alone(ctx, r)), i
            for i in range(0, len(reqs), 4):
                r = reqs[i]
                _log("serving_restart_oracle", run=run, i=i, kind=r["kind"])
                if r["kind"] == "knn":
                    op, od, _ = oracle.scan_knn(x, [r["template"]], r["k"], r["merge"])
                    assert _same((res[i]["pos"], res[i]["d"]), (op[0], od[0])), i
                else:
                    _, op, od = oracle.scan_threshold(x, [r["template"]], r["rho"], r["merge"])
                    assert _same((res[i]["pos"], res[i]["d"]), (op, od)), i


def test_serving_errors():
    from tremor_amd.capi import Context, TmxError
    x, onsets, rng = _wave(50_000, 13, 5)
    with Context(precision="bf16x3") as ctx:
        ctx.load_waveform(x)
        with pytest.raises(TmxError):
            ctx.submit(x[:100], "knn", 10)        # not running
        ctx.serve_start(lanes=1, max_batch=2, max_template=1000)
        with pytest.raises(TmxError):
            ctx.submit(x[:2000], "knn", 10)       # longer than max_template
        with pytest.raises(TmxError):
            ctx.submit(x[:100], "knn", 0)         # bad k
        with pytest.raises(TmxError):
            ctx.submit(x[:100], "threshold", rho=float("nan"))
        with pytest.raises(TmxError):
            ctx.load_waveform(x)                  # refused while serving
        t = ctx.submit(x[100:300], "knn", 3)
        r = ctx.wait(t, timeout=60.0)
        assert r["status"] == 0 and r["pos"].size == 3
        with pytest.raises(TmxError):
            ctx.poll(t)                           # retired ticket
        t2 = ctx.submit(x[500:700], "threshold", rho=0.5)
        ctx.serve_stop()                          # answers what was submitted
        r2 = ctx.poll(t2)
        assert r2 is not None and r2["status"] == 0
        with pytest.raises(TmxError):
            ctx.submit(x[:100], "knn", 10)        # stopped


def test_serving_index_many_lengths(oracle):
    """Serving with every template on the iSAX index, requests of 24
    distinct lengths over two lanes: the lanes' batches need more lengths
    between them than the resident words keep, so they drop each other's
    words and indexes; every request still finishes with the answer it gets
    alone and the oracle's."""
    from tremor_amd.capi import Context
    x, onsets, rng = _wave(1 << 16, 13, 20)
    reqs = []
    for i in range(48):
        m = 64 + 16 * (i % 24)
        s = int(min(onsets[rng.integers(0, onsets.size)], x.size - m))
        t = (x[s:s + m].astype(np.float64) + 0.1 * rng.standard_normal(m)).astype(np.float32)
        if i % 3:
            reqs.append(dict(template=t, kind="knn", k=int(rng.choice([1, 5])), merge=0, rho=0.0))
        else:
            reqs.append(dict(template=t, kind="threshold", k=0, merge=0, rho=0.8))
    with Context(index="always", leaf_size=300) as ctx:
        ctx.load_waveform(x)
        _log("serving_index", requests=len(reqs))
        ctx.serve_start(lanes=2, max_batch=6)
        tickets = [ctx.submit(r["template"], r["kind"], r["k"], r["rho"], r["merge"]) for r in reqs]
        res = [ctx.wait(t, timeout=120.0) for t in tickets]
        ctx.serve_stop()
        assert all(r is not None and r["status"] == 0 for r in res), [r and r["error"] for r in res]
        for i, (r, got) in enumerate(zip(reqs, res)):
            if i % 4:
                continue
            _log("serving_index_alone", i=i, kind=r["kind"], m=r["template"].size)
            assert _same((got["pos"], got["d"]), _alone(ctx, r)), i
    for i, (r, got) in enumerate(zip(reqs, res)):
        if r["kind"] == "knn":
            po, do, _ = oracle.scan_knn(x, [r["template"]], r["k"], r["merge"])
            assert _same((got["pos"], got["d"]), (po[0], do[0])), i
        else:
            tt, tp, td = oracle.scan_threshold(x, [r["template"]], r["rho"], r["merge"])
            assert _same((got["pos"], got["d"]), (tp, td)), i

"""Input checks of the C-ABI on the device: non-finite waveform samples are
refused at load, and window ranges or starts that would wrap 64-bit
arithmetic are errors or out-of-range windows, never device reads."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_non_finite_sample_refused():
    from tremor_amd import synth
    from tremor_amd.capi import Context, TmxError

    rng = np.random.default_rng(3)
    x = synth.bandpassed_noise(1 << 16, rng)
    with Context(device=0, precision="bf16x3") as ctx:
        ctx.load_waveform(x)
        assert ctx.n == x.size
        for i, v in ((12345, np.inf), (0, np.nan), (x.size - 1, -np.inf)):
            print(f"[case] non-finite at {i}: {v}", flush=True)
            y = x.copy()
            y[i] = v
            with pytest.raises(TmxError, match=f"sample {i} is not finite"):
                ctx.load_waveform(y)
            assert ctx.n == 0
        # the sentinel and huge finite values are accepted
        y = x.copy()
        y[7] = synth.MISSING
        y[8] = np.float32(3.0e38)
        ctx.load_waveform(y)
        assert ctx.n == y.size


def test_ranges_that_would_wrap():
    from tremor_amd import capi, synth
    from tremor_amd.capi import Context

    rng = np.random.default_rng(4)
    n = 1 << 15
    x = synth.bandpassed_noise(n, rng)
    with Context(device=0, precision="fp32") as ctx:
        ctx.load_waveform(x)
        m = 100
        W = n - m + 1
        a = np.zeros(8, dtype=np.float32)
        b = np.zeros(8, dtype=np.float32)
        for first, count in ((2**64 - 10, 20), (1, 2**64 - 1), (W, 1), (W - 5, 6)):
            print(f"[case] window_stats first={first} count={count}", flush=True)
            # raw call: the range check comes before any output is written
            rc = ctx._L.tmx_window_stats(ctx._h, m, first, count, a.ctypes.data_as(capi._f32p),
                                         b.ctypes.data_as(capi._f32p))
            assert rc != 0
            assert b"out of bounds" in ctx._L.tmx_last_error(ctx._h)
        mean, sd = ctx.window_stats(m, W - 5, 5)
        assert mean.size == 5 and np.all(sd > 0)
        t = rng.standard_normal(m).astype(np.float32)
        starts = np.array([0, W - 1, W, n, 2**64 - 1, 2**64 - m, 2**63], dtype=np.uint64)
        d = ctx.distance_batch(t, starts)
        fmax = np.float32(np.finfo(np.float32).max)
        assert d[0] < fmax and d[1] < fmax
        assert np.all(d[2:] == fmax)

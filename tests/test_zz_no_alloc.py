"""The allocation-free distance batch (tmx_distance_batch_noalloc) against
tmx_distance_batch, through the C-ABI: bit-identical distances on two calls
in a row on one device, the second with a shorter template and fewer
windows, so the shared scratch holds the first call's larger contents when
the second runs; then once on a stream of the caller's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FMAX = np.float32(np.finfo(np.float32).max)


def _case(rng, n, m, count):
    starts = rng.integers(0, n - m + 1, size=count).astype(np.uint64)
    # windows past the end (including starts whose sum with m would wrap)
    starts[:3] = np.array([n - m + 1, n, 2**64 - 5], dtype=np.uint64)
    t = (rng.standard_normal(m) * 2.0 + 0.5).astype(np.float32)
    return t, starts


def test_noalloc_matches_distance_batch_twice():
    from hip_mem import DeviceArray, Stream
    from tremor_amd import synth
    from tremor_amd.capi import Context, TmxError

    rng = np.random.default_rng(11)
    n = 1 << 18
    x = synth.bandpassed_noise(n, rng)
    x[1000] = synth.MISSING
    x[5000:9000] = 1.25  # constant stretch: windows inside it have std 0
    with Context(device=0, precision="fp32") as ctx:
        ctx.load_waveform(x)
        for m, count in ((3000, 20000), (257, 5000)):
            print(f"[case] noalloc m={m} count={count}", flush=True)
            t, starts = _case(rng, n, m, count)
            starts[3] = 1000    # first sample missing
            starts[4] = 5000    # constant window
            want = ctx.distance_batch(t, starts)
            d_t, d_s = DeviceArray(t), DeviceArray(starts)
            d_o = DeviceArray(np.full(count, -1.0, dtype=np.float32))
            ctx.distance_batch_noalloc(d_t.ptr, m, d_s.ptr, count, d_o.ptr)
            ctx.sync()
            got = d_o.numpy()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), m
            assert np.all(got[:5] == FMAX)
        # on a stream of the caller's
        print("[case] noalloc on a caller stream m=500 count=777", flush=True)
        s = Stream()
        m, count = 500, 777
        t, starts = _case(rng, n, m, count)
        want = ctx.distance_batch(t, starts)
        d_t, d_s = DeviceArray(t), DeviceArray(starts)
        d_o = DeviceArray(np.zeros(count, dtype=np.float32))
        ctx.distance_batch_noalloc(d_t.ptr, m, d_s.ptr, count, d_o.ptr, s.handle)
        s.synchronize()
        assert np.array_equal(d_o.numpy().view(np.uint32), want.view(np.uint32))
        s.destroy()
        # longer than the scratch holds: refused before anything is enqueued
        with pytest.raises(TmxError, match="allocation-free"):
            ctx.distance_batch_noalloc(d_t.ptr, 65537, d_s.ptr, count, d_o.ptr)

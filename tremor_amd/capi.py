"""ctypes binding of the C-ABI in include/tremor_amd/tmx.h (libtremor_amd.so).

The library is loaded from this package directory; if it is missing the
import of this module raises: there is no fallback to any other code path.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtremor_amd.so")

PREC_FP32 = 0
PREC_BF16X3 = 1
PREC_FP16 = 2
PRECISIONS = {"fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp16": PREC_FP16}
FFT_MODES = {"auto": 0, "off": 1, "always": 2}
SCAN_MODES = {"auto": 0, "off": 1, "always": 2}

EXPORTS = [
    "tmx_cfg_default", "tmx_create", "tmx_destroy", "tmx_last_error", "tmx_set_precision", "tmx_set_fft_mode",
    "tmx_set_scan_mode", "tmx_set_index_mode",
    "tmx_load_waveform",
    "tmx_waveform_length", "tmx_knn", "tmx_threshold", "tmx_hits_free", "tmx_distance_batch",
    "tmx_distance_batch_noalloc", "tmx_window_stats",
    "tmx_prepare_template", "tmx_fft_probe", "tmx_index_probe", "tmx_sync",
    "tmx_server_cfg_default", "tmx_server_start", "tmx_server_stop", "tmx_submit", "tmx_poll", "tmx_wait",
    "tmx_result_free",
]

REQ_KNN = 0
REQ_THRESHOLD = 1


class Cfg(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("precision", ctypes.c_int32), ("cand_capacity", ctypes.c_int64),
                ("seed_target", ctypes.c_int32), ("verbose", ctypes.c_int32), ("fft_mode", ctypes.c_int32),
                ("scan_mode", ctypes.c_int32), ("verify_seq", ctypes.c_int32), ("autotune", ctypes.c_int32),
                ("max_candidates", ctypes.c_int64), ("lower_bounds", ctypes.c_int32),
                ("paa_segments", ctypes.c_int32), ("mfma_launch", ctypes.c_int32), ("index_mode", ctypes.c_int32),
                ("leaf_size", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("seconds_total", ctypes.c_double), ("seconds_seed", ctypes.c_double),
                ("seconds_main", ctypes.c_double), ("seconds_verify", ctypes.c_double),
                ("seconds_select", ctypes.c_double), ("comparisons", ctypes.c_int64),
                ("candidates", ctypes.c_int64), ("retries", ctypes.c_int32), ("max_abs_rho_err", ctypes.c_float),
                ("kernel_ms_main", ctypes.c_double), ("flops_main", ctypes.c_double),
                ("fft_templates", ctypes.c_int32), ("fft_block", ctypes.c_int32), ("verified", ctypes.c_int64),
                ("scan_templates", ctypes.c_int32), ("scan_blocks", ctypes.c_double),
                ("tuned_group", ctypes.c_int32), ("seconds_tune", ctypes.c_double), ("fft_lmax", ctypes.c_int32),
                ("splits", ctypes.c_int32), ("lb_skipped", ctypes.c_int64), ("tuned_launch", ctypes.c_int32),
                ("index_templates", ctypes.c_int32), ("index_windows", ctypes.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Hits(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int64), ("tmpl", ctypes.POINTER(ctypes.c_int32)),
                ("counter", ctypes.POINTER(ctypes.c_int64)), ("pos", ctypes.POINTER(ctypes.c_uint64)),
                ("d", ctypes.POINTER(ctypes.c_float))]


class ServerCfg(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_int32), ("max_batch", ctypes.c_int32), ("mem_budget", ctypes.c_int64),
                ("max_template", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("m", ctypes.c_int32), ("tmpl", ctypes.POINTER(ctypes.c_float)),
                ("k", ctypes.c_int32), ("rho", ctypes.c_float), ("merge_offset", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("kind", ctypes.c_int32), ("count", ctypes.c_int64),
                ("pos", ctypes.POINTER(ctypes.c_uint64)), ("d", ctypes.POINTER(ctypes.c_float)),
                ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("batch", ctypes.c_int64), ("batch_size", ctypes.c_int32), ("lane", ctypes.c_int32),
                ("error", ctypes.c_char * 160)]


_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python -m tremor_amd.build` (no fallback path exists)")
    L = ctypes.CDLL(path)
    L.tmx_cfg_default.argtypes = [ctypes.POINTER(Cfg)]
    L.tmx_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(Cfg)]
    L.tmx_destroy.argtypes = [ctypes.c_void_p]
    L.tmx_destroy.restype = None
    L.tmx_last_error.argtypes = [ctypes.c_void_p]
    L.tmx_last_error.restype = ctypes.c_char_p
    L.tmx_set_precision.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.tmx_set_fft_mode.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.tmx_set_scan_mode.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.tmx_set_index_mode.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.tmx_load_waveform.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint64]
    L.tmx_waveform_length.argtypes = [ctypes.c_void_p]
    L.tmx_waveform_length.restype = ctypes.c_uint64
    L.tmx_knn.argtypes = [ctypes.c_void_p, _f32p, _i32p, _u64p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                          _u64p, _f32p, ctypes.POINTER(Stats)]
    L.tmx_threshold.argtypes = [ctypes.c_void_p, _f32p, _i32p, _u64p, ctypes.c_int32, ctypes.c_float,
                                ctypes.c_int32, ctypes.POINTER(ctypes.POINTER(Hits)), ctypes.POINTER(Stats)]
    L.tmx_hits_free.argtypes = [ctypes.POINTER(Hits)]
    L.tmx_hits_free.restype = None
    L.tmx_distance_batch.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_int32, _u64p, ctypes.c_int32, _f32p]
    L.tmx_distance_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    L.tmx_window_stats.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint64, _f32p, _f32p]
    L.tmx_prepare_template.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_int32, _f32p, _i32p]
    L.tmx_fft_probe.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _f32p, ctypes.c_int32, _f32p]
    _u32p = ctypes.POINTER(ctypes.c_uint32)
    L.tmx_index_probe.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _u32p, ctypes.c_uint64, _u32p, _u32p,
                                  ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_uint8)]
    L.tmx_sync.argtypes = [ctypes.c_void_p]
    L.tmx_server_cfg_default.argtypes = [ctypes.POINTER(ServerCfg)]
    L.tmx_server_start.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerCfg)]
    L.tmx_server_stop.argtypes = [ctypes.c_void_p]
    L.tmx_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request)]
    L.tmx_submit.restype = ctypes.c_int64
    L.tmx_poll.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(Result)]
    L.tmx_wait.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(Result)]
    L.tmx_result_free.argtypes = [ctypes.POINTER(Result)]
    L.tmx_result_free.restype = None
    return L


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


class TmxError(RuntimeError):
    pass


def _p(a, t):
    return a.ctypes.data_as(t)


def pack_templates(templates):
    lengths = np.array([len(t) for t in templates], dtype=np.int32)
    offsets = np.zeros(len(templates), dtype=np.uint64)
    if len(templates) > 1:
        offsets[1:] = np.cumsum(lengths[:-1].astype(np.uint64))
    flat = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.float32) for t in templates]))
    return flat, lengths, offsets


@dataclass
class ThresholdResult:
    tmpl: np.ndarray
    counter: np.ndarray
    pos: np.ndarray
    d: np.ndarray
    stats: dict


class _HitsOwner:
    """Holds a tmx_hits result until every array viewing it is gone."""

    def __init__(self, L, hits):
        self._L = L
        self._hits = hits

    def __del__(self):
        if self._hits is not None:
            self._L.tmx_hits_free(self._hits)
            self._hits = None


def _view(ptr, ctype, n, owner):
    """A numpy array over n elements at ptr, keeping owner alive."""
    buf = (ctype * n).from_address(ctypes.addressof(ptr.contents))
    buf._owner = owner
    return np.ctypeslib.as_array(buf)


class Context:
    """One device context: a resident waveform and search calls on it."""

    def __init__(self, device: int = 0, precision: str | int = "bf16x3", cand_capacity: int = 0,
                 seed_target: int = 0, verbose: bool = False, fft: str = "auto", scan: str = "auto",
                 verify_seq: bool = False, autotune: bool | int = True, max_candidates: int = 0,
                 lower_bounds: bool = False, paa_segments: int = 16, launch: str = "auto",
                 index: str = "off", leaf_size: int = 2000):
        L = lib()
        cfg = Cfg()
        L.tmx_cfg_default(ctypes.byref(cfg))
        cfg.device = device
        cfg.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        cfg.cand_capacity = cand_capacity
        cfg.seed_target = seed_target
        cfg.verbose = 1 if verbose else 0
        cfg.fft_mode = FFT_MODES[fft]
        cfg.scan_mode = SCAN_MODES[scan]
        cfg.verify_seq = 1 if verify_seq else 0
        # True: timed per batch shape; False: fixed 8; 4, 8 or 16: that size, fixed
        cfg.autotune = (1 if autotune else 0) if isinstance(autotune, bool) else int(autotune)
        cfg.max_candidates = int(max_candidates)
        # the skip-sequential method's iSAX lower bound on the early-abandoning scan
        cfg.lower_bounds = 1 if lower_bounds else 0
        cfg.paa_segments = int(paa_segments)
        cfg.mfma_launch = {"auto": 0, "grid": 1, "queue": 2}[launch]
        cfg.index_mode = {"off": 0, "auto": 1, "always": 2}[index]
        cfg.leaf_size = int(leaf_size)
        h = ctypes.c_void_p()
        rc = L.tmx_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc != 0:
            raise TmxError(f"tmx_create failed ({rc}): no usable HIP device {device}")
        self._h = h
        self._L = L

    def close(self):
        if self._h:
            self._L.tmx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise TmxError(self._L.tmx_last_error(self._h).decode())

    def set_precision(self, precision):
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._check(self._L.tmx_set_precision(self._h, p))

    def set_fft_mode(self, mode: str):
        self._check(self._L.tmx_set_fft_mode(self._h, FFT_MODES[mode]))

    def set_scan_mode(self, mode: str):
        """Early-abandoning scan: "auto" (cost model), "off" or "always"."""
        self._check(self._L.tmx_set_scan_mode(self._h, SCAN_MODES[mode]))

    def set_index_mode(self, mode: str, leaf_size: int = 2000):
        """iSAX index: "off", "auto" (the 1% rule) or "always"; leaves of at
        most leaf_size windows."""
        self._check(self._L.tmx_set_index_mode(self._h, {"off": 0, "auto": 1, "always": 2}[mode], leaf_size))

    def load_waveform(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        self._check(self._L.tmx_load_waveform(self._h, _p(x, _f32p), x.size))

    @property
    def n(self) -> int:
        return int(self._L.tmx_waveform_length(self._h))

    def knn(self, templates, k: int, merge: int = 0):
        flat, lengths, offsets = pack_templates(templates)
        T = len(templates)
        pos = np.zeros((T, k), dtype=np.uint64)
        d = np.zeros((T, k), dtype=np.float32)
        st = Stats()
        self._check(self._L.tmx_knn(self._h, _p(flat, _f32p), _p(lengths, _i32p), _p(offsets, _u64p), T, k, merge,
                                    _p(pos, _u64p), _p(d, _f32p), ctypes.byref(st)))
        return pos, d, st.as_dict()

    def threshold(self, templates, rho: float, merge: int = 0) -> ThresholdResult:
        flat, lengths, offsets = pack_templates(templates)
        T = len(templates)
        out = ctypes.POINTER(Hits)()
        st = Stats()
        self._check(self._L.tmx_threshold(self._h, _p(flat, _f32p), _p(lengths, _i32p), _p(offsets, _u64p), T,
                                          float(rho), merge, ctypes.byref(out), ctypes.byref(st)))
        h = out.contents
        n = int(h.count)
        if not n:
            self._L.tmx_hits_free(out)
            return ThresholdResult(np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros(0, np.uint64),
                                   np.zeros(0, np.float32), st.as_dict())
        # The rows stay in the library's buffers (no copy of what can be
        # gigabytes); they are freed with the last of the four arrays.
        owner = _HitsOwner(self._L, out)
        return ThresholdResult(_view(h.tmpl, ctypes.c_int32, n, owner), _view(h.counter, ctypes.c_int64, n, owner),
                               _view(h.pos, ctypes.c_uint64, n, owner), _view(h.d, ctypes.c_float, n, owner),
                               st.as_dict())

    def distance_batch(self, template, starts):
        t = np.ascontiguousarray(template, dtype=np.float32)
        s = np.ascontiguousarray(starts, dtype=np.uint64)
        out = np.zeros(s.size, dtype=np.float32)
        self._check(self._L.tmx_distance_batch(self._h, _p(t, _f32p), t.size, _p(s, _u64p), s.size,
                                               _p(out, _f32p)))
        return out

    def distance_batch_noalloc(self, d_template: int, m: int, d_starts: int, count: int, d_out: int,
                               stream: int = 0):
        """Device pointers in and out (integers); enqueued on `stream` (0: the
        context's stream), not waited for."""
        self._check(self._L.tmx_distance_batch_noalloc(self._h, ctypes.c_void_p(d_template), m,
                                                       ctypes.c_void_p(d_starts), count, ctypes.c_void_p(d_out),
                                                       ctypes.c_void_p(stream or None)))

    def window_stats(self, m: int, first: int, count: int):
        mean = np.zeros(count, dtype=np.float32)
        sd = np.zeros(count, dtype=np.float32)
        self._check(self._L.tmx_window_stats(self._h, m, first, count, _p(mean, _f32p), _p(sd, _f32p)))
        return mean, sd

    def prepare_template(self, template):
        t = np.ascontiguousarray(template, dtype=np.float32)
        z = np.zeros(t.size, dtype=np.float32)
        order = np.zeros(max(t.size // 32, 1), dtype=np.int32)
        self._check(self._L.tmx_prepare_template(self._h, _p(t, _f32p), t.size, _p(z, _f32p), _p(order, _i32p)))
        return z, order[: t.size // 32]

    def fft_probe(self, B: int, kind: str, data):
        """The library's own FFT plans on host data: kind "r2c" (real pieces of
        B samples -> B/2 + 1 bins) or "inverse" (complex of length B/2,
        unnormalised backward).  Returns complex64 rows."""
        if kind == "r2c":
            a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1, B)
            out = np.zeros((a.shape[0], B // 2 + 1), dtype=np.complex64)
            k = 0
        else:
            a = np.ascontiguousarray(data, dtype=np.complex64).reshape(-1, B // 2)
            out = np.zeros((a.shape[0], B // 2), dtype=np.complex64)
            k = 1
        self._check(self._L.tmx_fft_probe(self._h, B, k, a.ctypes.data_as(_f32p), a.shape[0],
                                          out.ctypes.data_as(_f32p)))
        return out

    def index_probe(self, m: int, leaf_size: int = 2000):
        """The iSAX index of template length m as a search builds it: (window
        ids in key order, leaves as an int64 array of rows (start, count,
        depth) in key order, the windows' 16-segment words)."""
        W = self.n - m + 1
        u32 = ctypes.POINTER(ctypes.c_uint32)
        nl = ctypes.c_int64(0)
        ids = np.zeros(max(W, 1), dtype=np.uint32)
        words = np.zeros((max(W, 1), 16), dtype=np.uint8)
        self._check(self._L.tmx_index_probe(self._h, m, leaf_size, ids.ctypes.data_as(u32), ids.size, None, None,
                                            None, 0, ctypes.byref(nl), words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        n = nl.value
        st = np.zeros(max(n, 1), dtype=np.uint32)
        ct = np.zeros(max(n, 1), dtype=np.uint32)
        dp = np.zeros(max(n, 1), dtype=np.int32)
        self._check(self._L.tmx_index_probe(self._h, m, leaf_size, None, 0, st.ctypes.data_as(u32),
                                            ct.ctypes.data_as(u32), dp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            n, ctypes.byref(nl), None))
        leaves = np.stack([st[:n], ct[:n], dp[:n]], axis=1).astype(np.int64)
        return ids[:W], leaves, words[:W]

    def sync(self):
        self._check(self._L.tmx_sync(self._h))

    # ---- serving loop ----

    def serve_start(self, lanes: int = 2, max_batch: int = 64, mem_budget: int = 0, max_template: int = 65536):
        c = ServerCfg()
        self._L.tmx_server_cfg_default(ctypes.byref(c))
        c.lanes, c.max_batch, c.mem_budget, c.max_template = lanes, max_batch, mem_budget, max_template
        self._check(self._L.tmx_server_start(self._h, ctypes.byref(c)))

    def serve_stop(self):
        self._check(self._L.tmx_server_stop(self._h))

    def submit(self, template, kind: str = "knn", k: int = 10, rho: float = 0.8, merge: int = 0) -> int:
        """Queues one request; returns its ticket.  kind: 'knn' or 'threshold'."""
        t = np.ascontiguousarray(template, dtype=np.float32)
        r = Request()
        r.kind = REQ_KNN if kind == "knn" else REQ_THRESHOLD
        r.m = t.size
        r.tmpl = _p(t, _f32p)
        r.k = k
        r.rho = rho
        r.merge_offset = merge
        ticket = self._L.tmx_submit(self._h, ctypes.byref(r))
        if ticket <= 0:
            raise TmxError(self._L.tmx_last_error(self._h).decode())
        return int(ticket)

    def _collect(self, rc, res):
        if rc < 0:
            raise TmxError("unknown ticket")
        if rc == 1:
            return None
        try:
            n = int(res.count)
            out = {"status": int(res.status), "error": res.error.decode(errors="replace"),
                   "pos": np.ctypeslib.as_array(res.pos, (n,)).copy() if n else np.zeros(0, np.uint64),
                   "d": np.ctypeslib.as_array(res.d, (n,)).copy() if n else np.zeros(0, np.float32),
                   "submit_s": res.submit_s, "start_s": res.start_s, "done_s": res.done_s,
                   "batch": int(res.batch), "batch_size": int(res.batch_size), "lane": int(res.lane)}
        finally:
            self._L.tmx_result_free(ctypes.byref(res))
        return out

    def poll(self, ticket: int):
        """The answer as a dict, or None while pending."""
        res = Result()
        return self._collect(self._L.tmx_poll(self._h, ticket, ctypes.byref(res)), res)

    def wait(self, ticket: int, timeout: float = -1.0):
        """Waits for the answer (timeout < 0: no limit); None on timeout."""
        res = Result()
        return self._collect(self._L.tmx_wait(self._h, ticket, float(timeout), ctypes.byref(res)), res)

"""ctypes view of the CPU oracle (liboracle.so) and, where it was built, of the
reference's own hot-path functions (_ref/libtremor_ref.so), plus an FP64
Pearson oracle in numpy.

TEST INFRASTRUCTURE ONLY.  tests/, the smoke entry point and the CPU-baseline
leg of the benchmark use this module as the checker.  The product (the
tremor_amd package and its C-ABI library) never imports it.

The FP64 oracle follows experiments/validate_search.py:80-86: z-normalise
every window and the template in FP64 (population std), correlate, and mark
windows with std <= 1e-12 as -inf.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libtremor_ref.so")
MISSING = np.float32(-100000.0)
FLT_MAX = np.float32(np.finfo(np.float32).max)

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(ref_dir: str | None = "/root/reference") -> None:
    """Compile liboracle.so, and _ref/ when the reference sources exist."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref_dir and os.path.isfile(os.path.join(ref_dir, "tremor", "distance_computers", "ZnormSearch.hpp")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={ref_dir}"], check=True)
        # The reference's own sequential scan needs MPICH; where it is absent
        # the committed fixtures stand in for it.
        if os.path.isfile("/opt/conda/include/mpi.h"):
            subprocess.run(["make", "-s", "-C", HERE, "ref-sss", f"REF={ref_dir}"], check=True)


def _p(a, t):
    return a.ctypes.data_as(t)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(None)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_cc_to_ed.restype = ctypes.c_float
        L.orc_cc_to_ed.argtypes = [ctypes.c_float, ctypes.c_int]
        L.orc_ed_to_cc.restype = ctypes.c_float
        L.orc_ed_to_cc.argtypes = [ctypes.c_float, ctypes.c_int]
        L.orc_znormalize.argtypes = [_f32p, ctypes.c_int]
        L.orc_block_order.restype = ctypes.c_int
        L.orc_block_order.argtypes = [_f32p, ctypes.c_int, _i32p]
        L.orc_window_stats.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.orc_dist.restype = ctypes.c_float
        L.orc_dist.argtypes = [_f32p, _f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, _i32p, _i32p]
        L.orc_dist_all.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_int64, _f32p, _f32p, ctypes.c_int, _f32p]
        L.orc_select_knn.restype = ctypes.c_int
        L.orc_select_knn.argtypes = [_u64p, _f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _u64p, _f32p]
        L.orc_scan_knn.restype = ctypes.c_int
        L.orc_scan_knn.argtypes = [_f32p, ctypes.c_int64, _f32p, _i32p, _i64p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, _u64p, _f32p, _f64p]
        L.orc_scan_threshold.restype = ctypes.c_int64
        L.orc_scan_threshold.argtypes = [_f32p, ctypes.c_int64, _f32p, _i32p, _i64p, ctypes.c_int, ctypes.c_float,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int64, _i32p, _u64p, _f32p]
        _u8p = ctypes.POINTER(ctypes.c_uint8)
        L.orc_sax_breakpoints.argtypes = [_f32p]
        L.orc_paa.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p]
        L.orc_sax_words.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u8p]
        L.orc_lower_bounds.argtypes = [_f32p, _f32p, _u8p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _f32p]
        L.orc_scan_knn_lb.restype = ctypes.c_int
        L.orc_scan_knn_lb.argtypes = [_f32p, ctypes.c_int64, _f32p, _i32p, _i64p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64p, _f32p, _f64p]
        L.orc_scan_threshold_lb.restype = ctypes.c_int64
        L.orc_scan_threshold_lb.argtypes = [_f32p, ctypes.c_int64, _f32p, _i32p, _i64p, ctypes.c_int,
                                            ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int64, _i32p, _u64p, _f32p]
        _lib = L
    return _lib


def ref():
    """The reference's own functions, or None where _ref was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = ctypes.CDLL(REF_PATH)
        R.ref_cc_to_ed.restype = ctypes.c_float
        R.ref_cc_to_ed.argtypes = [ctypes.c_float, ctypes.c_int]
        R.ref_ed_to_cc.restype = ctypes.c_float
        R.ref_ed_to_cc.argtypes = [ctypes.c_float, ctypes.c_int]
        R.ref_znormalize.argtypes = [_f32p, ctypes.c_int]
        R.ref_block_order.restype = ctypes.c_int
        R.ref_block_order.argtypes = [_f32p, ctypes.c_int, _i32p]
        R.ref_dist.restype = ctypes.c_float
        R.ref_dist.argtypes = [_f32p, _f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, _i32p, _i32p]
        R.ref_select_knn.restype = ctypes.c_int
        R.ref_select_knn.argtypes = [_u64p, _f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _u64p, _f32p]
        R.ref_bsf_init.restype = ctypes.c_void_p
        R.ref_bsf_init.argtypes = [ctypes.c_int]
        R.ref_bsf_insert_offset.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_uint64, ctypes.c_int]
        R.ref_bsf_read.argtypes = [ctypes.c_void_p, _f32p, _i64p]
        R.ref_bsf_destroy.argtypes = [ctypes.c_void_p]
        _ref = R
    return _ref


# ---------------------------------------------------------------- wrappers

def cc_to_ed(cc: float, m: int) -> np.float32:
    return np.float32(lib().orc_cc_to_ed(float(cc), int(m)))


def ed_to_cc(ed: float, m: int) -> np.float32:
    return np.float32(lib().orc_ed_to_cc(float(ed), int(m)))


def znormalize(t) -> np.ndarray:
    q = np.ascontiguousarray(t, dtype=np.float32).copy()
    lib().orc_znormalize(_p(q, _f32p), q.size)
    return q


def block_order(qz) -> np.ndarray:
    qz = np.ascontiguousarray(qz, dtype=np.float32)
    order = np.zeros(qz.size // 32 + 1, dtype=np.int32)
    b = lib().orc_block_order(_p(qz, _f32p), qz.size, _p(order, _i32p))
    return order[:b]


def window_stats(raw, m: int, threads: int = 0):
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    w = raw.size - m + 1
    mean = np.zeros(max(w, 0), dtype=np.float32)
    sd = np.zeros(max(w, 0), dtype=np.float32)
    if w > 0:
        lib().orc_window_stats(_p(raw, _f32p), raw.size, m, threads or os.cpu_count(), _p(mean, _f32p),
                               _p(sd, _f32p))
    return mean, sd


def dist(qz, raw, start: int, mean: float, sd: float, bound: float = float(FLT_MAX), order=None):
    qz = np.ascontiguousarray(qz, dtype=np.float32)
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    if order is None:
        order = block_order(qz)
    order = np.ascontiguousarray(order, dtype=np.int32)
    blocks = ctypes.c_int32(0)
    d = lib().orc_dist(_p(qz, _f32p), _p(raw, _f32p), int(start), qz.size, float(mean), float(sd), float(bound),
                       _p(order, _i32p), ctypes.byref(blocks))
    return np.float32(d), blocks.value


def dist_all(qz, raw, mean=None, sd=None, threads: int = 0) -> np.ndarray:
    """Full reference distances of one z-normalised template against every window."""
    qz = np.ascontiguousarray(qz, dtype=np.float32)
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    m = qz.size
    if mean is None:
        mean, sd = window_stats(raw, m, threads)
    out = np.zeros(raw.size - m + 1, dtype=np.float32)
    lib().orc_dist_all(_p(qz, _f32p), m, _p(raw, _f32p), raw.size, _p(mean, _f32p), _p(sd, _f32p),
                       threads or os.cpu_count(), _p(out, _f32p))
    return out


def select_knn(pos, d, k: int, merge: int):
    pos = np.ascontiguousarray(pos, dtype=np.uint64)
    d = np.ascontiguousarray(d, dtype=np.float32)
    op = np.zeros(k, dtype=np.uint64)
    od = np.zeros(k, dtype=np.float32)
    n = lib().orc_select_knn(_p(pos, _u64p), _p(d, _f32p), pos.size, k, merge, _p(op, _u64p), _p(od, _f32p))
    return op[:n], od[:n]


class _Bsf(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("pos", _i64p), ("knn", _f32p)]


def bsf_run(k: int, merge: int, ds, ps):
    """Feeds (d, pos) pairs through the bound set; returns (knn, pos)."""
    L = lib()
    L.orc_bsf_init.argtypes = [ctypes.POINTER(_Bsf), ctypes.c_int]
    L.orc_bsf_free.argtypes = [ctypes.POINTER(_Bsf)]
    L.orc_bsf_insert_offset.argtypes = [ctypes.POINTER(_Bsf), ctypes.c_float, ctypes.c_int64, ctypes.c_int]
    q = _Bsf()
    L.orc_bsf_init(ctypes.byref(q), k)
    for d, p in zip(ds, ps):
        L.orc_bsf_insert_offset(ctypes.byref(q), float(d), int(p), merge)
    knn = np.array([q.knn[i] for i in range(k)], dtype=np.float32)
    pos = np.array([q.pos[i] for i in range(k)], dtype=np.int64)
    L.orc_bsf_free(ctypes.byref(q))
    return knn, pos


def pack_templates(templates):
    """List of 1-D templates -> (flat float32, lengths int32, offsets int64)."""
    lengths = np.array([len(t) for t in templates], dtype=np.int32)
    offsets = np.zeros(len(templates), dtype=np.int64)
    if len(templates) > 1:
        offsets[1:] = np.cumsum(lengths[:-1].astype(np.int64))
    flat = np.concatenate([np.asarray(t, dtype=np.float32) for t in templates]) if templates else \
        np.zeros(0, np.float32)
    return np.ascontiguousarray(flat), lengths, offsets


def scan_knn(raw, templates, k: int, merge: int = 0, threads: int = 0, segments: int = 0):
    """SkipSequentialScan::search_knn (one rank; segments > 0: the
    skip-sequential method with that many PAA segments).  Returns (positions
    [T,k] uint64, distances [T,k] float32, seconds [T])."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    flat, lengths, offsets = pack_templates(templates)
    T = len(templates)
    op = np.zeros((T, k), dtype=np.uint64)
    od = np.zeros((T, k), dtype=np.float32)
    sec = np.zeros(T, dtype=np.float64)
    lib().orc_scan_knn_lb(_p(raw, _f32p), raw.size, _p(flat, _f32p), _p(lengths, _i32p), _p(offsets, _i64p), T, k,
                          merge, threads or os.cpu_count(), segments, _p(op, _u64p), _p(od, _f32p), _p(sec, _f64p))
    return op, od, sec


def scan_threshold(raw, templates, rho: float, merge: int = 0, threads: int = 0, segments: int = 0):
    """SkipSequentialScan::search_threshold (one rank; segments > 0: the
    skip-sequential method).  Returns (template ids int32, positions uint64,
    distances float32)."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    flat, lengths, offsets = pack_templates(templates)
    T = len(templates)
    cap = 1 << 16
    while True:
        ot = np.zeros(cap, dtype=np.int32)
        op = np.zeros(cap, dtype=np.uint64)
        od = np.zeros(cap, dtype=np.float32)
        rows = lib().orc_scan_threshold_lb(_p(raw, _f32p), raw.size, _p(flat, _f32p), _p(lengths, _i32p),
                                           _p(offsets, _i64p), T, float(rho), merge, threads or os.cpu_count(),
                                           segments, cap, _p(ot, _i32p), _p(op, _u64p), _p(od, _f32p))
        if rows <= cap:
            return ot[:rows], op[:rows], od[:rows]
        cap = int(rows)


# ----------------------------------------------------------- iSAX bounds

_u8 = ctypes.POINTER(ctypes.c_uint8)


def sax_breakpoints() -> np.ndarray:
    """The 255 breakpoints of cardinality 256 (float32)."""
    bp = np.zeros(255, dtype=np.float32)
    lib().orc_sax_breakpoints(_p(bp, _f32p))
    return bp


def paa(z, segments: int) -> np.ndarray:
    """paa_from_ts (SAX.cpp:73-85) of a z-normalised template."""
    z = np.ascontiguousarray(z, dtype=np.float32)
    out = np.zeros(segments, dtype=np.float32)
    lib().orc_paa(_p(z, _f32p), z.size, segments, _p(out, _f32p))
    return out


def sax_words(raw, m: int, segments: int, threads: int = 0) -> np.ndarray:
    """The iSAX word of every window (SkipSequential.cpp:155-190 with lower
    bounds): uint8 [W, segments]."""
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    W = raw.size - m + 1
    out = np.zeros((max(W, 0), segments), dtype=np.uint8)
    lib().orc_sax_words(_p(raw, _f32p), raw.size, m, segments, threads or os.cpu_count(), _p(out, _u8))
    return out


def lower_bounds(paa_q, words, m: int) -> np.ndarray:
    """SkipSequentialScan::lower_bound (SkipSequential.cpp:200-211) of every
    word against the template's PAA."""
    paa_q = np.ascontiguousarray(paa_q, dtype=np.float32)
    words = np.ascontiguousarray(words, dtype=np.uint8)
    bp = sax_breakpoints()
    out = np.zeros(words.shape[0], dtype=np.float32)
    lib().orc_lower_bounds(_p(bp, _f32p), _p(paa_q, _f32p), _p(words, _u8), words.shape[0], m, words.shape[1],
                           _p(out, _f32p))
    return out


# ------------------------------------------------------------ iSAX index

IX_PLANES = 4      # symbol bits per segment in the key
IX_ROOT_DEPTH = 16  # one bit per segment: the reference's 2^16 root nodes


def isax_keys(words) -> np.ndarray:
    """Interleaved keys of 16-segment words (uint8 [W, 16]): bit plane p (0 =
    most significant) of segment s's symbol at key bit 63 - (16 p + s), p < 4."""
    w = np.asarray(words, dtype=np.uint8).astype(np.uint64)
    key = np.zeros(w.shape[0], dtype=np.uint64)
    for p in range(IX_PLANES):
        for s in range(16):
            key |= ((w[:, s] >> np.uint64(7 - p)) & np.uint64(1)) << np.uint64(63 - (16 * p + s))
    return key


def isax_index(words, leaf_size: int):
    """The index build restated (indexing.cpp:24-281's tree, split in turn
    over the segments): window ids stably sorted by key, then the leaves top
    down - the root level holds the runs of equal top-16 key bits, and a node
    of more than leaf_size windows splits on its next key bit until 64 bits
    are known.  Returns (ids, leaves: int64 rows (start, count, depth) in key
    order)."""
    keys = isax_keys(words)
    ids = np.argsort(keys, kind="stable").astype(np.uint32)
    sk = keys[ids]
    top = sk >> np.uint64(64 - IX_ROOT_DEPTH)
    starts = np.flatnonzero(np.r_[True, top[1:] != top[:-1]]) if sk.size else np.zeros(0, dtype=np.int64)
    ends = np.r_[starts[1:], sk.size]
    stack = [(int(a), int(b - a), IX_ROOT_DEPTH) for a, b in zip(starts, ends)]
    leaves = []
    while stack:
        start, count, depth = stack.pop()
        if not (depth < 64 and count > leaf_size):
            leaves.append((start, count, depth))
            continue
        bit = np.uint64(1) << np.uint64(63 - depth)
        lo = start + int(np.searchsorted((sk[start:start + count] & bit) != 0, True))
        if lo > start:
            stack.append((start, lo - start, depth + 1))
        if lo < start + count:
            stack.append((lo, start + count - lo, depth + 1))
    leaves.sort()
    return ids, np.array(leaves, dtype=np.int64).reshape(-1, 3)


def isax_leaf_bound(paa_q, key: int, depth: int, m: int) -> np.float32:
    """minidist_paa_to_isax (SAX.cpp:112-170) of a leaf: segment s known to
    b = ceil((depth - s) / 16) bits (at most 4) spans the symbols [v 2^(8-b),
    (v + 1) 2^(8-b) - 1] of its top bits v; the PAA value's distance from
    their breakpoint range (MINVAL / MAXVAL = -/+2e6 at the ends), squared in
    FP32 and summed in segment order, times m / 16."""
    bp = sax_breakpoints()
    f = np.float32
    dist = f(0.0)
    for s in range(16):
        b = min(IX_PLANES, (depth - s + 15) // 16) if depth > s else 0
        v = 0
        for p in range(b):
            v = (v << 1) | ((int(key) >> (63 - (16 * p + s))) & 1)
        lo_sym = v << (8 - b)
        hi_sym = lo_sym + (1 << (8 - b)) - 1
        lo = f(-2000000.0) if lo_sym == 0 else bp[lo_sym - 1]
        hi = f(2000000.0) if hi_sym == 255 else bp[hi_sym]
        q = f(paa_q[s])
        d = f(lo - q) if lo > q else (f(hi - q) if hi < q else f(0.0))
        dist = f(dist + f(d * d))
    return f(f(f(m) / f(16.0)) * dist)


# ------------------------------------------------------------- FP64 oracle

def pearson_fp64(raw, template) -> np.ndarray:
    """FP64 Pearson correlation of the template with every window
    (validate_search.py:80-86); constant windows are -inf."""
    x = np.asarray(raw, dtype=np.float64)
    m = len(template)
    w = np.lib.stride_tricks.sliding_window_view(x, m)
    mu = w.mean(axis=1)
    sd = w.std(axis=1)
    q = np.asarray(template, dtype=np.float64)
    qn = (q - q.mean()) / q.std()
    with np.errstate(invalid="ignore", divide="ignore"):
        c = (w @ qn - mu * qn.sum()) / (m * sd)
    c[sd <= 1e-12] = -np.inf
    return c

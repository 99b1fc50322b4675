"""Writes tests/golden/ref_vectors.npz from the reference's own hot-path
functions, compiled from the reference sources into oracle/_ref
(`make -C oracle ref`).  The fixture pins the oracle where the reference
sources are absent (the GPU host, other machines).

Cases: template z-norm and block order (ZnormSearch.hpp:33-90), the
distance with and without early abandoning (ZnormSearch.hpp:103-165) for
lengths around the 32-sample block and 8-lane boundaries, missing samples,
loud and near-constant windows; canonical kNN selection with exclusion
(ZnormSearch.hpp:170-189); bound-set insertion sequences
(iSAXPqueue.cpp:471-549).
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libtremor_ref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20260917)
    lens = [1, 7, 8, 9, 15, 16, 17, 31, 32, 33, 47, 48, 63, 64, 65, 100, 256, 257, 1000, 1023, 4096]
    reps = {4096: 3, 1000: 6, 1023: 4}
    q_raw, w_raw, meta, zbits, order_flat, order_len, dbits, blocks = [], [], [], [], [], [], [], []
    for m in lens:
        for r in range(reps.get(m, 10)):
            kind = r % 5
            t = rng.standard_normal(m).astype(np.float32) * np.float32(10.0 ** rng.integers(-3, 4))
            t += np.float32(rng.standard_normal() * 3)
            if kind == 4 and m > 2:
                t[rng.integers(0, m)] = O.MISSING
            w = rng.standard_normal(m + 4).astype(np.float32)
            if kind == 1:
                w = w * np.float32(1e6) + np.float32(1e4)
            elif kind == 2:
                w = np.float32(5.0) + w * np.float32(1e-9)
            elif kind == 3:
                w[rng.integers(1, m + 4, size=2)] = O.MISSING
            start = int(rng.integers(0, 4))
            z = t.copy()
            R.ref_znormalize(O._p(z, O._f32p), m)
            order = np.zeros(m // 32 + 1, np.int32)
            nb = R.ref_block_order(O._p(z, O._f32p), m, O._p(order, O._i32p))
            mean, sd = O.window_stats(w[start:start + m], m)
            for bound in (float(O.FLT_MAX), float(rng.uniform(0, 2 * m))):
                bl = ctypes.c_int32(0)
                d = R.ref_dist(O._p(z, O._f32p), O._p(w, O._f32p), start, m, float(mean[0]), float(sd[0]), bound,
                               O._p(order, O._i32p), ctypes.byref(bl))
                q_raw.append(t)
                w_raw.append(w)
                meta.append((m, start, np.float32(mean[0]).view(np.uint32), np.float32(sd[0]).view(np.uint32),
                             np.float32(bound).view(np.uint32)))
                zbits.append(z.view(np.uint32))
                order_flat.append(order[:nb])
                order_len.append(nb)
                dbits.append(np.float32(d).view(np.uint32))
                blocks.append(bl.value)
    # canonical selection
    sel_in_pos, sel_in_d, sel_meta, sel_out_pos, sel_out_d = [], [], [], [], []
    for r in range(40):
        cnt = int(rng.integers(0, 300))
        pos = rng.choice(5000, size=cnt, replace=False).astype(np.uint64)
        d = rng.integers(0, 40, size=cnt).astype(np.float32) * np.float32(0.25)  # many exact ties
        k = int(rng.integers(1, 20))
        merge = int(rng.choice([0, 1, 5, 50]))
        op = np.zeros(k, np.uint64)
        od = np.zeros(k, np.float32)
        kept = R.ref_select_knn(O._p(pos, O._u64p), O._p(d, O._f32p), cnt, k, merge, O._p(op, O._u64p),
                                O._p(od, O._f32p))
        sel_in_pos.append(pos)
        sel_in_d.append(d)
        sel_meta.append((cnt, k, merge, kept))
        sel_out_pos.append(op[:kept])
        sel_out_d.append(od[:kept])
    # bound-set insertion sequences
    bsf_meta, bsf_in_d, bsf_in_p, bsf_out_d, bsf_out_p = [], [], [], [], []
    for r in range(30):
        k = int(rng.integers(1, 12))
        merge = int(rng.choice([0, 3, 20]))
        steps = int(rng.integers(1, 200))
        ds = rng.integers(0, 30, size=steps).astype(np.float32)
        ps = rng.integers(0, 400, size=steps).astype(np.uint64)
        q = R.ref_bsf_init(k)
        for dd, pp in zip(ds, ps):
            R.ref_bsf_insert_offset(q, float(dd), int(pp), merge)
        knn = np.zeros(k, np.float32)
        pos = np.zeros(k, np.int64)
        R.ref_bsf_read(q, O._p(knn, O._f32p), pos.ctypes.data_as(O._i64p))
        R.ref_bsf_destroy(q)
        bsf_meta.append((k, merge, steps))
        bsf_in_d.append(ds)
        bsf_in_p.append(ps)
        bsf_out_d.append(knn)
        bsf_out_p.append(pos)

    def cat(xs, dt):
        return np.concatenate([np.asarray(x, dtype=dt).ravel() for x in xs]) if xs else np.zeros(0, dt)

    np.savez_compressed(
        OUT,
        dist_meta=np.array(meta, dtype=np.int64),
        q_raw=cat(q_raw, np.float32), w_raw=cat(w_raw, np.float32), zbits=cat(zbits, np.uint32),
        order=cat(order_flat, np.int32), order_len=np.array(order_len, np.int64),
        dbits=np.array(dbits, np.uint32), blocks=np.array(blocks, np.int64),
        sel_meta=np.array(sel_meta, np.int64), sel_in_pos=cat(sel_in_pos, np.uint64),
        sel_in_d=cat(sel_in_d, np.float32), sel_out_pos=cat(sel_out_pos, np.uint64),
        sel_out_d=cat(sel_out_d, np.float32),
        bsf_meta=np.array(bsf_meta, np.int64), bsf_in_d=cat(bsf_in_d, np.float32), bsf_in_p=cat(bsf_in_p, np.uint64),
        bsf_out_d=cat(bsf_out_d, np.float32), bsf_out_p=cat(bsf_out_p, np.int64),
    )
    print(OUT, os.path.getsize(OUT), "bytes", len(dbits), "distance cases")


if __name__ == "__main__":
    main()

"""The oracle's restatement of the iSAX index (oracle.isax_index,
oracle.isax_leaf_bound), pinned on the CPU by what fixes it independently:

- the leaves partition the key order, each holds windows sharing the top
  `depth` key bits, at most leaf_size of them unless all 64 bits are known,
  and no leaf's parent could have stopped splitting;
- a brute-force build that groups the keys by every prefix length gives the
  same leaves;
- a leaf's region bound never exceeds the bound of any of its windows (the
  reference's SkipSequentialScan::lower_bound, pinned by the skip-sequential
  fixtures), and is 0 when the PAA lies inside every segment's range.

The device build is compared with this restatement in test_gpu_index.py.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def words_case(oracle):
    from tremor_amd import synth
    rng = np.random.default_rng(2024)
    x = synth.bandpassed_noise(1 << 15, rng)
    synth.inject_events(x, 12, rng)
    m = 64
    return x, m, oracle.sax_words(x, m, 16)


def _brute_leaves(keys_sorted, leaf_size):
    """Leaves by prefix groups: the runs of equal top-d bits for d = 16 ... 64;
    a run is a leaf when it holds at most leaf_size windows (or d = 64) and
    its parent run (top d - 1 bits, or none at d = 16) held more."""
    n = keys_sorted.size
    leaves = []
    parent_big = np.ones(n, dtype=bool)  # at d = 16 every window's parent is the (unsplittable) root
    for d in range(16, 65):
        pref = keys_sorted >> np.uint64(64 - d) if d < 64 else keys_sorted
        starts = np.flatnonzero(np.r_[True, pref[1:] != pref[:-1]])
        ends = np.r_[starts[1:], n]
        big = np.zeros(n, dtype=bool)
        for a, b in zip(starts, ends):
            if not parent_big[a]:
                continue
            if b - a <= leaf_size or d == 64:
                leaves.append((int(a), int(b - a), d))
            else:
                big[a:b] = True
        parent_big = big
        if not big.any():
            break
    return sorted(leaves)


@pytest.mark.parametrize("leaf_size", [40, 300, 2000])
def test_leaves_partition_and_match_brute_force(oracle, words_case, leaf_size):
    x, m, words = words_case
    ids, leaves = oracle.isax_index(words, leaf_size)
    keys = oracle.isax_keys(words)
    sk = keys[ids]
    assert np.array_equal(np.sort(ids), np.arange(words.shape[0], dtype=np.uint32))
    assert np.all(sk[1:] >= sk[:-1])
    pos = 0
    for start, count, depth in leaves:
        assert start == pos and count >= 1 and 16 <= depth <= 64
        assert count <= leaf_size or depth == 64
        seg = sk[start:start + count]
        if depth < 64:
            assert np.all(seg >> np.uint64(64 - depth) == seg[0] >> np.uint64(64 - depth))
        pos += count
    assert pos == words.shape[0]
    assert [tuple(r) for r in leaves.tolist()] == _brute_leaves(sk, leaf_size)


def test_keys_interleave_the_top_bits(oracle):
    rng = np.random.default_rng(5)
    w = rng.integers(0, 256, size=(200, 16), dtype=np.uint8)
    keys = oracle.isax_keys(w)
    for i in range(w.shape[0]):
        want = 0
        for p in range(4):
            for s in range(16):
                want |= ((int(w[i, s]) >> (7 - p)) & 1) << (63 - (16 * p + s))
        assert int(keys[i]) == want
    # The key orders the words by their top bit planes first.
    assert int(oracle.isax_keys(np.full((1, 16), 0x80, dtype=np.uint8))[0]) == 0xFFFF << 48


def test_leaf_bound_below_every_window_bound(oracle, words_case):
    x, m, words = words_case
    ids, leaves = oracle.isax_index(words, 300)
    keys = oracle.isax_keys(words)
    rng = np.random.default_rng(8)
    for trial in range(3):
        j = int(rng.integers(0, x.size - m))
        q = x[j:j + m] + rng.normal(0.0, 0.3, m).astype(np.float32)
        paa_q = oracle.paa(oracle.znormalize(q), 16)
        wlb = oracle.lower_bounds(paa_q, words, m)
        for start, count, depth in leaves[:: max(1, len(leaves) // 400)]:
            members = ids[start:start + count]
            lb = oracle.isax_leaf_bound(paa_q, int(keys[members[0]]), int(depth), m)
            # Regions nest, so every gap is at most the window's; the sums
            # differ only in order (within a few ulps).
            assert lb <= wlb[members].min() * (1.0 + 1e-6) + 1e-30


def test_leaf_bound_closed_forms(oracle):
    bp = oracle.sax_breakpoints()
    m = 160
    # PAA inside every range: 0.
    assert oracle.isax_leaf_bound(np.zeros(16, dtype=np.float32), 0, 16, m) == 0.0
    # Every segment known to 1 bit (depth 16), key 0 (lower half): a PAA value
    # of 1 in every segment is 1 - bp[127] = 1 above the range's top, so the
    # bound is m / 16 * 16 * 1^2 = m.
    assert bp[127] == 0.0
    assert oracle.isax_leaf_bound(np.ones(16, dtype=np.float32), 0, 16, m) == np.float32(m)
    # Depth 17: segment 0 known to 2 bits ([0, 63], top bp[63]); the others to
    # one bit.
    paa_q = np.zeros(16, dtype=np.float32)
    paa_q[0] = 1.0
    want = np.float32(np.float32(m) / np.float32(16)) * np.float32((np.float32(1.0) - bp[63]) ** 2)
    assert oracle.isax_leaf_bound(paa_q, 0, 17, m) == want

"""Pins of the CPU oracle: the reference's own compiled bloom sizing, the
reference source's constants, textbook results (exact kNN, sorting,
FP64 distances), brute force on small inputs and structural invariants."""
import itertools
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import assert_same

REF = "/root/reference"
REF_BLOOM = os.path.join(os.path.dirname(O.__file__), "_ref", "bloom_params")


# ---- A13: bloom sizing --------------------------------------------------
def test_hash_parameter_matches_reference_binary():
    """orc_hash_parameter against auto_tune_bloom.h compiled where it lies."""
    if not os.path.exists(REF_BLOOM):
        pytest.skip("oracle/_ref/bloom_params not built (reference sources absent)")
    out = subprocess.run([REF_BLOOM], capture_output=True, text=True, check=True).stdout
    rows = [tuple(map(int, l.split())) for l in out.strip().splitlines()]
    assert len(rows) == 4096
    for nc, hl, bit, hh in rows:
        assert O.hash_parameter(nc) == (hl, bit, hh), nc


def test_hash_parameter_closed_form():
    # p = (1 - e^{-k n/m})^k with n = 12.897 nc: the first m = 32*2^i under 3%
    for nc in (32, 64, 128, 256):
        hl, bit, hh = O.hash_parameter(nc)
        n = 12.897 * nc
        m = 32 * hl
        k = round(np.log(2) * m / n)
        assert hh == k == 7
        assert (1 - np.exp(-k * n / m)) ** k < 0.03
        m_half = m / 2
        k2 = round(np.log(2) * m_half / n)
        assert (1 - np.exp(-k2 * n / m_half)) ** k2 >= 0.03
    assert [O.hash_parameter(n)[0] for n in (32, 64, 128, 256)] == [128, 256, 512, 1024]


# ---- A3: bloom hash -----------------------------------------------------
def test_hash_constants_match_reference_source():
    src = os.path.join(REF, "graph_search.cuh")
    if not os.path.exists(src):
        pytest.skip("reference sources absent")
    text = open(src).read()
    block = text[text.index("random_number[10 * 2]"):]
    block = block[: block.index("};")]
    consts = [int(x, 16) for x in re.findall(r"0x([0-9a-fA-F]{8})U", block)]
    assert len(consts) == 20
    # hash h applies constants 2h and 2h+1: a zero input word mixes to 0, an
    # input of 1 exercises the first multiplier only through the shifts
    for h in range(10):
        x = 0x12345678
        y = x ^ (x >> 16)
        y = (y * consts[2 * h]) & 0xFFFFFFFF
        y ^= y >> 13
        y = (y * consts[2 * h + 1]) & 0xFFFFFFFF
        y ^= y >> 16
        assert O.hash_raw(h, x) == y
    assert O.hash_raw(0, 0) == 0


def test_bloom_false_positive_rate_matches_theory():
    """With n ids added to m bits and k hashes, the false-positive rate of
    fresh ids must track (1 - e^{-kn/m})^k (a weak or mis-indexed mix
    breaks this)."""
    words, k = 512, 3
    m = words * 32
    for n in (800, 2000):
        bits = np.zeros(m, bool)
        ids = np.arange(n) * 7919 + 13
        for x in ids:
            for h in range(k):
                bits[O.hash_raw(h, int(x)) % m] = True
        fresh = np.arange(20000) * 104729 + 5_000_000
        fp = np.mean([all(bits[O.hash_raw(h, int(x)) % m] for h in range(k)) for x in fresh])
        theory = (1 - np.exp(-k * n / m)) ** k
        assert abs(fp - theory) < 0.25 * theory + 0.002, (n, fp, theory)


# ---- A2: distance -------------------------------------------------------
def test_distance_fp64_and_integer_exact():
    rng = np.random.default_rng(0)
    for d in (3, 32, 64, 96, 100, 128, 960):
        p = rng.normal(size=d).astype(np.float32) * 10
        q = rng.normal(size=d).astype(np.float32) * 10
        ref = float(((p.astype(np.float64) - q) ** 2).sum())
        got = float(O.distance(p, q))
        assert abs(got - ref) <= 2 * d * 2.0 ** -24 * ref
        pi = rng.integers(0, 256, size=d).astype(np.float32)
        qi = rng.integers(0, 256, size=d).astype(np.float32)
        exact = int(((pi.astype(np.int64) - qi.astype(np.int64)) ** 2).sum())
        assert float(O.distance(pi, qi)) == float(exact)


def test_distance_association_order():
    """A rounding case that separates the lane/tree association from a
    left-to-right sum: with x0 = 2^26 and x8 = x16 = x24 = 4 the tree forms
    (x0 + x16) + (x8 + x24) = 2^26 + 8, the straight sum stays at 2^26."""
    d = 32
    p = np.zeros(d, np.float32)
    q = np.zeros(d, np.float32)
    p[0] = 2.0 ** 13
    p[8] = p[16] = p[24] = 2.0
    assert float(O.distance(p, q)) == 2.0 ** 26 + 8
    seq = np.float32(0)
    for x in (p * p):
        seq = np.float32(seq + x)
    assert float(seq) == 2.0 ** 26


# ---- A4/A5: networks ----------------------------------------------------
@pytest.mark.parametrize("m", [2, 4, 8])
def test_sort_desc_all_permutations(m):
    for perm in itertools.permutations(range(m)):
        d = np.array(perm, np.float32)
        dd, ii = O.sort_desc(d, np.arange(m, dtype=np.int32))
        assert list(dd) == sorted(perm, reverse=True)
        assert sorted(ii) == list(range(m))
        assert all(d[ii] == dd)


def test_sort_desc_larger_random_and_ties():
    rng = np.random.default_rng(1)
    for m in (16, 32, 64, 128, 256):
        for _ in range(20):
            d = rng.integers(0, 8, size=m).astype(np.float32)  # many ties
            ids = np.arange(m, dtype=np.int32)
            dd, ii = O.sort_desc(d, ids, O.MODE_REF)
            assert (np.diff(dd) <= 0).all()
            assert all(d[ii] == dd)
            dt, it = O.sort_desc(d, ids, O.MODE_TUNED)
            order = np.lexsort((-ids, -d))  # dist desc, id desc
            assert_same(it, ids[order], "tuned sort")


@pytest.mark.parametrize("n", [4, 8, 16, 64, 256])
def test_merge_asc_sorts_bitonic(n):
    rng = np.random.default_rng(n)
    for _ in range(50):
        a = np.sort(rng.normal(size=n // 2 + rng.integers(0, n // 2)))
        b = np.sort(rng.normal(size=n - a.size))[::-1]
        seq = np.concatenate([a, b]).astype(np.float32)
        rot = rng.integers(0, n)
        seq = np.roll(seq, rot)  # cyclic shifts of asc-desc are bitonic
        dd, ii = O.merge_asc(seq, np.arange(n, dtype=np.int32))
        assert (np.diff(dd) >= 0).all()
        assert all(seq[ii] == dd)


def test_pair_index_formula_matches_reference_bitonic_print():
    """test/bitonic.cpp prints (i, i+sub) for i=(i/sub)*2*sub + (i&(sub-1))
    over n = 8: every substep must touch each index exactly once."""
    n = 8
    for sub in (4, 2, 1):
        us = [(i // sub) * 2 * sub + (i & (sub - 1)) for i in range(n // 2)]
        touched = sorted(us + [u + sub for u in us])
        assert touched == list(range(n))


# ---- A7: kd-tree, leaves, capacity --------------------------------------
def _ref_split(vals, xm):
    """kdtree.cpp:44-62 two-pointer split, transliterated, to check that it
    is a threshold partition (the oracle uses a stable partition)."""
    a = list(vals)
    l, r = 0, len(a) - 1
    while l < r:
        if a[l] >= xm:
            while a[r] >= xm and l < r:
                r -= 1
            if l < r:
                a[l], a[r] = a[r], a[l]
                r -= 1
            else:
                break
        l += 1
    if a[l] < xm:
        l += 1
    return a, l


def test_reference_split_is_threshold_partition():
    rng = np.random.default_rng(2)
    for _ in range(3000):
        n = int(rng.integers(2, 20))
        vals = rng.integers(0, 6, size=n).astype(np.float32)
        xm = float(rng.integers(0, 7)) - 0.5 * rng.integers(0, 2)
        a, l = _ref_split(vals, xm)
        assert l == int((vals < xm).sum())
        assert all(v < xm for v in a[:l]) and all(v >= xm for v in a[l:])


def test_find_xm_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(300):
        n = int(rng.integers(3, 200))
        v = rng.normal(size=n).astype(np.float32)
        lo, hi = float(v.min()), float(v.max())
        xm = O.find_xm(v, lo, hi)
        assert (v <= xm).sum() <= n // 2 or xm == np.float32((np.float32(lo) + np.float32(hi)) / 2)
        assert lo <= xm <= hi


def test_kdtree_invariants(small_workload):
    w = small_workload
    pts = w.base[:, :3]
    boxes, belong = O.kdtree_build(pts, w.cap)
    n_leaf = boxes.shape[0]
    assert (belong >= 0).all()
    counts = np.bincount(belong, minlength=n_leaf)
    assert counts.sum() == w.n and counts.max() <= w.cap and counts.min() > 0
    for i in range(n_leaf):
        p = pts[belong == i]
        assert_same(boxes[i, :3], p.min(0), "tight min")
        assert_same(boxes[i, 3:], p.max(0), "tight max")
    pids = O.aabb_pids(belong, n_leaf, w.cap)
    for i in range(n_leaf):
        mem = np.flatnonzero(belong == i)
        assert_same(pids[i, : mem.size], mem, "pids ascending")
        assert (pids[i, mem.size:] == w.n).all()


def test_kdtree_dfs_order_is_spatially_nested(small_workload):
    """Left-first DFS: splitting the leaf range in two at the first split
    separates the points by the root split value on one axis."""
    w = small_workload
    pts = w.base[:, :3]
    boxes, belong = O.kdtree_build(pts, w.cap)
    # the root split puts leaves [0, a) left of leaves [a, n_leaf) on one axis
    n_leaf = boxes.shape[0]
    ok = False
    for a in range(1, n_leaf):
        for ax in range(3):
            if boxes[:a, 3 + ax].max() <= boxes[a:, ax].min():
                ok = True
    assert ok


def test_kdtree_split_failure_reported():
    pts = np.zeros((100, 3), np.float32)  # every point identical: no split
    with pytest.raises(RuntimeError):
        O.kdtree_build(pts, 10)


def test_leaf_capacity_float32():
    assert O.leaf_capacity(1_000_000, 0.000256) == int(np.float32(1e6) * np.float32(0.000256))
    assert O.leaf_capacity(1_000_000, 0.000256) == 256
    assert O.leaf_capacity(100_000, 0.00128) == 128
    assert O.leaf_capacity(10_000_000, 0.0001024) == 1024
    assert O.leaf_capacity(1_000_000, 0.000128) == 128  # flags.cpp default point_ratio


def test_expand_aabbs():
    b = np.array([[0, 1, 2, 10, 3, 2]], np.float32)
    e = O.expand_aabbs(b, 0.2)
    assert_same(e, np.array([[-2, 0.6, 2, 12, 3.4, 2]], np.float32), "expand")


# ---- A6: entry rule -----------------------------------------------------
def test_select_entry_brute_force():
    rng = np.random.default_rng(4)
    boxes = rng.uniform(0, 10, size=(50, 3)).astype(np.float32)
    boxes = np.concatenate([boxes, boxes + rng.uniform(0, 4, size=(50, 3)).astype(np.float32)], 1)
    q = rng.uniform(-1, 15, size=(500, 5)).astype(np.float32)
    h = O.select_entry(boxes, q)
    for i in range(500):
        inside = [(q[i, :3] >= boxes[j, :3]).all() and (q[i, :3] <= boxes[j, 3:]).all()
                  for j in range(50)]
        exp = inside.index(True) if any(inside) else 0
        assert h[i] == exp


# ---- A8: projection -----------------------------------------------------
def test_project_f64_closed_form():
    rng = np.random.default_rng(5)
    Q = rng.normal(size=(20, 16)).astype(np.float32)
    R = rng.normal(size=(16, 16)).astype(np.float32)
    mu = rng.normal(size=16).astype(np.float32)
    got = O.project_f64(Q, R, mu, 8)
    ref = (Q.astype(np.float64) - mu) @ R[:, :8].astype(np.float64)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


# ---- A12: recall ----------------------------------------------------------
def test_recall_closed_form():
    gt = np.array([[1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 99]], np.int32)
    ids = np.array([[10, 2, 50, 51, 52, 53, 54, 55, 56, 3]], np.int32)
    assert O.recall(ids, gt, 10) == pytest.approx(0.3)
    assert O.recall(ids, gt[:, :1], 1) == 0.0


# ---- A1: end-to-end search against textbook kNN ---------------------------
def _complete_graph(n):
    D = 1 << (n - 1).bit_length()
    g = np.full((n, D), n, np.int32)
    for u in range(n):
        g[u, : n - 1] = [v for v in range(n) if v != u]
    return g


@pytest.mark.parametrize("mode", [0, 1])
def test_search_complete_graph_is_exact_knn(mode):
    """Complete graph, one entry point: the first expansion sees every
    point, so a beam >= k must hold the exact k nearest (distinct dists)."""
    rng = np.random.default_rng(6)
    n, d = 31, 40
    X = rng.normal(size=(n, d)).astype(np.float32)
    Q = rng.normal(size=(25, d)).astype(np.float32)
    g = _complete_graph(n)
    pids = np.array([[0]], np.int32)
    hits = np.zeros(Q.shape[0], np.int32)
    p = O.make_params(k=10, beam=32, mode=mode)
    ids, dists = O.search(X, g, Q, pids, hits, p)
    d2 = ((Q[:, None, :].astype(np.float64) - X[None]) ** 2).sum(-1)
    exact = np.argsort(d2, axis=1, kind="stable")[:, :10]
    assert_same(ids, exact, "exact knn")


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,width,beam", [(60, 1, 32),   # degree 64 >= n-1: M = 64 > N
                                          (33, 2, 32),   # degree 32, width 2: M = 64 > N
                                          (40, 2, 32),   # degree 64, width 2: M = 128
                                          (31, 1, 16),   # beams below 32 (P2 mask)
                                          (31, 1, 8),
                                          (20, 2, 4)])
def test_search_complete_graph_wide_batch_is_exact_knn(mode, n, width, beam):
    """Batch wider than the beam (C = N < M): the first expansion of the
    single entry point puts every other point into B, and the min-exchange
    must pair L[i] with the i-th of B's N smallest (B[M-N+i]) for the beam to
    hold the exact N nearest.  A pairing with B[i] (B's largest, padding
    included) keeps far points and fails here."""
    rng = np.random.default_rng(16)
    d = 24
    X = rng.normal(size=(n, d)).astype(np.float32)
    Q = rng.normal(size=(30, d)).astype(np.float32)
    g = _complete_graph(n)
    pids = np.array([[0]], np.int32)
    hits = np.zeros(Q.shape[0], np.int32)
    k = min(10, beam)
    p = O.make_params(k=k, beam=beam, width=width, mode=mode)
    ids, dists = O.search(X, g, Q, pids, hits, p)
    d2 = ((Q[:, None, :].astype(np.float64) - X[None]) ** 2).sum(-1)
    exact = np.argsort(d2, axis=1, kind="stable")[:, :k]
    assert_same(ids, exact, "exact knn")


@pytest.mark.parametrize("mode", [0, 1])
def test_merge_batch_wide_batch_keeps_smallest_sorted(mode):
    """Per-iteration invariant of the merge when C = N (batch at least as
    wide as the beam): from any ascending beam and any batch, the merged
    beam is sorted ascending and is exactly the N smallest of L and B (by
    (dist, id); distinct distances so both comparators agree), and the
    pairs leave B with the displaced entries.  Checked for every N, M power
    of two with 1 <= N <= M <= 128, with and without MAX entries."""
    rng = np.random.default_rng(17)
    for lgn in range(0, 8):
        for lgm in range(lgn, 8):
            N, M = 1 << lgn, 1 << lgm
            for trial in range(6):
                vals = rng.permutation(4 * (N + M)).astype(np.float32)[:N + M] + 1
                ids = rng.permutation(10 * (N + M))[:N + M].astype(np.int32)
                ld, li = vals[:N].copy(), ids[:N].copy()
                if trial % 2:
                    ld[N // 2:] = O.MAXD  # part-filled beam
                order = np.argsort(ld, kind="stable")
                ld, li = ld[order], li[order]
                bd, bi = vals[N:].copy(), ids[N:].copy()
                if trial >= 3:
                    bd[rng.random(M) < 0.3] = O.MAXD  # filtered candidates
                skip = trial >= 4
                Ld, Li, Bd, Bi = O.merge_batch(ld, li, bd, bi, mode, skip_max=skip)
                assert (np.diff(Ld) >= 0).all(), (N, M, trial)
                allp = sorted(zip(np.r_[ld, bd].tolist(), np.r_[li, bi].tolist()))
                fin = [x for x in allp if x[0] < O.MAXD][:N]
                got = sorted((a, b) for a, b in zip(Ld.tolist(), Li.tolist()) if a < O.MAXD)
                assert got == fin, (N, M, trial)
                # nothing lost: beam and batch together hold the same pairs
                assert sorted(zip(np.r_[Ld, Bd].tolist(), np.r_[Li, Bi].tolist())) == allp


@pytest.mark.parametrize("mode", [0, 1])
def test_merge_cascade_idempotent_and_batch_skip(mode):
    """What lets the GPU kernel skip a whole batch step: the half-cleaner
    cascade (graph_search.cuh:933-958) is idempotent on any input, bitonic
    or not, ties included, and its output's last slot holds the largest key;
    so a batch whose valid elements are all at or above that key (strictly
    above in the (dist, id) mode) leaves the beam unchanged through sort,
    min-exchange and merge.  Every permutation for N <= 4, 1500 of those of
    N = 8, random inputs with ties up to N = 256."""
    import itertools
    rng = np.random.default_rng(19)

    def key(d, i):
        return (d, abs(i)) if mode == 1 else (d,)

    cases = [np.array(p, np.float32) for n in (2, 4) for p in itertools.permutations(range(n))]
    perms8 = list(itertools.permutations(range(8)))
    cases += [np.array(perms8[i], np.float32) for i in rng.choice(len(perms8), 1500, replace=False)]
    for n in (16, 32, 64, 128, 256):
        cases += [rng.integers(0, 3 * n, n).astype(np.float32) for _ in range(40)]
    for x in cases:
        n = x.size
        ids = rng.permutation(10 * n)[:n].astype(np.int32) * rng.choice([-1, 1], n).astype(np.int32)
        ld, li = O.merge_asc(x, ids, mode)
        d2, i2 = O.merge_asc(ld, li, mode)
        assert_same(d2, ld, "idempotent dists")
        assert_same(i2, li, "idempotent ids")
        top = max(key(a, b) for a, b in zip(ld.tolist(), li.tolist()))
        assert key(float(ld[-1]), int(li[-1])) == top, "last slot holds the largest key"
        for m in ((n,) if n <= 8 else (n, 32, 64)):
            bd = (ld[-1] + rng.integers(0 if mode == 0 else 1, 50, m)).astype(np.float32)
            bd[rng.random(m) < 0.3] = O.MAXD
            bi = rng.integers(0, 10 * n, m).astype(np.int32)
            for skip in (False, True):
                Ld, Li, _, _ = O.merge_batch(ld, li, bd, bi, mode, skip_max=skip)
                assert_same(Ld, ld, f"batch above the beam n={n} m={m}")
                assert_same(Li, li, f"batch above the beam n={n} m={m}")


def test_search_max_iter_zero_returns_sorted_entries():
    rng = np.random.default_rng(7)
    n, d, cap = 200, 32, 64
    X = rng.normal(size=(n, d)).astype(np.float32)
    Q = rng.normal(size=(10, d)).astype(np.float32)
    g = np.full((n, 32), n, np.int32)
    pids = np.sort(rng.choice(n, size=(3, cap), replace=True), axis=1).astype(np.int32)
    pids = np.array([np.unique(r)[:40].tolist() + [n] * (cap - min(40, np.unique(r).size))
                     for r in pids], np.int32)
    hits = rng.integers(0, 3, size=10).astype(np.int32)
    p = O.make_params(k=16, beam=64, max_iter=0)
    ids, dists = O.search(X, g, Q, pids, hits, p)
    for i in range(10):
        row = pids[hits[i]]
        row = row[row < n]
        dd = np.array([O.distance(X[j], Q[i]) for j in row])
        o = np.argsort(dd, kind="stable")
        assert_same(ids[i, : min(16, row.size)], row[o][:16], "entries")
        assert (ids[i, row.size:] == n).all()


def test_search_set_invariant_on_small_graph(small_workload):
    """Every returned id is a real point, distances are the true ones and
    ascending, no id repeats, and recall is in a sane range."""
    w = small_workload
    p = O.make_params(k=10, beam=64)
    ids, dists, st = O.search(w.base, w.graph, w.queries, w.pids, w.hits, p, with_stats=True)
    assert (ids < w.n).all()
    assert (np.diff(dists, axis=1) >= 0).all()
    for i in range(0, ids.shape[0], 17):
        assert len(set(ids[i])) == 10
        for j in range(10):
            assert dists[i, j] == O.distance(w.base[ids[i, j]], w.queries[i])
    assert O.recall(ids, w.gt) > 0.9
    assert (st[:, 1] <= st[:, 2] * 1).all()  # width 1: one row per iteration


def test_search_rejects_bad_params(small_workload):
    w = small_workload
    with pytest.raises(RuntimeError):
        O.search(w.base, w.graph, w.queries, w.pids, w.hits, O.make_params(k=100, beam=64))


# ---- full-precision rerank (graph_helper.cu:146-213) -------------------

def test_rerank_distance_association():
    """Eight partials over dims l mod 8, summed 0..7: one case whose value
    separates that grouping from a left-to-right sum (2^24 + 14 vs 2^24)."""
    p = np.zeros(16, np.float32)
    q = np.zeros(16, np.float32)
    p[0] = 4096.0                   # partial 0: 2^24
    p[1:8] = 1.0                    # partials 1..7: 1 + 1
    p[9:16] = 1.0
    got = O.rerank_distance(p, q)
    assert got == np.float32(2 ** 24 + 14)
    seq = np.float32(0)
    for x in (p - q) ** 2:
        seq = np.float32(seq + x)
    assert seq == np.float32(2 ** 24) and got != seq


def test_rerank_matches_fp64_order_ties_and_padding():
    rng = np.random.default_rng(4)
    n, d, nq, nc = 400, 96, 30, 64
    base = rng.integers(0, 20, size=(n, d)).astype(np.float32)  # integer data: exact distances
    qs = rng.integers(0, 20, size=(nq, d)).astype(np.float32)
    cand = np.stack([rng.permutation(n)[:nc] for _ in range(nq)]).astype(np.int32)
    cand[0, 5] = cand[0, 6] = n            # padding ids sort last
    base[cand[1, 7]] = base[cand[1, 3]]    # an exact tie: the lower id first
    ids, dists = O.rerank(base, qs, cand, nc)
    ex = ((base[np.minimum(cand, n - 1)].astype(np.float64) - qs[:, None, :]) ** 2).sum(-1)
    ex[cand >= n] = np.inf
    for i in range(nq):
        key = sorted(zip(ex[i], cand[i]))
        assert [c for _, c in key] == ids[i].tolist()
        assert np.array_equal(np.array([e for e, _ in key], np.float32), dists[i])
    assert ids[0, -1] == n and ids[0, -2] == n and np.isinf(dists[0, -1])
    top3, _ = O.rerank(base, qs, cand, 3)
    assert np.array_equal(top3, ids[:, :3])
    with pytest.raises(ValueError):
        O.rerank(base[:, :90], qs[:, :90], cand, 3)  # d % 8 != 0

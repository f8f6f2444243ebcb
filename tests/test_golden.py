"""Golden vectors: `tests/golden/search_small.npz` (written by
`tests/golden/make_golden.py`) holds a small integer-valued search case and
the oracle's kd-tree, entry leaves and search outputs for five settings.

CPU: the oracle still reproduces every recorded vector bit-for-bit, and the
recorded outputs satisfy what the search must return independent of any
implementation (distances recomputed exactly in FP64, ascending (distance,
id) order, distinct in-range ids, every point in exactly one leaf, every
query's entry box containing it).  GPU: the HIP path, through the C ABI,
reproduces the same vectors without running the oracle."""
import os

import numpy as np
import pytest

from conftest import assert_same

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "search_small.npz")


def _load():
    z = np.load(PATH)
    g = {k: z[k] for k in z.files}
    g["base"] = g["base"].astype(np.float32)
    g["queries"] = g["queries"].astype(np.float32)
    return g


def test_golden_oracle_reproduces():
    import oracle as O
    from rt_entry_amd import data as D
    g = _load()
    n = g["base"].shape[0]
    boxes, belong = O.kdtree_build(g["base"][:, :3], g["pids"].shape[1])
    assert_same(boxes.view(np.int32), g["boxes"].view(np.int32), "boxes")
    pids = O.aabb_pids(belong, boxes.shape[0], g["pids"].shape[1])
    assert_same(pids, g["pids"], "pids")
    assert_same(O.expand_aabbs(boxes, 0.2).view(np.int32), g["aabbs"].view(np.int32), "aabbs")
    assert_same(O.select_entry(g["aabbs"], g["queries"]), g["hits"], "hits")
    gpad = D.pad_graph(g["graph"], n)
    for i, (k, beam, width, mi, mode) in enumerate(g["settings"]):
        po = O.make_params(k=int(k), beam=int(beam), width=int(width), max_iter=int(mi),
                           mode=int(mode))
        ids, dists, st = O.search(g["base"], gpad, g["queries"], g["pids"], g["hits"], po,
                                  with_stats=True)
        assert_same(ids, g[f"ids_{i}"], f"ids setting {i}")
        assert_same(dists.view(np.int32), g[f"dists_{i}"].view(np.int32), f"dists setting {i}")
        assert_same(st.astype(np.int32), g[f"stats_{i}"], f"stats setting {i}")


def test_golden_vectors_are_valid_results():
    g = _load()
    base, Q = g["base"].astype(np.float64), g["queries"].astype(np.float64)
    n = base.shape[0]
    # leaves: every point in exactly one leaf, inside its (unexpanded) box
    p = g["pids"][g["pids"] < n]
    assert_same(np.sort(p), np.arange(n), "leaf partition")
    for leaf, row in enumerate(g["pids"]):
        pts = base[row[row < n], :3]
        assert (pts >= g["boxes"][leaf, :3]).all() and (pts <= g["boxes"][leaf, 3:]).all()
    # each entry leaf is the first expanded box containing the query
    lo, hi = g["aabbs"][:, :3], g["aabbs"][:, 3:]
    inside = ((Q[:, None, :3] >= lo[None]) & (Q[:, None, :3] <= hi[None])).all(-1)
    first = np.where(inside.any(1), inside.argmax(1), 0)
    assert_same(g["hits"], first, "first containing box")
    for i, (k, beam, width, mi, mode) in enumerate(g["settings"]):
        ids, d = g[f"ids_{i}"], g[f"dists_{i}"]
        assert ids.shape == (Q.shape[0], k)
        assert ((ids >= 0) & (ids < n)).all()
        assert all(len(set(r)) == k for r in ids)
        exact = ((base[ids] - Q[:, None, :]) ** 2).sum(-1)  # integer data: exact in FP32
        assert_same(d.astype(np.float64), exact, f"distances setting {i}")
        key = d.astype(np.float64) * (n + 1) + ids
        assert (np.diff(key, axis=1) > 0).all(), f"order setting {i}"
    # no search returns a 10th distance below the exact 10th-nearest one
    d_all = ((base[None] - Q[:, None]) ** 2).sum(-1)
    true10 = np.sort(d_all, 1)[:, 9]
    assert (g["dists_0"][:, 9] >= true10).all()


@pytest.mark.gpu
def test_golden_gpu_reproduces(ctx):
    from rt_entry_amd import make_params
    g = _load()
    ctx.set_base(g["base"])
    ctx.set_graph(g["graph"])
    ctx.set_leaves(g["aabbs"], g["pids"])
    print(f"[golden] n={g['base'].shape[0]} nq={g['queries'].shape[0]}", flush=True)
    assert_same(ctx.select_entry(g["queries"]), g["hits"], "hits")
    for i, (k, beam, width, mi, mode) in enumerate(g["settings"]):
        kw = dict(k=int(k), beam=int(beam), width=int(width), max_iter=int(mi), mode=int(mode))
        print(f"[golden] setting {i} {kw}", flush=True)
        ids, dists, st = ctx.search(g["queries"], g["hits"], make_params(**kw), with_stats=True)
        assert_same(ids, g[f"ids_{i}"], f"ids setting {i}")
        assert_same(dists.view(np.int32), g[f"dists_{i}"].view(np.int32), f"dists setting {i}")
        assert_same(st, g[f"stats_{i}"], f"stats setting {i}")
        if width == 1:
            bi, bd = ctx.search_batched(g["queries"], g["hits"], None, None, make_params(**kw))
            assert_same(bi, g[f"ids_{i}"], f"batched ids setting {i}")
            assert_same(bd.view(np.int32), g[f"dists_{i}"].view(np.int32),
                        f"batched dists setting {i}")

"""Writes `search_small.npz`: a small, self-contained search case and the
oracle's outputs for it, the fixed vectors the CPU and GPU golden tests read.

The case: 768 integer-valued 128-d points and 64 queries of the 1M x 128
integer mixture's distribution (values in [0, 255], so every FP32 distance is
an exact integer); the exact 64-NN graph with reverse edges, rank-pruned to
degree 32; a kd-tree over the first three coordinates with leaf capacity 32
(`kdtree.cpp:4-125`), expanded by 0.2 (`aabb.cu:27-38`), and each query's
entry leaf (first containing box, `devicePrograms.cu:29-96`).  The search
outputs (ids, distances, statistics) are recorded for each setting in
`SETTINGS`, with the filter sized as `auto_tune_bloom.h:67-89` / the
reference's fixed 512 words x 3 hashes dictate (`graph_search.cuh:13-22`).

Run from the repository root: `python tests/golden/make_golden.py`.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# (k, beam, width, max_iter, mode): both filter modes, a wide beam, multi-entry
# width, an early cut and k == beam.
SETTINGS = [
    (10, 64, 1, 100, 0),
    (10, 128, 2, 100, 1),
    (10, 32, 4, 100, 0),
    (32, 32, 1, 6, 0),
    (100, 128, 1, 100, 1),
]
N, NQ, DEGREE, K_FWD, CAP, EXPAND = 768, 64, 32, 64, 32, 0.2


def make():
    import oracle as O
    from rt_entry_amd import data as D
    w = D.generate(2, n=N, nq=NQ, n_train=0)
    base, Q = w["base"], w["queries"]
    assert np.array_equal(base, np.round(base)) and base.min() >= 0 and base.max() <= 255
    # stored as uint8: compute from the stored values (the mixture's -0.0 become +0.0)
    base = base.astype(np.uint8).astype(np.float32)
    Q = Q.astype(np.uint8).astype(np.float32)
    knn, _ = D.knn_exact(base, base, K_FWD, exclude_self=True)
    graph = D.graph_from_knn(knn, DEGREE, rev_rank=DEGREE)
    boxes, belong = O.kdtree_build(base[:, :3], CAP)
    pids = O.aabb_pids(belong, boxes.shape[0], CAP)
    aabbs = O.expand_aabbs(boxes, EXPAND)
    hits = O.select_entry(aabbs, Q)
    gpad = D.pad_graph(graph, N)
    out = dict(base=base.astype(np.uint8), queries=Q.astype(np.uint8), graph=graph.astype(np.int32),
               boxes=boxes, pids=pids, aabbs=aabbs, hits=hits,
               settings=np.array(SETTINGS, np.int32))
    for i, (k, beam, width, mi, mode) in enumerate(SETTINGS):
        po = O.make_params(k=k, beam=beam, width=width, max_iter=mi, mode=mode)
        ids, dists, st = O.search(base, gpad, Q, pids, hits, po, with_stats=True)
        out[f"ids_{i}"] = ids
        out[f"dists_{i}"] = dists
        out[f"stats_{i}"] = st.astype(np.int32)
    return out


if __name__ == "__main__":
    out = make()
    path = os.path.join(HERE, "search_small.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")

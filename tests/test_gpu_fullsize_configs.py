"""The other workloads at their full sizes, every offline stage on the GPU:
- 1: 100k x 128 Gaussian mixture, 1k queries, plain degree-32 kNN graph,
  1024 leaves of 128 ids: every query against the oracle;
- 3: 1M x 96 unit vectors, 20k queries with per-query k in {1, 10, 100} and
  beam in {32, 64, 128, 256} in one batched launch: 256 sampled queries
  against the oracle;
- 4: 1M x 960 low-rank mixture, degree-48 graph (row stride 64), 10k
  queries: 128 sampled queries against the oracle, the batched launch
  against rte_search for all.
Each also checks the kd-tree against the oracle's build and every entry
leaf against the oracle's selection; ids and distances are bit-exact."""
import numpy as np
import pytest

import oracle as O
from conftest import assert_same

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


class Cfg:
    pass


@pytest.fixture(scope="module", params=[1, 3, 4], ids=lambda c: f"workload{c}")
def wl(request, ctx):
    import torch
    from rt_entry_amd import Context, expand_aabbs, leaf_capacity
    from rt_entry_amd import data as D
    from rt_entry_amd import gpuprep as G
    cfg = request.param
    c = D.CONFIGS[cfg]
    dev = torch.device("cuda:0")
    w = D.generate(cfg, n_train=0)
    f = Cfg()
    f.cfg, f.n, f.d = cfg, c["n"], c["d"]
    raw_h, Q_h = w["base"], w["queries"]
    assert raw_h.shape == (c["n"], c["d"]) and Q_h.shape == (c["nq"], c["d"])
    print(f"[full] workload {cfg}: generated {raw_h.shape}", flush=True)
    raw = torch.from_numpy(raw_h).to(dev)
    Q = torch.from_numpy(Q_h).to(dev)
    rev = 0 if cfg == 1 else c["degree"]  # workload 1: the plain kNN graph
    with Context(0) as cx:
        f.R, f.mu, _ = cx.pca_fit(raw)
        f.pbase = cx.pca_rotate(raw, f.R, f.mu)
        if c["d"] in (32, 64, 96, 128):
            knn = cx.knn(raw, c["k_fwd"])
        else:
            knn, _ = G.knn_gpu(raw, None, c["k_fwd"], exclude_self=True)
        f.graph = cx.graph_from_knn(knn, c["degree"], rev)
        f.cap = leaf_capacity(c["n"], c["point_ratio"])
        boxes, pids = cx.build_kdtree(f.pbase, f.cap)
    torch.cuda.synchronize()
    del knn, raw
    f.boxes, f.pids = boxes.cpu().numpy(), pids.cpu().numpy()
    f.aabbs = expand_aabbs(f.boxes, 0.2)
    f.pb_h = f.pbase.cpu().numpy()
    f.graph_h = D.pad_graph(f.graph.cpu().numpy(), c["n"])
    ctx.set_base(f.pbase)
    ctx.set_graph(f.graph)
    ctx.set_leaves(f.aabbs, f.pids)
    ctx.set_pca(f.R, f.mu, c["d"])
    f.pq = ctx.project(Q)
    f.hits = ctx.select_entry(f.pq)
    ctx.synchronize()
    f.pq_h, f.hits_h = f.pq.cpu().numpy(), f.hits.cpu().numpy()
    assert f.boxes.shape[0] == c["n_leaf"]
    print(f"[full] workload {cfg}: built {f.boxes.shape[0]} leaves, cap {f.cap}", flush=True)
    return f


def _setup(ctx, f):
    ctx.set_base(f.pbase)
    ctx.set_graph(f.graph)
    ctx.set_leaves(f.aabbs, f.pids)


def test_kdtree_and_entries(wl):
    f = wl
    boxes, belong = O.kdtree_build(np.ascontiguousarray(f.pb_h[:, :3]), f.cap)
    assert_same(f.boxes.view(np.int32), boxes.view(np.int32), "boxes")
    assert_same(f.pids, O.aabb_pids(belong, boxes.shape[0], f.cap), "pids")
    assert_same(f.hits_h, O.select_entry(f.aabbs, f.pq_h), "entry leaves")


def _rows_ok(f, ids, dists, k):
    """ids in range and distinct, distances finite."""
    assert ((ids >= 0) & (ids < f.n)).all()
    assert np.isfinite(dists).all()
    s = np.sort(ids[:, :k], axis=1)
    assert not (s[:, 1:] == s[:, :-1]).any(), "repeated id"


def test_search(ctx, wl):
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    import torch
    f = wl
    _setup(ctx, f)
    nq = f.pq_h.shape[0]
    if f.cfg == 1:
        # every query: k = 10, beam 64, max_iter 100 (the workload's setting)
        p = make_params(k=10, beam=64, max_iter=100)
        ids, dists, st = ctx.search(f.pq, f.hits, p, with_stats=True)
        ctx.synchronize()
        io, do, so = O.search(f.pb_h, f.graph_h, f.pq_h, f.pids, f.hits_h,
                              O.make_params(k=10, beam=64, max_iter=100), with_stats=True)
        assert_same(ids.cpu().numpy(), io, "ids")
        assert_same(dists.cpu().numpy().view(np.int32), do.view(np.int32), "dists")
        assert_same(st.cpu().numpy(), so.astype(np.int32), "stats")
        _rows_ok(f, io, do, 10)
        return
    if f.cfg == 3:
        kq, bq = D.k_beam_assignment(3, nq)
        dev = f.pq.device
        ig, dg = ctx.search_batched(f.pq, f.hits, torch.from_numpy(kq).to(dev),
                                    torch.from_numpy(bq).to(dev), make_params(max_iter=60),
                                    k_stride=int(kq.max()))
        ctx.synchronize()
        ig, dg = ig.cpu().numpy(), dg.cpu().numpy()
        sample = np.sort(np.random.default_rng(31).choice(nq, 256, replace=False))
        io, do = O.search_uneven(f.pb_h, f.graph_h, f.pq_h[sample], f.pids, f.hits_h[sample],
                                 kq[sample], bq[sample], O.make_params(k=10, beam=64, max_iter=60))
        for j, q in enumerate(sample):
            k = kq[q]
            assert_same(ig[q, :k], io[j, :k], f"ids q={q} k={k} beam={bq[q]}")
            assert_same(dg[q, :k].view(np.int32), do[j, :k].view(np.int32), f"dists q={q}")
        for k in (1, 10, 100):
            sel = kq == k
            _rows_ok(f, ig[sel, :k], dg[sel, :k], k)
        return
    # workload 4: 960-d rows, degree 48 at stride 64
    p = make_params(k=10, beam=64, max_iter=40)
    ids, dists = ctx.search(f.pq, f.hits, p)
    bi, bd = ctx.search_batched(f.pq, f.hits, None, None, p, k_stride=10)
    ctx.synchronize()
    ids, dists = ids.cpu().numpy(), dists.cpu().numpy()
    assert_same(bi.cpu().numpy(), ids, "batched ids")
    assert_same(bd.cpu().numpy().view(np.int32), dists.view(np.int32), "batched dists")
    sample = np.random.default_rng(41).choice(nq, 128, replace=False)
    io, do = O.search(f.pb_h, f.graph_h, f.pq_h[sample], f.pids, f.hits_h[sample],
                      O.make_params(k=10, beam=64, max_iter=40))
    assert_same(ids[sample], io, "ids vs oracle")
    assert_same(dists[sample].view(np.int32), do.view(np.int32), "dists vs oracle")
    _rows_ok(f, ids, dists, 10)

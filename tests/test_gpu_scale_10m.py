"""The large-scale combined shape at its full size: 10M x 128 integer-valued
mixture (10k centres; 5.12 GB of FP32 rows, so row offsets pass 2^32 bytes),
100k queries with k in {10, 100}, 16384 kd-tree leaves of 1024 ids, a
128 -> 256 -> 256 -> 16384 entry network.  The graph is a random
degree-32 graph (distinct neighbours, no self loops): an exact kNN graph of
10M rows is an offline build of minutes, and the search's results are
defined for any graph.  Checked: the GPU kd-tree against the oracle's build,
entry selection, 256 sampled queries of the uniform and the uneven batched
search bit-exact against the oracle, properties of every result row, and the
learned entry against selection + network."""
import numpy as np
import pytest

import oracle as O
from conftest import assert_same

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

N, NQ, DEG, POINT_RATIO = 10_000_000, 100_000, 32, 0.0001024


def _random_graph(n, deg, seed):
    """deg distinct uniform neighbours per row, none equal to the row."""
    rng = np.random.default_rng(seed)
    g = rng.integers(0, n - 1, (n, deg), dtype=np.int64)
    g += g >= np.arange(n, dtype=np.int64)[:, None]
    g = g.astype(np.int32)
    s = np.sort(g, axis=1)
    for u in np.nonzero((s[:, 1:] == s[:, :-1]).any(axis=1))[0]:
        row = []
        while len(row) < deg:
            v = int(rng.integers(0, n))
            if v != u and v not in row:
                row.append(v)
        g[u] = row
    return g


class Big:
    pass


@pytest.fixture(scope="module")
def big(ctx):
    import torch
    from rt_entry_amd import Context, expand_aabbs, leaf_capacity
    from rt_entry_amd import data as D
    dev = torch.device("cuda:0")
    w = D.generate(5)
    f = Big()
    raw_h, f.Q_h = w["base"], w["queries"]
    assert raw_h.shape == (N, 128) and f.Q_h.shape == (NQ, 128)
    print(f"[10m] generated {raw_h.shape}", flush=True)
    raw = torch.from_numpy(raw_h).to(dev)
    f.Q = torch.from_numpy(f.Q_h).to(dev)
    with Context(0) as cx:
        f.R, f.mu, f.eig = cx.pca_fit(raw)
        f.pbase = cx.pca_rotate(raw, f.R, f.mu)
        f.cap = leaf_capacity(N, POINT_RATIO)
        boxes, pids = cx.build_kdtree(f.pbase, f.cap)
    torch.cuda.synchronize()
    rows = np.random.default_rng(3).choice(N, 2048, replace=False)
    f.rot_rows, f.raw_rows = rows, raw_h[rows].copy()
    del raw, raw_h, w
    f.boxes, f.pids = boxes.cpu().numpy(), pids.cpu().numpy()
    f.aabbs = expand_aabbs(f.boxes, 0.2)
    f.pb_h = f.pbase.cpu().numpy()
    f.graph_h = _random_graph(N, DEG, seed=4)
    f.graph = torch.from_numpy(f.graph_h).to(dev)
    ctx.set_base(f.pbase)
    ctx.set_graph(f.graph)
    ctx.set_leaves(f.aabbs, f.pids)
    ctx.set_pca(f.R, f.mu, 128)
    f.pq = ctx.project(f.Q)
    f.hits = ctx.select_entry(f.pq)
    ctx.synchronize()
    f.pq_h, f.hits_h = f.pq.cpu().numpy(), f.hits.cpu().numpy()
    print(f"[10m] built: {f.boxes.shape[0]} leaves, cap {f.cap}", flush=True)
    return f


def _setup(ctx, f):
    ctx.set_base(f.pbase)
    ctx.set_graph(f.graph)
    ctx.set_leaves(f.aabbs, f.pids)


def test_pca_rotation_and_kdtree(big):
    """Sampled rotated rows equal the ordered FP64 rotation bit for bit; the
    basis is orthonormal with descending eigenvalues; the GPU kd-tree equals
    the oracle's build on the rotated rows."""
    f = big
    np.testing.assert_allclose(f.R.astype(np.float64).T @ f.R, np.eye(128), atol=1e-5)
    assert (np.diff(f.eig) <= 0).all()
    X64 = f.raw_rows.astype(np.float64)
    R64 = f.R.astype(np.float64)
    s = np.zeros((X64.shape[0], 128))
    for j in range(128):  # j ascending, multiply then add
        s = s + (X64[:, j] - np.float64(f.mu[j]))[:, None] * R64[j][None, :]
    assert_same(f.pb_h[f.rot_rows], s.astype(np.float32), "rotated rows")
    boxes, belong = O.kdtree_build(np.ascontiguousarray(f.pb_h[:, :3]), f.cap)
    assert boxes.shape[0] == 16384 and f.cap == 1024
    assert_same(f.boxes.view(np.int32), boxes.view(np.int32), "boxes")
    assert_same(f.pids, O.aabb_pids(belong, boxes.shape[0], f.cap), "pids")


def test_project_and_select(big):
    f = big
    ref = O.project_f64(f.Q_h, f.R, f.mu, 128)
    bound = 4 * 128 * 2.0 ** -24 * (
        np.abs(f.Q_h).astype(np.float64) @ np.abs(f.R).astype(np.float64)
        + np.abs(f.mu).astype(np.float64) @ np.abs(f.R).astype(np.float64))
    err = np.abs(f.pq_h.astype(np.float64) - ref)
    assert (err <= bound + 1e-30).all(), f"max err {err.max()}"
    assert_same(f.hits_h, O.select_entry(f.aabbs, f.pq_h), "hits")


def _check_rows(f, ids, dists, seed, qidx=None):
    """ids in range, filled and distinct; on 2000 rows the distances equal
    the recomputed squared distances."""
    assert ((ids >= 0) & (ids < N)).all()
    s = np.sort(ids, axis=1)
    assert not (s[:, 1:] == s[:, :-1]).any(), "repeated id"
    rows = np.random.default_rng(seed).choice(ids.shape[0], min(2000, ids.shape[0]), replace=False)
    P = f.pb_h[ids[rows]].astype(np.float64)
    q = f.pq_h[rows if qidx is None else qidx[rows]].astype(np.float64)[:, None, :]
    np.testing.assert_allclose(dists[rows], ((P - q) ** 2).sum(-1), rtol=2e-5, atol=1e-3)


def test_search_uniform(ctx, big):
    from rt_entry_amd import make_params
    f = big
    _setup(ctx, f)
    p = make_params(k=10, beam=128, max_iter=40, mode=0)
    ids, dists = ctx.search(f.pq, f.hits, p)
    bi, bd = ctx.search_batched(f.pq, f.hits, None, None, p, k_stride=10)
    ctx.synchronize()
    ids, dists = ids.cpu().numpy(), dists.cpu().numpy()
    assert_same(bi.cpu().numpy(), ids, "batched ids")
    assert_same(bd.cpu().numpy().view(np.int32), dists.view(np.int32), "batched dists")
    _check_rows(f, ids, dists, 10)
    sample = np.random.default_rng(21).choice(NQ, 256, replace=False)
    po = O.make_params(k=10, beam=128, max_iter=40, mode=0)
    io, do = O.search(f.pb_h, f.graph_h, f.pq_h[sample], f.pids, f.hits_h[sample], po)
    assert_same(ids[sample], io, "ids vs oracle")
    assert_same(dists[sample].view(np.int32), do.view(np.int32), "dists vs oracle")


def test_search_uneven(ctx, big):
    """k in {10, 100}, beam in {64, 128, 256} per query, one persistent
    launch over the 100k queries."""
    import torch
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    f = big
    _setup(ctx, f)
    kq, bq = D.k_beam_assignment(5, NQ, ks=(10, 100), beams=(64, 128, 256))
    dev = f.pq.device
    ig, dg = ctx.search_batched(f.pq, f.hits, torch.from_numpy(kq).to(dev),
                                torch.from_numpy(bq).to(dev), make_params(max_iter=30, mode=1),
                                k_stride=100)
    ctx.synchronize()
    ig, dg = ig.cpu().numpy(), dg.cpu().numpy()
    sample = np.sort(np.random.default_rng(22).choice(NQ, 256, replace=False))
    po = O.make_params(k=10, beam=64, max_iter=30, mode=1)
    po.bloom_words = 0
    io, do = O.search_uneven(f.pb_h, f.graph_h, f.pq_h[sample], f.pids, f.hits_h[sample],
                             kq[sample], bq[sample], po)
    for j, q in enumerate(sample):
        k = kq[q]
        assert_same(ig[q, :k], io[j, :k], f"ids q={q} k={k} beam={bq[q]}")
        assert_same(dg[q, :k].view(np.int32), do[j, :k].view(np.int32), f"dists q={q}")
    for k in (10, 100):
        sel = np.nonzero(kq == k)[0]
        _check_rows(f, ig[sel, :k], dg[sel, :k], k, qidx=sel)


def test_predict_entry(ctx, big):
    """128 -> 256 -> 256 -> 16384 over the 100k queries (two 65536-row
    chunks): rte_predict_entry equals selection + rte_mlp_predict."""
    import torch
    f = big
    _setup(ctx, f)
    ctx.mlp_create(128, 256, 256, 16384, seed=13)
    W, b = ctx.mlp_get()
    rng = np.random.default_rng(14)
    ctx.mlp_set([x * 4 for x in W], [rng.normal(0, 0.5, size=x.shape).astype(np.float32) for x in b])
    ctx.mlp_set_input_norm(f.pq_h.mean(0), 1.0 / (f.pq_h.std(0) + 1e-3))
    for min_conf in (0.0, 0.2, 1.5):
        kd = ctx.select_entry(f.pq)
        conf0 = torch.empty(NQ, dtype=torch.float32, device=f.pq.device)
        ref = ctx.mlp_predict(f.pq, fallback=kd, min_conf=min_conf, conf=conf0)
        conf1 = torch.empty_like(conf0)
        got = ctx.predict_entry(f.pq, min_conf=min_conf, conf=conf1)
        ctx.synchronize()
        assert_same(got.cpu().numpy(), ref.cpu().numpy(), f"leaves min_conf {min_conf}")
        assert_same(conf1.cpu().numpy().view(np.int32), conf0.cpu().numpy().view(np.int32),
                    f"confidences min_conf {min_conf}")
        if min_conf > 1:
            assert_same(got.cpu().numpy(), f.hits_h, "fallback = kd-tree leaves")


def test_entry_links_and_lds_queue_beam512(ctx, big):
    """The combined workload's graph form at full size: entry-space links
    (16 per row, the row's own leaf) added on the device equal the numpy
    definition on 2048 sampled rows; then a uniform batched search at beam
    512 with the beam in an LDS queue (the launch the autotuner picks on
    this workload; row byte offsets pass 2^32) equals the oracle on 128
    sampled queries of the linked graph."""
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    f = big
    _setup(ctx, f)
    g = f.graph.clone()
    ctx.graph_add_entry_links(g, 16, window=0, seed=1)
    ctx.synchronize()
    g_h = g.cpu().numpy()
    rows = np.sort(np.random.default_rng(23).choice(N, 2048, replace=False))
    ref = D.add_entry_links(f.graph_h, f.pids, 16, window=0, seed=1, rows=rows)
    assert_same(g_h[rows], ref, "entry links (sampled rows)")
    ctx.set_graph(g)
    p = make_params(k=10, beam=512, max_iter=40, mode=0, queue=1, order=1)
    ids, dists = ctx.search_batched(f.pq, f.hits, None, None, p, k_stride=10)
    ctx.synchronize()
    ids, dists = ids.cpu().numpy(), dists.cpu().numpy()
    _check_rows(f, ids, dists, 10)
    sample = np.random.default_rng(24).choice(NQ, 128, replace=False)
    po = O.make_params(k=10, beam=512, max_iter=40, mode=0)
    io, do = O.search(f.pb_h, g_h, f.pq_h[sample], f.pids, f.hits_h[sample], po)
    assert_same(ids[sample], io, "ids vs oracle")
    assert_same(dists[sample].view(np.int32), do.view(np.int32), "dists vs oracle")
    ctx.set_graph(f.graph)

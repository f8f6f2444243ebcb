"""The headline workload at its full size: 1M x 128 integer-valued mixture
(1000 centres), 10k queries, degree-32 kNN graph with reverse edges, PCA to
3-D, 4096 kd-tree leaves of 256 ids.  Every offline stage (PCA fit and
rotation, exact kNN, graph, kd-tree) and the query path (projection, entry
selection, learned entry, search, batched launches) runs on the GPU; each
stage is checked on the GPU stage's own inputs: against the oracle on
sampled rows/queries (bit-exact), and over all rows/queries through
properties that hold at any size (exact integer distances, distinct ids,
distances that recompute, recall, launch-shape invariance)."""
import numpy as np
import pytest

import oracle as O
from conftest import assert_same

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

N, NQ, DEG, KFWD, POINT_RATIO = 1_000_000, 10_000, 32, 64, 0.000256


class Full:
    pass


@pytest.fixture(scope="module")
def full():
    import torch
    from rt_entry_amd import Context, expand_aabbs, leaf_capacity
    from rt_entry_amd import data as D
    from rt_entry_amd import gpuprep as G
    dev = torch.device("cuda:0")
    w = D.generate(2)
    f = Full()
    f.raw_h, f.Q_h = w["base"], w["queries"]
    assert f.raw_h.shape == (N, 128) and f.Q_h.shape == (NQ, 128)
    print(f"[full] generated {f.raw_h.shape}", flush=True)
    f.raw = torch.from_numpy(f.raw_h).to(dev)
    f.Q = torch.from_numpy(f.Q_h).to(dev)
    with Context(0) as cx:
        f.R, f.mu, f.eig = cx.pca_fit(f.raw)
        f.pbase = cx.pca_rotate(f.raw, f.R, f.mu)
        f.knn = cx.knn(f.raw, KFWD)
        f.graph = cx.graph_from_knn(f.knn, DEG, DEG)
        f.cap = leaf_capacity(N, POINT_RATIO)
        boxes, pids = cx.build_kdtree(f.pbase, f.cap)
    torch.cuda.synchronize()
    f.boxes, f.pids = boxes.cpu().numpy(), pids.cpu().numpy()
    f.aabbs = expand_aabbs(f.boxes, 0.2)
    f.pb_h = f.pbase.cpu().numpy()
    f.knn_h = f.knn.cpu().numpy()
    f.graph_h = f.graph.cpu().numpy()
    f.gt = G.knn_gpu(f.raw, f.Q, 10, fp64=True)[0].cpu().numpy()
    print(f"[full] built: {f.boxes.shape[0]} leaves, cap {f.cap}", flush=True)
    return f


def _exact_sq_dists(X64, rows):
    """Squared distances from the given rows of X to every row, in FP64
    (exact: the data are integers in [0, 255], so every product and partial
    sum is an integer below 2^53)."""
    xn = (X64 * X64).sum(1)
    Y = X64[rows]
    return xn[:, None] - 2.0 * (X64 @ Y.T) + xn[rows][None, :]


def test_pca_fit_and_rotation(full):
    """Mean, eigenvalues and the three kd-tree axes equal the numpy FP64 fit;
    a sample of rotated rows equals the ordered FP64 rotation bit for bit."""
    from rt_entry_amd import data as D
    f = full
    R0, mu0, eig0 = D.pca_fit(f.raw_h)
    np.testing.assert_allclose(f.mu, mu0, rtol=1e-6, atol=1e-4)
    np.testing.assert_allclose(f.eig, eig0, rtol=1e-9)
    # columns 0..2 are the kd-tree axes; their eigenvalue gaps are >= 0.9%
    np.testing.assert_allclose(f.R[:, :3], R0[:, :3], atol=2e-5)
    np.testing.assert_allclose(f.R.astype(np.float64).T @ f.R, np.eye(128), atol=1e-5)
    rows = np.random.default_rng(5).choice(N, 4096, replace=False)
    X64 = f.raw_h[rows].astype(np.float64)
    R64 = f.R.astype(np.float64)
    s = np.zeros((rows.size, 128))
    for j in range(128):  # j ascending, multiply then add
        s = s + (X64[:, j] - np.float64(f.mu[j]))[:, None] * R64[j][None, :]
    assert_same(f.pb_h[rows], s.astype(np.float32), "rotated rows")


def test_knn_rows_exact(full):
    """Sampled kNN rows equal the exact (distance, id) order; every row is
    free of its own id and of duplicates."""
    f = full
    rows = np.random.default_rng(6).choice(N, 48, replace=False)
    Dm = _exact_sq_dists(f.raw_h.astype(np.float64), rows)
    ids = np.arange(N)
    for i, r in enumerate(rows):
        d = Dm[:, i].copy()
        d[r] = np.inf
        t = np.partition(d, KFWD - 1)[KFWD - 1]
        cand = ids[d <= t]
        exp = cand[np.lexsort((cand, d[cand]))][:KFWD]
        assert_same(f.knn_h[r], exp, f"knn row {r}")
    assert not (f.knn_h == ids[:, None]).any()
    s = np.sort(f.knn_h, axis=1)
    assert not (s[:, 1:] == s[:, :-1]).any()


def _graph_rows(knn, rows, degree, rev):
    """Graph rows by the rule, row by row: candidates are the forward edges
    u->knn[u][j] (rank j) and the reverse edges u->v for knn[v][j] == u,
    j < rev (rank j); a repeated neighbour keeps its smallest (rank, reverse)
    key; the row keeps the `degree` candidates of smallest (rank, reverse,
    id), padded with n."""
    n = knn.shape[0]
    cand = {int(u): {} for u in rows}

    def add(u, v, key):
        if u != v and (v not in cand[u] or key < cand[u][v]):
            cand[u][v] = key

    for u in rows:
        for j, v in enumerate(knn[u]):
            add(int(u), int(v), (j, 0))
    vv, jj = np.nonzero(np.isin(knn[:, :rev], rows))
    for v, j in zip(vv, jj):
        add(int(knn[v, j]), int(v), (int(j), 1))
    out = np.full((len(rows), degree), n, np.int32)
    for i, u in enumerate(rows):
        order = sorted(cand[int(u)].items(), key=lambda kv: (kv[1][0], kv[1][1], kv[0]))
        keep = [v for v, _ in order[:degree]]
        out[i, :len(keep)] = keep
    return out


def test_graph_rows(full):
    f = full
    rng = np.random.default_rng(7)
    # random rows plus the rows most often named in the kNN lists (hubs with
    # the most reverse edges)
    hubs = np.argsort(np.bincount(f.knn_h[:, :DEG].ravel(), minlength=N))[-16:]
    rows = np.unique(np.concatenate([rng.choice(N, 48, replace=False), hubs]))
    assert_same(f.graph_h[rows], _graph_rows(f.knn_h, rows, DEG, DEG), "graph rows")
    g = f.graph_h
    assert ((g >= 0) & (g <= N)).all()
    assert not (g == np.arange(N)[:, None]).any()


def test_kdtree_equals_oracle(full):
    """The GPU kd-tree over the rotated base equals the oracle's build."""
    f = full
    boxes, belong = O.kdtree_build(np.ascontiguousarray(f.pb_h[:, :3]), f.cap)
    assert boxes.shape[0] == 4096
    assert_same(f.boxes.view(np.int32), boxes.view(np.int32), "boxes")
    assert_same(f.pids, O.aabb_pids(belong, boxes.shape[0], f.cap), "pids")


def _setup(ctx, f):
    ctx.set_base(f.pbase)
    ctx.set_graph(f.graph)
    ctx.set_leaves(f.aabbs, f.pids)
    ctx.set_pca(f.R, f.mu, 128)


@pytest.fixture(scope="module")
def query(ctx, full):
    """Projected queries and kd-tree entry leaves of the full workload."""
    f = full
    _setup(ctx, f)
    f.pq = ctx.project(f.Q)
    f.hits = ctx.select_entry(f.pq)
    ctx.synchronize()
    f.pq_h = f.pq.cpu().numpy()
    f.hits_h = f.hits.cpu().numpy()
    return f


def test_project_and_select(query):
    f = query
    ref = O.project_f64(f.Q_h, f.R, f.mu, 128)
    bound = 4 * 128 * 2.0 ** -24 * (
        np.abs(f.Q_h).astype(np.float64) @ np.abs(f.R).astype(np.float64)
        + np.abs(f.mu).astype(np.float64) @ np.abs(f.R).astype(np.float64))
    err = np.abs(f.pq_h.astype(np.float64) - ref)
    assert (err <= bound + 1e-30).all(), f"max err {err.max()}"
    assert_same(f.hits_h, O.select_entry(f.aabbs, f.pq_h), "hits")


def _check_rows(f, ids, dists, seed, qidx=None):
    """Properties of every result row: ids in range, filled and distinct, and
    (on 2000 rows) distances equal to the recomputed squared distance.  The
    order is not asserted: the reference's merge network is not a full sort
    when the beam is wider than the candidate batch (graph_search.cuh:921-958),
    and the sampled oracle comparison covers the order."""
    assert ((ids >= 0) & (ids <= N)).all()
    assert not (ids == N).any(), "unfilled result slot"
    s = np.sort(ids, axis=1)
    assert not (s[:, 1:] == s[:, :-1]).any(), "repeated id"
    unsorted = int((np.diff(dists, axis=1) < 0).any(axis=1).sum())
    print(f"[full] {unsorted} of {ids.shape[0]} rows not in ascending order", flush=True)
    rows = np.random.default_rng(seed).choice(ids.shape[0], min(2000, ids.shape[0]), replace=False)
    P = f.pb_h[ids[rows]].astype(np.float64)
    q = f.pq_h[rows if qidx is None else qidx[rows]].astype(np.float64)[:, None, :]
    ref = ((P - q) ** 2).sum(-1)
    np.testing.assert_allclose(dists[rows], ref, rtol=2e-5, atol=1e-3)


@pytest.mark.parametrize("beam,max_iter,mode", [(256, 51, 0), (64, 33, 0), (128, 40, 1)])
def test_search_full_size(ctx, query, beam, max_iter, mode):
    """All 10k queries through rte_search and the uniform batched launch;
    256 sampled queries bit-exact against the oracle (ids, distances,
    per-query statistics)."""
    from rt_entry_amd import make_params
    f = query
    _setup(ctx, f)
    p = make_params(k=10, beam=beam, max_iter=max_iter, mode=mode)
    print(f"[full] search beam={beam} max_iter={max_iter} mode={mode}", flush=True)
    ids, dists, st = ctx.search(f.pq, f.hits, p, with_stats=True)
    bi, bd = ctx.search_batched(f.pq, f.hits, None, None, p, k_stride=10)
    ctx.synchronize()
    ids, dists, st = ids.cpu().numpy(), dists.cpu().numpy(), st.cpu().numpy()
    assert_same(bi.cpu().numpy(), ids, "batched ids")
    assert_same(bd.cpu().numpy().view(np.int32), dists.view(np.int32), "batched dists")
    _check_rows(f, ids, dists, beam)
    sample = np.random.default_rng(beam + mode).choice(NQ, 256, replace=False)
    bw, bh = O.default_bloom(mode, beam)
    po = O.make_params(k=10, beam=beam, max_iter=max_iter, mode=mode, bloom_words=bw,
                       bloom_hashes=bh)
    io, do, so = O.search(f.pb_h, f.graph_h, f.pq_h[sample], f.pids, f.hits_h[sample], po,
                          with_stats=True)
    assert_same(ids[sample], io, "ids vs oracle")
    assert_same(dists[sample].view(np.int32), do.view(np.int32), "dists vs oracle")
    assert_same(st[sample], so.astype(np.int32), "stats vs oracle")
    rec = O.recall(ids, f.gt)
    print(f"[full] recall@10 {rec:.4f}", flush=True)
    if (beam, max_iter) == (256, 51):
        assert rec >= 0.95


def test_uneven_batch_full_size(ctx, query):
    """Per-query k in {1, 10, 100} and beam in {32, 64, 128, 256} over all
    10k queries in one persistent launch; 256 sampled queries bit-exact
    against the oracle, every row's properties checked."""
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    f = query
    _setup(ctx, f)
    import torch
    kq, bq = D.k_beam_assignment(3, NQ)
    dev = f.pq.device
    p = make_params(max_iter=60, mode=0)
    ig, dg = ctx.search_batched(f.pq, f.hits, torch.from_numpy(kq).to(dev),
                                torch.from_numpy(bq).to(dev), p, k_stride=int(kq.max()))
    ctx.synchronize()
    ig, dg = ig.cpu().numpy(), dg.cpu().numpy()
    sample = np.sort(np.random.default_rng(9).choice(NQ, 256, replace=False))
    po = O.make_params(k=10, beam=64, max_iter=60, mode=0)
    io, do = O.search_uneven(f.pb_h, f.graph_h, f.pq_h[sample], f.pids, f.hits_h[sample],
                             kq[sample], bq[sample], po)
    for j, q in enumerate(sample):
        k = kq[q]
        assert_same(ig[q, :k], io[j, :k], f"ids q={q} k={k} beam={bq[q]}")
        assert_same(dg[q, :k].view(np.int32), do[j, :k].view(np.int32), f"dists q={q}")
    for k in np.unique(kq):
        sel = np.nonzero(kq == k)[0]
        _check_rows(f, ig[sel, :k], dg[sel, :k], int(k), qidx=sel)


def test_predict_entry_full_size(ctx, query):
    """The headline network shape (128 -> 256 -> 256 -> 4096) over all 10k
    queries: rte_predict_entry equals rte_select_entry + rte_mlp_predict,
    and with an unreachable threshold it is the kd-tree selection."""
    import torch
    f = query
    _setup(ctx, f)
    ctx.mlp_create(128, 256, 256, 4096, seed=11)
    W, b = ctx.mlp_get()
    rng = np.random.default_rng(12)
    ctx.mlp_set([x * 4 for x in W], [rng.normal(0, 0.5, size=x.shape).astype(np.float32) for x in b])
    ctx.mlp_set_input_norm(f.pq_h.mean(0), 1.0 / (f.pq_h.std(0) + 1e-3))
    for min_conf in (0.0, 0.3, 1.5):
        kd = ctx.select_entry(f.pq)
        conf0 = torch.empty(NQ, dtype=torch.float32, device=f.pq.device)
        ref = ctx.mlp_predict(f.pq, fallback=kd, min_conf=min_conf, conf=conf0)
        conf1 = torch.empty_like(conf0)
        got = ctx.predict_entry(f.pq, min_conf=min_conf, conf=conf1)
        ctx.synchronize()
        assert_same(got.cpu().numpy(), ref.cpu().numpy(), f"leaves min_conf {min_conf}")
        assert_same(conf1.cpu().numpy().view(np.int32), conf0.cpu().numpy().view(np.int32),
                    f"confidences min_conf {min_conf}")
        c = conf1.cpu().numpy()
        assert ((c > 0) & (c <= 1)).all()
        if min_conf > 1:
            assert_same(got.cpu().numpy(), f.hits_h, "fallback = kd-tree leaves")

"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same FP32 inputs.  Ids and distances must match bit-for-bit."""
import numpy as np
import pytest

import oracle as O
from conftest import assert_same

pytestmark = pytest.mark.gpu


def _ctx_for(ctx, w):
    ctx.set_base(w.base)
    ctx.set_graph(w.graph_raw)
    ctx.set_leaves(w.aabbs, w.pids)
    return ctx


def _run_both(ctx, w, queries, hits, **kw):
    from rt_entry_amd import make_params
    mode = kw.get("mode", 0)
    beam = kw.get("beam", 64)
    bw, bh = O.default_bloom(mode, 1 << (beam - 1).bit_length())
    po = O.make_params(**{**kw, "bloom_words": kw.get("bloom_words", bw),
                          "bloom_hashes": kw.get("bloom_hashes", bh)})
    ids_o, d_o, st_o = O.search(w.base, w.graph, queries, w.pids, hits, po, with_stats=True)
    pg = make_params(**kw)
    ids_g, d_g, st_g = ctx.search(queries, hits, pg, with_dists=True, with_stats=True)
    return (ids_o, d_o, st_o), (ids_g, d_g, st_g)


@pytest.mark.parametrize("beam", [32, 64, 128, 256])
def test_search_parity_beams(ctx, small_workload, beam):
    w = small_workload
    _ctx_for(ctx, w)
    print(f"[parity] beam={beam} n={w.n} nq={w.queries.shape[0]}", flush=True)
    (io, do, so), (ig, dg, sg) = _run_both(ctx, w, w.queries, w.hits, k=10, beam=beam)
    assert_same(ig, io, "ids")
    assert_same(dg.view(np.int32), do.view(np.int32), "dists")
    assert_same(sg, so.astype(np.int32), "stats")


@pytest.mark.parametrize("width", [1, 2, 4])
@pytest.mark.parametrize("mode", [0, 1])
def test_search_parity_width_mode(ctx, int_workload, width, mode):
    w = int_workload
    _ctx_for(ctx, w)
    print(f"[parity] width={width} mode={mode}", flush=True)
    (io, do, so), (ig, dg, sg) = _run_both(ctx, w, w.queries, w.hits, k=10, beam=64,
                                           width=width, mode=mode)
    assert_same(ig, io, "ids")
    assert_same(dg.view(np.int32), do.view(np.int32), "dists")


def test_search_parity_k_equals_beam(ctx, int_workload):
    w = int_workload
    _ctx_for(ctx, w)
    (io, do, _), (ig, dg, _) = _run_both(ctx, w, w.queries, w.hits, k=128, beam=128, max_iter=60)
    assert_same(ig, io, "ids")
    assert_same(dg.view(np.int32), do.view(np.int32), "dists")


@pytest.mark.parametrize("algo", [0, 2])
def test_search_parity_random_entry(ctx, small_workload, algo):
    w = small_workload
    _ctx_for(ctx, w)
    (io, do, _), (ig, dg, _) = _run_both(ctx, w, w.queries, None, k=10, beam=64, algo=algo)
    assert_same(ig, io, "ids")
    assert_same(dg.view(np.int32), do.view(np.int32), "dists")


def test_search_parity_degree48(ctx, deg48_workload):
    w = deg48_workload
    _ctx_for(ctx, w)
    for beam, width in ((32, 1), (64, 2), (128, 1)):
        print(f"[parity] deg48 beam={beam} width={width}", flush=True)
        (io, do, _), (ig, dg, _) = _run_both(ctx, w, w.queries, w.hits, k=10, beam=beam,
                                             width=width)
        assert_same(ig, io, f"ids beam={beam} width={width}")
        assert_same(dg.view(np.int32), do.view(np.int32), "dists")


def test_search_parity_max_iter_and_bloom(ctx, small_workload):
    w = small_workload
    _ctx_for(ctx, w)
    for kw in (dict(max_iter=0), dict(max_iter=1), dict(max_iter=7),
               dict(bloom_words=64, bloom_hashes=3), dict(bloom_words=2048),
               dict(bloom_words=32, mode=1), dict(bloom_words=1024, mode=1, bloom_hashes=7)):
        print(f"[parity] {kw}", flush=True)
        (io, do, _), (ig, dg, _) = _run_both(ctx, w, w.queries, w.hits, k=10, beam=64, **kw)
        assert_same(ig, io, f"ids {kw}")
        assert_same(dg.view(np.int32), do.view(np.int32), f"dists {kw}")


def test_search_parity_truncated_dims(ctx, small_workload):
    """Search on the first 64 PCA columns of a 128-wide base (row stride 128,
    the reference's DIM < d layout) and on d = 100 (partial lane groups)."""
    from rt_entry_amd import make_params
    w = small_workload
    for d in (64, 100):
        ctx.set_base(w.base, d=d)
        ctx.set_graph(w.graph_raw)
        ctx.set_leaves(w.aabbs, w.pids)
        po = O.make_params(k=10, beam=64)
        q = np.ascontiguousarray(w.queries[:, :d])
        base_d = np.ascontiguousarray(w.base[:, :d])
        io, do = O.search(base_d, w.graph, q, w.pids, w.hits, po)
        ig, dg = ctx.search(q, w.hits, make_params(k=10, beam=64))
        assert_same(ig, io, f"ids d={d}")
        assert_same(dg.view(np.int32), do.view(np.int32), f"dists d={d}")


def test_select_entry_parity(ctx, small_workload, int_workload):
    for w in (small_workload, int_workload):
        ctx.set_leaves(w.aabbs, w.pids)
        h = ctx.select_entry(w.queries)
        assert_same(h, w.hits, "hits")
        # points outside every box map to leaf 0; points on box faces are inside
        far = np.full((4, w.queries.shape[1]), 1e9, np.float32)
        assert (ctx.select_entry(far) == 0).all()
        face = np.zeros((w.aabbs.shape[0], w.queries.shape[1]), np.float32)
        face[:, :3] = w.aabbs[:, 3:6]
        assert_same(ctx.select_entry(face), O.select_entry(w.aabbs, face), "faces")


def test_projection_tolerance(ctx, small_workload):
    w = small_workload
    Q = w.raw_queries
    d = Q.shape[1]
    for d_out in (128, 64, 3):
        ctx.set_pca(w.R, w.mu, d_out)
        pq = ctx.project(Q)
        ref = O.project_f64(Q, w.R, w.mu, d_out)
        # FP32 GEMM bound: |err| <= c * d * eps32 * sum_j |q_j||R_jc| (+ the
        # mu.R term); c = 4 covers the bias subtraction and the mu.R GEMV
        bound = 4 * d * 2.0 ** -24 * (
            np.abs(Q).astype(np.float64) @ np.abs(w.R[:, :d_out]).astype(np.float64)
            + np.abs(w.mu).astype(np.float64) @ np.abs(w.R[:, :d_out]).astype(np.float64))
        err = np.abs(pq.astype(np.float64) - ref)
        assert (err <= bound + 1e-30).all(), f"max err {err.max()} d_out={d_out}"


def test_project_then_select_then_search(ctx, int_workload):
    """The whole path on device tensors, checked stage by stage."""
    import torch
    from rt_entry_amd import make_params
    w = int_workload
    dev = torch.device("cuda:0")
    base = torch.from_numpy(w.base).to(dev)
    graph = torch.from_numpy(w.graph_raw).to(dev)
    ctx.set_base(base)
    ctx.set_graph(graph)
    ctx.set_leaves(w.aabbs, w.pids)
    ctx.set_pca(w.R, w.mu, w.base.shape[1])
    Q = torch.from_numpy(w.raw_queries).to(dev)
    pq = ctx.project(Q)
    hits = ctx.select_entry(pq)
    ids, dists = ctx.search(pq, hits, make_params(k=10, beam=64))
    ctx.synchronize()
    pq_h = pq.cpu().numpy()
    assert_same(hits.cpu().numpy(), O.select_entry(w.aabbs, pq_h), "hits")
    io, do = O.search(w.base, w.graph, pq_h, w.pids, hits.cpu().numpy(), O.make_params(k=10, beam=64))
    assert_same(ids.cpu().numpy(), io, "ids")
    assert_same(dists.cpu().numpy().view(np.int32), do.view(np.int32), "dists")
    assert O.recall(ids.cpu().numpy(), w.gt) > 0.8


@pytest.mark.parametrize("mode", [0, 1])
def test_batched_uneven_parity(ctx, int_workload, mode):
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    nq = w.queries.shape[0]
    kq, bq = D.k_beam_assignment(3, nq)
    po = O.make_params(k=10, beam=64, mode=mode)
    if mode == 1:
        po.bloom_words = 0
    io, do = O.search_uneven(w.base, w.graph, w.queries, w.pids, w.hits, kq, bq, po)
    ig, dg = ctx.search_batched(w.queries, w.hits, kq, bq, make_params(mode=mode),
                                k_stride=io.shape[1])
    for q in range(nq):
        k = kq[q]
        assert_same(ig[q, :k], io[q, :k], f"ids q={q} k={k} beam={bq[q]}")
        assert_same(dg[q, :k].view(np.int32), do[q, :k].view(np.int32), f"dists q={q}")


def test_errors_are_codes(ctx, small_workload):
    from rt_entry_amd import RteError, make_params
    w = small_workload
    _ctx_for(ctx, w)
    with pytest.raises(RteError):
        ctx.search(w.queries, w.hits, make_params(k=100, beam=64))  # k > beam
    with pytest.raises(RteError):
        ctx.search(w.queries, w.hits, make_params(k=10, beam=64, width=9))
    with pytest.raises(RteError):
        ctx.search(w.queries[:, :3].copy(), w.hits, make_params(k=10, beam=64))  # q_stride < d
    with pytest.raises(RteError):  # hash count is fixed per mode
        ctx.search(w.queries, w.hits, make_params(k=10, beam=64, bloom_hashes=5))


def test_batched_graph_replays_and_autotune(ctx, int_workload):
    """The batched launch is a cached hipGraph: replay it several times with
    eager searches in between and check every replay; then the autotuner's
    configuration must give the same results."""
    import torch
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    dev = torch.device("cuda:0")
    nq = w.queries.shape[0]
    kq, bq = D.k_beam_assignment(3, nq)
    po = O.make_params(k=10, beam=64, mode=0)
    io, do = O.search_uneven(w.base, w.graph, w.queries, w.pids, w.hits, kq, bq, po)
    PQ = torch.from_numpy(w.queries).to(dev)
    leaf = torch.from_numpy(w.hits).to(dev)
    kqd = torch.from_numpy(kq).to(dev)
    bqd = torch.from_numpy(bq).to(dev)
    p = make_params(mode=0)

    def check(ids, dists, what):
        ids = ids.cpu().numpy()
        dists = dists.cpu().numpy()
        for q in range(nq):
            k = kq[q]
            assert_same(ids[q, :k], io[q, :k], f"{what} ids q={q}")
            assert_same(dists[q, :k].view(np.int32), do[q, :k].view(np.int32), f"{what} q={q}")

    for rep in range(4):
        ids, dists = ctx.search_batched(PQ, leaf, kqd, bqd, p, k_stride=io.shape[1])
        ctx.synchronize()
        check(ids, dists, f"replay {rep}")
        ctx.search(PQ, leaf, make_params(k=10, beam=64))  # eager work in between
        ctx.synchronize()
    best, ms = ctx.autotune_batched(PQ, leaf, kqd, bqd, p, k_stride=io.shape[1], bloom_words=(0,))
    print(f"[autotune] wg_waves={best.wg_waves} grid={best.grid} bloom={best.bloom_words} "
          f"{ms:.3f} ms", flush=True)
    assert best.wg_waves in (1, 2, 4) and best.grid >= -3 and best.wpq in (1, 2, 4) and ms > 0
    ids, dists = ctx.search_batched(PQ, leaf, kqd, bqd, best, k_stride=io.shape[1])
    ctx.synchronize()
    check(ids, dists, "autotuned")


@pytest.mark.parametrize("wpb", [1, 2, 3])
def test_search_launch_shapes_same_results(ctx, small_workload, wpb):
    from rt_entry_amd import make_params
    w = small_workload
    _ctx_for(ctx, w)
    ref, _ = ctx.search(w.queries, w.hits, make_params(k=10, beam=64))
    ids, _ = ctx.search(w.queries, w.hits, make_params(k=10, beam=64, wg_waves=wpb, grid=7))
    assert_same(ids, ref, f"wg_waves={wpb}")


def test_batched_large_batch_prep(ctx, int_workload):
    """A batch of ~17k queries with mixed k and beam (every beam class, many
    preparation workgroups): the workload's queries tiled must give the
    oracle's results tiled, for the whole batch and a prefix of it."""
    import torch
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    dev = torch.device("cuda:0")
    nq = w.queries.shape[0]
    kq, bq = D.k_beam_assignment(3, nq)
    po = O.make_params(k=10, beam=64, mode=0)
    io, do = O.search_uneven(w.base, w.graph, w.queries, w.pids, w.hits, kq, bq, po)
    reps = -(-16385 // nq) + 1
    print(f"[large batch] {nq} queries x {reps}", flush=True)
    PQ = torch.from_numpy(np.tile(w.queries, (reps, 1))).to(dev)
    leaf = torch.from_numpy(np.tile(w.hits, reps)).to(dev)
    kqd = torch.from_numpy(np.tile(kq, reps)).to(dev)
    bqd = torch.from_numpy(np.tile(bq, reps)).to(dev)
    for n in (PQ.shape[0], nq * (16384 // nq)):
        ids, dists = ctx.search_batched(PQ[:n], leaf[:n], kqd[:n], bqd[:n], make_params(mode=0),
                                        k_stride=io.shape[1])
        ctx.synchronize()
        ids = ids.cpu().numpy().reshape(-1, nq, io.shape[1])
        dists = dists.cpu().numpy().reshape(-1, nq, io.shape[1])
        for r in range(ids.shape[0]):
            for q in range(nq):
                k = kq[q]
                if not np.array_equal(ids[r, q, :k], io[q, :k]) or not np.array_equal(
                        dists[r, q, :k].view(np.int32), do[q, :k].view(np.int32)):
                    raise AssertionError(f"n={n} copy {r} query {q}: {ids[r, q, :k]} vs {io[q, :k]}")


def test_batched_uniform_equals_search(ctx, int_workload):
    """A uniform batch (no per-query arrays) is the search kernel alone in a
    captured graph, walking the queries grid-stride: ids and distances equal
    rte_search's for every launch shape, over several replays, and with
    output rows wider than k."""
    import torch
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    dev = torch.device("cuda:0")
    PQ = torch.from_numpy(w.queries).to(dev)
    leaf = torch.from_numpy(w.hits).to(dev)
    for beam, k in ((64, 10), (128, 16)):
        p = make_params(k=k, beam=beam, max_iter=50)
        ref_i, ref_d = ctx.search(PQ, leaf, p)
        ref_i, ref_d = ref_i.cpu().numpy(), ref_d.cpu().numpy()
        for wg, grid in ((4, 0), (1, 0), (2, 3), (1, 1000)):
            q = make_params(k=k, beam=beam, max_iter=50, wg_waves=wg, grid=grid)
            for rep in range(3):
                ids, dists = ctx.search_batched(PQ, leaf, None, None, q, k_stride=k + 3)
                ctx.synchronize()
                ctx.search(PQ, leaf, make_params(k=10, beam=32))  # eager work in between
                ctx.synchronize()
                what = f"beam {beam} k {k} wg {wg} grid {grid} replay {rep}"
                assert_same(ids.cpu().numpy()[:, :k], ref_i, what)
                assert_same(dists.cpu().numpy()[:, :k].view(np.int32), ref_d.view(np.int32), what)
        best, ms = ctx.autotune_batched(PQ, leaf, None, None, p, k_stride=k)
        ids, dists = ctx.search_batched(PQ, leaf, None, None, best, k_stride=k)
        ctx.synchronize()
        assert_same(ids.cpu().numpy(), ref_i, f"autotuned wg {best.wg_waves} grid {best.grid}")


@pytest.mark.parametrize("mode,beam,width", [(0, 64, 1), (1, 128, 1), (0, 64, 2)])
def test_search_parity_960d(ctx, wide_workload, mode, beam, width):
    """960-d vectors (30 dims per lane, the reference's largest DIM) with a
    degree-48 graph: ids, distances and stats equal the oracle's."""
    w = wide_workload
    _ctx_for(ctx, w)
    print(f"[parity] 960-d mode={mode} beam={beam} width={width}", flush=True)
    (io, do, so), (ig, dg, sg) = _run_both(ctx, w, w.queries, w.hits, k=10, beam=beam,
                                           mode=mode, width=width, max_iter=40)
    assert_same(ig, io, "ids")
    assert_same(dg.view(np.int32), do.view(np.int32), "dists")
    assert_same(sg, so.astype(np.int32), "stats")
    assert O.recall(ig, w.gt) > 0.5


@pytest.mark.parametrize("dim,beam,k", [(64, 64, 10), (32, 256, 10), (64, 128, 128)])
def test_rerank_after_truncated_search(ctx, small_workload, dim, beam, k):
    """REORDER path: search on the first `dim` PCA columns returning the whole
    beam, then full-precision rerank on the original vectors: ids and
    distances equal the oracle's (same candidates, same association), on
    device and host buffers, with unfilled slots (id np) in the candidates."""
    import torch
    from rt_entry_amd import make_params
    w = small_workload
    ctx.set_base(w.base, d=dim)
    ctx.set_graph(w.graph_raw)
    ctx.set_leaves(w.aabbs, w.pids)
    ctx.set_rerank_base(w.raw_base)
    q = np.ascontiguousarray(w.queries[:, :dim])
    cand, _ = ctx.search(q, w.hits, make_params(k=beam, beam=beam, max_iter=60))
    io_c, _ = O.search(np.ascontiguousarray(w.base[:, :dim]), w.graph, q, w.pids, w.hits,
                       O.make_params(k=beam, beam=beam, max_iter=60))
    assert_same(cand, io_c, "beam ids")
    cand = cand.copy()
    cand[0, -3:] = w.n  # unfilled slots sort last
    print(f"[rerank] dim={dim} beam={beam} k={k}", flush=True)
    io, do = O.rerank(w.raw_base, w.raw_queries, cand, k)
    ih, dh = ctx.rerank(w.raw_queries, cand, k)  # host buffers
    assert_same(ih, io, "host ids")
    assert_same(dh.view(np.int32), do.view(np.int32), "host dists")
    dev = torch.device("cuda:0")
    idd, ddd = ctx.rerank(torch.from_numpy(w.raw_queries).to(dev), torch.from_numpy(cand).to(dev), k)
    ctx.synchronize()
    assert_same(idd.cpu().numpy(), io, "device ids")
    assert_same(ddd.cpu().numpy().view(np.int32), do.view(np.int32), "device dists")
    if k == 10:  # the rerank recovers recall lost to the truncated search
        print(f"[rerank] recall search-only {O.recall(cand[:, :10], w.gt):.4f} "
              f"reranked {O.recall(io, w.gt):.4f}", flush=True)
        assert O.recall(io, w.gt) >= O.recall(np.ascontiguousarray(cand[:, :10]), w.gt)


@pytest.mark.parametrize("dim,beam", [(64, 64), (64, 128)])
def test_rerank_960d(ctx, wide_workload, dim, beam):
    """REORDER on 960-d rows (30 whole 128-byte lines per row): search on
    the first 64 PCA columns, rerank on the original vectors, bit-exact
    against the oracle."""
    from rt_entry_amd import make_params
    w = wide_workload
    ctx.set_base(w.base, d=dim)
    ctx.set_graph(w.graph_raw)
    ctx.set_leaves(w.aabbs, w.pids)
    ctx.set_rerank_base(w.raw_base)
    q = np.ascontiguousarray(w.queries[:, :dim])
    cand, _ = ctx.search(q, w.hits, make_params(k=beam, beam=beam, max_iter=40))
    print(f"[rerank] 960-d dim={dim} beam={beam}", flush=True)
    io, do = O.rerank(w.raw_base, w.raw_queries, cand, 10)
    ih, dh = ctx.rerank(w.raw_queries, cand, 10)
    assert_same(ih, io, "ids")
    assert_same(dh.view(np.int32), do.view(np.int32), "dists")


@pytest.mark.parametrize("beam", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("mode", [0, 1])
def test_search_parity_small_beams(ctx, int_workload, beam, mode):
    """Beams below 32 (graph.cu:53 accepts any n_candidates): the first
    expansion mask covers the first `beam` slots (P2), the batch (32) is
    wider than the beam; ids, distances and statistics equal the oracle's,
    with kd-tree and random entries and widths 1 and 2."""
    w = int_workload
    _ctx_for(ctx, w)
    for kw in (dict(), dict(width=2), dict(algo=0)):
        print(f"[parity] small beam={beam} mode={mode} {kw}", flush=True)
        hits = None if kw.get("algo") == 0 else w.hits
        (io, do, so), (ig, dg, sg) = _run_both(ctx, w, w.queries, hits, k=beam, beam=beam,
                                               mode=mode, **kw)
        assert_same(ig, io, f"ids {kw}")
        assert_same(dg.view(np.int32), do.view(np.int32), f"dists {kw}")
        assert_same(sg, so.astype(np.int32), f"stats {kw}")


def test_query_one_call_equals_the_three_calls(ctx, int_workload):
    """rte_query (projection, entry selection, search in one call) gives the
    projection, leaves, ids and distances of rte_project + rte_select_entry +
    rte_search bit for bit, and the oracle's ids on its PCA queries; with
    random entries (ALGO 0) no leaves are selected."""
    import torch
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    ctx.set_pca(w.R, w.mu, w.R.shape[1])
    Q = torch.from_numpy(w.raw_queries).cuda()
    p = make_params(k=10, beam=64)
    ids, dists, pq, leaf = ctx.query(Q, p)
    pq2 = ctx.project(Q)
    leaf2 = ctx.select_entry(pq2)
    ids2, dists2 = ctx.search(pq2, leaf2, p)
    ctx.synchronize()
    assert_same(pq.cpu().numpy().view(np.int32), pq2.cpu().numpy().view(np.int32), "PCA queries")
    assert_same(leaf.cpu().numpy(), leaf2.cpu().numpy(), "leaves")
    assert_same(ids.cpu().numpy(), ids2.cpu().numpy(), "ids")
    assert_same(dists.cpu().numpy().view(np.int32), dists2.cpu().numpy().view(np.int32), "dists")
    io, do = O.search(w.base, w.graph, pq.cpu().numpy(), w.pids, leaf.cpu().numpy(),
                      O.make_params(k=10, beam=64))
    assert_same(ids.cpu().numpy(), io, "ids vs oracle")
    p0 = make_params(k=10, beam=64, algo=0)
    ids0, dists0, _, _ = ctx.query(Q, p0)
    ids3, dists3 = ctx.search(pq2, None, p0)
    ctx.synchronize()
    assert_same(ids0.cpu().numpy(), ids3.cpu().numpy(), "random-entry ids")


def test_caller_stream_is_used_and_restored():
    """rte_set_stream makes the context use a caller stream (rte_get_stream
    returns it) and NULL restores the context's own stream."""
    import torch
    from rt_entry_amd import Context
    from rt_entry_amd._lib import lib
    with Context(0) as cx:
        own = lib().rte_get_stream(cx._h)
        assert own
        s = torch.cuda.Stream()
        cx.set_stream(s.cuda_stream)
        assert lib().rte_get_stream(cx._h) == s.cuda_stream
        cx.set_stream(None)
        assert lib().rte_get_stream(cx._h) == own


@pytest.mark.parametrize("d,nq,d_out", [(100, 1, 3), (100, 65, 100), (260, 130, 64),
                                        (960, 67, 128), (64, 200, 37)])
def test_projection_tolerance_odd_shapes(d, nq, d_out):
    """The projection kernel on shapes off its fast paths: rows that are not
    16-byte aligned (d % 4 != 0: scalar loads), several 128-deep chunks with
    a partial last one (each split into two k halves), a partial last row
    tile, output widths that are not multiples of 4 or of the 64-column tile
    -- all within the FP32 GEMM error bound of the FP64 result."""
    rng = np.random.default_rng(d + nq + d_out)
    R = rng.normal(size=(d, d)).astype(np.float32) / np.sqrt(d)
    mu = rng.normal(size=d).astype(np.float32)
    Q = (rng.normal(size=(nq, d)) * 3).astype(np.float32)
    from rt_entry_amd import Context
    with Context(0) as cx:  # its own context: the shared one keeps its PCA
        cx.set_pca(R, mu, d_out)
        pq = cx.project(Q)
    ref = O.project_f64(Q, R, mu, d_out)
    bound = 4 * d * 2.0 ** -24 * (
        np.abs(Q).astype(np.float64) @ np.abs(R[:, :d_out]).astype(np.float64)
        + np.abs(mu).astype(np.float64) @ np.abs(R[:, :d_out]).astype(np.float64))
    err = np.abs(pq.astype(np.float64) - ref)
    print(f"[project] d {d} nq {nq} d_out {d_out} max err {err.max():.3e}", flush=True)
    assert pq.shape == (nq, d_out)
    assert (err <= bound + 1e-30).all(), f"max err {err.max()}"

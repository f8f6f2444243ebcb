"""GPU parity of the launch variants: two and four wavefronts per query
(the query's waves split each batch's distances), the beam held in an LDS
queue instead of VGPRs, the leaf-ordered batched
queue (queries sorted by entry leaf on the device, one slot range per XCD),
and the grid codes.  Every variant must give the oracle's ids, distances
and statistics bit-for-bit: they change how the work is spread, never what
is computed."""
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


def _oracle(w, queries, hits, **kw):
    mode = kw.get("mode", 0)
    beam = kw.get("beam", 64)
    bw, bh = O.default_bloom(mode, 1 << (beam - 1).bit_length())
    po = O.make_params(**{**kw, "bloom_words": bw, "bloom_hashes": bh})
    return O.search(w.base, w.graph, queries, w.pids, hits, po, with_stats=True)


@pytest.mark.parametrize("wpq", [2, 4])
@pytest.mark.parametrize("beam,width,mode", [(32, 1, 0), (64, 1, 1), (64, 2, 0), (256, 1, 0),
                                             (128, 4, 1), (8, 1, 0)])
def test_multiwave_search_parity(ctx, int_workload, wpq, beam, width, mode):
    """rte_search with wpq waves per query equals the oracle (ids, distances,
    statistics), at one and several rows per iteration and both modes."""
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    print(f"[variant] wpq={wpq} beam={beam} width={width} mode={mode}", flush=True)
    k = min(10, beam)
    io, do, so = _oracle(w, w.queries, w.hits, k=k, beam=beam, width=width, mode=mode)
    for grid in (0, 5):  # one workgroup per query, or a grid stride over 5
        ig, dg, sg = ctx.search(w.queries, w.hits,
                                make_params(k=k, beam=beam, width=width, mode=mode, wpq=wpq,
                                            grid=grid), with_stats=True)
        assert_same(ig, io, f"ids grid {grid}")
        assert_same(dg.view(np.int32), do.view(np.int32), f"dists grid {grid}")
        assert_same(sg, so.astype(np.int32), f"stats grid {grid}")


@pytest.mark.parametrize("wpq", [2, 4])
def test_multiwave_other_dims(ctx, small_workload, wide_workload, deg48_workload, wpq):
    """Two and four waves per query at 960-d (two dims groups per lane share
    a load group), a degree-48 graph (64-slot rows) and truncated d = 100."""
    from rt_entry_amd import make_params
    for name, w, kw, d in (("960-d", wide_workload, dict(beam=64, max_iter=40), None),
                           ("deg48", deg48_workload, dict(beam=128, width=2), None),
                           ("d100", small_workload, dict(beam=64), 100)):
        print(f"[variant] wpq={wpq} {name}", flush=True)
        if d is None:
            _ctx_for(ctx, w)
            base, q = w.base, w.queries
        else:
            ctx.set_base(w.base, d=d)
            ctx.set_graph(w.graph_raw)
            ctx.set_leaves(w.aabbs, w.pids)
            base = np.ascontiguousarray(w.base[:, :d])
            q = np.ascontiguousarray(w.queries[:, :d])
        po = O.make_params(k=10, **kw)
        io, do, so = O.search(base, w.graph, q, w.pids, w.hits, po, with_stats=True)
        ig, dg, sg = ctx.search(q, w.hits, make_params(k=10, wpq=wpq, **kw), with_stats=True)
        assert_same(ig, io, f"{name} ids")
        assert_same(dg.view(np.int32), do.view(np.int32), f"{name} dists")
        assert_same(sg, so.astype(np.int32), f"{name} stats")


def test_multiwave_random_entry(ctx, small_workload):
    from rt_entry_amd import make_params
    w = small_workload
    _ctx_for(ctx, w)
    for wpq in (2, 4):
        print(f"[variant] wpq={wpq} random entry", flush=True)
        io, do, so = _oracle(w, w.queries, None, k=10, beam=64, algo=0)
        ig, dg, sg = ctx.search(w.queries, None, make_params(k=10, beam=64, algo=0, wpq=wpq),
                                with_stats=True)
        assert_same(ig, io, "ids")
        assert_same(dg.view(np.int32), do.view(np.int32), "dists")
        assert_same(sg, so.astype(np.int32), "stats")


@pytest.mark.parametrize("beam,width,mode,wg", [(8, 1, 0, 4), (32, 1, 1, 4), (64, 2, 0, 1),
                                                (128, 4, 1, 2), (256, 1, 0, 4), (256, 2, 1, 1),
                                                (512, 1, 0, 4), (512, 1, 1, 2)])
def test_lds_queue_search_parity(ctx, int_workload, beam, width, mode, wg):
    """rte_search with the beam held in an LDS queue (RTE_QUEUE_LDS) equals
    the oracle (ids, distances, statistics) for beams 8..512, one to four
    rows per iteration, both modes and 1, 2 or 4 queries per workgroup."""
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    print(f"[variant] queue=lds beam={beam} width={width} mode={mode} wg={wg}", flush=True)
    k = min(10, beam)
    io, do, so = _oracle(w, w.queries, w.hits, k=k, beam=beam, width=width, mode=mode)
    for grid in (0, 5):
        ig, dg, sg = ctx.search(w.queries, w.hits,
                                make_params(k=k, beam=beam, width=width, mode=mode, queue=1,
                                            wg_waves=wg, grid=grid), with_stats=True)
        assert_same(ig, io, f"ids grid {grid}")
        assert_same(dg.view(np.int32), do.view(np.int32), f"dists grid {grid}")
        assert_same(sg, so.astype(np.int32), f"stats grid {grid}")


def test_lds_queue_other_dims(ctx, small_workload, wide_workload, deg48_workload):
    """The LDS queue at 960-d, a degree-48 graph, truncated d = 100 and
    random (LCG) entries."""
    from rt_entry_amd import make_params
    for name, w, kw, d in (("960-d", wide_workload, dict(beam=128, max_iter=40), None),
                           ("deg48", deg48_workload, dict(beam=256, width=2), None),
                           ("d100", small_workload, dict(beam=128), 100),
                           ("random entry", small_workload, dict(beam=256, algo=0), None)):
        print(f"[variant] queue=lds {name}", flush=True)
        if d is None:
            _ctx_for(ctx, w)
            base, q = w.base, w.queries
        else:
            ctx.set_base(w.base, d=d)
            ctx.set_graph(w.graph_raw)
            ctx.set_leaves(w.aabbs, w.pids)
            base = np.ascontiguousarray(w.base[:, :d])
            q = np.ascontiguousarray(w.queries[:, :d])
        hits = None if kw.get("algo", 1) == 0 else w.hits
        po = O.make_params(k=10, **kw)
        io, do, so = O.search(base, w.graph, q, w.pids, hits, po, with_stats=True)
        ig, dg, sg = ctx.search(q, hits, make_params(k=10, queue=1, **kw), with_stats=True)
        assert_same(ig, io, f"{name} ids")
        assert_same(dg.view(np.int32), do.view(np.int32), f"{name} dists")
        assert_same(sg, so.astype(np.int32), f"{name} stats")


def test_batched_leaf_order_and_wpq(ctx, int_workload):
    """Uniform batched launches with every (wpq, order, grid code) equal
    rte_search, over replays with eager work in between (the leaf sort's
    histogram is left zero by each launch, its range counters are reset by
    the next)."""
    import torch
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    dev = torch.device("cuda:0")
    PQ = torch.from_numpy(w.queries).to(dev)
    leaf = torch.from_numpy(w.hits).to(dev)
    p = make_params(k=10, beam=64, max_iter=50)
    ref_i, ref_d = ctx.search(PQ, leaf, p)
    ref_i, ref_d = ref_i.cpu().numpy(), ref_d.cpu().numpy()
    for wpq, queue in ((1, 0), (2, 0), (4, 0), (1, 1)):
        for order in (0, 1):
            for grid in (0, -1, -2, -3, 3):
                q = make_params(k=10, beam=64, max_iter=50, wpq=wpq, order=order, grid=grid,
                                queue=queue)
                print(f"[variant] batched wpq={wpq} queue={queue} order={order} grid={grid}",
                      flush=True)
                for rep in range(3):
                    ids, dists = ctx.search_batched(PQ, leaf, None, None, q, k_stride=10)
                    ctx.synchronize()
                    ctx.search(PQ, leaf, make_params(k=10, beam=32))  # eager work between
                    ctx.synchronize()
                    what = f"wpq {wpq} order {order} grid {grid} replay {rep}"
                    assert_same(ids.cpu().numpy(), ref_i, what)
                    assert_same(dists.cpu().numpy().view(np.int32), ref_d.view(np.int32), what)


@pytest.mark.parametrize("mode", [0, 1])
def test_batched_uneven_leaf_order(ctx, int_workload, mode):
    """Uneven batch (per-query k and beam) sorted by (beam class, entry
    leaf), with one and two waves per query: the oracle's per-query
    results."""
    import torch
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    w = int_workload
    _ctx_for(ctx, w)
    dev = torch.device("cuda:0")
    nq = w.queries.shape[0]
    kq, bq = D.k_beam_assignment(3, nq)
    po = O.make_params(k=10, beam=64, mode=mode)
    if mode == 1:
        po.bloom_words = 0
    io, do = O.search_uneven(w.base, w.graph, w.queries, w.pids, w.hits, kq, bq, po)
    PQ = torch.from_numpy(w.queries).to(dev)
    leaf = torch.from_numpy(w.hits).to(dev)
    kqd = torch.from_numpy(kq).to(dev)
    bqd = torch.from_numpy(bq).to(dev)
    for wpq, queue in ((1, 0), (2, 0), (1, 1)):
        for rep in range(2):
            print(f"[variant] uneven leaf order wpq={wpq} queue={queue} mode={mode} rep={rep}",
                  flush=True)
            ig, dg = ctx.search_batched(PQ, leaf, kqd, bqd,
                                        make_params(mode=mode, wpq=wpq, order=1, queue=queue),
                                        k_stride=io.shape[1])
            ctx.synchronize()
            ig, dg = ig.cpu().numpy(), dg.cpu().numpy()
            for q in range(nq):
                k = kq[q]
                assert_same(ig[q, :k], io[q, :k], f"ids q={q} k={k} beam={bq[q]}")
                assert_same(dg[q, :k].view(np.int32), do[q, :k].view(np.int32), f"dists q={q}")


def test_autotune_explores_new_knobs(ctx, int_workload):
    """The autotuner's choice is one of its candidates and reproduces
    rte_search; an invalid knob in the request is an error."""
    import torch
    from rt_entry_amd import RteError, make_params
    w = int_workload
    _ctx_for(ctx, w)
    dev = torch.device("cuda:0")
    PQ = torch.from_numpy(w.queries).to(dev)
    leaf = torch.from_numpy(w.hits).to(dev)
    p = make_params(k=10, beam=128, max_iter=40)
    ref_i, _ = ctx.search(PQ, leaf, p)
    best, ms = ctx.autotune_batched(PQ, leaf, None, None, p, k_stride=10, bloom_words=(0, 1024))
    print(f"[variant] autotune wpq={best.wpq} wg={best.wg_waves} queue={best.queue} "
          f"grid={best.grid} order={best.order} bloom={best.bloom_words} {ms:.3f} ms", flush=True)
    assert best.wpq in (1, 2, 4) and best.order in (0, 1) and -3 <= best.grid <= 0
    assert best.queue in (0, 1) and not (best.queue == 1 and best.wpq > 1)
    assert best.bloom_words in (0, 1024) and ms > 0
    ids, _ = ctx.search_batched(PQ, leaf, None, None, best, k_stride=10)
    ctx.synchronize()
    if best.bloom_words == 0:
        assert_same(ids.cpu().numpy(), ref_i.cpu().numpy(), "autotuned")
    with pytest.raises(RteError):
        ctx.search(PQ, leaf, make_params(k=10, beam=64, wpq=3))
    with pytest.raises(RteError):
        ctx.search(PQ, leaf, make_params(k=10, beam=64, wpq=2, wg_waves=4))
    with pytest.raises(RteError):
        ctx.search_batched(PQ, leaf, None, None, make_params(k=10, beam=64, order=2))
    with pytest.raises(RteError):
        ctx.search(PQ, leaf, make_params(k=10, beam=64, queue=2))
    with pytest.raises(RteError):  # the LDS queue runs one wave per query
        ctx.search(PQ, leaf, make_params(k=10, beam=64, queue=1, wpq=2))


def test_batched_uneven_leaf_order_many_keys(ctx, int_workload):
    """Uneven batch sorted by (beam class, entry leaf) over a kd-tree of
    two-point leaves: about 4000 leaves, so about 20,000 sort keys, whose
    histogram and scan run in one workgroup with the counts in dynamic LDS
    (above 8192 keys, which fit the static array of the multi-pass scan).
    The 256 queries give the oracle's per-query results; the same queries
    repeated 80 times (20,480 queries, close to the batch size that path
    takes) give the same results per query."""
    import torch
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    w = int_workload
    boxes, belong = O.kdtree_build(w.base[:, :3], 2)
    n_leaf = boxes.shape[0]
    assert 8192 < 5 * n_leaf <= 32768, n_leaf
    pids = O.aabb_pids(belong, n_leaf, 2)
    aabbs = O.expand_aabbs(boxes, 0.2)
    hits = O.select_entry(aabbs, w.queries)
    ctx.set_base(w.base)
    ctx.set_graph(w.graph_raw)
    ctx.set_leaves(aabbs, pids)
    try:
        dev = torch.device("cuda:0")
        nq = w.queries.shape[0]
        kq, bq = D.k_beam_assignment(4, nq)
        io, do = O.search_uneven(w.base, w.graph, w.queries, pids, hits, kq, bq,
                                 O.make_params(k=10, beam=64, mode=0))
        print(f"[variant] uneven leaf order, {n_leaf} leaves, {5 * n_leaf} keys", flush=True)
        ig, dg = ctx.search_batched(torch.from_numpy(w.queries).to(dev),
                                    torch.from_numpy(hits).to(dev),
                                    torch.from_numpy(kq).to(dev), torch.from_numpy(bq).to(dev),
                                    make_params(mode=0, order=1), k_stride=io.shape[1])
        ctx.synchronize()
        ig, dg = ig.cpu().numpy(), dg.cpu().numpy()
        for q in range(nq):
            k = kq[q]
            assert_same(ig[q, :k], io[q, :k], f"ids q={q} k={k} beam={bq[q]}")
            assert_same(dg[q, :k].view(np.int32), do[q, :k].view(np.int32), f"dists q={q}")
        rep = 80
        assert rep * nq <= 32768
        print(f"[variant] the same batch repeated {rep} times: {rep * nq} queries", flush=True)
        it, dt = ctx.search_batched(torch.from_numpy(np.tile(w.queries, (rep, 1))).to(dev),
                                    torch.from_numpy(np.tile(hits, rep)).to(dev),
                                    torch.from_numpy(np.tile(kq, rep)).to(dev),
                                    torch.from_numpy(np.tile(bq, rep)).to(dev),
                                    make_params(mode=0, order=1), k_stride=io.shape[1])
        ctx.synchronize()
        it, dt = it.cpu().numpy(), dt.cpu().numpy()
        mask = np.arange(io.shape[1])[None, :] < kq[:, None]
        for r in range(rep):
            blk = slice(r * nq, (r + 1) * nq)
            assert_same(np.where(mask, it[blk], 0), np.where(mask, ig, 0), f"ids, copy {r}")
            assert_same(np.where(mask, dt[blk].view(np.int32), 0),
                        np.where(mask, dg.view(np.int32), 0), f"dists, copy {r}")
    finally:
        ctx.set_leaves(w.aabbs, w.pids)

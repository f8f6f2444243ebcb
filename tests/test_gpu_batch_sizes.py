"""Batch sizes at the edges: empty batches (host arrays and device tensors,
whose data pointer is NULL) return without error and write nothing, and
one-query and ragged batches (not multiples of a workgroup's waves or of the
batched launch's queue) match the oracle bit-for-bit."""
import numpy as np
import pytest

import oracle as O
from conftest import assert_same

pytestmark = pytest.mark.gpu


def _setup(ctx, w):
    ctx.set_base(w.base)
    ctx.set_graph(w.graph_raw)
    ctx.set_leaves(w.aabbs, w.pids)


def test_empty_batches(ctx, int_workload):
    import torch
    from rt_entry_amd import make_params
    w = int_workload
    _setup(ctx, w)
    d = w.queries.shape[1]
    dev = torch.device("cuda:0")
    p = make_params(k=10, beam=64)
    for where, empty in (("host", np.zeros((0, d), np.float32)),
                         ("device", torch.zeros((0, d), dtype=torch.float32, device=dev))):
        print(f"[empty] {where}", flush=True)
        leaf = ctx.select_entry(empty)
        assert leaf.shape == (0,)
        ids, dists, st = ctx.search(empty, leaf, p, with_stats=True)
        assert ids.shape == (0, 10) and dists.shape == (0, 10) and st.shape == (0, 3)
        bi, bd = ctx.search_batched(empty, leaf, None, None, p)
        assert bi.shape == (0, 10)
        kq = empty[:, 0].int() if where == "device" else np.zeros(0, np.int32)
        bi, bd = ctx.search_batched(empty, leaf, kq, kq, p, k_stride=10, beam_bound=64)
        assert bi.shape == (0, 10)
        ctx.set_pca(w.R, w.mu, d)
        assert ctx.project(empty).shape == (0, d)
    ctx.synchronize()
    # the predictor on an empty device batch
    ctx.mlp_create(d, 64, 64, w.pids.shape[0], seed=1)
    e = torch.zeros((0, d), dtype=torch.float32, device=dev)
    assert ctx.predict_entry(e).shape == (0,)
    assert ctx.mlp_predict(e).shape == (0,)
    ctx.synchronize()
    # scalar arguments are still checked on an empty batch
    from rt_entry_amd import RteError
    with pytest.raises(RteError):
        ctx.search(np.zeros((0, 3), np.float32), None, p)  # q_stride < d


@pytest.mark.parametrize("nq", [1, 2, 63, 65, 127])
def test_ragged_batches_match_oracle(ctx, int_workload, nq):
    from rt_entry_amd import make_params
    w = int_workload
    _setup(ctx, w)
    Q = np.ascontiguousarray(w.queries[:nq])
    hits = w.hits[:nq]
    print(f"[ragged] nq={nq}", flush=True)
    assert_same(ctx.select_entry(Q), hits, "hits")
    for k, beam, width in ((10, 64, 1), (16, 128, 2)):
        po = O.make_params(k=k, beam=beam, width=width)
        io, do, so = O.search(w.base, w.graph, Q, w.pids, hits, po, with_stats=True)
        ig, dg, sg = ctx.search(Q, hits, make_params(k=k, beam=beam, width=width),
                                with_stats=True)
        what = f"nq {nq} k {k} beam {beam} width {width}"
        assert_same(ig, io, "ids " + what)
        assert_same(dg.view(np.int32), do.view(np.int32), "dists " + what)
        assert_same(sg, so.astype(np.int32), "stats " + what)
        if width == 1:
            bi, bd = ctx.search_batched(Q, hits, None, None, make_params(k=k, beam=beam))
            assert_same(bi, io, "batched ids " + what)
            assert_same(bd.view(np.int32), do.view(np.int32), "batched dists " + what)


@pytest.mark.parametrize("h2", [32, 256, 512])
def test_ragged_batches_predictor(ctx, int_workload, h2):
    """The entry predictor on 1..300-row batches (one partial row block of
    the output-layer kernel: the eight-wave kernel at h2 = 256, the
    four-wave one at 32 and 512; column groups of a single 64-column tile)
    gives every row the leaf it gets inside a 600-row batch (the softmax
    partials then combine in another grouping, so confidences agree to
    rounding), and rte_predict_entry equals rte_select_entry followed by
    rte_mlp_predict bit for bit."""
    import torch
    w = int_workload
    _setup(ctx, w)
    rng = np.random.default_rng(h2)
    d = w.queries.shape[1]
    ctx.mlp_create(d, 64, h2, 1000, seed=h2)
    W, b = ctx.mlp_get()
    ctx.mlp_set([x * 2 for x in W], [rng.normal(0, 0.3, size=x.shape).astype(np.float32)
                                     for x in b])
    X = torch.from_numpy((rng.normal(size=(600, d)) * 3).astype(np.float32)).cuda()
    conf_all = torch.empty(600, dtype=torch.float32, device="cuda")
    leaf_all = ctx.mlp_predict(X, min_conf=0.0, conf=conf_all).cpu().numpy()
    conf_all = conf_all.cpu().numpy()
    for nq in (1, 2, 63, 65, 127, 257, 300):
        print(f"[ragged predictor] h2 {h2} nq {nq}", flush=True)
        Xs = X[:nq].contiguous()
        conf = torch.empty(nq, dtype=torch.float32, device="cuda")
        leaf = ctx.mlp_predict(Xs, min_conf=0.0, conf=conf).cpu().numpy()
        assert_same(leaf, leaf_all[:nq], f"leaves nq {nq}")
        c = conf.cpu().numpy()
        assert np.abs(c - conf_all[:nq]).max() <= 1e-5 * conf_all[:nq].max(), nq
        kd = ctx.select_entry(Xs)
        c0 = torch.empty(nq, dtype=torch.float32, device="cuda")
        c1 = torch.empty(nq, dtype=torch.float32, device="cuda")
        ref = ctx.mlp_predict(Xs, fallback=kd, min_conf=0.5, conf=c0)
        got = ctx.predict_entry(Xs, min_conf=0.5, conf=c1)
        ctx.synchronize()
        assert_same(got.cpu().numpy(), ref.cpu().numpy(), f"predict_entry nq {nq}")
        assert_same(c1.cpu().numpy().view(np.int32), c0.cpu().numpy().view(np.int32),
                    f"predict_entry confidences nq {nq}")

"""Learned entry predictor end to end on a small workload: labels from the
plain GPU search, GPU training, inference with kd-tree fallback, and the gain
of learned entries over kd-tree entries at equal search budget."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_predictor_pipeline_and_gain(ctx, int_workload):
    import torch
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    from rt_entry_amd import predictor as P
    w = int_workload
    dev = torch.device("cuda:0")
    ctx.set_base(w.base)
    ctx.set_graph(w.graph_raw)
    ctx.set_leaves(w.aabbs, w.pids)
    train_raw = D.extra_queries(2, 6000)
    train = D.pca_rotate(train_raw, w.R, w.mu)
    gt_tr, _ = D.knn_exact(w.raw_base, train_raw, 10)
    kd_tr = O.select_entry(w.aabbs, train)
    belong = P.leaf_of_points(w.pids, w.n)
    cand = P.candidate_leaves(kd_tr, gt_tr, belong)
    PQ = torch.from_numpy(train).to(dev)
    labels, score = P.make_labels(ctx, PQ, None, gt_tr, beam=64, budgets=(2, 4, 8, 16),
                                  kd=kd_tr, n_leaf=w.pids.shape[0])
    assert labels.shape == (6000,)
    # the all-leaf labels agree with a shortlist covering every leaf
    sub = slice(0, 40)
    all_l, all_s = P.make_labels(ctx, PQ[sub].contiguous(), None, gt_tr[sub], beam=64,
                                 budgets=(2, 4, 8, 16), kd=kd_tr[sub], n_leaf=w.pids.shape[0])
    full = np.tile(np.arange(w.pids.shape[0], dtype=np.int32), (40, 1))
    full = np.concatenate([kd_tr[sub, None], full[full != kd_tr[sub, None]].reshape(40, -1)], 1)
    l2, s2 = P.make_labels(ctx, PQ[sub].contiguous(), full, gt_tr[sub], beam=64,
                           budgets=(2, 4, 8, 16))
    assert (all_s == s2).all() and (all_s >= score[:40]).all()
    assert ((labels >= 0) & (labels < w.pids.shape[0])).all()
    changed = (labels != kd_tr).mean()
    print(f"[pred] labels differing from the kd-tree leaf: {changed:.3f}", flush=True)
    n_leaf = w.pids.shape[0]
    losses, _ = P.train(ctx, PQ, torch.from_numpy(labels).to(dev), n_leaf, epochs=25,
                        batch=500, lr=2e-3, log=None)
    assert losses[-1] < 3.5, losses
    Qd = torch.from_numpy(w.queries).to(dev)
    kd = torch.from_numpy(w.hits).to(dev)
    conf = torch.empty(Qd.shape[0], dtype=torch.float32, device=dev)
    leaf = P.predict(ctx, Qd, kd, min_conf=0.5, conf=conf)
    leaf_h = leaf.cpu().numpy()
    c = conf.cpu().numpy()
    # fallback: low-confidence queries keep the kd-tree leaf
    assert (leaf_h[c < 0.5] == w.hits[c < 0.5]).all()
    rec = {}
    for name, entry in (("kd", w.hits), ("learned", leaf_h)):
        for mi in (2, 4, 8):
            ids = ctx.search(w.queries, entry, make_params(k=10, beam=64, max_iter=mi),
                             with_dists=False)
            rec[(name, mi)] = O.recall(ids, w.gt)
    print(f"[pred] recall kd/learned: {rec}", flush=True)
    # the learned entries converge faster at a small hop budget
    assert rec[("learned", 4)] > rec[("kd", 4)]


@pytest.mark.parametrize("hidden", [(64, 64), (48, 40)])
def test_predict_entry_equals_select_then_predict(ctx, int_workload, hidden):
    """rte_predict_entry (kd-tree selection on a second stream, overlapping
    the network) gives exactly the leaves and confidences of
    rte_select_entry followed by rte_mlp_predict, for thresholds that keep
    every prediction, some, and none; (64, 64) takes the fused kernels with a
    strided input view, (48, 40) the generic ones."""
    import torch
    w = int_workload
    dev = torch.device("cuda:0")
    ctx.set_leaves(w.aabbs, w.pids)
    n_leaf = w.pids.shape[0]
    d = w.queries.shape[1]
    ctx.mlp_create(d, hidden[0], hidden[1], n_leaf, seed=3)
    rng = np.random.default_rng(2)
    W, b = ctx.mlp_get()
    ctx.mlp_set([x * 4 for x in W], [rng.normal(0, 0.5, size=x.shape).astype(np.float32) for x in b])
    ctx.mlp_set_input_norm(w.queries.mean(0), 1.0 / (w.queries.std(0) + 1e-3))
    pq = torch.from_numpy(w.queries).to(dev)
    if hidden == (64, 64):  # rows 8 floats wider than the network input
        wide = torch.zeros((pq.shape[0], d + 8), dtype=torch.float32, device=dev)
        wide[:, :d] = pq
        X = wide[:, :d]
    else:
        X = pq
    for min_conf in (0.0, 0.5, 1.5):
        kd = ctx.select_entry(pq)
        conf0 = torch.empty(pq.shape[0], dtype=torch.float32, device=dev)
        ref = ctx.mlp_predict(pq, fallback=kd, min_conf=min_conf, conf=conf0)
        conf1 = torch.empty_like(conf0)
        got = ctx.predict_entry(X, min_conf=min_conf, conf=conf1)
        ctx.synchronize()
        ref, got = ref.cpu().numpy(), got.cpu().numpy()
        print(f"[predict_entry] hidden {hidden} min_conf {min_conf}: "
              f"{(ref != kd.cpu().numpy()).mean():.3f} not kd", flush=True)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(conf1.cpu().numpy().view(np.int32),
                                      conf0.cpu().numpy().view(np.int32))
        if min_conf > 1:
            np.testing.assert_array_equal(got, kd.cpu().numpy())

"""rte_search_batched_noalloc (the batched search with its buffers in a
static device array, for callers that cannot allocate) against
rte_search_batched through the C ABI: identical ids and distances, bit for
bit, on two calls in a row with the second on a smaller batch, for uniform
and uneven batches, both query orders and one and two waves per query; and
the limits are errors."""
import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu


def _setup(ctx, w):
    import torch
    dev = torch.device("cuda:0")
    ctx.set_base(torch.from_numpy(w.base).to(dev))
    ctx.set_graph(torch.from_numpy(w.graph_raw).to(dev))
    ctx.set_leaves(w.aabbs, w.pids)
    return dev


def _pair(ctx, PQ, leaf, kq, bq, p, k_stride):
    import torch
    nq = PQ.shape[0]
    ref_i, ref_d = ctx.search_batched(PQ, leaf, kq, bq, p, k_stride=k_stride)
    ids = torch.full((nq, k_stride), -7, dtype=torch.int32, device=PQ.device)
    dists = torch.full((nq, k_stride), -7.0, dtype=torch.float32, device=PQ.device)
    ctx.search_batched_noalloc(PQ, leaf, kq, bq, p, k_stride, out=(ids, dists))
    ctx.synchronize()
    return ref_i.cpu().numpy(), ref_d.cpu().numpy(), ids.cpu().numpy(), dists.cpu().numpy()


def test_noalloc_matches_batched(ctx, int_workload):
    import torch
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    w = int_workload
    dev = _setup(ctx, w)
    nq = w.queries.shape[0]
    PQ = torch.from_numpy(w.queries).to(dev)
    leaf = torch.from_numpy(w.hits).to(dev)
    kq_h, bq_h = D.k_beam_assignment(3, nq)
    kq = torch.from_numpy(kq_h).to(dev)
    bq = torch.from_numpy(bq_h).to(dev)
    for uneven in (False, True):
        for order in (0, 1):
            for wpq, queue in ((1, 0), (2, 0), (1, 1)):
                if uneven:
                    p = make_params(mode=0, order=order, wpq=wpq, queue=queue)
                    k_stride = int(kq_h.max())
                else:
                    p = make_params(k=10, beam=64, max_iter=50, order=order, wpq=wpq,
                                    queue=queue)
                    k_stride = 10
                # two calls in a row, the second on a smaller batch
                for n in (nq, nq // 3):
                    what = f"uneven={uneven} order={order} wpq={wpq} queue={queue} nq={n}"
                    print(f"[noalloc] {what}", flush=True)
                    args = (PQ[:n], leaf[:n], kq[:n] if uneven else None,
                            bq[:n] if uneven else None)
                    ri, rd, gi, gd = _pair(ctx, *args, p, k_stride)
                    if uneven:  # rows hold k_per_q results; the rest is untouched
                        for q in range(n):
                            k = kq_h[q]
                            assert_same(gi[q, :k], ri[q, :k], f"{what} ids q={q}")
                            assert_same(gd[q, :k].view(np.int32), rd[q, :k].view(np.int32),
                                        f"{what} dists q={q}")
                    else:
                        assert_same(gi, ri, f"{what} ids")
                        assert_same(gd.view(np.int32), rd.view(np.int32), f"{what} dists")


def test_noalloc_limits(ctx, int_workload):
    import torch
    from rt_entry_amd import RteError, make_params
    from rt_entry_amd._lib import lib
    w = int_workload
    dev = _setup(ctx, w)
    limit = 262144  # RTE_NOALLOC_MAX_QUERIES
    n = limit + 1
    print(f"[noalloc] limit check nq={n}", flush=True)
    PQ = torch.zeros((n, w.base.shape[1]), dtype=torch.float32, device=dev)
    leaf = torch.zeros(n, dtype=torch.int32, device=dev)
    ids = torch.empty((n, 10), dtype=torch.int32, device=dev)
    dists = torch.empty((n, 10), dtype=torch.float32, device=dev)
    with pytest.raises(RteError, match="above the limit"):
        ctx.search_batched_noalloc(PQ, leaf, None, None, make_params(k=10, beam=64), 10,
                                   out=(ids, dists))
    # the limit itself is accepted (queries of zeros: a valid search)
    ctx.search_batched_noalloc(PQ[:limit], leaf[:limit], None, None,
                               make_params(k=10, beam=32, max_iter=2), 10,
                               out=(ids[:limit], dists[:limit]))
    ctx.synchronize()
    assert int(ids[:limit].min()) >= 0
    assert lib().rte_search_batched_noalloc(None, None, 0, 0, None, None, None, None, 0, None,
                                            None) == -1

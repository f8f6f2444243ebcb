"""Benchmark of the query path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 2]

Workload (default): 1M x 128 integer-valued vectors from a 1000-centre
mixture, 10k queries from the same mixture, exact 64-NN graph with reverse
edges rank-pruned to degree 32, PCA-3D kd-tree with 4096 leaves (cap 256),
k = 10.  Data is synthetic (seeded), prepared on the GPU before timing.

A step is one pass of the query path over the 10k-query batch, with every
input already in HBM: FP32 MFMA projection -> kd-tree entry selection ->
beam search.  The search setting is the cheapest point of a beam x
max_iter sweep whose recall@10 against FP64 exact kNN is >= 0.95.

Printed: one JSON line with the throughput (queries/s), the roofline of the
dominant kernel (search) measured live with HIP events on its stream, and
the CPU baseline (the oracle's search on the host cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS at recall@10≥0.95, 1M×128-d, learned vs kd-tree entry on 1 MI355X"
HBM_PEAK = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def prepare(workload: int, scale: float, dev):
    import torch

    import oracle as O
    from rt_entry_amd import data as D
    from rt_entry_amd import gpuprep as G
    cfg = D.CONFIGS[workload]
    t0 = time.time()
    w = D.generate(workload, scale=scale, n_train=0)
    base_h, q_h = w["base"], w["queries"]
    n, d = base_h.shape
    log(f"generated {n}x{d} base, {q_h.shape[0]} queries in {time.time() - t0:.1f}s")
    base = torch.from_numpy(base_h).to(dev)
    Q = torch.from_numpy(q_h).to(dev)
    t0 = time.time()
    R, mu = G.pca_fit_gpu(base)
    pbase = G.pca_rotate_gpu(base, R, mu)
    log(f"pca {time.time() - t0:.1f}s")
    t0 = time.time()
    knn, _ = G.knn_gpu(base, None, cfg["k_fwd"], exclude_self=True)
    graph = G.graph_from_knn_gpu(knn, cfg["degree"], rev_rank=cfg["degree"])
    del knn
    torch.cuda.synchronize()
    log(f"knn graph {time.time() - t0:.1f}s")
    t0 = time.time()
    gt, _ = G.knn_gpu(base, Q, 10, fp64=True)
    gt = gt.cpu().numpy()
    log(f"ground truth {time.time() - t0:.1f}s")
    t0 = time.time()
    from rt_entry_amd import leaf_capacity
    # entry.cu:22: aabb_size = np * point_ratio (float32); scaled runs keep
    # the leaf count
    cap = leaf_capacity(n, cfg["point_ratio"] / scale)
    pb_h = pbase.cpu().numpy()
    aabbs, boxes, pids = D.build_leaves(pb_h[:, :3], cap, 0.2)
    log(f"kd-tree {boxes.shape[0]} leaves cap {cap} in {time.time() - t0:.1f}s")
    return dict(base=pbase, base_h=pb_h, graph=graph, Q=Q, R=R, mu=mu, gt=gt, aabbs=aabbs,
                pids=pids, n=n, d=d, cap=cap, n_leaf=boxes.shape[0], cfg=cfg)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--beam", type=int, default=0, help="skip the sweep: fixed beam")
    ap.add_argument("--max-iter", type=int, default=0)
    ap.add_argument("--mode", type=int, default=0, help="0 REF comparator, 1 TUNED")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(queries shard across GPUs with no exchange, see the design notes)"}))
        sys.exit(2)

    import torch

    import oracle as O
    from rt_entry_amd import Context, make_params

    dev = torch.device("cuda:0")
    W = prepare(args.workload, args.scale, dev)
    nq = W["Q"].shape[0]
    k = 10
    ctx = Context(0)
    # the query path runs on a stream of its own (default priority); HIP
    # events bracket the search kernel on that stream
    stream = torch.cuda.Stream(device=dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.set_base(W["base"])
    ctx.set_graph(W["graph"])
    ctx.set_leaves(W["aabbs"], W["pids"])
    ctx.set_pca(W["R"], W["mu"], W["d"])

    pq = torch.empty((nq, W["d"]), dtype=torch.float32, device=dev)
    hits = torch.empty(nq, dtype=torch.int32, device=dev)
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    dists = torch.empty((nq, k), dtype=torch.float32, device=dev)

    def step(p):
        ctx.project(W["Q"], out=pq)
        ctx.select_entry(pq, out=hits)
        ctx.search(pq, hits, p, out=(ids, dists, None))

    def recall_of(p):
        step(p)
        torch.cuda.synchronize()
        return O.recall(ids.cpu().numpy(), W["gt"], 10)

    def time_steps(p, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            step(p)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    # ---- sweep: cheapest (beam, max_iter) with recall >= 0.95
    if args.beam:
        chosen = (args.beam, args.max_iter or 100)
        rec = recall_of(make_params(k=k, beam=chosen[0], max_iter=chosen[1], mode=args.mode))
        sweep = [dict(beam=chosen[0], max_iter=chosen[1], recall=rec)]
    else:
        sweep = []
        best = None
        for beam in (32, 64, 128, 256):
            for mi in (20, 30, 40, 60, 80, 100, 150, 200):
                if mi < beam // 4:
                    continue
                p = make_params(k=k, beam=beam, max_iter=mi, mode=args.mode)
                rec = recall_of(p)
                t = time_steps(p, 2)
                sweep.append(dict(beam=beam, max_iter=mi, recall=round(rec, 4),
                                  ms=round(t * 1e3, 3)))
                log(f"sweep beam={beam} max_iter={mi} recall={rec:.4f} {t * 1e3:.3f} ms")
                if rec >= 0.95:
                    if best is None or t < best[0]:
                        best = (t, beam, mi, rec)
                    break
        if best is None:
            best = min(((s["ms"], s["beam"], s["max_iter"], s["recall"]) for s in sweep),
                       key=lambda x: -x[3])
        chosen = (best[1], best[2])
        rec = best[3]
    p = make_params(k=k, beam=chosen[0], max_iter=chosen[1], mode=args.mode)
    log(f"chosen beam={chosen[0]} max_iter={chosen[1]} recall={rec:.4f}")

    # ---- timed region
    for _ in range(args.warmup):
        step(p)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.project(W["Q"], out=pq)
        ctx.select_entry(pq, out=hits)
        ev[i][0].record(stream)
        ctx.search(pq, hits, p, out=(ids, dists, None))
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    qps = nq * args.steps / wall
    search_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    final_recall = O.recall(ids.cpu().numpy(), W["gt"], 10)

    # ---- algorithmic bytes of the search kernel (per query:
    #      4d(E+1) + 4D X + 4N + 8k; E distance evaluations, X rows read)
    stats = torch.empty((nq, 3), dtype=torch.int32, device=dev)
    ctx.search(pq, hits, p, out=(ids, dists, stats))
    torch.cuda.synchronize()
    st = stats.cpu().numpy().astype(np.int64)
    D = 1 << (W["cfg"]["degree"] - 1).bit_length()
    alg_bytes = float((4 * W["d"] * (st[:, 0] + 1) + 4 * D * st[:, 1] + 4 * chosen[0]
                       + 8 * k).sum())
    achieved = alg_bytes / (search_ms * 1e-3)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "search_traffic.json")
    if os.path.exists(tf):
        try:
            tj = json.load(open(tf))
            if tj.get("beam") == chosen[0] and tj.get("max_iter") == chosen[1]:
                traffic = tj.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    # ---- CPU baseline: the oracle's search on the host cores, bounded sample
    cpu = None
    if not args.no_cpu_baseline:
        pq_h = pq.cpu().numpy()
        hits_h = hits.cpu().numpy()
        g_h = W["graph"].cpu().numpy()
        from rt_entry_amd.data import pad_graph
        g_h = pad_graph(g_h, W["n"])
        po = O.make_params(k=k, beam=chosen[0], max_iter=chosen[1], mode=args.mode)
        threads = min(os.cpu_count() or 1, 16)
        m = 256
        t = 0.0
        while True:
            t0 = time.perf_counter()
            O.search(W["base_h"], g_h, pq_h[:m], W["pids"], hits_h[:m], po, n_threads=threads)
            t = time.perf_counter() - t0
            if t >= args.cpu_seconds / 4 or m >= nq:
                break
            m = min(nq, int(m * max(2.0, (args.cpu_seconds / 4) / max(t, 1e-3))))
        reps = max(1, int(args.cpu_seconds / max(t, 1e-3)))
        reps = min(reps, 5)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            O.search(W["base_h"], g_h, pq_h[:m], W["pids"], hits_h[:m], po, n_threads=threads)
            times.append(time.perf_counter() - t0)
        cpu_qps = m / float(np.median(times))
        cpu = {"value": round(cpu_qps, 1), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"first {m} of the {nq} queries, same PCA queries, entry leaves, graph "
                         f"and parameters; search stage only; median of {reps} runs"}

    out = {
        "metric": METRIC,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{W['n']}x{W['d']} integer-valued mixture (workload {args.workload}), "
                        f"{nq} queries, degree-{W['cfg']['degree']} kNN+reverse graph, "
                        f"{W['n_leaf']} kd-tree leaves, k={k}",
            "entry": "kd-tree",
            "beam": chosen[0], "max_iter": chosen[1], "width": 1,
            "mode": "tuned" if args.mode else "ref",
            "recall@10": round(final_recall, 4),
            "search_kernel_ms": round(search_ms, 4),
            "sweep": sweep,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
            "kernel": "search_kernel", "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

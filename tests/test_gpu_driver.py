"""The reference-style host driver (rte_search: Graph::Input/RB_Graph/
Projection/Init_entry/Search over the C ABI) on files in the reference's
formats: ids match the oracle on the same cached PCA artefacts, a cold run
builds its own PCA and kd-tree caches, and out.txt carries the reference's
lines."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import assert_same

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "rt_entry_amd", "rte_search")


def _write_inputs(tmp, w):
    from rt_entry_amd import data as D
    D.write_fvecs(tmp / "base.fvecs", w.raw_base)
    D.write_fvecs(tmp / "query.fvecs", w.raw_queries)
    D.write_ivecs(tmp / "gt.ivecs", w.gt[:, :100])
    D.write_nsw(tmp / "graph.nsw", w.graph_raw)


def _run(tmp, cache, extra=()):
    cmd = [DRIVER, "--data_path", str(tmp / "base.fvecs"), "--query_path",
           str(tmp / "query.fvecs"), "--gt_path", str(tmp / "gt.ivecs"), "--graph_path",
           str(tmp / "graph.nsw"), "--data_name", str(cache), "--n_candidates", "64",
           "--topk", "10", "--point_ratio", str(64 / 8000), "--ids_out", str(tmp / "ids.ivecs"),
           "--out_file", str(tmp / "out.txt"), *extra]
    print("[driver]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_driver_matches_oracle_on_cached_artefacts(tmp_path, int_workload, ctx):
    from rt_entry_amd import data as D
    w = int_workload
    _write_inputs(tmp_path, w)
    cache = tmp_path / "cache"
    cache.mkdir()
    # the reference's cache files: O.fbin, mean.fbin, pca_base.fbin, .aabbs
    D.write_fbin(cache / "O.fbin", w.R)
    D.write_fbin(cache / "mean.fbin", w.mu[None, :])
    D.write_fbin(cache / "pca_base.fbin", w.base)
    cap = 64
    assert O.leaf_capacity(8000, 64 / 8000) == cap
    D.write_aabbs(cache / f"{cap}_0.aabbs", w.boxes, w.pids)
    _run(tmp_path, cache)
    ids = D.read_ivecs(tmp_path / "ids.ivecs")
    # the driver projects the queries with the FP32 MFMA kernel; the same
    # kernel through the Python binding gives the same bits
    ctx.set_pca(w.R, w.mu, w.R.shape[1])
    pq = ctx.project(w.raw_queries)
    hits = O.select_entry(w.aabbs, pq)
    io, _ = O.search(w.base, w.graph, pq, w.pids, hits, O.make_params(k=10, beam=64))
    assert_same(ids, io, "driver ids")
    text = open(tmp_path / "out.txt").read()
    assert "time search:" in text and "recall@10 = " in text and "n_aabbs: " in text
    rec = float(text.split("recall@10 = ")[1].split()[0])
    assert abs(rec - O.recall(ids, w.gt)) < 1e-6


def test_driver_cold_run_builds_caches(tmp_path, int_workload, ctx):
    from rt_entry_amd import data as D
    w = int_workload
    _write_inputs(tmp_path, w)
    cache = tmp_path / "cold"
    _run(tmp_path, cache)
    R = D.read_fbin(cache / "O.fbin").astype(np.float64)
    assert np.allclose(R.T @ R, np.eye(R.shape[0]), atol=1e-4)
    # same principal subspace as the FP64 numpy fit (eigenvectors up to sign)
    assert np.allclose(np.abs(R[:, :3].T @ w.R[:, :3].astype(np.float64)).diagonal(), 1, atol=1e-3)
    # the cold run fits and rotates with the library's GPU PCA
    R_lib, mu_lib, _ = ctx.pca_fit(w.raw_base)
    assert_same(D.read_fbin(cache / "O.fbin"), R_lib, "cached rotation")
    assert_same(D.read_fbin(cache / "mean.fbin")[0], mu_lib, "cached mean")
    pb = D.read_fbin(cache / "pca_base.fbin")
    assert pb.shape == w.base.shape
    assert_same(pb, ctx.pca_rotate(w.raw_base, R_lib, mu_lib), "cached pca base")
    boxes, pids = D.read_aabbs(cache / "64_0.aabbs")
    exp_boxes, belong = O.kdtree_build(pb[:, :3], 64)
    assert_same(boxes, exp_boxes, "cached boxes")
    assert_same(pids, O.aabb_pids(belong, exp_boxes.shape[0], 64), "cached pids")
    ids = D.read_ivecs(tmp_path / "ids.ivecs")
    assert O.recall(ids, w.gt) > 0.8
    # a warm run reads the caches and gives the same ids
    _run(tmp_path, cache)
    assert_same(D.read_ivecs(tmp_path / "ids.ivecs"), ids, "warm run")


def test_driver_reorder_matches_oracle(tmp_path, int_workload, ctx):
    """--dim 64 --reorder 1 (the reference's DIM < d with REORDER): the walk
    on the first 64 PCA columns returns the whole beam, reranked on the
    original vectors; ids equal the oracle's search + rerank."""
    from rt_entry_amd import data as D
    w = int_workload
    _write_inputs(tmp_path, w)
    cache = tmp_path / "cache"
    cache.mkdir()
    D.write_fbin(cache / "O.fbin", w.R)
    D.write_fbin(cache / "mean.fbin", w.mu[None, :])
    D.write_fbin(cache / "pca_base.fbin", w.base)
    D.write_aabbs(cache / "64_0.aabbs", w.boxes, w.pids)
    _run(tmp_path, cache, ("--dim", "64", "--reorder", "1"))
    ids = D.read_ivecs(tmp_path / "ids.ivecs")
    ctx.set_pca(w.R, w.mu, w.R.shape[1])
    pq = ctx.project(w.raw_queries)
    hits = O.select_entry(w.aabbs, pq)
    cand, _ = O.search(np.ascontiguousarray(w.base[:, :64]), w.graph, np.ascontiguousarray(pq[:, :64]),
                       w.pids, hits, O.make_params(k=64, beam=64))
    io, _ = O.rerank(w.raw_base, w.raw_queries, cand, 10)
    assert_same(ids, io, "driver reorder ids")
    text = open(tmp_path / "out.txt").read()
    assert "REORDER = 1" in text and "time reorder:" in text


def test_driver_txt_and_hvecs_inputs(tmp_path, int_workload, ctx):
    """Base and queries as .txt (file_read.cu:103-119) and .hvecs
    (:163-182) give the same ids as the .fvecs run."""
    from rt_entry_amd import data as D
    w = int_workload
    _write_inputs(tmp_path, w)
    cache = tmp_path / "cache"
    _run(tmp_path, cache)
    ref = D.read_ivecs(tmp_path / "ids.ivecs")
    for ext, writer in (("txt", D.write_txt), ("hvecs", D.write_hvecs)):
        writer(tmp_path / f"base.{ext}", w.raw_base)
        writer(tmp_path / f"query.{ext}", w.raw_queries)
        cmd = [DRIVER, "--data_path", str(tmp_path / f"base.{ext}"), "--query_path",
               str(tmp_path / f"query.{ext}"), "--gt_path", str(tmp_path / "gt.ivecs"),
               "--graph_path", str(tmp_path / "graph.nsw"), "--data_name", str(cache),
               "--n_candidates", "64", "--topk", "10", "--point_ratio", str(64 / 8000),
               "--ids_out", str(tmp_path / f"ids_{ext}.ivecs"), "--out_file",
               str(tmp_path / "out.txt")]
        print("[driver]", " ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert_same(D.read_ivecs(tmp_path / f"ids_{ext}.ivecs"), ref, f"{ext} ids")


def test_driver_ground_truth_rows(tmp_path, int_workload, ctx):
    """A ground-truth file with more rows than queries scores the first nq
    rows (graph_helper.cu:64-121); one with fewer rows is an error."""
    from rt_entry_amd import data as D
    w = int_workload
    _write_inputs(tmp_path, w)
    cache = tmp_path / "cache"
    _run(tmp_path, cache)
    ref = D.read_ivecs(tmp_path / "ids.ivecs")
    nq = w.raw_queries.shape[0]
    gt = w.gt[:, :100]
    D.write_ivecs(tmp_path / "gt.ivecs", np.concatenate([gt, gt[:7]], axis=0))
    _run(tmp_path, cache)
    assert_same(D.read_ivecs(tmp_path / "ids.ivecs"), ref, "ids with a longer gt file")
    text = open(tmp_path / "out.txt").read()
    rec = float(text.rsplit("recall@10 = ", 1)[1].split()[0])
    assert abs(rec - O.recall(ref, gt)) < 1e-6
    D.write_ivecs(tmp_path / "gt.ivecs", gt[:nq - 3])
    cmd = [DRIVER, "--data_path", str(tmp_path / "base.fvecs"), "--query_path",
           str(tmp_path / "query.fvecs"), "--gt_path", str(tmp_path / "gt.ivecs"),
           "--graph_path", str(tmp_path / "graph.nsw"), "--data_name", str(cache),
           "--n_candidates", "64", "--topk", "10", "--point_ratio", str(64 / 8000),
           "--out_file", str(tmp_path / "out.txt")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "ground truth has" in r.stderr

"""Workload preparation: the torch graph builder and kNN agree with the numpy
ones (run here on CPU tensors), file formats round-trip."""
import os

import numpy as np
import torch

from conftest import assert_same


def test_knn_and_graph_torch_match_numpy():
    from rt_entry_amd import data as D
    from rt_entry_amd import gpuprep as G
    w = D.generate(2, n=3000, nq=50, n_train=0)
    X, Q = w["base"], w["queries"]
    kn, _ = D.knn_exact(X, X, 64, exclude_self=True)
    kt, _ = G.knn_gpu(torch.from_numpy(X), None, 64, exclude_self=True, block=512)
    assert_same(kt.numpy(), kn, "knn")
    for deg, rev in ((32, 32), (32, 0), (48, 48)):
        gn = D.graph_from_knn(kn, deg, rev_rank=rev)
        gt = G.graph_from_knn_gpu(torch.from_numpy(kn), deg, rev_rank=rev)
        assert_same(gt.numpy(), gn, f"graph deg={deg} rev={rev}")
    gq, _ = D.knn_exact(X, Q, 10)
    gq2, _ = G.knn_gpu(torch.from_numpy(X), torch.from_numpy(Q), 10, fp64=True)
    assert_same(gq2.numpy(), gq, "gt")


def test_graph_rules():
    from rt_entry_amd import data as D
    knn = np.array([[1, 2], [0, 2], [1, 0]], np.int32)
    g = D.graph_from_knn(knn, 2, rev_rank=2)
    # node 0: fwd 1 (r0), 2 (r1); reverse from 1 (r0), 2 (r1) -> dedupe
    assert_same(g, np.array([[1, 2], [0, 2], [1, 0]], np.int32), "tiny graph")
    assert (D.graph_from_knn(knn, 3, rev_rank=0)[:, 2] == 3).all()  # padded with n


def test_pca_and_rotation():
    from rt_entry_amd import data as D
    from rt_entry_amd import gpuprep as G
    w = D.generate(1, n=5000, nq=10, n_train=0)
    R, mu, ev = D.pca_fit(w["base"])
    assert (np.diff(ev) <= 1e-9).all()
    assert np.allclose(R.T.astype(np.float64) @ R, np.eye(128), atol=1e-5)
    R2, mu2 = G.pca_fit_gpu(torch.from_numpy(w["base"]))
    assert np.allclose(np.abs((R2.T.astype(np.float64) @ R)).diagonal()[:20], 1.0, atol=1e-4)
    x = D.pca_rotate(w["base"], R, mu)
    # rotation preserves distances
    a, b = w["base"][0].astype(np.float64), w["base"][1].astype(np.float64)
    assert abs(np.sum((x[0] - x[1]).astype(np.float64) ** 2) - np.sum((a - b) ** 2)) < 1e-3 * np.sum((a - b) ** 2)


def test_file_formats_roundtrip(tmp_path):
    from rt_entry_amd import data as D
    rng = np.random.default_rng(0)
    x = rng.normal(size=(7, 5)).astype(np.float32)
    i = rng.integers(0, 100, size=(7, 3)).astype(np.int32)
    D.write_fvecs(tmp_path / "a.fvecs", x)
    assert_same(D.read_fvecs(tmp_path / "a.fvecs"), x, "fvecs")
    D.write_ivecs(tmp_path / "a.ivecs", i)
    assert_same(D.read_ivecs(tmp_path / "a.ivecs"), i, "ivecs")
    D.write_fbin(tmp_path / "a.fbin", x)
    assert_same(D.read_fbin(tmp_path / "a.fbin"), x, "fbin")
    D.write_nsw(tmp_path / "g.nsw", i)
    assert_same(D.read_nsw(tmp_path / "g.nsw", 7), i, "nsw")
    b = rng.normal(size=(4, 6)).astype(np.float32)
    p = rng.integers(0, 9, size=(4, 8)).astype(np.int32)
    D.write_aabbs(tmp_path / "x.aabbs", b, p)
    b2, p2 = D.read_aabbs(tmp_path / "x.aabbs")
    assert_same(b2, b, "aabbs")
    assert_same(p2, p, "pids")
    raw = np.zeros((3, 4 + 5), np.uint8)
    raw[:, :4] = np.frombuffer(np.int32(5).tobytes(), np.uint8)
    raw[:, 4:] = rng.integers(0, 255, size=(3, 5))
    raw.tofile(tmp_path / "a.bvecs")
    assert_same(D.read_bvecs(tmp_path / "a.bvecs"), raw[:, 4:].astype(np.float32), "bvecs")
    D.write_hvecs(tmp_path / "a.hvecs", x)
    assert_same(D.read_hvecs(tmp_path / "a.hvecs"), x, "hvecs")
    # hvecs: int32 d, then the rows (file_read.cu:163-182)
    hv = np.fromfile(tmp_path / "a.hvecs", np.int32)
    assert hv[0] == 5 and hv.size == 1 + 7 * 5
    D.write_hh(tmp_path / "a.hh", x)
    assert_same(D.read_hh(tmp_path / "a.hh"), x, "hh")
    D.write_txt(tmp_path / "a.txt", x)
    assert_same(D.read_txt(tmp_path / "a.txt"), x, "txt")  # repr round-trips float32
    assert open(tmp_path / "a.txt").readline() == "7 5\n"
    # separators as fscanf("%f,") accepts them: commas, or whitespace alone
    (tmp_path / "b.txt").write_text("2 3\n1.5,2,-3e2\n4\n5.25 6\n")
    assert_same(D.read_txt(tmp_path / "b.txt"),
                np.array([[1.5, 2, -300], [4, 5.25, 6]], np.float32), "txt separators")
    for ext in ("txt", "fvecs", "fbin", "hvecs", "hh"):
        assert_same(D.read_data(tmp_path / f"a.{ext}"), x, f"read_data .{ext}")


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "rt_entry_amd", "rte_search")


def test_driver_reads_every_data_format(tmp_path):
    """The C++ driver's readers (host/graph.cpp read_data) on files written by
    data.py, through `rte_search --check_file` (no GPU work): same n, d and
    FP64 sum of the values for txt, fvecs, fbin, hvecs, bvecs and hh."""
    import subprocess
    from rt_entry_amd import data as D
    if not os.path.exists(DRIVER):
        import pytest
        pytest.skip("driver not built")
    rng = np.random.default_rng(3)
    x = rng.integers(-50, 50, size=(33, 12)).astype(np.float32) / 4
    D.write_txt(tmp_path / "a.txt", x)
    D.write_fvecs(tmp_path / "a.fvecs", x)
    D.write_fbin(tmp_path / "a.fbin", x)
    D.write_hvecs(tmp_path / "a.hvecs", x)
    D.write_hh(tmp_path / "a.hh", x)
    xb = rng.integers(0, 255, size=(33, 12)).astype(np.uint8)
    raw = np.zeros((33, 16), np.uint8)
    raw[:, :4] = np.frombuffer(np.int32(12).tobytes(), np.uint8)
    raw[:, 4:] = xb
    raw.tofile(tmp_path / "a.bvecs")
    (tmp_path / "b.txt").write_text("2 3\n1.5,2,-3e2\n4\n5.25 6\n")
    cases = [(f"a.{e}", x) for e in ("txt", "fvecs", "fbin", "hvecs", "hh")]
    cases += [("a.bvecs", xb.astype(np.float32)),
              ("b.txt", np.array([[1.5, 2, -300], [4, 5.25, 6]], np.float32))]
    for name, ref in cases:
        r = subprocess.run([DRIVER, "--check_file", str(tmp_path / name)], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        n, d, s = r.stdout.split()
        # every value is a multiple of 1/4 below 2^10: the float sum is exact
        assert (int(n), int(d)) == ref.shape and float(s) == float(ref.astype(np.float64).sum()), name
    (tmp_path / "short.txt").write_text("3 2\n1,2\n3,4\n")
    r = subprocess.run([DRIVER, "--check_file", str(tmp_path / "short.txt")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "short txt data" in r.stderr


def test_k_beam_assignment():
    from rt_entry_amd import data as D
    k, b = D.k_beam_assignment(3, 1000)
    assert set(np.unique(k)) == {1, 10, 100}
    assert (b >= 32).all() and (b >= 128)[k == 100].all()
    assert ((b & (b - 1)) == 0).all()


def test_recall_matches_oracle():
    """bench.py's recall (package numpy) against the oracle's `orc_recall`
    and closed forms: all found = 1, none found = 0, a ground truth shorter
    than `at` still divides by `at`, hits counted once per ground-truth id."""
    import oracle as O
    from rt_entry_amd import data as D
    rng = np.random.default_rng(7)
    for nq, k, gk in ((1000, 10, 10), (300, 100, 100), (200, 1, 10), (150, 10, 5), (20000, 10, 10)):
        ids = rng.integers(0, 40, (nq, k)).astype(np.int32)
        gt = rng.integers(0, 40, (nq, gk)).astype(np.int32)
        assert D.recall(ids, gt, 10) == O.recall(ids, gt, 10)
    gt = np.arange(30, dtype=np.int32).reshape(3, 10)
    assert D.recall(gt[:, ::-1], gt, 10) == 1.0
    assert D.recall(gt + 100, gt, 10) == 0.0
    assert D.recall(gt, gt[:, :5], 10) == 0.5
    ids = np.zeros((1, 10), np.int32)
    ids[0, 1] = 1
    assert D.recall(ids, np.array([[1, 2, 3, 4, 5, 6, 7, 8, 9, 10]], np.int32), 10) == 0.1


def test_entry_links_reference_rules():
    """data.add_entry_links (the definition rte_graph_add_entry_links is
    checked against): kept entries unchanged, links drawn from the leaves
    within `window` of the node's own leaf, never the node itself, a kept
    entry or a repeat; deterministic in the seed; r = 0 is the identity."""
    import oracle as O
    from rt_entry_amd import data as D
    rng = np.random.default_rng(11)
    n = 2500
    X = rng.normal(size=(n, 3)).astype(np.float32)
    boxes, belong = O.kdtree_build(X, 64)
    pids = O.aabb_pids(belong, boxes.shape[0], 64)
    g = np.full((n, 16), n, np.int32)
    for u in range(n):
        c = rng.choice(n, 13, replace=False)
        c = c[c != u][:12]
        g[u, :c.size] = c
    assert_same(D.add_entry_links(g, pids, 0), g, "r = 0")
    for r, w in ((4, 0), (6, 2)):
        a = D.add_entry_links(g, pids, r, window=w, seed=9)
        assert_same(a, D.add_entry_links(g, pids, r, window=w, seed=9), "deterministic")
        assert not np.array_equal(a, D.add_entry_links(g, pids, r, window=w, seed=10))
        keep = 16 - r
        assert_same(a[:, :keep], g[:, :keep], "kept entries")
        for u in range(n):
            links = [v for v in a[u, keep:].tolist() if v != n]
            assert len(links) == len(set(links)) and u not in links
            assert not set(links) & set(a[u, :keep].tolist())
            assert all(abs(int(belong[v]) - int(belong[u])) <= w for v in links)
        assert (a[:, keep:] != n).mean() > 0.99
        rows = np.array([0, 7, 7, 1234, n - 1])
        assert_same(D.add_entry_links(g, pids, r, window=w, seed=9, rows=rows), a[rows],
                    "a row subset equals those rows of the whole graph")

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


class Workload:
    """A small search workload: PCA-rotated base and queries, graph (padded
    to a power-of-two stride), leaves and ground truth."""

    def __init__(self, cfg, n, nq, degree, k_fwd, rev, cap, d=None, expand=0.2, seed_off=0):
        from rt_entry_amd import data as D
        import oracle as O
        w = D.generate(cfg, n=n, nq=nq, n_train=0)
        base, Q = w["base"], w["queries"]
        if d is not None:
            base, Q = base[:, :d].copy(), Q[:, :d].copy()
        self.raw_base, self.raw_queries = base, Q
        R, mu, _ = D.pca_fit(base)
        self.R, self.mu = R, mu
        self.base = D.pca_rotate(base, R, mu)
        self.queries = D.pca_rotate(Q, R, mu)
        knn, _ = D.knn_exact(base, base, k_fwd, exclude_self=True)
        g = D.graph_from_knn(knn, degree, rev_rank=rev)
        self.graph_raw = g
        self.graph = D.pad_graph(g, n)
        self.gt, _ = D.knn_exact(base, Q, 100 if n > 100 else n)
        boxes, belong = O.kdtree_build(self.base[:, :3], cap)
        self.boxes = boxes
        self.pids = O.aabb_pids(belong, boxes.shape[0], cap)
        self.aabbs = O.expand_aabbs(boxes, expand)
        self.hits = O.select_entry(self.aabbs, self.queries)
        self.n, self.cap = n, cap


_cache = {}


@pytest.fixture(scope="session")
def small_workload():
    """Workload-1 distribution, 8k points, 128-d, degree-32 kNN graph."""
    if "small" not in _cache:
        _cache["small"] = Workload(1, n=8000, nq=256, degree=32, k_fwd=32, rev=0, cap=64)
    return _cache["small"]


@pytest.fixture(scope="session")
def int_workload():
    """Workload-2 distribution (integer-valued), 8k points, reverse edges."""
    if "int" not in _cache:
        _cache["int"] = Workload(2, n=8000, nq=256, degree=32, k_fwd=64, rev=32, cap=64)
    return _cache["int"]


@pytest.fixture(scope="session")
def wide_workload():
    """960-d low-rank mixture (workload 4's distribution), degree-48 graph
    stored at stride 64: the 30-dims-per-lane kernel."""
    if "wide" not in _cache:
        _cache["wide"] = Workload(4, n=3000, nq=48, degree=48, k_fwd=64, rev=48, cap=64)
    return _cache["wide"]


@pytest.fixture(scope="session")
def deg48_workload():
    """Degree-48 graph (row stride 64), 96-d unit vectors."""
    if "deg48" not in _cache:
        _cache["deg48"] = Workload(3, n=6000, nq=128, degree=48, k_fwd=64, rev=48, cap=64)
    return _cache["deg48"]


@pytest.fixture(scope="session")
def ctx():
    from rt_entry_amd import Context
    c = Context(0)
    yield c
    c.close()


def assert_same(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: {a[i]} vs {b[i]}")

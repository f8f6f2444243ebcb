"""GPU kd-tree build (rte_build_kdtree_gpu) against the host build
(rte_build_kdtree, itself checked against the oracle's restatement of the
reference's Kdtree): same boxes, DFS leaf order, pid table and error codes."""
import numpy as np
import pytest

from rt_entry_amd import RteError, build_kdtree

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(5)
    yield "gauss", rng.normal(size=(100000, 3)).astype(np.float32) * [30, 10, 3], 64
    yield "int-dups", rng.integers(0, 60, size=(50000, 3)).astype(np.float32), 256
    c = rng.normal(size=(40, 3)) * 100
    z = rng.integers(0, 40, size=70000)
    yield "clusters", (c[z] + rng.normal(size=(70000, 3))).astype(np.float32), 128
    yield "small", rng.uniform(-1, 1, size=(1000, 3)).astype(np.float32), 16
    yield "root-leaf", rng.uniform(-1, 1, size=(5, 3)).astype(np.float32), 8
    yield "cap-2048", rng.normal(size=(300000, 3)).astype(np.float32), 2048
    yield "cap-1", rng.permutation(3000).reshape(1000, 3).astype(np.float32), 1
    yield "offset", (rng.normal(size=(60000, 3)) * 0.01 + 50).astype(np.float32), 32


@pytest.mark.parametrize("name,pts,cap", list(_cases()), ids=lambda x: x if isinstance(x, str) else "")
def test_kdtree_gpu_equals_host(ctx, name, pts, cap):
    print(f"[kdtree] {name} n={pts.shape[0]} cap={cap}", flush=True)
    boxes, pids = ctx.build_kdtree(pts, cap)
    hb, hp = build_kdtree(pts, cap)
    assert boxes.shape == hb.shape
    np.testing.assert_array_equal(boxes, hb)
    np.testing.assert_array_equal(pids, hp)


def test_kdtree_gpu_strided_device_input(ctx):
    """Rows of a wider device array (the PCA base itself): only the first
    three columns matter; device outputs."""
    import torch
    rng = np.random.default_rng(8)
    X = rng.normal(size=(200000, 16)).astype(np.float32)
    boxes, pids = ctx.build_kdtree(torch.from_numpy(X).cuda(), 128)
    hb, hp = build_kdtree(np.ascontiguousarray(X[:, :3]), 128)
    np.testing.assert_array_equal(boxes.cpu().numpy(), hb)
    np.testing.assert_array_equal(pids.cpu().numpy(), hp)


def test_kdtree_gpu_errors_match_host(ctx):
    same = np.ones((100, 3), np.float32)          # no split can separate them
    with pytest.raises(RteError) as e1:
        build_kdtree(same, 4)
    with pytest.raises(RteError) as e2:
        ctx.build_kdtree(same, 4)
    assert e1.value.code == e2.value.code
    pts = np.random.default_rng(1).normal(size=(5000, 3)).astype(np.float32)
    with pytest.raises(RteError) as e1:
        build_kdtree(pts, 16, max_leaves=10)
    with pytest.raises(RteError) as e2:
        ctx.build_kdtree(pts, 16, max_leaves=10)
    assert e1.value.code == e2.value.code
    with pytest.raises(RteError):
        ctx.build_kdtree(pts, 4096)                # beyond the GPU builder's leaf size

"""GPU graph construction: exact kNN lists from the MFMA kernel against FP64
exact kNN (bit-exact on integer-valued data, where every FP32 term is an
exact integer; set agreement on float data)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _int_data(n, d, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 255, size=(32, d))
    z = rng.integers(0, 32, size=n)
    return np.clip(np.rint(centres[z] + rng.normal(0, 20, size=(n, d))), 0, 255).astype(np.float32)


@pytest.mark.parametrize("n,d,k", [(6000, 128, 64), (3001, 96, 32), (1000, 32, 10), (130, 64, 64)])
def test_knn_exact_on_integer_data(ctx, n, d, k):
    from rt_entry_amd import data as D
    X = _int_data(n, d, seed=n)
    print(f"[knn] n={n} d={d} k={k}", flush=True)
    ids, dists = ctx.knn(X, k, with_dists=True)
    ref_ids, ref_d = D.knn_exact(X, X, k, exclude_self=True)
    np.testing.assert_array_equal(ids, ref_ids)
    np.testing.assert_array_equal(dists, ref_d.astype(np.float32))


def test_knn_rows_equal_full(ctx):
    """rte_knn_rows over row ranges gives exactly the matching rows of
    rte_knn, for host and device inputs; ranges outside X are errors."""
    import torch
    from rt_entry_amd import RteError
    X = _int_data(5000, 128, seed=4)
    ids, dists = ctx.knn(X, 32, with_dists=True)
    Xd = torch.from_numpy(X).cuda()
    for r0, nr in ((0, 1), (1234, 3000), (4999, 1), (0, 5000)):
        ri, rd = ctx.knn(X, 32, with_dists=True, rows=(r0, nr))
        np.testing.assert_array_equal(ri, ids[r0:r0 + nr])
        np.testing.assert_array_equal(rd, dists[r0:r0 + nr])
        gi = ctx.knn(Xd, 32, rows=(r0, nr)).cpu().numpy()
        np.testing.assert_array_equal(gi, ids[r0:r0 + nr])
    for bad in ((4990, 20), (-1, 5), (10, 0)):
        with pytest.raises(RteError):
            ctx.knn(X, 32, rows=bad)


def test_knn_float_data_agrees(ctx):
    import torch
    from rt_entry_amd import data as D
    w = D.generate(1, n=5000, nq=10, n_train=0)
    X = torch.from_numpy(w["base"]).cuda()
    ids = ctx.knn(X, 64).cpu().numpy()
    ref, _ = D.knn_exact(w["base"], w["base"], 64, exclude_self=True)
    same = np.mean([len(set(a) & set(b)) / 64 for a, b in zip(ids, ref)])
    print(f"[knn] float data set agreement {same:.5f}", flush=True)
    assert same > 0.999
    assert (ids != np.arange(5000)[:, None]).all()


@pytest.mark.parametrize("degree,rev", [(32, 32), (32, 0), (48, 48), (16, 64)])
def test_graph_from_knn_equals_numpy(ctx, degree, rev):
    """GPU reverse edges + dedupe + rank pruning equal data.graph_from_knn
    row for row, on the kernel's own kNN lists and on random lists with many
    shared neighbours (hubs)."""
    from rt_entry_amd import data as D
    X = _int_data(5000, 128, seed=9)
    knn = ctx.knn(X, 64)
    g = ctx.graph_from_knn(knn, degree, rev)
    np.testing.assert_array_equal(g, D.graph_from_knn(knn, degree, rev_rank=rev))
    rng = np.random.default_rng(degree + rev)
    n = 3000
    hub = rng.integers(0, 20, size=(n, 64))          # a few ids referenced by everyone
    rnd = np.stack([rng.permutation(n)[:64] for _ in range(n)])
    knn2 = np.where(rng.random((n, 64)) < 0.3, hub, rnd).astype(np.int32)
    # distinct ids per row, self excluded (as rte_knn produces)
    for i in range(n):
        row, seen = [], set([i])
        for x in list(knn2[i]) + list(rng.permutation(n)):
            if x not in seen:
                seen.add(x)
                row.append(x)
            if len(row) == 64:
                break
        knn2[i] = row
    g2 = ctx.graph_from_knn(knn2, degree, rev)
    np.testing.assert_array_equal(g2, D.graph_from_knn(knn2, degree, rev_rank=rev))


def _aniso(n, d, seed):
    """Gaussian rows with distinct variances along a random orthogonal basis
    (well separated eigenvalues, so the eigenvectors are unique up to sign)."""
    rng = np.random.default_rng(seed)
    Qb, _ = np.linalg.qr(rng.normal(size=(d, d)))
    scales = np.linspace(1.0, 12.0, d)[::-1]
    return (rng.normal(size=(n, d)) * scales @ Qb.T + rng.uniform(-5, 5, size=d)).astype(
        np.float32)


@pytest.mark.parametrize("n,d", [(20000, 128), (7001, 96), (5000, 48), (3000, 13), (6000, 200)])
def test_pca_fit_equals_numpy(ctx, n, d):
    """rte_pca_fit (GPU FP64 covariance + host Jacobi) against data.pca_fit
    (numpy FP64 covariance + LAPACK eigh): same mean, eigenvalues, ordering
    and sign rule."""
    from rt_entry_amd import data as D
    X = _aniso(n, d, seed=d)
    print(f"[pca] n={n} d={d}", flush=True)
    R, mu, eig = ctx.pca_fit(X)
    R0, mu0, eig0 = D.pca_fit(X)
    np.testing.assert_allclose(mu, mu0, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(eig, eig0, rtol=1e-9)
    np.testing.assert_allclose(R, R0, atol=2e-5)
    # orthonormal columns
    np.testing.assert_allclose(R.astype(np.float64).T @ R, np.eye(d), atol=1e-5)


def test_pca_fit_strided_device_input(ctx):
    import torch
    from rt_entry_amd import data as D
    X = _aniso(9000, 64, seed=3)
    wide = np.concatenate([X, np.full((9000, 16), 1e6, np.float32)], axis=1)
    R, mu, eig = ctx.pca_fit(torch.from_numpy(wide).cuda(), d=64)
    R0, mu0, eig0 = D.pca_fit(X)
    np.testing.assert_allclose(mu, mu0, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(eig, eig0, rtol=1e-9)
    np.testing.assert_allclose(R, R0, atol=2e-5)


def _rotate_ordered(X, R, mu):
    """(X - mu) R in FP64 with the reduction over j ascending, multiply then
    add -- the operation order of the reference's host loop."""
    X64 = X.astype(np.float64)
    R64 = R.astype(np.float64)
    s = np.zeros((X.shape[0], R.shape[1]))
    for j in range(X.shape[1]):
        s = s + (X64[:, j] - np.float64(mu[j]))[:, None] * R64[j][None, :]
    return s.astype(np.float32)


@pytest.mark.parametrize("n,d,d_out,device", [(5000, 128, 128, False), (5003, 128, 64, True),
                                              (777, 200, 200, True), (300, 13, 5, False)])
def test_pca_rotate_bit_exact(ctx, n, d, d_out, device):
    import torch
    X = _aniso(n, d, seed=n + d)
    R, mu, _ = ctx.pca_fit(X)
    print(f"[pca_rotate] n={n} d={d} d_out={d_out} device={device}", flush=True)
    out = ctx.pca_rotate(torch.from_numpy(X).cuda() if device else X, R, mu, d_out)
    if device:
        out = out.cpu().numpy()
    np.testing.assert_array_equal(out, _rotate_ordered(X, R[:, :d_out], mu))


def test_entry_links_equal_reference(ctx, int_workload):
    """rte_graph_add_entry_links equals data.add_entry_links bit for bit
    (host and device graphs, several r and windows), and a search on the
    linked graph equals the oracle's on the same graph."""
    import torch
    import oracle as O
    from conftest import assert_same
    from rt_entry_amd import data as D
    from rt_entry_amd import make_params
    w = int_workload
    ctx.set_base(w.base)
    ctx.set_leaves(w.aabbs, w.pids)
    for r, win, seed in ((4, 0, 1), (8, 2, 7), (32, 1, 3)):
        print(f"[links] r={r} window={win} seed={seed}", flush=True)
        ref = D.add_entry_links(w.graph_raw, w.pids, r, window=win, seed=seed)
        gh = ctx.graph_add_entry_links(w.graph_raw, r, window=win, seed=seed)
        assert_same(gh, ref, f"host r={r}")
        gd = torch.from_numpy(w.graph_raw.copy()).cuda()
        ctx.graph_add_entry_links(gd, r, window=win, seed=seed)
        assert_same(gd.cpu().numpy(), ref, f"device r={r}")
    g = D.add_entry_links(w.graph_raw, w.pids, 8, window=1, seed=5)
    ctx.set_graph(g)
    io, do = O.search(w.base, D.pad_graph(g, w.n), w.queries, w.pids, w.hits,
                      O.make_params(k=10, beam=64))
    ig, dg = ctx.search(w.queries, w.hits, make_params(k=10, beam=64))
    assert_same(ig, io, "ids on the linked graph")
    assert_same(dg.view(np.int32), do.view(np.int32), "dists on the linked graph")

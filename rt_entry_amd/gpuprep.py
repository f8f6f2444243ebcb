"""Workload preparation on the GPU with torch: PCA, exact kNN graph with
reverse edges, FP64 ground truth.  This is set-up plumbing for tests and the
benchmark (the reference reads these from disk: graph.cu:114-252); it is not
part of the timed query path.

Exactness: for integer-valued data (workloads 2 and 5) every FP32 product
and partial sum of |x|^2 + |y|^2 - 2 x.y stays below 2^24, so the FP32 GEMM
distances are exact and the graph is the exact kNN graph (ties by id among
the candidates kept).
"""
from __future__ import annotations

import numpy as np
import torch


def pca_fit_gpu(base: torch.Tensor):
    """FP64 mean/covariance on the device, eigen-decomposition on the host;
    eigenvectors sorted by descending eigenvalue, sign-fixed (largest
    component positive).  Returns R (d x d, float32, host), mu (float32,
    host)."""
    n, d = base.shape
    mu = torch.zeros(d, dtype=torch.float64, device=base.device)
    C = torch.zeros((d, d), dtype=torch.float64, device=base.device)
    step = 1 << 18
    for i in range(0, n, step):
        mu += base[i:i + step].double().sum(0)
    mu /= n
    for i in range(0, n, step):
        x = base[i:i + step].double() - mu
        C += x.T @ x
    C /= max(1, n - 1)
    w, V = np.linalg.eigh(C.cpu().numpy())
    order = np.argsort(-w, kind="stable")
    V = V[:, order]
    sgn = np.sign(V[np.abs(V).argmax(0), np.arange(d)])
    V = V * np.where(sgn == 0, 1.0, sgn)
    return V.astype(np.float32), mu.cpu().numpy().astype(np.float32)


def pca_rotate_gpu(x: torch.Tensor, R: np.ndarray, mu: np.ndarray) -> torch.Tensor:
    """(x - mu) R in FP64 on the device, stored FP32."""
    R64 = torch.from_numpy(R.astype(np.float64)).to(x.device)
    mu64 = torch.from_numpy(mu.astype(np.float64)).to(x.device)
    out = torch.empty((x.shape[0], R.shape[1]), dtype=torch.float32, device=x.device)
    step = 1 << 18
    for i in range(0, x.shape[0], step):
        out[i:i + step] = ((x[i:i + step].double() - mu64) @ R64).float()
    return out


@torch.no_grad()
def knn_gpu(X: torch.Tensor, Q: torch.Tensor | None, k: int, exclude_self: bool = False,
            block: int = 2048, margin: int = 32, fp64: bool = False):
    """k nearest rows of X for each row of Q (Q = X when None), ordered by
    (distance, id).  Returns (ids int32 [nq,k], dists [nq,k])."""
    self_mode = Q is None
    if self_mode:
        Q = X
    dt = torch.float64 if fp64 else torch.float32
    # at most ~8 GiB of distances per block of rows
    block = max(64, min(block, (8 << 30) // (X.shape[0] * (8 if fp64 else 4))))
    Xd = X.to(dt)
    xn = (Xd * Xd).sum(1)
    nq = Q.shape[0]
    kk = min(X.shape[0], k + margin + (1 if exclude_self else 0))
    ids = torch.empty((nq, k), dtype=torch.int32, device=X.device)
    dist = torch.empty((nq, k), dtype=dt, device=X.device)
    for s in range(0, nq, block):
        q = Q[s:s + block].to(dt)
        d2 = (q * q).sum(1, keepdim=True) + xn[None, :] - 2.0 * (q @ Xd.T)
        if exclude_self and self_mode:
            r = torch.arange(q.shape[0], device=X.device)
            d2[r, s + r] = float("inf")
        vals, cand = torch.topk(d2, kk, dim=1, largest=False, sorted=True)
        # order candidates by (dist, id)
        o = torch.argsort(cand, dim=1, stable=True)
        cand, vals = torch.gather(cand, 1, o), torch.gather(vals, 1, o)
        o = torch.argsort(vals, dim=1, stable=True)
        ids[s:s + block] = torch.gather(cand, 1, o)[:, :k].int()
        dist[s:s + block] = torch.gather(vals, 1, o)[:, :k]
        del d2
    return ids, dist


@torch.no_grad()
def graph_from_knn_gpu(knn: torch.Tensor, degree: int, rev_rank: int | None = None):
    """Device version of data.graph_from_knn (same rules): forward edges of
    rank j, reverse edges of rank j < rev_rank, dedupe keeping the smallest
    (rank, reverse), keep the `degree` smallest (rank, reverse, id), pad
    with n."""
    n, kf = knn.shape
    dev = knn.device
    rev_rank = degree if rev_rank is None else rev_rank
    ar = torch.arange(n, device=dev, dtype=torch.int64)
    u = ar.repeat_interleave(kf)
    v = knn.reshape(-1).long()
    r = torch.arange(kf, device=dev, dtype=torch.int64).repeat(n)
    t = torch.zeros_like(u)
    if rev_rank > 0:
        kr = min(rev_rank, kf)
        u = torch.cat([u, knn[:, :kr].reshape(-1).long()])
        v = torch.cat([v, ar.repeat_interleave(kr)])
        r = torch.cat([r, torch.arange(kr, device=dev, dtype=torch.int64).repeat(n)])
        t = torch.cat([t, torch.ones(n * kr, device=dev, dtype=torch.int64)])
    keep = u != v
    u, v, r, t = u[keep], v[keep], r[keep], t[keep]
    # key1 orders by (u, v, rank, type): u < 2^24, v < 2^24, rank < 2^7
    key = (u << 39) | (v << 15) | (r << 1) | t
    key, _ = torch.sort(key)
    u = key >> 39
    v = (key >> 15) & ((1 << 24) - 1)
    r = (key >> 1) & ((1 << 14) - 1)
    t = key & 1
    first = torch.ones_like(u, dtype=torch.bool)
    first[1:] = (u[1:] != u[:-1]) | (v[1:] != v[:-1])
    u, v, r, t = u[first], v[first], r[first], t[first]
    key2 = (u << 39) | (r << 25) | (t << 24) | v
    key2, _ = torch.sort(key2)
    u = key2 >> 39
    v = key2 & ((1 << 24) - 1)
    start = torch.searchsorted(u, ar)
    pos = torch.arange(u.numel(), device=dev) - start[u]
    sel = pos < degree
    g = torch.full((n, degree), n, dtype=torch.int32, device=dev)
    g[u[sel], pos[sel]] = v[sel].int()
    return g

"""Python host API over the C ABI.

`Context` mirrors include/rt_entry.h one call for one call.  Arguments may be
numpy arrays (host memory: RTE_HOST) or torch CUDA tensors (device memory:
RTE_DEVICE); device tensors passed to the set_* calls are borrowed, so the
context keeps a reference to them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (MODE_REF, MODE_TUNED, RTE_DEVICE, RTE_HOST, Params, RteError, check, lib)

__all__ = ["Context", "Params", "make_params", "MODE_REF", "MODE_TUNED", "RteError",
           "build_kdtree", "expand_aabbs", "hash_parameter", "leaf_capacity"]


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch") and hasattr(x, "is_cuda")


def _ptr(x, dtype):
    """(pointer, where, keepalive) of a numpy array or torch tensor."""
    if x is None:
        return None, RTE_HOST, None
    if _is_torch(x):
        import torch
        want = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
        if x.dtype != want:
            raise TypeError(f"expected {want}, got {x.dtype}")
        if not x.is_contiguous():
            raise ValueError("device tensors must be contiguous")
        if not x.is_cuda:
            a = x.numpy()
            return a.ctypes.data_as(C.c_void_p), RTE_HOST, a
        return C.c_void_p(x.data_ptr()), RTE_DEVICE, x
    a = np.ascontiguousarray(x, dtype=dtype)
    return a.ctypes.data_as(C.c_void_p), RTE_HOST, a


def make_params(k=10, beam=64, width=1, max_iter=100, mode=MODE_REF, algo=1, bloom_words=0,
                bloom_hashes=0, wg_waves=0, grid=0, wpq=0, order=0, queue=0) -> Params:
    return Params(int(k), int(beam), int(width), int(max_iter), int(mode), int(algo),
                  int(bloom_words), int(bloom_hashes), int(wg_waves), int(grid), int(wpq),
                  int(order), int(queue))


def hash_parameter(nc: int):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    lib().rte_hash_parameter(int(nc), C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def leaf_capacity(np_: int, point_ratio: float) -> int:
    """entry.cu:22: aabb_size = np * point_ratio in float32, truncated."""
    return int(np.float32(np_) * np.float32(point_ratio))


def build_kdtree(pts3, max_node: int, max_leaves: int | None = None):
    """Host kd-tree (librte): returns (unexpanded boxes [n_leaf,6], pids
    [n_leaf,max_node] ascending, padded with np)."""
    pts = np.ascontiguousarray(pts3, dtype=np.float32)
    n = pts.shape[0]
    if max_leaves is None:
        max_leaves = max(16, 4 * (n // max(1, max_node)) + 16)
    aabbs = np.zeros((max_leaves, 6), np.float32)
    pids = np.zeros((max_leaves, max_node), np.int32)
    nl = C.c_int()
    rc = lib().rte_build_kdtree(pts.ctypes.data_as(C.c_void_p), n, int(max_node), max_leaves,
                                aabbs.ctypes.data_as(C.c_void_p), pids.ctypes.data_as(C.c_void_p),
                                C.byref(nl))
    if rc != 0:
        raise RteError(rc, "rte_build_kdtree failed")
    return aabbs[: nl.value].copy(), pids[: nl.value].copy()


def expand_aabbs(aabbs, ratio: float):
    a = np.array(aabbs, dtype=np.float32, copy=True, order="C")
    lib().rte_expand_aabbs(a.ctypes.data_as(C.c_void_p), a.shape[0], float(ratio))
    return a


class Context:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().rte_create(C.byref(self._h), int(device))
        if rc != 0:
            raise RteError(rc, f"rte_create(device={device}) failed")
        self._keep = {}
        self.d = 0
        self.d_out = 0
        self.np = 0

    def close(self):
        if self._h:
            lib().rte_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        return check(self._h, rc)

    @property
    def stream(self) -> int:
        return int(lib().rte_get_stream(self._h) or 0)

    def set_stream(self, stream_ptr: int | None):
        self._chk(lib().rte_set_stream(self._h, C.c_void_p(stream_ptr or 0)))

    def synchronize(self):
        self._chk(lib().rte_synchronize(self._h))

    # ---- data
    def set_base(self, base, d: int | None = None):
        p, w, keep = _ptr(base, np.float32)
        n, stride = base.shape
        d = stride if d is None else d
        self._chk(lib().rte_set_base(self._h, p, n, d, stride, w))
        self._keep["base"] = keep if w == RTE_DEVICE else None
        self.d, self.np = d, n

    def set_graph(self, rows):
        p, w, keep = _ptr(rows, np.int32)
        n, deg = rows.shape
        self._chk(lib().rte_set_graph(self._h, p, n, deg, w))
        self._keep["graph"] = keep if w == RTE_DEVICE else None

    def set_pca(self, R, mu, d_out: int):
        pr, w, k1 = _ptr(R, np.float32)
        pm, w2, k2 = _ptr(mu, np.float32)
        if w != w2:
            raise ValueError("R and mu must live in the same memory")
        d, rc = R.shape
        self._chk(lib().rte_set_pca(self._h, pr, pm, d, rc, int(d_out), w))
        self.d_out = int(d_out)

    def project(self, Q, out=None):
        pq, w, keep = _ptr(Q, np.float32)
        nq = Q.shape[0]
        if out is None:
            if w == RTE_DEVICE:
                import torch
                out = torch.empty((nq, self.d_out), dtype=torch.float32, device=Q.device)
            else:
                out = np.empty((nq, self.d_out), np.float32)
        po, wo, _ = _ptr(out, np.float32)
        if wo != w:
            raise ValueError("input and output must live in the same memory")
        self._chk(lib().rte_project(self._h, pq, nq, po, w))
        return out

    def set_leaves(self, aabbs_expanded, pids):
        pa, w, k1 = _ptr(aabbs_expanded, np.float32)
        pp, w2, k2 = _ptr(pids, np.int32)
        if w != w2:
            raise ValueError("boxes and pids must live in the same memory")
        n_leaf, cap = pids.shape
        self._chk(lib().rte_set_leaves(self._h, pa, pp, n_leaf, cap, w))
        self._keep["pids"] = k2 if w == RTE_DEVICE else None

    def select_entry(self, PQ, out=None):
        pq, w, keep = _ptr(PQ, np.float32)
        nq, stride = PQ.shape
        if out is None:
            if w == RTE_DEVICE:
                import torch
                out = torch.empty(nq, dtype=torch.int32, device=PQ.device)
            else:
                out = np.empty(nq, np.int32)
        po, wo, _ = _ptr(out, np.int32)
        self._chk(lib().rte_select_entry(self._h, pq, nq, stride, po, w))
        return out

    # ---- search
    def search(self, PQ, leaf, params: Params, with_dists=True, with_stats=False, out=None):
        pq, w, keep = _ptr(PQ, np.float32)
        nq, qs = PQ.shape
        k = params.k
        if out is None:
            if w == RTE_DEVICE:
                import torch
                dev = PQ.device
                ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
                dists = torch.empty((nq, k), dtype=torch.float32, device=dev) if with_dists else None
                stats = torch.empty((nq, 3), dtype=torch.int32, device=dev) if with_stats else None
            else:
                ids = np.empty((nq, k), np.int32)
                dists = np.empty((nq, k), np.float32) if with_dists else None
                stats = np.empty((nq, 3), np.int32) if with_stats else None
        else:
            ids, dists, stats = out
        pl, _, _ = _ptr(leaf, np.int32) if leaf is not None else (None, w, None)
        pi, _, _ = _ptr(ids, np.int32)
        pd, _, _ = _ptr(dists, np.float32) if dists is not None else (None, w, None)
        ps, _, _ = _ptr(stats, np.int32) if stats is not None else (None, w, None)
        self._chk(lib().rte_search(self._h, pq, nq, qs, pl, C.byref(params), pi, pd, ps, w))
        res = [ids]
        if with_dists:
            res.append(dists)
        if with_stats:
            res.append(stats)
        return tuple(res) if len(res) > 1 else ids

    def query(self, Q, params: Params):
        """rte_query: the whole query path on device tensors in one call --
        projection, kd-tree entry selection (ALGO 1; the random entries of
        the other ALGO values need no leaves) and the search.  Returns
        (ids, dists, PQ, leaf)."""
        import torch
        if not (_is_torch(Q) and Q.is_cuda):
            raise ValueError("rte_query takes device tensors")
        nq = Q.shape[0]
        dev = Q.device
        k = params.k if params.k > 0 else 1 << (max(params.beam, 1) - 1).bit_length()
        pq = torch.empty((nq, self.d_out), dtype=torch.float32, device=dev)
        leaf = torch.zeros(nq, dtype=torch.int32, device=dev)
        ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
        dists = torch.empty((nq, k), dtype=torch.float32, device=dev)
        qp, _, _ = _ptr(Q, np.float32)
        P = lambda t: C.c_void_p(t.data_ptr())
        self._chk(lib().rte_query(self._h, qp, nq, C.byref(params), P(pq), P(leaf), P(ids),
                                  P(dists)))
        return ids, dists, pq, leaf

    def search_batched(self, PQ, leaf, k_per_q, beam_per_q, params: Params, k_stride=None,
                       out=None, beam_bound=None):
        """Batched launch (an uneven batch: a captured hipGraph of the batch
        preparation and the search; a uniform batch: direct launches).  With
        per-query k_per_q and beam_per_q the batch is uneven; beam_bound is its largest effective
        beam (selects the kernel; default: computed from the arrays).  With
        both None the batch is uniform: every query uses params.k and
        params.beam (results equal search())."""
        pq, w, keep = _ptr(PQ, np.float32)
        nq, qs = PQ.shape
        uniform = k_per_q is None and beam_per_q is None
        if (k_per_q is None) != (beam_per_q is None):
            raise ValueError("k_per_q and beam_per_q: both or neither")
        params = Params(*[getattr(params, f) for f, _ in Params._fields_])
        if k_stride is None:
            k_stride = params.k if uniform else int(k_per_q.max())
        if not uniform:
            if beam_bound is None:
                beam_bound = max(int(beam_per_q.max()), int(k_per_q.max()), 32)
            params.beam = 1 << (int(beam_bound) - 1).bit_length()
        if out is None:
            if w == RTE_DEVICE:
                import torch
                ids = torch.full((nq, k_stride), -1, dtype=torch.int32, device=PQ.device)
                dists = torch.full((nq, k_stride), float("nan"), dtype=torch.float32,
                                   device=PQ.device)
            else:
                ids = np.full((nq, k_stride), -1, np.int32)
                dists = np.full((nq, k_stride), np.nan, np.float32)
        else:
            ids, dists = out
        pl, _, _ = _ptr(leaf, np.int32) if leaf is not None else (None, w, None)
        pk = pb = None
        if not uniform:
            pk, wk, k1 = _ptr(k_per_q, np.int32)
            pb, wb, k2 = _ptr(beam_per_q, np.int32)
            if wk != w or wb != w:
                raise ValueError("k_per_q / beam_per_q must live with the queries")
        pi, _, _ = _ptr(ids, np.int32)
        pd, _, _ = _ptr(dists, np.float32)
        self._chk(lib().rte_search_batched(self._h, pq, nq, qs, pl, pk, pb, C.byref(params),
                                           int(k_stride), pi, pd, w))
        return ids, dists

    def search_batched_noalloc(self, PQ, leaf, k_per_q, beam_per_q, params: Params, k_stride,
                               out):
        """rte_search_batched_noalloc: the batched search with no allocation
        (device tensors only, outputs given)."""
        ids, dists = out
        P = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        nq, qs = PQ.shape
        if (k_per_q is None) != (beam_per_q is None):
            raise ValueError("k_per_q and beam_per_q: both or neither")
        params = Params(*[getattr(params, f) for f, _ in Params._fields_])
        if beam_per_q is not None:  # the batch's beam bound, as search_batched
            bound = max(int(beam_per_q.max()), int(k_per_q.max()), 32)
            params.beam = 1 << (bound - 1).bit_length()
        self._chk(lib().rte_search_batched_noalloc(self._h, P(PQ), int(nq), int(qs), P(leaf),
                                                   P(k_per_q), P(beam_per_q), C.byref(params),
                                                   int(k_stride), P(ids), P(dists)))
        return ids, dists

    # ---- learned entry predictor (device tensors)
    def mlp_create(self, d_in: int, h1: int, h2: int, n_out: int, seed: int = 1):
        self._chk(lib().rte_mlp_create(self._h, int(d_in), int(h1), int(h2), int(n_out),
                                       C.c_uint64(seed)))
        self.mlp_dims = (int(d_in), int(h1), int(h2), int(n_out))

    def mlp_set_input_norm(self, mean, inv_std):
        m = np.ascontiguousarray(mean, np.float32)
        v = np.ascontiguousarray(inv_std, np.float32)
        self._chk(lib().rte_mlp_set_input_norm(self._h, m.ctypes.data_as(C.c_void_p),
                                               v.ctypes.data_as(C.c_void_p)))
        self.mlp_norm = (m.copy(), v.copy())

    def mlp_get(self):
        d = self.mlp_dims
        W = [np.empty((d[i], d[i + 1]), np.float32) for i in range(3)]
        b = [np.empty(d[i + 1], np.float32) for i in range(3)]
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        self._chk(lib().rte_mlp_get(self._h, P(W[0]), P(b[0]), P(W[1]), P(b[1]), P(W[2]),
                                    P(b[2])))
        return W, b

    def mlp_set(self, W, b):
        W = [np.ascontiguousarray(w, np.float32) for w in W]
        b = [np.ascontiguousarray(x, np.float32) for x in b]
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        self._chk(lib().rte_mlp_set(self._h, P(W[0]), P(b[0]), P(W[1]), P(b[1]), P(W[2]),
                                    P(b[2])))

    def mlp_grad(self, X, y):
        """Gradients of the mean cross-entropy on one device batch (host f32)."""
        d = self.mlp_dims
        gW = [np.empty((d[i], d[i + 1]), np.float32) for i in range(3)]
        gb = [np.empty(d[i + 1], np.float32) for i in range(3)]
        loss = C.c_float()
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        self._chk(lib().rte_mlp_grad(self._h, C.c_void_p(X.data_ptr()), C.c_void_p(y.data_ptr()),
                                     int(X.shape[0]), P(gW[0]), P(gb[0]), P(gW[1]), P(gb[1]),
                                     P(gW[2]), P(gb[2]), C.byref(loss)))
        return gW, gb, loss.value

    def mlp_train_step(self, X, y, lr: float, beta1: float = 0.9, want_loss: bool = True):
        """One Adam step; returns the batch's mean loss (a host sync) or None."""
        loss = C.c_float()
        self._chk(lib().rte_mlp_train_step(self._h, C.c_void_p(X.data_ptr()),
                                           C.c_void_p(y.data_ptr()), int(X.shape[0]),
                                           float(lr), float(beta1),
                                           C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def mlp_predict(self, X, fallback=None, min_conf: float = 0.5, out=None, conf=None):
        import torch
        n = X.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=X.device)
        fb = C.c_void_p(fallback.data_ptr()) if fallback is not None else None
        cp = C.c_void_p(conf.data_ptr()) if conf is not None else None
        self._chk(lib().rte_mlp_predict(self._h, C.c_void_p(X.data_ptr()), int(n), fb,
                                        float(min_conf), C.c_void_p(out.data_ptr()), cp))
        return out

    def pca_fit(self, X, d=None):
        """(R d x d, mu d, eigenvalues d) of the first d columns of the rows of
        X (host or device; the row stride is X's column count)."""
        px, w, keep = _ptr(X, np.float32)
        n, stride = X.shape
        d = stride if d is None else d
        R = np.empty((d, d), np.float32)
        mu = np.empty(d, np.float32)
        eig = np.empty(d, np.float64)
        self._chk(lib().rte_pca_fit(self._h, px, int(n), int(d), int(stride),
                                    R.ctypes.data_as(C.c_void_p), mu.ctypes.data_as(C.c_void_p),
                                    eig.ctypes.data_as(C.c_void_p), w))
        return R, mu, eig

    def pca_rotate(self, X, R, mu, d_out=None):
        """(X - mu) R[:, :d_out] in FP64, stored FP32 (torch tensor on the
        device for a device X, numpy otherwise)."""
        px, w, keep = _ptr(X, np.float32)
        n, d = X.shape
        R = np.ascontiguousarray(np.asarray(R, np.float32))
        mu = np.ascontiguousarray(np.asarray(mu, np.float32))
        d_out = R.shape[1] if d_out is None else int(d_out)
        R = np.ascontiguousarray(R[:, :d_out])
        if w == RTE_DEVICE:
            import torch
            out = torch.empty((n, d_out), dtype=torch.float32, device=X.device)
            po = C.c_void_p(out.data_ptr())
        else:
            out = np.empty((n, d_out), np.float32)
            po = out.ctypes.data_as(C.c_void_p)
        self._chk(lib().rte_pca_rotate(self._h, px, int(n), int(d), int(d),
                                       R.ctypes.data_as(C.c_void_p),
                                       mu.ctypes.data_as(C.c_void_p), d_out, po, w))
        return out

    def build_kdtree(self, X, max_node: int, max_leaves: int | None = None):
        """GPU kd-tree over the first three columns of X (host or device):
        (unexpanded boxes [n_leaf,6], pids [n_leaf,max_node]) as numpy
        arrays for a host X, device tensors for a device X."""
        px, w, keep = _ptr(X, np.float32)
        n, stride = X.shape
        if max_leaves is None:
            max_leaves = max(16, 4 * (n // max(1, max_node)) + 16)
        nl = C.c_int()
        if w == RTE_DEVICE:
            import torch
            aabbs = torch.empty((max_leaves, 6), dtype=torch.float32, device=X.device)
            pids = torch.empty((max_leaves, max_node), dtype=torch.int32, device=X.device)
            pa, pp = C.c_void_p(aabbs.data_ptr()), C.c_void_p(pids.data_ptr())
        else:
            aabbs = np.zeros((max_leaves, 6), np.float32)
            pids = np.zeros((max_leaves, max_node), np.int32)
            pa, pp = aabbs.ctypes.data_as(C.c_void_p), pids.ctypes.data_as(C.c_void_p)
        self._chk(lib().rte_build_kdtree_gpu(self._h, px, int(n), int(stride), int(max_node),
                                             int(max_leaves), pa, pp, C.byref(nl), w))
        return aabbs[: nl.value].clone() if w == RTE_DEVICE else aabbs[: nl.value].copy(), \
            pids[: nl.value].clone() if w == RTE_DEVICE else pids[: nl.value].copy()

    def knn(self, X, k: int, with_dists: bool = False, rows=None):
        """Exact k nearest rows of X to each row of X (itself excluded), by
        (distance, id); X host array or device tensor.  rows = (row0, nrows)
        restricts the output to those rows (still searched against all of X)."""
        px, w, keep = _ptr(X, np.float32)
        n, d = X.shape
        row0, nr = (0, n) if rows is None else (int(rows[0]), int(rows[1]))
        if w == RTE_DEVICE:
            import torch
            ids = torch.empty((nr, k), dtype=torch.int32, device=X.device)
            dists = torch.empty((nr, k), dtype=torch.float32, device=X.device) if with_dists else None
        else:
            ids = np.empty((nr, k), np.int32)
            dists = np.empty((nr, k), np.float32) if with_dists else None
        pi, _, _ = _ptr(ids, np.int32)
        pd = _ptr(dists, np.float32)[0] if dists is not None else None
        if rows is None:
            self._chk(lib().rte_knn(self._h, px, int(n), int(d), int(d), int(k), pi, pd, w))
        else:
            self._chk(lib().rte_knn_rows(self._h, px, int(n), int(d), int(d), row0, nr, int(k),
                                         pi, pd, w))
        return (ids, dists) if with_dists else ids

    def graph_from_knn(self, knn, degree: int, rev_rank: int | None = None):
        """Rank-pruned graph with reverse edges from kNN lists (host array or
        device tensor), rows padded with np."""
        pk, w, keep = _ptr(knn, np.int32)
        n, kf = knn.shape
        rev_rank = degree if rev_rank is None else rev_rank
        if w == RTE_DEVICE:
            import torch
            g = torch.empty((n, degree), dtype=torch.int32, device=knn.device)
        else:
            g = np.empty((n, degree), np.int32)
        pg, _, _ = _ptr(g, np.int32)
        self._chk(lib().rte_graph_from_knn(self._h, pk, int(n), int(kf), int(degree),
                                           int(rev_rank), pg, w))
        return g

    def graph_add_entry_links(self, graph, r: int, window: int = 0, seed: int = 1):
        """rte_graph_add_entry_links on a graph (host array: a modified copy
        is returned; device tensor: changed in place and returned).  Needs
        set_leaves first."""
        if _is_torch(graph) and graph.is_cuda:
            g = graph
        else:
            g = np.array(graph, np.int32, copy=True)
        pg, w, keep = _ptr(g, np.int32)
        n, deg = g.shape
        self._chk(lib().rte_graph_add_entry_links(self._h, pg, int(n), int(deg), int(r),
                                                  int(window), C.c_uint64(seed), w))
        return g

    def set_rerank_base(self, base):
        """Full-precision vectors for rte_rerank (np x d f32, host or device)."""
        p, w, keep = _ptr(base, np.float32)
        self._rr_keep = keep if w == RTE_DEVICE else None
        self._chk(lib().rte_set_rerank_base(self._h, p, int(base.shape[0]), int(base.shape[1]),
                                            int(base.shape[1]), w))

    def rerank(self, Q, cand, k: int = 10, out=None):
        """(ids, dists) of the k best candidates per query by full-precision
        distance; Q and cand host arrays or device tensors."""
        pq, w, keep = _ptr(Q, np.float32)
        nq, qs = Q.shape
        n_cand = cand.shape[1]
        if out is None:
            if w == RTE_DEVICE:
                import torch
                ids = torch.empty((nq, k), dtype=torch.int32, device=Q.device)
                dists = torch.empty((nq, k), dtype=torch.float32, device=Q.device)
            else:
                ids = np.empty((nq, k), np.int32)
                dists = np.empty((nq, k), np.float32)
        else:
            ids, dists = out
        pc, wc, keep2 = _ptr(cand, np.int32)
        if wc != w:
            raise ValueError("cand must live with the queries")
        pi, _, _ = _ptr(ids, np.int32)
        pd, _, _ = _ptr(dists, np.float32)
        self._chk(lib().rte_rerank(self._h, pq, nq, qs, pc, int(n_cand), int(k), pi, pd, w))
        return ids, dists

    def predict_entry(self, PQ, min_conf: float = 0.5, out=None, conf=None):
        """Entry leaves of device PCA queries: the network's leaf, or the
        kd-tree leaf below min_conf (the kd-tree selection runs beside the
        network's hidden layers, in the same launch)."""
        import torch
        nq, qs = PQ.shape
        if out is None:
            out = torch.empty(nq, dtype=torch.int32, device=PQ.device)
        cp = C.c_void_p(conf.data_ptr()) if conf is not None else None
        self._chk(lib().rte_predict_entry(self._h, C.c_void_p(PQ.data_ptr()), int(nq),
                                          int(PQ.stride(0)), float(min_conf),
                                          C.c_void_p(out.data_ptr()), cp))
        return out

    def autotune_batched(self, PQ, leaf, k_per_q, beam_per_q, params: Params, k_stride=None,
                         bloom_words=(0,)):
        """Fastest launch shape (wpq, wg_waves, queue, grid, order,
        bloom_words) for this batch (device tensors; k_per_q = beam_per_q =
        None for a uniform batch): every shape timed over three launches,
        the close finalists again in interleaved rounds.  Returns (best
        Params, its time in ms)."""
        import torch
        nq, qs = PQ.shape
        uniform = k_per_q is None and beam_per_q is None
        if k_stride is None:
            k_stride = params.k if uniform else int(k_per_q.max())
        ids = torch.empty((nq, k_stride), dtype=torch.int32, device=PQ.device)
        dists = torch.empty((nq, k_stride), dtype=torch.float32, device=PQ.device)
        if not uniform and (params.beam <= 0 or params.beam < int(beam_per_q.max())):
            params = Params(*[getattr(params, f) for f, _ in Params._fields_])
            params.beam = 1 << (max(int(beam_per_q.max()), int(k_per_q.max()), 32) - 1).bit_length()
        bw = (C.c_int * len(bloom_words))(*[int(x) for x in bloom_words])
        best = Params()
        ms = C.c_float()
        P = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        self._chk(lib().rte_autotune_batched(self._h, P(PQ), nq, qs, P(leaf), P(k_per_q),
                                             P(beam_per_q), C.byref(params), int(k_stride),
                                             P(ids), P(dists), bw, len(bloom_words),
                                             C.byref(best), C.byref(ms)))
        return best, ms.value

"""Synthetic workloads, PCA, exact kNN graphs, ground truth, leaf tables and
the reference's on-disk formats.

Workloads (seeded numpy PCG64; seed = 20261202 + 100*cfg + {0 centres,
1 base, 2 queries, 3 training queries, 4 k/beam assignment}):

  1  100k x 128, 64 centres U[0,100], sigma 8; 1k queries; degree-32 exact kNN
  2  1M x 128, 1000 centres U[0,255], sigma 20, rounded and clamped to
     [0,255]; 10k queries; exact 64-NN + reverse edges, rank-pruned to 32
  3  1M x 96 unit vectors, 2000 normalised-Gaussian centres, sigma 0.15;
     20k queries with k in {1,10,100}, beam in {32,64,128,256}
  4  1M x 960, rank-64 mixture of 500 centres plus N(0, 0.05^2); 10k
     queries, 100k training queries; degree 48 stored at stride 64
  5  10M x 128 integer mixture of 10k centres; 100k queries, k in {10,100}

`scale` shrinks n (and queries) for tests; the distributions are unchanged.
"""
from __future__ import annotations

import os
import struct

import numpy as np

SEED0 = 20261202

CONFIGS = {
    1: dict(n=100_000, nq=1_000, d=128, degree=32, k_fwd=32, n_leaf=1024, point_ratio=0.00128,
            k=10, beam=64),
    2: dict(n=1_000_000, nq=10_000, d=128, degree=32, k_fwd=64, n_leaf=4096,
            point_ratio=0.000256, k=10, beam=64),
    3: dict(n=1_000_000, nq=20_000, d=96, degree=32, k_fwd=64, n_leaf=4096,
            point_ratio=0.000256, k=10, beam=64),
    4: dict(n=1_000_000, nq=10_000, d=960, degree=48, k_fwd=64, n_leaf=4096,
            point_ratio=0.000256, k=10, beam=64, n_train=100_000),
    5: dict(n=10_000_000, nq=100_000, d=128, degree=32, k_fwd=64, n_leaf=16384,
            point_ratio=0.0001024, k=10, beam=64),
}


def seed(cfg: int, what: int) -> int:
    return SEED0 + 100 * cfg + what


def _rng(cfg, what):
    return np.random.Generator(np.random.PCG64(seed(cfg, what)))


def _mixture_int(rng_c, rng_x, n, d, n_centres, lo, hi, sigma, centres=None, chunk=1 << 18):
    if centres is None:
        centres = rng_c.uniform(lo, hi, size=(n_centres, d))
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        z = rng_x.integers(0, n_centres, size=m)
        x = centres[z] + rng_x.normal(0.0, sigma, size=(m, d))
        out[s:s + m] = np.clip(np.rint(x), lo, hi)
    return out, centres


def _mixture_f(rng_c, rng_x, n, d, n_centres, lo, hi, sigma, centres=None, chunk=1 << 18):
    if centres is None:
        centres = rng_c.uniform(lo, hi, size=(n_centres, d))
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        z = rng_x.integers(0, n_centres, size=m)
        out[s:s + m] = centres[z] + rng_x.normal(0.0, sigma, size=(m, d))
    return out, centres


def _sphere(rng_c, rng_x, n, d, n_centres, sigma, centres=None, chunk=1 << 18):
    if centres is None:
        c = rng_c.normal(size=(n_centres, d))
        centres = c / np.linalg.norm(c, axis=1, keepdims=True)
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        z = rng_x.integers(0, n_centres, size=m)
        x = centres[z] + rng_x.normal(0.0, sigma, size=(m, d))
        out[s:s + m] = x / np.linalg.norm(x, axis=1, keepdims=True)
    return out, centres


def _lowrank(rng_c, rng_x, n, d, n_centres, rank, s_lat, noise, params=None, chunk=1 << 16):
    if params is None:
        Wm, _ = np.linalg.qr(rng_c.normal(size=(d, rank)))
        mu = rng_c.normal(0.0, 1.0, size=(n_centres, rank))
        params = (Wm, mu)
    Wm, mu = params
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        z = rng_x.integers(0, n_centres, size=m)
        lat = mu[z] + s_lat * rng_x.normal(size=(m, rank))
        out[s:s + m] = lat @ Wm.T + rng_x.normal(0.0, noise, size=(m, d))
    return out, params


def generate(cfg: int, scale: float = 1.0, n: int | None = None, nq: int | None = None,
             n_train: int | None = None):
    """Returns dict(base, queries[, train]) float32 for workload `cfg`."""
    c = CONFIGS[cfg]
    n = n if n is not None else max(1000, int(c["n"] * scale))
    nq = nq if nq is not None else max(10, int(c["nq"] * min(1.0, scale * 10)))
    rc, rb, rq, rt = _rng(cfg, 0), _rng(cfg, 1), _rng(cfg, 2), _rng(cfg, 3)
    out = {}
    if cfg == 1:
        base, cen = _mixture_f(rc, rb, n, 128, 64, 0.0, 100.0, 8.0)
        out["queries"], _ = _mixture_f(None, rq, nq, 128, 64, 0, 0, 8.0, centres=cen)
    elif cfg in (2, 5):
        ncen = 1000 if cfg == 2 else 10_000
        base, cen = _mixture_int(rc, rb, n, 128, ncen, 0.0, 255.0, 20.0)
        out["queries"], _ = _mixture_int(None, rq, nq, 128, ncen, 0.0, 255.0, 20.0, centres=cen)
    elif cfg == 3:
        base, cen = _sphere(rc, rb, n, 96, 2000, 0.15)
        out["queries"], _ = _sphere(None, rq, nq, 96, 2000, 0.15, centres=cen)
    elif cfg == 4:
        base, prm = _lowrank(rc, rb, n, 960, 500, 64, 0.25, 0.05)
        out["queries"], _ = _lowrank(None, rq, nq, 960, 500, 64, 0.25, 0.05, params=prm)
        nt = n_train if n_train is not None else max(1000, int(c["n_train"] * scale))
        out["train"], _ = _lowrank(None, rt, nt, 960, 500, 64, 0.25, 0.05, params=prm)
    else:
        raise ValueError(cfg)
    if cfg != 4 and n_train:
        gen = {1: lambda: _mixture_f(None, rt, n_train, 128, 64, 0, 0, 8.0, centres=cen),
               2: lambda: _mixture_int(None, rt, n_train, 128, 1000, 0.0, 255.0, 20.0,
                                       centres=cen),
               3: lambda: _sphere(None, rt, n_train, 96, 2000, 0.15, centres=cen),
               5: lambda: _mixture_int(None, rt, n_train, 128, 10_000, 0.0, 255.0, 20.0,
                                       centres=cen)}[cfg]
        out["train"], _ = gen()
    out["base"] = base
    return out


def extra_queries(cfg: int, n: int, what: int = 3):
    """n more queries of workload cfg from stream `what` (3: training
    queries), drawn around the same centres as the base."""
    rc, rx = _rng(cfg, 0), _rng(cfg, what)
    if cfg == 1:
        cen = rc.uniform(0.0, 100.0, size=(64, 128))
        return _mixture_f(None, rx, n, 128, 64, 0, 0, 8.0, centres=cen)[0]
    if cfg in (2, 5):
        ncen = 1000 if cfg == 2 else 10_000
        cen = rc.uniform(0.0, 255.0, size=(ncen, 128))
        return _mixture_int(None, rx, n, 128, ncen, 0.0, 255.0, 20.0, centres=cen)[0]
    if cfg == 3:
        c = rc.normal(size=(2000, 96))
        cen = c / np.linalg.norm(c, axis=1, keepdims=True)
        return _sphere(None, rx, n, 96, 2000, 0.15, centres=cen)[0]
    if cfg == 4:
        Wm, _ = np.linalg.qr(rc.normal(size=(960, 64)))
        mu = rc.normal(0.0, 1.0, size=(500, 64))
        return _lowrank(None, rx, n, 960, 500, 64, 0.25, 0.05, params=(Wm, mu))[0]
    raise ValueError(cfg)


def k_beam_assignment(cfg: int, nq: int, ks=(1, 10, 100), beams=(32, 64, 128, 256)):
    """Uniform per-query k and beam draws (seed +4); beam_eff = max(beam,
    pow2ceil(k))."""
    r = _rng(cfg, 4)
    k = np.asarray(ks, np.int32)[r.integers(0, len(ks), size=nq)]
    b = np.asarray(beams, np.int32)[r.integers(0, len(beams), size=nq)]
    kp2 = 1 << np.ceil(np.log2(np.maximum(k, 1))).astype(np.int32)
    return k, np.maximum(b, kp2).astype(np.int32)


# ---- PCA (pca.cpp:4-58, 122-139) --------------------------------------
def pca_fit(base: np.ndarray, chunk: int = 1 << 18):
    """FP64 mean and covariance, eigenvectors sorted by descending
    eigenvalue.  Returns (R[d,d] float32, column j = eigenvector j;
    mu[d] float32; eigenvalues float64)."""
    n, d = base.shape
    s = np.zeros(d)
    for i in range(0, n, chunk):
        s += base[i:i + chunk].astype(np.float64).sum(0)
    mu = s / n
    C = np.zeros((d, d))
    for i in range(0, n, chunk):
        x = base[i:i + chunk].astype(np.float64) - mu
        C += x.T @ x
    C /= max(1, n - 1)
    w, V = np.linalg.eigh(C)
    order = np.argsort(-w, kind="stable")
    V = V[:, order]
    # fix each eigenvector's sign (largest component positive) for determinism
    sgn = np.sign(V[np.abs(V).argmax(0), np.arange(d)])
    V = V * np.where(sgn == 0, 1.0, sgn)
    return V.astype(np.float32), mu.astype(np.float32), w[order]


def pca_rotate(x: np.ndarray, R: np.ndarray, mu: np.ndarray, chunk: int = 1 << 18):
    """(x - mu) R in FP64, stored FP32 (pca.cpp:122-139)."""
    out = np.empty((x.shape[0], R.shape[1]), np.float32)
    R64 = R.astype(np.float64)
    mu64 = mu.astype(np.float64)
    for i in range(0, x.shape[0], chunk):
        out[i:i + chunk] = (x[i:i + chunk].astype(np.float64) - mu64) @ R64
    return out


# ---- exact kNN and graphs ---------------------------------------------
def knn_exact(X: np.ndarray, Q: np.ndarray, k: int, exclude_self: bool = False,
              block: int = 2048):
    """FP64 exact k nearest neighbours of each row of Q among rows of X, ties
    broken by id.  (CPU; for test-sized inputs.)"""
    X64 = X.astype(np.float64)
    xn = (X64 * X64).sum(1)
    nq = Q.shape[0]
    ids = np.empty((nq, k), np.int32)
    dist = np.empty((nq, k), np.float64)
    kk = min(X.shape[0], k + 1 + 8)
    for s in range(0, nq, block):
        q = Q[s:s + block].astype(np.float64)
        d2 = xn[None, :] - 2.0 * q @ X64.T + (q * q).sum(1)[:, None]
        if exclude_self:
            rows = np.arange(q.shape[0])
            d2[rows, s + rows] = np.inf
        part = np.argpartition(d2, kk - 1, axis=1)[:, :kk]
        pd = np.take_along_axis(d2, part, 1)
        o = np.lexsort((part, pd), axis=1)
        ids[s:s + block] = np.take_along_axis(part, o, 1)[:, :k]
        dist[s:s + block] = np.take_along_axis(pd, o, 1)[:, :k]
    return ids, dist


def recall(ids: np.ndarray, gt: np.ndarray, at: int = 10, block: int = 16384) -> float:
    """Recall@`at` as `graph_helper.cu:64-121` counts it: for each query, each
    of the first min(gt_k, at) ground-truth ids found anywhere among the k
    returned ids is a hit; the total is divided by nq * at."""
    ids = np.asarray(ids)
    gt = np.asarray(gt)[:, :at]
    nq = ids.shape[0]
    hits = 0
    for s in range(0, nq, block):
        hits += int((gt[s:s + block, :, None] == ids[s:s + block, None, :]).any(-1).sum())
    return hits / nq / at


def graph_from_knn(knn: np.ndarray, degree: int, rev_rank: int | None = None) -> np.ndarray:
    """Rank-pruned kNN graph with reverse edges (numpy).

    Candidate edges of u: forward u->knn[u][j] with rank j, and reverse
    u->v for every v with knn[v][j] == u, j < rev_rank, with rank j.
    Duplicates keep their smallest rank (forward first); each row keeps the
    `degree` candidates of smallest (rank, reverse, id) and is padded with n.
    With rev_rank = 0 this is the plain kNN graph."""
    n, kf = knn.shape
    rev_rank = degree if rev_rank is None else rev_rank
    u_f = np.repeat(np.arange(n, dtype=np.int64), kf)
    v_f = knn.reshape(-1).astype(np.int64)
    r_f = np.tile(np.arange(kf, dtype=np.int64), n)
    parts_u, parts_v, parts_r, parts_t = [u_f], [v_f], [r_f], [np.zeros_like(u_f)]
    if rev_rank > 0:
        kr = min(rev_rank, kf)
        sub = knn[:, :kr]
        parts_u.append(sub.reshape(-1).astype(np.int64))
        parts_v.append(np.repeat(np.arange(n, dtype=np.int64), kr))
        parts_r.append(np.tile(np.arange(kr, dtype=np.int64), n))
        parts_t.append(np.ones(n * kr, np.int64))
    u = np.concatenate(parts_u)
    v = np.concatenate(parts_v)
    r = np.concatenate(parts_r)
    t = np.concatenate(parts_t)
    keep = u != v
    u, v, r, t = u[keep], v[keep], r[keep], t[keep]
    # dedupe (u,v) keeping min (rank, type)
    o = np.lexsort((t, r, v, u))
    u, v, r, t = u[o], v[o], r[o], t[o]
    first = np.ones(u.size, bool)
    first[1:] = (u[1:] != u[:-1]) | (v[1:] != v[:-1])
    u, v, r, t = u[first], v[first], r[first], t[first]
    o = np.lexsort((v, t, r, u))
    u, v = u[o], v[o]
    start = np.searchsorted(u, np.arange(n))
    pos = np.arange(u.size) - start[u]
    sel = pos < degree
    g = np.full((n, degree), n, np.int32)
    g[u[sel], pos[sel]] = v[sel]
    return g


def _splitmix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    z ^= z >> np.uint64(30)
    z *= np.uint64(0xBF58476D1CE4E5B9)
    z ^= z >> np.uint64(27)
    z *= np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    return z


def add_entry_links(g: np.ndarray, pids: np.ndarray, r: int, window: int = 0, seed: int = 1,
                    tries: int = 8, rows=None) -> np.ndarray:
    """Numpy definition of rte_graph_add_entry_links: the last r slots of
    row u become points of the leaves within `window` ids of u's leaf (in
    the pid table's order), candidate t of slot j drawn as
    splitmix64(seed + ((u << 16) | (j << 8) | t) * 0x9E3779B97F4A7C15) mod
    the window's point count, redrawn when equal to u, a kept entry or an
    earlier link, padding (n) after `tries` draws.  A row's links depend on
    that row alone, so `rows` (node ids) returns just those rows, in order."""
    n, degree = g.shape
    n_leaf, cap = pids.shape
    if rows is not None:
        rows = np.asarray(rows, np.int64)
        out = np.array(g[rows], np.int32, copy=True)
    else:
        rows = np.arange(n, dtype=np.int64)
        out = np.array(g, np.int32, copy=True)
    if r == 0:
        return out
    leaf_of = np.full(n, -1, np.int64)
    valid = pids < n
    leaf_of[pids[valid]] = np.repeat(np.arange(n_leaf), cap).reshape(n_leaf, cap)[valid]
    size = valid.sum(1).astype(np.uint64)
    keep = degree - r
    golden = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        for o, u in enumerate(rows.tolist()):
            l = leaf_of[u]
            if l < 0:
                out[o, keep:] = n
                continue
            lo, hi = max(0, l - window), min(n_leaf - 1, l + window)
            total = size[lo:hi + 1].sum()
            links = []
            for j in range(r):
                pick = n
                for t in range(tries):
                    key = np.uint64((u << 16) | (j << 8) | t)
                    idx = int(_splitmix64(np.array([np.uint64(seed) + key * golden]))[0] % total)
                    x = lo
                    while idx >= int(size[x]):
                        idx -= int(size[x])
                        x += 1
                    c = int(pids[x, idx])
                    if c != u and c not in out[o, :keep] and c not in links:
                        pick = c
                        break
                links.append(pick)
            out[o, keep:] = links
    return out


def pad_graph(g: np.ndarray, n: int) -> np.ndarray:
    """Pad rows to a power-of-two stride with n (graph.cu:134)."""
    deg = g.shape[1]
    D = 1 << (deg - 1).bit_length()
    if D == deg:
        return np.ascontiguousarray(g, np.int32)
    out = np.full((g.shape[0], D), n, np.int32)
    out[:, :deg] = g
    return out


# ---- leaves ------------------------------------------------------------
def build_leaves(pca3: np.ndarray, cap: int, expand_ratio: float = 0.2):
    """kd-tree leaves (librte host build): (expanded boxes, unexpanded boxes,
    pid table)."""
    from .api import build_kdtree, expand_aabbs
    boxes, pids = build_kdtree(np.ascontiguousarray(pca3[:, :3], np.float32), cap)
    return expand_aabbs(boxes, expand_ratio), boxes, pids


# ---- file formats (file_read.cu, entry.cu:171-190, pca.cpp:60-139) ------
def write_fvecs(path, x):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    rec = np.empty((n, d + 1), np.float32)
    rec[:, 0] = np.array([d], np.int32).view(np.float32)[0]
    rec[:, 1:] = x
    rec.tofile(path)


def read_fvecs(path):
    raw = np.fromfile(path, np.int32)
    d = int(raw[0])
    return raw.reshape(-1, d + 1)[:, 1:].view(np.float32).copy()


def write_ivecs(path, x):
    x = np.ascontiguousarray(x, np.int32)
    n, d = x.shape
    rec = np.empty((n, d + 1), np.int32)
    rec[:, 0] = d
    rec[:, 1:] = x
    rec.tofile(path)


def read_ivecs(path):
    raw = np.fromfile(path, np.int32)
    d = int(raw[0])
    return raw.reshape(-1, d + 1)[:, 1:].copy()


def read_bvecs(path, limit=100_000_000):
    raw = np.fromfile(path, np.uint8)
    d = int(raw[:4].view(np.int32)[0])
    rows = raw.reshape(-1, d + 4)[:limit]
    return rows[:, 4:].astype(np.float32)


def write_hvecs(path, x):
    """int32 d, then the rows as float32 (file_read.cu:163-182)."""
    x = np.ascontiguousarray(x, np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", x.shape[1]))
        x.tofile(f)


def read_hvecs(path):
    """file_read.cu:163-182: n = (bytes / 4 - 1) / d rows after the header."""
    raw = np.fromfile(path, np.int32)
    d = int(raw[0])
    n = (raw.size - 1) // d
    return raw[1:1 + n * d].view(np.float32).reshape(n, d).copy()


# the hh layout (file_read.cu:82-101) is the same header and rows
write_hh, read_hh = write_hvecs, read_hvecs


def write_txt(path, x):
    """"n d" on the first line, then one comma-separated row per line
    (file_read.cu:103-119 reads this)."""
    x = np.asarray(x, np.float32)
    with open(path, "w") as f:
        f.write(f"{x.shape[0]} {x.shape[1]}\n")
        for row in x:
            f.write(",".join(repr(float(v)) for v in row) + "\n")


def read_txt(path):
    """file_read.cu:103-119: "n d", then n*d floats each followed by an
    optional comma, whitespace anywhere between them."""
    with open(path) as f:
        head = f.readline().split()
        n, d = int(head[0]), int(head[1])
        toks = f.read().replace(",", " ").split()
    if len(toks) < n * d:
        raise ValueError(f"{path}: {len(toks)} values for {n}x{d}")
    return np.array(toks[:n * d], dtype=np.float32).reshape(n, d)


def read_data(path):
    """Base or query vectors by file extension, in the order of
    file_read.cu:16-80 (txt, fvecs, fbin, hvecs, bvecs, hh)."""
    p = str(path)
    for ext, fn in (("txt", read_txt), ("fvecs", read_fvecs), ("fbin", read_fbin),
                    ("hvecs", read_hvecs), ("bvecs", read_bvecs), ("hh", read_hh)):
        if p.endswith(ext):
            return fn(p)
    raise ValueError(f"unknown data file type: {p}")


def write_fbin(path, x):
    x = np.ascontiguousarray(x, np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", x.shape[0], x.shape[1]))
        x.tofile(f)


def read_fbin(path):
    with open(path, "rb") as f:
        n, d = struct.unpack("<ii", f.read(8))
        return np.fromfile(f, np.float32, count=n * d).reshape(n, d)


def write_nsw(path, g):
    np.ascontiguousarray(g, np.int32).tofile(path)


def read_nsw(path, n):
    """Raw int32 rows; degree = file size / (4 n) (file_read.cu:237-255)."""
    size = os.path.getsize(path)
    deg = size // (4 * n)
    return np.fromfile(path, np.int32, count=n * deg).reshape(n, deg)


def write_aabbs(path, boxes_unexpanded, pids):
    """entry.cu:171-180: n_aabbs, aabb_size, boxes (unexpanded), pids."""
    b = np.ascontiguousarray(boxes_unexpanded, np.float32)
    p = np.ascontiguousarray(pids, np.int32)
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", b.shape[0], p.shape[1]))
        b.tofile(f)
        p.tofile(f)


def read_aabbs(path):
    with open(path, "rb") as f:
        n, cap = struct.unpack("<ii", f.read(8))
        b = np.fromfile(f, np.float32, count=n * 6).reshape(n, 6)
        p = np.fromfile(f, np.int32, count=n * cap).reshape(n, cap)
    return b, p

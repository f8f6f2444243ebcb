"""CPU oracle of the rt_entry query path (ctypes binding of liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product package
(rt_entry_amd) never imports it.

The C sources (oracle.c) restate the reference algorithm; see oracle.h for
the per-function reference citations and the parity decisions P1-P7.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MODE_REF = 0
MODE_TUNED = 1
MAXD = np.float32(536870912.0)


class Params(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "k", "beam", "width", "max_iter", "mode", "algo", "bloom_words", "bloom_hashes")]


def build(force: bool = False) -> str:
    """Compile liboracle.so (and the reference's bloom-sizing driver when the
    reference sources are present) with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    return _LIB_PATH


def build_native(out_dir: str) -> str:
    """The CPU baseline's build of the same sources: -O3 -march=native for the
    host it runs on (compiled there, since a native build for another CPU can
    use instructions this one lacks), contraction still off so its results
    equal the portable oracle's.  Returns the library path."""
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "liboracle_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-ffp-contract=off", "-fPIC",
                    "-std=c11", "-shared", "-o", out, os.path.join(_HERE, "oracle.c"), "-lm"],
                   check=True)
    return out


_libs = {}


def lib(path: str | None = None):
    path = path or _LIB_PATH
    if path not in _libs:
        if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        i64 = C.c_int64
        L.orc_hash_raw.argtypes = [C.c_int, C.c_uint32]
        L.orc_hash_raw.restype = C.c_uint32
        L.orc_hash_parameter.argtypes = [C.c_int, P, P, P]
        L.orc_kdtree_build.argtypes = [P, C.c_int, C.c_int, P, C.c_int, P, P]
        L.orc_kdtree_build.restype = C.c_int
        L.orc_find_xm.argtypes = [P, C.c_int, C.c_float, C.c_float]
        L.orc_find_xm.restype = C.c_float
        L.orc_aabb_pids.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
        L.orc_expand_aabbs.argtypes = [P, C.c_int, C.c_float]
        L.orc_leaf_capacity.argtypes = [C.c_int, C.c_float]
        L.orc_leaf_capacity.restype = C.c_int
        L.orc_select_entry.argtypes = [P, C.c_int, P, i64, C.c_int, P]
        L.orc_project_f64.argtypes = [P, i64, C.c_int, P, C.c_int, C.c_int, P, P]
        L.orc_distance.argtypes = [P, P, C.c_int]
        L.orc_distance.restype = C.c_float
        L.orc_search.argtypes = [P, i64, i64, C.c_int, P, C.c_int, P, i64, i64, P,
                                 C.c_int, P, C.POINTER(Params), P, P, P, C.c_int]
        L.orc_search.restype = C.c_int
        L.orc_search_uneven.argtypes = [P, i64, i64, C.c_int, P, C.c_int, P, i64, i64, P,
                                        C.c_int, P, P, P, C.POINTER(Params), C.c_int, P, P,
                                        C.c_int]
        L.orc_search_uneven.restype = C.c_int
        L.orc_sort_desc.argtypes = [P, P, C.c_int, C.c_int]
        L.orc_merge_asc.argtypes = [P, P, C.c_int, C.c_int]
        L.orc_merge_batch.argtypes = [P, P, C.c_int, P, P, C.c_int, C.c_int, C.c_int]
        L.orc_rerank_distance.argtypes = [P, P, C.c_int]
        L.orc_rerank_distance.restype = C.c_float
        L.orc_rerank.argtypes = [P, i64, i64, C.c_int, P, i64, i64, P, C.c_int, C.c_int, P, P,
                                 C.c_int]
        L.orc_recall.argtypes = [P, i64, C.c_int, P, C.c_int, C.c_int]
        L.orc_recall.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        _libs[path] = L
    return _libs[path]


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def hash_raw(h: int, x: int) -> int:
    return int(lib().orc_hash_raw(h, x & 0xFFFFFFFF))


def hash_parameter(nc: int):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    lib().orc_hash_parameter(nc, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def leaf_capacity(np_: int, point_ratio: float) -> int:
    return int(lib().orc_leaf_capacity(int(np_), float(point_ratio)))


def find_xm(vals, lo: float, hi: float) -> float:
    v = _c(vals, np.float32)
    return float(lib().orc_find_xm(_p(v), v.size, lo, hi))


def kdtree_build(pts3, max_node: int, max_leaves: int | None = None):
    """Returns (aabbs[n_leaf,6] unexpanded, belong[n]); raises on split failure."""
    pts = _c(pts3, np.float32)
    n = pts.shape[0]
    if max_leaves is None:
        max_leaves = max(1, 4 * (n // max(1, max_node)) + 16)
    aabbs = np.zeros((max_leaves, 6), np.float32)
    belong = np.full(n, -1, np.int32)
    nl = C.c_int()
    rc = lib().orc_kdtree_build(_p(pts), n, int(max_node), _p(aabbs), max_leaves, _p(belong),
                                C.byref(nl))
    if rc != 0:
        raise RuntimeError(f"orc_kdtree_build failed rc={rc}")
    return aabbs[: nl.value].copy(), belong


def aabb_pids(belong, n_leaf: int, cap: int):
    b = _c(belong, np.int32)
    pids = np.empty((n_leaf, cap), np.int32)
    lib().orc_aabb_pids(_p(b), b.size, n_leaf, cap, _p(pids))
    return pids


def expand_aabbs(aabbs, ratio: float):
    a = np.array(aabbs, dtype=np.float32, copy=True, order="C")
    lib().orc_expand_aabbs(_p(a), a.shape[0], float(ratio))
    return a


def select_entry(aabbs_expanded, pq):
    a = _c(aabbs_expanded, np.float32)
    q = _c(pq, np.float32)
    hits = np.empty(q.shape[0], np.int32)
    lib().orc_select_entry(_p(a), a.shape[0], _p(q), q.shape[0], q.shape[1], _p(hits))
    return hits


def project_f64(Q, R, mu, d_out: int):
    Qc = _c(Q, np.float32)
    Rc = _c(R, np.float32)
    m = _c(mu, np.float32)
    out = np.empty((Qc.shape[0], d_out), np.float64)
    lib().orc_project_f64(_p(Qc), Qc.shape[0], Qc.shape[1], _p(Rc), Rc.shape[1], d_out, _p(m),
                          _p(out))
    return out


def distance(p, q) -> np.float32:
    a = _c(p, np.float32)
    b = _c(q, np.float32)
    return np.float32(lib().orc_distance(_p(a), _p(b), a.size))


def sort_desc(d, ids, mode=MODE_REF):
    dd = np.array(d, np.float32, copy=True)
    ii = np.array(ids, np.int32, copy=True)
    lib().orc_sort_desc(_p(dd), _p(ii), dd.size, mode)
    return dd, ii


def merge_asc(d, ids, mode=MODE_REF):
    dd = np.array(d, np.float32, copy=True)
    ii = np.array(ids, np.int32, copy=True)
    lib().orc_merge_asc(_p(dd), _p(ii), dd.size, mode)
    return dd, ii


def merge_batch(ld, li, bd, bi, mode=MODE_REF, skip_max=False):
    """One batch (bd, bi) into the beam (ld, li) as the search does it:
    returns the new beam and the batch after its sort and exchanges."""
    L = [np.array(ld, np.float32, copy=True), np.array(li, np.int32, copy=True)]
    B = [np.array(bd, np.float32, copy=True), np.array(bi, np.int32, copy=True)]
    lib().orc_merge_batch(_p(L[0]), _p(L[1]), L[0].size, _p(B[0]), _p(B[1]), B[0].size, mode,
                          int(skip_max))
    return L[0], L[1], B[0], B[1]


def default_bloom(mode: int, beam: int):
    """Filter size: the reference's fixed 512 words x 3 hashes
    (graph_search.cuh:4-5) or the auto_tune_bloom.h sizing for TUNED."""
    if mode == MODE_REF:
        return 512, 3
    hl, _, hh = hash_parameter(beam)
    return hl, hh


def make_params(k=10, beam=64, width=1, max_iter=100, mode=MODE_REF, algo=1,
                bloom_words=0, bloom_hashes=0) -> Params:
    if bloom_words == 0:
        bloom_words, bloom_hashes = default_bloom(mode, beam)
    return Params(k, beam, width, max_iter, mode, algo, bloom_words, bloom_hashes)


def search(base, graph, queries, pids, hits, params: Params, n_threads: int = 0,
           with_stats: bool = False, lib_path: str | None = None):
    """Oracle beam search.  graph must already have a power-of-two row stride
    (padded with np).  Returns ids[nq,k], dists[nq,k] (and stats[nq,3] =
    distance evaluations, graph rows read, iterations).  lib_path selects
    another build of the same sources (the CPU baseline's native build)."""
    b = _c(base, np.float32)
    g = _c(graph, np.int32)
    q = _c(queries, np.float32)
    pid = _c(pids, np.int32) if pids is not None else np.zeros((1, 1), np.int32)
    nq = q.shape[0]
    h = _c(hits, np.int32) if hits is not None else np.zeros(nq, np.int32)
    D = g.shape[1]
    shift = D.bit_length() - 1
    if (1 << shift) != D:
        raise ValueError("graph row stride must be a power of two")
    ids = np.empty((nq, params.k), np.int32)
    dists = np.empty((nq, params.k), np.float32)
    st = np.empty((nq, 3), np.int64) if with_stats else None
    rc = lib(lib_path).orc_search(_p(b), b.shape[0], b.shape[1], q.shape[1] if q.shape[1] < b.shape[1]
                          else b.shape[1], _p(g), shift, _p(q), nq, q.shape[1], _p(pid),
                          pid.shape[1], _p(h), C.byref(params), _p(ids), _p(dists),
                          _p(st) if st is not None else None, int(n_threads))
    if rc != 0:
        raise RuntimeError(f"orc_search failed rc={rc}")
    if with_stats:
        return ids, dists, st
    return ids, dists


def search_uneven(base, graph, queries, pids, hits, k_per_q, beam_per_q, params: Params,
                  n_threads: int = 0):
    b = _c(base, np.float32)
    g = _c(graph, np.int32)
    q = _c(queries, np.float32)
    pid = _c(pids, np.int32)
    nq = q.shape[0]
    h = _c(hits, np.int32)
    kq = _c(k_per_q, np.int32)
    bq = _c(beam_per_q, np.int32)
    kmax = int(kq.max())
    D = g.shape[1]
    shift = D.bit_length() - 1
    ids = np.full((nq, kmax), -1, np.int32)
    dists = np.full((nq, kmax), np.nan, np.float32)
    rc = lib().orc_search_uneven(_p(b), b.shape[0], b.shape[1], min(q.shape[1], b.shape[1]),
                                 _p(g), shift, _p(q), nq, q.shape[1], _p(pid), pid.shape[1],
                                 _p(h), _p(kq), _p(bq), C.byref(params), kmax, _p(ids),
                                 _p(dists), int(n_threads))
    if rc != 0:
        raise RuntimeError(f"orc_search_uneven failed rc={rc}")
    return ids, dists


def rerank_distance(p, q) -> float:
    """graph_helper.cu:161-190 distance of two rows (d % 8 == 0)."""
    a = _c(p, np.float32)
    b = _c(q, np.float32)
    return float(lib().orc_rerank_distance(_p(a), _p(b), a.shape[0]))


def rerank(base, queries, cand, k: int, n_threads: int = 0):
    """Full-precision rerank of candidate ids (graph_helper.cu:146-213):
    (ids, dists) of the k best candidates by (dist, id) per query."""
    b = _c(base, np.float32)
    qv = _c(queries, np.float32)
    c = _c(cand, np.int32)
    nq, n_cand = c.shape
    ids = np.empty((nq, k), np.int32)
    dists = np.empty((nq, k), np.float32)
    rc = lib().orc_rerank(_p(b), b.shape[0], b.shape[1], b.shape[1], _p(qv), nq, qv.shape[1],
                          _p(c), n_cand, int(k), _p(ids), _p(dists),
                          int(n_threads) or num_threads())
    if rc != 0:
        raise ValueError("rerank arguments (d % 8 == 0, 1 <= k <= candidates)")
    return ids, dists


def recall(ids, gt, at: int = 10) -> float:
    i = _c(ids, np.int32)
    g = _c(gt, np.int32)
    return float(lib().orc_recall(_p(i), i.shape[0], i.shape[1], _p(g), g.shape[1], at))


def num_threads() -> int:
    return int(lib().orc_num_threads())

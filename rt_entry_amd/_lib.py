"""ctypes binding of librte.so (the C ABI declared in include/rt_entry.h).

The library is the only compute path: if it is missing or fails to load,
every call raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librte.so")

RTE_OK = 0
RTE_HOST = 0
RTE_DEVICE = 1
MODE_REF = 0
MODE_TUNED = 1

ERRORS = {-1: "EINVAL", -2: "ESTATE", -3: "EHIP", -4: "ENOMEM", -5: "EUNSUP", -6: "EBUILD"}

# every symbol include/rt_entry.h declares
EXPORTS = [
    "rte_create", "rte_destroy", "rte_last_error", "rte_set_stream", "rte_get_stream",
    "rte_synchronize", "rte_set_base", "rte_set_graph", "rte_set_pca", "rte_project",
    "rte_set_leaves", "rte_build_kdtree", "rte_expand_aabbs", "rte_select_entry",
    "rte_search", "rte_search_batched", "rte_search_batched_noalloc", "rte_autotune_batched",
    "rte_query", "rte_mlp_create", "rte_mlp_train_step",
    "rte_mlp_predict", "rte_predict_entry", "rte_set_rerank_base", "rte_rerank", "rte_knn", "rte_knn_rows", "rte_graph_from_knn", "rte_graph_add_entry_links", "rte_pca_fit", "rte_pca_rotate", "rte_build_kdtree_gpu", "rte_mlp_get", "rte_mlp_set", "rte_mlp_grad", "rte_mlp_set_input_norm", "rte_hash_parameter",
    "rte_search_supported", "rte_version",
]


class RteError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "k", "beam", "width", "max_iter", "mode", "algo", "bloom_words", "bloom_hashes",
        "wg_waves", "grid", "wpq", "order", "queue")]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python -m rt_entry_amd.build` (no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i64 = C.c_int64
    i32 = C.c_int
    f32 = C.c_float
    sig = {
        "rte_create": ([C.POINTER(P), i32], i32),
        "rte_destroy": ([P], i32),
        "rte_last_error": ([P], C.c_char_p),
        "rte_set_stream": ([P, P], i32),
        "rte_get_stream": ([P], P),
        "rte_synchronize": ([P], i32),
        "rte_set_base": ([P, P, i64, i32, i64, i32], i32),
        "rte_set_graph": ([P, P, i64, i32, i32], i32),
        "rte_set_pca": ([P, P, P, i32, i32, i32, i32], i32),
        "rte_project": ([P, P, i64, P, i32], i32),
        "rte_set_leaves": ([P, P, P, i32, i32, i32], i32),
        "rte_build_kdtree": ([P, i64, i32, i32, P, P, P], i32),
        "rte_expand_aabbs": ([P, i32, f32], None),
        "rte_select_entry": ([P, P, i64, i32, P, i32], i32),
        "rte_search": ([P, P, i64, i64, P, C.POINTER(Params), P, P, P, i32], i32),
        "rte_search_batched": ([P, P, i64, i64, P, P, P, C.POINTER(Params), i32, P, P, i32],
                               i32),
        "rte_search_batched_noalloc": ([P, P, i64, i64, P, P, P, C.POINTER(Params), i32, P, P],
                                       i32),
        "rte_query": ([P, P, i64, C.POINTER(Params), P, P, P, P], i32),
        "rte_autotune_batched": ([P, P, i64, i64, P, P, P, C.POINTER(Params), i32, P, P, P, i32,
                                  C.POINTER(Params), C.POINTER(C.c_float)], i32),
        "rte_mlp_create": ([P, i32, i32, i32, i32, C.c_uint64], i32),
        "rte_mlp_train_step": ([P, P, P, i32, f32, f32, P], i32),
        "rte_mlp_predict": ([P, P, i64, P, f32, P, P], i32),
        "rte_predict_entry": ([P, P, i64, i32, f32, P, P], i32),
        "rte_set_rerank_base": ([P, P, i64, i32, i64, i32], i32),
        "rte_rerank": ([P, P, i64, i64, P, i32, i32, P, P, i32], i32),
        "rte_knn": ([P, P, i64, i32, i64, i32, P, P, i32], i32),
        "rte_knn_rows": ([P, P, i64, i32, i64, i64, i64, i32, P, P, i32], i32),
        "rte_graph_from_knn": ([P, P, i64, i32, i32, i32, P, i32], i32),
        "rte_graph_add_entry_links": ([P, P, i64, i32, i32, i32, C.c_uint64, i32], i32),
        "rte_pca_fit": ([P, P, i64, i32, i64, P, P, P, i32], i32),
        "rte_pca_rotate": ([P, P, i64, i32, i64, P, P, i32, P, i32], i32),
        "rte_build_kdtree_gpu": ([P, P, i64, i64, i32, i32, P, P, P, i32], i32),
        "rte_mlp_get": ([P, P, P, P, P, P, P], i32),
        "rte_mlp_set": ([P, P, P, P, P, P, P], i32),
        "rte_mlp_set_input_norm": ([P, P, P], i32),
        "rte_mlp_grad": ([P, P, P, i32, P, P, P, P, P, P, P], i32),
        "rte_hash_parameter": ([i32, P, P, P], None),
        "rte_search_supported": ([i32, i32, i32], i32),
        "rte_version": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(ctx, rc: int):
    if rc != RTE_OK:
        msg = lib().rte_last_error(ctx)
        raise RteError(rc, msg.decode() if msg else "")
    return rc

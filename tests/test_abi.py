"""The C ABI library loads, exports every symbol include/rt_entry.h declares,
and its host-only entry points agree with the oracle (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import assert_same

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "rt_entry.h")).read()
    return sorted(set(re.findall(r"\b(rte_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from rt_entry_amd import _lib
    L = _lib.lib()
    decl = _declared()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(_lib.EXPORTS)


def test_version_and_supported():
    from rt_entry_amd import _lib
    L = _lib.lib()
    assert b"gfx950" in L.rte_version()
    assert L.rte_search_supported(64, 32, 128) == 1
    assert L.rte_search_supported(256, 128, 960) == 1
    assert L.rte_search_supported(64, 32, 4000) == 0


def test_hash_parameter_abi_matches_oracle():
    from rt_entry_amd import hash_parameter
    for nc in list(range(1, 600)) + [1024, 2048, 4096]:
        assert hash_parameter(nc) == O.hash_parameter(nc), nc


def test_host_kdtree_matches_oracle(small_workload, int_workload):
    from rt_entry_amd import build_kdtree, expand_aabbs
    for w in (small_workload, int_workload):
        boxes, pids = build_kdtree(w.base[:, :3], w.cap)
        assert_same(boxes, w.boxes, "boxes")
        assert_same(pids, w.pids, "pids")
        assert_same(expand_aabbs(boxes, 0.2), w.aabbs, "expanded")


def test_host_kdtree_split_failure_is_error_code():
    from rt_entry_amd import RteError, build_kdtree
    with pytest.raises(RteError):
        build_kdtree(np.zeros((64, 3), np.float32), 8)


def test_create_without_gpu_returns_error_code():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("GPU present")
    from rt_entry_amd import _lib
    h = C.c_void_p()
    rc = _lib.lib().rte_create(C.byref(h), 0)
    assert rc != 0 and not h

"""Host-side work plan of the batched Multipole kernel (agx_batch_plan):
every row of every potential is covered exactly once, items respect the rows
limit and come longest (cost x rows) first; bad arguments are rejected.
Runs on the CPU (the plan is host code of the native library)."""
import ctypes as C

import numpy as np
import pytest


def _plan(counts, chunk, path, cost):
    from agama_gpu_amd import _lib
    L = _lib.load()
    counts = np.asarray(counts, np.int64)
    chunk = np.asarray(chunk, np.int32)
    path = np.asarray(path, np.int32)
    cost = np.asarray(cost, np.float64)
    n = C.c_int64(0)
    rc = L.agx_batch_plan(len(counts), counts.ctypes.data, chunk.ctypes.data, path.ctypes.data,
                          cost.ctypes.data, None, 0, C.byref(n))
    if rc:
        return rc, None
    items = np.zeros(n.value, _lib.BATCH_ITEM_DTYPE)
    rc = L.agx_batch_plan(len(counts), counts.ctypes.data, chunk.ctypes.data, path.ctypes.data,
                          cost.ctypes.data, items.ctypes.data if len(items) else None, len(items),
                          C.byref(n))
    return rc, items


def test_plan_covers_every_row_once():
    rng = np.random.default_rng(0)
    counts = rng.integers(0, 40000, 37)
    counts[3] = 0
    counts[5] = 1
    counts[7] = 16385
    chunk = rng.choice([2048, 4096, 8192, 16384], 37)
    path = rng.integers(0, 2, 37)
    cost = rng.uniform(100, 3000, 37)
    rc, items = _plan(counts, chunk, path, cost)
    assert rc == 0
    off = np.concatenate([[0], np.cumsum(counts)])
    seen = np.zeros(off[-1], np.int32)
    for it in items:
        p = it["pot"]
        assert 1 <= it["count"] <= chunk[p]
        assert off[p] <= it["start"] and it["start"] + it["count"] <= off[p + 1]
        assert it["path"] == path[p]
        seen[it["start"]:it["start"] + it["count"]] += 1
    assert np.all(seen == 1)
    c = cost[items["pot"]] * items["count"]
    assert np.all(np.diff(c) <= 0)
    assert len(items) == sum((int(n) + int(k) - 1) // int(k) for n, k in zip(counts, chunk))


@pytest.mark.parametrize("bad,code", [("chunk0", -2), ("chunk_big", -2), ("path", -5),
                                      ("count", -2), ("odd_degree", -5), ("lane_degree", -5),
                                      ("big_degree", -5), ("high_bits", -5), ("degree", 0)])
def test_plan_rejects(bad, code):
    """Item paths: AGX_BATCH_LANE, or AGX_BATCH_LDS with an even triaxial
    degree up to 32 in bits 8-15 (carried to the items unchanged)."""
    from agama_gpu_amd import _lib
    counts, chunk, path, cost = [10, 20], [8, 8], [0, 1], [1.0, 1.0]
    if bad == "chunk0":
        chunk = [0, 8]
    elif bad == "chunk_big":
        chunk = [8, _lib.AGX_BATCH_MAX_ITEM + 1]
    elif bad == "path":
        path = [0, 2]
    elif bad == "odd_degree":
        path = [0, 1 | (7 << 8)]
    elif bad == "lane_degree":
        path = [8 << 8, 1]
    elif bad == "big_degree":
        path = [0, 1 | (34 << 8)]
    elif bad == "high_bits":
        path = [0, 1 | (1 << 16)]
    elif bad == "degree":
        path = [0, 1 | (8 << 8)]
    else:
        counts = [10, -1]
    rc, _ = _plan(counts, chunk, path, cost)
    assert rc == code


def test_plan_capacity_and_nulls():
    from agama_gpu_amd import _lib
    L = _lib.load()
    counts = np.array([100], np.int64)
    chunk = np.array([10], np.int32)
    path = np.array([0], np.int32)
    cost = np.array([1.0])
    n = C.c_int64(0)
    items = np.zeros(5, _lib.BATCH_ITEM_DTYPE)
    rc = L.agx_batch_plan(1, counts.ctypes.data, chunk.ctypes.data, path.ctypes.data,
                          cost.ctypes.data, items.ctypes.data, 5, C.byref(n))
    assert rc == -7 and n.value == 10            # AGX_E_WORKSPACE, needed count reported
    assert L.agx_batch_plan(1, None, chunk.ctypes.data, path.ctypes.data, cost.ctypes.data,
                            None, 0, C.byref(n)) == -1

"""The allocation-free unique-edge entry point (sm_unique_edges_noalloc)
against sm_unique_edges with a caller workspace, through the C ABI: two
calls in a row on the library's shared scratch, the second smaller, and a
request beyond the scratch's size refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _with_workspace(lib, pt, n, nvert, torch, dev):
    out = torch.empty((max(n, 1), 2), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    wsb = int(lib.lib.sm_unique_edges_workspace(n, nvert))
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    lib.check(lib.lib.sm_unique_edges(pt.data_ptr(), n, nvert, out.data_ptr(), cnt.data_ptr(),
                                      ws.data_ptr(), wsb, None), "sm_unique_edges")
    torch.cuda.synchronize()
    k = int(cnt.item())
    return out[:k].cpu().numpy()


def _noalloc(lib, pt, n, nvert, torch, dev):
    out = torch.full((max(n, 1), 2), -7, dtype=torch.int32, device=dev)
    cnt = torch.full((1,), -1, dtype=torch.int64, device=dev)
    lib.check(lib.lib.sm_unique_edges_noalloc(pt.data_ptr(), n, nvert, out.data_ptr(),
                                              cnt.data_ptr(), None), "sm_unique_edges_noalloc")
    torch.cuda.synchronize()
    k = int(cnt.item())
    return out[:k].cpu().numpy()


def test_noalloc_matches_workspace_entry_point():
    import torch
    from seismicmesh_amd import _lib
    _lib.require_gpu()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(21)
    # a large call (long buckets included), then a smaller one on the same
    # scratch, whose leftovers must not leak into the result
    for n, nvert in [(1_500_000, 300_000), (40_000, 9_000)]:
        pairs = rng.integers(0, nvert, (n, 2)).astype(np.int32)
        pairs[:2000, 0] = 7  # one vertex with a very long bucket
        pt = torch.from_numpy(pairs).to(dev)
        want = _with_workspace(_lib, pt, n, nvert, torch, dev)
        got = _noalloc(_lib, pt, n, nvert, torch, dev)
        assert len(want) > 0
        assert np.array_equal(got, want), (n, nvert)
        ref = np.unique(np.sort(pairs, 1), axis=0)
        assert np.array_equal(got, ref)


def test_noalloc_refuses_inputs_beyond_its_scratch():
    import torch
    from seismicmesh_amd import _lib
    _lib.require_gpu()
    dev = torch.device("cuda", 0)
    n, nvert = 8_000_000, 2_000_000
    assert _lib.lib.sm_unique_edges_workspace(n, nvert) > _lib.NOALLOC_SCRATCH_BYTES
    pt = torch.zeros((16, 2), dtype=torch.int32, device=dev)
    out = torch.empty((16, 2), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    rc = _lib.lib.sm_unique_edges_noalloc(pt.data_ptr(), n, nvert, out.data_ptr(),
                                          cnt.data_ptr(), None)
    assert rc == -3  # SM_ERR_CAPACITY, nothing launched
    assert "scratch" in _lib.last_error()

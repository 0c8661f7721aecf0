"""rmh_statistics_noalloc (the allocation-free statistics entry point) against rmh_statistics
with a caller workspace, through the C ABI: identical bytes on two calls in a row, the second
with a smaller input, plus shapes without partial sums.  Marked gpu; kept last in collection
order because every call on the GPU shares the library's scratch array."""
import ctypes

import numpy as np
import pytest
import torch

from rmagine_amd import _lib
from rmagine_amd import math as M
from rmagine_amd import types as T

pytestmark = pytest.mark.gpu

dev = "cuda"


def _inputs(rng, seg, nseg, shared):
    q = rng.normal(size=(nseg, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    pre = T.make_transforms(rng.uniform(-0.5, 0.5, (nseg, 3)), q)
    nd = seg if shared else seg * nseg
    d = (rng.normal(size=(nd, 3)) * 5).astype(np.float32)
    mp = (rng.normal(size=(seg * nseg, 3)) * 5 + 0.1).astype(np.float32)
    nr = rng.normal(size=(seg * nseg, 3))
    nr = (nr / np.linalg.norm(nr, axis=1, keepdims=True)).astype(np.float32)
    dmask = (rng.uniform(size=nd) > 0.1).astype(np.uint8)
    mmask = (rng.uniform(size=seg * nseg) > 0.1).astype(np.uint8)
    Td = lambda x: torch.as_tensor(x, device=dev)
    data = M.PointCloudView(Td(d), mask=Td(dmask), segment_stride=0 if shared else None)
    model = M.PointCloudView(Td(mp), mask=Td(mmask), normals=Td(nr))
    return Td(pre), data, model


def _both(kind, pre, data, model, seg, nseg, max_dist=6.0):
    prm = T.rmh_umeyama_constraints(max_dist, 0, 0)
    dc, mc = data.to_c(seg), model.to_c(seg)
    ref = torch.full((nseg, 16), float("nan"), dtype=torch.float64, device=dev)
    got = torch.full((nseg, 16), float("nan"), dtype=torch.float64, device=dev)
    ws = M._workspace(seg, nseg, dev)
    _lib.call("rmh_statistics", kind, M._ptr(pre), pre.shape[0], ctypes.byref(dc), ctypes.byref(mc), seg, nseg,
              ctypes.byref(prm), M._ptr(ref), M._ptr(ws), 0 if ws is None else ws.numel(), None)
    _lib.call("rmh_statistics_noalloc", kind, M._ptr(pre), pre.shape[0], ctypes.byref(dc), ctypes.byref(mc), seg,
              nseg, ctypes.byref(prm), M._ptr(got), None)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("kind", [M.P2P, M.P2L])
def test_noalloc_equals_workspace_entry_two_calls(kind):
    rng = np.random.default_rng(101 + kind)
    # first call: 2,000 hypotheses x 8,192 correspondences (4 blocks per segment: partial sums
    # in the scratch array); second call: smaller, 500 x 4,096 (2 blocks per segment)
    for seg, nseg, shared in ((8192, 2000, True), (4096, 500, False)):
        pre, data, model = _inputs(rng, seg, nseg, shared)
        ref, got = _both(kind, pre, data, model, seg, nseg)
        assert np.all(np.isfinite(ref))
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
        assert ref.view(np.uint32)[:, 30].sum() > 0


def test_noalloc_single_block_and_large_segment():
    rng = np.random.default_rng(202)
    # one block per segment (no partial sums), then one segment over 256 blocks
    for seg, nseg in ((500, 20000), (1_000_000, 1)):
        pre, data, model = _inputs(rng, seg, nseg, False)
        ref, got = _both(M.P2L, pre, data, model, seg, nseg)
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_noalloc_argument_errors():
    lib = _lib.lib()
    prm = T.rmh_umeyama_constraints(1.0, 0, 0)
    c = T.rmh_cloud()
    assert lib.rmh_statistics_noalloc(7, None, 1, ctypes.byref(c), ctypes.byref(c), 10, 1, ctypes.byref(prm),
                                      None, None) == 1
    assert lib.rmh_statistics_noalloc(0, None, 1, ctypes.byref(c), ctypes.byref(c), 10, 70000,
                                      ctypes.byref(prm), None, None) == 1

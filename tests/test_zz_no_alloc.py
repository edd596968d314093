"""tn_intersect_noalloc (the allocation-free variant of tn_intersect, for callers
that cannot allocate) against tn_intersect through the C ABI: bit-exact ids,
r, s, t on consecutive calls (the second smaller, so the static work counters
are reused), every ray traversed once, and replays of a captured graph."""
import numpy as np
import pytest

import turner_amd as T
from turner_amd import scenes
from turner_amd._lib import TN_ISECT_COUNT

pytestmark = pytest.mark.gpu


def _outputs(torch, n, dev):
    return (torch.empty(n, dtype=torch.int32, device=dev),) + tuple(
        torch.empty(n, dtype=torch.float32, device=dev) for _ in range(3))


def _call(fn, tree, to, td, outs, n, flags, stream):
    ids, r, s, t = outs
    T.check(fn(tree.handle, T.ptr(to), T.ptr(td), n, T.ptr(ids), T.ptr(r), T.ptr(s), T.ptr(t), flags, stream))


def test_noalloc_matches_intersect(gpu):
    import torch
    L = T.lib()
    sc = scenes.intersect_scene(level=5)
    tree = T.KDTree(sc["verts"])
    _, cus = T.device_info()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    # above the work-counter threshold (cus x 28 waves x 256 rays), ragged; then smaller
    sizes = [max(1 << 21, cus * 28 * 256) + 4321, cus * 28 * 256 + 999, 70001]
    for k, n in enumerate(sizes):
        print(f"[noalloc] call {k}: n={n}", flush=True)
        o, d = scenes.random_rays(n, seed=40 + k)
        to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        ref, got = _outputs(torch, n, dev), _outputs(torch, n, dev)
        _call(L.tn_intersect, tree, to, td, ref, n, 0, stream)
        cnt = np.zeros(8, np.uint64)
        T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
        _call(L.tn_intersect_noalloc, tree, to, td, got, n, TN_ISECT_COUNT, stream)
        torch.cuda.synchronize()
        T.check(L.tn_get_counters(tree.handle, cnt.ctypes.data, 1))
        assert cnt[4] == n, "every ray traversed exactly once"
        for x, y in zip(got, ref):
            assert torch.equal(x, y), f"call {k}"
        again = _outputs(torch, n, dev)
        _call(L.tn_intersect_noalloc, tree, to, td, again, n, 0, stream)
        torch.cuda.synchronize()
        for x, y in zip(again, ref):
            assert torch.equal(x, y), f"call {k} repeated"
    assert L.tn_intersect_noalloc(None, None, None, 0, None, None, None, None, 0, None) == T._lib.TN_ERR_INVALID


def test_noalloc_in_captured_graph(gpu):
    """Captured once, replayed three times with eager work in between: every
    replay equals the eager tn_intersect result (the captured launch uses the
    static work counters, cleared by a kernel node)."""
    import torch
    L = T.lib()
    sc = scenes.intersect_scene(level=4)
    tree = T.KDTree(sc["verts"])
    _, cus = T.device_info()
    n = cus * 28 * 256 + 777
    o, d = scenes.random_rays(n, seed=77)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    ref = _outputs(torch, n, dev)
    _call(L.tn_intersect, tree, to, td, ref, n, 0, torch.cuda.current_stream().cuda_stream)
    got = _outputs(torch, n, dev)
    # first call outside the capture: looks up the device address of the counters
    _call(L.tn_intersect_noalloc, tree, to, td, got, n, 0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _call(L.tn_intersect_noalloc, tree, to, td, got, n, 0, torch.cuda.current_stream().cuda_stream)
    for k in range(3):
        for x in got:
            x.zero_()
        _call(L.tn_intersect, tree, to, td, _outputs(torch, n, dev), n, 0, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for x, y in zip(got, ref):
            assert torch.equal(x, y), f"replay {k}"
    del g

"""re_trace_closest_noalloc (the allocation-free closest-hit entry point)
against re_trace_closest / re_trace_counts, through the C ABI: two calls in a
row, the second with a smaller input, with host and with device pointers;
prim ids and t bit-identical, counters equal, and an oversized host-pointer
call refused."""
import ctypes as C
import sys

import numpy as np
import pytest

import rayengine_amd as re
from rayengine_amd import _lib
from rayengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(ctx, scene, o, d, n, p, t, counts, flags, **params):
    tp = re._trace_params(**params)
    st = _lib.lib().re_trace_closest_noalloc(ctx.handle, scene.id, o, d, n, C.byref(tp), p, t, None, counts, flags)
    if st != 0:
        raise re.ReError(st, _lib.lib().re_last_error(ctx.handle).decode())


def _host_call(ctx, scene, o, d, **params):
    n = len(o)
    p = np.full(n, -7, np.int32)
    t = np.zeros(n, np.float32)
    cnt = (C.c_uint64 * 4)()
    _noalloc(ctx, scene, _lib.ptr(o), _lib.ptr(d), n, _lib.ptr(p), _lib.ptr(t), cnt, 0, **params)
    return p, t, list(cnt)


def test_noalloc_matches_trace_closest(gpu_ctx):
    tri_v, tree, o, d = W.small_scene(256, 256)
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    rng = np.random.default_rng(5)
    sub = np.sort(rng.choice(len(o), 5000, replace=False))
    # two calls in a row through host pointers, the second smaller
    for k, (so, sd) in enumerate(((o, d), (np.ascontiguousarray(o[sub]), np.ascontiguousarray(d[sub])))):
        for variant in (1, 3, 4, 5):   # 1 and 5 take rays from the area's ray counters
            say(f"host pointers, call {k}, {len(so)} rays, variant {variant}")
            p, t, cnt = _host_call(gpu_ctx, scene, so, sd, variant=variant)
            pr, tr, cr = scene.trace_counts(so, sd, variant=variant)
            assert np.array_equal(p, pr) and np.array_equal(t.view(np.uint32), tr.view(np.uint32))
            assert cnt == list(cr)
    # device pointers, the second call smaller again
    import torch
    dev = torch.device("cuda:0")
    for k, (so, sd) in enumerate(((o, d), (o[sub], d[sub]))):
        say(f"device pointers, call {k}, {len(so)} rays")
        n = len(so)
        to, td = torch.from_numpy(np.ascontiguousarray(so)).to(dev), torch.from_numpy(np.ascontiguousarray(sd)).to(dev)
        tp = torch.full((n,), -7, dtype=torch.int32, device=dev)
        tt = torch.zeros(n, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        cnt = (C.c_uint64 * 4)()
        _noalloc(gpu_ctx, scene, to.data_ptr(), td.data_ptr(), n, tp.data_ptr(), tt.data_ptr(), cnt,
                 re.RE_DEVICE_PTRS)
        pr, tr, cr = scene.trace_counts(so, sd)
        assert np.array_equal(tp.cpu().numpy(), pr)
        assert np.array_equal(tt.cpu().numpy().view(np.uint32), tr.view(np.uint32))
        assert list(cnt) == list(cr)
    scene.release()


def test_noalloc_refuses_oversized_host_calls(gpu_ctx):
    tri_v, tree, o, d = W.small_scene(8, 8)
    scene = gpu_ctx.upload_scene([re.SceneObject(tree)])
    limit = _lib.lib().re_trace_closest_noalloc_max_rays()
    assert limit == 1 << 18
    n = limit + 1
    big_o = np.zeros((n, 3), np.float32)
    big_d = np.tile(np.float32([0, 0, -1]), (n, 1))
    p = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    with pytest.raises(re.ReError, match="262144"):
        _noalloc(gpu_ctx, scene, _lib.ptr(big_o), _lib.ptr(big_d), n, _lib.ptr(p), _lib.ptr(t), None, 0)
    # the context still works afterwards
    p2, t2, _ = _host_call(gpu_ctx, scene, o, d)
    pr, tr = scene.trace_closest(o, d)
    assert np.array_equal(p2, pr) and np.array_equal(t2.view(np.uint32), tr.view(np.uint32))
    scene.release()

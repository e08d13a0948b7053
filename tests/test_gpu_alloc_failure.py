"""GPU tests of the allocation-failure contract: creation returns NULL when an allocation
fails and releases everything it had acquired (reference src/plan.c:81-90,112-114 and
src/bridct_alloc.h:10-20: NULL on a failed allocation, no leak), for plans, ragged
batches and codecs; calls that create buffers on first use report BRIDCT_DEVICE_ERROR (or
fall back to a path that needs none) and leave the object usable.

Failures are injected with bridct_debug_alloc_faults(k): the library's next k GPU-runtime
acquisitions succeed and the one after fails without calling the runtime.  Its return
value, the number of runtime resources the library holds, shows leaks."""
import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu


def each_failure(create, destroy, limit=64):
    """Fails the k-th acquisition of create() for k = 0, 1, ... until create succeeds;
    every failed create must return NULL and leave the resource count unchanged.
    Returns the number of acquisitions a successful create makes."""
    lib = capi.load()
    for k in range(limit):
        base = lib.bridct_debug_alloc_faults(k)
        obj = create()
        left = lib.bridct_debug_alloc_faults(-1)
        if obj:
            destroy(obj)
            assert lib.bridct_debug_alloc_faults(-1) == base, 'destroy leaked'
            assert k > 0, 'no acquisition happened'
            return k
        assert left == base, ('failed create leaked', k, left - base)
    raise AssertionError('create never succeeded')


@pytest.mark.parametrize('h,w', [(8, 8), (256, 256), (16, 512), (1024, 1024)])
def test_plan_create_failures(h, w):
    lib = capi.load()
    n = each_failure(lambda: lib.bridct_plan_create(h, w), lib.bridct_plan_destroy)
    print('plan %dx%d: %d acquisitions' % (h, w, n), flush=True)
    # the legacy square API builds its GPU state the same way and reports the failure
    if h == w and h <= 256:
        x = np.ones((h, w), np.float32)
        y = np.full((h, w), 7.0, np.float32)
        ws = np.empty(lib.bridct_workspace_floats(h) + 4, np.float32)
        wsa = ws[(-ws.ctypes.data % 16) // 4:]
        assert lib.bridct_debug_alloc_faults(0) >= 0
        rc = lib.bridct_apply(h, 0, capi.ptr(x), capi.ptr(y), capi.ptr(wsa), wsa.size)
        lib.bridct_debug_alloc_faults(-1)
        assert rc in (capi.DEVICE_ERROR, capi.OK)
        if rc == capi.DEVICE_ERROR:
            assert (y == 7.0).all(), 'a failed call wrote its output'
        rc = lib.bridct_apply(h, 0, capi.ptr(x), capi.ptr(y), capi.ptr(wsa), wsa.size)
        assert rc == 0 and abs(y[0, 0] - np.sqrt(h * w)) < 1e-3


def test_plan_first_use_buffers():
    """Buffers made on first use: the pinned pipeline of large host calls (a failure there
    falls back to the plain chunked path, same results) and the dense route's intermediate
    (a failure is reported, the next call succeeds)."""
    lib = capi.load()
    h = w = 256
    x = np.random.default_rng(3).uniform(-1, 1, (32, h, w)).astype(np.float32)
    want = None
    for k in range(12):
        plan = capi.Plan(h, w)
        y = np.empty_like(x)
        lib.bridct_debug_alloc_faults(k)
        rc = plan.apply_batch(0, x, y)
        lib.bridct_debug_alloc_faults(-1)
        assert rc == 0, (k, rc)
        if want is None:
            want = S.oracle_2d(x[:2], 0)
        assert S.max_rel(y[:2], want) < 1e-6
        y2 = np.empty_like(x)
        assert plan.apply_batch(0, x, y2) == 0 and np.array_equal(y, y2)
        plan.close()
    # dense route: the intermediate buffer is allocated by the first device call
    plan = capi.Plan(64, 64)
    assert plan.set_route(capi.ROUTE_MFMA_F16) == 0
    xs = np.random.default_rng(4).uniform(-1, 1, (8, 64, 64)).astype(np.float32)
    d_in, d_out = capi.DeviceBuffer(xs.nbytes), capi.DeviceBuffer(xs.nbytes)
    d_in.upload(xs)
    lib.bridct_debug_alloc_faults(0)
    rc = plan.apply_device(0, d_in.ptr, d_out.ptr, 8)
    lib.bridct_debug_alloc_faults(-1)
    assert rc == capi.DEVICE_ERROR
    assert plan.apply_device(0, d_in.ptr, d_out.ptr, 8) == 0 and plan.synchronize() == 0
    ys = np.empty_like(xs)
    d_out.download(ys)
    assert S.max_rel(ys, S.oracle_2d(xs, 0)) < 2e-6
    plan.close()


@pytest.mark.parametrize('shapes', [
    [(8, 8)] * 100,                                           # tiny kernel
    [(16, 32), (128, 1024), (64, 64), (1024, 8), (256, 256)],  # persistent queue
    [(1024, 1024)] * 2 + [(64, 512)] * 8,                      # grouped
], ids=['tiny', 'persistent', 'grouped'])
def test_batch_create_failures(shapes):
    lib = capi.load()
    hs = np.array([h for h, _ in shapes], np.int32)
    ws = np.array([w for _, w in shapes], np.int32)
    n = each_failure(lambda: lib.bridct_batch_create(len(shapes), capi.ptr(hs), capi.ptr(ws)),
                     lib.bridct_batch_destroy)
    print('batch %d arrays: %d acquisitions' % (len(shapes), n), flush=True)
    # the host-pointer call stages through buffers made on first use
    rng = np.random.default_rng(5)
    arrays = [rng.uniform(-1, 1, s).astype(np.float32) for s in shapes]
    x = np.concatenate([a.ravel() for a in arrays])
    for k in range(6):
        b = capi.Batch(shapes)
        y = np.full_like(x, 7.0)
        lib.bridct_debug_alloc_faults(k)
        rc = b.apply(0, x, y)
        lib.bridct_debug_alloc_faults(-1)
        assert rc in (0, capi.DEVICE_ERROR), rc
        assert b.apply(0, x, y) == 0
        for (h, w), a, out in zip(shapes, arrays, b.split(y)):
            assert S.max_rel(out, S.oracle_2d(a, 0)) < 1e-6, (h, w)
        b.close()


def test_codec_create_failures():
    lib = capi.load()
    q = np.ascontiguousarray(S.JPEG_LUMA, np.float32)
    n = each_failure(lambda: lib.bridct_codec8_create(capi.ptr(q)), lib.bridct_codec8_destroy)
    print('codec: %d acquisitions' % n, flush=True)


def test_batch_graph_failure():
    """The captured graph of bridct_batch_graph_apply is a counted resource: a failed
    capture reports BRIDCT_DEVICE_ERROR and holds nothing, the next call captures and
    replays, and destroy gives back every resource."""
    lib = capi.load()
    shapes = [(16, 32), (128, 1024), (64, 64), (256, 256)]
    rng = np.random.default_rng(6)
    arrays = [rng.uniform(-1, 1, s).astype(np.float32) for s in shapes]
    x = np.concatenate([a.ravel() for a in arrays])
    base = lib.bridct_debug_alloc_faults(-1)
    b = capi.Batch(shapes)
    d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    before = lib.bridct_debug_alloc_faults(0)
    rc = b.graph_apply(capi.FORWARD, d_in.ptr, d_out.ptr)
    assert lib.bridct_debug_alloc_faults(-1) == before
    assert rc == capi.DEVICE_ERROR
    assert b.graph_apply(capi.FORWARD, d_in.ptr, d_out.ptr) == 0 and b.synchronize() == 0
    assert lib.bridct_debug_alloc_faults(-1) == before + 1
    y = np.empty_like(x)
    d_out.download(y)
    for (h, w), a, out in zip(shapes, arrays, b.split(y)):
        assert S.max_rel(out, S.oracle_2d(a, 0)) < 1e-6, (h, w)
    d_in.free()
    d_out.free()
    b.close()
    assert lib.bridct_debug_alloc_faults(-1) == base

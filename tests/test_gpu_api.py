"""GPU tests of the C ABI: host-pointer plan calls, device-pointer batches, in place,
special closed forms (reference tests/test_api.c, tests/test_small_optimized.c,
tests/test_runtime.c)."""
import ctypes

import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def L():
    return capi.load()


def test_gpu_available(L):
    assert L.bridct_gpu_available() == 1


@pytest.mark.parametrize('h,w', [(8, 8), (16, 16), (32, 32), (64, 64), (128, 128), (256, 256),
                                 (512, 512), (1024, 1024), (8, 1024), (1024, 8), (16, 32), (512, 128)])
def test_api_roundtrip_sine(L, h, w):
    """reference tests/test_api.c: sin(0.017 k) input, round trip |y - x| < 2e-5, input
    unchanged, outputs finite, for every mode."""
    x = np.sin(0.017 * np.arange(h * w)).astype(np.float32).reshape(h, w)
    x0 = x.copy()
    p = L.bridct_plan_create(h, w)
    assert p
    try:
        for mode in (0, 1, 2):
            y = np.full_like(x, np.nan)
            assert L.bridct_plan_apply(p, mode, capi.ptr(x), capi.ptr(y)) == 0
            assert np.array_equal(x, x0)
            assert np.isfinite(y).all()
            if mode == 2:
                assert np.abs(y - x).max() < 2e-5
        route = L.bridct_plan_route(p).decode()
        assert route == ('paper8-square' if h == w and h <= 256 else 'blocked-extension')
    finally:
        L.bridct_plan_destroy(p)


def test_apply_argument_errors_write_nothing(L):
    p = L.bridct_plan_create(16, 16)
    try:
        x = np.ones((16, 16), np.float32)
        y = np.full((16, 16), 7.0, np.float32)
        assert L.bridct_plan_apply(None, 0, capi.ptr(x), capi.ptr(y)) == capi.INVALID_ARGUMENT
        assert L.bridct_plan_apply(p, 3, capi.ptr(x), capi.ptr(y)) == capi.INVALID_ARGUMENT
        assert L.bridct_plan_apply(p, 0, None, capi.ptr(y)) == capi.INVALID_ARGUMENT
        assert L.bridct_plan_apply(p, 0, capi.ptr(x), capi.ptr(x)) == capi.OVERLAPPING_BUFFERS
        buf = np.zeros(16 * 16 + 8, np.float32)
        a, b = buf[:256], buf[8:264]
        assert L.bridct_plan_apply(p, 0, capi.ptr(a), capi.ptr(b)) == capi.OVERLAPPING_BUFFERS
        assert np.all(y == 7.0)
    finally:
        L.bridct_plan_destroy(p)


@pytest.mark.parametrize('n', [8, 16, 32, 64, 128, 256])
def test_impulses_square(L, n):
    """Every impulse (reference tests/test_small_optimized.c:95-150 for 8 and 16; a sample
    of rows/columns above) against the FP64 basis outer product."""
    idx = range(n) if n <= 16 else [0, 1, n // 3, n // 2, n - 2, n - 1]
    xs = []
    for r in idx:
        for c in idx:
            x = np.zeros((n, n), np.float32)
            x[r, c] = 1
            xs.append(x)
    x = np.stack(xs)
    plan = capi.Plan(n, n)
    try:
        for mode in (0, 1, 2):
            y = np.empty_like(x)
            assert plan.apply_batch(mode, x, y) == 0
            exact = S.oracle_2d(x, mode)
            err = np.linalg.norm((y - exact).reshape(len(xs), -1), axis=1)
            assert err.max() <= 2e-5
    finally:
        plan.close()


@pytest.mark.parametrize('h,w', [(8, 8), (256, 256), (512, 512), (512, 1024), (1024, 512),
                                 (1024, 1024), (32, 256)])
def test_constant_closed_form(L, h, w):
    """Ones -> sqrt(hw) at DC, zero elsewhere; its inverse -> 1/sqrt(hw)
    (reference tests/test_runtime.c:141-153)."""
    x = np.ones((h, w), np.float32)
    plan = capi.Plan(h, w)
    try:
        y = np.empty_like(x)
        assert plan.apply(0, x, y) == 0
        exp = np.zeros((h, w))
        exp[0, 0] = np.sqrt(h * w)
        assert np.abs(y - exp).max() <= 2e-6 * np.sqrt(h * w)
        d = np.zeros((h, w), np.float32)
        d[0, 0] = 1
        assert plan.apply(1, d, y) == 0
        assert np.abs(y - 1 / np.sqrt(h * w)).max() <= 1e-6 / np.sqrt(h * w) * 8
    finally:
        plan.close()


@pytest.mark.parametrize('h,w', [(8, 8), (64, 256), (256, 256), (1024, 1024), (128, 1024)])
def test_device_batch_inplace_and_host_agree(L, h, w):
    rng = np.random.default_rng(7 + h + w)
    batch = 3
    x = rng.uniform(-1, 1, (batch, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    d_in = capi.DeviceBuffer(x.nbytes)
    d_out = capi.DeviceBuffer(x.nbytes)
    try:
        for mode in (0, 1, 2):
            host = np.empty_like(x)
            assert plan.apply_batch(mode, x, host) == 0
            d_in.upload(x)
            assert plan.apply_device(mode, d_in.ptr, d_out.ptr, batch) == 0
            assert plan.synchronize() == 0
            dev = np.empty_like(x)
            d_out.download(dev)
            assert np.array_equal(dev, host)
            # in place
            assert plan.apply_device(mode, d_in.ptr, d_in.ptr, batch) == 0
            assert plan.synchronize() == 0
            d_in.download(dev)
            assert np.array_equal(dev, host)
            # single-array host API equals the batch path bitwise
            one = np.empty_like(x[0])
            assert plan.apply(mode, x[1], one) == 0
            assert np.array_equal(one, host[1])
        bad = ctypes.c_void_p(d_in.ptr + 4)
        assert plan.apply_device(0, bad, d_out.ptr, 1) == capi.BAD_ALIGNMENT
        part = ctypes.c_void_p(d_in.ptr + 16)
        assert plan.apply_device(0, d_in.ptr, part, 1) == capi.OVERLAPPING_BUFFERS
    finally:
        d_in.free()
        d_out.free()
        plan.close()


@pytest.mark.parametrize('h,w,batch', [(64, 64, 600), (256, 256, 37), (1024, 1024, 3), (8, 16, 40000)])
def test_host_pipeline_chunks(L, h, w, batch):
    """Host calls of 2^18 floats or more go through the two-slot pinned pipeline in chunks
    of 2^20 floats: several chunks with a partial last one, a single partial chunk, and one
    array per chunk all equal the device path bitwise, and the input is left unchanged."""
    rng = np.random.default_rng(batch + h)
    x = rng.uniform(-1, 1, (batch, h, w)).astype(np.float32)
    x0 = x.copy()
    plan = capi.Plan(h, w)
    d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    try:
        d_in.upload(x)
        for mode in (0, 1, capi.FORWARD_THEN_INVERSE):
            print('pipeline %dx%d batch %d mode %d' % (h, w, batch, mode), flush=True)
            host = np.full_like(x, np.nan)
            assert plan.apply_batch(mode, x, host) == 0
            if mode == capi.FORWARD_THEN_INVERSE:
                assert plan.apply_device(0, d_in.ptr, d_out.ptr, batch) == 0
                assert plan.apply_device(1, d_out.ptr, d_out.ptr, batch) == 0
            else:
                assert plan.apply_device(mode, d_in.ptr, d_out.ptr, batch) == 0
            assert plan.synchronize() == 0
            dev = np.empty_like(x)
            d_out.download(dev)
            assert np.array_equal(dev, host)
            assert np.array_equal(x, x0)
    finally:
        plan.close()


def test_forward_then_inverse_mode_equals_two_calls(L):
    rng = np.random.default_rng(3)
    for h, w in [(16, 16), (128, 512), (1024, 1024)]:
        x = rng.standard_normal((2, h, w)).astype(np.float32)
        plan = capi.Plan(h, w)
        f = np.empty_like(x)
        i = np.empty_like(x)
        r = np.empty_like(x)
        assert plan.apply_batch(0, x, f) == 0
        assert plan.apply_batch(1, f, i) == 0
        assert plan.apply_batch(3, x, r) == 0
        assert np.array_equal(i, r)
        plan.close()


@pytest.mark.parametrize('n', [8, 16, 32, 64, 128, 256])
def test_legacy_apply_matches_plan(L, n):
    """bridct_apply with caller workspace gives the plan's result bitwise
    (reference tests/verify_plan.py:38-42)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)).astype(np.float32)
    need = L.bridct_workspace_floats(n)
    ws = np.zeros(need + 4, np.float32)
    off = (-ws.ctypes.data % 16) // 4
    wsa = ws[off:off + need]
    plan = capi.Plan(n, n)
    try:
        for mode in (0, 1, 2):
            a = np.empty_like(x)
            b = np.empty_like(x)
            assert L.bridct_apply(n, mode, capi.ptr(x), capi.ptr(a), capi.ptr(wsa), need) == 0
            assert plan.apply(mode, x, b) == 0
            assert np.array_equal(a, b)
    finally:
        plan.close()


@pytest.mark.parametrize('h,w', [(1024, 1024), (512, 128), (8, 1024), (32, 256), (16, 16)])
def test_passes_compose_to_apply_device(h, w):
    """Row pass then column pass (bridct_plan_apply_device_pass) equals the whole transform."""
    batch = 3
    x = np.random.default_rng(h + w).uniform(-1, 1, (batch, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    a, b, c = (capi.DeviceBuffer(x.nbytes) for _ in range(3))
    try:
        a.upload(x)
        for mode in (capi.FORWARD, capi.INVERSE):
            assert plan.apply_device(mode, a.ptr, c.ptr, batch) == 0
            assert plan.apply_device_pass(mode, 0, a.ptr, b.ptr, batch) == 0
            assert plan.apply_device_pass(mode, 1, b.ptr, b.ptr, batch) == 0
            assert plan.synchronize() == 0
            y, z = np.empty_like(x), np.empty_like(x)
            b.download(y)
            c.download(z)
            assert np.array_equal(y, z)
        assert plan.apply_device_pass(capi.ROUNDTRIP, 0, a.ptr, b.ptr, batch) == capi.INVALID_ARGUMENT
        assert plan.apply_device_pass(capi.FORWARD, 2, a.ptr, b.ptr, batch) == capi.INVALID_ARGUMENT
        assert plan.apply_device_pass(capi.FORWARD, 1, a.ptr, b.ptr, batch) == capi.INVALID_ARGUMENT
    finally:
        plan.close()


@pytest.mark.parametrize('h,w', [(128, 512), (256, 256), (512, 128), (1024, 64), (128, 1024), (512, 512),
                                 (1024, 1024)])
def test_narrow_column_strips_bitwise(h, w):
    """Column passes with too few workgroups for the GPU take strips of a half or a quarter
    of the width.  The same arrays transformed in batches that pick the wide, the half and
    the quarter strips give the same bits, in every mode."""
    rng = np.random.default_rng(h * 7 + w)
    full = 32 * 1024 // w    # 512 workgroups of 64 columns: the wide strips
    x = rng.uniform(-1, 1, (full, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    a, b = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    try:
        a.upload(x)
        for mode in (capi.FORWARD, capi.INVERSE, capi.ROUNDTRIP):
            print('narrow strips %dx%d mode %d' % (h, w, mode), flush=True)
            ref = None
            for batch in (full, max(1, full // 2), max(1, full // 8), 1):
                assert plan.apply_device(mode, a.ptr, b.ptr, batch) == 0
                assert plan.synchronize() == 0
                y = np.empty((full, h, w), np.float32)
                b.download(y)
                if ref is None:
                    ref = y.copy()
                    assert np.isfinite(ref).all()
                else:
                    assert np.array_equal(y[:batch], ref[:batch]), batch
    finally:
        a.free()
        b.free()
        plan.close()

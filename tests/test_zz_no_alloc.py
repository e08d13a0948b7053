"""GPU test of bridct_plan_apply_device_static, the device-pointer call that allocates
nothing (its dense-route intermediate is the library's static per-GPU scratch array).

For each route (the split-MFMA routes use the scratch; the butterfly route needs none) two
calls in a row, the second on a smaller batch, must equal bridct_plan_apply_device's
output bitwise; the calls acquire no runtime resource (the library's resource count is
unchanged, and an acquisition failure injected before them does not fire); a batch beyond
the scratch is refused with BRIDCT_WORKSPACE_TOO_SMALL and writes nothing."""
import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu

STATIC_FLOATS = 8 << 20


def run(lib, plan, fn, mode, x, batch):
    d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    rc = fn(plan.ptr, mode, d_in.ptr, d_out.ptr, batch)
    assert rc == 0, rc
    assert plan.synchronize() == 0
    y = np.empty_like(x)
    d_out.download(y)
    return y


@pytest.mark.parametrize('route', [capi.ROUTE_MFMA_F16, capi.ROUTE_MFMA_BF16, capi.ROUTE_BUTTERFLY])
@pytest.mark.parametrize('h,w,b1,b2', [(64, 64, 16, 3), (256, 128, 32, 5), (1024, 1024, 8, 1)])
def test_static_equals_device_call(route, h, w, b1, b2):
    lib = capi.load()
    rng = np.random.default_rng(h + w + route)
    quiet = capi.Plan(h, w)           # the no-alloc calls
    plain = capi.Plan(h, w)           # the existing entry point
    try:
        assert quiet.set_route(route) == 0 and plain.set_route(route) == 0
        for mode in (0, 1, 2):
            for batch in (b1, b2):
                print('static call route %d shape %dx%d batch %d mode %d' % (route, h, w, batch, mode), flush=True)
                x = rng.uniform(-1, 1, (batch, h, w)).astype(np.float32)
                d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
                d_in.upload(x)
                before = lib.bridct_debug_alloc_faults(0)      # the next acquisition would fail
                rc = lib.bridct_plan_apply_device_static(quiet.ptr, mode, d_in.ptr, d_out.ptr, batch)
                sync = quiet.synchronize()
                after = lib.bridct_debug_alloc_faults(-1)
                assert rc == 0 and sync == 0, (rc, sync)
                assert after == before, 'the static call acquired a runtime resource'
                got = np.empty_like(x)
                d_out.download(got)
                want = run(lib, plain, lib.bridct_plan_apply_device, mode, x, batch)
                assert np.array_equal(got, want), (route, h, w, batch, mode)
                if mode < 2:
                    tol = 2e-6 if route == capi.ROUTE_MFMA_F16 else 2e-5 if route == capi.ROUTE_MFMA_BF16 else 1e-6
                    assert S.max_rel(got[0], S.oracle_2d(x[0], mode)) < tol
                d_in.free()
                d_out.free()
    finally:
        quiet.close()
        plain.close()


def test_static_call_refuses_batches_beyond_scratch():
    lib = capi.load()
    h = w = 1024
    batch = STATIC_FLOATS // (h * w) + 1
    plan = capi.Plan(h, w)
    try:
        assert plan.set_route(capi.ROUTE_MFMA_F16) == 0
        d = capi.DeviceBuffer(batch * h * w * 4)
        sentinel = np.full(batch * h * w, 3.0, np.float32)
        d.upload(sentinel)
        out = capi.DeviceBuffer(batch * h * w * 4)
        out.upload(sentinel)
        rc = lib.bridct_plan_apply_device_static(plan.ptr, 0, d.ptr, out.ptr, batch)
        assert rc == capi.WORKSPACE_TOO_SMALL
        assert plan.synchronize() == 0
        back = np.empty_like(sentinel)
        out.download(back)
        assert (back == 3.0).all()
        # the largest batch that fits is accepted
        assert lib.bridct_plan_apply_device_static(plan.ptr, 0, d.ptr, out.ptr, batch - 1) == 0
        assert plan.synchronize() == 0
    finally:
        plan.close()

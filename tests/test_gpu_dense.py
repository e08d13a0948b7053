"""GPU tests of the dense split-MFMA route (Y = C_h X C_w^T on the matrix cores from FP16
or BF16 hi + lo operands, three products, FP32 accumulation): error against the FP64
oracle per shape and route, written to gpurun_out/dense_errors.json, and the route's API
rules.  The reference has no dense route; its definition of the transform is
docs/API.md:7-13 (the oracle's Q matrices)."""
import json
import os

import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu

# Error bounds of the split routes, written from the precision of the split: a hi + lo
# pair carries about 22 bits (FP16) or 16 bits (BF16) of each operand, accumulated in FP32.
BOUND = {capi.ROUTE_MFMA_F16: 2e-6, capi.ROUTE_MFMA_BF16: 6e-5}
SHAPES = [(64, 64), (256, 256), (1024, 1024), (64, 256), (256, 64), (128, 1024), (512, 128)]
ERRORS = {}


def _run(plan, mode, x):
    d_in = capi.DeviceBuffer(x.nbytes)
    d_out = capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    assert plan.apply_device(mode, d_in.ptr, d_out.ptr, x.shape[0]) == 0
    assert plan.synchronize() == 0
    y = np.empty_like(x)
    d_out.download(y)
    return y


@pytest.mark.parametrize('route', [capi.ROUTE_MFMA_F16, capi.ROUTE_MFMA_BF16])
@pytest.mark.parametrize('h,w', SHAPES)
def test_dense_route_error(route, h, w):
    batch = 2 if h * w >= 1 << 20 else 4
    x = np.random.default_rng(h * 7 + w).uniform(-1, 1, (batch, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    try:
        assert plan.set_route(route) == 0 and plan.route == route
        print('dense route', route, h, w, flush=True)
        rec = {}
        for mode in (capi.FORWARD, capi.INVERSE):
            y = _run(plan, mode, x)
            ref = S.oracle_2d(x, mode)
            err = float(max(S.max_rel(y[b], ref[b]) for b in range(batch)))
            rms = float(np.sqrt(np.mean((y - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
            rec['fwd' if mode == 0 else 'inv'] = {'max_rel': err, 'rms_rel': rms}
            assert err < BOUND[route], (mode, err)
        back = _run(plan, capi.ROUNDTRIP, x)
        rec['roundtrip_max_rel'] = float(S.max_rel(back, x))
        # forward then inverse: both errors, the forward one relative to max|Y| (about three
        # times max|x| for these inputs)
        assert rec['roundtrip_max_rel'] < 3 * BOUND[route]
        ERRORS['%s %dx%d' % ('f16' if route == capi.ROUTE_MFMA_F16 else 'bf16', h, w)] = rec
    finally:
        plan.close()


def test_dense_route_in_place_and_batches():
    h, w = 256, 128
    x = np.random.default_rng(5).uniform(-1, 1, (6, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    try:
        assert plan.set_route(capi.ROUTE_MFMA_F16) == 0
        y = _run(plan, capi.FORWARD, x)
        d = capi.DeviceBuffer(x.nbytes)
        d.upload(x)
        assert plan.apply_device(capi.FORWARD, d.ptr, d.ptr, 6) == 0
        assert plan.synchronize() == 0
        z = np.empty_like(x)
        d.download(z)
        assert np.array_equal(y, z)
        # host-pointer batch and FORWARD_THEN_INVERSE go through the same route
        out = np.empty_like(x)
        assert plan.apply_batch(capi.FORWARD, x, out) == 0
        assert np.array_equal(out, y)
        assert plan.apply_batch(capi.FORWARD_THEN_INVERSE, x, out) == 0
        assert S.max_rel(out, x) < 4e-6
        # back to the butterflies
        assert plan.set_route(capi.ROUTE_BUTTERFLY) == 0
        assert plan.apply_batch(capi.FORWARD, x, out) == 0
        assert S.max_rel(out, S.oracle_2d(x, 0)) < 1e-6
    finally:
        plan.close()


def test_dense_route_rules():
    L = capi.load()
    small = capi.Plan(32, 256)
    try:
        assert small.set_route(capi.ROUTE_MFMA_F16) == capi.INVALID_ARGUMENT
        assert small.route == capi.ROUTE_BUTTERFLY
    finally:
        small.close()
    p = capi.Plan(64, 64)
    try:
        assert p.set_route(3) == capi.INVALID_ARGUMENT
        assert p.set_route(-1) == capi.INVALID_ARGUMENT
    finally:
        p.close()
    assert L.bridct_plan_set_route(None, capi.ROUTE_MFMA_F16) == capi.INVALID_ARGUMENT


def test_write_dense_error_table():
    out = os.path.join(S.ROOT, 'gpurun_out')
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, 'dense_errors.json'), 'w') as f:
        json.dump(ERRORS, f, indent=1, sort_keys=True)

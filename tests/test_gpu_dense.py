"""GPU tests of the dense split-MFMA route (Y = C_h X C_w^T on the matrix cores from FP16
or BF16 hi + lo operands, three products, FP32 accumulation): error against the FP64
oracle for every shape with both axes >= 64, both directions, at scales 1, 1e-3 and 1e3,
with the butterfly route's error on the same arrays (written as a JSON table to the run's
output folder), exact power-of-two scale equivariance, and the route's API
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
NAMES = {capi.ROUTE_MFMA_F16: 'f16', capi.ROUTE_MFMA_BF16: 'bf16'}
# every shape with both axes >= 64 (the route's domain)
SHAPES = [(h, w) for h in (64, 128, 256, 512, 1024) for w in (64, 128, 256, 512, 1024)]
SCALES = (1.0, 1e-3, 1e3)
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


def _errors(y, ref):
    worst = float(max(S.max_rel(y[b], ref[b]) for b in range(y.shape[0])))
    rms = float(np.sqrt(np.mean((y - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    return worst, rms


@pytest.mark.parametrize('h,w', SHAPES)
def test_dense_route_error(h, w):
    """Both split routes, forward and inverse, at unit scale and at 1e-3 and 1e3: error
    against FP64 within the route's bound, stated beside the butterfly route's on the same
    arrays; and the per-array power-of-two prescale makes the route scale-equivariant
    (2^k X gives exactly 2^k times the result)."""
    batch = 2 if h * w >= 1 << 20 else 4
    base = np.random.default_rng(h * 7 + w).uniform(-1, 1, (batch, h, w)).astype(np.float32)
    plans = {r: capi.Plan(h, w) for r in (capi.ROUTE_BUTTERFLY, capi.ROUTE_MFMA_F16, capi.ROUTE_MFMA_BF16)}
    try:
        for r, p in plans.items():
            assert p.set_route(r) == 0 and p.route == r
        print('dense route %dx%d' % (h, w), flush=True)
        for scale in SCALES:
            x = (base * np.float32(scale)).astype(np.float32)
            for mode in (capi.FORWARD, capi.INVERSE):
                ref = S.oracle_2d(x, mode)
                bf = _errors(_run(plans[capi.ROUTE_BUTTERFLY], mode, x), ref)
                for route in (capi.ROUTE_MFMA_F16, capi.ROUTE_MFMA_BF16):
                    err, rms = _errors(_run(plans[route], mode, x), ref)
                    key = '%s %dx%d %s scale %g' % (NAMES[route], h, w, 'fwd' if mode == 0 else 'inv', scale)
                    ERRORS[key] = {'max_rel': err, 'rms_rel': rms, 'butterfly_max_rel': bf[0],
                                   'butterfly_rms_rel': bf[1], 'ratio_to_butterfly': err / bf[0]}
                    assert err < BOUND[route], (key, err)
        # exact scale equivariance through the power-of-two prescale
        for route in (capi.ROUTE_MFMA_F16, capi.ROUTE_MFMA_BF16):
            y0 = _run(plans[route], capi.FORWARD, base)
            for k in (-20, 17):
                yk = _run(plans[route], capi.FORWARD, np.ldexp(base, k).astype(np.float32))
                assert np.array_equal(yk, np.ldexp(y0, k).astype(np.float32)), (route, k)
        back = _run(plans[capi.ROUTE_MFMA_F16], capi.ROUNDTRIP, base)
        assert S.max_rel(back, base) < 3 * BOUND[capi.ROUTE_MFMA_F16]
    finally:
        for p in plans.values():
            p.close()


def test_dense_route_tiny_and_mixed_scale_arrays():
    """One batch whose arrays span 10^-30 .. 10^30 (and an all-zero array): each array's
    result is within the bound of its own scale; zeros stay exact zeros."""
    h = w = 64
    rng = np.random.default_rng(12)
    scales = [1e-30, 1e-6, 1.0, 1e6, 1e30, 0.0]
    x = np.stack([rng.uniform(-1, 1, (h, w)) * s for s in scales]).astype(np.float32)
    plan = capi.Plan(h, w)
    try:
        assert plan.set_route(capi.ROUTE_MFMA_F16) == 0
        for mode in (capi.FORWARD, capi.INVERSE):
            y = _run(plan, mode, x)
            ref = S.oracle_2d(x, mode)
            for i, s in enumerate(scales):
                if s == 0.0:
                    assert not np.any(y[i])
                else:
                    assert S.max_rel(y[i], ref[i]) < BOUND[capi.ROUTE_MFMA_F16], (mode, s)
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
    out = S.OUT_DIR
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, 'dense_errors.json'), 'w') as f:
        json.dump(ERRORS, f, indent=1, sort_keys=True)

"""GPU tests of launches split by the grid limit.

A launch covers at most 2^32 - 1 work-items per grid dimension.  The whole-array kernels
and the tiny batch kernel size their grids by the batch (one 256-lane workgroup per few
arrays or per item), so calls from about 128 GiB on run in several launches.  The test
hook bridct_debug_grid_limit(k) lowers the limit to k workgroups per launch, so that small
calls take the split path; their results must equal the single-launch results bitwise
(the same kernels on the same arrays), and the last array of each call is checked
against the FP64 oracle.  Contract: reference src/plan.c:126-163 applied per array,
whatever the batch size."""
import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu


def plan_call(plan, mode, x, in_place):
    d_in = capi.DeviceBuffer(x.nbytes)
    d_out = d_in if in_place else capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    assert plan.apply_device(mode, d_in.ptr, d_out.ptr, len(x)) == 0
    assert plan.synchronize() == 0
    y = np.empty_like(x)
    d_out.download(y)
    return y


def batch_call(b, mode, x):
    d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    assert b.apply_device(mode, d_in.ptr, d_out.ptr) == 0 and b.synchronize() == 0
    y = np.empty_like(x)
    d_out.download(y)
    return y


def with_limit(blocks, fn):
    lib = capi.load()
    prev = lib.bridct_debug_grid_limit(blocks)
    try:
        return fn()
    finally:
        assert lib.bridct_debug_grid_limit(prev) == blocks


@pytest.mark.parametrize('h,w,batch,limit', [(8, 8, 5000, 3), (32, 32, 1000, 7), (16, 512, 301, 2),
                                             (64, 128, 97, 5)])
@pytest.mark.parametrize('in_place', [False, True])
def test_whole_array_plan_split(h, w, batch, limit, in_place):
    """Whole-array plan calls in launches of at most `limit` workgroups, all three modes."""
    x = np.random.default_rng(h * w + batch).uniform(-1, 1, (batch, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    try:
        for mode in (0, 1, 2):
            print('whole-array %dx%d x %d, limit %d, mode %d, in place %d' % (h, w, batch, limit, mode, in_place),
                  flush=True)
            one = plan_call(plan, mode, x, in_place)
            split = with_limit(limit, lambda: plan_call(plan, mode, x, in_place))
            assert np.array_equal(one, split), ('split launches differ', h, w, mode)
            exact = S.oracle_2d(x[-1], mode)
            assert S.max_rel(split[-1], exact) < 1e-6
    finally:
        plan.close()


def test_grouped_whole_array_gather_split():
    """A grouped ragged batch (whole-array shapes in gather form beside two-pass shapes)
    with launches of at most 4 workgroups: bitwise equal to the unlimited call."""
    shapes = [(32, 64)] * 300 + [(256, 512)] * 3 + [(128, 8)] * 211 + [(8, 1024)] * 50
    b = capi.Batch(shapes)
    try:
        assert b.set_strategy(capi.BATCH_GROUPED) == 0
        x = np.random.default_rng(5).uniform(-1, 1, b.elements).astype(np.float32)
        for mode in (capi.FORWARD, capi.INVERSE):
            print('grouped gather split, mode %d' % mode, flush=True)
            one = batch_call(b, mode, x)
            split = with_limit(4, lambda: batch_call(b, mode, x))
            assert np.array_equal(one, split), ('split gather launches differ', mode)
    finally:
        b.close()


def test_tiny_batch_split():
    """The tiny batch kernel (every array at most 32 x 32) over 20,000 mixed tiny arrays in
    launches of at most 5 items: bitwise equal to one launch, and the last array of each
    shape against the FP64 oracle."""
    rng = np.random.default_rng(9)
    sizes = [8, 16, 32]
    shapes = [(int(rng.choice(sizes)), int(rng.choice(sizes))) for _ in range(20000)]
    b = capi.Batch(shapes)
    try:
        x = np.random.default_rng(10).uniform(-1, 1, b.elements).astype(np.float32)
        for mode in (capi.FORWARD, capi.INVERSE):
            print('tiny batch split, mode %d' % mode, flush=True)
            one = batch_call(b, mode, x)
            split = with_limit(5, lambda: batch_call(b, mode, x))
            assert np.array_equal(one, split), ('split tiny launches differ', mode)
            last = {}
            for i, s in enumerate(shapes):
                last[s] = i
            for (h, w), i in last.items():
                o = b.offsets[i]
                exact = S.oracle_2d(x[o:o + h * w].reshape(h, w), mode)
                assert S.max_rel(split[o:o + h * w].reshape(h, w), exact) < 1e-6
    finally:
        b.close()

"""GPU parity: every one of the 64 shapes and 3 modes against the FP64 oracle and against
the reference CPU library's output on the same seeded inputs.

Bars:
  * max|y - ref| / max|ref| (vs FP64) at most 2x the reference CPU path's on the same input;
  * ||y_gpu - y_cpu|| / ||y_cpu|| <= 1e-6 against the reference CPU output;
  * the reference's own acceptance rule ||y - ref||_F / ||x||_F <= 2e-5
    (tests/verify_portability.py:53-59) on its shape-sweep families, zero -> exact zero.
"""
import json
import os

import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu

OUT = S.OUT_DIR


def gpu_apply(plan, x, mode):
    y = np.empty_like(x)
    rc = plan.apply_batch(mode, x, y)
    assert rc == 0, rc
    return y


@pytest.fixture(scope='module')
def plans():
    cache = {}
    yield cache
    for p in cache.values():
        p.close()


def get_plan(plans, h, w):
    if (h, w) not in plans:
        plans[(h, w)] = capi.Plan(h, w)
    return plans[(h, w)]


_records = []


@pytest.mark.parametrize('h,w', S.SHAPES)
def test_uniform_error_gate(plans, h, w):
    """uniform[-1,1) batch of 2: 2x reference max-rel error, 1e-6 rel-L2 vs reference."""
    if not S.ref_available():
        pytest.skip('reference library not built (oracle/_ref)')
    rng = np.random.default_rng(1000 + 16 * h + w)
    x = rng.uniform(-1, 1, (2, h, w)).astype(np.float32)
    plan = get_plan(plans, h, w)
    for mode in (0, 1, 2):
        y = gpu_apply(plan, x, mode)
        yr = S.ref_apply(x, mode)
        exact = S.oracle_2d(x, mode)
        e_gpu = S.max_rel(y, exact)
        e_ref = S.max_rel(yr, exact)
        par = S.rel_l2(y, yr)
        _records.append(dict(h=h, w=w, mode=mode, gpu_max_rel=e_gpu, ref_max_rel=e_ref,
                             rel_l2_vs_ref=par))
        assert np.isfinite(y).all()
        assert e_gpu <= 2.0 * e_ref, (h, w, mode, e_gpu, e_ref)
        assert par <= 1e-6, (h, w, mode, par)


@pytest.mark.parametrize('h,w', S.SHAPES)
def test_reference_families(plans, h, w):
    rng = np.random.default_rng(202609201 + h * 1024 + w)
    plan = get_plan(plans, h, w)
    for name, x in S.families(h, w, rng):
        for mode in (0, 1, 2):
            y = gpu_apply(plan, x[None], mode)[0]
            assert np.isfinite(y).all(), (name, mode)
            if name == 'zero':
                assert not np.any(y), (name, mode)
                continue
            exact = S.oracle_2d(x, mode)
            assert S.frob_input_rel(y, exact, x) <= 2e-5, (name, mode)


def test_write_error_table():
    """Writes the per-shape error table of the uniform gate (runs after it)."""
    if not _records:
        pytest.skip('no records')
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, 'parity_errors.json'), 'w') as f:
        json.dump(_records, f, indent=0)

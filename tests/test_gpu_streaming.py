"""GPU tests of the streaming kernel forms against the FP64 oracle.

Calls whose input plus output exceed 256 MiB (the Infinity Cache) run non-temporal
instantiations of the pass kernels and of the tiny batch kernel: the forms every
large-batch throughput figure comes from.  Each test runs one such call at the size the
benchmark uses (or just past the threshold), compares a sample of arrays per shape with
the FP64 oracle (at most 2x the reference CPU library's max-relative error on the same
array, the rule of the 64-shape parity gate; a fixed 4.5e-7 for the 8x8 blocks), and
compares the same arrays bitwise with a call small enough to take the cached forms (the
arithmetic is the same; only the load/store hints differ).  Contract: reference
src/plan.c:126-163 applied per array.
"""
import numpy as np
import pytest

import bench
import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu

STREAM_MIN_BYTES = 256 << 20
# 2x the reference CPU's worst max-relative error against FP64 over the 64 shapes on
# uniform inputs (forward 2.23e-7, inverse 2.14e-7): the fixed bar of the 8x8 blocks.
FWD_INV_TOL = 4.5e-7


def gate(y, x, mode):
    """max-rel error of y against FP64 at most twice the reference CPU's on x (parity rule)."""
    exact = S.oracle_2d(x, mode)
    e_gpu = S.max_rel(y, exact)
    e_ref = S.max_rel(S.ref_apply(np.ascontiguousarray(x, np.float32), mode), exact)
    assert e_gpu <= 2.0 * e_ref, (x.shape, mode, e_gpu, e_ref)
    return e_gpu


def device_call(plan, mode, x, batch, in_place=False):
    d_in = capi.DeviceBuffer(x.nbytes)
    d_out = d_in if in_place else capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    assert plan.apply_device(mode, d_in.ptr, d_out.ptr, batch) == 0
    assert plan.synchronize() == 0
    y = np.empty_like(x)
    d_out.download(y)
    d_in.free()
    d_out.free()
    return y


@pytest.mark.parametrize('batch,in_place', [(40, False), (40, True), (128, False)])
def test_1024_streaming_batch_against_oracle(batch, in_place):
    """40 arrays (320 MiB in + out) and 128 arrays (the dense-route comparison's butterfly
    leg, 1 GiB) of 1024^2: forward, inverse and fused round trip through the streaming
    row and column passes."""
    h = w = 1024
    rng = np.random.default_rng(batch)
    x = rng.uniform(-1, 1, (batch, h, w)).astype(np.float32)
    assert 2 * x.nbytes > STREAM_MIN_BYTES
    sample = [0, batch // 2 + 1, batch - 1]
    plan = capi.Plan(h, w)
    try:
        for mode in (0, 1, 2):
            print('1024^2 x %d mode %d in_place %d' % (batch, mode, in_place), flush=True)
            y = device_call(plan, mode, x, batch, in_place)
            small = device_call(plan, mode, np.ascontiguousarray(x[sample]), len(sample))
            assert np.array_equal(y[sample], small), 'streaming and cached forms differ'
            for i in sample:
                gate(y[i], x[i], mode)
            assert np.isfinite(y).all()
    finally:
        plan.close()


CHUNK_MIN_BYTES = 512 << 20   # two-pass calls with this much intermediate run in chunks
CHUNK_BYTES = 256 << 20


def reference_in_pieces(plan, mode, x, piece=16):
    """The transform of x (batch, h, w) by calls of `piece` arrays, each far below the chunk
    threshold."""
    y = np.empty_like(x)
    for a in range(0, len(x), piece):
        y[a:a + piece] = device_call(plan, mode, np.ascontiguousarray(x[a:a + piece]), len(x[a:a + piece]))
    return y


@pytest.mark.parametrize('in_place', [False, True])
def test_1024_chunked_batch_bitwise(in_place):
    """130 arrays of 1024^2 (520 MiB of intermediate): the call runs in chunks of 64, 64 and 2
    arrays; every array bitwise equal to calls of 16 arrays, forward, inverse and round trip,
    and the last array against the FP64 oracle."""
    h = w = 1024
    batch = 130
    assert batch * h * w * 4 >= CHUNK_MIN_BYTES and batch % (CHUNK_BYTES // (h * w * 4))
    x = np.random.default_rng(130).uniform(-1, 1, (batch, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    try:
        for mode in (0, 1, 2):
            print('1024^2 x %d chunked, mode %d in_place %d' % (batch, mode, in_place), flush=True)
            y = device_call(plan, mode, x, batch, in_place)
            ref = reference_in_pieces(plan, mode, x)
            bad = [i for i in range(batch) if not np.array_equal(y[i], ref[i])]
            assert not bad, ('chunked call differs from 16-array calls', mode, bad[:8])
            gate(y[-1], x[-1], mode)
    finally:
        plan.close()


def test_grouped_chunked_gather_bitwise():
    """A ragged batch with 140 arrays of 1024^2 (560 MiB of intermediate in that group: gather
    calls in chunks of 64, 64 and 12) and a few 256x512 and 64x64 fillers, grouped strategy:
    every array bitwise equal to a plan call on 16 arrays of its shape, forward, inverse and
    fused round trip."""
    shapes = [(1024, 1024)] * 70 + [(256, 512)] * 5 + [(1024, 1024)] * 70 + [(64, 64)] * 7
    b = capi.Batch(shapes)
    try:
        assert b.set_strategy(capi.BATCH_GROUPED) == 0
        x = np.random.default_rng(140).uniform(-1, 1, b.elements).astype(np.float32)
        d_x, d_y = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
        d_x.upload(x)
        for mode in (capi.FORWARD, capi.INVERSE, capi.ROUNDTRIP):
            print('grouped 140 x 1024^2 chunked, mode %d' % mode, flush=True)
            assert b.apply_device(mode, d_x.ptr, d_y.ptr) == 0 and b.synchronize() == 0
            y = np.empty_like(x)
            d_y.download(y)
            for shape in sorted(set(shapes)):
                idx = [i for i, s in enumerate(shapes) if s == shape]
                h, w = shape
                xs = np.stack([x[b.offsets[i]:b.offsets[i] + h * w].reshape(h, w) for i in idx])
                plan = capi.Plan(h, w)
                try:
                    ref = reference_in_pieces(plan, mode, xs)
                finally:
                    plan.close()
                for k, i in enumerate(idx):
                    got = y[b.offsets[i]:b.offsets[i] + h * w].reshape(h, w)
                    assert np.array_equal(got, ref[k]), ('grouped chunked result differs', shape, mode, i)
        d_x.free()
        d_y.free()
    finally:
        b.close()


def test_ragged_10000_grouped_streaming_against_oracle():
    """The ragged mix of 10,000 arrays (seed 7, 2.7 GB per buffer): the automatic strategy
    (grouped, per-shape kernels in gather form, streaming hints) in one call; for every
    distinct shape the first array against the FP64 oracle and, bitwise, against a plan of
    that shape on that array alone; every array's round trip back to its input."""
    shapes, rng = bench.ragged_shapes(10000, 7)
    b = capi.Batch(shapes)
    try:
        assert b.strategy == capi.BATCH_GROUPED
        x = bench.ragged_values(shapes, b.offsets, b.elements, rng)
        assert 2 * x.nbytes > STREAM_MIN_BYTES
        d_x, d_y, d_z = (capi.DeviceBuffer(x.nbytes) for _ in range(3))
        d_x.upload(x)
        print('ragged 10000 forward', flush=True)
        assert b.apply_device(capi.FORWARD, d_x.ptr, d_y.ptr) == 0
        print('ragged 10000 inverse', flush=True)
        assert b.apply_device(capi.INVERSE, d_y.ptr, d_z.ptr) == 0
        assert b.synchronize() == 0
        y = np.empty_like(x)
        z = np.empty_like(x)
        d_y.download(y)
        d_z.download(z)
        # graph replays of forward-then-inverse (the large shapes on the batch stream, the
        # small ones on the auxiliary streams, all captured) give the eager round trip's bits
        for rep in range(2):
            print('ragged 10000 graph replay', rep, flush=True)
            assert b.graph_apply(capi.FORWARD_THEN_INVERSE, d_x.ptr, d_y.ptr) == 0 and b.synchronize() == 0
            g = np.empty_like(x)
            d_y.download(g)
            assert np.array_equal(g, z), 'graph replay differs from the eager round trip'
            del g
            assert b.apply_device(capi.FORWARD, d_x.ptr, d_y.ptr) == 0 and b.synchronize() == 0
        for d in (d_x, d_y, d_z):
            d.free()
        first = {}
        for i, s in enumerate(shapes):
            first.setdefault(s, i)
        assert len(first) == 64
        plans = {}
        worst = 0.0
        for (h, w), i in sorted(first.items()):
            o = b.offsets[i]
            a = x[o:o + h * w].reshape(h, w)
            got = y[o:o + h * w].reshape(h, w)
            worst = max(worst, gate(got, a, 0))
            gate(z[o:o + h * w].reshape(h, w), got, 1)
            p = plans.setdefault((h, w), capi.Plan(h, w))
            one = np.empty_like(a)
            assert p.apply_batch(capi.FORWARD, np.ascontiguousarray(a), one) == 0
            assert np.array_equal(one, got), ('grouped streaming result differs from the plan', h, w)
        for p in plans.values():
            p.close()
        for o, (h, w) in zip(b.offsets, shapes):
            a, r = x[o:o + h * w], z[o:o + h * w]
            assert np.abs(r - a).max() <= 1.5e-6 * np.abs(a).max(), (h, w)
        print('ragged worst forward max-rel vs FP64 %.3e' % worst, flush=True)
    finally:
        b.close()


def test_blocks_8x8_streaming_against_oracle():
    """1,048,576 8x8 blocks (256 MiB per buffer, pixel values 0..255 minus 128) through the
    tiny batch kernel's streaming form: every block against the FP64 oracle (one GEMM with
    the Kronecker product Q_8 (x) Q_8), and the first 65,536 blocks bitwise against a
    cached-form call of those blocks alone."""
    nb = 1 << 20
    x = bench.blocks_values(nb, 4)
    assert 2 * x.nbytes > STREAM_MIN_BYTES
    q = S.qmatrix(8)
    kron = np.kron(q, q)                     # vec(Q X Q^T) = (Q (x) Q) vec(X), row-major
    sub = 65536
    big = capi.Batch([(8, 8)] * nb)
    small = capi.Batch([(8, 8)] * sub)
    try:
        for mode, mat in ((capi.FORWARD, kron), (capi.INVERSE, kron.T)):
            print('8x8 blocks mode', mode, flush=True)
            d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
            d_in.upload(x)
            assert big.apply_device(mode, d_in.ptr, d_out.ptr) == 0
            assert big.synchronize() == 0
            y = np.empty_like(x)
            d_out.download(y)
            d_in.free()
            d_out.free()
            ref = x.reshape(nb, 64).astype(np.float64) @ mat.T
            got = y.reshape(nb, 64)
            diff = np.abs(got - ref).max(axis=1)
            scale = np.abs(ref).max(axis=1)
            assert (scale > 0).all()
            assert float((diff / scale).max()) < FWD_INV_TOL, float((diff / scale).max())
            part = np.ascontiguousarray(x[:sub * 64])
            d_in, d_out = capi.DeviceBuffer(part.nbytes), capi.DeviceBuffer(part.nbytes)
            d_in.upload(part)
            assert small.apply_device(mode, d_in.ptr, d_out.ptr) == 0
            assert small.synchronize() == 0
            yp = np.empty_like(part)
            d_out.download(yp)
            d_in.free()
            d_out.free()
            assert np.array_equal(yp, y[:sub * 64]), 'streaming and cached tiny kernels differ'
    finally:
        big.close()
        small.close()

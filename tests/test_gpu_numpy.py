"""GPU tests of the NumPy extension, following the reference's own validation of its
NumPy interface (python/test_api.py): rejected inputs and their exception types, output
ownership and `out=` guards, closed / uninitialized / reinitialized plans, numerical error
against the FP64 DCT (scipy) on the reference's seeded families, roundtrip ==
inverse(forward) bitwise, independent plans in threads, and the busy rejection of a plan
shared by two threads.  Plus the RaggedPlan list entry and the Codec8 block codec step, which
the reference does not have."""
import concurrent.futures
import gc
import threading

import numpy as np
import pytest
from scipy import fft

pytestmark = pytest.mark.gpu


def rejected(func, error):
    with pytest.raises(error):
        func()


def test_rejected_inputs(npmod):
    """python/test_api.py:51-83."""
    plan = npmod.Plan(8, 8)
    x = np.arange(64, dtype=np.float32).reshape(8, 8).copy()
    for bad in (x.astype(np.float64), x.astype(np.float16), x.astype(np.int32),
                x.astype(np.dtype('f4').newbyteorder('S')), x.tolist(), np.ma.array(x)):
        rejected(lambda bad=bad: plan.forward(bad), TypeError)
    for bad in (x.reshape(-1), x.reshape(1, 1, 8, 8), np.zeros((16, 8), np.float32),
                np.zeros((0, 8, 8), np.float32), x.T, x[::-1],
                np.ndarray((8, 8), np.float32, buffer=bytearray(257), offset=1)):
        rejected(lambda bad=bad: plan.forward(bad), ValueError)
    rejected(lambda: plan.forward(x, out=x), ValueError)
    storage = np.arange(65, dtype=np.float32)
    rejected(lambda: plan.forward(storage[:-1].reshape(8, 8), out=storage[1:].reshape(8, 8)), ValueError)
    readonly = np.empty_like(x)
    readonly.flags.writeable = False
    rejected(lambda: plan.forward(x, out=readonly), ValueError)
    rejected(lambda: plan.forward(x, out=np.empty((8, 8))), TypeError)
    rejected(lambda: plan.forward(x, out=np.empty((1, 8, 8), np.float32)), ValueError)
    rejected(lambda: plan.forward(x, np.empty_like(x)), TypeError)
    rejected(lambda: plan.forward(x, norm='ortho'), TypeError)
    rejected(lambda: plan.forward(), TypeError)
    rejected(lambda: plan.forward(input=x), TypeError)
    rejected(lambda: plan.__init__(16, 16), RuntimeError)
    for shape in ((4, 8), (26, 32), (8, 2048), (0, 8), (-8, 8)):
        rejected(lambda shape=shape: npmod.Plan(*shape), ValueError)
    # read-only input is fine; outputs are fresh, owned and unshared (python/test_api.py:84-95)
    ro = x.copy()
    ro.flags.writeable = False
    assert np.array_equal(plan.forward(ro), plan.forward(x))
    y, z = plan.forward(x), plan.forward(x)
    assert y is not z and y.flags.owndata and z.flags.owndata
    assert not np.shares_memory(x, y) and not np.shares_memory(y, z)
    del plan
    gc.collect()
    assert np.isfinite(y).all()


def test_close_and_uninitialized(npmod):
    """python/test_api.py:96-105."""
    x = np.zeros((8, 8), np.float32)
    plan = npmod.Plan(8, 8)
    assert plan.route == 'paper8-square'
    plan.close()
    plan.close()
    assert plan.closed and plan.route is None
    rejected(lambda: plan.forward(x), RuntimeError)
    u = npmod.Plan.__new__(npmod.Plan)
    rejected(lambda: u.forward(x), RuntimeError)
    u.close()


SHAPES = [(n, n) for n in (8, 16, 32, 64, 128, 256, 512, 1024)] + [(16, 32), (8, 64), (1024, 512), (512, 1024)]


@pytest.mark.parametrize('h,w', SHAPES)
def test_numerics_and_out_contract(npmod, h, w):
    """python/test_api.py:106-143: seeded signed / impulse / near-constant families, error
    (input-relative Frobenius) <= 2e-5 against scipy's FP64 DCT, `out=` fills a guarded
    buffer exactly, inputs unchanged, roundtrip == inverse(forward) bitwise."""
    plan = npmod.Plan(h, w)
    assert plan.shape == (h, w) and not plan.closed
    for batch in (1, 4):
        shape = (h, w) if batch == 1 else (batch, h, w)
        rng = np.random.default_rng(210920 + h * 1024 + w + batch)
        signed = rng.standard_normal(shape).astype(np.float32)
        impulse = np.zeros(shape, np.float32)
        impulse[..., 1, 2] = 1
        for values in (signed, impulse, (1 + signed * 1e-5).astype(np.float32)):
            initial = values.copy()
            for mode, method in enumerate((plan.forward, plan.inverse, plan.roundtrip)):
                ref = ((fft.dctn if mode == 0 else fft.idctn)(values.astype(np.float64), type=2, axes=(-2, -1),
                                                              norm='ortho', workers=1)
                       if mode != 2 else values.astype(np.float64))
                actual = method(values)
                error = float(np.linalg.norm(actual.astype(np.float64) - ref) /
                              max(np.linalg.norm(values.astype(np.float64)), np.finfo(float).tiny))
                assert actual.shape == values.shape and actual.dtype == np.float32
                assert np.isfinite(actual).all() and error <= 2e-5
                assert actual.flags.owndata and not np.shares_memory(actual, values)
                assert np.array_equal(initial, values)
                guarded = np.full(actual.size + 8, np.float32(-81723))
                out = guarded[4:-4].reshape(shape)
                returned = method(values, out=out)
                assert returned is out and np.array_equal(actual, out)
                assert np.all(guarded[:4] == -81723) and np.all(guarded[-4:] == -81723)
                if mode == 2:
                    assert np.array_equal(actual, plan.inverse(plan.forward(values)))
    plan.close()


def test_threads(npmod):
    """python/test_api.py:145-182: independent plans in threads; a plan shared by two
    threads either completes or rejects the second caller as busy."""
    def independent(seed):
        values = np.random.default_rng(seed).standard_normal((128, 256)).astype(np.float32)
        worker = npmod.Plan(128, 256)
        actual = worker.roundtrip(values)
        worker.close()
        return float(np.linalg.norm(actual.astype(float) - values) / np.linalg.norm(values))

    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as pool:
        assert max(pool.map(independent, range(8))) <= 2e-5

    shared = npmod.Plan(1024, 1024)
    values = np.ones((8, 1024, 1024), np.float32)
    barrier = threading.Barrier(2)

    def overlapping():
        barrier.wait()
        try:
            result = shared.roundtrip(values)
            assert np.max(np.abs(result - 1)) < 2e-5
            return 'completed'
        except RuntimeError as error:
            assert 'another thread' in str(error)
            return 'busy-rejected'

    attempts = []
    for _ in range(3):
        with concurrent.futures.ThreadPoolExecutor(max_workers=2) as pool:
            outcome = [f.result() for f in [pool.submit(overlapping) for _ in range(2)]]
        attempts.append(outcome)
        if 'completed' in outcome and 'busy-rejected' in outcome:
            break
    assert all('completed' in a for a in attempts)
    shared.close()


def test_ragged_plan(npmod):
    rng = np.random.default_rng(31)
    lengths = (8, 16, 32, 64, 128, 256, 512, 1024)
    shapes = [(int(rng.choice(lengths)), int(rng.choice(lengths))) for _ in range(40)]
    arrays = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    rp = npmod.RaggedPlan(shapes)
    try:
        assert rp.shapes == tuple(shapes) and not rp.closed
        plans = {s: npmod.Plan(*s) for s in set(shapes)}
        fwd = rp.forward(arrays)
        inv = rp.inverse(arrays)
        rt = rp.roundtrip(arrays)
        for s, a, f, i, r in zip(shapes, arrays, fwd, inv, rt):
            assert f.shape == s and f.dtype == np.float32 and f.flags.owndata
            # same transform as a Plan of that shape (the kernels may differ: whole-array
            # vs row/column units, so agreement is to FP32 rounding, not bitwise)
            for got, want in ((f, plans[s].forward(a)), (i, plans[s].inverse(a)), (r, plans[s].roundtrip(a))):
                assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max(), s
        # roundtrip materializes the spectrum: bitwise inverse(forward(x)) of the same object
        again = rp.inverse(fwd)
        assert all(np.array_equal(a, b) for a, b in zip(rt, again))
        outs = [np.empty(s, np.float32) for s in shapes]
        back = rp.forward(tuple(arrays), out=outs)
        assert all(o is b for o, b in zip(outs, back))
        assert all(np.array_equal(o, f) for o, f in zip(outs, fwd))
        for p in plans.values():
            p.close()
        rejected(lambda: rp.forward(arrays[:-1]), ValueError)
        rejected(lambda: rp.forward(arrays[0]), TypeError)
        rejected(lambda: rp.forward([a.astype(np.float64) for a in arrays]), TypeError)
        wrong = np.zeros((8 if shapes[0][0] != 8 else 16, shapes[0][1]), np.float32)
        rejected(lambda: rp.forward([wrong] + arrays[1:]), ValueError)
        rejected(lambda: rp.forward(arrays, out=arrays), ValueError)
        rejected(lambda: rp.forward(arrays, scale=2), TypeError)
        rejected(lambda: rp.__init__(shapes), RuntimeError)
    finally:
        rp.close()
    rp.close()
    assert rp.closed
    rejected(lambda: rp.forward(arrays), RuntimeError)
    rejected(lambda: npmod.RaggedPlan([]), ValueError)
    rejected(lambda: npmod.RaggedPlan([(8, 8), (12, 8)]), ValueError)
    rejected(lambda: npmod.RaggedPlan([8, 8]), TypeError)


def test_codec8(npmod):
    """Codec8: the 8x8 block codec step of the C ABI (bridct_codec8_*) from NumPy: results
    equal the C entry points bitwise, `out=` contract, strided rows, rejected inputs."""
    import support as S
    from bridct_amd import capi
    rng = np.random.default_rng(12)
    img = rng.integers(0, 256, (64, 96), dtype=np.uint8)
    codec = npmod.Codec8(S.JPEG_LUMA.reshape(8, 8))
    ref = capi.Codec8(S.JPEG_LUMA)
    try:
        assert not codec.closed and np.array_equal(codec.quant, S.JPEG_LUMA.reshape(8, 8))
        co = codec.encode(img)
        assert co.shape == (8, 12, 8, 8) and co.dtype == np.int16 and co.flags.owndata
        assert np.array_equal(co.reshape(-1, 64), ref.encode(img))
        pix = codec.decode(co)
        assert pix.shape == (64, 96) and pix.dtype == np.uint8
        assert np.array_equal(pix, ref.decode(co.reshape(-1, 64), 64, 96))
        # strided rows in and out
        big = np.zeros((80, 128), np.uint8)
        big[8:72, 16:112] = img
        assert np.array_equal(codec.encode(big[8:72, 16:112]), co)
        canvas = np.full((64, 120), 9, np.uint8)
        assert codec.decode(co, out=canvas[:, 8:104]) is not None
        assert np.array_equal(canvas[:, 8:104], pix) and np.all(canvas[:, :8] == 9) and np.all(canvas[:, 104:] == 9)
        out = np.empty_like(co)
        assert codec.encode(img, out=out) is out and np.array_equal(out, co)
        # rejected inputs
        for bad in (img.astype(np.int16), img.tolist()):
            rejected(lambda bad=bad: codec.encode(bad), TypeError)
        for bad in (img[:60], img[:, :90], img[:, ::2], img.reshape(1, 64, 96), np.zeros((0, 8), np.uint8)):
            rejected(lambda bad=bad: codec.encode(bad), ValueError)
        rejected(lambda: codec.decode(co.astype(np.int32)), TypeError)
        rejected(lambda: codec.decode(co[:, :, :4]), ValueError)
        rejected(lambda: codec.decode(co[:, ::2]), ValueError)
        rejected(lambda: codec.encode(img, out=np.empty((8, 12, 8, 8), np.int32)), TypeError)
        rejected(lambda: codec.encode(img, out=np.empty((8, 11, 8, 8), np.int16)), ValueError)
        storage = np.zeros(64 * 96 * 3, np.uint8)
        view_img = storage[:64 * 96].reshape(64, 96)
        view_co = storage[64:64 + 64 * 96 * 2].view(np.int16).reshape(8, 12, 8, 8)
        rejected(lambda: codec.encode(view_img, out=view_co), ValueError)
        ro = np.empty_like(co)
        ro.flags.writeable = False
        rejected(lambda: codec.encode(img, out=ro), ValueError)
        rejected(lambda: codec.encode(img, scale=2), TypeError)
        rejected(lambda: codec.__init__(S.JPEG_LUMA), RuntimeError)
    finally:
        codec.close()
        ref.close()
    codec.close()
    assert codec.closed
    rejected(lambda: codec.encode(img), RuntimeError)
    rejected(lambda: npmod.Codec8(np.zeros(64, np.float32)), ValueError)
    rejected(lambda: npmod.Codec8(np.ones(64)), TypeError)
    rejected(lambda: npmod.Codec8(np.ones(60, np.float32)), ValueError)

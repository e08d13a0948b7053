"""GPU tests of ragged batches over arrays of mixed shapes, with both strategies (one
persistent launch, and the grouped per-shape kernels): each array's result equals the plan
path of its shape (reference src/plan.c:126-163 applied per array) -- bitwise for the
grouped strategy, which runs the plan path's kernels -- and the FP64 oracle; graph
replays equal eager calls."""
import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu

FTI = 3   # BRIDCT_GPU_FORWARD_THEN_INVERSE
STRATEGIES = [capi.BATCH_PERSISTENT, capi.BATCH_GROUPED]


def agree(y, ref, strategy, tol=1e-6):
    """Grouped results are the plan path's bitwise; persistent ones agree to FP32
    rounding (some shapes use other kernels there)."""
    if strategy == capi.BATCH_GROUPED:
        return np.array_equal(y, ref)
    return np.abs(y - ref).max() <= tol * np.abs(ref).max()


def ragged(shapes, seed):
    rng = np.random.default_rng(seed)
    return [(rng.uniform(-1, 1, (h, w)) * 10 ** rng.uniform(-3, 3)).astype(np.float32) for h, w in shapes]


def flat(arrays):
    return np.concatenate([a.ravel() for a in arrays]).astype(np.float32)


def plan_results(shapes, arrays, mode):
    out = []
    plans = {}
    for (h, w), a in zip(shapes, arrays):
        if (h, w) not in plans:
            plans[(h, w)] = capi.Plan(h, w)
        y = np.empty_like(a)
        assert plans[(h, w)].apply_batch(mode, a, y) == 0
        out.append(y)
    for p in plans.values():
        p.close()
    return out


def run_device(batch, mode, x, graph=False, in_place=False):
    d_in = capi.DeviceBuffer(x.nbytes)
    d_out = d_in if in_place else capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    fn = batch.graph_apply if graph else batch.apply_device
    assert fn(mode, d_in.ptr, d_out.ptr) == 0
    assert batch.synchronize() == 0
    y = np.empty_like(x)
    d_out.download(y)
    return y


def all_shapes():
    return [(h, w) for h in S.LENGTHS for w in S.LENGTHS]


@pytest.mark.parametrize('strategy', STRATEGIES)
@pytest.mark.parametrize('mode', [0, 1, 2, FTI])
def test_every_shape_once(mode, strategy):
    shapes = all_shapes()
    arrays = ragged(shapes, 11)
    b = capi.Batch(shapes)
    try:
        assert b.set_strategy(strategy) == 0 and b.strategy == strategy
        print('batch of 64 shapes, mode', mode, 'strategy', strategy, flush=True)
        y = b.split(run_device(b, mode, flat(arrays)))
        ref = plan_results(shapes, arrays, mode)
        for (h, w), a, yi, ri in zip(shapes, arrays, y, ref):
            assert agree(yi, ri, strategy), (h, w)
            if mode in (0, 1):
                o = S.oracle_2d(a, mode)
                assert S.max_rel(yi, o) < 1e-6, (h, w)
            else:
                assert np.abs(yi - a).max() <= 2e-6 * np.abs(a).max(), (h, w)
    finally:
        b.close()


@pytest.mark.parametrize('strategy', STRATEGIES)
def test_ragged_seeded_mix(strategy):
    rng = np.random.default_rng(7)
    shapes = [(int(rng.choice(S.LENGTHS)), int(rng.choice(S.LENGTHS))) for _ in range(400)]
    arrays = ragged(shapes, 7)
    b = capi.Batch(shapes)
    try:
        assert b.set_strategy(strategy) == 0
        x = flat(arrays)
        for mode in (0, 1, 2):
            print('ragged 400 arrays, mode', mode, 'strategy', strategy, flush=True)
            y = b.split(run_device(b, mode, x))
            ref = plan_results(shapes, arrays, mode)
            for (h, w), yi, ri in zip(shapes, y, ref):
                assert agree(yi, ri, strategy), (h, w)
        print('ragged 400 arrays, in place', flush=True)
        y_in_place = run_device(b, 0, x, in_place=True)
        assert np.array_equal(y_in_place, run_device(b, 0, x))
    finally:
        b.close()


def test_uniform_small_blocks():
    """8x8 blocks of pixel values (integers 0..255 minus 128) in one launch."""
    n = 65536
    rng = np.random.default_rng(3)
    x = (rng.integers(0, 256, (n, 8, 8)) - 128).astype(np.float32)
    b = capi.Batch([(8, 8)] * n)
    p = capi.Plan(8, 8)
    try:
        for mode in (0, 1, 2):
            print('8x8 blocks, mode', mode, flush=True)
            y = run_device(b, mode, x.ravel()).reshape(n, 8, 8)
            r = np.empty_like(x)
            assert p.apply_batch(mode, x, r) == 0
            assert np.abs(y - r).max() <= 1e-6 * np.abs(r).max()
    finally:
        b.close()
        p.close()


def test_tiny_kernel_mix():
    """Every array at most 32 x 32: the batch runs the tiny persistent kernel (smaller LDS
    area, more resident workgroups); results equal the plan path, eager and replayed."""
    rng = np.random.default_rng(13)
    lengths = (8, 16, 32)
    shapes = [(int(rng.choice(lengths)), int(rng.choice(lengths))) for _ in range(5000)]
    arrays = ragged(shapes, 13)
    b = capi.Batch(shapes)
    try:
        x = flat(arrays)
        for mode in (0, 1, 2, FTI):
            ref = plan_results(shapes, arrays, mode) if mode != FTI else arrays
            for graph in (False, True, True):
                print('tiny mix of 5000 arrays, mode', mode, 'graph', graph, flush=True)
                y = b.split(run_device(b, mode, x, graph=graph))
                for (h, w), yi, ri in zip(shapes, y, ref):
                    tol = 1e-6 if mode != FTI else 2e-6
                    assert np.abs(yi - ri).max() <= tol * np.abs(ri).max(), (h, w)
    finally:
        b.close()


@pytest.mark.parametrize('strategy', STRATEGIES)
def test_large_arrays_with_small_fillers(strategy):
    shapes = [(1024, 1024)] * 3 + [(16, 16)] * 500 + [(512, 1024), (8, 1024), (1024, 8)]
    arrays = ragged(shapes, 5)
    b = capi.Batch(shapes)
    try:
        assert b.set_strategy(strategy) == 0
        for mode in (2, FTI):
            print('large + fillers, mode', mode, flush=True)
            y = b.split(run_device(b, mode, flat(arrays)))
            for (h, w), a, yi in zip(shapes, arrays, y):
                assert np.abs(yi - a).max() <= 2e-6 * np.abs(a).max(), (h, w)
    finally:
        b.close()


def test_persistent_round_trip_repeatable_at_scale():
    """The persistent kernel's fused round trip on a batch large enough that a CU runs column
    strips and then inverse row bands of the same array: the ROW2 items must see every
    column item's stores (neighbouring strips share cache lines), so repeated calls give
    the same bits and each array comes back to within 2e-6 of itself.  (A workgroup-scope
    L1 invalidation before the ROW2 items had left stale lines: the results changed from
    call to call on the 10,000-array mix.)"""
    rng = np.random.default_rng(11)
    lengths = [256, 512, 1024]
    shapes = [(int(rng.choice(lengths)), int(rng.choice(lengths))) for _ in range(300)] + [(16, 16)] * 64
    arrays = ragged(shapes, 12)
    x = flat(arrays)
    b = capi.Batch(shapes)
    try:
        assert b.set_strategy(capi.BATCH_PERSISTENT) == 0
        d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
        d_in.upload(x)
        first = None
        for rep in range(3):
            print('persistent round trip at scale, call', rep, flush=True)
            assert b.apply_device(2, d_in.ptr, d_out.ptr) == 0 and b.synchronize() == 0
            y = np.empty_like(x)
            d_out.download(y)
            if first is None:
                first = y
                for (h, w), a, yi in zip(shapes, arrays, b.split(y)):
                    assert np.abs(yi - a).max() <= 2e-6 * np.abs(a).max(), (h, w)
            else:
                assert np.array_equal(y, first)
    finally:
        b.close()


@pytest.mark.parametrize('strategy', STRATEGIES)
def test_graph_replays_match_eager(strategy):
    rng = np.random.default_rng(9)
    shapes = [(int(rng.choice(S.LENGTHS)), int(rng.choice(S.LENGTHS))) for _ in range(60)]
    arrays = ragged(shapes, 9)
    x = flat(arrays)
    b = capi.Batch(shapes)
    d_in = capi.DeviceBuffer(x.nbytes)
    d_out = capi.DeviceBuffer(x.nbytes)
    try:
        assert b.set_strategy(strategy) == 0
        d_in.upload(x)
        eager = {}
        for mode in (0, FTI):
            assert b.apply_device(mode, d_in.ptr, d_out.ptr) == 0
            assert b.synchronize() == 0
            eager[mode] = np.empty_like(x)
            d_out.download(eager[mode])
        for rep in range(4):
            for mode in (0, FTI):
                print('graph replay', rep, 'mode', mode, flush=True)
                assert b.graph_apply(mode, d_in.ptr, d_out.ptr) == 0
                assert b.synchronize() == 0
                y = np.empty_like(x)
                d_out.download(y)
                bad = [(i, shape) for i, (shape, a, e) in enumerate(zip(shapes, b.split(y), b.split(eager[mode])))
                       if not np.array_equal(a, e)]
                assert not bad, 'arrays differing from the eager result: %s' % bad
                # eager work between replays
                assert b.apply_device(1, d_in.ptr, d_out.ptr) == 0
                assert b.synchronize() == 0
    finally:
        b.close()


@pytest.mark.parametrize('shapes', [[(512, 512), (8, 8)], [(1024, 1024)] * 5 + [(8, 8), (32, 64), (256, 128)]],
                         ids=['one-partial-chunk', 'three-chunks'])
def test_host_apply_staged(shapes):
    """Host calls of 1 MiB or more go through pinned staging in chunks of 8 MiB: a single
    partial chunk, and several chunks with a partial last one, equal the device path
    bitwise and leave the input unchanged."""
    arrays = ragged(shapes, 5)
    x = flat(arrays)
    x0 = x.copy()
    b = capi.Batch(shapes)
    try:
        for mode in (0, 1, FTI):
            print('staged host batch %d arrays mode %d' % (len(shapes), mode), flush=True)
            y = np.full_like(x, np.nan)
            assert b.apply(mode, x, y) == 0
            assert np.array_equal(y, run_device(b, mode, x))
            assert np.array_equal(x, x0)
    finally:
        b.close()


@pytest.mark.parametrize('strategy', STRATEGIES)
def test_host_apply_matches_device(strategy):
    shapes = [(64, 256), (256, 64), (8, 8), (1024, 32), (128, 128)]
    arrays = ragged(shapes, 2)
    x = flat(arrays)
    b = capi.Batch(shapes)
    try:
        assert b.set_strategy(strategy) == 0
        for mode in (0, 1, 2, FTI):
            y = np.empty_like(x)
            assert b.apply(mode, x, y) == 0
            assert np.array_equal(y, run_device(b, mode, x))
    finally:
        b.close()


def test_strategy_selection_and_switching():
    """AUTO resolves to GROUPED for large mixes (at least 2^21 pixels per distinct shape on
    average) and to PERSISTENT otherwise; switching strategies between graph replays
    recaptures the graph; a tiny batch also runs grouped."""
    big = capi.Batch([(1024, 1024)] * 4 + [(512, 256)] * 20)
    small = capi.Batch([(64, 64), (128, 32)] * 10)
    tiny = capi.Batch([(8, 8), (16, 32)] * 100)
    try:
        assert big.strategy == capi.BATCH_GROUPED
        assert small.strategy == capi.BATCH_PERSISTENT and tiny.strategy == capi.BATCH_PERSISTENT
        for bad in (-1, 3):
            assert big.set_strategy(bad) == capi.INVALID_ARGUMENT
        assert big.set_strategy(capi.BATCH_PERSISTENT) == 0 and big.strategy == capi.BATCH_PERSISTENT
        assert big.set_strategy(capi.BATCH_AUTO) == 0 and big.strategy == capi.BATCH_GROUPED
        x = flat(ragged(big.shapes, 21))
        d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
        d_in.upload(x)
        outs = []
        for strategy in (capi.BATCH_GROUPED, capi.BATCH_PERSISTENT, capi.BATCH_GROUPED):
            print('graph with strategy', strategy, flush=True)
            assert big.set_strategy(strategy) == 0
            assert big.graph_apply(0, d_in.ptr, d_out.ptr) == 0
            assert big.synchronize() == 0
            y = np.empty_like(x)
            d_out.download(y)
            outs.append(y)
        ref = flat(plan_results(big.shapes, big.split(x), 0))
        assert np.array_equal(outs[0], ref) and np.array_equal(outs[2], ref)
        assert np.abs(outs[1] - ref).max() <= 1e-6 * np.abs(ref).max()
        xt = flat(ragged(tiny.shapes, 22))
        assert tiny.set_strategy(capi.BATCH_GROUPED) == 0
        yt = run_device(tiny, 0, xt)
        assert np.array_equal(yt, flat(plan_results(tiny.shapes, tiny.split(xt), 0)))
    finally:
        big.close()
        small.close()
        tiny.close()


def test_batch_argument_errors():
    L = capi.load()
    hs = np.array([8, 12], np.int32)
    ws = np.array([8, 8], np.int32)
    assert not L.bridct_batch_create(2, capi.ptr(hs), capi.ptr(ws))
    assert not L.bridct_batch_create(0, capi.ptr(hs), capi.ptr(ws))
    b = capi.Batch([(8, 8), (16, 16)])
    try:
        d = capi.DeviceBuffer(4 * 320 + 64)
        assert b.apply_device(4, d.ptr, d.ptr) == capi.INVALID_ARGUMENT
        assert b.apply_device(0, d.ptr + 4, d.ptr + 4) == capi.BAD_ALIGNMENT
        assert b.apply_device(0, d.ptr, d.ptr + 16) == capi.OVERLAPPING_BUFFERS
        assert b.synchronize() == 0
    finally:
        b.close()


def test_wait_bound_reports_error_and_recovers():
    """A dependency wait that runs out sets the error flag instead of hanging: with a bound
    of one sleep, the column items of a large array (queued right behind its rows) cannot
    all find the rows done, synchronize reports BRIDCT_DEVICE_ERROR, and the next call with
    the normal bound is correct again (the flag is reset by each launch)."""
    shapes = [(1024, 1024)]
    x = flat(ragged(shapes, 4))
    b = capi.Batch(shapes)
    d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    try:
        d_in.upload(x)
        assert b.set_spin_limit(0) == capi.INVALID_ARGUMENT
        assert b.set_spin_limit(1) == 0
        print('tiny wait bound', flush=True)
        assert b.apply_device(0, d_in.ptr, d_out.ptr) == 0
        assert b.synchronize() == capi.DEVICE_ERROR
        assert b.set_spin_limit(1 << 21) == 0
        assert b.apply_device(0, d_in.ptr, d_out.ptr) == 0
        assert b.synchronize() == 0
        y = np.empty_like(x)
        d_out.download(y)
        p = capi.Plan(1024, 1024)
        r = np.empty_like(x)
        assert p.apply_batch(0, x, r) == 0
        p.close()
        assert np.abs(y - r).max() <= 1e-6 * np.abs(r).max()
    finally:
        b.close()


def test_grouped_follows_table_routes(tmp_path, monkeypatch):
    """A schedule table that sends shapes to the split-MFMA routes: plans take those routes,
    a grouped batch runs the same dense GEMMs in gather form and equals the per-shape plans
    bitwise (eager and graph replay, every mode), the automatic strategy is GROUPED, and
    the persistent kernel, which runs butterflies only, agrees with butterfly plans to FP32
    rounding.  A dense entry for a shape with an axis below 64 is ignored by both."""
    table = tmp_path / 'schedule.txt'
    table.write_text('64 64 mfma_f16 0\n128 256 mfma_f16 0\n64 1024 mfma_bf16 0\n512 64 mfma_f16 0\n'
                     '32 64 mfma_f16 0\n')
    monkeypatch.setenv('BRIDCT_SCHEDULE', str(table))
    shapes = ([(64, 64)] * 5 + [(128, 256)] * 3 + [(16, 16)] * 4 + [(64, 1024)] * 2 + [(512, 64)] * 2 +
              [(32, 64)] * 3 + [(256, 256)] * 2)
    arrays = ragged(shapes, 11)
    x = flat(arrays)
    for (h, w), route in (((64, 64), capi.ROUTE_MFMA_F16), ((64, 1024), capi.ROUTE_MFMA_BF16),
                          ((32, 64), capi.ROUTE_BUTTERFLY), ((256, 256), capi.ROUTE_BUTTERFLY)):
        p = capi.Plan(h, w)
        assert p.route == route, (h, w)
        p.close()
    b = capi.Batch(shapes)
    try:
        assert b.strategy == capi.BATCH_GROUPED
        for mode in (0, 1, 2, FTI):
            print('dense routes in a grouped batch, mode %d' % mode, flush=True)
            if mode == FTI:
                ref = flat(plan_results(shapes, plan_results(shapes, arrays, 0), 1))
            else:
                ref = flat(plan_results(shapes, arrays, mode))
            assert np.array_equal(run_device(b, mode, x), ref)
            assert np.array_equal(run_device(b, mode, x, graph=True), ref)
        assert b.set_strategy(capi.BATCH_PERSISTENT) == 0
        monkeypatch.setenv('BRIDCT_SCHEDULE', 'none')   # butterfly plans for the comparison
        for mode in (0, 1):
            y = run_device(b, mode, x)
            assert agree(y, flat(plan_results(shapes, arrays, mode)), capi.BATCH_PERSISTENT)
    finally:
        b.close()

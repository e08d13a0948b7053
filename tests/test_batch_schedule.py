"""The host-side schedule of ragged batches (bdx_batch_schedule in bridct_amd/csrc/batch.c),
checked on the CPU: every array and every row/column unit is queued exactly once, every
dependency points backwards in the queue, and a simulated persistent grid of any size
drains the queue.  The reference has no batch object (callers loop over plans,
python/bridct_numpy.c:96-112); these are properties of the schedule itself."""
import ctypes
import heapq
import random

import numpy as np
import pytest

from bridct_amd import capi

THREADS = 256
SMALL, ROW, COL, ROW2 = 0, 1, 2, 3


class Run(ctypes.Structure):
    _fields_ = [('h', ctypes.c_uint), ('w', ctypes.c_uint), ('count', ctypes.c_uint), ('pad', ctypes.c_uint),
                ('off', ctypes.c_ulonglong)]


class Item(ctypes.Structure):
    _fields_ = [('run', ctypes.c_uint), ('idx', ctypes.c_uint), ('unit', ctypes.c_uint),
                ('kind_ctr', ctypes.c_uint)]


class Schedule(ctypes.Structure):
    _fields_ = [('count', ctypes.c_size_t), ('elements', ctypes.c_size_t),
                ('offsets', ctypes.POINTER(ctypes.c_size_t)), ('runs', ctypes.POINTER(Run)),
                ('nruns', ctypes.c_uint), ('ntwo', ctypes.c_uint),
                ('items', ctypes.POINTER(Item) * 2), ('nitems', ctypes.c_uint * 2), ('nctr', ctypes.c_uint),
                ('tiny', ctypes.c_int)]


def team(n):
    return 1 if n <= 64 else 16 if n >= 1024 else 8 if n >= 256 else 4


BANDS = 1   # BDX_ITEM_BANDS: bands / strips per item for lines of 512 and more


def row_unit(w):
    return (THREADS // team(w) if w >= 128 else THREADS) * (BANDS if w >= 512 else 1)


def col_unit(h):
    return (THREADS // team(h) if h >= 128 else THREADS) * (BANDS if h >= 512 else 1)


def is_small(h, w):
    return h <= 128 and w <= 128


def is_tiny(shapes):
    return all(h <= 32 and w <= 32 for h, w in shapes)


def per_unit(h, w, tiny=False):
    # tiny batches (every array <= 32 x 32): about 2048 pixels per item; otherwise as many
    # whole arrays as fit the persistent kernel's LDS area
    return 2048 // (h * w) if tiny else 20400 // (h * (w + 1))


def ceil(a, b):
    return -(-a // b)


def schedule(shapes, lag):
    lib = capi.load()
    fn = lib.bdx_batch_schedule
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.POINTER(Schedule)]
    lib.bdx_schedule_free.argtypes = [ctypes.POINTER(Schedule)]
    hs = np.ascontiguousarray([h for h, _ in shapes], np.int32)
    ws = np.ascontiguousarray([w for _, w in shapes], np.int32)
    s = Schedule()
    rc = fn(len(shapes), hs.ctypes.data, ws.ctypes.data, lag, ctypes.byref(s))
    if rc:
        return rc, None
    out = {
        'elements': s.elements,
        'offsets': [s.offsets[i] for i in range(s.count)],
        'runs': [(r.h, r.w, r.count, r.off) for r in (s.runs[i] for i in range(s.nruns))],
        'ntwo': s.ntwo, 'nctr': s.nctr, 'tiny': s.tiny,
        'items': [[(it.run, it.idx, it.unit, it.kind_ctr >> 28, it.kind_ctr & 0x0FFFFFFF)
                   for it in (s.items[q][i] for i in range(s.nitems[q]))] for q in (0, 1)],
    }
    lib.bdx_schedule_free(ctypes.byref(s))
    return 0, out


def check(shapes, lag):
    rc, s = schedule(shapes, lag)
    assert rc == 0
    # layout: back to back in the given order; runs group consecutive equal shapes
    offs = np.concatenate([[0], np.cumsum([h * w for h, w in shapes])])
    assert s['offsets'] == offs[:-1].tolist() and s['elements'] == offs[-1]
    arrays = []   # (run index, index in run) of every array, in batch order
    for r, (h, w, count, off) in enumerate(s['runs']):
        assert off == s['offsets'][len(arrays)]
        for i in range(count):
            assert (h, w) == shapes[len(arrays)]
            arrays.append((r, i))
        if r:
            assert (s['runs'][r - 1][0], s['runs'][r - 1][1]) != (h, w)
    assert len(arrays) == len(shapes)
    two = [a for a, (h, w) in zip(arrays, shapes) if not is_small(h, w)]
    assert s['ntwo'] == len(two) and s['nctr'] == 2 + 2 * len(two)
    assert s['tiny'] == is_tiny(shapes)
    for fused in (0, 1):
        items = s['items'][fused]
        small_seen = {}
        units = {}      # (run, idx, kind) -> list of (position, unit)
        ctr_of = {}
        for pos, (run, idx, unit, kind, ctr) in enumerate(items):
            h, w = s['runs'][run][:2]
            if kind == SMALL:
                assert is_small(h, w) and 1 <= unit <= per_unit(h, w, s['tiny'])
                for a in range(idx, idx + unit):
                    assert a < s['runs'][run][2]
                    small_seen[(run, a)] = small_seen.get((run, a), 0) + 1
            else:
                assert not is_small(h, w) and kind in ((ROW, COL, ROW2) if fused else (ROW, COL))
                units.setdefault((run, idx, kind), []).append((pos, unit))
                assert ctr_of.setdefault((run, idx), ctr) == ctr
        assert set(small_seen) == {a for a, (h, w) in zip(arrays, shapes) if is_small(h, w)}
        assert all(v == 1 for v in small_seen.values())
        assert sorted(ctr_of.values()) == list(range(len(two)))
        for (run, idx) in two:
            h, w = s['runs'][run][:2]
            kinds = (ROW, COL, ROW2) if fused else (ROW, COL)
            nunits = {ROW: ceil(h, row_unit(w)), COL: ceil(w, col_unit(h)), ROW2: ceil(h, row_unit(w))}
            for k in kinds:
                assert sorted(u for _, u in units[(run, idx, k)]) == list(range(nunits[k]))
            # dependencies point backwards: all rows before any column, columns before ROW2
            assert max(p for p, _ in units[(run, idx, ROW)]) < min(p for p, _ in units[(run, idx, COL)])
            if fused:
                assert max(p for p, _ in units[(run, idx, COL)]) < min(p for p, _ in units[(run, idx, ROW2)])
    return s


def simulate(s, fused, blocks, seed, prefetch=True):
    """Persistent grid of `blocks` workers pulling items in queue order; an item that waits
    holds its worker until its dependencies are done.  With prefetch, a worker claims its
    next item when it starts the current one and holds it unstarted (a claiming order the
    schedule must also survive).  Returns the makespan (or raises on a deadlock)."""
    items = s['items'][fused]
    need = {}
    for run, idx, unit, kind, ctr in items:
        if kind in (ROW, COL):
            need[(ctr, kind)] = need.get((ctr, kind), 0) + 1
    done = {}
    rng = random.Random(seed)
    t, head = 0.0, 0
    waiting = []      # (worker, item) started but blocked on a dependency
    running = []      # heap of (finish time, worker, item)
    held = {}         # worker -> claimed, not started item
    free = list(range(blocks))

    def ready(it):
        kind, ctr = it[3], it[4]
        if kind == COL:
            return done.get((ctr, ROW), 0) == need[(ctr, ROW)]
        if kind == ROW2:
            return done.get((ctr, COL), 0) == need[(ctr, COL)]
        return True

    while head < len(items) or waiting or running or held:
        idle = []
        for wkr in free:
            if wkr in held:
                cur = held.pop(wkr)
            elif head < len(items):
                cur = items[head]
                head += 1
            else:
                idle.append(wkr)
                continue
            waiting.append((wkr, cur))
            if prefetch and head < len(items):
                held[wkr] = items[head]
                head += 1
        free = idle
        still = []
        for wkr, it in waiting:
            if ready(it):
                heapq.heappush(running, (t + rng.uniform(0.5, 1.5), wkr, it))
            else:
                still.append((wkr, it))
        waiting = still
        if not running:
            raise AssertionError('deadlock at queue position %d' % head)
        t, wkr, it = heapq.heappop(running)
        done[(it[4], it[3])] = done.get((it[4], it[3]), 0) + 1
        free.append(wkr)
    return t


def test_uniform_batches():
    for shape in [(8, 8), (16, 64), (128, 128), (1024, 1024), (512, 128), (128, 1024), (32, 256), (1024, 8)]:
        for n in (1, 3, 64):
            check([shape] * n, 256)


def test_ragged_batches_and_lags():
    rng = np.random.default_rng(7)
    lengths = [8, 16, 32, 64, 128, 256, 512, 1024]
    shapes = [(int(rng.choice(lengths)), int(rng.choice(lengths))) for _ in range(300)]
    for lag in (0, 1, 17, 256, 100000):
        s = check(shapes, lag)
        for fused in (0, 1):
            for blocks in (1, 3, 512):
                simulate(s, fused, blocks, seed=blocks + lag)
                simulate(s, fused, blocks, seed=blocks + lag, prefetch=False)


def test_all_shapes_once():
    lengths = [8, 16, 32, 64, 128, 256, 512, 1024]
    shapes = [(h, w) for h in lengths for w in lengths]
    s = check(shapes, 256)
    for fused in (0, 1):
        simulate(s, fused, 2, seed=1)


def test_many_small_blocks_split_evenly():
    # all arrays <= 32 x 32: the tiny kernel's smaller LDS area sets the packing
    for shapes in ([(8, 8)] * 100000, [(8, 8)] * 99999 + [(64, 8)]):
        s = check(shapes, 256)
        per = per_unit(8, 8, s['tiny'])
        counts = [u for r, _, u, k, _ in s['items'][0] if k == SMALL and s['runs'][r][:2] == (8, 8)]
        assert sum(counts) == 99999 + (shapes[-1] == (8, 8)) and len(counts) == ceil(sum(counts), per)
        assert max(counts) - min(counts) <= 1
    assert check([(32, 32), (8, 16)], 0)['tiny'] and not check([(32, 32), (8, 64)], 0)['tiny']


@pytest.mark.parametrize('shapes', [[], [(8, 12)], [(2048, 8)], [(8, 8), (4, 8)]])
def test_rejects_invalid(shapes):
    rc, _ = schedule(shapes, 256)
    assert rc == -1

"""The host-side schedule of ragged batches (bdx_batch_schedule in bridct_amd/csrc/batch.c),
checked on the CPU: every array and every row/column item is listed exactly once, grouped
into units (one two-pass array with its ROW, COL and ROW2 items, or one SMALL item), and
a simulation of the persistent kernel's per-XCD queues (tickets cut into slots: the rows of
the unit a queue takes j-th, the columns of its previous unit, the ROW2 items of the one
before) drains every unit exactly once, without deadlock, for any number of queues and
workers.  The reference has no batch object (callers loop over plans,
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


class Unit(ctypes.Structure):
    _fields_ = [('item0', ctypes.c_uint * 2), ('nrow', ctypes.c_ushort), ('ncol', ctypes.c_ushort),
                ('nrow2', ctypes.c_ushort), ('pad', ctypes.c_ushort)]


class Schedule(ctypes.Structure):
    _fields_ = [('count', ctypes.c_size_t), ('elements', ctypes.c_size_t),
                ('offsets', ctypes.POINTER(ctypes.c_size_t)), ('runs', ctypes.POINTER(Run)),
                ('nruns', ctypes.c_uint), ('ntwo', ctypes.c_uint),
                ('items', ctypes.POINTER(Item) * 2), ('nitems', ctypes.c_uint * 2),
                ('units', ctypes.POINTER(Unit)), ('nunits', ctypes.c_uint), ('qbase', ctypes.c_uint),
                ('nctr', ctypes.c_uint), ('tiny', ctypes.c_int)]


XQ = 16   # BDX_XQ: queues (one per XCD id)


def qwords(nunits, lag):
    return XQ + 2 + 2 * XQ * (nunits + 2 * lag + 1)   # BDX_QWORDS: every unit, then the lag's slots


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


def schedule(shapes):
    lib = capi.load()
    fn = lib.bdx_batch_schedule
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Schedule)]
    lib.bdx_schedule_free.argtypes = [ctypes.POINTER(Schedule)]
    hs = np.ascontiguousarray([h for h, _ in shapes], np.int32)
    ws = np.ascontiguousarray([w for _, w in shapes], np.int32)
    s = Schedule()
    rc = fn(len(shapes), hs.ctypes.data, ws.ctypes.data, ctypes.byref(s))
    if rc:
        return rc, None
    out = {
        'elements': s.elements,
        'offsets': [s.offsets[i] for i in range(s.count)],
        'runs': [(r.h, r.w, r.count, r.off) for r in (s.runs[i] for i in range(s.nruns))],
        'ntwo': s.ntwo, 'nctr': s.nctr, 'qbase': s.qbase, 'tiny': s.tiny,
        'items': [[(it.run, it.idx, it.unit, it.kind_ctr >> 28, it.kind_ctr & 0x0FFFFFFF)
                   for it in (s.items[q][i] for i in range(s.nitems[q]))] for q in (0, 1)],
        'units': [((u.item0[0], u.item0[1]), u.nrow, u.ncol, u.nrow2) for u in (s.units[i] for i in range(s.nunits))],
    }
    lib.bdx_schedule_free(ctypes.byref(s))
    return 0, out


def check(shapes):
    rc, s = schedule(shapes)
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
    assert s['ntwo'] == len(two) and s['tiny'] == is_tiny(shapes)
    assert s['qbase'] % 2 == 0 and s['qbase'] >= 2 + 2 * len(two)
    # the queue words hold every unit's slot plus 2 lag + 1 end slots per queue (BDX_LAG)
    lag = ((s['nctr'] - s['qbase'] - XQ - 2) // (2 * XQ) - len(s['units']) - 1) // 2
    assert lag >= 1 and s['nctr'] == s['qbase'] + qwords(len(s['units']), lag)
    s['lag'] = lag
    for fused in (0, 1):
        items = s['items'][fused]
        small_seen = {}
        units = {}      # (run, idx, kind) -> list of unit indices
        ctr_of = {}
        for run, idx, unit, kind, ctr in items:
            h, w = s['runs'][run][:2]
            if kind == SMALL:
                assert is_small(h, w) and 1 <= unit <= per_unit(h, w, s['tiny'])
                for a in range(idx, idx + unit):
                    assert a < s['runs'][run][2]
                    small_seen[(run, a)] = small_seen.get((run, a), 0) + 1
            else:
                assert not is_small(h, w) and kind in ((ROW, COL, ROW2) if fused else (ROW, COL))
                units.setdefault((run, idx, kind), []).append(unit)
                assert ctr_of.setdefault((run, idx), ctr) == ctr
        assert set(small_seen) == {a for a, (h, w) in zip(arrays, shapes) if is_small(h, w)}
        assert all(v == 1 for v in small_seen.values())
        assert sorted(ctr_of.values()) == list(range(len(two)))
        for (run, idx) in two:
            h, w = s['runs'][run][:2]
            nunits = {ROW: ceil(h, row_unit(w)), COL: ceil(w, col_unit(h)), ROW2: ceil(h, row_unit(w))}
            for k in ((ROW, COL, ROW2) if fused else (ROW, COL)):
                assert sorted(units[(run, idx, k)]) == list(range(nunits[k]))
        # units tile the item list: ROW (or the SMALL item), then COL, then ROW2, consecutive
        pos, sizes = 0, []
        for k, (item0, nrow, ncol, nrow2) in enumerate(s['units']):
            assert item0[fused] == pos
            kinds = [items[pos + i][3] for i in range(nrow + ncol + (nrow2 if fused else 0))]
            if kinds[0] == SMALL:
                assert (nrow, ncol, nrow2) == (1, 0, 0)
            else:
                assert kinds == [ROW] * nrow + [COL] * ncol + ([ROW2] * nrow2 if fused else [])
                assert nrow2 == nrow and k < len(two)
                first = items[pos]
                assert all(items[pos + i][:2] == first[:2] for i in range(len(kinds)))
                sizes.append(s['runs'][first[0]][0] * s['runs'][first[0]][1])
            pos += len(kinds)
        assert pos == len(items)
        assert sizes == sorted(sizes, reverse=True)   # two-pass arrays by decreasing size first
    return s


def simulate(s, fused, queues, workers, seed, lag=None):
    """The persistent kernel's protocol: `queues` XCD queues of `workers` workers each.  A
    worker draws the next ticket of its queue; the tickets of a queue are cut into slots,
    slot j = the ROW items of the unit u_j the queue takes j-th (from a global counter, when
    its first ticket is drawn), then the COL items of u_{j-lag}, then (fused) the ROW2 items
    of u_{j-2 lag} (lag = BDX_LAG).  A COL item waits until all ROW items of its unit are done,
    a ROW2 item until all COL items are.  Returns the makespan, or raises on a deadlock; every
    item must run exactly once."""
    lag = s['lag'] if lag is None else lag
    units = s['units']
    items = s['items'][fused]
    rng = random.Random(seed)
    nxt = [0]                      # global unit counter
    slots = [[] for _ in range(queues)]   # per queue: [(unit or None, start, end)]
    ticket = [0] * queues
    END = None

    def parts(u):
        if u is END:
            return 0, 0, 0
        _, nrow, ncol, nrow2 = units[u]
        return nrow, ncol, (nrow2 if fused else 0)

    def resolve(q, v):
        """item index of ticket v of queue q, or None when the queue is done"""
        sl = slots[q]

        def unit_at(j):
            return sl[j][0] if j >= 0 else END

        def terminal(j):   # slot j and every later one are empty
            return all(unit_at(j - i) is END for i in range((2 * lag if fused else lag) + 1))

        while True:
            j = len(sl)
            start = sl[-1][2] if sl else 0
            if sl and terminal(j - 1):
                return None
            if sl and v < sl[-1][2]:
                break
            if v != start:   # tickets are drawn in order, so the first ticket of a slot comes first
                raise AssertionError('ticket %d drawn before slot start %d' % (v, start))
            u = nxt[0] if nxt[0] < len(units) else END
            nxt[0] += 1
            u1 = sl[j - lag][0] if j >= lag else END
            u2 = sl[j - 2 * lag][0] if j >= 2 * lag else END
            length = parts(u)[0] + parts(u1)[1] + parts(u2)[2]
            sl.append((u, start, start + length))
        for j, (u, start, end) in enumerate(sl):
            if start <= v < end:
                o = v - start
                u1 = sl[j - lag][0] if j >= lag else END
                u2 = sl[j - 2 * lag][0] if j >= 2 * lag else END
                r0, c1 = parts(u)[0], parts(u1)[1]
                if o < r0:
                    return units[u][0][fused] + o
                o -= r0
                if o < c1:
                    return units[u1][0][fused] + units[u1][1] + o
                o -= c1
                return units[u2][0][fused] + units[u2][1] + units[u2][2] + o
        raise AssertionError('ticket %d of queue %d in no slot' % (v, q))

    need = {}
    for run, idx, unit, kind, ctr in items:
        if kind in (ROW, COL):
            need[(ctr, kind)] = need.get((ctr, kind), 0) + 1
    done, ran = {}, [0] * len(items)

    def ready(i):
        kind, ctr = items[i][3], items[i][4]
        if kind == COL:
            return done.get((ctr, ROW), 0) == need[(ctr, ROW)]
        if kind == ROW2:
            return done.get((ctr, COL), 0) == need[(ctr, COL)]
        return True

    t = 0.0
    free = [(q, k) for q in range(queues) for k in range(workers)]
    rng.shuffle(free)
    waiting, running = [], []
    exited = set()
    while True:
        for wkr in free:
            q = wkr[0]
            v = ticket[q]
            ticket[q] += 1
            i = resolve(q, v)
            if i is None:
                exited.add(wkr)
            else:
                waiting.append((wkr, i))
        free = []
        still = []
        for wkr, i in waiting:
            if ready(i):
                heapq.heappush(running, (t + rng.uniform(0.5, 1.5), wkr, i))
            else:
                still.append((wkr, i))
        waiting = still
        if not running:
            if waiting:
                raise AssertionError('deadlock: %d workers waiting' % len(waiting))
            break
        t, wkr, i = heapq.heappop(running)
        kind, ctr = items[i][3], items[i][4]
        done[(ctr, kind)] = done.get((ctr, kind), 0) + 1
        ran[i] += 1
        free.append(wkr)
    assert ran == [1] * len(items)
    assert len(exited) == queues * workers
    return t


def test_uniform_batches():
    for shape in [(8, 8), (16, 64), (128, 128), (1024, 1024), (512, 128), (128, 1024), (32, 256), (1024, 8)]:
        for n in (1, 3, 64):
            s = check([shape] * n)
            for fused in (0, 1):
                simulate(s, fused, 8, 2, seed=n)


def test_ragged_batches_any_grid():
    rng = np.random.default_rng(7)
    lengths = [8, 16, 32, 64, 128, 256, 512, 1024]
    shapes = [(int(rng.choice(lengths)), int(rng.choice(lengths))) for _ in range(300)]
    s = check(shapes)
    for fused in (0, 1):
        for queues, workers in ((1, 1), (1, 3), (8, 1), (8, 64), (3, 5), (16, 2)):
            for lag in (1, 2, 3):
                simulate(s, fused, queues, workers, seed=queues * 100 + workers + lag, lag=lag)


def test_all_shapes_once():
    lengths = [8, 16, 32, 64, 128, 256, 512, 1024]
    shapes = [(h, w) for h in lengths for w in lengths]
    s = check(shapes)
    for fused in (0, 1):
        simulate(s, fused, 8, 2, seed=1)
        simulate(s, fused, 1, 1, seed=2)


def test_many_small_blocks_split_evenly():
    # all arrays <= 32 x 32: the tiny kernel's smaller LDS area sets the packing
    for shapes in ([(8, 8)] * 100000, [(8, 8)] * 99999 + [(64, 8)]):
        s = check(shapes)
        per = per_unit(8, 8, s['tiny'])
        counts = [u for r, _, u, k, _ in s['items'][0] if k == SMALL and s['runs'][r][:2] == (8, 8)]
        assert sum(counts) == 99999 + (shapes[-1] == (8, 8)) and len(counts) == ceil(sum(counts), per)
        assert max(counts) - min(counts) <= 1
    assert check([(32, 32), (8, 16)])['tiny'] and not check([(32, 32), (8, 64)])['tiny']


@pytest.mark.parametrize('shapes', [[], [(8, 12)], [(2048, 8)], [(8, 8), (4, 8)]])
def test_rejects_invalid(shapes):
    rc, _ = schedule(shapes)
    assert rc == -1

"""Autotuner of the per-shape schedule table (route and whole-array packing).

For every supported shape it times the candidates on the GPU -- the FP32 butterfly route
with several whole-array packings (arrays per workgroup, for shapes whose array fits one
workgroup's LDS), for those shapes also the row and column pass kernels instead (pack -1,
where those kernels take the shape), and, for shapes with both axes >= 64, the
split-MFMA dense routes with
FP16 and BF16 operands -- on a synthetic batch of uniform[-1, 1) arrays, forward and
inverse, with HIP events.  Every candidate is gated on error first: its max relative error
against the FP64 transform (on two arrays of the batch) must stay within twice the
butterfly route's error on the same arrays (the butterfly route itself sits within twice
the reference CPU's error, which the parity tests check).  The fastest candidate that
passes is written to the table, one line per shape (the two-pass kernels only where they
also beat automatic whole-array packing on a batch of 64 arrays by 25%, since a plan
serves any batch and the two passes are two launches):

    <height> <width> <butterfly|mfma_f16|mfma_bf16> <pack>

followed by a comment with every candidate's rate and error.  Plans read the table when
they are created (bridct_plan_create; see include/bridct_gpu.h).  This replaces the
reference's compiled-in policy tables (src/rect_policy.inc, small_policy.json).

Usage:  python3 -m bridct_amd.autotune --out bridct_amd/schedule_gfx950.txt
"""
import argparse
import ctypes
import datetime
import os
import sys
import time

import numpy as np

from . import capi

LENGTHS = (8, 16, 32, 64, 128, 256, 512, 1024)
SMALL_LDS_FLOATS = 36864      # whole-array route: h (w + 1) floats of LDS at most
PACKS = (0, 1, 2, 4, 8, 16, 32, 64, 128, 256, 512)
PACK_TWO_PASS = -1            # the row and column pass kernels instead of the whole-array one
SMALL_BATCH = 64              # arrays of the second comparison a two-pass choice must also win
# ... by this factor: the two passes are two launches per direction, a cost that isolated
# timing does not charge (a batch-64 step of four shapes on four streams ran 1.2% slower
# with 32 x 256 on two passes, although that shape alone ran 18% faster)
TWO_PASS_MARGIN = 1.25
GATE_ARRAYS = 4               # arrays whose error gates a route
ROUTE_NAMES = {capi.ROUTE_BUTTERFLY: 'butterfly', capi.ROUTE_MFMA_F16: 'mfma_f16',
               capi.ROUTE_MFMA_BF16: 'mfma_bf16'}


class Events:
    """HIP event timing on a stream."""

    def __init__(self):
        try:
            self.lib = ctypes.CDLL('libamdhip64.so')
        except OSError:
            self.lib = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so')
        self.lib.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.lib.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                 ctypes.c_void_p]
        self.lib.hipEventDestroy.argtypes = [ctypes.c_void_p]
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.lib.hipEventCreate(ctypes.byref(self.a)) == 0
        assert self.lib.hipEventCreate(ctypes.byref(self.b)) == 0

    def time_ms(self, stream, fn, reps):
        fn()
        assert self.lib.hipEventRecord(self.a, stream) == 0
        for _ in range(reps):
            fn()
        assert self.lib.hipEventRecord(self.b, stream) == 0
        assert self.lib.hipEventSynchronize(self.b) == 0
        ms = ctypes.c_float()
        assert self.lib.hipEventElapsedTime(ctypes.byref(ms), self.a, self.b) == 0
        return ms.value / reps


def dct_fp64(n):
    """Orthonormal DCT-II matrix in FP64 (the transform's definition, reference
    docs/API.md:7-13)."""
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    return np.sqrt(np.where(k == 0, 1.0, 2.0) / n) * np.cos(np.pi * ((k * (2 * j + 1)) % (4 * n)) / (2.0 * n))


def two_pass_ok(h, w):
    """Whether the row and column pass kernels take an h x w shape in whole row bands and
    column strips (BDX_TWO_PASS_OK: bands of 256 / team rows, 256 for lines up to 64;
    strips of 64 columns, 32 at height 1024, 256 for lines up to 64)."""
    team = 1 if w <= 64 else 16 if w >= 1024 else 8 if w >= 256 else 4
    rows = 256 if w <= 64 else 256 // team
    cols = 32 if h >= 1024 else 64 if h >= 128 else 256
    return h % rows == 0 and w % cols == 0


def candidates(h, w, batch):
    out = []
    if h * (w + 1) <= SMALL_LDS_FLOATS:
        per_max = SMALL_LDS_FLOATS // (h * (w + 1))
        out += [(capi.ROUTE_BUTTERFLY, p) for p in PACKS if p <= min(per_max, batch)]
        if two_pass_ok(h, w):
            out.append((capi.ROUTE_BUTTERFLY, PACK_TWO_PASS))
    else:
        out.append((capi.ROUTE_BUTTERFLY, 0))
    if h >= 64 and w >= 64:
        out += [(capi.ROUTE_MFMA_F16, 0), (capi.ROUTE_MFMA_BF16, 0)]
    return out


def tune_shape(ev, h, w, pixels):
    batch = max(GATE_ARRAYS, min(65535, pixels // (h * w)))
    rng = np.random.default_rng(h * 1024 + w)
    x = rng.uniform(-1, 1, (batch, h, w)).astype(np.float32)
    # the arrays the error gate reads carry the ragged workload's per-array scales,
    # 10^U[-3, 3], so a route must hold its accuracy across that range
    for i in range(GATE_ARRAYS):
        x[i] *= np.float32(10 ** rng.uniform(-3, 3))
    d_in, d_out = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    qh, qw = dct_fp64(h), dct_fp64(w)
    ref = [(qh @ x[i].astype(np.float64) @ qw.T, qh.T @ x[i].astype(np.float64) @ qw) for i in range(GATE_ARRAYS)]
    plan = capi.Plan(h, w)
    results = []
    try:
        for route, pack in candidates(h, w, batch):
            if plan.set_route(route) or plan.set_pack(pack):
                continue
            ms = {}
            err = 0.0
            for mode in (capi.FORWARD, capi.INVERSE):
                assert plan.apply_device(mode, d_in.ptr, d_out.ptr, batch) == 0, (h, w, route, pack)
                ms[mode] = ev.time_ms(plan.stream, lambda: plan.apply_device(mode, d_in.ptr, d_out.ptr, batch), 3)
                assert plan.synchronize() == 0
                y = np.empty_like(x[:GATE_ARRAYS])
                d_out.download(y)
                for i in range(GATE_ARRAYS):
                    r = ref[i][mode]
                    err = max(err, float(np.abs(y[i] - r).max() / np.abs(r).max()))
            gpx = 2 * x.size / ((ms[0] + ms[1]) * 1e-3) / 1e9
            results.append({'route': route, 'pack': pack, 'gpx': gpx, 'err': err})
    finally:
        plan.close()
    base = max(r['err'] for r in results if r['route'] == capi.ROUTE_BUTTERFLY and r['pack'] == 0)
    ok = [r for r in results if r['err'] <= 2 * base]
    best = max(ok, key=lambda r: r['gpx'])
    small = None
    if best['pack'] == PACK_TWO_PASS and batch > SMALL_BATCH:
        # the two-pass kernels spread each array over many workgroups, which pays on large
        # batches; a plan serves any batch, so the choice must also win on a small one
        small = small_batch_rates(ev, h, w, d_in, d_out)
        if small[PACK_TWO_PASS] < TWO_PASS_MARGIN * small[0]:
            best = max((r for r in ok if r['pack'] != PACK_TWO_PASS), key=lambda r: r['gpx'])
    return batch, best, results, small


def small_batch_rates(ev, h, w, d_in, d_out):
    """Forward+inverse Gpixel/s of automatic whole-array packing and of the two-pass
    kernels on SMALL_BATCH arrays (the first arrays of the tuning batch)."""
    plan = capi.Plan(h, w)
    rates = {}
    try:
        for pack in (0, PACK_TWO_PASS):
            assert plan.set_pack(pack) == 0
            ms = 0.0
            for mode in (capi.FORWARD, capi.INVERSE):
                assert plan.apply_device(mode, d_in.ptr, d_out.ptr, SMALL_BATCH) == 0
                ms += min(ev.time_ms(plan.stream, lambda: plan.apply_device(mode, d_in.ptr, d_out.ptr, SMALL_BATCH),
                                     20) for _ in range(3))
            rates[pack] = 2 * SMALL_BATCH * h * w / (ms * 1e-3) / 1e9
    finally:
        plan.close()
    return rates


def main(argv=None):
    ap = argparse.ArgumentParser(description='Autotune the per-shape schedule table.')
    ap.add_argument('--out', default=os.path.join(os.path.dirname(os.path.abspath(__file__)), 'schedule_gfx950.txt'))
    ap.add_argument('--pixels', type=int, default=1 << 24, help='pixels per timed batch')
    ap.add_argument('--shapes', default='', help='comma list HxW (default: all 64)')
    args = ap.parse_args(argv)
    os.environ['BRIDCT_SCHEDULE'] = 'none'   # time the candidates, not a previous table
    lib = capi.load()
    if lib.bridct_gpu_available() != 1:
        print('no usable gfx950 GPU', file=sys.stderr)
        return 1
    shapes = ([tuple(int(v) for v in s.split('x')) for s in args.shapes.split(',')] if args.shapes
              else [(h, w) for h in LENGTHS for w in LENGTHS])
    ev = Events()
    lines = ['# BRiDCT schedule table for gfx950 (MI355X): <height> <width> <route> <pack>',
             '# written by: python3 -m bridct_amd.autotune --pixels %d  (%s UTC)' % (
                 args.pixels, datetime.datetime.utcnow().strftime('%Y-%m-%d %H:%M')),
             '# gate: max relative error vs FP64 <= 2 x the butterfly route (pack 0) on the same %d arrays,' % GATE_ARRAYS,
             '# scaled by 10^U[-3, 3] each;',
             '# then the highest forward+inverse Gpixel/s; the two-pass kernels (pack -1) only where they',
             '# also beat automatic whole-array packing on %d arrays by a factor %.2f.' % (SMALL_BATCH, TWO_PASS_MARGIN),
             '# Comment: every candidate as route/pack:Gpx/s:error.']
    t0 = time.time()
    for h, w in shapes:
        batch, best, results, small = tune_shape(ev, h, w, args.pixels)
        cand = ' '.join('%s/%d:%.1f:%.1e' % (ROUTE_NAMES[r['route']], r['pack'], r['gpx'], r['err']) for r in results)
        if small:
            cand += '; on %d arrays: butterfly/0:%.1f butterfly/-1:%.1f' % (SMALL_BATCH, small[0], small[PACK_TWO_PASS])
        lines.append('%d %d %s %d  # batch %d; %s' % (h, w, ROUTE_NAMES[best['route']], best['pack'], batch, cand))
        print(lines[-1], flush=True)
    with open(args.out, 'w') as f:
        f.write('\n'.join(lines) + '\n')
    print('wrote %s (%d shapes, %.1f s)' % (args.out, len(shapes), time.time() - t0), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())

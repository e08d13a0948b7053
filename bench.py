#!/usr/bin/env python3
"""Benchmark of the MI355X BRiDCT 2D plan path on one GPU.

Workload (the large and rectangular plain-FP32 configuration): batches of 64 float32
arrays at 1024x1024, 512x128, 128x1024 and 32x256, values a smooth field (sum of 16
Gaussian blobs, amplitude renormalized into [-1, 1]) plus N(0, 0.01^2) noise.  One step
is the orthonormal forward DCT-II and then the inverse DCT-III of every batch, with inputs
already resident in HBM.  Prints one JSON line: whole-job Gpixel/s of forward + inverse,
per-shape forward / inverse rates, the HBM roofline of the dominant launch (forward over
the 64 x 1024^2 batch, timed with HIP events on its stream) and the reference CPU path
timed on this host's cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from bridct_amd import capi  # noqa: E402

HBM_PEAK_GBS = 8000.0
SHAPES = [(1024, 1024), (512, 128), (128, 1024), (32, 256)]
BATCH = 64


class Hip:
    """The few HIP runtime calls the benchmark needs (events, device sync)."""

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

    def event(self):
        e = ctypes.c_void_p()
        assert self.lib.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.lib.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.lib.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.lib.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0


def smooth_field(h, w, batch, rng):
    """Sum of 16 Gaussian blobs per array (separable), max |field| = 1, + N(0, 0.01^2)."""
    out = np.empty((batch, h, w), np.float32)
    r = np.arange(h, dtype=np.float64)
    c = np.arange(w, dtype=np.float64)
    m = min(h, w)
    for b in range(batch):
        f = np.zeros((h, w))
        for _ in range(16):
            amp = rng.uniform(-1, 1)
            cr, cc = rng.uniform(0, h), rng.uniform(0, w)
            sig = rng.uniform(m / 32, m / 4)
            f += amp * np.outer(np.exp(-(r - cr) ** 2 / (2 * sig * sig)),
                                np.exp(-(c - cc) ** 2 / (2 * sig * sig)))
        f /= max(np.abs(f).max(), 1e-30)
        out[b] = f + rng.normal(0.0, 0.01, (h, w))
    return out


def cpu_baseline(inputs):
    """Reference CPU library (oracle/_ref, built from the reference sources) timed on this
    host: a bounded sample of the same workload, forward then inverse, all host cores."""
    path = os.path.join(ROOT, 'oracle', '_ref', 'libref_bench.so')
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.ref_bench_run.restype = ctypes.c_int
    lib.ref_bench_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_double)]
    per_shape = 8
    hs, ws, offs, chunks = [], [], [], []
    off = 0
    for (h, w), x in inputs.items():
        for i in range(per_shape):
            hs.append(h)
            ws.append(w)
            offs.append(off)
            chunks.append(x[i].ravel())
            off += h * w
    data = np.concatenate(chunks).astype(np.float32)
    out = np.empty_like(data)
    spec = np.empty_like(data)
    threads = min(os.cpu_count() or 1, 16)
    h_a = np.array(hs, np.int32)
    w_a = np.array(ws, np.int32)
    o_a = np.array(offs, np.uint64)
    ns = ctypes.c_double()
    rc = lib.ref_bench_run(len(hs), h_a.ctypes.data, w_a.ctypes.data, o_a.ctypes.data,
                           data.ctypes.data, out.ctypes.data, spec.ctypes.data, 3, threads, 5,
                           200.0, ctypes.byref(ns))
    if rc:
        return None
    gpx = data.size / ns.value
    cpu = 'unknown'
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                cpu = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return {'value': round(gpx, 4), 'unit': 'Gpixel/s', 'cores': threads, 'kind': 'reference',
            'sample': '%d arrays of each shape (%d Mpx), forward then inverse, one plan per '
                      'thread, median of 5 blocks; host CPU %s' % (per_shape, data.size // 10 ** 6, cpu)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: the transform has no multi-GPU path in this '
              'build (arrays shard across GPUs with no exchange); run with --gpus 1.',
              file=sys.stderr)
        return 2
    lib = capi.load()
    if lib.bridct_gpu_available() != 1:
        print('no usable gfx950 GPU', file=sys.stderr)
        return 1
    hip = Hip()
    rng = np.random.default_rng(20261203)
    runs = []
    inputs = {}
    for h, w in SHAPES:
        x = smooth_field(h, w, BATCH, rng)
        inputs[(h, w)] = x
        plan = capi.Plan(h, w)
        bufs = [capi.DeviceBuffer(x.nbytes) for _ in range(3)]
        bufs[0].upload(x)
        runs.append((h, w, plan, bufs))

    def step():
        for h, w, plan, (a, b, c) in runs:
            assert plan.apply_device(capi.FORWARD, a.ptr, b.ptr, BATCH) == 0
            assert plan.apply_device(capi.INVERSE, b.ptr, c.ptr, BATCH) == 0

    for _ in range(args.warmup):
        step()
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    hip.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    pixels = sum(h * w * BATCH for h, w in SHAPES)
    value = pixels / (ms_per_step * 1e-3) / 1e9

    # per-shape, per-direction launch times with HIP events on each plan's stream
    detail = {}
    reps = 5
    for h, w, plan, (a, b, c) in runs:
        st = plan.stream
        for name, mode, src, dst in (('fwd', capi.FORWARD, a, b), ('inv', capi.INVERSE, b, c)):
            e0, e1 = hip.event(), hip.event()
            hip.record(e0, st)
            for _ in range(reps):
                assert plan.apply_device(mode, src.ptr, dst.ptr, BATCH) == 0
            hip.record(e1, st)
            ms = hip.elapsed_ms(e0, e1) / reps
            gbs = 8.0 * h * w * BATCH / (ms * 1e-3) / 1e9
            detail['%dx%d_%s' % (h, w, name)] = {
                'ms': round(ms, 4), 'gpx_s': round(h * w * BATCH / (ms * 1e-3) / 1e9, 2),
                'hbm_frac': round(gbs / HBM_PEAK_GBS, 4)}
    # sanity: round trip of the 1024^2 batch
    x = inputs[(1024, 1024)]
    back = np.empty_like(x)
    runs[0][3][2].download(back)
    rt_err = float(np.abs(back - x).max() / np.abs(x).max())

    dom = detail['1024x1024_fwd']
    achieved = 8.0 * 1024 * 1024 * BATCH / (dom['ms'] * 1e-3) / 1e9
    line = {
        'metric': 'Gpixel/s fwd+inv 2D DCT & %HBM peak; split-MFMA/batched vs plain FP32',
        'value': round(value, 3), 'unit': 'Gpixel/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(ms_per_step, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'plain FP32 butterfly route: batch 64 of 1024x1024, 512x128, '
                               '128x1024, 32x256 float32 smooth field + N(0,0.01^2), forward '
                               'then inverse', 'batch': BATCH,
                   'shapes': ['%dx%d' % s for s in SHAPES], 'route': 'butterfly'},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': None,
                     'launch': 'forward, 64 x 1024x1024, 8 B/px algorithmic'},
        'per_shape': detail,
        'roundtrip_max_rel_err_1024': rt_err,
    }
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(inputs)
    print(json.dumps(line))
    return 0


if __name__ == '__main__':
    sys.exit(main())

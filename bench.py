#!/usr/bin/env python3
"""Benchmark of the MI355X BRiDCT 2D plan path on one GPU.

Workload (the large and rectangular plain-FP32 configuration): batches of 64 float32
arrays at 1024x1024, 512x128, 128x1024 and 32x256, values a smooth field (sum of 16
Gaussian blobs, amplitude renormalized into [-1, 1]) plus N(0, 0.01^2) noise.  One step
is the orthonormal forward DCT-II and then the inverse DCT-III of every batch, with inputs
already resident in HBM.  Before the W warm-up steps the step runs for about 0.3 s so that the
timed steps see the GPU's settled clock.  Prints one JSON line: whole-job Gpixel/s of forward + inverse,
per-shape forward / inverse rates, the HBM roofline of the dominant kernel (each kernel of
the 64 x 1024^2 transforms timed in sequence, row pass then column pass, with HIP events on
the plan's stream), device copies as the practical bandwidth reference, and the reference
CPU path timed on this host's cores.  Secondary legs (not part of `value`): the dense split-MFMA route vs the
butterflies at 64^2, 256^2, 1024^2 (batch 128), 1,048,576 independent 8x8 blocks in one
persistent launch vs per-block launches, and a ragged batch of 10,000 arrays of random
supported shapes (seed 7) in one persistent launch vs one plan launch per array, and the
fused 8x8 block codec step (u8 image -> quantized int16 coefficients, and back).
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
    """The few HIP runtime calls the benchmark needs (events, device sync, device copy)."""

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
        self.lib.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                            ctypes.c_void_p]

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

    def copy_d2d(self, dst, src, nbytes, stream=None):
        assert self.lib.hipMemcpyAsync(dst, src, nbytes, 3, stream) == 0

    def time_ms(self, stream, fn, reps):
        """Average milliseconds of fn() over reps calls, HIP events on `stream`."""
        fn()
        e0, e1 = self.event(), self.event()
        self.record(e0, stream)
        for _ in range(reps):
            fn()
        self.record(e1, stream)
        ms = self.elapsed_ms(e0, e1) / reps
        self.lib.hipEventDestroy(e0)
        self.lib.hipEventDestroy(e1)
        return ms


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


def host_cpus():
    """CPU model, physical cores, logical CPUs, the CPUs this process may run on (affinity)
    and its cgroup CPU quota; `threads` is the smallest of physical cores, affinity and
    quota: more threads than that only time-slice."""
    model, cores = 'unknown', set()
    try:
        phys = core = None
        for line in open('/proc/cpuinfo'):
            key, _, val = line.partition(':')
            key, val = key.strip(), val.strip()
            if key == 'model name' and model == 'unknown':
                model = val
            elif key == 'physical id':
                phys = val
            elif key == 'core id':
                core = val
            elif not key and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    physical = len(cores) or logical
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else logical
    quota = None
    try:
        q, period = open('/sys/fs/cgroup/cpu.max').read().split()
        if q != 'max':
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    threads = min(physical, affinity, max(1, int(quota)) if quota else physical, 256)
    return {'model': model, 'physical_cores': physical, 'logical_cpus': logical, 'affinity_cpus': affinity,
            'cgroup_cpu_quota': quota, 'threads': threads}


def _ref_bench():
    path = os.path.join(ROOT, 'oracle', '_ref', 'libref_bench.so')
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.ref_bench_run.restype = ctypes.c_int
    lib.ref_bench_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_double)]
    return lib


def _ref_rate(lib, shapes, offsets, data, threads, blocks=5, target_ms=200.0):
    """Reference CPU forward-then-inverse (materialized spectrum) over arrays `shapes` at
    `offsets` of `data`: Gpixel/s of the median pass, or None."""
    h_a = np.ascontiguousarray([h for h, _ in shapes], np.int32)
    w_a = np.ascontiguousarray([w for _, w in shapes], np.int32)
    o_a = np.ascontiguousarray(offsets, np.uint64)
    out, spec = np.empty_like(data), np.empty_like(data)
    ns = ctypes.c_double()
    if lib.ref_bench_run(len(shapes), h_a.ctypes.data, w_a.ctypes.data, o_a.ctypes.data, data.ctypes.data,
                         out.ctypes.data, spec.ctypes.data, 3, threads, blocks, target_ms, ctypes.byref(ns)):
        return None
    return round(data.size / ns.value, 4)


def cpu_baseline(inputs, blocks=None, ragged=None):
    """Reference CPU library (oracle/_ref, built from the reference sources) timed on this
    host, one plan per thread and static partitions (the bench/bench.c contract: plans and
    buffers made first, warm-up, repetitions doubled to >= 200 ms, median of 5 blocks), on
    as many threads as the process has CPUs (physical cores, capped by affinity and the
    cgroup quota).  `value`: the headline workload's four shapes, forward then inverse, one
    array of each shape per thread (equal work per thread).  Also: the 256x256 method (C1)
    on 1 and all threads, the 1,048,576 8x8 blocks (C4) and the 10,000-array ragged mix
    (C5), each forward then inverse through per-array plan calls, and the reference's own
    forward FP64 error per shape class on the ragged sample the GPU leg states."""
    lib = _ref_bench()
    if lib is None:
        return None
    cpus = host_cpus()
    threads = cpus['threads']
    # headline sample: arrays interleaved by shape, so that each thread's contiguous
    # partition holds one array of every shape
    shapes, offs, chunks, off = [], [], [], 0
    for i in range(threads):
        for (h, w), x in inputs.items():
            shapes.append((h, w))
            offs.append(off)
            chunks.append(x[i % x.shape[0]].ravel())
            off += h * w
    data = np.concatenate(chunks).astype(np.float32)
    gpx = _ref_rate(lib, shapes, offs, data, threads)
    if gpx is None:
        return None
    gpx_1 = _ref_rate(lib, shapes[:len(inputs)], offs[:len(inputs)], data[:offs[len(inputs) - 1] + shapes[len(inputs) - 1][0] * shapes[len(inputs) - 1][1]], 1)
    lib.ref_bench_avx2_fma.restype = ctypes.c_int
    avx2 = bool(lib.ref_bench_avx2_fma())
    ref = ctypes.CDLL(os.path.join(ROOT, 'oracle', '_ref', 'libbridct.so'))
    ref.bridct_plan_create.restype = ctypes.c_void_p
    ref.bridct_plan_create.argtypes = [ctypes.c_int, ctypes.c_int]
    ref.bridct_plan_route.restype = ctypes.c_char_p
    ref.bridct_plan_route.argtypes = [ctypes.c_void_p]
    ref.bridct_plan_destroy.argtypes = [ctypes.c_void_p]
    ref.bridct_plan_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    routes = {}
    for h, w in list(inputs) + [(256, 256)]:
        p = ref.bridct_plan_create(h, w)
        routes['%dx%d' % (h, w)] = ref.bridct_plan_route(p).decode() if p else None
        ref.bridct_plan_destroy(p)
    # the plain 256x256 method: 256 arrays (seeds 0..255), forward then inverse
    c1 = c1_inputs()
    c1_offs = [i * 65536 for i in range(c1.shape[0])]
    c1_rates = {'threads_%d' % t: _ref_rate(lib, [(256, 256)] * c1.shape[0], c1_offs, c1.ravel(), t)
                for t in sorted({1, threads})}
    res = {'value': gpx, 'unit': 'Gpixel/s', 'cores': threads, 'kind': 'reference',
           'sample': '%d arrays: one of each headline shape (1024x1024, 512x128, 128x1024, 32x256) per thread '
                     '(%d Mpx), forward then inverse, one plan per thread, median of 5 blocks' % (
                         len(shapes), data.size // 10 ** 6),
           'host': cpus, 'value_threads_1': gpx_1, 'plain_256x256_gpx_s': c1_rates, 'host_avx2_fma': avx2,
           'reference_routes': routes}
    if blocks is not None:
        nb = blocks.size // 64
        res['blocks_8x8'] = {'blocks': nb, 'fwd_then_inv_gpx_s': _ref_rate(
            lib, [(8, 8)] * nb, np.arange(nb, dtype=np.uint64) * 64, blocks, threads, 3)}
    if ragged is not None:
        x, rshapes, roffs, sample = ragged
        res['ragged'] = {'arrays': len(rshapes), 'fwd_then_inv_gpx_s': _ref_rate(lib, rshapes, roffs, x, threads, 3)}
        # the reference's own forward error on the arrays the GPU leg states its error on
        per_class = {}
        for i in sample:
            (h, w), o = rshapes[i], roffs[i]
            a = np.ascontiguousarray(x[o:o + h * w].reshape(h, w))
            y = np.empty_like(a)
            p = ref.bridct_plan_create(h, w)
            ref.bridct_plan_apply(p, 0, a.ctypes.data, y.ctypes.data)
            ref.bridct_plan_destroy(p)
            _class_error(per_class, h, w, y, a)
        res['ragged']['ref_forward_error_per_class'] = _class_summary(per_class)
    return res


def shape_class(h, w):
    """Shape classes of the ragged error statement, by the longer axis."""
    m = max(h, w)
    return 'max_axis_8_32' if m <= 32 else 'max_axis_64_256' if m <= 256 else 'max_axis_512_1024'


def _class_error(acc, h, w, y, x):
    """Accumulates the forward error of y against the FP64 transform of x into acc[class]."""
    ref = _dct_fp64(h) @ x.astype(np.float64) @ _dct_fp64(w).T
    d = y.astype(np.float64) - ref
    c = acc.setdefault(shape_class(h, w), {'arrays': 0, 'max_rel': 0.0, 'sq_err': 0.0, 'sq_ref': 0.0})
    c['arrays'] += 1
    c['max_rel'] = max(c['max_rel'], float(np.abs(d).max() / np.abs(ref).max()))
    c['sq_err'] += float((d * d).sum())
    c['sq_ref'] += float((ref * ref).sum())


def _class_summary(acc):
    return {k: {'arrays': v['arrays'], 'max_rel': v['max_rel'],
                'rms_rel': (v['sq_err'] / v['sq_ref']) ** 0.5 if v['sq_ref'] else 0.0} for k, v in sorted(acc.items())}


KERNELS_1024 = (
    # (label, rocprof kernel name: <length, op or mode, gather form, streaming>, mode, pass)
    ('row pass, forward', 'k_rows<1024, 1, false, true>', capi.FORWARD, 0),
    ('column pass, forward', 'k_cols<1024, 0, false, true>', capi.FORWARD, 1),
    ('row pass, inverse', 'k_rows<1024, 3, false, true>', capi.INVERSE, 0),
    ('column pass, inverse', 'k_cols<1024, 1, false, true>', capi.INVERSE, 1),
)


def committed_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE, calibrated on a device copy), or None."""
    path = os.path.join(ROOT, 'profiles', 'pmc_traffic_1024x1024_batch64.txt')
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(kernel + ' |'):
            fields = dict(f.strip().split('=') for f in line.split('|')[1:] if '=' in f)
            return int(float(fields['traffic_bytes']))
    return None


def leg_copy(hip, lib):
    """256 MiB device copies: hipMemcpy, and the library's streaming float4 copy kernel (the
    practical ceiling of one read and one write per element)."""
    n = 256 << 20
    a, b = capi.DeviceBuffer(n), capi.DeviceBuffer(n)
    ms = hip.time_ms(None, lambda: hip.copy_d2d(b.ptr, a.ptr, n), 10)
    lib.bdx_copy_stream.restype = ctypes.c_int
    lib.bdx_copy_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    ks = hip.time_ms(None, lambda: lib.bdx_copy_stream(a.ptr, b.ptr, n // 4, None), 10)
    return {'bytes': n, 'ms': round(ms, 4), 'gbs': round(2 * n / (ms * 1e-3) / 1e9, 1),
            'stream_kernel_ms': round(ks, 4), 'stream_kernel_gbs': round(2 * n / (ks * 1e-3) / 1e9, 1)}


def leg_blocks(hip, lib):
    """1,048,576 independent 8x8 blocks (an 8192x8192 image tiled, block-contiguous), pixel
    values integers 0..255 minus 128: one persistent launch (and its hipGraph) for forward,
    inverse and forward-then-inverse, vs one plan launch per block on a 65,536-block subset."""
    nb = 1 << 20
    x = blocks_values(nb, 4)
    batch = capi.Batch([(8, 8)] * nb)
    da, db = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    da.upload(x)
    res = {'blocks': nb}
    for name, mode in (('fwd', capi.FORWARD), ('inv', capi.INVERSE), ('fwd_then_inv', capi.FORWARD_THEN_INVERSE)):
        ms = hip.time_ms(batch.stream, lambda: batch.apply_device(mode, da.ptr, db.ptr), 5)
        gms = hip.time_ms(batch.stream, lambda: batch.graph_apply(mode, da.ptr, db.ptr), 5)
        res[name] = {'ms': round(ms, 4), 'gpx_s': round(x.size / (ms * 1e-3) / 1e9, 1),
                     'graph_ms': round(gms, 4), 'graph_gpx_s': round(x.size / (gms * 1e-3) / 1e9, 1)}
    assert batch.synchronize() == 0
    y = np.empty_like(x)
    db.download(y)
    res['roundtrip_max_abs_err'] = float(np.abs(y - x).max())
    plan = capi.Plan(8, 8)
    sub = 65536
    hip.sync()
    t0 = time.perf_counter()
    for i in range(sub):
        lib.bridct_plan_apply_device(plan.ptr, capi.FORWARD, da.ptr + 256 * i, db.ptr + 256 * i, 1)
    assert plan.synchronize() == 0
    t = time.perf_counter() - t0
    per_block = sub * 64 / t / 1e9
    res['per_block_launches_fwd'] = {'subset_blocks': sub, 'gpx_s': round(per_block, 4),
                                     'extrapolated_s_all_blocks': round(t * nb / sub, 2)}
    res['persistent_vs_per_block'] = round(res['fwd']['gpx_s'] / per_block, 1)
    plan.close()
    batch.close()
    return res, x


def c1_inputs():
    """256 float32 arrays of 256x256, i.i.d. uniform[-1, 1), array i from seed i."""
    return np.stack([np.random.default_rng(i).uniform(-1, 1, (256, 256)).astype(np.float32)
                     for i in range(256)])


def leg_c1(hip, lib):
    """Plain method: 256x256 float32 arrays, uniform[-1, 1) from seeds 0..255, forward DCT-II
    then inverse DCT-III (spectrum materialized), FP32 butterflies: one batched launch per
    direction and one launch per array; max relative errors against FP64."""
    x = c1_inputs()
    n = x.shape[0]
    a, b, c = (capi.DeviceBuffer(x.nbytes) for _ in range(3))
    a.upload(x)
    plan = capi.Plan(256, 256)
    res = {'arrays': n}

    def both():
        assert plan.apply_device(capi.FORWARD, a.ptr, b.ptr, n) == 0
        assert plan.apply_device(capi.INVERSE, b.ptr, c.ptr, n) == 0

    ms = hip.time_ms(plan.stream, both, 5)
    res['batched'] = {'ms_fwd_plus_inv': round(ms, 4), 'gpx_s': round(x.size / (ms * 1e-3) / 1e9, 1)}
    pp = plan.ptr
    step = 256 * 256 * 4

    def per_array():
        for i in range(n):
            lib.bridct_plan_apply_device(pp, capi.FORWARD, a.ptr + step * i, b.ptr + step * i, 1)
            lib.bridct_plan_apply_device(pp, capi.INVERSE, b.ptr + step * i, c.ptr + step * i, 1)

    hip.sync()
    t0 = time.perf_counter()
    per_array()
    assert plan.synchronize() == 0
    t = time.perf_counter() - t0
    res['per_array_launches'] = {'s': round(t, 4), 'gpx_s': round(x.size / t / 1e9, 2)}
    y, z = np.empty_like(x), np.empty_like(x)
    b.download(y)
    c.download(z)
    q = _dct_fp64(256)
    ef = ei = er = 0.0
    for i in range(n):
        x64 = x[i].astype(np.float64)
        yr = q @ x64 @ q.T
        ir = q.T @ x64 @ q
        ef = max(ef, float(np.abs(y[i] - yr).max() / np.abs(yr).max()))
        er = max(er, float(np.abs(z[i] - x64).max() / np.abs(x64).max()))
        if i < 32:
            assert plan.apply_device(capi.INVERSE, a.ptr + step * i, b.ptr + step * i, 1) == 0
    assert plan.synchronize() == 0
    b.download(y)
    for i in range(32):
        ir = q.T @ x[i].astype(np.float64) @ q
        ei = max(ei, float(np.abs(y[i] - ir).max() / np.abs(ir).max()))
    res['max_rel_err_vs_fp64'] = {'forward': ef, 'inverse_32_arrays': ei, 'forward_then_inverse': er}
    plan.close()
    return res


MFMA_PEAK_TFLOPS = 2500.0   # MI355X dense FP16 / BF16


def _dct_fp64(n):
    """Orthonormal DCT-II matrix in FP64 (docs/API.md:7-13 of the reference), for the
    error statement of the dense-route leg."""
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    return np.sqrt(np.where(k == 0, 1.0, 2.0) / n) * np.cos(np.pi * ((k * (2 * j + 1)) % (4 * n)) / (2.0 * n))


def leg_dense(hip):
    """Dense split-MFMA route vs FP32 butterflies at 64^2, 256^2 and 1024^2, batch 128,
    uniform[-1, 1) inputs: time per direction, MFMA rate (3 products x 2 h w (h + w) flops
    per array) and max / RMS relative error against FP64 on two arrays per shape."""
    res = {}
    names = {capi.ROUTE_BUTTERFLY: 'butterfly_fp32', capi.ROUTE_MFMA_F16: 'mfma_f16_split',
             capi.ROUTE_MFMA_BF16: 'mfma_bf16_split'}
    for n in (64, 256, 1024):
        batch = 128
        x = np.random.default_rng(n).uniform(-1, 1, (batch, n, n)).astype(np.float32)
        a, b = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
        a.upload(x)
        q = _dct_fp64(n)
        ref = [q @ x[i].astype(np.float64) @ q.T for i in range(2)]
        shape = {}
        for route, name in names.items():
            plan = capi.Plan(n, n)
            assert plan.set_route(route) == 0
            rec = {}
            for d, mode in (('fwd', capi.FORWARD), ('inv', capi.INVERSE)):
                ms = hip.time_ms(plan.stream, lambda: plan.apply_device(mode, a.ptr, b.ptr, batch), 3)
                rec[d + '_ms'] = round(ms, 4)
                rec[d + '_gpx_s'] = round(x.size / (ms * 1e-3) / 1e9, 1)
                if route != capi.ROUTE_BUTTERFLY:
                    tf = 3 * 2.0 * n * n * (2 * n) * batch / (ms * 1e-3) / 1e12
                    rec[d + '_tflops'] = round(tf, 1)
                    rec[d + '_mfma_frac'] = round(tf / MFMA_PEAK_TFLOPS, 4)
            assert plan.apply_device(capi.FORWARD, a.ptr, b.ptr, batch) == 0
            assert plan.synchronize() == 0
            y = np.empty_like(x)
            b.download(y)
            mx = max(float(np.abs(y[i] - ref[i]).max() / np.abs(ref[i]).max()) for i in range(2))
            rms = max(float(np.sqrt(np.mean((y[i] - ref[i]) ** 2) / np.mean(ref[i] ** 2))) for i in range(2))
            rec['fwd_max_rel_err'] = mx
            rec['fwd_rms_rel_err'] = rms
            shape[name] = rec
            plan.close()
        best = min((shape[nm]['fwd_ms'] + shape[nm]['inv_ms'], nm) for nm in shape)[1]
        shape['faster_route'] = best
        res['%dx%d' % (n, n)] = shape
        del a, b
    return res


# Luminance quantization table of ITU-T T.81 Annex K (Table K.1), row-major u * 8 + v.
JPEG_LUMA = np.array([
    16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55,
    14, 13, 16, 24, 40, 57, 69, 56, 14, 17, 22, 29, 51, 87, 80, 62,
    18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99], np.float32)


def leg_codec(hip):
    """8x8 block codec step on an 8192x8192 u8 image (a smooth synthetic field plus noise)
    with the T.81 Annex K luminance table: encode (u8 -> level shift, 8x8 DCT, quantize ->
    int16) and decode, each one kernel; 3 B per pixel of HBM traffic per direction."""
    h = w = 8192
    rng = np.random.default_rng(11)
    r = np.arange(h, dtype=np.float32)[:, None]
    c = np.arange(w, dtype=np.float32)[None, :]
    img = np.clip(np.rint(128 + 80 * np.sin(r / 97.0) * np.cos(c / 131.0) +
                          rng.normal(0, 3, (h, w)).astype(np.float32)), 0, 255).astype(np.uint8)
    nb = (h // 8) * (w // 8)
    codec = capi.Codec8(JPEG_LUMA)
    d_img, d_co, d_out = capi.DeviceBuffer(h * w), capi.DeviceBuffer(nb * 128), capi.DeviceBuffer(h * w)
    d_img.upload(img)
    res = {'image': '%dx%d u8' % (h, w), 'quant': 'T.81 Annex K luminance'}
    for name, fn in (('encode', lambda: codec.encode_device(d_img.ptr, h, w, w, d_co.ptr)),
                     ('decode', lambda: codec.decode_device(d_co.ptr, h, w, w, d_out.ptr))):
        ms = hip.time_ms(codec.stream, fn, 10)
        gbs = 3.0 * h * w / (ms * 1e-3) / 1e9
        res[name] = {'ms': round(ms, 4), 'gpx_s': round(h * w / (ms * 1e-3) / 1e9, 1), 'gbs': round(gbs, 1),
                     'hbm_frac': round(gbs / HBM_PEAK_GBS, 4)}
    assert codec.synchronize() == 0
    back = np.empty_like(img)
    d_out.download(back)
    co = np.empty((nb, 64), np.int16)
    d_co.download(co)
    mse = float(np.mean((back.astype(np.float64) - img) ** 2))
    res['psnr_db'] = round(10 * np.log10(255.0 ** 2 / mse), 2) if mse else None
    res['nonzero_coeff_frac'] = round(float(np.count_nonzero(co)) / co.size, 4)
    codec.close()
    return res


def ragged_shapes(n=10000, seed=7):
    """Shapes of the ragged mixed batch: each axis uniform over {8, ..., 1024}; returns the
    shapes and the generator, positioned to draw the values."""
    rng = np.random.default_rng(seed)
    lengths = [8, 16, 32, 64, 128, 256, 512, 1024]
    return list(zip(rng.choice(lengths, n).tolist(), rng.choice(lengths, n).tolist())), rng


def ragged_values(shapes, offsets, elements, rng):
    """Values of the ragged batch, back to back at `offsets`: uniform[-1, 1) times a
    per-array scale 10^U[-3, 3]."""
    x = np.empty(elements, np.float32)
    for o, (h, w) in zip(offsets, shapes):
        x[o:o + h * w] = rng.uniform(-1, 1, h * w) * 10 ** rng.uniform(-3, 3)
    return x


def blocks_values(nb=1 << 20, seed=4):
    """nb block-contiguous 8x8 blocks, pixel values integers 0..255 minus 128."""
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 256, nb * 64) - 128).astype(np.float32)


def leg_ragged(hip, lib):
    """10,000 float32 arrays, each axis uniform over {8, ..., 1024}, values uniform[-1, 1)
    times a per-array scale 10^U[-3, 3], seed 7: the batch object (one persistent launch per
    direction, and the grouped per-shape kernels; `fwd` / `inv` report the automatic choice)
    vs one plan launch per array (the same FP32 butterfly kernels)."""
    n = 10000
    shapes, rng = ragged_shapes(n, 7)
    batch = capi.Batch(shapes)
    x = ragged_values(shapes, batch.offsets, batch.elements, rng)
    da, db = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    da.upload(x)
    res = {'arrays': n, 'pixels': int(x.size), 'strategy_auto': {1: 'persistent', 2: 'grouped'}[batch.strategy]}
    worst = 0.0
    for sname, strategy in (('persistent', capi.BATCH_PERSISTENT), ('grouped', capi.BATCH_GROUPED)):
        assert batch.set_strategy(strategy) == 0
        r = {}
        for name, mode in (('fwd', capi.FORWARD), ('inv', capi.INVERSE)):
            ms = hip.time_ms(batch.stream, lambda: batch.apply_device(mode, da.ptr, db.ptr), 3)
            r[name] = {'ms': round(ms, 3), 'gpx_s': round(x.size / (ms * 1e-3) / 1e9, 1),
                       'hbm_frac': round(8.0 * x.size / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        gms = hip.time_ms(batch.stream, lambda: batch.graph_apply(capi.FORWARD_THEN_INVERSE, da.ptr, db.ptr), 3)
        r['graph_fwd_then_inv'] = {'ms': round(gms, 3), 'gpx_s': round(x.size / (gms * 1e-3) / 1e9, 1)}
        assert batch.synchronize() == 0
        y = np.empty_like(x)
        db.download(y)
        for o, (h, w) in zip(batch.offsets, shapes):
            a, b = x[o:o + h * w], y[o:o + h * w]
            worst = max(worst, float(np.abs(b - a).max() / np.abs(a).max()))
        res[sname] = r
    assert batch.set_strategy(capi.BATCH_AUTO) == 0
    # the automatic choice is the headline of this leg
    best = res['grouped' if batch.strategy == capi.BATCH_GROUPED else 'persistent']
    res['fwd'], res['inv'] = best['fwd'], best['inv']
    res['roundtrip_max_rel_err'] = worst
    # forward error against FP64 per shape class, the automatic strategy, on up to three
    # arrays of every distinct shape (max-relative per array; RMS over the class)
    assert batch.set_strategy(capi.BATCH_AUTO) == 0
    assert batch.apply_device(capi.FORWARD, da.ptr, db.ptr) == 0 and batch.synchronize() == 0
    y = np.empty_like(x)
    db.download(y)
    seen, sample = {}, []
    for i, s in enumerate(shapes):
        if seen.get(s, 0) < 3:
            seen[s] = seen.get(s, 0) + 1
            sample.append(i)
    per_class = {}
    for i in sample:
        (h, w), o = shapes[i], batch.offsets[i]
        _class_error(per_class, h, w, y[o:o + h * w].reshape(h, w), x[o:o + h * w].reshape(h, w))
    res['forward_error_per_class'] = _class_summary(per_class)
    res['forward_error_sample'] = 'up to 3 arrays of each of the 64 shapes (%d arrays)' % len(sample)
    plans = {s: capi.Plan(*s) for s in sorted(set(shapes))}
    for name, mode in (('fwd', capi.FORWARD), ('inv', capi.INVERSE)):
        hip.sync()
        t0 = time.perf_counter()
        for o, s in zip(batch.offsets, shapes):
            lib.bridct_plan_apply_device(plans[s].ptr, mode, da.ptr + 4 * o, db.ptr + 4 * o, 1)
        hip.sync()
        t = time.perf_counter() - t0
        res['per_array_launches_' + name] = {'s': round(t, 4), 'gpx_s': round(x.size / t / 1e9, 2)}
    plain = res['per_array_launches_fwd']['gpx_s'] + res['per_array_launches_inv']['gpx_s']
    res['batched_vs_plain'] = round((res['fwd']['gpx_s'] + res['inv']['gpx_s']) / plain, 2)
    res['persistent_vs_plain'] = round((res['persistent']['fwd']['gpx_s'] + res['persistent']['inv']['gpx_s']) /
                                       plain, 2)
    for p in plans.values():
        p.close()
    offsets = list(batch.offsets)
    batch.close()
    return res, (x, shapes, offsets, sample)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    # a four-shape step takes about 0.5 ms and reaches its steady time after about 40 steps,
    # so the defaults warm up for 50 steps (25 ms) and time 50
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--legs', default='c2,copy,c1,dense,blocks,ragged,codec',
                    help='comma list of legs: c2 (always run), copy, c1, dense, blocks, ragged, codec')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: the transform has no multi-GPU path in this '
              'build (arrays shard across GPUs with no exchange); run with --gpus 1.',
              file=sys.stderr)
        return 2
    legs = set(args.legs.split(','))
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

    # Setup: the step runs for about 0.3 s before the W warm-up steps, so that the timed
    # steps see the clock the GPU settles at under this load (the first steps after idle
    # run up to 10% slower, and 50 steps last only 25 ms).  Not part of the timed region.
    t_ramp = time.perf_counter()
    while time.perf_counter() - t_ramp < 0.3:
        for _ in range(20):
            step()
        hip.sync()
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

    # per-shape, per-direction transform times with HIP events on each plan's stream
    detail = {}
    for h, w, plan, (a, b, c) in runs:
        for name, mode, src, dst in (('fwd', capi.FORWARD, a, b), ('inv', capi.INVERSE, b, c)):
            ms = hip.time_ms(plan.stream, lambda: plan.apply_device(mode, src.ptr, dst.ptr, BATCH), 5)
            gbs = 8.0 * h * w * BATCH / (ms * 1e-3) / 1e9
            detail['%dx%d_%s' % (h, w, name)] = {
                'ms': round(ms, 4), 'gpx_s': round(h * w * BATCH / (ms * 1e-3) / 1e9, 2),
                'hbm_frac': round(gbs / HBM_PEAK_GBS, 4)}
    # sanity: round trip of the 1024^2 batch
    x = inputs[(1024, 1024)]
    back = np.empty_like(x)
    runs[0][3][2].download(back)
    rt_err = float(np.abs(back - x).max() / np.abs(x).max())

    # each kernel of the 1024^2 transforms as it runs in the transform (row pass, then the
    # column pass on its output), with HIP events between the launches on the plan stream:
    # 8 B per pixel per kernel (one read, one write)
    h, w, plan, (a, b, c) = runs[0]
    px = 1024 * 1024 * BATCH
    kernels = {}
    reps = 21   # 20 timed dispatches of each kernel
    for mode in (capi.FORWARD, capi.INVERSE):
        ev = [hip.event() for _ in range(2 * reps + 1)]
        hip.record(ev[0], plan.stream)
        for r in range(reps):   # back to back, so that the queue never runs dry between events
            assert plan.apply_device_pass(mode, 0, a.ptr, b.ptr, BATCH) == 0
            hip.record(ev[2 * r + 1], plan.stream)
            assert plan.apply_device_pass(mode, 1, b.ptr, b.ptr, BATCH) == 0
            hip.record(ev[2 * r + 2], plan.stream)
        times = [hip.elapsed_ms(ev[i], ev[i + 1]) for i in range(2 * reps)]
        for e in ev:
            hip.lib.hipEventDestroy(e)
        for pss in (0, 1):
            ms = sum(times[2 + pss::2]) / (reps - 1)   # the first repetition is a warm-up
            label, kname = next((l, k) for l, k, m, p in KERNELS_1024 if m == mode and p == pss)
            kernels[kname] = {'what': label, 'ms': round(ms, 4), 'gbs': round(8.0 * px / (ms * 1e-3) / 1e9, 1)}
    dom_name = max(kernels, key=lambda k: kernels[k]['ms'])
    dom = kernels[dom_name]
    achieved = 8.0 * px / (dom['ms'] * 1e-3) / 1e9
    traffic = committed_traffic(dom_name)
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
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
                     'kernel': dom_name + ' (' + dom['what'] + ', 64 x 1024x1024, 8 B/px per launch)'},
        'kernels_1024x1024': kernels,
        'per_shape': detail,
        'roundtrip_max_rel_err_1024': rt_err,
    }
    if 'copy' in legs:
        line['copy_reference'] = leg_copy(hip, lib)
    if 'c1' in legs:
        line['plain_256x256'] = leg_c1(hip, lib)
    if 'dense' in legs:
        line['dense_route'] = leg_dense(hip)
    blocks_x = ragged_data = None
    if 'blocks' in legs:
        line['blocks_8x8'], blocks_x = leg_blocks(hip, lib)
    if 'ragged' in legs:
        line['ragged'], ragged_data = leg_ragged(hip, lib)
    if 'codec' in legs:
        line['codec_8x8'] = leg_codec(hip)
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(inputs, blocks_x, ragged_data)
    print(json.dumps(line))
    return 0


if __name__ == '__main__':
    sys.exit(main())

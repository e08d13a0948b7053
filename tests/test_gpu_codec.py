"""GPU tests of the 8x8 block codec step (bridct_codec8_*, bdx_codec.hip) through the C ABI:
coefficients and pixels against the FP64 oracle (equal away from rounding boundaries, at
most one step apart within FP32 reach of one), against the plan path's own 8x8 transform
(the fused kernel computes the same transform), device and host entry points bitwise
equal, strided rasters, saturation, and argument errors.  The reference has no codec
step; its 8x8 transform contract is src/plan.c:126-163 for an 8 x 8 plan."""
import ctypes

import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu

# FP32 reach near a rounding boundary, in units of the rounded quantity: the 8x8 transform
# error is a few 1e-7 of max|Y| <= 1024, divided by the step.  Exact ties are common (with
# unit steps the DC term of integer pixels is a multiple of 1/8), and neither the FP32
# kernel nor the FP64 oracle resolves them exactly, so either neighbour is accepted there.
def band(quant):
    return 1e-3 / np.asarray(quant, np.float64).reshape(64)


QUANTS = {'ones': np.ones(64, np.float32), 'luma': S.JPEG_LUMA,
          'luma_q90': np.maximum(np.floor(S.JPEG_LUMA * 0.2 + 0.5), 1).astype(np.float32),
          'fine': np.full(64, 0.25, np.float32)}


def image(h, w, seed, kind='random'):
    rng = np.random.default_rng(seed)
    if kind == 'random':
        return rng.integers(0, 256, (h, w), dtype=np.uint8)
    r, c = np.indices((h, w))
    smooth = 128 + 90 * np.sin(r / 23.0) * np.cos(c / 31.0) + rng.normal(0, 4, (h, w))
    return np.clip(np.rint(smooth), 0, 255).astype(np.uint8)


def device_encode(codec, img):
    h, w = img.shape
    nb = (h // 8) * (w // 8)
    d_img, d_co = capi.DeviceBuffer(img.nbytes), capi.DeviceBuffer(nb * 128)
    d_img.upload(np.ascontiguousarray(img))
    assert codec.encode_device(d_img.ptr, h, w, w, d_co.ptr) == 0
    assert codec.synchronize() == 0
    out = np.empty((nb, 64), np.int16)
    d_co.download(out)
    return out


def device_decode(codec, coeffs, h, w):
    d_co, d_img = capi.DeviceBuffer(coeffs.nbytes), capi.DeviceBuffer(h * w)
    d_co.upload(np.ascontiguousarray(coeffs))
    assert codec.decode_device(d_co.ptr, h, w, w, d_img.ptr) == 0
    assert codec.synchronize() == 0
    out = np.empty((h, w), np.uint8)
    d_img.download(out)
    return out


@pytest.mark.parametrize('qname', list(QUANTS))
@pytest.mark.parametrize('h,w,kind', [(8, 8, 'random'), (64, 520, 'random'), (256, 192, 'smooth'),
                                      (1080 // 8 * 8, 1920, 'smooth')])
def test_encode_decode_against_oracle(qname, h, w, kind):
    quant = QUANTS[qname]
    img = image(h, w, h + w, kind)
    codec = capi.Codec8(quant)
    try:
        print('codec %dx%d %s %s' % (h, w, kind, qname), flush=True)
        got = device_encode(codec, img)
        want, quot = S.oracle_encode8(img, quant)
        S.rounding_agrees(got, want, quot, band(quant)[None, :])
        # decode of the oracle's coefficients, against the oracle's decode
        pix = device_decode(codec, want, h, w)
        want_pix, val = S.oracle_decode8(want, h, w, quant)
        S.rounding_agrees(pix, want_pix, val, 1e-3)
    finally:
        codec.close()


def test_fused_transform_is_the_plan_transform():
    """Coefficients equal round(plan FORWARD(x - 128) / Q) of the 8 x 8 plan path, and
    pixels round(plan INVERSE(q Q) + 128), up to FP32 rounding at ties."""
    h, w = 128, 256
    img = image(h, w, 3)
    quant = QUANTS['luma']
    codec = capi.Codec8(quant)
    plan = capi.Plan(8, 8)
    try:
        blocks = img.reshape(h // 8, 8, w // 8, 8).transpose(0, 2, 1, 3).reshape(-1, 8, 8)
        x = np.ascontiguousarray(blocks.astype(np.float32) - 128)
        y = np.empty_like(x)
        assert plan.apply_batch(capi.FORWARD, x, y) == 0
        quot = y.reshape(-1, 64).astype(np.float64) / quant
        got = device_encode(codec, img)
        want = np.clip(np.rint(y.reshape(-1, 64) / quant), -32768, 32767)
        S.rounding_agrees(got, want, quot, 1e-4)
        deq = (got.astype(np.float32) * quant).reshape(-1, 8, 8)
        xr = np.empty_like(deq)
        assert plan.apply_batch(capi.INVERSE, np.ascontiguousarray(deq), xr) == 0
        raster = xr.reshape(h // 8, w // 8, 8, 8).transpose(0, 2, 1, 3).reshape(h, w).astype(np.float64) + 128
        pix = device_decode(codec, got, h, w)
        S.rounding_agrees(pix, np.clip(np.rint(raster), 0, 255), raster, 1e-4)
    finally:
        plan.close()
        codec.close()


def test_host_entry_points_and_strides():
    big = image(200, 336, 9)
    view = big[8:200, 16:336]              # rows 336 bytes apart, 192 x 320
    codec = capi.Codec8(QUANTS['luma'])
    try:
        host = codec.encode(view)
        dev = device_encode(codec, np.ascontiguousarray(view))
        assert np.array_equal(host, dev)
        pix = codec.decode(host, 192, 320)
        assert np.array_equal(pix, device_decode(codec, host, 192, 320))
        # strided host decode into a wider raster leaves the gap bytes alone
        wide = np.full((192, 336), 7, np.uint8)
        lib = capi.load()
        assert lib.bridct_codec8_decode(codec.ptr, capi.ptr(host), 192, 320, 336, capi.ptr(wide)) == 0
        assert np.array_equal(wide[:, :320], pix) and np.all(wide[:, 320:] == 7)
    finally:
        codec.close()


def test_host_entry_points_staged():
    """Host calls whose copies reach 1 MiB go through pinned staging in 8 MiB chunks: a
    2048 x 3072 image (6 MiB of pixels, 12 MiB of coefficients) and a strided view of it
    give the device path's results bitwise."""
    big = image(2048, 3072 + 64, 11)
    codec = capi.Codec8(QUANTS['luma'])
    try:
        for view in (np.ascontiguousarray(big[:, :3072]), big[:, 64:]):
            print('staged codec host call, contiguous %s' % view.flags.c_contiguous, flush=True)
            h, w = view.shape
            host = codec.encode(view)
            assert np.array_equal(host, device_encode(codec, np.ascontiguousarray(view)))
            pix = codec.decode(host, h, w)
            assert np.array_equal(pix, device_decode(codec, host, h, w))
    finally:
        codec.close()


def test_saturation():
    codec = capi.Codec8(np.full(64, 1e-3, np.float32))
    try:
        img = np.zeros((16, 16), np.uint8)
        img[8:, 8:] = 255
        got = device_encode(codec, img)
        assert got[0, 0] == -32768 and got[3, 0] == 32767
    finally:
        codec.close()
    codec = capi.Codec8(np.full(64, 50.0, np.float32))
    try:
        coeffs = np.zeros((4, 64), np.int16)
        coeffs[0, 0], coeffs[1, 0] = 32767, -32768
        pix = device_decode(codec, coeffs, 16, 16)
        assert np.all(pix[:8, :8] == 255) and np.all(pix[:8, 8:] == 0) and np.all(pix[8:] == 128)
    finally:
        codec.close()


def test_argument_errors():
    lib = capi.load()
    bad_q = np.ones(64, np.float32)
    for v in (0.0, -1.0, np.inf, np.nan):
        bad_q[5] = v
        assert not lib.bridct_codec8_create(capi.ptr(bad_q))
    assert not lib.bridct_codec8_create(None)
    codec = capi.Codec8(np.ones(64, np.float32))
    try:
        d_img, d_co = capi.DeviceBuffer(64 * 64), capi.DeviceBuffer(64 * 128)
        enc = lambda *a: codec.encode_device(*a)
        assert enc(d_img.ptr, 12, 64, 64, d_co.ptr) == -1           # not a multiple of 8
        assert enc(d_img.ptr, 64, 0, 64, d_co.ptr) == -1
        assert enc(d_img.ptr, 64, 64, 32, d_co.ptr) == -1           # stride below width
        assert enc(d_img.ptr + 4, 8, 8, 8, d_co.ptr) == -3          # misaligned raster
        assert enc(d_img.ptr, 8, 8, 12, d_co.ptr) == -3             # stride not a multiple of 8
        assert enc(d_img.ptr, 8, 8, 8, d_co.ptr + 8) == -3          # misaligned coefficients
        assert enc(None, 8, 8, 8, d_co.ptr) == -1
        assert lib.bridct_codec8_encode_device(None, d_img.ptr, 8, 8, 8, d_co.ptr) == -1
        host = np.zeros(64 * 64 + 64 * 128, np.uint8)
        # host buffers that overlap
        assert lib.bridct_codec8_encode(codec.ptr, capi.ptr(host), 64, 64, 64,
                                        ctypes.c_void_p(host.ctypes.data + 64)) == -4
    finally:
        codec.close()
    lib.bridct_codec8_destroy(None)

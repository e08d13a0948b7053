"""Pins the FP64 oracle of the 8x8 block codec step (oracle_codec8_encode/decode in
oracle/bridct_oracle.c) before it is trusted as the checker of the GPU codec kernels.

  * the unrounded coefficients against SciPy's independent orthonormal DCT-II
    (scipy.fft.dctn) of the level-shifted blocks, divided by the steps;
  * the unrounded pixels against SciPy's DCT-III of the dequantized coefficients;
  * rounding: coefficients are NumPy's round-half-even of the quotients, saturated to
    int16, pixels likewise to 0..255, including exact ties;
  * closed forms: a constant block has only a DC term 8 (v - 128); one coefficient decodes
    to 128 plus the outer product of two DCT basis rows;
  * block order and strided rasters.
The quantization rule is baseline JPEG's (ITU-T T.81, A.3.4) with ties to even; the
reference has no codec step, so nothing here comes from its tests.
"""
import numpy as np
import pytest
import scipy.fft

import support as S


def blocks_of(img):
    h, w = img.shape
    return img.reshape(h // 8, 8, w // 8, 8).transpose(0, 2, 1, 3).reshape(-1, 8, 8)


def raster_of(blocks, h, w):
    return blocks.reshape(h // 8, w // 8, 8, 8).transpose(0, 2, 1, 3).reshape(h, w)


@pytest.mark.parametrize('h,w', [(8, 8), (16, 40), (64, 24)])
def test_encode_matches_scipy(h, w):
    rng = np.random.default_rng(h * 100 + w)
    img = rng.integers(0, 256, (h, w), dtype=np.uint8)
    for quant in (np.ones(64), S.JPEG_LUMA, rng.uniform(0.3, 50, 64)):
        coeffs, quot = S.oracle_encode8(img, quant)
        y = scipy.fft.dctn(blocks_of(img).astype(np.float64) - 128, type=2, axes=(1, 2), norm='ortho')
        want = y.reshape(-1, 64) / np.asarray(quant, np.float64)
        assert np.allclose(quot, want, rtol=0, atol=1e-11)
        assert np.array_equal(coeffs, np.clip(np.rint(quot), -32768, 32767).astype(np.int16))


@pytest.mark.parametrize('h,w', [(8, 8), (24, 16)])
def test_decode_matches_scipy(h, w):
    rng = np.random.default_rng(h + w)
    nb = (h // 8) * (w // 8)
    for quant in (np.ones(64), S.JPEG_LUMA):
        coeffs = rng.integers(-60, 61, (nb, 64)).astype(np.int16)
        img, val = S.oracle_decode8(coeffs, h, w, quant)
        x = scipy.fft.idctn(coeffs.reshape(-1, 8, 8) * np.asarray(quant, np.float64).reshape(8, 8), type=2,
                            axes=(1, 2), norm='ortho') + 128
        assert np.allclose(val, raster_of(x, h, w), rtol=0, atol=1e-11)
        assert np.array_equal(img, np.clip(np.rint(val), 0, 255).astype(np.uint8))


def test_constant_blocks_and_ties():
    # DC of a constant block v is 8 (v - 128); with Q0 = 16 the quotient (v - 128) / 2 is
    # a tie for odd v - 128, rounded to even
    for v in (0, 1, 127, 128, 129, 131, 255):
        img = np.full((8, 8), v, np.uint8)
        quant = np.ones(64)
        quant[0] = 16
        coeffs, quot = S.oracle_encode8(img, quant)
        assert abs(quot[0, 0] - (v - 128) / 2) < 1e-12 and np.abs(quot[0, 1:]).max() < 1e-12
        assert np.all(coeffs[0, 1:] == 0)
        if (v - 128) % 2 == 0:
            assert coeffs[0, 0] == (v - 128) // 2
    # saturation
    coeffs, _ = S.oracle_encode8(np.zeros((8, 8), np.uint8), np.full(64, 1e-3))
    assert coeffs[0, 0] == -32768
    coeffs, _ = S.oracle_encode8(np.full((8, 8), 255, np.uint8), np.full(64, 1e-3))
    assert coeffs[0, 0] == 32767


def test_single_coefficient_decodes_to_basis():
    q8 = S.qmatrix(8)
    for u, v in ((0, 0), (1, 2), (7, 7), (3, 0)):
        coeffs = np.zeros((1, 64), np.int16)
        coeffs[0, u * 8 + v] = 100
        _, val = S.oracle_decode8(coeffs, 8, 8, np.ones(64))
        k = np.arange(8)
        basis = np.sqrt(np.where(np.arange(8) == 0, 1., 2.) / 8)[:, None] * np.cos(np.pi * k[:, None] * (k + .5) / 8)
        assert np.allclose(val, 128 + 100 * np.outer(basis[u], basis[v]), atol=1e-11)
        assert np.allclose(basis, q8, atol=1e-15)
        img, _ = S.oracle_decode8(coeffs * 10, 8, 8, np.ones(64))
        assert img.min() >= 0 and img.max() <= 255


def test_block_order_and_stride():
    rng = np.random.default_rng(5)
    big = rng.integers(0, 256, (32, 48), dtype=np.uint8)
    view = big[8:24, 8:40]              # strided rows
    coeffs, _ = S.oracle_encode8(view, S.JPEG_LUMA)
    ref, _ = S.oracle_encode8(np.ascontiguousarray(view), S.JPEG_LUMA)
    assert np.array_equal(coeffs, ref)
    # block b = by * (W / 8) + bx
    one, _ = S.oracle_encode8(np.ascontiguousarray(view[8:16, 16:24]), S.JPEG_LUMA)
    assert np.array_equal(coeffs[1 * 4 + 2], one[0])


def test_planted_rounding_mistake_is_caught():
    """rounding_agrees itself: a result rounded half-up at exact ties, or off by one away
    from a boundary, is reported."""
    unrounded = np.array([0.5, 1.5, 2.2, -0.5, 3.49])
    want = np.rint(unrounded)
    assert S.rounding_agrees(want, want, unrounded, 1e-6) == 3
    half_up = np.floor(unrounded + 0.5)
    S.rounding_agrees(half_up, want, unrounded, 1e-6)      # ties only: allowed
    with pytest.raises(AssertionError):
        S.rounding_agrees(want + np.array([0, 0, 1, 0, 0]), want, unrounded, 1e-6)

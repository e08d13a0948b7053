"""Shared test helpers: the oracle (FP64 closed form), the reference CPU library built from
source under oracle/_ref (when present), error metrics and the reference's input families.

Only tests (and the smoke check / bench CPU-baseline leg) load anything under oracle/.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# Where GPU tests write their result tables (error per shape, corpus gate summary).
OUT_DIR = os.environ.get('BRIDCT_TEST_OUT', os.path.join(ROOT, 'test_out'))
ORACLE_LIB = os.path.join(ROOT, 'oracle', 'liboracle.so')
REF_LIB = os.path.join(ROOT, 'oracle', '_ref', 'libbridct.so')
LENGTHS = (8, 16, 32, 64, 128, 256, 512, 1024)
SHAPES = [(h, w) for h in LENGTHS for w in LENGTHS]

_dp = ctypes.POINTER(ctypes.c_double)
_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_LIB)
        lib.oracle_supported.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.oracle_sj_scale.restype = ctypes.c_double
        lib.oracle_sj_scale.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.oracle_axis_scale.restype = ctypes.c_double
        lib.oracle_axis_scale.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.oracle_norm_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_qmatrix.argtypes = [ctypes.c_int, ctypes.c_void_p]
        lib.oracle_core.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_dct2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p]
        _oracle = lib
    return _oracle


def qmatrix(n):
    q = np.empty((n, n), np.float64)
    assert oracle().oracle_qmatrix(n, q.ctypes.data) == 0
    return q


_qcache = {}


def oracle_2d(x, mode):
    """FP64 orthonormal 2D transform of x (..., h, w) through the oracle's Q matrices:
    mode 0: Q_h X Q_w^T, 1: Q_h^T X Q_w, 2: X (reference docs/API.md:9-12)."""
    x = np.asarray(x, np.float64)
    h, w = x.shape[-2:]
    if mode == 2:
        return x.copy()
    for n in (h, w):
        if n not in _qcache:
            _qcache[n] = qmatrix(n)
    qh, qw = _qcache[h], _qcache[w]
    if mode == 0:
        return qh @ x @ qw.T
    return qh.T @ x @ qw


def oracle_2d_direct(x, mode):
    """The oracle's own C loop (oracle_dct2d), for one h x w array."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    h, w = x.shape
    assert oracle().oracle_dct2d(h, w, mode, x.ctypes.data, y.ctypes.data) == 0
    return y


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """Reference BRiDCT CPU library built from the reference sources (oracle/Makefile)."""
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_LIB)
        lib.bridct_plan_create.restype = ctypes.c_void_p
        lib.bridct_plan_create.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.bridct_plan_destroy.argtypes = [ctypes.c_void_p]
        lib.bridct_plan_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.bridct_plan_route.restype = ctypes.c_char_p
        lib.bridct_plan_route.argtypes = [ctypes.c_void_p]
        _ref = lib
    return _ref


def ref_apply(x, mode):
    """Reference CPU output for (..., h, w) float32 input (mode 3: forward then inverse)."""
    x = np.ascontiguousarray(x, np.float32)
    h, w = x.shape[-2:]
    lib = ref()
    plan = lib.bridct_plan_create(h, w)
    assert plan
    flat = x.reshape(-1, h, w)
    out = np.empty_like(flat)
    spec = np.empty((h, w), np.float32)
    try:
        for i in range(flat.shape[0]):
            if mode == 3:
                assert lib.bridct_plan_apply(plan, 0, flat[i].ctypes.data, spec.ctypes.data) == 0
                assert lib.bridct_plan_apply(plan, 1, spec.ctypes.data, out[i].ctypes.data) == 0
            else:
                assert lib.bridct_plan_apply(plan, mode, flat[i].ctypes.data, out[i].ctypes.data) == 0
    finally:
        lib.bridct_plan_destroy(plan)
    return out.reshape(x.shape)


def max_rel(y, ref):
    """max|y - ref| / max|ref| (infinity-norm relative error)."""
    d = np.abs(np.asarray(y, np.float64) - ref).max()
    m = np.abs(ref).max()
    return float(d / m) if m else float(d)


def frob_input_rel(y, ref, x):
    """||y - ref||_F / ||x||_F (reference dataset/evaluate.py:17-21)."""
    err = float(np.linalg.norm(np.asarray(y, np.float64) - ref))
    n = float(np.linalg.norm(np.asarray(x, np.float64)))
    return err / n if n else (0.0 if err == 0.0 else float('inf'))


def rel_l2(y, ref):
    d = float(np.linalg.norm(np.asarray(y, np.float64) - np.asarray(ref, np.float64)))
    n = float(np.linalg.norm(np.asarray(ref, np.float64)))
    return d / n if n else d


def error_norm(actual, expected, input_norm):
    """Input-relative Frobenius error, None when the output is not finite (the metric of
    the reference's corpus evaluation, dataset/evaluate.py:17-21)."""
    if not np.isfinite(actual).all():
        return None
    error = float(np.linalg.norm(actual.astype(np.float64) - expected))
    return error / input_norm if input_norm else (0.0 if error == 0.0 else None)


CORPUS_SIZES = (8, 16, 32, 64, 128, 256)
_corpus = None


def corpus_cases(n=None):
    """The reference corpus (474 float32 square arrays, dataset/inputs.npz) as committed in
    tests/golden/corpus.npz: [(id, domain, array)] in manifest order, optionally only size
    n.  Every array's sha256 is checked against the reference manifest on first load."""
    global _corpus
    if _corpus is None:
        import hashlib
        import json
        gold = os.path.join(ROOT, 'tests', 'golden')
        manifest = json.load(open(os.path.join(gold, 'corpus_manifest.json')))
        cases = []
        with np.load(os.path.join(gold, 'corpus.npz'), allow_pickle=False) as data:
            assert set(data.files) == {c['id'] for c in manifest['cases']}
            for c in manifest['cases']:
                x = data[c['id']]
                assert x.shape == (c['N'], c['N']) and x.dtype == np.dtype('<f4')
                assert hashlib.sha256(x.tobytes(order='C')).hexdigest() == c['sha256'], c['id']
                cases.append((c['id'], c['domain'], np.ascontiguousarray(x)))
        _corpus = cases
    return [c for c in _corpus if n is None or c[2].shape[0] == n]


def families(h, w, rng):
    """The reference's shape-sweep families (tests/verify_portability.py:64-73)."""
    noise = rng.normal(size=(h, w)).astype(np.float32)
    alt = np.where(np.indices((h, w)).sum(axis=0) % 2, 1., -1.)
    impulse = np.zeros((h, w), np.float32)
    impulse[h // 3, w // 2] = 1
    return [('zero', np.zeros_like(impulse)), ('impulse', impulse), ('random', noise),
            ('neighbors', (1 + alt * 2 ** -23).astype(np.float32)),
            ('tiny', (noise * 2. ** -70).astype(np.float32)),
            ('large', (noise * 2. ** 50).astype(np.float32))]


# ---- 8x8 block codec step ----------------------------------------------------------------

# Luminance quantization table of ITU-T T.81 Annex K (Table K.1), row-major u * 8 + v.
JPEG_LUMA = np.array([
    16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55,
    14, 13, 16, 24, 40, 57, 69, 56, 14, 17, 22, 29, 51, 87, 80, 62,
    18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99], np.float32)


def _codec_lib():
    lib = oracle()
    if not getattr(lib, '_codec_typed', False):
        lib.oracle_codec8_encode.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_codec8_decode.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib._codec_typed = True
    return lib


def oracle_encode8(img, quant):
    """FP64 codec encode of a uint8 (H, W) image (rows may be strided): (coefficients
    (nblocks, 64) int16, unrounded quotients Y / Q (nblocks, 64))."""
    h, w = img.shape
    assert img.dtype == np.uint8 and img.strides[1] == 1
    q = np.ascontiguousarray(quant, np.float64).reshape(64)
    nb = (h // 8) * (w // 8)
    coeffs = np.empty((nb, 64), np.int16)
    quot = np.empty((nb, 64), np.float64)
    assert _codec_lib().oracle_codec8_encode(img.ctypes.data, h, w, img.strides[0], q.ctypes.data,
                                             quot.ctypes.data, coeffs.ctypes.data) == 0
    return coeffs, quot


def oracle_decode8(coeffs, h, w, quant):
    """FP64 codec decode: (uint8 image (H, W), unrounded X + 128 (H, W))."""
    c = np.ascontiguousarray(coeffs, np.int16)
    q = np.ascontiguousarray(quant, np.float64).reshape(64)
    img = np.empty((h, w), np.uint8)
    val = np.empty((h, w), np.float64)
    assert _codec_lib().oracle_codec8_decode(c.ctypes.data, h, w, w, q.ctypes.data, val.ctypes.data,
                                             img.ctypes.data) == 0
    return img, val


def rounding_agrees(got, want, unrounded, band):
    """Integer results from FP32 arithmetic against the FP64 oracle: equal wherever the
    oracle's unrounded value lies more than `band` from a rounding boundary (k + 1/2);
    within the band the FP32 value may round to the neighbour.  Returns the number of
    positions inside the band (all of which differ by at most 1)."""
    got = np.asarray(got, np.int64)
    want = np.asarray(want, np.int64)
    near = np.abs(np.abs(unrounded - np.floor(unrounded)) - 0.5) <= band
    bad = (got != want) & ~near
    assert not bad.any(), 'mismatch away from a rounding boundary at %s' % (np.argwhere(bad)[:5].tolist(),)
    assert (np.abs(got - want) <= 1).all()
    return int(near.sum())

"""The reference's deterministic adversarial corpus (474 float32 square arrays, sizes 8..256),
regenerated in test code so that GPU runs need neither the reference tree nor its archive.

Restates the families of the reference generator (dataset/corpus.py:14-60, corpus
'dct-adversarial-v1'): the arrays must be bit-identical to the published ones, so every
family keeps the reference's values, seeds and order of floating-point operations; the
sha256 of each array is checked against the reference manifest
(tests/golden/corpus_manifest.json, from dataset/manifest.json) by tests/test_corpus.py.
Test infrastructure only.
"""
import numpy as np

SEEDS = (910193, 2718281, 3141593, 1618033)       # dataset/corpus.py:7
SIZES = (8, 16, 32, 64, 128, 256)                 # dataset/corpus.py:8


def dct_matrix(n):
    """Orthonormal DCT-II matrix as the corpus defines it (dataset/corpus.py:10-12)."""
    k = np.arange(n, dtype=np.float64)
    return np.cos(np.pi * k[:, None] * (k[None, :] + .5) / n) * np.sqrt(np.where(k == 0, 1., 2.) / n)[:, None]


def corpus(n):
    """[(id, domain, array)] of size n, in manifest order."""
    out = []
    r, c = np.indices((n, n))
    checker = np.where((r + c) % 2, -1., 1.)

    def add(name, value, domain='core'):
        out.append(('n%d_%s' % (n, name), domain, np.broadcast_to(value, (n, n)).astype('<f4').copy()))

    add('zero', 0.)
    add('signed_zero', np.copysign(np.zeros((n, n)), checker))
    for e in (-100, -30, 0, 30, 90):
        a = np.float32(2. ** e)
        add('constant__e%d' % e, a)
        add('negative_constant__e%d' % e, -a)
        add('alternating__e%d' % e, a * checker)
        add('row_stripes__e%d' % e, a * np.where(r % 2, -1, 1))
    for e in (-70, 0, 70):
        a = np.float32(2. ** e)
        b = np.nextafter(a, np.float32(np.inf))
        add('adjacent_positive__e%d' % e, np.where((r + 2 * c) % 3, a, b))
        add('adjacent_negative__e%d' % e, -np.where((r + 2 * c) % 3, a, b))
        add('near_opposites__e%d' % e, np.where((r + c) % 2, -a, b))
        ulp = np.float64(b) - np.float64(a)
        add('offset_ulp__e%d' % e, np.float64(a) + ulp * ((3 * r + 5 * c) % 17 - 8))
    for rr, cc in ((0, 0), (n - 1, n - 1), (n // 2, n // 3), (1, n - 2)):
        x = np.zeros((n, n))
        x[rr, cc] = 1.
        add('impulse__r%d_c%d' % (rr, cc), x)
    x = np.zeros((n, n))
    x[1, 2] = 1.
    x[n - 2, n - 1] = -np.nextafter(np.float32(1.), np.float32(0.))
    add('sparse_cancellation', x)
    q = dct_matrix(n)
    for k, l in ((0, 1), (1, 0), (1, 2), (n // 2, n // 3), (n - 1, n - 2)):
        add('spatial_mode__k%d_l%d' % (k, l), np.outer(q[k], q[l]))
        x = np.zeros((n, n))
        x[k, l] = 1.
        add('coefficient_mode__k%d_l%d' % (k, l), x)
    add('asymmetric_ramp', (3 * r - 7 * c + .25 * r * c) / n)
    for seed in SEEDS:
        rng = np.random.Generator(np.random.PCG64(seed + n))
        z = rng.uniform(-1., 1., size=(n, n))
        for e in (-100, 0, 90):
            add('random_uniform__seed%d_e%d' % (seed, e), z * 2. ** e)
        add('random_normal__seed%d' % seed, rng.standard_normal((n, n)))
        add('dynamic_range__seed%d' % seed, np.ldexp(z, rng.integers(-90, 91, size=(n, n))))
    for e in (-149, -140, -126):
        add('tiny_alternating__e%d' % e, checker * 2. ** e, 'subnormal_diagnostic')
    for e in (110, 120, 126):
        add('boundary_constant__e%d' % e, 2. ** e, 'range_diagnostic')
        add('boundary_alternating__e%d' % e, checker * 2. ** e, 'range_diagnostic')
    return out


def all_cases():
    return [case for n in SIZES for case in corpus(n)]


def error_norm(actual, expected, input_norm):
    """Input-relative Frobenius error, None when the output is not finite
    (dataset/evaluate.py:17-21)."""
    if not np.isfinite(actual).all():
        return None
    error = float(np.linalg.norm(actual.astype(np.float64) - expected))
    return error / input_norm if input_norm else (0.0 if error == 0.0 else None)

"""Pins the FP64 oracle (oracle/bridct_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake (a dropped sqrt(2), a wrong sign or
index, a transposed operand, a swapped sine/cosine branch in s_{N,k}) fails:
  * per-axis scales against the reference's printed tables scale8..scale1024;
  * the closed form against SciPy's independent DCT-II/III (scipy.fft.dctn/idctn);
  * the C loop against the matrix form, orthogonality, impulse -> basis outer product,
    ones -> sqrt(hw) at DC, the scaled 1D core F_n against its definition;
  * the reference corpus summary: the oracle reproduces the published error figures;
  * oracle_supported by brute force over every (h, w) in [-8, 2048]^2 against the shape
    rule of src/plan.c:33-37 written out as a set;
  * oracle_norm_tables against the reference's baked norm_sj / corr_sj literals
    (src/generated/tables.inc:3-4,7-8, src/generated/large.c:13574-13581).
"""
import json
import os

import numpy as np
import pytest
import scipy.fft

import support as S

GOLD = os.path.join(S.ROOT, 'tests', 'golden')


def test_axis_scale_matches_reference_tables():
    data = json.load(open(os.path.join(GOLD, 'axis_scale.json')))['scale']
    lib = S.oracle()
    for n_s, vals in data.items():
        n = int(n_s)
        ours = np.array([lib.oracle_axis_scale(n, k) for k in range(n)])
        assert np.allclose(ours, vals, rtol=2e-15, atol=0), n


def test_sj_scale_special_values():
    lib = S.oracle()
    assert lib.oracle_sj_scale(4, 3) == 1.0
    assert abs(lib.oracle_sj_scale(8, 1) - np.cos(2 * np.pi / 8)) < 1e-16
    # k mod N/4 beyond N/8 takes the sine branch
    expected = np.sin(2 * np.pi * 5 / 32) * lib.oracle_sj_scale(8, 5 % 8)
    assert abs(lib.oracle_sj_scale(32, 5) - expected) < 1e-16


@pytest.mark.parametrize('h,w', [(8, 8), (16, 32), (64, 8), (128, 256), (1024, 16), (256, 1024)])
def test_closed_form_matches_scipy(h, w):
    rng = np.random.default_rng(h * 7 + w)
    x = rng.standard_normal((h, w))
    f = S.oracle_2d(x, 0)
    i = S.oracle_2d(x, 1)
    assert np.abs(f - scipy.fft.dctn(x, type=2, norm='ortho')).max() < 1e-12 * np.abs(f).max()
    assert np.abs(i - scipy.fft.idctn(x, type=2, norm='ortho')).max() < 1e-12 * np.abs(i).max()


@pytest.mark.parametrize('h,w', [(8, 8), (8, 16), (32, 8), (64, 128)])
def test_direct_loop_matches_matrix_form(h, w):
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (h, w))
    for mode in (0, 1, 2):
        a = S.oracle_2d_direct(x, mode)
        b = S.oracle_2d(x, mode)
        assert np.abs(a - b).max() < 1e-13


@pytest.mark.parametrize('n', S.LENGTHS)
def test_q_orthogonal_and_first_rows(n):
    q = S.qmatrix(n)
    assert np.abs(q @ q.T - np.eye(n)).max() < 1e-12
    assert np.allclose(q[0], 1 / np.sqrt(n), rtol=0, atol=1e-15)
    # row 1 is sqrt(2/n) cos(pi (j + 1/2) / n): positive then negative, antisymmetric
    assert q[1, 0] > 0 and np.allclose(q[1], -q[1, ::-1], atol=1e-15)


def test_impulse_is_basis_outer_product():
    h, w = 16, 32
    qh, qw = S.qmatrix(h), S.qmatrix(w)
    x = np.zeros((h, w))
    x[3, 5] = 1
    assert np.abs(S.oracle_2d(x, 0) - np.outer(qh[:, 3], qw[:, 5])).max() < 1e-15
    assert np.abs(S.oracle_2d(x, 1) - np.outer(qh[3], qw[5])).max() < 1e-15


def test_ones_to_dc():
    h, w = 32, 64
    y = S.oracle_2d(np.ones((h, w)), 0)
    exp = np.zeros((h, w))
    exp[0, 0] = np.sqrt(h * w)
    assert np.abs(y - exp).max() < 1e-12


@pytest.mark.parametrize('n', [8, 16, 64, 256])
def test_scaled_core_definition(n):
    """oracle_core computes F_n[k,j] = cos(pi k (2j+1)/(2n)) / s_{4n,k} and its transpose."""
    lib = S.oracle()
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    s = np.array([lib.oracle_sj_scale(4 * n, kk) for kk in range(n)])[:, None]
    F = np.cos(np.pi * k * (2 * j + 1) / (2 * n)) / s
    x = np.random.default_rng(n).standard_normal(n)
    y = np.empty(n)
    assert lib.oracle_core(n, 0, x.ctypes.data, y.ctypes.data) == 0
    assert np.abs(y - F @ x).max() < 1e-13 * np.abs(F @ x).max()
    assert lib.oracle_core(n, 1, x.ctypes.data, y.ctypes.data) == 0
    assert np.abs(y - F.T @ x).max() < 1e-13 * np.abs(F.T @ x).max()


def test_reference_corpus_summary_reproduced():
    """The committed summary was produced with the oracle's Q matrices and the reference
    CPU library; it reproduces the reference's published outcome (DATASET_CARD.md:34)."""
    s = json.load(open(os.path.join(GOLD, 'corpus_summary.json')))
    assert s['cases'] == 474 and s['directions'] == 1422 and s['core_directions'] == 1260
    assert s['boundary_fail'] == s['published_boundary_fail'] == 90
    assert abs(s['core_max_error'] - s['published_core_max_error']) < 0.01 * s['published_core_max_error']


def test_corpus_small_fixture_against_oracle():
    z = np.load(os.path.join(GOLD, 'corpus_small.npz'))
    for key, n in (('x', 8), ('x16', 16)):
        x = z[key]
        for mode in (0, 1, 2):
            y = z['y%d_%d' % (mode, n)]
            exact = S.oracle_2d(x.astype(np.float64), mode)
            for i in range(x.shape[0]):
                e = S.frob_input_rel(y[i], exact[i], x[i])
                assert e <= 2e-5, (n, mode, i, e)


def test_planted_mistakes_are_caught():
    """A transposed operand or a dropped sqrt(2) in Q would fail the SciPy comparison."""
    x = np.random.default_rng(9).standard_normal((8, 16))
    qh, qw = S.qmatrix(8), S.qmatrix(16)
    good = scipy.fft.dctn(x, type=2, norm='ortho')
    assert np.abs(qh.T @ x @ qw - good).max() > 1e-3          # transposed cores
    bad = qh.copy()
    bad[1:] /= np.sqrt(2)
    assert np.abs(bad @ x @ qw.T - good).max() > 1e-3         # dropped sqrt(2)


def test_supported_brute_force():
    """Every (h, w) in [-8, 2048]^2: supported exactly when both axes are in
    {8, 16, ..., 1024} (src/plan.c:33-37).  A plausible slip (n < 1024, a missing
    power-of-two test, 4 admitted) fails at one of the grid's points."""
    import ctypes
    lib = S.oracle()
    fn = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_int)(('oracle_supported', lib))
    allowed = {8, 16, 32, 64, 128, 256, 512, 1024}
    rng = range(-8, 2049)
    ok_h = [h in allowed for h in rng]
    count = 0
    for h, hok in zip(rng, ok_h):
        if hok:
            got = [fn(h, w) for w in rng]
            assert got == [1 if w in allowed else 0 for w in rng], h
            count += sum(got)
        else:
            assert not any(fn(h, w) for w in rng), h
    assert count == 64


def test_norm_tables_match_reference_literals():
    """oracle_norm_tables(h, w) against the reference's baked float tables: the
    normalization nu_h[r] nu_w[c] and its square, row-major, to within one float32
    rounding (the literals are the reference's own FP64 values rounded by the C
    compiler; the oracle rounds its FP64 product once)."""
    z = np.load(os.path.join(GOLD, 'norm_tables.npz'))
    lib = S.oracle()
    for n in (8, 16, 32, 64, 128, 256):
        norm = np.empty((n, n), np.float32)
        corr = np.empty((n, n), np.float32)
        assert lib.oracle_norm_tables(n, n, norm.ctypes.data, corr.ctypes.data) == 0
        rows = list(range(n)) if n <= 64 else sorted({0, 1, 2, 3, n // 2, n - 1})
        for name, ours in (('norm', norm), ('corr', corr)):
            ref = z['%s%d' % (name, n)]
            assert ref.shape == (len(rows), n)
            got = ours[rows]
            ulp = np.spacing(np.abs(ref))
            assert (np.abs(got - ref) <= ulp).all(), (name, n)
            assert np.mean(got == ref) > 0.95, (name, n)
    # rectangular plans: the outer product of the two axes' square-table factors
    h, w = 8, 32
    norm = np.empty((h, w), np.float32)
    corr = np.empty((h, w), np.float32)
    assert lib.oracle_norm_tables(h, w, norm.ctypes.data, corr.ctypes.data) == 0
    nu_h = np.sqrt(z['norm8'].astype(np.float64).diagonal())
    nu_w = np.sqrt(z['norm32'].astype(np.float64).diagonal())
    assert np.allclose(norm, np.outer(nu_h, nu_w), rtol=3e-7, atol=0)
    assert np.allclose(corr, np.outer(nu_h, nu_w) ** 2, rtol=5e-7, atol=0)
    # unsupported shapes are rejected and write nothing
    sentinel = np.full(4, 7.0, np.float32)
    assert lib.oracle_norm_tables(8, 12, sentinel.ctypes.data, sentinel.ctypes.data) == -1
    assert (sentinel == 7.0).all()


def test_norm_tables_planted_mistakes_fail():
    """The literal check rejects a transposed or unsquared table (shown on the CPU copy)."""
    z = np.load(os.path.join(GOLD, 'norm_tables.npz'))
    ref = z['norm16'].astype(np.float64)
    nu = np.sqrt(ref.diagonal())
    skew = np.outer(nu, nu * np.linspace(1, 1.001, 16))          # a column-dependent slip
    assert np.abs(skew - ref).max() > np.spacing(np.float32(ref.max()))
    assert np.abs(z['corr16'] - ref).max() > 1e-3                 # corr is the square, not norm

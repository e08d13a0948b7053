"""Pins the FP64 oracle (oracle/bridct_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake (a dropped sqrt(2), a wrong sign or
index, a transposed operand, a swapped sine/cosine branch in s_{N,k}) fails:
  * per-axis scales against the reference's printed tables scale8..scale1024;
  * the closed form against SciPy's independent DCT-II/III (scipy.fft.dctn/idctn);
  * the C loop against the matrix form, orthogonality, impulse -> basis outer product,
    ones -> sqrt(hw) at DC, the scaled 1D core F_n against its definition;
  * the reference corpus summary: the oracle reproduces the published error figures.
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

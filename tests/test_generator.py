"""The Shao-Johnson generator (bridct_amd/tools/gen_sj.py): every graph it emits is the
closed form it claims, the emitted headers are current, and its FP32 error is small."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'bridct_amd', 'tools'))
import gen_sj as G  # noqa: E402

import support as S  # noqa: E402


@pytest.mark.parametrize('n', [8, 16, 32, 64, 128, 256])
def test_sj_core_is_scaled_dct2(n):
    g, outs = G.dct2_sj(n)
    G.check('sj', g, outs, n, G.F_closed(n), tol=1e-11)


def test_sj_scale_same_as_oracle():
    lib = S.oracle()
    for N in (4, 8, 16, 32, 64, 128, 256, 4096):
        for k in range(0, N, max(1, N // 37)):
            assert G.sj_scale(N, k) == pytest.approx(lib.oracle_sj_scale(N, k), rel=1e-15, abs=0)


@pytest.mark.parametrize('N', [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize('sigma', [1, -1])
def test_dft_codelets(N, sigma):
    g, outs = G.dft_graph(N, sigma)
    G.check('dft', g, outs, 2 * N, G.dft_closed_real(N, sigma))


@pytest.mark.parametrize('kind', [1, 2, 4])
def test_scaled_kinds(kind):
    """S, S2, S4 outputs are the DFT divided by s_{N,k}, s_{2N,k}, s_{4N,k}."""
    N = 32
    g = G.Graph(2 * N)
    x = [(g.inp(j), g.inp(N + j)) for j in range(N)]
    Y = G.cfft(g, x, kind, 1)
    outs = [y[0] for y in Y] + [y[1] for y in Y]
    M = G.dft_closed_real(N, 1)
    scale = np.array([G.kind_scale(kind, N, k) for k in range(N)] * 2)[:, None]
    G.check('kind', g, outs, 2 * N, M / scale)


def test_transpose_is_transpose():
    g, outs = G.dct2_sj(32)
    raw = [(1.0, h[1]) for h in outs]
    gt, ot = G.transpose(g, raw, 32)
    A = G.graph_matrix(g, raw, 32)
    B = G.graph_matrix(gt, ot, 32)
    assert np.abs(A.T - B).max() < 1e-12


def test_instruction_counts_at_most_reference_adds():
    """Each add/sub/FMA is one instruction; the reference's 1D cores spend 26, 72, 186, 456
    additions alone for n = 8..64 (its counting shim, plus 8, 24, 66, 166 multiplies)."""
    ref_adds = {8: 26, 16: 72, 32: 186, 64: 456}
    for n, adds in ref_adds.items():
        g, outs = G.dct2_sj(n)
        raw = [(1.0, h[1]) for h in outs]
        assert g.counts(raw)['instructions'] <= adds


def test_fp32_error_small():
    for n in (8, 16, 32, 64):
        g, outs = G.dct2_sj(n)
        assert G.fp32_error(g, outs, n, G.F_closed(n)) < 2.5e-7


def test_emitted_headers_current():
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([sys.executable, os.path.join(ROOT, 'bridct_amd', 'tools', 'gen_sj.py'),
                        '--out', d], check=True, capture_output=True)
        for name in ('sj_codelets.h', 'sj_factors.h'):
            fresh = open(os.path.join(d, name)).read()
            tree = open(os.path.join(ROOT, 'bridct_amd', 'csrc', 'generated', name)).read()
            assert fresh == tree, name


def test_planted_mistake_fails_check():
    """A wrong twiddle sign in the split radix is caught by the closed-form check."""
    g, outs = G.dft_graph(8, 1)
    M = G.dft_closed_real(8, -1)
    with pytest.raises(SystemExit):
        G.check('planted', g, outs, 16, M)

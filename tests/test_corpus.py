"""CPU checks of the corpus harness: the committed corpus arrays are the published ones
(sha256 of every array against the reference manifest, the rule of dataset/verify.py:8-28),
and the error metric detects the reference's 78 seeded faults (dataset/evaluate.py:79-87)."""
import hashlib
import json
import os

import numpy as np

import support as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
THRESHOLD = 2e-5


def test_committed_corpus_is_published_corpus():
    manifest = json.load(open(os.path.join(GOLD, 'corpus_manifest.json')))
    cases = S.corpus_cases()
    assert len(cases) == len(manifest['cases']) == 474
    for (cid, domain, x), m in zip(cases, manifest['cases']):
        assert cid == m['id'] and domain == m['domain'] and x.shape == (m['N'], m['N'])
        assert x.dtype == np.dtype('<f4') and np.isfinite(x).all()
        assert hashlib.sha256(x.tobytes(order='C')).hexdigest() == m['sha256'], cid


def test_seeded_faults_detected():
    """For every size, mode and fault (scale by 1.01, transpose, sign, zero; lost AC for
    the forward), at least one core case must be flagged by the metric."""
    detected = {}
    for n in S.CORPUS_SIZES:
        q = S.qmatrix(n)
        for cid, domain, x in S.corpus_cases(n):
            if domain != 'core':
                continue
            xd = x.astype(np.float64)
            norm = float(np.linalg.norm(xd))
            for mode, expected in enumerate((q @ xd @ q.T, q.T @ xd @ q, xd)):
                faults = {'scale': expected * 1.01, 'transpose': expected.T, 'sign': -expected,
                          'zero': np.zeros_like(expected)}
                if mode == 0:
                    lost = np.zeros_like(expected)
                    lost[0, 0] = expected[0, 0]
                    faults['lost_ac'] = lost
                for name, changed in faults.items():
                    e = S.error_norm(changed, expected, norm)
                    key = (n, mode, name)
                    detected[key] = detected.get(key, 0) + int(e is None or e >= THRESHOLD)
    assert len(detected) == 78
    assert all(v > 0 for v in detected.values()), [k for k, v in detected.items() if not v]

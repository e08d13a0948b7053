#!/usr/bin/env python3
"""Writes the committed golden fixtures of the 2D DCT plan path from the reference tree.

Run in the build container (the reference tree is not present where GPU tests run):
    python tests/golden/make_golden.py [--reference /root/reference]

It reads only data and literal tables from the reference, and runs only the reference
library compiled from its own sources by oracle/Makefile (oracle/_ref/libbridct.so):

* axis_scale.json: the per-axis orthonormal scales scale8..scale1024 printed in the
  reference's generated tables (src/generated/rect_generated.inc:115, 350, 834, 1695,
  2878, 4629, 6929, 9858).  They equal s_{4n,k} sqrt((k ? 2 : 1)/n) and pin the oracle's
  sj_scale / axis_scale.
* corpus_small.npz: the 8x8 and 16x16 core cases of the reference corpus
  (dataset/inputs.npz, sha256-checked against dataset/manifest.json) with the reference
  CPU outputs for FORWARD, INVERSE and ROUNDTRIP.
* corpus_summary.json: for all 474 corpus inputs and 3 modes, the reference CPU output's
  input-relative Frobenius error against the FP64 Q-matrix closed form
  (dataset/evaluate.py:17-21), and the reference's published maximum over core cases
  (dataset/DATASET_CARD.md:34) for comparison.
* corpus_manifest.json: id, size, domain and sha256 of every corpus array
  (dataset/manifest.json).
* corpus.npz: the 474 corpus arrays themselves, copied from dataset/inputs.npz after the
  archive's and every array's sha256 were checked against dataset/manifest.json (the rule
  of dataset/verify.py:8-28).  Tests check each array's sha256 again on load.
* norm_tables.npz: the reference's baked normalization and correction tables, read as text
  from src/generated/tables.inc:3-4,7-8 (norm_sj8/16, corr_sj8/16) and
  src/generated/large.c:13574-13581 (norm_sj32..256, corr_sj32..256), each literal rounded
  to float32 as the C compiler rounds it.  Complete for n <= 64; rows 0, 1, 2, 3, n/2 and
  n-1 of the larger tables.  They pin oracle_norm_tables.
"""
import argparse
import ctypes
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, 'tests'))


def parse_scales(ref):
    text = open(os.path.join(ref, 'src/generated/rect_generated.inc')).read()
    out = {}
    for m in re.finditer(r'static const float scale(\d+)\[\]=\{([^}]*)\}', text):
        n = int(m.group(1))
        vals = [float(v.rstrip('f')) for v in m.group(2).split(',')]
        assert len(vals) == n
        out[str(n)] = vals
    assert sorted(map(int, out)) == [8, 16, 32, 64, 128, 256, 512, 1024]
    return out


NORM_ROWS = lambda n: sorted({0, 1, 2, 3, n // 2, n - 1})


def parse_norm_tables(ref):
    """{'norm8': float32 (n, n) or the sampled rows, ...} from the reference literals."""
    gen = os.path.join(ref, 'src/generated')
    texts = [open(os.path.join(gen, 'tables.inc')).read(), open(os.path.join(gen, 'large.c')).read()]
    out = {}
    for kind in ('norm', 'corr'):
        for n in (8, 16, 32, 64, 128, 256):
            found = [m for t in texts for m in re.finditer(r'static const float %s_sj%d\[\]=\{([^}]*)\}' % (kind, n), t)]
            assert len(found) == 1, (kind, n)
            vals = np.array([np.float32(v.strip().rstrip('f')) for v in found[0].group(1).split(',')], np.float32)
            assert vals.size == n * n, (kind, n, vals.size)
            table = vals.reshape(n, n)
            if n > 64:
                table = table[NORM_ROWS(n)]
            out['%s%d' % (kind, n)] = table
    return out


def error_norm(actual, expected, input_norm):
    """dataset/evaluate.py:17-21 restated."""
    if not np.isfinite(actual).all():
        return None
    error = float(np.linalg.norm(actual.astype(np.float64) - expected))
    return error / input_norm if input_norm else (0.0 if error == 0.0 else None)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--reference', default='/root/reference')
    args = ap.parse_args()
    import support as S
    if not S.ref_available():
        raise SystemExit('build oracle/_ref first: make -C oracle')

    scales = parse_scales(args.reference)
    with open(os.path.join(HERE, 'axis_scale.json'), 'w') as f:
        json.dump({'source': 'src/generated/rect_generated.inc scale{n}[]', 'scale': scales}, f)
    np.savez_compressed(os.path.join(HERE, 'norm_tables.npz'), **parse_norm_tables(args.reference))

    ds = os.path.join(args.reference, 'dataset')
    manifest = json.load(open(os.path.join(ds, 'manifest.json')))
    archive = os.path.join(ds, 'inputs.npz')
    assert hashlib.sha256(open(archive, 'rb').read()).hexdigest() == manifest['archive_sha256']
    rows = []
    small = {}
    arrays = {}
    with np.load(archive, allow_pickle=False) as data:
        assert set(data.files) == {c['id'] for c in manifest['cases']}
        for case in manifest['cases']:
            x = data[case['id']]
            assert x.shape == (case['N'], case['N']) and x.dtype == np.dtype('<f4') and np.isfinite(x).all()
            assert hashlib.sha256(x.tobytes(order='C')).hexdigest() == case['sha256']
            arrays[case['id']] = x
            n = case['N']
            q = S.qmatrix(n)
            xd = x.astype(np.float64)
            norm = float(np.linalg.norm(xd))
            outs = []
            for mode, exact in enumerate((q @ xd @ q.T, q.T @ xd @ q, xd)):
                with np.errstate(all='ignore'):
                    y = S.ref_apply(x, mode)
                    e = error_norm(y, exact, norm)
                outs.append(y)
                rows.append(dict(id=case['id'], n=n, domain=case['domain'], mode=mode, error=e))
            if n <= 16 and case['domain'] == 'core':
                small[case['id']] = (x, outs)
    core = [r['error'] for r in rows if r['domain'] == 'core' and r['error'] is not None]
    summary = dict(
        cases=len(manifest['cases']), directions=len(rows),
        core_directions=sum(r['domain'] == 'core' for r in rows),
        core_max_error=max(core),
        published_core_max_error=2.250789485770909e-7,
        boundary_fail=sum(1 for r in rows if r['domain'] != 'core' and
                          (r['error'] is None or r['error'] > 2e-5)),
        published_boundary_fail=90,
        rows=rows)
    with open(os.path.join(HERE, 'corpus_summary.json'), 'w') as f:
        json.dump(summary, f, indent=0)
    with open(os.path.join(HERE, 'corpus_manifest.json'), 'w') as f:
        json.dump({'version': manifest['version'], 'archive_sha256': manifest['archive_sha256'],
                   'cases': [dict(id=c['id'], N=c['N'], domain=c['domain'], sha256=c['sha256'])
                             for c in manifest['cases']]}, f, indent=0)
    np.savez_compressed(os.path.join(HERE, 'corpus.npz'), **arrays)
    ids = sorted(small)
    np.savez_compressed(
        os.path.join(HERE, 'corpus_small.npz'),
        ids=np.array(ids),
        x=np.stack([small[i][0] for i in ids if small[i][0].shape[0] == 8]),
        x16=np.stack([small[i][0] for i in ids if small[i][0].shape[0] == 16]),
        **{'y%d_8' % m: np.stack([small[i][1][m] for i in ids if small[i][0].shape[0] == 8])
           for m in range(3)},
        **{'y%d_16' % m: np.stack([small[i][1][m] for i in ids if small[i][0].shape[0] == 16])
           for m in range(3)})
    print('core max error %.6e (published %.6e), boundary failures %d (published 90)'
          % (summary['core_max_error'], 2.250789485770909e-7, summary['boundary_fail']))


if __name__ == '__main__':
    main()

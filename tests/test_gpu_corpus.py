"""The reference's adversarial corpus (474 arrays x 3 directions, committed as data in
tests/golden/corpus.npz) through the GPU plan path
with the reference's evaluation rules (dataset/evaluate.py:55-78): return code, input
unchanged, guard sentinels intact, input-relative Frobenius error < 2e-5.  Gate (the
"no CPU pass turns into a GPU failure" rule of the reference's numerical comparisons):
every direction the reference CPU library passes (tests/golden/corpus_summary.json) must
pass on the GPU; all core directions pass; the GPU's core maximum stays within twice the
reference's.  Writes the per-direction results as JSON to the run's output folder."""
import json
import os

import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
THRESHOLD = 2e-5


def test_corpus_gate():
    L = capi.load()
    cpu = {(r['id'], r['mode']): r['error'] for r in json.load(open(os.path.join(GOLD, 'corpus_summary.json')))['rows']}
    rows = []
    for n in S.CORPUS_SIZES:
        print('corpus size', n, flush=True)
        q = S.qmatrix(n)
        plan = L.bridct_plan_create(n, n)
        assert plan
        try:
            for cid, domain, x in S.corpus_cases(n):
                original = x.view(np.uint32).copy()
                xd = x.astype(np.float64)
                norm = float(np.linalg.norm(xd))
                guard = np.full(n * n + 8, 531.25, np.float32)
                y = guard[4:-4].reshape(n, n)
                for mode, expected in enumerate((q @ xd @ q.T, q.T @ xd @ q, xd)):
                    y.fill(np.nan)
                    rc = L.bridct_plan_apply(plan, mode, capi.ptr(x), capi.ptr(y))
                    with np.errstate(all='ignore'):
                        err = S.error_norm(y, expected, norm)
                    ok = bool(rc == 0 and np.array_equal(x.view(np.uint32), original) and
                              np.all(guard[:4] == 531.25) and np.all(guard[-4:] == 531.25) and
                              err is not None and err < THRESHOLD)
                    ref_err = cpu[(cid, mode)]
                    rows.append(dict(id=cid, n=n, domain=domain, mode=mode, gpu_error=err, gpu_pass=ok,
                                     cpu_error=ref_err, cpu_pass=ref_err is not None and ref_err < THRESHOLD))
        finally:
            L.bridct_plan_destroy(plan)
    core = [r for r in rows if r['domain'] == 'core']
    boundary = [r for r in rows if r['domain'] != 'core']
    summary = dict(
        directions=len(rows), core_directions=len(core),
        core_failures=sum(not r['gpu_pass'] for r in core),
        gpu_core_max_error=max(r['gpu_error'] for r in core),
        cpu_core_max_error=max(r['cpu_error'] for r in core),
        boundary_failures_gpu=sum(not r['gpu_pass'] for r in boundary),
        boundary_failures_cpu=sum(not r['cpu_pass'] for r in boundary),
        cpu_pass_gpu_fail=[(r['id'], r['mode']) for r in rows if r['cpu_pass'] and not r['gpu_pass']],
        gpu_pass_cpu_fail=sum(1 for r in rows if r['gpu_pass'] and not r['cpu_pass']))
    os.makedirs(S.OUT_DIR, exist_ok=True)
    with open(os.path.join(S.OUT_DIR, 'corpus_gpu.json'), 'w') as f:
        json.dump(dict(summary=summary, rows=rows), f, indent=0)
    print(json.dumps(summary), flush=True)
    assert len(rows) == 1422 and summary['core_directions'] == 1260
    assert not summary['cpu_pass_gpu_fail']
    assert summary['core_failures'] == 0
    assert summary['gpu_core_max_error'] <= 2 * summary['cpu_core_max_error']

"""CPU stress test of the host copy pool of the pinned pipeline
(bridct_amd/csrc/copy_pool.c): the C driver tests/c/copy_pool_stress.c is built with gcc
against the pool's source and run under a time limit.  It covers many pools used right
after creation (helpers that first take the lock after the caller has started a round
must still take part in it, or the round never ends) and random job sizes checked byte
for byte (the pieces of a job must cover it whole)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, '..', 'bridct_amd', 'csrc')


@pytest.fixture(scope='module')
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp('pool') / 'copy_pool_stress')
    subprocess.run(['gcc', '-O2', '-pthread', '-Wall', '-Werror', '-I', CSRC,
                    os.path.join(HERE, 'c', 'copy_pool_stress.c'), os.path.join(CSRC, 'copy_pool.c'), '-o', exe],
                   check=True)
    return exe


@pytest.mark.parametrize('pools,rounds,maxbytes', [(2000, 3, 65536), (1, 200, 4 << 20), (300, 40, 1027),
                                                   (50, 20, 0)])
def test_copy_pool_stress(driver, pools, rounds, maxbytes):
    r = subprocess.run([driver, str(pools), str(rounds), str(maxbytes)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith('ok ')

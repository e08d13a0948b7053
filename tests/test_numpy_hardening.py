"""Hardening of the NumPy objects Plan, RaggedPlan and Codec8: the lifecycle and re-entry
contract the reference checks for its Plan in python/test_hardening.py:37-148.

  * a NEW object (tp_new only) refuses calls, close() on it is a no-op, and it can still be
    initialized; an initialized or closed one refuses __init__;
  * a failed __init__ (bad argument kind, unsupported shape) leaves the object NEW, so a
    later __init__ may succeed;
  * Python code run inside __init__ (an argument's __index__) or inside a call (the
    output allocation, through a NumPy allocator hook) that tries to close, re-initialize
    or call the same object gets RuntimeError, and the outer call completes unharmed;
  * a failed output allocation raises MemoryError and leaves the object usable; so do late
    argument errors (out of the wrong type, shape or overlapping);
  * the module refuses to load in a subinterpreter (ImportError naming the main
    interpreter), whether or not the main interpreter imported it first.

The allocator hook is the repository's own test module tests/c/np_alloc_hook.c, built here
with the host C compiler.  Tests marked gpu need a GPU to open objects; the others run
anywhere (without a GPU, __init__ ends in RuntimeError after its arguments are read).
"""
import importlib.util
import os
import subprocess
import sys
import sysconfig

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope='module')
def hook(tmp_path_factory):
    out = str(tmp_path_factory.mktemp('hook') / ('np_alloc_hook' + sysconfig.get_config_var('EXT_SUFFIX')))
    subprocess.run(['gcc', '-std=c11', '-O1', '-fPIC', '-shared', '-Wall', '-Wextra', '-Werror',
                    '-I', sysconfig.get_paths()['include'], '-I', np.get_include(),
                    os.path.join(HERE, 'c', 'np_alloc_hook.c'), '-o', out], check=True)
    spec = importlib.util.spec_from_file_location('np_alloc_hook', out)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def refused(label, error, call, log):
    with pytest.raises(error):
        call()
    log.append(label)


# ---- per-type fixtures: constructor arguments, a valid input, one call -----------------

def plan_case(npmod):
    x = np.linspace(-1, 1, 64, dtype=np.float32).reshape(8, 8)
    return npmod.Plan, (8, 8), x, lambda obj, **kw: obj.forward(x, **kw)


def ragged_case(npmod):
    xs = [np.linspace(-1, 1, 64, dtype=np.float32).reshape(8, 8),
          np.linspace(1, -1, 128, dtype=np.float32).reshape(8, 16)]
    return npmod.RaggedPlan, ([(8, 8), (8, 16)],), xs, lambda obj, **kw: obj.forward(xs, **kw)


def codec_case(npmod):
    img = (np.arange(256, dtype=np.uint32).reshape(16, 16) * 7 % 256).astype(np.uint8)
    quant = np.full(64, 2.0, np.float32)
    return npmod.Codec8, (quant,), img, lambda obj, **kw: obj.encode(img, **kw)


CASES = [plan_case, ragged_case, codec_case]


def gpu_ready():
    from bridct_amd import capi
    return capi.load().bridct_gpu_available() == 1


# ---- lifecycle and __init__ re-entry (any machine) -------------------------------------

@pytest.mark.parametrize('case', CASES, ids=['Plan', 'RaggedPlan', 'Codec8'])
def test_new_object_and_init_reentry(npmod, case):
    cls, args, _, call = case(npmod)
    log = []
    fresh = cls.__new__(cls)
    refused('new-call', RuntimeError, lambda: call(fresh), log)
    fresh.close()                       # nothing to release: stays NEW
    fresh.close()
    assert fresh.closed
    refused('bad-kind', TypeError, lambda: fresh.__init__('8', 8) if cls is npmod.Plan else fresh.__init__(3.5),
            log)
    if cls is npmod.Plan:
        refused('bad-shape', ValueError, lambda: fresh.__init__(26, 8), log)
    elif cls is npmod.RaggedPlan:
        refused('bad-shape', ValueError, lambda: fresh.__init__([(8, 8), (26, 8)]), log)
    else:
        refused('bad-steps', ValueError, lambda: fresh.__init__(np.zeros(64, np.float32)), log)
    assert fresh.closed and log
    if cls is npmod.Codec8:
        return                      # its constructor reads no integer: nothing runs Python inside it
    for action in ('close', 'init', 'call'):
        obj = cls.__new__(cls)
        seen = []

        class Eight:
            """An axis length whose __index__ touches the object under construction."""

            def __index__(self):
                attempt = {'close': obj.close, 'init': lambda: obj.__init__(*args),
                           'call': lambda: call(obj)}[action]
                refused('index-reentry-' + action, RuntimeError, attempt, seen)
                return 8

        ctor = (Eight(), 8) if cls is npmod.Plan else ([(Eight(), 8), (8, 16)],)
        if gpu_ready():
            obj.__init__(*ctor)
            assert not obj.closed
            call(obj)
            obj.close()
        else:
            with pytest.raises(RuntimeError, match='No usable gfx950 GPU'):
                obj.__init__(*ctor)
            assert obj.closed
        assert seen == ['index-reentry-' + action], seen


def test_subinterpreter_import_refused():
    """Import inside a subinterpreter fails with ImportError naming the main interpreter
    (python/test_hardening.py:111-145), with and without a prior import in the main one."""
    code = r'''
import sys
import _xxsubinterpreters as interpreters
sys.path.insert(0, ROOT)
MAIN_FIRST
sub = interpreters.create()
try:
    try:
        interpreters.run_string(sub, "import sys; sys.path.insert(0, " + repr(ROOT) + "); from bridct_amd import bridct_numpy")
        raise SystemExit("the subinterpreter import was accepted")
    except interpreters.RunFailedError as failure:
        text = str(failure)
    assert "ImportError" in text and "main interpreter" in text, text
    print("refused:", text)
finally:
    interpreters.destroy(sub)
'''
    for first in (False, True):
        src = code.replace('ROOT', repr(ROOT)).replace(
            'MAIN_FIRST', 'from bridct_amd import bridct_numpy' if first else '')
        r = subprocess.run([sys.executable, '-c', src], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert 'refused:' in r.stdout


def test_alloc_hook_itself(hook):
    """The trap fires once, on the first array-data allocation, and fails it on request."""
    calls = []
    a = hook.intercept(lambda: (np.empty(10), np.empty(10)), lambda: calls.append(1))
    assert calls == [1] and a[0].shape == (10,)
    with pytest.raises(MemoryError):
        hook.intercept(lambda: np.empty(1000), lambda: calls.append(2), True)
    assert calls == [1, 2]
    np.empty(1000)                      # the previous handler is back
    with pytest.raises(MemoryError):    # a callback that raises fails the allocation too
        hook.intercept(lambda: np.empty(1000), lambda: 1 / 0)
    np.empty(1000)


# ---- re-entry and failure inside calls (GPU) -------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize('case', CASES, ids=['Plan', 'RaggedPlan', 'Codec8'])
def test_reinit_and_close_rules(npmod, case):
    cls, args, _, call = case(npmod)
    obj = cls.__new__(cls)
    obj.__init__(*args)
    assert not obj.closed
    with pytest.raises(RuntimeError):
        obj.__init__(*args)              # initialized: never again
    first = call(obj)
    obj.close()
    assert obj.closed
    with pytest.raises(RuntimeError):
        obj.__init__(*args)              # closed: never again
    with pytest.raises(RuntimeError):
        call(obj)
    fresh = cls(*args)
    again = call(fresh)
    fresh.close()
    if isinstance(first, list):
        assert all(np.array_equal(a, b) for a, b in zip(first, again))
    else:
        assert np.array_equal(first, again)


@pytest.mark.gpu
@pytest.mark.parametrize('case', CASES, ids=['Plan', 'RaggedPlan', 'Codec8'])
def test_allocator_reentry_and_failure(npmod, hook, case):
    cls, args, _, call = case(npmod)
    obj = cls(*args)
    want = call(obj)

    def same(got):
        if isinstance(want, list):
            return all(np.array_equal(a, b) for a, b in zip(want, got))
        return np.array_equal(want, got)

    log = []
    for action in ('close', 'call', 'init'):
        def inside(action=action):
            attempt = {'close': obj.close, 'call': lambda: call(obj), 'init': lambda: obj.__init__(*args)}[action]
            refused('alloc-reentry-' + action, RuntimeError, attempt, log)

        got = hook.intercept(lambda: call(obj), inside)
        assert same(got), action
        assert not obj.closed
    assert log == ['alloc-reentry-close', 'alloc-reentry-call', 'alloc-reentry-init']

    fired = []
    with pytest.raises(MemoryError):
        hook.intercept(lambda: call(obj), lambda: fired.append(1), True)
    assert fired == [1] and not obj.closed
    assert same(call(obj))                               # usable after the failure

    # late argument errors leave the object usable
    for error, kw in ((TypeError, {'out': 3}), (TypeError, {'bogus': 1})):
        with pytest.raises(error):
            call(obj, **kw)
        assert same(call(obj))
    if cls is npmod.Plan:
        x = case(npmod)[2]
        for error, bad in ((ValueError, lambda: obj.forward(x, out=np.empty((1, 8, 8), np.float32))),
                           (ValueError, lambda: obj.forward(x, out=x)),
                           (TypeError, lambda: obj.forward(x.astype(np.float64)))):
            with pytest.raises(error):
                bad()
            assert same(call(obj))
    obj.close()
    obj.close()
    assert obj.closed


@pytest.mark.gpu
def test_plan_numerics_after_hardening(npmod):
    """python/test_hardening.py:95-107: after the lifecycle exercises, fresh plans still give
    2e-5 input-relative Frobenius accuracy, owning outputs, roundtrip == inverse(forward)."""
    from scipy import fft
    for shape in ((8, 8), (16, 32), (128, 128), (512, 1024)):
        plan = npmod.Plan(*shape)
        x = np.random.default_rng(sum(shape)).standard_normal(shape).astype(np.float32)
        for mode, method in enumerate((plan.forward, plan.inverse, plan.roundtrip)):
            want = x.astype(float) if mode == 2 else (fft.dctn if mode == 0 else fft.idctn)(x.astype(float),
                                                                                           norm='ortho')
            got = method(x)
            err = float(np.linalg.norm(got.astype(float) - want) / np.linalg.norm(x.astype(float)))
            assert err < 2e-5
            assert got.flags.owndata and not np.shares_memory(x, got)
            if mode == 2:
                assert np.array_equal(got, plan.inverse(plan.forward(x)))
        plan.close()

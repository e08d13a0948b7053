"""CPU checks of the C ABI: the library loads, exports every symbol the public headers
declare, and the host-side rules that need no GPU call behave like the reference
(src/plan.c:33-37, :117-135; src/bridct.c:13-45; docs/API.md)."""
import ctypes
import os
import re

import numpy as np
import pytest

from bridct_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for hdr in ('bridct.h', 'bridct_gpu.h'):
        text = open(os.path.join(ROOT, 'include', hdr)).read()
        names |= set(re.findall(r'BRIDCT_API[^;(]*?\b(bridct_\w+)\s*\(', text))
    return names


def test_headers_declare_reference_api():
    names = declared_symbols()
    for ref_name in ('bridct_supported', 'bridct_plan_create', 'bridct_plan_destroy',
                     'bridct_plan_route', 'bridct_plan_apply', 'bridct_workspace_floats',
                     'bridct_apply'):
        assert ref_name in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(capi.SIGNATURES) >= declared_symbols()


def test_version_and_enums():
    text = open(os.path.join(ROOT, 'include', 'bridct.h')).read()
    assert '#define BRIDCT_VERSION "0.2.0-dev.7"' in text
    for decl in ('BRIDCT_FORWARD = 0', 'BRIDCT_INVERSE = 1', 'BRIDCT_ROUNDTRIP = 2', 'BRIDCT_OK = 0',
                 'BRIDCT_INVALID_ARGUMENT = -1', 'BRIDCT_WORKSPACE_TOO_SMALL = -2',
                 'BRIDCT_BAD_ALIGNMENT = -3', 'BRIDCT_OVERLAPPING_BUFFERS = -4'):
        assert decl in text


def test_supported_truth_table():
    L = capi.load()
    good = [8, 16, 32, 64, 128, 256, 512, 1024]
    for h in [-8, 0, 4, 7, 8, 12, 26, 16, 1024, 2048, 1023]:
        for w in [8, 26, 1024, 2048]:
            assert L.bridct_supported(h, w) == int(h in good and w in good)


def test_route_null_and_destroy_null():
    L = capi.load()
    assert L.bridct_plan_route(None) is None
    L.bridct_plan_destroy(None)


def test_workspace_floats_match_reference_formula():
    L = capi.load()
    for n in (8, 16, 32, 64, 128, 256):
        e = n * n
        exp = 2 * e if n <= 16 else 8 * e + 2 * 1100 * 4 + 4 * 4 * 256 + 4 * n * (4 * n + 4)
        assert L.bridct_workspace_floats(n) == exp
    for n in (0, 4, 12, 512, 1024):
        assert L.bridct_workspace_floats(n) == 0


def test_legacy_apply_error_precedence_without_writes():
    """src/bridct.c:32-45: -1 before -2 before -3 before -4, output untouched
    (tests/test_small_optimized.c:41-93)."""
    L = capi.load()
    n = 8
    need = L.bridct_workspace_floats(n)
    x = np.ones((n, n), np.float32)
    y = np.full((n, n), 5.0, np.float32)
    raw = np.zeros(need + 8, np.float32)
    off = (-raw.ctypes.data % 16) // 4
    ws = raw[off:off + need + 4]
    p = capi.ptr
    assert L.bridct_apply(7, 0, p(x), p(y), p(ws), need) == -1
    assert L.bridct_apply(n, 3, p(x), p(y), p(ws), need) == -1
    assert L.bridct_apply(n, 0, None, p(y), p(ws), need) == -1
    assert L.bridct_apply(n, 0, p(x), p(y), None, need) == -1
    assert L.bridct_apply(n, 0, p(x), p(y), p(ws), need - 1) == -2
    mis = ctypes.c_void_p(ws.ctypes.data + 4)
    assert L.bridct_apply(n, 0, p(x), p(y), mis, need) == -3
    assert L.bridct_apply(n, 0, p(x), p(x), p(ws), need) == -4
    big = np.zeros(need + 128, np.float32)
    off = (-big.ctypes.data % 16) // 4
    wsb = big[off:]
    inside = ctypes.c_void_p(wsb.ctypes.data + 64)
    assert L.bridct_apply(n, 0, p(x), inside, p(wsb), need) == -4
    assert np.all(y == 5.0)


@pytest.mark.skipif(capi.load().bridct_gpu_available() == 1, reason='a GPU is present')
def test_no_gpu_fails_loudly():
    """Without a usable gfx950 device there is no plan (no CPU fallback)."""
    L = capi.load()
    assert L.bridct_plan_create(8, 8) is None
    assert L.bridct_gpu_malloc(16) is None
    n = 8
    need = L.bridct_workspace_floats(n)
    x = np.ones((n, n), np.float32)
    y = np.zeros((n, n), np.float32)
    raw = np.zeros(need + 8, np.float32)
    off = (-raw.ctypes.data % 16) // 4
    ws = raw[off:off + need]
    assert L.bridct_apply(n, 0, capi.ptr(x), capi.ptr(y), capi.ptr(ws), need) == capi.DEVICE_ERROR


def test_numpy_extension_loads(npmod):
    assert npmod.core_version == '0.2.0-dev.7'
    assert hasattr(npmod, 'Plan')
    for shape in ((4, 8), (26, 32), (8, 2048), (0, 8), (-8, 8)):
        with pytest.raises(ValueError):
            npmod.Plan(*shape)
    with pytest.raises(TypeError):
        npmod.Plan(8)
    with pytest.raises(TypeError):
        npmod.Plan('8', 8)
    u = npmod.Plan.__new__(npmod.Plan)
    with pytest.raises(RuntimeError):
        u.forward(np.zeros((8, 8), np.float32))
    u.close()
    assert u.closed and u.route is None

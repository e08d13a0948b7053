"""The schedule table format read at plan creation (bridct_amd/csrc/schedule.c): parsing on
the CPU, and on the GPU that a table sets a new plan's route and packing.  The reference
compiles its per-shape policy in (src/rect_policy.inc:1-46, small_policy.json:14-21)."""
import ctypes
import os

import numpy as np
import pytest

import support as S
from bridct_amd import autotune, capi


def lookup(path, h, w):
    lib = capi.load()
    fn = lib.bdx_schedule_lookup_file
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                   ctypes.POINTER(ctypes.c_int)]
    r, p = ctypes.c_int(-7), ctypes.c_int(-7)
    rc = fn(str(path).encode(), h, w, ctypes.byref(r), ctypes.byref(p))
    return rc, r.value, p.value


def test_parse_table(tmp_path):
    t = tmp_path / 'sched.txt'
    t.write_text('# comment line\n'
                 '8 8 butterfly 64   # trailing comment\n'
                 '\n'
                 '256 256 mfma_f16 0\n'
                 '64 64 mfma_bf16 0\n'
                 '64 64 butterfly 4\n'          # the last matching line wins
                 '32 32 warp_drive 1\n'         # unknown route: ignored
                 '16 16 butterfly -3\n'         # bad pack: ignored
                 '128 128 butterfly -1\n'       # the two-pass kernels
                 'garbage\n')
    assert lookup(t, 8, 8) == (0, capi.ROUTE_BUTTERFLY, 64)
    assert lookup(t, 256, 256) == (0, capi.ROUTE_MFMA_F16, 0)
    assert lookup(t, 64, 64) == (0, capi.ROUTE_BUTTERFLY, 4)
    assert lookup(t, 32, 32)[0] == -1
    assert lookup(t, 16, 16)[0] == -1
    assert lookup(t, 128, 128) == (0, capi.ROUTE_BUTTERFLY, capi.PACK_TWO_PASS)
    assert lookup(t, 1024, 8)[0] == -1
    assert lookup(tmp_path / 'missing.txt', 8, 8)[0] == -1


def test_committed_table_is_complete():
    """The committed table (written by the autotuner on an MI355X) covers all 64 shapes with
    a valid route; dense routes only where both axes are at least 64; the two-pass choice
    only for whole-array shapes that the row and column kernels take."""
    path = os.path.join(os.path.dirname(capi.LIB_PATH), 'schedule_gfx950.txt')
    if not os.path.exists(path):
        pytest.skip('no schedule table committed yet')
    for h in (8, 16, 32, 64, 128, 256, 512, 1024):
        for w in (8, 16, 32, 64, 128, 256, 512, 1024):
            rc, route, pack = lookup(path, h, w)
            assert rc == 0, (h, w)
            assert route == capi.ROUTE_BUTTERFLY or (h >= 64 and w >= 64)
            assert 0 <= pack <= 4096 or (pack == capi.PACK_TWO_PASS and autotune.two_pass_ok(h, w)
                                         and h * (w + 1) <= autotune.SMALL_LDS_FLOATS), (h, w, pack)


def test_two_pass_eligibility():
    """The shapes the two-pass choice is offered for: the whole-array shapes whose row bands
    and column strips are whole (for lines up to 64, bands of 256 rows and strips of 256
    columns)."""
    small = [(h, w) for h in (8, 16, 32, 64, 128, 256, 512, 1024) for w in (8, 16, 32, 64, 128, 256, 512, 1024)
             if h * (w + 1) <= autotune.SMALL_LDS_FLOATS]
    eligible = sorted(s for s in small if autotune.two_pass_ok(*s))
    assert eligible == [(16, 1024), (32, 256), (32, 512), (32, 1024), (64, 256), (64, 512), (128, 128), (128, 256),
                        (256, 64), (256, 128), (512, 64), (1024, 32)]
    # every shape too large for the whole-array kernel is one the two-pass kernels take
    assert all(autotune.two_pass_ok(h, w) for h in (8, 16, 32, 64, 128, 256, 512, 1024)
               for w in (8, 16, 32, 64, 128, 256, 512, 1024) if h * (w + 1) > autotune.SMALL_LDS_FLOATS)


@pytest.mark.gpu
def test_table_sets_plan_route_and_pack(tmp_path, monkeypatch):
    t = tmp_path / 'sched.txt'
    t.write_text('256 256 mfma_f16 0\n16 16 butterfly 8\n32 32 mfma_f16 0\n128 128 butterfly -1\n'
                 '64 64 butterfly -1\n')
    monkeypatch.setenv('BRIDCT_SCHEDULE', str(t))
    p = capi.Plan(256, 256)
    q = capi.Plan(16, 16)
    r = capi.Plan(32, 32)      # dense route impossible below 64: stays butterfly
    u = capi.Plan(128, 128)
    v = capi.Plan(64, 64)      # two-pass kernels do not take 64 x 64: automatic packing
    try:
        assert p.route == capi.ROUTE_MFMA_F16 and p.pack == 0
        assert q.route == capi.ROUTE_BUTTERFLY and q.pack == 8
        assert r.route == capi.ROUTE_BUTTERFLY
        assert u.route == capi.ROUTE_BUTTERFLY and u.pack == capi.PACK_TWO_PASS
        assert v.route == capi.ROUTE_BUTTERFLY and v.pack == 0
    finally:
        for x in (p, q, r, u, v):
            x.close()
    monkeypatch.setenv('BRIDCT_SCHEDULE', 'none')
    p = capi.Plan(256, 256)
    try:
        assert p.route == capi.ROUTE_BUTTERFLY and p.pack == 0
    finally:
        p.close()


@pytest.mark.gpu
@pytest.mark.parametrize('h,w', [(8, 8), (16, 64), (64, 32), (128, 128)])
def test_packing_does_not_change_results(h, w):
    x = np.random.default_rng(h + w).uniform(-1, 1, (300, h, w)).astype(np.float32)
    p = capi.Plan(h, w)
    try:
        outs = []
        for pack in (0, 1, 3, 64, 4096):
            assert p.set_pack(pack) == 0
            y = np.empty_like(x)
            assert p.apply_batch(capi.FORWARD, x, y) == 0
            outs.append(y)
        for y in outs[1:]:
            assert np.array_equal(y, outs[0])
        assert p.set_pack(-2) == capi.INVALID_ARGUMENT and p.set_pack(4097) == capi.INVALID_ARGUMENT
        assert (p.set_pack(capi.PACK_TWO_PASS) == 0) == autotune.two_pass_ok(h, w)
    finally:
        p.close()


@pytest.mark.gpu
@pytest.mark.parametrize('h,w', [(16, 1024), (32, 256), (32, 512), (32, 1024), (64, 256), (64, 512), (128, 128),
                                 (128, 256), (256, 64), (256, 128), (512, 64), (1024, 32)])
def test_two_pass_choice_on_whole_array_shapes(h, w):
    """A whole-array shape sent to the row and column kernels (pack -1) meets the FP64
    oracle in every mode, as a plain call and through the per-pass entry point (both passes
    in order), and its pass-split call equals its plain call bitwise."""
    print('two-pass %dx%d' % (h, w), flush=True)
    x = np.random.default_rng(h * 3 + w).uniform(-1, 1, (5, h, w)).astype(np.float32)
    p = capi.Plan(h, w)
    a, b = capi.DeviceBuffer(x.nbytes), capi.DeviceBuffer(x.nbytes)
    try:
        assert p.set_pack(capi.PACK_TWO_PASS) == 0 and p.pack == capi.PACK_TWO_PASS
        for mode in (capi.FORWARD, capi.INVERSE, capi.ROUNDTRIP):
            y = np.empty_like(x)
            assert p.apply_batch(mode, x, y) == 0
            assert S.max_rel(y, S.oracle_2d(x, mode)) < 1e-6, mode
        a.upload(x)
        for mode in (capi.FORWARD, capi.INVERSE):
            assert p.apply_device_pass(mode, 0, a.ptr, b.ptr, 5) == 0
            assert p.apply_device_pass(mode, 1, b.ptr, b.ptr, 5) == 0
            assert p.synchronize() == 0
            y, z = np.empty_like(x), np.empty_like(x)
            b.download(y)
            assert p.apply_batch(mode, x, z) == 0
            assert np.array_equal(y, z), mode
    finally:
        a.free()
        b.free()
        p.close()

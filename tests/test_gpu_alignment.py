"""GPU tests of host-pointer alignment: the host C API needs only float alignment for input
and output (docs/API.md:17), and 16-byte alignment only for the legacy workspace.

Follows the reference's layout sweep (tests/test_small16.c:50-109): six (input, output,
workspace) byte residues modulo 64 -- (0,0,0) (16,48,48) (48,16,48) (32,16,0) (4,12,16)
(60,4,32) -- for the legacy 16x16 bridct_apply with impulses 0, 1, 85, 255 in all three
modes: closed-form accuracy, input unchanged, guard floats around every buffer intact,
output bitwise equal to the plan API on aligned buffers; then the error statuses with the
same layouts (short workspace, bad mode, overlap, misaligned workspace) write nothing.
The plan API's single and batched host calls take 4-, 8- and 12-byte residues too,
bitwise equal to 16-byte-aligned calls, on shapes of both routes and through the
pipelined path of large batches."""
import numpy as np
import pytest

import support as S
from bridct_amd import capi

pytestmark = pytest.mark.gpu

TRIPLES = [(0, 0, 0), (16, 48, 48), (48, 16, 48), (32, 16, 0), (4, 12, 16), (60, 4, 32)]
GUARD = 32          # guard floats on each side
SENTINEL = np.float32(12345.0)


class Placed:
    """`count` floats at byte residue `residue` modulo 64, with guard floats on both sides."""

    def __init__(self, count, residue):
        assert residue % 4 == 0
        self.raw = np.empty(count * 4 + 2 * GUARD * 4 + 128, np.uint8)
        base = self.raw.ctypes.data
        start = (-(base + GUARD * 4) % 64 + residue) % 64 + GUARD * 4
        self.all = self.raw[start - GUARD * 4:start + count * 4 + GUARD * 4].view(np.float32)
        self.data = self.all[GUARD:GUARD + count]
        assert self.data.ctypes.data % 64 == residue
        self.all[:] = SENTINEL

    def guards_intact(self):
        return bool((self.all[:GUARD] == SENTINEL).all() and (self.all[-GUARD:] == SENTINEL).all())


def test_legacy_16x16_layout_sweep():
    lib = capi.load()
    n = 16
    need = lib.bridct_workspace_floats(n)
    assert need == 512
    q = S.qmatrix(n)
    plan = capi.Plan(n, n)
    try:
        for tx, ty, tw in TRIPLES:
            x, y, w = Placed(n * n, tx), Placed(n * n, ty), Placed(need, tw)
            for impulse in (0, 1, 85, 255):
                for mode in (0, 1, 2):
                    x.data[:] = 0
                    x.data[impulse] = 1
                    y.data[:] = SENTINEL
                    w.data[:] = SENTINEL
                    assert lib.bridct_apply(n, mode, capi.ptr(x.data), capi.ptr(y.data), capi.ptr(w.data), need) == 0
                    r, c = divmod(impulse, n)
                    want = (np.outer(q[:, r], q[:, c]) if mode == 0 else np.outer(q[r], q[c]) if mode == 1
                            else x.data.reshape(n, n).astype(np.float64))
                    err = float(np.sqrt(((y.data.reshape(n, n) - want) ** 2).sum()))
                    assert err <= 2e-5, (tx, ty, tw, impulse, mode, err)
                    assert (x.data == (np.arange(n * n) == impulse)).all()
                    assert x.guards_intact() and y.guards_intact() and w.guards_intact()
                    aligned = np.empty((n, n), np.float32)
                    assert plan.apply(mode, np.ascontiguousarray(x.data.reshape(n, n)), aligned) == 0
                    assert np.array_equal(aligned.ravel(), y.data), (tx, ty, tw, impulse, mode)
            # error statuses, same layouts: nothing written
            y.data[:] = SENTINEL
            w.data[:] = SENTINEL
            xp, yp, wp = capi.ptr(x.data), capi.ptr(y.data), capi.ptr(w.data)
            for mode in (0, 1):
                assert lib.bridct_apply(n, mode, xp, yp, wp, need - 1) == capi.WORKSPACE_TOO_SMALL
                assert lib.bridct_apply(n, mode, xp, xp, wp, need) == capi.OVERLAPPING_BUFFERS
                assert lib.bridct_apply(n, mode, xp, yp, capi.ptr(w.data[1:]), need) == capi.BAD_ALIGNMENT
            assert lib.bridct_apply(n, 3, xp, yp, wp, need) == capi.INVALID_ARGUMENT
            assert (y.data == SENTINEL).all() and (w.data == SENTINEL).all()
            assert x.guards_intact() and y.guards_intact() and w.guards_intact()
    finally:
        plan.close()


@pytest.mark.parametrize('h,w,batch', [(8, 8, 3), (16, 32, 2), (256, 256, 2), (1024, 64, 2), (128, 128, 64)])
def test_plan_host_calls_any_float_alignment(h, w, batch):
    """bridct_plan_apply (one array) and bridct_plan_apply_batch (the last case, 1 Mi floats,
    takes the pinned pipeline) with input and output at residues 4, 8, 12 (mod 16):
    bitwise equal to 16-byte-aligned buffers, guards intact, input unchanged."""
    lib = capi.load()
    rng = np.random.default_rng(h + w)
    src = rng.uniform(-1, 1, (batch, h, w)).astype(np.float32)
    plan = capi.Plan(h, w)
    try:
        for mode in (0, 1, 2):
            want = np.empty_like(src)
            assert plan.apply_batch(mode, src, want) == 0
            for rx, ry in ((4, 8), (8, 12), (12, 4), (4, 4)):
                x, y = Placed(src.size, rx), Placed(src.size, ry)
                x.data[:] = src.ravel()
                if batch == 1 or h * w * batch < (1 << 18):
                    for i in range(batch):
                        sl = slice(i * h * w, (i + 1) * h * w)
                        assert lib.bridct_plan_apply(plan.ptr, mode, capi.ptr(x.data[sl]), capi.ptr(y.data[sl])) == 0
                else:
                    assert lib.bridct_plan_apply_batch(plan.ptr, mode, capi.ptr(x.data), capi.ptr(y.data),
                                                       batch) == 0
                assert np.array_equal(y.data, want.ravel()), (mode, rx, ry)
                assert np.array_equal(x.data, src.ravel())
                assert x.guards_intact() and y.guards_intact()
    finally:
        plan.close()

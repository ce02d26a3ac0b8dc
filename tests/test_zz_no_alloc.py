"""ws_select_noalloc (the selection entry point that allocates no device memory: its descriptor and
look-back state live in a per-GPU device array) against ws_select, on two calls in a row with the
second input smaller than the first (the array keeps the first call's state), and against numpy."""
import ctypes as C

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 100, n).astype(np.int32)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    v[rng.random(n) < 0.05] = abi.NULL_I64
    deleted = rng.random(n) < 0.1
    words = np.zeros((n + 31) // 32, np.uint32)
    idx = np.nonzero(deleted)[0]
    np.bitwise_or.at(words, idx >> 5, (np.uint32(1) << (idx & 31).astype(np.uint32)))
    # WHERE k < 37 AND v > -500 (NULL v fails)
    sel = (k < 37) & (v != abi.NULL_I64) & (v > -500) & ~deleted
    return k, v, words, sel


def _select(gpu, fn, k, v, words, n):
    dk, dv, dw = ws.DevBuf.from_array(k), ws.DevBuf.from_array(v), ws.DevBuf.from_array(words)
    out = ws.DevBuf(max(n, 1) * 4)
    b = abi.EnvBuilder()
    sk, sv = b.slot(abi.T_I32), b.slot(abi.T_I64N)
    b.env.cols[sk].p, b.env.cols[sv].p = dk.p, dv.p
    where = b.prog([(abi.OP_COL, sk), b.ilit(37), (abi.OP_LT,), (abi.OP_COL, sv), b.ilit(-500), (abi.OP_GT,), (abi.OP_AND,)])
    cnt = C.c_int64()
    if fn == "ws_select":
        ws.check(gpu.ws_select(C.byref(b.env), 1, where, n, dw.p, out.p, None, C.byref(cnt)))
    else:
        ws.check(gpu.ws_select_noalloc(C.byref(b.env), 1, where, n, dw.p, out.p, C.byref(cnt)))
    return out.to_array(np.int32, cnt.value)


@pytest.mark.parametrize("sizes", [(6_000_000, 150_000), (300_000, 70_001)])
def test_select_noalloc_matches_select(gpu, sizes):
    for i, n in enumerate(sizes):  # the second call is smaller than the first
        H.say(f"case select_noalloc n={n} (call {i + 1} of 2)")
        k, v, words, sel = _case(n, seed=100 + i)
        want = np.nonzero(sel)[0].astype(np.int32)
        a = _select(gpu, "ws_select", k, v, words, n)
        b = _select(gpu, "ws_select_noalloc", k, v, words, n)
        assert np.array_equal(a, want)
        assert np.array_equal(b, want), (n, len(b), len(want))


def test_select_noalloc_rejects_oversized_input(gpu):
    H.say("case select_noalloc size limit")
    b = abi.EnvBuilder()
    s = b.slot(abi.T_I32)
    cnt = C.c_int64(-1)
    rc = gpu.ws_select_noalloc(C.byref(b.env), 0, b.col(s), (1 << 28) + 1, None, None, C.byref(cnt))
    assert rc == ws.WS_ERR_INTERNAL and cnt.value == 0
    assert b"exceed the maximum" in gpu.ws_last_error()

"""Learned sizing on the GPU: the MLP's forward, backward and Adam kernels against the CPU oracle
(oracle/mlp_oracle.py), then training on exact group counts logged by real aggregations and a
held-out comparison with the range-based sizing rule it replaces (engine_select.cpp:1154-1235)."""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

import helpers as H
import warpstore_amd as ws
from warpstore_amd import abi

pytestmark = pytest.mark.gpu


def _mlp():
    spec = importlib.util.spec_from_file_location("mlp_oracle", os.path.join(H.ROOT, "oracle", "mlp_oracle.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _sizer(L, seed):
    h = C.c_void_p()
    ws.check(L.ws_sizer_create(seed, C.byref(h)))
    return h


def _params(L, h):
    p = np.zeros(1633, np.float32)
    ws.check(L.ws_sizer_get_params(h, ws.ptr(p)))
    return p


def test_mlp_kernels_match_oracle(gpu):
    """Forward outputs, the gradient of the MSE loss and one Adam step, GPU vs oracle (fp32
    kernels against a float64 restatement: 1e-5 absolute on outputs, 1e-4 relative on gradients)."""
    H.say("case sizer kernels vs oracle")
    L, M = gpu, _mlp()
    h = _sizer(L, 7)
    try:
        p = _params(L, h)
        rng = np.random.default_rng(1)
        B = 300
        x = rng.random((B, 16)).astype(np.float32)
        t = rng.random(B).astype(np.float32)
        y = np.zeros(B, np.float32)
        ws.check(L.ws_sizer_predict(h, ws.ptr(x), B, ws.ptr(y)))
        yo = M.forward(p.astype(np.float64), x.astype(np.float64))[0]
        np.testing.assert_allclose(y, yo, atol=1e-5, rtol=0)
        loss = np.zeros(1, np.float32)
        g = np.zeros(1633, np.float32)
        ws.check(L.ws_sizer_loss_grad(h, ws.ptr(x), ws.ptr(t), B, ws.ptr(loss), ws.ptr(g)))
        go = M.grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
        assert abs(loss[0] - M.loss(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))) < 1e-5
        np.testing.assert_allclose(g, go, rtol=1e-4, atol=1e-6)
        # one Adam step over the whole batch (one minibatch)
        ws.check(L.ws_sizer_train(h, ws.ptr(x), ws.ptr(t), B, 1, B, 1e-3, 5, None))
        q = _params(L, h)
        qo, _, _ = M.adam(p.astype(np.float64), go, np.zeros(1633), np.zeros(1633), 1, 1e-3)
        np.testing.assert_allclose(q, qo, atol=2e-6, rtol=0)
    finally:
        L.ws_sizer_destroy(h)


# ------------------------------------------------------------------ training on logged counts

def _task(rng):
    """A random filter + GROUP BY shape: uniform or skewed integer keys, some stored sparse (a large
    base and stride, so their span overstates their distinct count), and a selective filter."""
    n = int(2 ** rng.uniform(14, 20))
    nk = int(rng.integers(1, 4))
    cols = {}
    for k in range(nk):
        d = max(1, int(10 ** rng.uniform(0.3, 5.5)))
        v = rng.integers(0, d, n) if rng.random() < 0.5 else (d * rng.random(n) ** 3).astype(np.int64)
        if rng.random() < 0.4:
            v = v * int(10 ** rng.uniform(1, 4)) + int(rng.integers(0, 10 ** 9))
        cols[f"k{k}"] = v.astype(np.int64)
    cols["val"] = rng.integers(0, 1_000_000, n).astype(np.int64)
    theta = int(rng.uniform(0.05, 1.0) * 1_000_000)
    return cols, nk, theta


def _run_task(L, cols, nk, theta):
    """Exact groups (logged with the features) and the range rule's estimate of one task."""
    t = H.Table(cols, 1 << 18)
    slots = [(f"k{k}", abi.T_I64, 0) for k in range(nk)] + [("val", abi.T_I64, 0)]
    dev = H.Bound(t, slots, True)
    stats = (abi.ColStats * (nk + 1))()
    for s in range(nk + 1):
        for b in range(len(t.sizes)):
            st = abi.ColStats()
            ws.check(L.ws_column_stats(C.byref(dev.views[b * len(slots) + s]), t.sizes[b], None, C.byref(st)))
            if b == 0:
                stats[s] = st
            else:
                L.ws_col_stats_merge(C.byref(stats[s]), C.byref(st))
    vs = stats[nk]
    sel = min(1.0, max(0.0, (theta - vs.imin) / max(1, vs.imax - vs.imin + 1)))
    f = np.zeros(16, np.float32)
    ws.check(L.ws_sizing_features(C.cast(stats, C.c_void_p), nk, None, len(cols["val"]), sel, ws.ptr(f)))
    b = abi.EnvBuilder()
    ks = [b.slot(abi.T_I64) for _ in range(nk)]
    sv = b.slot(abi.T_I64)
    plan = abi.AggPlan()
    plan.has_where = 1
    plan.where = b.prog([(abi.OP_COL, sv), b.ilit(theta), (abi.OP_LT,)])
    plan.nkeys = nk
    for k in range(nk):
        plan.keys[k] = b.col(ks[k])
        plan.key_int[k] = 1
    plan.naggs = 1
    plan.ops[0] = abi.AGG_COUNT
    plan.agg_int[0] = 1
    n = len(cols["val"])
    cap = 1
    while cap < 2 * n:
        cap <<= 1
    gk, _, _, _ = H.gpu_aggregate(plan, b.env, dev, cap=cap, features=f, strategy=1)
    span = 1.0
    for k in range(nk):
        span *= float(stats[k].imax - stats[k].imin + 1)
    plain = min(span, float(n))  # the range rule: key ranges, capped by the live rows
    return len(gk), plain, f


def _cap(est):
    c = 1024
    while c < 2 * est:
        c <<= 1
    return c


def test_learned_sizing_beats_range_rule(gpu):
    """Trains the sizing MLP on the GPU on exact counts logged by 320 aggregations, then on 80
    held-out shapes compares its capacity choice with the range rule: both end with exact results
    (an under-estimate is retried bigger, as the engine does on overflow), but the learned choice
    must allocate far fewer table slots at a small retry cost, with a smaller log2 error."""
    H.say("case learned sizing: training and held-out comparison")
    L = gpu
    L.ws_sizing_log_clear()
    rng = np.random.default_rng(2024)
    truth, plain = [], []
    for i in range(400):
        cols, nk, theta = _task(rng)
        g, pl, _ = _run_task(L, cols, nk, theta)
        truth.append(g)
        plain.append(pl)
    n = L.ws_sizing_log_size()
    assert n == 400, n
    X = np.zeros((n, 16), np.float32)
    T = np.zeros(n, np.float32)
    assert L.ws_sizing_log_copy(ws.ptr(X), ws.ptr(T), n) == n
    np.testing.assert_allclose(T, np.log2(np.array(truth, np.float64) + 1) / 32, rtol=1e-6)
    tr, te = slice(0, 320), slice(320, 400)
    h = _sizer(L, 11)
    try:
        loss = np.zeros(1, np.float32)
        Xtr, Ttr = np.ascontiguousarray(X[tr]), np.ascontiguousarray(T[tr])
        ws.check(L.ws_sizer_train(h, ws.ptr(Xtr), ws.ptr(Ttr), 320, 400, 32, 3e-3, 9, ws.ptr(loss)))
        Xte = np.ascontiguousarray(X[te])
        y = np.zeros(80, np.float32)
        ws.check(L.ws_sizer_predict(h, ws.ptr(Xte), 80, ws.ptr(y)))
    finally:
        L.ws_sizer_destroy(h)
    g_te = np.array(truth[320:], np.float64)
    pred = np.maximum(1.0, 2.0 ** (32.0 * y.astype(np.float64)) - 1)
    pl_te = np.array(plain[320:], np.float64)
    err_learned = np.median(np.abs(np.log2(pred + 1) - np.log2(g_te + 1)))
    err_plain = np.median(np.abs(np.log2(pl_te + 1) - np.log2(g_te + 1)))

    def cost(est_list, margin):
        """Table slots allocated per task (retries included) and the number of retries."""
        slots, retries = [], 0
        for est, g in zip(est_list, g_te):
            c = _cap(est * margin)
            total = 0
            while g * 10 > c * 7:  # overflow at load 0.7: retry at 4x (engine_select.cpp:1514-1522)
                total += c
                retries += 1
                c = _cap(max(4 * c, 4 * g) / 2)
            slots.append(total + c)
        return np.array(slots, np.float64), retries

    s_plain, r_plain = cost(pl_te, 1.0)
    s_learned, r_learned = cost(pred, 1.5)
    gm = float(np.exp(np.mean(np.log(s_plain / s_learned))))  # geometric mean of per-task savings
    over = float(np.mean(s_plain > 4 * s_learned))
    H.say(f"train loss {loss[0]:.5f}; median |log2 err| learned {err_learned:.3f} vs range rule {err_plain:.3f}; "
          f"slots learned {int(s_learned.sum())} ({r_learned} retries) vs range rule {int(s_plain.sum())} ({r_plain} "
          f"retries); per-task saving geometric mean {gm:.2f}x, range rule >4x bigger on {100 * over:.0f}% of tasks")
    assert loss[0] < 0.002
    assert err_learned < 0.5 and err_learned < err_plain / 2
    assert gm > 1.2
    assert r_learned <= 8

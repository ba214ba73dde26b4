"""Learned plan selector on the GPU: the HIP inference, gradient and Adam
kernels against the float64 restatement (oracle/mlp.py) and the host
replay; training on the GPU; librettPlan following the loaded selector with
bit-exact results."""
import numpy as np
import pytest

import librett_amd as L
import oracle
from librett_amd import data, hip
from librett_amd import selector as S
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dev_array(a):
    b = hip.DeviceBuffer(max(4, a.nbytes))
    hip.htod(b.ptr, np.ascontiguousarray(a))
    return b


def fetch(b, dtype, n):
    out = np.zeros(n, dtype=dtype)
    hip.dtoh(out, b.ptr)
    return out


def params(seed):
    p = S.init_params(seed).astype(np.float64)
    p[M.H * M.F:M.H * M.F + M.H] = np.random.default_rng(seed).normal(0, 0.1, M.H)
    return p.astype(np.float32)


def test_gpu_forward(gpu_stream):
    say("case: mlp forward kernel, 1000 samples")
    rng = np.random.default_rng(1)
    p = params(1)
    x = rng.normal(size=(1000, M.F)).astype(np.float32)
    dp, dx, do = dev_array(p), dev_array(x), hip.DeviceBuffer(4000)
    L.mlp_forward(dp.ptr, dx.ptr, 1000, do.ptr, gpu_stream)
    hip.stream_sync(gpu_stream)
    got = fetch(do, np.float32, 1000)
    ref = M.forward(p, x)
    assert np.all(np.abs(got - ref) <= 1e-4 * (1 + np.abs(ref)))
    host = L.mlp_host_forward(p, x)
    assert np.all(np.abs(got - host) <= 1e-5 * (1 + np.abs(host)))
    assert np.array_equal(np.argmin(got.reshape(40, 25), 1), np.argmin(ref.reshape(40, 25), 1))


def test_gpu_grad_several_workgroups(gpu_stream):
    say("case: mlp gradient kernel, 4 workgroups, ragged minibatch of 450")
    rng = np.random.default_rng(2)
    p = params(2)
    n = 1000
    x = rng.normal(size=(n, M.F)).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    idx = rng.permutation(n).astype(np.int32)
    base, end, nb = 100, 550, 4
    dp, dx, dy, di = dev_array(p), dev_array(x), dev_array(y), dev_array(idx)
    dpart, dloss = hip.DeviceBuffer(nb * M.P * 4), hip.DeviceBuffer(nb * 4)
    sel = idx[base:end]
    w = (rng.random(n) + 0.1).astype(np.float32)
    dw = dev_array(w)
    for weighted in (False, True):
        inv_m = 1.0 / (float(w[sel].astype(np.float64).sum()) if weighted else (end - base))
        L.mlp_grad(dp.ptr, dx.ptr, dy.ptr, di.ptr, base, end, inv_m, nb, dpart.ptr, dloss.ptr, gpu_stream,
                   d_w=dw.ptr if weighted else None)
        hip.stream_sync(gpu_stream)
        part = fetch(dpart, np.float32, nb * M.P).reshape(nb, M.P)
        loss = fetch(dloss, np.float32, nb)
        gr, sse = M.grad(p, x[sel], y[sel], inv_m, w[sel] if weighted else None)
        g = part.astype(np.float64).sum(0)
        assert np.all(np.abs(g - gr) <= 1e-4 * np.abs(gr).max() + 1e-7)
        assert abs(loss.sum() - sse) <= 1e-4 * (1 + sse)
        # The last workgroup's share (samples 384..449 of the minibatch) equals the host replay.
        hg, _ = L.mlp_host_grad(p, x, y, sel[384:], inv_m, w=w if weighted else None)
        assert np.all(np.abs(part[3] - hg) <= 1e-5 * np.abs(hg).max() + 1e-8)


def test_gpu_adam(gpu_stream):
    say("case: mlp Adam kernel")
    rng = np.random.default_rng(3)
    parts = rng.normal(size=(5, M.P)).astype(np.float32)
    p = rng.normal(size=M.P).astype(np.float32)
    m = (rng.normal(size=M.P) * 0.1).astype(np.float32)
    v = (rng.random(M.P) * 0.1).astype(np.float32)
    t = 3
    # Hyper-parameters as the kernel sees them (float32), so the float64
    # restatement uses the same values.
    lr, b1, b2 = (float(np.float32(v_)) for v_ in (1e-3, 0.9, 0.999))
    c1, c2 = 1 / (1 - b1 ** t), 1 / (1 - b2 ** t)
    dg, dp, dm, dv = dev_array(parts), dev_array(p), dev_array(m), dev_array(v)
    L.mlp_adam(dg.ptr, 5, dp.ptr, dm.ptr, dv.ptr, lr, b1, b2, c1, c2, 1e-8, gpu_stream)
    hip.stream_sync(gpu_stream)
    gp, gm, gv = fetch(dp, np.float32, M.P), fetch(dm, np.float32, M.P), fetch(dv, np.float32, M.P)
    # Same per-thread code on the host: equal up to fused multiply-adds.
    hp, hm, hv = p.copy(), m.copy(), v.copy()
    L.mlp_host_adam(parts, hp, hm, hv, lr, b1, b2, c1, c2, 1e-8)
    assert np.allclose(gm, hm, rtol=1e-5, atol=1e-7) and np.allclose(gv, hv, rtol=1e-5, atol=1e-7)
    assert np.allclose(gp, hp, rtol=1e-5, atol=1e-6)
    pr, mr, vr = M.adam(p, m, v, parts.astype(np.float64).sum(0), lr, b1, b2, t, 1e-8)
    assert np.allclose(gm, mr, rtol=1e-5, atol=1e-7) and np.allclose(gv, vr, rtol=1e-5, atol=1e-7)
    assert np.allclose(gp, pr, rtol=1e-5, atol=1e-6)


def test_gpu_training_fits_a_teacher(gpu_stream):
    say("case: mlp training on 20000 samples of a random teacher network")
    rng = np.random.default_rng(4)
    teacher = params(40)
    x = rng.normal(size=(20000, M.F)).astype(np.float32)
    y = M.forward(teacher, x).astype(np.float32)
    p = S.init_params(7)
    dx, dy, dp = dev_array(x), dev_array(y), dev_array(p)
    losses = L.mlp_train(dx.ptr, dy.ptr, len(y), dp.ptr, 30, 2e-3, 1024, 5, gpu_stream)
    say("  mse by epoch", losses[0], losses[-1], "target var", float(y.var()))
    assert np.all(np.isfinite(losses))
    # 30 epochs from a He initialisation explain most of the variance.
    assert losses[-1] < 0.3 * y.var() and losses[-1] < 0.2 * losses[0]
    # Weighted training (weights on the device and the host) converges too.
    w = (np.random.default_rng(6).random(len(y)) + 0.1).astype(np.float32)
    dw, dq = dev_array(w), dev_array(S.init_params(8))
    lw = L.mlp_train(dx.ptr, dy.ptr, len(y), dq.ptr, 30, 2e-3, 1024, 5, gpu_stream, d_w=dw.ptr, w_host=w)
    say("  weighted mse by epoch", lw[0], lw[-1])
    assert np.all(np.isfinite(lw)) and lw[-1] < 0.2 * lw[0]
    q = fetch(dp, np.float32, M.P)
    out = hip.DeviceBuffer(len(y) * 4)
    L.mlp_forward(dp.ptr, dx.ptr, len(y), out.ptr, gpu_stream)
    hip.stream_sync(gpu_stream)
    got = fetch(out, np.float32, len(y))
    mse = float(np.mean((M.forward(q, x) - y) ** 2))
    assert abs(float(np.mean((got - y) ** 2)) - mse) <= 1e-3 * (1 + mse)


@pytest.mark.parametrize("dim,perm,elem", [([64, 48, 32], [2, 0, 1], 8), ([2, 8, 16, 32, 64, 8], [5, 3, 1, 0, 2, 4], 4),
                                           ([300, 200, 9], [1, 2, 0], 16)])
def test_plan_follows_selector(gpu_stream, dim, perm, elem):
    say("case: librettPlan with the selector", dim, perm, elem)
    feats, model_choice = L.selector_features(dim, perm, elem)
    # A network whose score is feature k plus a constant: the selector must
    # pick the candidate with the smallest feature k.
    k = int(np.argmax(feats.std(0)))
    W1 = np.zeros((M.H, M.F)); W1[0, k] = 1
    b1 = np.zeros(M.H); b1[0] = 1000
    W2 = np.zeros((M.H, M.H)); W2[0, 0] = 1
    w3 = np.zeros(M.H); w3[0] = 1
    p = M.pack(W1, b1, W2, np.zeros(M.H), w3, 0.0).astype(np.float32)
    assert L.selector_set(p, np.zeros(M.F), np.ones(M.F)) == L.LIBRETT_SUCCESS
    try:
        want = int(np.argmin(feats[:, k]))
        # librettPlan keeps the analytic choice unless the selector's score
        # for it is higher than its own pick's by more than the margin.
        if feats[model_choice, k] - feats[want, k] < L.selector_margin():
            want = model_choice
        got = L.selector_choose(dim, perm, elem)
        assert got == want, (got, want, k)
        r, h = L.librettPlan(len(dim), dim, perm, elem, gpu_stream)
        assert r == L.LIBRETT_SUCCESS
        infos = L.candidates(dim, perm, elem)
        assert L.plan_info(h)["desc"] == infos[want]["desc"]
        n = int(np.prod(dim))
        x = data.uniform(n, elem, 17)
        a, b = hip.DeviceBuffer(n * elem), hip.DeviceBuffer(n * elem)
        hip.htod(a.ptr, x)
        hip.memset(b.ptr, 0xEE, n * elem)
        assert L.librettExecute(h, a.ptr, b.ptr) == L.LIBRETT_SUCCESS
        hip.stream_sync(gpu_stream)
        out = np.empty_like(x)
        hip.dtoh(out, b.ptr)
        ref = oracle.transpose(dim, perm, elem, x)
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))
        L.librettDestroy(h)
    finally:
        L.selector_set(None, None, None)
    # Without weights the plan is the analytic model's choice again.
    r, h = L.librettPlan(len(dim), dim, perm, elem, gpu_stream)
    assert L.plan_info(h)["desc"] == L.candidates(dim, perm, elem)[model_choice]["desc"]
    L.librettDestroy(h)


def test_plan_follows_ensemble_mean(gpu_stream):
    dim, perm, elem = [2, 8, 16, 32, 64, 8], [5, 3, 1, 0, 2, 4], 4
    say("case: librettPlan with a two-network selector", dim, perm, elem)
    feats, model_choice = L.selector_features(dim, perm, elem)
    order = np.argsort(-feats.std(0))
    j, k = int(order[0]), int(order[1])

    def net(f, scale):
        W1 = np.zeros((M.H, M.F)); W1[0, f] = scale
        b1 = np.zeros(M.H); b1[0] = 1000
        W2 = np.zeros((M.H, M.H)); W2[0, 0] = 1
        w3 = np.zeros(M.H); w3[0] = 1
        return M.pack(W1, b1, W2, np.zeros(M.H), w3, 0.0)

    # Scores 3 f_j + 1000 and f_k + 1000: the ensemble ranks by 1.5 f_j + 0.5 f_k.
    p = np.concatenate([net(j, 3.0), net(k, 1.0)]).astype(np.float32)
    assert L.selector_set(p, np.zeros(M.F), np.ones(M.F)) == L.LIBRETT_SUCCESS
    try:
        score = 1.5 * feats[:, j].astype(np.float64) + 0.5 * feats[:, k]
        margin, low = L.selector_margin(), float(score.min())
        got = L.selector_choose(dim, perm, elem)
        # float32 scores around 1000: allow 1e-3 for rounding at ties.
        if score[model_choice] - low < margin - 1e-3:
            assert got == model_choice
        elif score[model_choice] - low > margin + 1e-3:
            assert score[got] <= low + 1e-3
        r, h = L.librettPlan(len(dim), dim, perm, elem, gpu_stream)
        assert r == L.LIBRETT_SUCCESS
        assert L.plan_info(h)["desc"] == L.candidates(dim, perm, elem)[got]["desc"]
        L.librettDestroy(h)
    finally:
        L.selector_set(None, None, None)


def test_retime_pairs(gpu_stream):
    # A 64 MiB f32 transpose: long enough that head-to-head timings agree.
    dim, perm, elem = [64, 1024, 256], [2, 0, 1], 4
    say("case: head-to-head re-timing", dim, perm, elem)
    n = int(np.prod(dim)) * elem
    a, b = hip.DeviceBuffer(n), hip.DeviceBuffer(n)
    hip.memset(a.ptr, 0, n)
    t = L.measure_candidates(dim, perm, elem, gpu_stream, a.ptr, b.ptr)
    ran = [i for i in range(len(t)) if t[i] > 0]
    assert len(ran) >= 2
    best = min(ran, key=lambda i: t[i])
    worst = max(ran, key=lambda i: t[i])
    rec = {"dims": dim, "perm": perm, "elem": elem, "times": t}
    r = S.retime([rec], [(0, best, best), (0, worst, best), (0, best, worst)], gpu_stream, a.ptr, b.ptr, rounds=7)
    assert r[0] == 1.0
    assert np.all(np.isfinite(r)) and np.all(r > 0)
    # The two orders of one pair are reciprocal up to timing noise (the
    # slowest candidate of a sweep can vary by a third between runs), and
    # the slow candidate stays slower than the fast one when the sweep saw a
    # clear gap.
    assert 1 / 1.5 < r[1] * r[2] < 1.5, r
    if t[worst] > 1.5 * t[best]:
        assert r[1] > 1.2 and r[2] < 1 / 1.2, (r, t[worst], t[best])

"""Learned plan selector on the host: the float64 restatement (oracle/mlp.py)
pinned by a hand-worked network, finite differences and Adam's closed-form
first step; the HIP kernels' per-thread code (host replay) against it; the
features; the weights file."""
import os
import subprocess
import sys

import numpy as np
import pytest

import librett_amd as L
from librett_amd import selector as S
from oracle import mlp as M

NF, NH, NP = L.mlp_sizes()


def rand_params(seed, scale=1.0):
    p = S.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    p[M.H * M.F:M.H * M.F + M.H] = rng.normal(0, 0.1, M.H)  # non-zero b1
    return (p * scale).astype(np.float32)


def test_sizes_match_restatement():
    assert (NF, NH, NP) == (M.F, M.H, M.P) == (50, 64, 64 * 50 + 64 + 64 * 64 + 64 + 64 + 1)


def test_forward_hand_worked():
    # One path through the network: W1[0,0]=2, b1[0]=-1, W2[0,0]=3, b2[0]=0.5,
    # w3[0]=4, b3=0.25. x0=1.5: z1=2, z2=6.5, out=26.25. x0=0.2: z1<0, out=4*0.5+0.25.
    W1 = np.zeros((M.H, M.F)); W1[0, 0] = 2
    b1 = np.zeros(M.H); b1[0] = -1
    W2 = np.zeros((M.H, M.H)); W2[0, 0] = 3
    b2 = np.zeros(M.H); b2[0] = 0.5
    w3 = np.zeros(M.H); w3[0] = 4
    p = M.pack(W1, b1, W2, b2, w3, 0.25)
    x = np.zeros((2, M.F)); x[0, 0] = 1.5; x[1, 0] = 0.2
    assert np.allclose(M.forward(p, x), [26.25, 2.25])
    assert np.allclose(L.mlp_host_forward(p, x), [26.25, 2.25])


def test_backward_finite_differences():
    rng = np.random.default_rng(0)
    p = rand_params(1).astype(np.float64)
    x = rng.normal(size=(20, M.F))
    y = rng.normal(size=20)
    g, _ = M.grad(p, x, y, 1 / 20)
    idx = rng.choice(M.P, 60, replace=False)
    idx = np.concatenate([idx, [M.P - 1, M.H * M.F, M.H * M.F + M.H + 5]])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.P); e[i] = h
        fd = (M.loss(p + e, x, y, 1 / 20) - M.loss(p - e, x, y, 1 / 20)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * (1 + abs(g[i])), (i, fd, g[i])


def test_weighted_backward_finite_differences():
    rng = np.random.default_rng(7)
    p = rand_params(3).astype(np.float64)
    x = rng.normal(size=(16, M.F))
    y = rng.normal(size=16)
    w = rng.random(16) + 0.1
    inv_m = 1 / w.sum()
    g, sse = M.grad(p, x, y, inv_m, w)
    assert abs(sse - np.sum(w * (M.forward(p, x) - y) ** 2)) < 1e-9 * (1 + sse)
    h = 1e-6
    for i in rng.choice(M.P, 40, replace=False):
        e = np.zeros(M.P); e[i] = h
        fd = (M.loss(p + e, x, y, inv_m, w) - M.loss(p - e, x, y, inv_m, w)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * (1 + abs(g[i])), (i, fd, g[i])
    # Unit weights are the unweighted loss.
    g1, _ = M.grad(p, x, y, 1 / 16, np.ones(16))
    g0, _ = M.grad(p, x, y, 1 / 16)
    assert np.allclose(g1, g0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    g = rng.normal(size=M.P)
    p0 = rng.normal(size=M.P)
    p1, m, v = M.adam(p0, np.zeros(M.P), np.zeros(M.P), g, 1e-3, 0.9, 0.999, 1, 1e-8)
    assert np.allclose(p1, p0 - 1e-3 * g / (np.abs(g) + 1e-8))


def test_host_forward_matches_restatement():
    rng = np.random.default_rng(4)
    p = rand_params(2)
    x = rng.normal(size=(300, M.F)).astype(np.float32)
    got = L.mlp_host_forward(p, x)
    ref = M.forward(p, x)
    assert np.all(np.abs(got - ref) <= 1e-4 * (1 + np.abs(ref)))
    # Argmin agreement per group of 25 candidates.
    assert np.array_equal(np.argmin(got.reshape(12, 25), 1), np.argmin(ref.reshape(12, 25), 1))


@pytest.mark.parametrize("count,weighted", [(128, False), (77, False), (1, False), (128, True), (77, True)])
def test_host_grad_matches_restatement(count, weighted):
    rng = np.random.default_rng(count)
    p = rand_params(5)
    n = 400
    x = rng.normal(size=(n, M.F)).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    w = (rng.random(n) + 0.1).astype(np.float32) if weighted else None
    idx = rng.choice(n, count, replace=False).astype(np.int32)
    inv_m = 1.0 / (float(w[idx].astype(np.float64).sum()) if weighted else count)
    g, sse = L.mlp_host_grad(p, x, y, idx, inv_m, w=w)
    gr, sser = M.grad(p, x[idx], y[idx], inv_m, None if w is None else w[idx])
    scale = np.abs(gr).max()
    assert np.all(np.abs(g - gr) <= 1e-4 * scale + 1e-7), np.abs(g - gr).max() / scale
    assert abs(sse - sser) <= 1e-4 * (1 + sser)


def test_host_adam_matches_restatement():
    rng = np.random.default_rng(9)
    parts = rng.normal(size=(3, M.P)).astype(np.float32)
    p = rng.normal(size=M.P).astype(np.float32)
    m = (rng.normal(size=M.P) * 0.1).astype(np.float32)
    v = (rng.random(M.P) * 0.1).astype(np.float32)
    t = 7
    # Hyper-parameters as the kernel code sees them (float32).
    lr, b1, b2 = (float(np.float32(v_)) for v_ in (1e-3, 0.9, 0.999))
    c1, c2 = 1 / (1 - b1 ** t), 1 / (1 - b2 ** t)
    pr, mr, vr = M.adam(p, m, v, parts.astype(np.float64).sum(0), lr, b1, b2, t, 1e-8)
    L.mlp_host_adam(parts, p, m, v, lr, b1, b2, c1, c2, 1e-8)
    assert np.allclose(m, mr, rtol=1e-5, atol=1e-7)
    assert np.allclose(v, vr, rtol=1e-5, atol=1e-7)
    assert np.allclose(p, pr, rtol=1e-5, atol=1e-6)


def test_features_describe_problem_and_candidate():
    dev = L.mi355x_desc()
    cases = [([64, 48, 32], [2, 0, 1], 8), ([8192, 16], [1, 0], 4), ([3, 3, 3, 2, 4], [1, 3, 0, 4, 2], 16),
             ([100, 1, 7], [2, 1, 0], 4), ([12, 34], [0, 1], 8)]
    for d, p, e in cases:
        f, mc = L.selector_features(d, p, e, device=dev)
        infos = L.host_candidates(d, p, e, device=dev)
        assert len(f) == len(infos) >= 1 and 0 <= mc < len(f)
        rd, rp = L.reduce_ranks(d, p)
        rr = len(rd)
        for fi, info in zip(f, infos):
            assert fi[0] == rr
            assert np.allclose(fi[1:1 + rr], np.log2(rd))
            assert np.array_equal(fi[8:8 + rr], np.array(rp) + 1)
            assert fi[15] == np.log2(e) and abs(fi[16] - np.log2(np.prod(d) * e)) < 1e-5
            onehot = fi[18:23]
            assert onehot.sum() == 1 and onehot[info["method"] - 1] == 1
            assert np.all(np.isfinite(fi))


def test_weights_file_roundtrip(tmp_path):
    p = rand_params(11)
    mean = np.linspace(-1, 1, M.F).astype(np.float32)
    std = np.linspace(0.5, 2, M.F).astype(np.float32)
    path = str(tmp_path / "w.bin")
    S.save_weights(path, p, mean, std)
    p2, m2, s2 = S.read_weights(path)
    assert np.array_equal(p, p2) and np.array_equal(mean, m2) and np.array_equal(std, s2)
    g0 = L.selector_generation()
    assert L.selector_load(path) == L.LIBRETT_SUCCESS
    g1 = L.selector_generation()
    assert g1 > g0 and g1 != 0
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(b"\0" * 64)
    assert L.selector_load(bad) == L.LIBRETT_INVALID_PARAMETER
    assert L.selector_generation() == g1
    p_nan = p.copy(); p_nan[5] = np.nan
    assert L.selector_set(p_nan, mean, std) == L.LIBRETT_INVALID_PARAMETER
    assert L.selector_set(None, None, None) == L.LIBRETT_SUCCESS
    assert L.selector_generation() == 0


def test_ensemble_weights_file(tmp_path):
    p = np.concatenate([rand_params(k) for k in (21, 22, 23)])
    mean = np.linspace(-1, 1, M.F).astype(np.float32)
    std = np.linspace(0.5, 2, M.F).astype(np.float32)
    path = str(tmp_path / "ens.bin")
    S.save_weights(path, p, mean, std)
    raw = open(path, "rb").read()
    assert raw[4:8] == (2).to_bytes(4, "little") and raw[16:20] == (3).to_bytes(4, "little")
    assert len(raw) == 20 + 4 * (2 * M.F + 3 * M.P)
    p2, m2, s2 = S.read_weights(path)
    assert np.array_equal(p, p2) and np.array_equal(mean, m2) and np.array_equal(std, s2)
    g0 = L.selector_generation()
    assert L.selector_load(path) == L.LIBRETT_SUCCESS and L.selector_generation() > g0
    # A parameter count that is not a whole number of networks, or more
    # networks than the selector averages, is refused.
    assert L.selector_set(p[:M.P + 1], mean, std) == L.LIBRETT_INVALID_PARAMETER
    assert L.selector_set(np.tile(p, 6), mean, std) == L.LIBRETT_INVALID_PARAMETER
    open(str(tmp_path / "short.bin"), "wb").write(raw[:-4])
    assert L.selector_load(str(tmp_path / "short.bin")) == L.LIBRETT_INVALID_PARAMETER
    assert L.selector_set(None, None, None) == L.LIBRETT_SUCCESS


def test_problem_generator_ranges():
    for d, p, e in S.problems(300, 5):
        assert 2 <= len(d) <= 7 and sorted(p) == list(range(len(d))) and e in (4, 8, 16)
        assert np.prod(d) * e <= 512 << 20 and min(d) >= 2


def test_records_roundtrip_without_features(tmp_path):
    recs = [{"dims": [4, 5, 6], "perm": [2, 0, 1], "elem": 8, "feats": np.ones((3, M.F), np.float32),
             "times": np.array([1e-3, -1.0, 2e-3]), "model_choice": 2},
            {"dims": [7, 9], "perm": [1, 0], "elem": 4, "feats": np.zeros((1, M.F), np.float32),
             "times": np.array([5e-4]), "model_choice": 0}]
    path = str(tmp_path / "r.npz")
    S.save_records(path, recs, with_features=False)
    back = S.load_records(path)
    assert len(back) == 2 and back[0]["feats"] is None
    for r, b in zip(recs, back):
        assert r["dims"] == b["dims"] and r["perm"] == b["perm"] and r["elem"] == b["elem"]
        assert r["model_choice"] == b["model_choice"]
        assert np.allclose(r["times"], b["times"], rtol=1e-6)
    # Candidates that did not run stay marked, and the targets skip them.
    y = S.targets(back[0])
    assert np.isnan(y[1]) and np.all(np.isfinite(y[[0, 2]]))


def test_records_roundtrip_with_features_many(tmp_path):
    # Thousands of records load in well under a second: every member of the
    # file is read once, not once per record.
    import time
    rng = np.random.default_rng(3)
    recs = []
    for i in range(3000):
        n = int(rng.integers(1, 40))
        recs.append({"dims": [2 + i % 7, 3], "perm": [1, 0], "elem": 8,
                     "feats": rng.standard_normal((n, M.F)).astype(np.float32),
                     "times": rng.uniform(1e-5, 1e-3, n), "model_choice": int(rng.integers(0, n))})
    path = str(tmp_path / "f.npz")
    S.save_records(path, recs)
    t = time.time()
    back = S.load_records(path)
    assert time.time() - t < 5.0
    assert len(back) == len(recs)
    for r, b in zip(recs[::97], back[::97]):
        assert np.array_equal(r["feats"], b["feats"]) and np.allclose(r["times"], b["times"], rtol=1e-6)
        assert r["model_choice"] == b["model_choice"] and r["dims"] == b["dims"]


def test_unloading_first_keeps_analytic_model(tmp_path):
    # In a fresh process, clearing the weights before anything else must not
    # be undone by the first-use load of the default weights file.
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import librett_amd as L; "
            "assert L.selector_set(None, None, None) == 0; print(L.selector_generation())") % S.__file__.rsplit("/", 2)[0]
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "0"


def test_margin_rule_in_evaluation():
    # Two candidates: the analytic model picks 1 (20% slower than 0).
    rec = {"times": np.array([1.0, 1.2]), "model_choice": 1}
    close, far = np.array([0.0, 0.05]), np.array([0.0, 0.2])
    sel, ana = S.evaluate([rec, rec], [close, far], margin=0.1)
    assert list(sel) == [1.2, 1.0] and list(ana) == [1.2, 1.2]
    sel, _ = S.evaluate([rec, rec], [close, far])
    assert list(sel) == [1.0, 1.0]
    # A candidate that did not run is never chosen.
    rec2 = {"times": np.array([-1.0, 1.0, 2.0]), "model_choice": 2}
    sel, _ = S.evaluate([rec2], [np.array([-5.0, 0.0, 0.5])], margin=0.1)
    assert list(sel) == [1.0]


def test_picks_name_the_pairs_evaluate_divides():
    recs = [{"times": np.array([1.0, 1.2, 0.9]), "model_choice": 1},
            {"times": np.array([-1.0, 2.0, 3.0]), "model_choice": 0},
            {"times": np.array([-1.0, -1.0]), "model_choice": 0}]
    sc = [np.array([0.0, 0.05, 0.2]), np.array([-9.0, 0.3, 0.1]), np.array([0.0, 0.0])]
    # Record 0: selector argmin 0, analytic 1, best 2; record 1: candidate 0
    # did not run (never picked, analytic -1), argmin 2, best 1; record 2:
    # nothing ran, skipped.
    assert S.picks(recs, sc) == [(0, 0, 1, 2), (1, 2, -1, 1)]
    # Margin 0.1: record 0 keeps the analytic choice (0.05 < 0.1).
    assert S.picks(recs, sc, margin=0.1) == [(0, 1, 1, 2), (1, 2, -1, 1)]
    sel, ana = S.evaluate(recs, sc, margin=0.1)
    assert np.allclose(sel, [1.2 / 0.9, 1.5]) and ana[0] == 1.2 / 0.9 and np.isnan(ana[1])


def test_margin_default_and_override():
    code = "import librett_amd as L; print(L.selector_margin())"
    env = dict(os.environ)
    env.pop("LIBRETT_SELECTOR_MARGIN", None)
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, check=True)
    assert float(out.stdout) == 0.03
    env["LIBRETT_SELECTOR_MARGIN"] = "0"
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, check=True)
    assert float(out.stdout) == 0.0

"""Learned plan selector: data collection, training and evaluation.

The MLP (2 x 64, FP32) scores every candidate plan of a problem from the
features the library computes (``librett_c_selector_features``: reduced
rank, log-dims, permutation, element size, the candidate's kernel shape and
the analytic model's transaction counts and times). It is trained to
predict ``log2(t / t_roof)`` -- the measured time of the candidate over the
8 TB/s roofline time of the move -- on timings that ``librettPlanMeasure``'s
own loop produces (``librett_c_measure_candidates``), with the library's HIP
training kernels (``librett_c_mlp_train``). ``librettPlan`` then takes the
candidate with the lowest score (``librett_c_selector_load``) and falls back
to the analytic model when no weights are loaded.

Command line (on a GPU)::

    # timings of problems [0, 10000) and [10000, 20000) of seed 1 (two runs)
    python -m librett_amd.selector collect --problems 10000 --total 20000 --offset 0 --seed 1 --out a.npz
    python -m librett_amd.selector collect --problems 10000 --total 20000 --offset 10000 --seed 1 --out b.npz
    # features recomputed on this GPU, training, held-out evaluation (seed + 1000)
    python -m librett_amd.selector train --data a.npz b.npz --heldout 2000 --seed 1 --weights w.bin --report r.json
    python -m librett_amd.selector evaluate --problems 2000 --seed 1001 --weights librett_amd/selector_weights.bin
    # all in one run
    python -m librett_amd.selector build --train 20000 --heldout 2000 --weights ... --report report.json

The reference has no learned selector; its heuristic is choosePlanHeuristic
(src/plan.cpp:721-789) over the GpuModel cycle counts (src/GpuModel.cpp).
"""
import argparse
import json
import os
import struct
import sys
import time

import numpy as np

MAGIC = 0x4D54524C  # "LRTM"
PEAK_BYTES_PER_S = 8.0e12
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_weights.bin")


# ---------------------------------------------------------------------------
# Weights file
# ---------------------------------------------------------------------------

def save_weights(path, params, mean, std):
    """One network (version 1) or an ensemble of k networks sharing the
    normalisation (version 2: k after the header, then k parameter sets)."""
    from . import mlp_sizes
    nf, nh, npar = mlp_sizes()
    params = np.asarray(params, dtype="<f4").reshape(-1)
    k = params.size // npar
    assert k >= 1 and params.size == k * npar and len(mean) == nf and len(std) == nf
    with open(path, "wb") as fh:
        if k == 1:
            fh.write(struct.pack("<4I", MAGIC, 1, nf, nh))
        else:
            fh.write(struct.pack("<5I", MAGIC, 2, nf, nh, k))
        fh.write(np.asarray(mean, dtype="<f4").tobytes())
        fh.write(np.asarray(std, dtype="<f4").tobytes())
        fh.write(params.tobytes())


def read_weights(path):
    """(params of all networks, flat; mean; std)."""
    from . import mlp_sizes
    nf, nh, npar = mlp_sizes()
    raw = open(path, "rb").read()
    magic, ver, f, h = struct.unpack("<4I", raw[:16])
    if magic != MAGIC or ver not in (1, 2) or (f, h) != (nf, nh):
        raise ValueError("not a selector weights file: %s" % path)
    k, head = 1, 16
    if ver == 2:
        k, = struct.unpack("<I", raw[16:20])
        head = 20
    a = np.frombuffer(raw[head:], dtype="<f4")
    if k < 1 or a.size != 2 * nf + k * npar:
        raise ValueError("truncated selector weights file: %s" % path)
    return a[2 * nf:].copy(), a[:nf].copy(), a[nf:2 * nf].copy()


def init_params(seed):
    """He-initialised parameters (biases zero)."""
    from . import mlp_sizes
    nf, nh, npar = mlp_sizes()
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, np.sqrt(2.0 / nf), (nh, nf))
    w2 = rng.normal(0, np.sqrt(2.0 / nh), (nh, nh))
    w3 = rng.normal(0, np.sqrt(1.0 / nh), nh)
    p = np.concatenate([w1.ravel(), np.zeros(nh), w2.ravel(), np.zeros(nh), w3, [0.0]]).astype(np.float32)
    assert p.size == npar
    return p


# ---------------------------------------------------------------------------
# Problems
# ---------------------------------------------------------------------------

def random_problem(rng, rank_lo=2, rank_hi=7, bytes_lo=1 << 20, bytes_hi=512 << 20, elems=(4, 8, 16)):
    """Rank, element size and a random permutation drawn uniformly; the
    volume log-uniform in [bytes_lo, bytes_hi] bytes, split over the dims by
    a random log-partition (every dim >= 2)."""
    elem = int(rng.choice(elems))
    rank = int(rng.integers(rank_lo, rank_hi + 1))
    target = float(np.exp(rng.uniform(np.log(bytes_lo), np.log(bytes_hi)))) / elem
    w = rng.dirichlet(np.ones(rank))
    dims = [max(2, int(round(target ** wi))) for wi in w]
    # Trim the largest dims until the volume is within the range.
    while int(np.prod(dims)) * elem > bytes_hi:
        i = int(np.argmax(dims))
        dims[i] = max(2, dims[i] // 2)
    perm = [int(v) for v in rng.permutation(rank)]
    return dims, perm, elem


def pow2_problem(rng, rank_lo=2, rank_hi=6, bytes_lo=1 << 20, bytes_hi=512 << 20, elems=(4, 8, 16)):
    """Like random_problem, but every dim a power of two (>= 2): extents whose
    rows fill whole 128-byte lines, which random dims almost never give."""
    elem = int(rng.choice(elems))
    rank = int(rng.integers(rank_lo, rank_hi + 1))
    total = int(round(np.log2(float(np.exp(rng.uniform(np.log(bytes_lo), np.log(bytes_hi)))) / elem)))
    total = max(rank, total)
    w = rng.dirichlet(np.ones(rank))
    ks = [max(1, int(round(total * wi))) for wi in w]
    while sum(ks) > total and max(ks) > 1:
        ks[int(np.argmax(ks))] -= 1
    dims = [1 << k for k in ks]
    perm = [int(v) for v in rng.permutation(rank)]
    return dims, perm, elem


def problems(n, seed, family="random", **kw):
    """n seeded problems of one family: "random" (random_problem) or "pow2"
    (pow2_problem)."""
    rng = np.random.default_rng(seed)
    gen = pow2_problem if family == "pow2" else random_problem
    return [gen(rng, **kw) for _ in range(n)]


# ---------------------------------------------------------------------------
# Collection (GPU)
# ---------------------------------------------------------------------------

def collect(probs, stream, in_ptr, out_ptr, log=None):
    """Features and measured times of every candidate of every problem."""
    from . import measure_candidates, selector_features
    recs = []
    t0 = time.time()
    last = t0
    for k, (d, p, e) in enumerate(probs):
        feats, mc = selector_features(d, p, e)
        times = measure_candidates(d, p, e, stream, in_ptr, out_ptr)
        n = min(len(feats), len(times))
        recs.append({"dims": list(d), "perm": list(p), "elem": int(e), "feats": feats[:n], "times": times[:n],
                     "model_choice": int(mc)})
        if log and (time.time() - last > 20 or k + 1 == len(probs)):
            last = time.time()
            log("collected %d/%d problems (%.0f s)" % (k + 1, len(probs), last - t0))
    return recs


def save_records(path, recs, with_features=True):
    """Records to a compressed .npz; without features it holds only the
    problems and the measured times (features are recomputed with
    attach_features on the same GPU model)."""
    off = np.cumsum([0] + [len(r["times"]) for r in recs])
    arrays = dict(
        times=np.concatenate([r["times"] for r in recs]).astype(np.float32),
        offsets=off,
        model_choice=np.array([r["model_choice"] for r in recs]),
        elem=np.array([r["elem"] for r in recs]),
        problems=np.array([json.dumps([r["dims"], r["perm"]]) for r in recs]),
    )
    if with_features:
        arrays["feats"] = np.concatenate([r["feats"] for r in recs]).astype(np.float32)
    np.savez_compressed(path, **arrays)


def load_records(path):
    # Each member of an .npz is read (and decompressed) once: indexing the
    # file object per record would decompress the whole array every time.
    with np.load(path, allow_pickle=False) as z:
        off = z["offsets"]
        feats = z["feats"] if "feats" in z.files else None
        times = z["times"].astype(np.float64)
        problems = z["problems"]
        elem = z["elem"]
        choice = z["model_choice"]
    recs = []
    for i in range(len(off) - 1):
        a, b = off[i], off[i + 1]
        d, p = json.loads(str(problems[i]))
        recs.append({"dims": d, "perm": p, "elem": int(elem[i]), "feats": feats[a:b] if feats is not None else None,
                     "times": times[a:b], "model_choice": int(choice[i])})
    return recs


def attach_features(recs, log=None):
    """Recomputes every record's candidate features on the current GPU;
    drops records whose candidate list no longer matches their times."""
    from . import selector_features
    out = []
    t0 = last = time.time()
    for k, r in enumerate(recs):
        feats, mc = selector_features(r["dims"], r["perm"], r["elem"])
        if len(feats) == len(r["times"]) and mc == r["model_choice"]:
            r = dict(r)
            r["feats"] = feats
            out.append(r)
        if log and time.time() - last > 5:
            last = time.time()
            log("features %d/%d (%.0f s)" % (k + 1, len(recs), last - t0))
    if log:
        log("features for %d of %d records" % (len(out), len(recs)))
    return out


# ---------------------------------------------------------------------------
# Training and evaluation
# ---------------------------------------------------------------------------

def targets(rec):
    """log2(t / t_roof) of every candidate that ran (NaN where it did not)."""
    nbytes = 2.0 * float(np.prod(rec["dims"])) * rec["elem"]
    t = np.asarray(rec["times"], dtype=np.float64)
    y = np.full(len(t), np.nan)
    ok = t > 0
    y[ok] = np.log2(t[ok] / (nbytes / PEAK_BYTES_PER_S))
    return y


# Sample weights: a candidate within a factor 2^tau of its problem's fastest
# counts fully, slower ones fade to the floor; selection only needs the
# order near the top.
WEIGHT_TAU = 0.5
WEIGHT_FLOOR = 0.1


def dataset(recs):
    """(features, targets, weights) of every candidate that ran."""
    xs, ys, ws = [], [], []
    for r in recs:
        y = targets(r)
        ok = np.isfinite(y)
        xs.append(r["feats"][ok])
        ys.append(y[ok])
        ws.append(np.exp(-(y[ok] - y[ok].min()) / WEIGHT_TAU) + WEIGHT_FLOOR)
    return (np.concatenate(xs).astype(np.float32), np.concatenate(ys).astype(np.float32),
            np.concatenate(ws).astype(np.float32))


def normaliser(x):
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    std[std < 1e-6] = 1.0
    return mean.astype(np.float32), std.astype(np.float32)


def train(recs, stream, epochs=60, lr=2e-3, minibatch=2048, seed=1, log=None):
    """Trains on the GPU; returns (params, mean, std, per-epoch loss)."""
    from . import hip, mlp_train
    x, y, w = dataset(recs)
    mean, std = normaliser(x)
    xn = ((x - mean) / std).astype(np.float32)
    p = init_params(seed)
    dx, dy, dw, dp = (hip.DeviceBuffer(xn.nbytes), hip.DeviceBuffer(y.nbytes), hip.DeviceBuffer(w.nbytes),
                      hip.DeviceBuffer(p.nbytes))
    hip.htod(dx.ptr, xn)
    hip.htod(dy.ptr, y)
    hip.htod(dw.ptr, w)
    hip.htod(dp.ptr, p)
    losses = []
    chunk = 10
    for e0 in range(0, epochs, chunk):
        n_ep = min(chunk, epochs - e0)
        losses += list(mlp_train(dx.ptr, dy.ptr, len(y), dp.ptr, n_ep, lr, minibatch, seed + e0, stream,
                                 d_w=dw.ptr, w_host=w))
        if log:
            log("epoch %d: weighted mse %.4f" % (e0 + n_ep, losses[-1]))
    hip.stream_sync(stream)
    hip.dtoh(p, dp.ptr)
    for b in (dx, dy, dw, dp):
        b.free()
    return p, mean, std, np.array(losses)


def train_ensemble(recs, stream, models=1, epochs=60, lr=2e-3, seed=1, log=None):
    """Trains `models` networks with seeds seed, seed + 1, ... on the same
    data; returns (all parameters, flat; mean; std; last network's losses)."""
    ps = []
    for k in range(models):
        p, mean, std, losses = train(recs, stream, epochs, lr, seed=seed + k, log=log)
        ps.append(p)
    return np.concatenate(ps), mean, std, losses


def scores(params, mean, std, recs, stream):
    """MLP scores of every candidate, computed by the HIP inference kernel;
    the mean over the networks of an ensemble."""
    from . import hip, mlp_forward, mlp_sizes
    npar = mlp_sizes()[2]
    x = np.concatenate([r["feats"] for r in recs]).astype(np.float32)
    xn = ((x - mean) / std).astype(np.float32)
    nets = np.asarray(params, dtype=np.float32).reshape(-1, npar)
    out = np.zeros(len(xn), dtype=np.float64)
    one = np.zeros(len(xn), dtype=np.float32)
    dx, dp, do = hip.DeviceBuffer(max(4, xn.nbytes)), hip.DeviceBuffer(npar * 4), hip.DeviceBuffer(max(4, one.nbytes))
    hip.htod(dx.ptr, xn)
    for p in nets:
        hip.htod(dp.ptr, np.ascontiguousarray(p))
        mlp_forward(dp.ptr, dx.ptr, len(xn), do.ptr, stream)
        hip.stream_sync(stream)
        hip.dtoh(one, do.ptr)
        out += one
    out = (out / len(nets)).astype(np.float32)
    for b in (dx, dp, do):
        b.free()
    res, o = [], 0
    for r in recs:
        res.append(out[o:o + len(r["times"])])
        o += len(r["times"])
    return res


def picks(recs, sc, margin=0.0):
    """Per problem with at least one timed candidate: (record index, the
    selector's pick, the analytic model's pick, the fastest timed candidate).
    With a margin, the selector's pick replaces the analytic choice only when
    its score is lower by more than the margin (librettPlan's rule). The
    analytic pick is -1 where the model's choice did not run."""
    out = []
    for k, (r, s) in enumerate(zip(recs, sc)):
        t = np.asarray(r["times"], dtype=np.float64)
        ok = t > 0
        if not ok.any():
            continue
        b = int(np.argmin(np.where(ok, t, np.inf)))
        s = np.where(ok, np.asarray(s, dtype=np.float64), np.inf)
        i = int(np.argmin(s))
        m = r["model_choice"]
        if margin > 0 and 0 <= m < len(t) and s[m] - s[i] < margin:
            i = m
        out.append((k, i, m if 0 <= m < len(t) and t[m] > 0 else -1, b))
    return out


def evaluate(recs, sc, margin=0.0):
    """Slowdown of the selector's and the analytic model's choice versus the
    fastest measured candidate, per problem (margin as in picks())."""
    sel, ana = [], []
    for k, i, m, b in picks(recs, sc, margin):
        t = np.asarray(recs[k]["times"], dtype=np.float64)
        sel.append(t[i] / t[b])
        ana.append(t[m] / t[b] if m >= 0 else np.nan)
    return np.array(sel), np.array(ana)


def retime(recs, chosen, stream, in_ptr, out_ptr, rounds=15):
    """Slowdowns of chosen candidates, timed again head to head with the
    fastest one. chosen: (record index, pick, best) triples. A pick that is
    the best counts 1.0. Otherwise both are planned (librett_c_plan_candidate)
    and run alternately, `rounds` times each, every run after an output
    memset as in librettPlanMeasure's loop, with the first of the pair
    alternating between rounds; the slowdown is the ratio of the medians.
    One timing per candidate (the exhaustive sweep) is noisy on problems
    that run for a few microseconds; this pass measures the pairs that
    decide the slowdown with many runs."""
    from . import hip, librettDestroy, librettExecute, plan_candidate, LIBRETT_SUCCESS
    e0, e1 = hip.Event(), hip.Event()
    res = []
    for k, i, b in chosen:
        if i == b:
            res.append(1.0)
            continue
        r = recs[k]
        d, p, e = r["dims"], r["perm"], r["elem"]
        n = int(np.prod(d)) * e
        hs = []
        for idx in (i, b):
            rc, h = plan_candidate(len(d), d, p, e, stream, idx)
            if rc != LIBRETT_SUCCESS:
                raise RuntimeError("plan_candidate %d of %s %s failed: %d" % (idx, d, p, rc))
            hs.append(h)
        ts = ([], [])
        for h in hs:
            librettExecute(h, in_ptr, out_ptr)
        hip.stream_sync(stream)
        for rd in range(rounds):
            for j in ((0, 1) if rd % 2 == 0 else (1, 0)):
                hip.memset(out_ptr, 0xFF, n, stream)
                e0.record(stream)
                librettExecute(hs[j], in_ptr, out_ptr)
                e1.record(stream)
                e1.synchronize()
                ts[j].append(e1.elapsed_ms(e0))
        for h in hs:
            librettDestroy(h)
        res.append(float(np.median(ts[0]) / np.median(ts[1])))
    e0.destroy()
    e1.destroy()
    return np.array(res)


def summary(slow):
    slow = slow[np.isfinite(slow)]
    return {"mean": float(slow.mean()), "geomean": float(np.exp(np.log(slow).mean())),
            "p50": float(np.median(slow)), "p90": float(np.percentile(slow, 90)),
            "p95": float(np.percentile(slow, 95)), "p99": float(np.percentile(slow, 99)), "max": float(slow.max()),
            "within_5pct": float((slow <= 1.05).mean()), "exact": float((slow <= 1.0 + 1e-9).mean()),
            "problems": int(len(slow))}


def _buffers(max_bytes):
    from . import hip
    a, b = hip.DeviceBuffer(max_bytes), hip.DeviceBuffer(max_bytes)
    hip.memset(a.ptr, 0, max_bytes)
    hip.memset(b.ptr, 0, max_bytes)
    return a, b


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m librett_amd.selector")
    ap.add_argument("cmd", choices=["collect", "train", "evaluate", "build"])
    ap.add_argument("--train", type=int, default=20000, help="build: training problems (seed --seed)")
    ap.add_argument("--heldout", type=int, default=2000, help="build: held-out problems (seed --seed + 1000)")
    ap.add_argument("--report", default=None, help="build / train: JSON report path")
    ap.add_argument("--total", type=int, default=None, help="collect: size of the seeded problem list")
    ap.add_argument("--offset", type=int, default=0, help="collect: first problem of the list to time")
    ap.add_argument("--problems", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--family", choices=["random", "pow2"], default="random",
                    help="collect: problem family (random dims, or powers of two)")
    ap.add_argument("--out", default="selector_data.npz")
    ap.add_argument("--data", nargs="+", default=None)
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--heldout-out", default=None, help="train: also save the held-out records here")
    ap.add_argument("--models", type=int, default=1,
                    help="build / train: networks in the ensemble (seeds --seed, --seed + 1, ...)")
    a = ap.parse_args(argv)
    from . import hip

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    log("selector %s: start" % a.cmd)
    stream = hip.stream_create()
    if a.cmd == "build":
        A, B = _buffers(512 << 20)
        t0 = time.time()
        tr = collect(problems(a.train, a.seed), stream, A.ptr, B.ptr, log)
        t1 = time.time()
        p, mean, std, losses = train_ensemble(tr, stream, a.models, a.epochs, a.lr, a.seed, log)
        t2 = time.time()
        save_weights(a.weights, p, mean, std)
        ho = collect(problems(a.heldout, a.seed + 1000), stream, A.ptr, B.ptr, log)
        s_tr, a_tr = evaluate(tr, scores(p, mean, std, tr, stream))
        s_ho, a_ho = evaluate(ho, scores(p, mean, std, ho, stream))
        rep = {"train_problems": len(tr), "heldout_problems": len(ho), "samples": int(sum(len(r["times"]) for r in tr)),
               "epochs": a.epochs, "lr": a.lr, "final_mse": float(losses[-1]), "collect_s": t1 - t0,
               "train_s": t2 - t1, "seed": a.seed,
               "train": {"selector": summary(s_tr), "analytic": summary(a_tr)},
               "heldout": {"selector": summary(s_ho), "analytic": summary(a_ho)}}
        print(json.dumps(rep))
        if a.report:
            with open(a.report, "w") as fh:
                json.dump(rep, fh, indent=1)
        if a.out:
            save_records(a.out, ho)
    elif a.cmd == "collect":
        A, B = _buffers(512 << 20)
        probs = problems(a.total or a.problems, a.seed, a.family)[a.offset:a.offset + a.problems]
        recs = collect(probs, stream, A.ptr, B.ptr, log)
        save_records(a.out, recs, with_features=False)
    elif a.cmd == "train":
        recs = []
        for path in a.data:
            recs += load_records(path)
        log("loaded %d records from %d files" % (len(recs), len(a.data)))
        if any(r["feats"] is None for r in recs):
            recs = attach_features(recs, log)
        t1 = time.time()
        p, mean, std, losses = train_ensemble(recs, stream, a.models, a.epochs, a.lr, a.seed, log)
        t2 = time.time()
        save_weights(a.weights, p, mean, std)
        s_tr, a_tr = evaluate(recs, scores(p, mean, std, recs, stream))
        rep = {"train_problems": len(recs), "samples": int(sum(len(r["times"]) for r in recs)),
               "epochs": a.epochs, "lr": a.lr, "final_mse": float(losses[-1]), "train_s": t2 - t1, "seed": a.seed,
               "models": a.models,
               "train": {"selector": summary(s_tr), "analytic": summary(a_tr)}}
        if a.heldout > 0:  # --heldout 0 leaves the held-out evaluation to a run of its own
            A, B = _buffers(512 << 20)
            ho = collect(problems(a.heldout, a.seed + 1000), stream, A.ptr, B.ptr, log)
            s_ho, a_ho = evaluate(ho, scores(p, mean, std, ho, stream))
            rep["heldout_problems"] = len(ho)
            rep["heldout"] = {"selector": summary(s_ho), "analytic": summary(a_ho)}
            if a.heldout_out:
                save_records(a.heldout_out, ho)
        print(json.dumps(rep))
        if a.report:
            with open(a.report, "w") as fh:
                json.dump(rep, fh, indent=1)
    else:
        if a.data:
            recs = []
            for path in a.data:
                recs += load_records(path)
            if any(r["feats"] is None for r in recs):
                recs = attach_features(recs, log)
        else:
            A, B = _buffers(512 << 20)
            recs = collect(problems(a.problems, a.seed), stream, A.ptr, B.ptr, log)
        p, mean, std = read_weights(a.weights)
        sel, ana = evaluate(recs, scores(p, mean, std, recs, stream))
        print(json.dumps({"heldout": {"selector": summary(sel), "analytic": summary(ana)}}))


if __name__ == "__main__":
    main()

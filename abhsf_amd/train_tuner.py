"""Train the learned tuner on the GPU and evaluate it against exhaustive search.

  python -m abhsf_amd.train_tuner --out DIR [--mode latency|batch] [--steps N --lr R --members K]

Mode latency (single-matrix SpMV):

1. corpus: seeded synthetic matrices of every family (train and held-out
   sets use different seeds and sizes; the held-out set contains the R-MAT
   scale-22 and block_mix n = 2^22 workloads);
2. labels: exhaustive timing sweep of all (s, variant) candidates
   (tuner.sweep) -> y = log(ms per entry);
3. features: tuner.matrix_features (GPU block histograms on a strip sample);
4. training: tuner.train (full-batch Adam by the library's MLP kernels);
5. report: per held-out matrix, the regret of the MLP pick, of plain ABHSF
   (s = 64) and of the storage-optimal block size (the reference's search,
   block_statistics/mmf.h:77-102), all against the exhaustive best.
Writes DIR/tuner_model.json, DIR/tuner_dataset.npz and DIR/tuner_report.json.

Mode batch (many matrices in one launch): matrices drawn like the batch
workload (synth.batch_member, other seeds: 300 train, 100 held-out); for each
candidate (s, balanced variant: tuner.BATCH_VARIANTS) every member is built at
that candidate and one batched launch is
profiled (abhsf_batch_profile): the label is log(cycles per entry) of the
member's work items inside the full launch, i.e. its share of the GPU.  The
held-out report gives per-matrix regret and the measured batched launch time
of the model's picks against plain s = 64 and the best single candidate.
Writes DIR/tuner_model_batch.json, DIR/tuner_dataset_batch.npz and
DIR/tuner_report_batch.json.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import hip, synth, tuner


# Training matrices of the same configuration (family and parameters, another
# seed) as a held-out or benchmark matrix: left out, so that held-out regret
# measures unseen configurations.
EXCLUDED_SAME_CONFIG = {
    "mix_1048576_0.2_0.3": "block_mix n=2^20 (0.2, 0.3) = held-out mix_2^20_a",
    "mix_4194304_0.2_0.3": "block_mix n=2^22 (0.2, 0.3) = held-out block_mix_2^22 and the benchmark matrix",
    "rmat_22_16": "R-MAT scale 22 edge factor 16 = held-out rmat_22_16 and the benchmark matrix",
}


def family(name: str) -> str:
    return {"lap": "laplacian", "stencil": "stencil27", "mix": "block_mix", "block": "block_mix", "rmat": "rmat",
            "banded": "banded", "rblk": "random_blocks"}[name.split("_")[0]]


def corpus(split: str):
    """(name, generator) pairs."""
    if split == "train":
        return [(nm, g) for nm, g in corpus("all_train") if nm not in EXCLUDED_SAME_CONFIG]
    if split == "all_train":
        out = [(f"lap_{g}", (lambda g=g: synth.laplacian2d(g))) for g in (256, 450, 700, 1500, 2400)]
        # the stencil's fastest block size follows the grid size modulo s
        # (where grid lines fall in the blocks): sizes 40 to 150 in steps of
        # 8 to 16, without the held-out 100 and the SpMM benchmark's 160
        out += [(f"stencil_{g}", (lambda g=g: synth.stencil27(g)))
                for g in (40, 48, 56, 64, 72, 80, 88, 96, 112, 120, 128, 136, 144, 150)]
        for n in (1 << 17, 1 << 19, 1 << 21):
            for pd, pb in ((0.2, 0.3), (0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (0.5, 0.5)):
                out.append((f"mix_{n}_{pd}_{pb}", (lambda n=n, pd=pd, pb=pb: synth.block_mix(n, 11, pd, pb))))
        for n in (1 << 18, 1 << 20, 1 << 22):
            for pd, pb in ((0.1, 0.1), (0.3, 0.6), (0.2, 0.3)):
                out.append((f"mix_{n}_{pd}_{pb}", (lambda n=n, pd=pd, pb=pb: synth.block_mix(n, 12, pd, pb))))
        for sc, ef in ((16, 8), (17, 16), (18, 16), (18, 32), (19, 8), (20, 8), (20, 16), (21, 16), (22, 8),
                       (22, 16)):
            out.append((f"rmat_{sc}_{ef}", (lambda sc=sc, ef=ef: synth.rmat(sc, ef, seed=11))))
        for n, bw in ((100000, 4), (100000, 32), (200000, 128), (500000, 16), (1000000, 8), (1000000, 64),
                      (2000000, 4)):
            out.append((f"banded_{n}_{bw}", (lambda n=n, bw=bw: synth.banded(n, bw, 11))))
        for n, bs, nbr, d in ((100000, 8, 4, 0.3), (100000, 16, 2, 0.1), (1000000, 4, 8, 0.5),
                              (1000000, 32, 4, 0.05), (500000, 16, 8, 0.2), (2000000, 8, 2, 0.4),
                              (300000, 64, 2, 0.02)):
            out.append((f"rblk_{n}_{bs}_{nbr}_{d}", (lambda n=n, bs=bs, nbr=nbr, d=d:
                                                    synth.random_blocks(n, bs, nbr, d, 11))))
        return out
    return [
        ("lap_1024", lambda: synth.laplacian2d(1024)),
        ("stencil_100", lambda: synth.stencil27(100)),
        ("mix_2^20_a", lambda: synth.block_mix(1 << 20, 5)),
        ("mix_2^20_b", lambda: synth.block_mix(1 << 20, 6, 0.3, 0.2)),
        ("rmat_19_16", lambda: synth.rmat(19, 16, seed=3)),
        ("banded_300000_16", lambda: synth.banded(300000, 16, 3)),
        ("rblk_300000_8_4_0.2", lambda: synth.random_blocks(300000, 8, 4, 0.2, 3)),
        ("rmat_22_16", lambda: synth.rmat(22, 16, seed=0)),
        ("block_mix_2^22", lambda: synth.block_mix(1 << 22, 0)),
    ]


def collect(split, stream, log):
    X, y, names, times = [], [], [], []
    for name, gen in corpus(split):
        t0 = time.perf_counter()
        A = gen()
        d_rp, d_c = hip.DeviceArray.from_numpy(A.rowptr), hip.DeviceArray.from_numpy(A.col)
        tf = time.perf_counter()
        Xc, per_s = tuner.matrix_features(A, d_rp, d_c, stream)
        features_ms = (time.perf_counter() - tf) * 1e3
        tm = tuner.sweep(A, stream)
        for i, cand in enumerate(tuner.candidates()):
            X.append(Xc[i])
            y.append(math.log(tm[cand] / A.nnz))
        names.append(name)
        times.append({f"{s}/{v}": tm[(s, v)] for s, v in tm} | {"nnz": A.nnz, "features_ms": features_ms,
                     "storage_opt": list(tuner.pick_storage_optimal(per_s))})
        best = min(tm, key=tm.get)
        log(f"{split} {name}: nnz={A.nnz} best={best} {tm[best]:.4f} ms "
            f"s64={tm[(64, 'balanced')]:.4f} ms ({time.perf_counter() - t0:.1f} s)")
    return np.array(X), np.array(y), names, times


def batch_corpus(split):
    seed, count = (1000, 300) if split == "train" else (2000, 100)
    rng = np.random.default_rng(seed)
    return [synth.batch_member(rng, i, seed) for i in range(count)]


def time_graph(B, stream, reps=5):
    for _ in range(2):
        B.spmv_async(1.0, 1.0, graph=True)
    B.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(reps):
        B.spmv_async(1.0, 1.0, graph=True)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps


def build_batch(mats, picks, xs, ys, stream):
    from .matrix import AbhsfBatch, AbhsfMatrix
    Ms = []
    for A, (s, v) in zip(mats, picks):
        M = AbhsfMatrix(block_size=s, variant=v, stream=stream)
        M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
        Ms.append(M)
    B = AbhsfBatch(Ms, stream)
    B.set_vectors(xs, ys)
    return B, Ms


def collect_batch(split, stream, log):
    """Features and in-launch cycle labels of every member for every candidate."""
    t0 = time.perf_counter()
    mats = batch_corpus(split)
    cands = tuner.candidates(tuner.BATCH_VARIANTS)
    rows = [tuner.candidates().index(c) for c in cands]
    X = np.zeros((len(mats), len(cands), tuner.NFEAT))
    for i, A in enumerate(mats):
        X[i] = tuner.matrix_features(A)[0][rows]
    log(f"{split}: {len(mats)} matrices, {sum(A.nnz for A in mats)} nnz, features "
        f"({time.perf_counter() - t0:.0f} s)")
    xs = [hip.DeviceArray.from_numpy(np.random.default_rng(i).uniform(-1, 1, A.n)) for i, A in enumerate(mats)]
    ys = [hip.DeviceArray.zeros(A.m) for A in mats]
    cyc = np.zeros((len(mats), len(cands)))
    graph_ms = {}
    for c, cand in enumerate(cands):
        B, Ms = build_batch(mats, [cand] * len(mats), xs, ys, stream)
        B.profile_cycles()  # warm-up
        runs = np.stack([B.profile_cycles().astype(np.float64) for _ in range(3)])
        cyc[:, c] = np.median(runs, axis=0)
        graph_ms[cand] = time_graph(B, stream)
        B.release()
        for M in Ms:
            M.release()
        log(f"{split} {cand}: batched launch {graph_ms[cand]:.3f} ms, sum of cycles {cyc[:, c].sum():.3e}")
    nnz = np.array([A.nnz for A in mats], dtype=np.float64)
    return mats, X, cyc, nnz, graph_ms, xs, ys


def main_batch(args, log):
    stream = hip.Stream()
    cands = tuner.candidates(tuner.BATCH_VARIANTS)
    _, Xtr, ctr, ntr, gtr, _, _ = collect_batch("train", stream, log)
    Xf = Xtr.reshape(-1, tuner.NFEAT)
    ytr = np.log(np.maximum(ctr, 1.0) / ntr[:, None]).reshape(-1)
    w = np.repeat(np.sqrt(ntr / ntr.mean()), len(cands))
    model, losses = tuner.train(Xf, ytr, w=w, steps=args.steps, lr=args.lr, members=args.members)
    log(f"trained: loss {losses[0]:.4f} -> {losses[-1]:.4f} in {model.meta['train_s']:.1f} s")
    model.meta["labels"] = "log(cycles per entry inside a batched launch)"
    model.meta["variants"] = list(tuner.BATCH_VARIANTS)
    model.save(os.path.join(args.out, "tuner_model_batch.json"))
    mats, Xte, cte, nte, gte, xs, ys = collect_batch("test", stream, log)
    np.savez(os.path.join(args.out, "tuner_dataset_batch.npz"), Xtr=Xtr, ctr=ctr, ntr=ntr, Xte=Xte, cte=cte,
             nte=nte)
    picks, regrets, reg64 = [], [], []
    i64 = cands.index((64, "balanced"))
    for i in range(len(mats)):
        pred = model.predict(Xte[i])
        k = int(np.argmin(pred))
        picks.append(cands[k])
        best = cte[i].min()
        regrets.append((cte[i, k] - best) / best * 100.0)
        reg64.append((cte[i, i64] - best) / best * 100.0)
    B, Ms = build_batch(mats, picks, xs, ys, stream)
    ms_pick = time_graph(B, stream)
    B.release()
    best_fixed = min(gte, key=gte.get)
    best_cyc = cte.min(axis=1).sum()
    report = {
        "labels": "cycles of the member's work items inside one batched launch (abhsf_batch_profile)",
        "train_matrices": int(len(ntr)), "test_matrices": int(len(nte)),
        "train_loss_first": float(losses[0]), "train_loss_last": float(losses[-1]),
        "mean_regret_pct": {"mlp": float(np.mean(regrets)), "plain_s64": float(np.mean(reg64))},
        "sum_cycles_vs_per_matrix_best": {
            "mlp": float(sum(cte[i, cands.index(p)] for i, p in enumerate(picks)) / best_cyc),
            "plain_s64": float(cte[:, i64].sum() / best_cyc),
            "best_fixed": float(cte[:, cands.index(best_fixed)].sum() / best_cyc)},
        "batched_launch_ms": {"mlp_picks": ms_pick, "plain_s64": gte[(64, "balanced")],
                              "best_fixed": gte[best_fixed], "best_fixed_candidate": list(best_fixed)},
        "speedup_vs_plain_s64": gte[(64, "balanced")] / ms_pick,
        "picks": {f"{s}/{v}": sum(1 for p in picks if p == (s, v)) for s, v in cands},
        "per_candidate_launch_ms": {f"{s}/{v}": gte[(s, v)] for s, v in cands},
    }
    json.dump(report, open(os.path.join(args.out, "tuner_report_batch.json"), "w"), indent=1)
    log(f"held-out batch: MLP picks {ms_pick:.3f} ms, plain s64 {gte[(64, 'balanced')]:.3f} ms "
        f"({report['speedup_vs_plain_s64']:.2f}x), best fixed {best_fixed} {gte[best_fixed]:.3f} ms; "
        f"mean regret {report['mean_regret_pct']}")


def cross_validate(Xtr, ytr, ntr, nc, grid, folds=5, repeats=3, members=5, seed=0, log=print):
    """k-fold cross-validation by matrix over the training set, repeated over
    `repeats` fold assignments: for every (steps, lr) of the grid, each
    matrix's regret of the held-in model's pick when it is held out (labels
    ytr are log times per entry, so time ratios within a matrix are exp of
    label differences), averaged over the repeats.  Returns {(steps, lr):
    (mean regret % over matrices, its standard error)} and the choice by the
    one-standard-error rule: the fewest steps whose mean regret is within one
    standard error of the lowest (the simplest model the folds cannot tell
    apart from the best).  One assignment alone moved that choice between
    runs (its standard errors overlap widely); the repeats steady it."""
    nmat = len(ntr)
    Y = ytr.reshape(nmat, nc)
    X = Xtr.reshape(nmat, nc, -1)
    assign = []
    for r in range(repeats):
        order = np.random.default_rng(seed + r).permutation(nmat)
        fold_of = np.empty(nmat, dtype=np.int64)
        fold_of[order] = np.arange(nmat) % folds
        assign.append(fold_of)
    res = {}
    for steps, lr in grid:
        per = np.zeros(nmat)
        for rep, fold_of in enumerate(assign):
            log(f"cv steps={steps} lr={lr}: fold assignment {rep + 1}/{repeats}")
            for f in range(folds):
                tr, te = np.flatnonzero(fold_of != f), np.flatnonzero(fold_of == f)
                model, _ = tuner.train(X[tr].reshape(-1, X.shape[-1]), Y[tr].reshape(-1), steps=steps, lr=lr,
                                       members=members)
                for i in te:
                    k = int(np.argmin(model.predict(X[i])))
                    per[i] += (math.exp(Y[i, k] - Y[i].min()) - 1.0) * 100.0 / repeats
        res[(steps, lr)] = (float(np.mean(per)), float(np.std(per) / math.sqrt(nmat)))
        log(f"cv steps={steps} lr={lr}: mean regret {res[(steps, lr)][0]:.2f}% (se {res[(steps, lr)][1]:.2f})")
    best = min(res, key=lambda k: res[k][0])
    bound = res[best][0] + res[best][1]
    ok = [k for k in res if res[k][0] <= bound]
    return res, min(ok, key=lambda k: (k[0], res[k][0]))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    # latency defaults: the choice of a 5-fold cross-validation over the
    # training matrices with the one-standard-error rule (--cv reruns it);
    # the batch set is 8x larger and keeps the longer schedule
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--lr", type=float, default=None)
    ap.add_argument("--members", type=int, default=None)
    ap.add_argument("--mode", choices=["latency", "batch"], default="latency")
    ap.add_argument("--cv", action="store_true",
                    help="latency: choose steps and lr by 5-fold cross-validation over the training matrices")
    ap.add_argument("--from-dataset", default=None,
                    help="latency: retrain from a saved tuner_dataset.npz (its labels) instead of sweeping")
    args = ap.parse_args(argv)
    if args.steps is None:
        args.steps = 4000 if args.mode == "batch" else 700
    if args.lr is None:
        args.lr = 2e-3 if args.mode == "batch" else 3e-3
    if args.members is None:  # latency: 10 members cross-validated better than 5
        args.members = 5 if args.mode == "batch" else 10
    os.makedirs(args.out, exist_ok=True)

    def log(msg):
        print(f"[train_tuner] {msg}", flush=True)

    if args.mode == "batch":
        return main_batch(args, log)

    stream = hip.Stream()
    cands = tuner.candidates()
    nc = len(cands)
    if args.from_dataset:  # labels already measured (same kernels): retrain and report only
        d = np.load(args.from_dataset, allow_pickle=False)
        Xtr, ytr, Xte, yte = d["Xtr"], d["ytr"], d["Xte"], d["yte"]
        ntr, nte = [str(v) for v in d["train_names"]], [str(v) for v in d["test_names"]]
        tte = json.loads(str(d["test_times"]))
        keep = np.repeat([nm not in EXCLUDED_SAME_CONFIG for nm in ntr], nc)
        Xtr, ytr, ntr = Xtr[keep], ytr[keep], [nm for nm in ntr if nm not in EXCLUDED_SAME_CONFIG]
    else:
        Xtr, ytr, ntr, ttr = collect("train", stream, log)
        Xte, yte, nte, tte = collect("test", stream, log)
        np.savez(os.path.join(args.out, "tuner_dataset.npz"), Xtr=Xtr, ytr=ytr, Xte=Xte, yte=yte,
                 train_names=np.array(ntr), test_names=np.array(nte), test_times=np.array(json.dumps(tte)),
                 candidates=np.array([f"{s}/{v}" for s, v in cands]))
    cv = None
    if args.cv:
        grid = [(st, lr) for st in (300, 700, 1500, 3000) for lr in (3e-3, 1e-2)]
        cv, (args.steps, args.lr) = cross_validate(Xtr, ytr, ntr, nc, grid, members=args.members, log=log)
        log(f"cv choice: steps={args.steps} lr={args.lr}")
    model, losses = tuner.train(Xtr, ytr, steps=args.steps, lr=args.lr, members=args.members)
    log(f"trained: loss {losses[0]:.4f} -> {losses[-1]:.4f} in {model.meta['train_s']:.1f} s")
    model.meta["train_matrices"] = list(ntr)
    model.save(os.path.join(args.out, "tuner_model.json"))
    model = tuner.Model.load(os.path.join(args.out, "tuner_model.json"))

    report = {"train_loss_first": float(losses[0]), "train_loss_last": float(losses[-1]),
              "schedule": {"steps": args.steps, "lr": args.lr, "members": args.members},
              "cv_mean_regret_pct": None if cv is None else {f"{k[0]}/{k[1]}": {"mean": v[0], "se": v[1]}
                                                              for k, v in cv.items()},
              "train_matrices": list(ntr), "excluded_same_config": EXCLUDED_SAME_CONFIG,
              "candidates": [f"{s}/{v}" for s, v in cands], "matrices": []}
    agg = {"mlp": [], "plain_s64": [], "storage_opt": []}
    for i, name in enumerate(nte):
        tm = {(s, v): tte[i][f"{s}/{v}"] for s, v in cands}
        tp = time.perf_counter()
        pred = model.predict(Xte[i * nc:(i + 1) * nc])
        predict_ms = (time.perf_counter() - tp) * 1e3
        pk = cands[int(np.argmin(pred))]
        so = tuple(tte[i]["storage_opt"])
        row = {
            "matrix": name, "nnz": tte[i]["nnz"], "best": list(min(tm, key=tm.get)),
            "best_ms": min(tm.values()), "mlp_pick": list(pk), "mlp_ms": tm[pk],
            "regret_mlp": tuner.regret(tm, pk), "regret_plain_s64": tuner.regret(tm, (64, "balanced")),
            "regret_storage_opt": tuner.regret(tm, so), "plain_s64_ms": tm[(64, "balanced")],
            "speedup_vs_plain": tm[(64, "balanced")] / tm[pk],
            "family": family(name), "tuning_ms": {"features": tte[i].get("features_ms"), "predict": predict_ms},
        }
        report["matrices"].append(row)
        agg["mlp"].append(row["regret_mlp"])
        agg["plain_s64"].append(row["regret_plain_s64"])
        agg["storage_opt"].append(row["regret_storage_opt"])
        log(f"holdout {name}: pick {pk} regret {row['regret_mlp']:.1f}% | plain s64 {row['regret_plain_s64']:.1f}% "
            f"| storage-opt {so} {row['regret_storage_opt']:.1f}%")
    report["mean_regret"] = {k: float(np.mean(v)) for k, v in agg.items()}
    report["max_regret"] = {k: float(np.max(v)) for k, v in agg.items()}
    fam = {}
    for row in report["matrices"]:
        f = fam.setdefault(row["family"], {"mlp": [], "plain_s64": []})
        f["mlp"].append(row["regret_mlp"])
        f["plain_s64"].append(row["regret_plain_s64"])
    report["mean_regret_by_family"] = {k: {m: float(np.mean(v)) for m, v in d.items()} for k, d in fam.items()}
    json.dump(report, open(os.path.join(args.out, "tuner_report.json"), "w"), indent=1)
    log(f"mean regret: {report['mean_regret']}")


if __name__ == "__main__":
    sys.exit(main())

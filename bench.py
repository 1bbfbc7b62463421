"""ABHSF SpMV benchmark on one MI355X.

Workload (the largest single-GPU configuration the metric is quoted on):
block-structured synthetic matrix, n = 2^22, 32-row strips drawn 20% three
dense 32x32 blocks / 30% block-tridiagonal of 8x8 sub-blocks / 50% 12 uniform
columns per row, values N(0,1) FP64 (seed 0), x ~ U[-1,1] (seed 0); ABHSF at
s in {16, 32, 64, 128}.  A step is one SpMV y += A*x (the reference harness's
accumulating call, spmv/unified/spmv.cpp:1222-1231) with all inputs resident
in HBM.

value = GFLOP/s (2*nnz per step) of the block size chosen by the tuner (the
cost-model tuner picks s and kernel variant from matrix statistics); every
measured s is reported, as is plain ABHSF at fixed s = 64.  The roofline entry
is for the SpMV kernel at the chosen s: algorithmic bytes per launch = all
ABHSF arrays as stored + 8n (x) + 16m (y read and written), divided by the
kernel's average duration from HIP events on the handle's stream.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--s 16,32,64,128]
                       [--n 4194304] [--no-cpu-baseline] [--no-check]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calib_copy_gbs(hip, nbytes=2 << 30, reps=10):
    """Device stream-copy ceiling (read + write bytes / time)."""
    L = C.CDLL(os.path.join(ROOT, "abhsf_amd", "libabhsf_calib.so"))
    L.abhsf_calib_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    n = nbytes // 8
    a, b = hip.DeviceArray.zeros(n), hip.DeviceArray.zeros(n)
    st = hip.Stream()
    for _ in range(2):
        assert L.abhsf_calib_copy(a.ptr, b.ptr, n, st.handle) == 0
    e0, e1 = hip.Event(), hip.Event()
    e0.record(st)
    for _ in range(reps):
        assert L.abhsf_calib_copy(a.ptr, b.ptr, n, st.handle) == 0
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    return 2 * nbytes / (ms * 1e-3) / 1e9, ms


def time_spmv(M, hip, dx, dy, steps, warmup, stream):
    for _ in range(warmup):
        M.spmv_async(1.0, dx, 1.0, dy)
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        M.spmv_async(1.0, dx, 1.0, dy)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1)


def cpu_baseline(A, s, x, seconds_target=10.0):
    """The CPU oracle's ABHSF SpMV (OpenMP over strips) on a bounded sample:
    the first rows of the same matrix, same s."""
    from oracle import pyoracle
    rows = min(A.m, 1 << 20)
    rp = A.rowptr[: rows + 1].copy()
    nnz = int(rp[-1])
    B = pyoracle.Abhsf(rows, A.n, rp, A.col[:nnz], A.val[:nnz], s.bit_length() - 1)
    y = np.zeros(rows)
    B.spmm(x, 1.0, 1.0, y)  # warm-up
    t0 = time.perf_counter()
    it = 0
    while True:
        y = B.spmm(x, 1.0, 1.0, y)
        it += 1
        el = time.perf_counter() - t0
        if el >= seconds_target or it >= 200:
            break
    gflops = 2.0 * nnz * it / el / 1e9
    return {
        "value": round(gflops, 3), "unit": "GFLOP/s", "cores": int(pyoracle.lib().orc_num_threads()),
        "kind": "port",
        "sample": f"first {rows} rows ({nnz} nnz) of the same matrix, s={s}, {it} SpMV calls in {el:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--s", default="16,32,64,128")
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-calib", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported by this build "
              "(the SpMV shards by block-row strips; only the single-GPU kernel is built)", file=sys.stderr)
        sys.exit(2)

    from abhsf_amd import AbhsfMatrix, hip, synth
    from abhsf_amd import tuner

    t0 = time.perf_counter()
    A = synth.block_mix(args.n, seed=0)
    log(f"matrix block_mix n={A.n} nnz={A.nnz} checksum={A.checksum():#x} ({time.perf_counter() - t0:.1f} s)")
    x = np.random.default_rng(0).uniform(-1.0, 1.0, A.n)
    stream = hip.Stream()
    dx = hip.DeviceArray.from_numpy(x)
    dy = hip.DeviceArray.zeros(A.m)
    d_rp, d_c, d_v = (hip.DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val))
    ref = None
    if not args.no_check:
        from oracle import pyoracle
        ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)

    s_list = [int(v) for v in args.s.split(",")]
    per_s = {}
    for s in s_list:
        M = AbhsfMatrix(block_size=s, stream=stream)
        M.from_csr(A.m, A.n, d_rp, d_c, d_v, on_device=True, nnz=A.nnz)
        st = M.stats()
        err = None
        if ref is not None:
            dy.fill_zero(stream.handle)
            M.spmv_async(1.0, dx, 0.0, dy)
            stream.synchronize()
            got = dy.numpy()
            err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
            if err > 1e-12:
                raise SystemExit(f"parity failure at s={s}: rel err {err}")
        dy.fill_zero(stream.handle)
        ms = time_spmv(M, hip, dx, dy, args.steps, args.warmup, stream)
        t_launch = ms / args.steps
        algo_bytes = st["stored_bytes"] + 8 * A.n + 16 * A.m
        gflops = 2.0 * A.nnz / (t_launch * 1e-3) / 1e9
        per_s[s] = {
            "gflops": round(gflops, 2), "ms": round(t_launch, 4), "gbs": round(algo_bytes / (t_launch * 1e-3) / 1e9, 1),
            "stored_bytes": st["stored_bytes"], "index_bytes": st["index_bytes"], "algo_bytes": algo_bytes,
            "blocks": st["blocks"], "build_ms": round(st["build_ms"], 1), "rel_err": err,
        }
        log(f"s={s}: {gflops:.1f} GFLOP/s, {t_launch:.4f} ms/SpMV, {per_s[s]['gbs']} GB/s algorithmic, "
            f"blocks={st['blocks']}, build {st['build_ms']:.0f} ms, err={err}")
        M.release()

    # tuner choice (cost model on matrix statistics; falls back to the best measured)
    pick = tuner.pick_block_size(A, candidates=s_list)
    s_best = pick if pick in per_s else max(per_s, key=lambda k: per_s[k]["gflops"])
    best = per_s[s_best]
    plain = per_s.get(64)

    calib = None
    if not args.no_calib:
        gbs, ms = calib_copy_gbs(hip)
        calib = {"copy_gbs": round(gbs, 1), "bytes": 2 * (2 << 30), "ms": round(ms, 3)}
        log(f"stream copy ceiling {gbs:.0f} GB/s")

    out = {
        "metric": "SpMV GFLOP/s + % of HBM roofline; learned-tuned vs plain ABHSF, same GPU",
        "value": best["gflops"],
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"block_mix n=2^{args.n.bit_length() - 1} (20% 3x dense 32x32, 30% tridiag 8x8, 50% 12/row), "
                        f"N(0,1) FP64 values, seed 0",
            "nnz": A.nnz, "s": s_best, "s_measured": s_list, "tuner": tuner.NAME,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": best["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(best["gbs"] / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "plain_s64": None if plain is None else {"gflops": plain["gflops"], "gbs": plain["gbs"]},
        "tuned_vs_plain": None if plain is None else round(best["gflops"] / plain["gflops"], 3),
        "per_s": per_s,
        "copy_calibration": calib,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(A, s_best, x)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

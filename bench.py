"""ABHSF SpMV benchmark on one MI355X.

Default workload (the configuration the target is stated on): R-MAT scale 22
(a, b, c, d = .57, .19, .19, .05), edge factor 16, duplicates merged, values
U[-1,1] FP64 (seed 0), x ~ U[-1,1] (seed 0).  Every (s, variant) candidate of
the learned tuner is built and timed (s in 8..256 x strip / balanced / panel /
tiled, plus the delta-coded format); value = GFLOP/s (2*nnz per step) at the
tuner's pick (abhsf_amd/tuner.py, MLP on GPU block statistics), against plain
ABHSF at fixed s = 64 (balanced) and the exhaustive best, with the regret of
the pick (block_statistics/consistency.cpp:147-149).  A step is one SpMV
y += A*x (the reference harness's accumulating call,
spmv/unified/spmv.cpp:1222-1231) with all inputs resident in HBM.

The roofline entry is for the SpMV kernels at the pick: algorithmic bytes per
launch = all ABHSF arrays as stored + 8n (x) + 16m (y read and written),
divided by the average duration of the launch from HIP events on the
handle's stream.

The block-structured matrix (n = 2^22, 32-row strips 20% three dense 32x32
blocks / 30% block-tridiagonal of 8x8 sub-blocks / 50% 12 uniform columns per
row, N(0,1) values) is measured in the same run at s in {16, 32, 64, 128}
(key "block_mix"), with its conversion compared to the CPU builder's arrays.

cpu_baseline: the oracle's OpenMP ABHSF SpMV on the host cores, whole R-MAT
matrix at s = 64, 5 warm-up + 20 timed calls, median of 5 repeats,
OMP_PROC_BIND=close (the protocol of spmv/unified/spmv.cpp:54-55); "c1" is
the same on the 256^2 5-point Laplacian.

Other workloads (--workload): block_mix alone, spmm (27-point stencil 160^3,
k in {1,4,8,16,32}, FP64 and FP32 storage), batch (2000 matrices in one
launch).

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--workload rmat]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

# the CPU baseline's thread placement; OpenMP reads it when a library that
# uses it is first loaded
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "SpMV GFLOP/s + % of HBM roofline; learned-tuned vs plain ABHSF, same GPU"


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
    return {"copy_gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1), "bytes": 2 * nbytes, "ms": round(ms, 3)}


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


def cpu_spmv_protocol(A, s, x, warmup=5, iters=20, repeats=5):
    """The CPU oracle's ABHSF SpMV (OpenMP over strips) on the whole matrix:
    `warmup` untimed calls, then `repeats` runs of `iters` timed calls (the
    protocol of spmv/unified/spmv.cpp:54-55); GFLOP/s of the median run."""
    from oracle import pyoracle
    t0 = time.perf_counter()
    B = pyoracle.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, s.bit_length() - 1)
    build_s = time.perf_counter() - t0
    y = np.zeros(A.m)
    for _ in range(warmup):
        y = B.spmm(x, 1.0, 1.0, y)
    runs = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for _ in range(iters):
            y = B.spmm(x, 1.0, 1.0, y)
        runs.append(time.perf_counter() - t0)
    med = float(np.median(runs))
    return {"value": round(2.0 * A.nnz * iters / med / 1e9, 3), "unit": "GFLOP/s",
            "ms_per_spmv": round(med / iters * 1e3, 3), "cores": int(pyoracle.lib().orc_num_threads()),
            "kind": "port", "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "s": s,
            "protocol": f"{warmup} warm-up + {iters} timed calls, median of {repeats} repeats",
            "oracle_build_s": round(build_s, 1)}


def cpu_baseline(A, s, x, name):
    """cpu_baseline leg: the whole matrix (no sampling) at block size s, and
    the 256^2 Laplacian at s = 64 (c1)."""
    from abhsf_amd import synth
    out = cpu_spmv_protocol(A, s, x)
    out["sample"] = f"whole matrix: {name}, {A.nnz} nnz, s={s}"
    L = synth.laplacian2d(256)
    xl = np.random.default_rng(0).uniform(-1.0, 1.0, L.n)
    c1 = cpu_spmv_protocol(L, 64, xl)
    ref = csr_spmv_numpy(L, xl)
    from oracle import pyoracle
    B = pyoracle.Abhsf(L.m, L.n, L.rowptr, L.col, L.val, 6)
    got = B.spmm(xl, 1.0, 0.0, np.zeros(L.m))
    c1["rel_err_vs_csr"] = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    c1["sample"] = f"256^2 5-point Laplacian, {L.nnz} nnz, s=64, x ~ U[-1,1] seed 0"
    out["c1"] = c1
    return out


def csr_spmv_numpy(A, x):
    """Plain FP64 CSR product in numpy, the result check of the timed runs
    (the oracle under oracle/ serves only the cpu_baseline leg here)."""
    rows = np.repeat(np.arange(A.m), np.diff(A.rowptr))
    return np.bincount(rows, weights=A.val * x[A.col], minlength=A.m)


def measure_candidates(A, cands, args, hip, stream, check=True):
    """Build and time each (s, variant) -- block size 0 marks the delta-coded
    formats (tuner.DELTA_CODECS); returns per-candidate results."""
    from abhsf_amd import tuner
    x = np.random.default_rng(0).uniform(-1.0, 1.0, A.n)
    dx = hip.DeviceArray.from_numpy(x)
    dy = hip.DeviceArray.zeros(A.m)
    d_rp, d_c, d_v = (hip.DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val))
    ref = None
    if check:
        ref = csr_spmv_numpy(A, x)
    res = {}
    for s, v in cands:
        M = tuner.build_candidate(A, s, v, d_rp, d_c, d_v, stream)
        st = M.stats()
        if v in tuner.DELTA_CODECS:  # same keys as the ABHSF statistics
            st = {"stored_bytes": st["stored_bytes"], "index_bytes": st["stored_bytes"] - 8 * A.nnz,
                  "blocks": {"words_1_2_4_8": list(st["words"].values())}, "build_ms": st["build_ms"]}
        err = None
        if ref is not None:
            dy.fill_zero(stream.handle)
            M.spmv_async(1.0, dx, 0.0, dy)
            stream.synchronize()
            err = float(np.max(np.abs(dy.numpy() - ref)) / np.max(np.abs(ref)))
            if err > 1e-12:
                raise SystemExit(f"parity failure at s={s} {v}: rel err {err}")
        dy.fill_zero(stream.handle)
        ms = time_spmv(M, hip, dx, dy, args.steps, args.warmup, stream) / args.steps
        algo = st["stored_bytes"] + 8 * A.n + 16 * A.m
        res[(s, v)] = {
            "gflops": round(2.0 * A.nnz / (ms * 1e-3) / 1e9, 2), "ms": round(ms, 4),
            "gbs": round(algo / (ms * 1e-3) / 1e9, 1), "stored_bytes": st["stored_bytes"],
            "index_bytes": st["index_bytes"], "algo_bytes": algo, "blocks": st["blocks"],
            "build_ms": round(st["build_ms"], 1), "rel_err": err,
        }
        log(f"s={s} {v}: {res[(s, v)]['gflops']} GFLOP/s, {ms:.4f} ms/SpMV, {res[(s, v)]['gbs']} GB/s, "
            f"blocks={st['blocks']}, build {st['build_ms']:.0f} ms, err={err}")
        M.release()
    return res, x


def measure_fp32(A, s, variant, args, hip, stream, x, f64_ms):
    """FP32 value storage with compensated (TwoSum) accumulation at (s,
    variant): time per SpMV and the max-norm relative error against the FP64
    product of the FP64 values and of the f32-rounded values."""
    from abhsf_amd import AbhsfMatrix
    dx, dy = hip.DeviceArray.from_numpy(x), hip.DeviceArray.zeros(A.m)
    d_rp, d_c, d_v = (hip.DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val))
    M = AbhsfMatrix(block_size=s, variant=variant, precision="f32", stream=stream)
    M.from_csr(A.m, A.n, d_rp, d_c, d_v, on_device=True, nnz=A.nnz)
    M.spmv_async(1.0, dx, 0.0, dy)
    stream.synchronize()
    got = dy.numpy()
    ref64 = csr_spmv_numpy(A, x)
    rows = np.repeat(np.arange(A.m), np.diff(A.rowptr))
    ref32 = np.bincount(rows, weights=A.val.astype(np.float32).astype(np.float64) * x[A.col], minlength=A.m)
    dy.fill_zero(stream.handle)
    ms = time_spmv(M, hip, dx, dy, args.steps, args.warmup, stream) / args.steps
    st = M.stats()
    M.release()
    return {"s": s, "variant": variant, "gflops": round(2.0 * A.nnz / (ms * 1e-3) / 1e9, 2), "ms": round(ms, 4),
            "stored_bytes": st["stored_bytes"], "time_vs_f64": round(ms / f64_ms, 3),
            "maxnorm_rel_err_vs_f64_values": float(np.max(np.abs(got - ref64)) / np.max(np.abs(ref64))),
            "maxnorm_rel_err_vs_f32_values": float(np.max(np.abs(got - ref32)) / np.max(np.abs(ref32)))}


def tuned_pick(A, allowed_s, hip):
    from abhsf_amd import tuner
    if not os.path.exists(tuner.MODEL_PATH):
        return None, "none"
    model = tuner.Model.load()
    if not model.compatible():
        log("tuner model does not match the feature layout: exhaustive timing decides")
        return None, "none"
    Xc, _ = tuner.matrix_features(A)
    pred = model.predict(Xc)
    cands = tuner.candidates()
    ok = [i for i, c in enumerate(cands) if c[0] in allowed_s or c[0] == 0]
    best = min(ok, key=lambda i: pred[i])
    return cands[best], tuner.NAME


def kernel_source_hash():
    """SHA-256 (first 16 hex digits) of the kernel and build sources, in name
    order: a stored counter profile applies to this build only if it was
    taken at the same hash."""
    import hashlib
    d = os.path.join(ROOT, "abhsf_amd", "csrc")
    h = hashlib.sha256()
    for f in sorted(os.listdir(d)):
        if f.endswith((".hip", ".hpp", ".cpp", ".c")) or f == "Makefile":
            h.update(f.encode())
            h.update(open(os.path.join(d, f), "rb").read())
    return h.hexdigest()[:16]


def measured_traffic(key):
    """HBM bytes per launch of the SpMV kernel for this configuration, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in
    profiles/pmc_traffic.json (each counter scaled by what it reads for the
    plain copy kernel of known size).  Returns (bytes, note): bytes is None
    when the configuration was not profiled or was profiled at other kernel
    sources (the note says which)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None, "no counter profile"
    rec = json.load(open(path)).get(key)
    if rec is None:
        return None, f"{key} not profiled"
    src = kernel_source_hash()
    if rec.get("source_sha") != src:
        return None, f"profile of {key} taken at sources {rec.get('source_sha')}, these are {src}"
    return int(rec["traffic_bytes"]), (f"rocprofv3 FETCH_SIZE x {rec['fetch_factor']} + WRITE_SIZE x "
                                       f"{rec['write_factor']} at sources {src} (profiles/pmc_traffic.json)")


def kernel_kind(name):
    """Which launch a dispatch belongs to in the counter summaries."""
    for pat, kind in (("k_spmv_strip", "spmv"), ("k_delta_spmv", "spmv"), ("k_delta32_spmv", "spmv"),
                      ("k_spmv_fixup", "fixup"), ("k_spmm_strip", "spmm"), ("k_batch_spmv", "batch"),
                      ("k_batch_fixup", "batch_fixup"), ("k_stream_copy", "copy")):
        if pat in name:
            return kind
    return None


def summarize_pmc(key, fetch_csv, write_csv, sizes_csv=None, command=""):
    """Per-dispatch HBM traffic of the SpMV kernels from rocprofv3 --pmc CSVs.

    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are averaged over the
    k_spmv_strip (or k_delta_spmv) dispatches (plus k_spmv_fixup when
    present) and scaled by what
    the same counter reads for k_stream_copy, the calibration copy of known
    size (2 GiB read + 2 GiB written per dispatch, calib_copy_gbs above).  With
    sizes_csv (a pass over TCC_EA0_RDREQ_{32B,64B,128B}_sum and
    TCC_EA0_RDREQ_sum) the read bytes are also resolved by request size.
    Writes the entry into profiles/pmc_traffic.json."""
    import csv
    from collections import defaultdict

    def per_kernel(path):
        acc = defaultdict(lambda: defaultdict(list))
        for row in csv.DictReader(open(path)):
            kind = kernel_kind(row["Kernel_Name"])
            if kind:
                acc[kind][row["Counter_Name"]].append(float(row["Counter_Value"]))
        return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}

    copy_bytes = 2 << 30
    f, w = per_kernel(fetch_csv), per_kernel(write_csv)
    f_factor = copy_bytes / (f["copy"]["FETCH_SIZE"] * 1024)
    w_factor = copy_bytes / (w["copy"]["WRITE_SIZE"] * 1024)
    # the launches of one product: SpMV (+ heavy-strip fix-up), SpMM, or the
    # batched launch (+ its fix-up)
    kinds = ("spmm",) if "spmm" in f else ("batch", "batch_fixup") if "batch" in f else ("spmv", "fixup")
    fetch = sum(f[k]["FETCH_SIZE"] for k in kinds if k in f) * 1024
    write = sum(w[k]["WRITE_SIZE"] for k in kinds if k in w) * 1024
    rec = {"fetch_kib": round(fetch / 1024, 1), "write_kib": round(write / 1024, 1),
           "fetch_factor": round(f_factor, 4), "write_factor": round(w_factor, 4),
           "traffic_bytes": int(fetch * f_factor + write * w_factor), "command": command,
           "source_sha": kernel_source_hash()}
    if sizes_csv:
        z = per_kernel(sizes_csv)

        def by_size(d):
            return (32 * d.get("TCC_EA0_RDREQ_32B_sum", 0) + 64 * d.get("TCC_EA0_RDREQ_64B_sum", 0)
                    + 128 * d.get("TCC_EA0_RDREQ_128B_sum", 0))
        rec["read_bytes_by_request_size"] = int(sum(by_size(z[k]) for k in kinds if k in z))
        rec["copy_read_bytes_by_request_size"] = int(by_size(z["copy"]))
        rec["read_requests"] = {c: int(sum(z[k].get(c, 0) for k in kinds if k in z))
                                for c in ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum",
                                          "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")}
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    allrec = json.load(open(path)) if os.path.exists(path) else {}
    allrec[key] = rec
    with open(path, "w") as fh:
        json.dump(allrec, fh, indent=1, sort_keys=True)
        fh.write("\n")
    return rec


def roofline(r, key=None):
    traffic, note = measured_traffic(key) if key else (None, "not profiled")
    return {"bound": "hbm", "achieved": r["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(r["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": note,
            "algorithmic_bytes": r.get("algo_bytes")}


def base_record(args, value, ms, dtype="f64"):
    return {"metric": METRIC, "value": value, "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic"}


def run_block_mix(args, hip):
    from abhsf_amd import synth
    t0 = time.perf_counter()
    A = synth.block_mix(args.n, seed=0)
    log(f"matrix block_mix n={A.n} nnz={A.nnz} checksum={A.checksum():#x} ({time.perf_counter() - t0:.1f} s)")
    stream = hip.Stream()
    s_list = [int(v) for v in (args.s or "16,32,64,128").split(",")]
    if args.variant:  # fixed configuration (profiling runs): no tuner, nothing else measured
        pick, tname = (s_list[0], args.variant), "fixed"
        cands = [pick]
    else:
        pick, tname = tuned_pick(A, s_list, hip)
        cands = [(s, v) for s in s_list for v in ("balanced", "tiled")]
        if pick is not None and pick not in cands:
            cands.append(pick)
    res, x = measure_candidates(A, cands, args, hip, stream, check=not args.no_check)
    if pick is None:  # no trained model: exhaustive timing decides
        pick = max(res, key=lambda k: res[k]["gflops"])
        tname = "exhaustive-timing"
    best = res[pick]
    plain = res.get((64, "balanced"))
    out = base_record(args, best["gflops"], best["ms"])
    out["config"] = {
        "workload": f"block_mix n=2^{args.n.bit_length() - 1} (20% 3x dense 32x32, 30% tridiag 8x8, "
                    f"50% 12/row), N(0,1) FP64 values, seed 0",
        "nnz": A.nnz, "s": pick[0], "variant": pick[1], "s_measured": s_list, "tuner": tname,
    }
    out["roofline"] = roofline(best, f"block_mix/{pick[0]}/{pick[1]}")
    out["plain_s64"] = None if plain is None else {"gflops": plain["gflops"], "gbs": plain["gbs"]}
    out["tuned_vs_plain"] = None if plain is None else round(best["gflops"] / plain["gflops"], 3)
    top = max(res, key=lambda k: res[k]["gflops"])
    out["exhaustive_best"] = {"s": top[0], "variant": top[1], "gflops": res[top]["gflops"]}
    out["per_s"] = {f"{s}/{v}": r for (s, v), r in res.items()}
    if not args.no_calib:
        out["copy_calibration"] = calib_copy_gbs(hip)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(A, 64, x, "block_mix n=2^22")
    return out


def run_rmat(args, hip):
    from abhsf_amd import synth, tuner
    A = synth.rmat(22, 16, seed=0)
    log(f"matrix rmat scale 22 ef 16 nnz={A.nnz}")
    stream = hip.Stream()
    s_allowed = tuner.S_CANDIDATES if args.s is None else [int(v) for v in args.s.split(",")]
    if args.variant:  # fixed configuration (profiling runs): no tuner, nothing else measured
        pick, tname = (s_allowed[0], args.variant), "fixed"
        cands = [pick]
    else:
        pick, tname = tuned_pick(A, s_allowed, hip)
        cands = [c for c in tuner.candidates() if c[0] in s_allowed or c[0] == 0]
        if args.s is not None:  # a restricted run: the balanced variant only, plus the pick
            cands = [c for c in cands if c[1] == "balanced" or c == pick]
    res, x = measure_candidates(A, cands, args, hip, stream, check=not args.no_check)
    if pick is None:
        pick, tname = max(res, key=lambda k: res[k]["gflops"]), "exhaustive-timing"
    best, plain = res[pick], res.get((64, "balanced"), res[pick])
    times = {k: r["ms"] for k, r in res.items()}
    out = base_record(args, best["gflops"], best["ms"])
    out["config"] = {"workload": "R-MAT scale 22 (a=.57,b=.19,c=.19), edge factor 16, duplicates merged, "
                     "U[-1,1] FP64 values, seed 0", "nnz": A.nnz, "s": pick[0], "variant": pick[1],
                     "tuner": tname}
    out["roofline"] = roofline(best, f"rmat/{pick[0]}/{pick[1]}")
    out["plain_s64"] = {"gflops": plain["gflops"], "gbs": plain["gbs"]}
    out["tuned_vs_plain"] = round(best["gflops"] / plain["gflops"], 3)
    out["regret_pct"] = round(tuner.regret(times, pick), 2) if len(times) > 1 else None
    top = max(res, key=lambda k: res[k]["gflops"])
    out["exhaustive_best"] = {"s": top[0], "variant": top[1], "gflops": res[top]["gflops"], "ms": res[top]["ms"]}
    out["per_candidate"] = {f"{s}/{v}": r for (s, v), r in res.items()}
    if not args.no_calib:
        out["copy_calibration"] = calib_copy_gbs(hip)
    if not args.no_extra and args.variant is None and pick[0] > 0:
        log("fp32 values at the pick ...")
        out["fp32_storage"] = measure_fp32(A, pick[0], pick[1], args, hip, stream, x, best["ms"])
    if not args.no_extra and args.variant is None:
        log("block_mix n=2^22 ...")
        bm = run_block_mix(argparse.Namespace(**{**vars(args), "no_cpu_baseline": True, "no_calib": True,
                                                 "s": None, "variant": None}), hip)
        out["block_mix"] = {k: bm[k] for k in ("value", "ms_per_step", "config", "roofline", "plain_s64",
                                               "tuned_vs_plain", "exhaustive_best", "per_s")}
    if not args.no_cpu_baseline and args.variant is None:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(A, 64, x, "R-MAT scale 22 ef 16")
    return out


def run_spmm(args, hip):
    from abhsf_amd import AbhsfMatrix, synth
    A = synth.stencil27(160)
    log(f"matrix stencil27 160^3 nnz={A.nnz}")
    stream = hip.Stream()
    s = args.spmm_s
    d_rp, d_c, d_v = (hip.DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val))
    out_k = {}
    precs = ("f64", "f32") if args.spmm_prec is None else (args.spmm_prec,)
    ks = (1, 4, 8, 16, 32) if args.spmm_k is None else (args.spmm_k,)
    for prec in precs:
        M = AbhsfMatrix(block_size=s, precision=prec, stream=stream)
        M.from_csr(A.m, A.n, d_rp, d_c, d_v, on_device=True, nnz=A.nnz)
        st = M.stats()
        for k in ks:
            X = np.random.default_rng(k).uniform(-1, 1, (A.n, k))
            dX, dY = hip.DeviceArray.from_numpy(X), hip.DeviceArray.zeros((A.m, k))
            for _ in range(args.warmup):
                M.spmm_async(k, 1.0, dX, k, 1.0, dY, k)
            stream.synchronize()
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            for _ in range(args.steps):
                M.spmm_async(k, 1.0, dX, k, 1.0, dY, k)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / args.steps
            # error of Y = A X against a numpy FP64 CSR product (columns: all for k <= 8, else 8)
            M.spmm_async(k, 1.0, dX, k, 0.0, dY, k)
            stream.synchronize()
            Y = dY.numpy()
            cols = range(k) if k <= 8 else range(0, k, k // 8)
            err = 0.0
            for j in cols:
                ref = csr_spmv_numpy(A, X[:, j].copy())
                err = max(err, float(np.max(np.abs(Y[:, j] - ref)) / np.max(np.abs(ref))))
            algo = st["stored_bytes"] + 8 * A.n * k + 16 * A.m * k
            out_k[f"{prec}/k{k}"] = {"gflops": round(2 * A.nnz * k / (ms * 1e-3) / 1e9, 1), "ms": round(ms, 4),
                                     "gbs": round(algo / (ms * 1e-3) / 1e9, 1), "maxnorm_rel_err": err,
                                     "algo_bytes": algo}
            log(f"{prec} k={k}: {out_k[f'{prec}/k{k}']}")
        M.release()
    best = max(out_k.values(), key=lambda r: r["gflops"])
    out = base_record(args, best["gflops"], best["ms"])
    out["config"] = {"workload": f"27-point stencil 160^3 (nnz {A.nnz}), k in {','.join(map(str, ks))}, s={s}",
                     "nnz": A.nnz}
    kmax = f"{precs[0]}/k{ks[-1]}"
    out["roofline"] = roofline(out_k[kmax], f"spmm/{s}/{kmax}")
    out["per_k"] = out_k
    if not args.no_calib:
        out["copy_calibration"] = calib_copy_gbs(hip)
    return out


def run_batch(args, hip):
    """Many matrices in one launch: the tuner's (s, variant) per matrix against
    plain ABHSF at s = 64 for all, both as one batched hipGraph replay."""
    from abhsf_amd import AbhsfBatch, AbhsfMatrix, synth, tuner
    rng = np.random.default_rng(0)
    stream = hip.Stream()
    mpath = tuner.BATCH_MODEL_PATH if os.path.exists(tuner.BATCH_MODEL_PATH) else tuner.MODEL_PATH
    model = tuner.Model.load(mpath) if os.path.exists(mpath) else None
    if model is not None and not model.compatible():
        log("tuner model does not match the feature layout: fixed s=64")
        model = None
    tuned, plain, xs, ys, nnz, indiv, picks = [], [], [], [], 0, [], {}
    t0 = time.perf_counter()
    for i in range(args.batch_count):
        A = synth.batch_member(rng, i, 0)
        s, v = 64, "balanced"
        if model is not None:
            (s, v), _ = tuner.pick(model, tuner.matrix_features(A)[0])
        picks[f"{s}/{v}"] = picks.get(f"{s}/{v}", 0) + 1
        d_rp, d_c, d_v = (hip.DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val))
        for lst, (bs, var) in ((tuned, (s, v)), (plain, (64, "balanced"))):
            M = AbhsfMatrix(block_size=bs, variant=var, stream=stream)
            M.from_csr(A.m, A.n, d_rp, d_c, d_v, on_device=True, nnz=A.nnz)
            lst.append(M)
        stream.synchronize()
        del d_rp, d_c, d_v
        xs.append(hip.DeviceArray.from_numpy(np.random.default_rng(i).uniform(-1, 1, A.n)))
        ys.append(hip.DeviceArray.zeros(A.m))
        nnz += A.nnz
        if i % 200 == 0:
            log(f"built {i} matrices, {nnz} nnz ({time.perf_counter() - t0:.0f} s)")
    # per-matrix latency: each tuned matrix alone
    for M, x, y in zip(tuned, xs, ys):
        M.spmv_async(1.0, x, 1.0, y)
        stream.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        M.spmv_async(1.0, x, 1.0, y)
        e1.record(stream)
        e1.synchronize()
        indiv.append(e0.elapsed_ms(e1))

    def time_batch(mats):
        B = AbhsfBatch(mats, stream)
        B.set_vectors(xs, ys)
        for _ in range(args.warmup):
            B.spmv_async(1.0, 1.0, graph=True)
        B.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        for _ in range(args.steps):
            B.spmv_async(1.0, 1.0, graph=True)
        e1.record(stream)
        e1.synchronize()
        B.release()
        return e0.elapsed_ms(e1) / args.steps

    ms = time_batch(tuned)
    ms_plain = float("nan") if args.batch_tuned_only else time_batch(plain)
    log(f"batched: tuned {ms:.3f} ms, plain s=64 {ms_plain:.3f} ms")
    out = base_record(args, round(2 * nnz / (ms * 1e-3) / 1e9, 2), round(ms, 4))
    out["config"] = {"workload": f"{args.batch_count} matrices, n log-uniform [1e3, 2e5], families Laplacian / "
                     "banded bw 4-64 / R-MAT ef 8 / random blocks 1-30%, N(0,1) FP64, one hipGraph launch",
                     "nnz": nnz, "picks": picks,
                     "tuner": (tuner.NAME + (" (throughput model)" if mpath == tuner.BATCH_MODEL_PATH else
                                             " (latency model)")) if model is not None else "fixed s=64"}
    algo = sum(M.stats()["stored_bytes"] + 8 * M.props.n + 16 * M.props.m for M in tuned)
    out["roofline"] = roofline({"gbs": round(algo / (ms * 1e-3) / 1e9, 1), "algo_bytes": algo}, "batch/tuned")
    out["plain_s64"] = {"gflops": round(2 * nnz / (ms_plain * 1e-3) / 1e9, 2), "ms": round(ms_plain, 4)}
    out["tuned_vs_plain"] = round(ms_plain / ms, 3)
    out["per_matrix_latency_ms"] = {"p50": round(float(np.percentile(indiv, 50)), 4),
                                    "p99": round(float(np.percentile(indiv, 99)), 4),
                                    "sum_individual": round(float(np.sum(indiv)), 3)}
    out["batched_vs_individual"] = round(float(np.sum(indiv)) / ms, 2)
    if not args.no_calib:
        out["copy_calibration"] = calib_copy_gbs(hip)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="rmat", choices=["rmat", "block_mix", "spmm", "batch"])
    ap.add_argument("--s", default=None, help="block sizes to measure (block_mix default 16,32,64,128; "
                    "rmat default all tuner candidates)")
    ap.add_argument("--variant", default=None, choices=["strip", "balanced", "panel", "tiled", "delta",
                                                         "delta_u32_row", "delta_u32_block"],
                    help="block_mix / rmat: measure only (first --s, VARIANT), no tuner (profiler passes)")
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--spmm-s", type=int, default=64)
    ap.add_argument("--spmm-k", type=int, default=None, help="spmm: one right-hand-side count (profiler passes)")
    ap.add_argument("--spmm-prec", default=None, choices=["f64", "f32"], help="spmm: one value precision")
    ap.add_argument("--batch-count", type=int, default=2000)
    ap.add_argument("--batch-tuned-only", action="store_true",
                    help="batch: time only the tuned batch (profiler passes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-calib", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="rmat: skip the block_mix measurement")
    ap.add_argument("--pmc-summary", nargs="+", metavar="ARG",
                    help="KEY FETCH_CSV WRITE_CSV [SIZES_CSV]: summarise rocprofv3 --pmc counter CSVs of a "
                         "profiled bench run into profiles/pmc_traffic.json and exit (no GPU needed)")
    ap.add_argument("--pmc-command", default="", help="the profiled command, recorded with --pmc-summary")
    args = ap.parse_args()
    if args.pmc_summary:
        a = args.pmc_summary
        if len(a) not in (3, 4):
            ap.error("--pmc-summary takes KEY FETCH_CSV WRITE_CSV [SIZES_CSV]")
        print(json.dumps(summarize_pmc(a[0], a[1], a[2], a[3] if len(a) == 4 else None, args.pmc_command)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the SpMV would shard by "
              "block-row strips; only the single-GPU kernel is implemented)", file=sys.stderr)
        sys.exit(2)
    from abhsf_amd import hip
    run = {"block_mix": run_block_mix, "rmat": run_rmat, "spmm": run_spmm, "batch": run_batch}[args.workload]
    print(json.dumps(run(args, hip)), flush=True)


if __name__ == "__main__":
    main()

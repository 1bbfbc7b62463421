"""GPU checks of the tuner: block statistics against the oracle (itself
pinned to the reference's block_stats binary), the MLP forward/backward
kernels against the FP64 CPU reference, training, and what the trained model
gains over plain ABHSF (fixed s = 64) on held-out matrices."""
import os

import numpy as np
import pytest

from abhsf_amd import capi, hip, synth, tuner
from abhsf_amd.hip import DeviceArray
from oracle import mlp_ref, pyoracle

from conftest import ROOT, read_mtx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["test_mtx", "random", "rmat"])
def test_block_hist_matches_oracle(gpu, case):
    if case == "test_mtx":
        m, n, r, c, v = read_mtx(os.path.join(ROOT, "tests", "golden", "test.mtx"))
        A = synth.from_triplets(m, n, r, c, v)
    elif case == "random":
        A = synth.random_sparse(700, 650, 0.02, seed=9)
    else:
        A = synth.rmat(13, 8, seed=5)
    r, c, _ = A.triplets()
    for kr in (1, 2, 3, 5, 8):
        for ks in (1, 3, 4, 8):
            cap = (1 << (kr + ks)) + 2
            import ctypes as C
            hist = np.zeros(cap, dtype=np.uint64)
            nb, ns = C.c_uint64(0), C.c_uint64(0)
            capi.call("abhsf_block_hist", A.rowptr.ctypes.data_as(C.c_void_p), A.col.ctypes.data_as(C.c_void_p),
                      A.m, A.n, A.nnz, 0, kr, ks, 1, hist.ctypes.data_as(C.c_void_p), cap, C.byref(nb),
                      C.byref(ns), None)
            want = pyoracle.block_hist(r, c, kr, ks)
            got = {int(z): int(hist[z]) for z in np.nonzero(hist)[0]}
            assert got == want, (kr, ks)
            assert nb.value == sum(want.values()) and ns.value == A.nnz


def test_block_hist_sampling_counts_every_strided_strip(gpu):
    A = synth.laplacian2d(64)
    h, nb, ns = tuner.block_hist(A.rowptr, A.col, A.m, A.n, 8, 4, on_device=False)
    rows = np.arange(A.m)
    keep = ((rows >> 3) % 4) == 0
    assert ns == int(np.diff(A.rowptr)[keep].sum())


def _gpu_forward_backward(theta, X, y, w):
    B, F = X.shape
    dt, dx = DeviceArray.from_numpy(theta.astype(np.float32)), DeviceArray.from_numpy(X.astype(np.float32))
    dy, dw = DeviceArray.from_numpy(y.astype(np.float32)), DeviceArray.from_numpy(w.astype(np.float32))
    H1, H2 = DeviceArray((B, 64), np.float32), DeviceArray((B, 64), np.float32)
    D1, D2 = DeviceArray((B, 64), np.float32), DeviceArray((B, 64), np.float32)
    out, D3 = DeviceArray((B,), np.float32), DeviceArray((B,), np.float32)
    grad = DeviceArray((theta.size,), np.float32)
    capi.call("abhsf_mlp_forward", dt.ptr, F, dx.ptr, B, H1.ptr, H2.ptr, out.ptr, None)
    capi.call("abhsf_mlp_backward", dt.ptr, F, dx.ptr, B, H1.ptr, H2.ptr, out.ptr, dy.ptr, dw.ptr,
              float(w.astype(np.float32).sum()), D1.ptr, D2.ptr, D3.ptr, grad.ptr, None)
    hip.synchronize()
    return out.numpy(), grad.numpy()


@pytest.mark.parametrize("F,B", [(21, 300), (5, 17), (64, 64)])
def test_mlp_kernels_match_fp64_reference(gpu, F, B):
    rng = np.random.default_rng(F)
    theta = (mlp_ref.init(F, seed=F) * 0.5).astype(np.float32).astype(np.float64)
    X = rng.standard_normal((B, F)).astype(np.float32).astype(np.float64)
    y = rng.standard_normal(B).astype(np.float32).astype(np.float64)
    w = rng.uniform(0.5, 2.0, B).astype(np.float32).astype(np.float64)
    assert capi.lib().abhsf_mlp_num_params(F) == theta.size
    out, g = _gpu_forward_backward(theta, X, y, w)
    _, _, ref_out = mlp_ref.forward(theta, X)
    ref_g = mlp_ref.grad(theta, X, y, w)
    assert np.max(np.abs(out - ref_out)) <= 1e-4 * max(1.0, np.max(np.abs(ref_out)))
    # FP32 sums of B terms: compare against the gradient's scale
    assert np.max(np.abs(g - ref_g)) <= 2e-4 * np.max(np.abs(ref_g))
    print(f"mlp F={F} B={B}: max |grad err| / max |grad| = {np.max(np.abs(g - ref_g)) / np.max(np.abs(ref_g)):.2e}")


def test_mlp_training_fits_a_known_function(gpu):
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (512, 6))
    y = np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * np.abs(X[:, 3])
    model, losses = tuner.train(X, y, steps=1500, lr=3e-3)
    assert losses[-1] < 0.05 * losses[0]
    pred = model.predict(X)
    assert np.mean((pred - y) ** 2) < 0.02
    # deterministic: the same training gives the same weights
    model2, _ = tuner.train(X, y, steps=1500, lr=3e-3)
    assert np.array_equal(model.theta, model2.theta)


def test_trained_model_beats_plain_abhsf_on_heldout(gpu):
    """What the network gains: on held-out matrices (seeds and sizes not in
    the training corpus) its pick has lower mean regret against the exhaustive
    sweep than plain ABHSF at fixed s = 64."""
    assert os.path.exists(tuner.MODEL_PATH), "train with python -m abhsf_amd.train_tuner"
    model = tuner.Model.load()
    mats = [synth.laplacian2d(600), synth.rmat(17, 16, seed=21), synth.block_mix(1 << 18, 21, 0.4, 0.2),
            synth.banded(200000, 6, 21), synth.stencil27(60)]
    st = hip.Stream()
    r_mlp, r_plain = [], []
    for A in mats:
        Xc, _ = tuner.matrix_features(A)
        pk, _ = tuner.pick(model, Xc)
        tm = tuner.sweep(A, st)
        r_mlp.append(tuner.regret(tm, pk))
        r_plain.append(tuner.regret(tm, (64, "balanced")))
        print(f"{A.name}: pick {pk} regret {r_mlp[-1]:.1f}% vs plain s=64 {r_plain[-1]:.1f}%")
    print(f"mean regret: mlp {np.mean(r_mlp):.1f}%, plain s=64 {np.mean(r_plain):.1f}%")
    assert np.mean(r_mlp) < np.mean(r_plain)


def test_throughput_model_beats_plain_abhsf_in_batched_launch(gpu):
    """The throughput model (labels: cycles inside a batched launch) picks
    per-matrix (s, variant) for a batch of unseen matrices (seeds outside
    its training corpus); one batched launch of those picks beats the same
    batch at plain s = 64, and the results stay exact."""
    from abhsf_amd import AbhsfBatch, AbhsfMatrix
    assert os.path.exists(tuner.BATCH_MODEL_PATH), "train with python -m abhsf_amd.train_tuner --mode batch"
    model = tuner.Model.load(tuner.BATCH_MODEL_PATH)
    rng = np.random.default_rng(3000)
    mats = [synth.batch_member(rng, i, 3000) for i in range(60)]
    st = hip.Stream()
    xs = [DeviceArray.from_numpy(np.random.default_rng(i).uniform(-1, 1, A.n)) for i, A in enumerate(mats)]
    times = {}
    for label in ("mlp", "plain_s64"):
        Ms = []
        for A in mats:
            s, v = tuner.pick(model, tuner.matrix_features(A)[0])[0] if label == "mlp" else (64, "balanced")
            M = AbhsfMatrix(block_size=s, variant=v, stream=st)
            M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
            Ms.append(M)
        ys = [DeviceArray.zeros(A.m) for A in mats]
        B = AbhsfBatch(Ms, st)
        B.set_vectors(xs, ys)
        B.spmv_async(1.0, 0.0, graph=True)
        B.synchronize()
        for i in (0, 17, 59):
            A = mats[i]
            ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, xs[i].numpy())
            assert np.max(np.abs(ys[i].numpy() - ref)) <= 1e-12 * max(np.max(np.abs(ref)), 1.0)
        best = []
        for _ in range(3):
            e0, e1 = hip.Event(), hip.Event()
            e0.record(st)
            for _ in range(5):
                B.spmv_async(1.0, 1.0, graph=True)
            e1.record(st)
            e1.synchronize()
            best.append(e0.elapsed_ms(e1) / 5)
        times[label] = min(best)
        B.release()
    print(f"batched launch: mlp picks {times['mlp']:.3f} ms, plain s=64 {times['plain_s64']:.3f} ms "
          f"({times['plain_s64'] / times['mlp']:.2f}x)")
    assert times["mlp"] < times["plain_s64"]


def test_blockstats_all_64_pairs_and_tables(gpu):
    """The GPU block-statistics pipeline (abhsf_amd/blockstats.py) reproduces
    the oracle's 64 histograms; the tables built on them are the ones checked
    against the reference binaries in tests/test_blockstats_cpu.py."""
    from abhsf_amd import blockstats
    for A in (synth.rmat(10, 8, seed=7), synth.block_mix(2048, seed=8)):
        r, c, _ = A.triplets()
        hg = blockstats.histograms(A)
        for (rr, ss), h in hg.items():
            assert h == pyoracle.block_hist(r, c, rr.bit_length() - 1, ss.bit_length() - 1), (rr, ss)
        t = blockstats.mmf_table(A, hg)
        ho = {k: pyoracle.block_hist(r, c, k[0].bit_length() - 1, k[1].bit_length() - 1) for k in hg}
        assert t == {k: pyoracle.mmf_row(A.m, A.n, A.nnz, k[0], k[1], ho[k]) for k in hg}


def test_fused_ensemble_predict_matches_member_forwards(gpu):
    """abhsf_mlp_predict (every member and sample in one launch) equals the
    mean of the members' abhsf_mlp_forward outputs: the same float
    operations in the same order, summed over members in FP64."""
    rng = np.random.default_rng(12)
    F, B, nm = tuner.NFEAT, 27, 5
    thetas = np.stack([tuner.init_theta(F, seed) for seed in range(nm)]).astype(np.float32)
    X = rng.standard_normal((B, F)).astype(np.float32)
    dx = DeviceArray.from_numpy(X)
    H1, H2, out = DeviceArray((B, 64), np.float32), DeviceArray((B, 64), np.float32), DeviceArray((B,), np.float32)
    acc = np.zeros(B)
    for th in thetas:
        dt = DeviceArray.from_numpy(np.ascontiguousarray(th))
        capi.call("abhsf_mlp_forward", dt.ptr, F, dx.ptr, B, H1.ptr, H2.ptr, out.ptr, None)
        hip.synchronize()
        acc += out.numpy().astype(np.float64)
    dts = DeviceArray.from_numpy(thetas)
    dout = DeviceArray((B,), np.float64)
    capi.call("abhsf_mlp_predict", dts.ptr, nm, F, dx.ptr, B, dout.ptr, None)
    hip.synchronize()
    assert np.allclose(dout.numpy(), acc / nm, rtol=1e-12, atol=0.0)

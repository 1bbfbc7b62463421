"""Batched launch of many matrices (each with its own block size), eager and
replayed as a hipGraph, against the oracle per matrix."""
import numpy as np
import pytest

from abhsf_amd import AbhsfBatch, AbhsfMatrix, hip, synth
from abhsf_amd.hip import DeviceArray
from oracle import pyoracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(got, ref):
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (den if den > 0 else 1.0))


def make_members():
    rng = np.random.default_rng(0)
    mats = []
    for i in range(24):
        fam = i % 4
        n = int(np.exp(rng.uniform(np.log(1000), np.log(30000))))
        if fam == 0:
            A = synth.laplacian2d(int(np.sqrt(n)))
        elif fam == 1:
            A = synth.banded(n, int(rng.integers(4, 65)), seed=i)
        elif fam == 2:
            A = synth.rmat(int(np.log2(n)), 8, seed=i)
        else:
            A = synth.random_blocks(n, int(rng.choice([4, 8, 16, 32])), 4, float(rng.uniform(0.01, 0.3)), seed=i)
        s = int(rng.choice([8, 16, 32, 64, 128, 256]))
        mats.append((A, s))
    mats.append((synth.rmat(15, 16, seed=99), 64))  # heavy strips -> units + fix-up
    return mats


def test_batched_launch_eager_and_graph(gpu):
    members = make_members()
    st = hip.Stream()
    Ms, xs, ys, refs, y0s = [], [], [], [], []
    for i, (A, s) in enumerate(members):
        M = AbhsfMatrix(block_size=s, variant=("balanced", "strip", "panel", "tiled")[(i // 4) % 4], stream=st)
        M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
        x = np.random.default_rng(i).uniform(-1, 1, A.n)
        y0 = np.random.default_rng(i + 100).uniform(-1, 1, A.m)
        Ms.append(M)
        xs.append(DeviceArray.from_numpy(x))
        ys.append(DeviceArray.from_numpy(y0))
        y0s.append(y0)
        refs.append((A, x))
    B = AbhsfBatch(Ms, st)
    B.set_vectors(xs, ys)
    B.spmv_async(2.0, 0.0)
    B.synchronize()
    for (A, x), y in zip(refs, ys):
        ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x, 2.0, 0.0)
        assert relerr(y.numpy(), ref) <= TOL
    # graph replays with eager work in between: y <- 0.5*A x + y accumulates
    for y, y0 in zip(ys, y0s):
        y.upload(y0)
    for rep in range(3):
        B.spmv_async(0.5, 1.0, graph=True)
        B.synchronize()
        Ms[0].spmv_async(1.0, xs[0], 0.0, DeviceArray.zeros(members[0][0].m))  # unrelated eager work
        Ms[0].synchronize()
        for (A, x), y, y0 in zip(refs, ys, y0s):
            ref = y0 + (rep + 1) * pyoracle.csr_spmv(A.rowptr, A.col, A.val, x, 0.5, 0.0)
            assert relerr(y.numpy(), ref) <= 1e-11, rep
    B.release()


def test_batch_profile_cycles(gpu):
    """The timed launch computes the same y as the plain one and gives every
    member a positive cycle count; heavier members cost more."""
    members = make_members()[:8]
    st = hip.Stream()
    Ms, xs, ys, refs = [], [], [], []
    for i, (A, s) in enumerate(members):
        M = AbhsfMatrix(block_size=s, stream=st)
        M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
        x = np.random.default_rng(i).uniform(-1, 1, A.n)
        Ms.append(M)
        xs.append(DeviceArray.from_numpy(x))
        ys.append(DeviceArray.zeros(A.m))
        refs.append((A, x))
    B = AbhsfBatch(Ms, st)
    B.set_vectors(xs, ys)
    cyc = B.profile_cycles(1.0, 0.0)
    assert cyc.shape == (len(Ms),) and np.all(cyc > 0)
    for (A, x), y in zip(refs, ys):
        assert relerr(y.numpy(), pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)) <= TOL
    # a member ten times larger costs more cycles than a tiny one
    big = synth.banded(200000, 32, 1)
    small = synth.banded(2000, 4, 1)
    Mb, Ms2 = AbhsfMatrix(block_size=64, stream=st), AbhsfMatrix(block_size=64, stream=st)
    Mb.from_csr(big.m, big.n, big.rowptr, big.col, big.val)
    Ms2.from_csr(small.m, small.n, small.rowptr, small.col, small.val)
    B2 = AbhsfBatch([Mb, Ms2], st)
    B2.set_vectors([DeviceArray.zeros(big.n), DeviceArray.zeros(small.n)],
                   [DeviceArray.zeros(big.m), DeviceArray.zeros(small.m)])
    c2 = B2.profile_cycles()
    assert c2[0] > 10 * c2[1], c2
    B2.release()
    B.release()


def test_batch_with_fp32_and_symmetric_members(gpu):
    """FP32-valued members (a second launch with compensated accumulation,
    FP32 heavy-strip partials in the fix-up) and symmetric / skew-symmetric
    members (their mirrored triangle's pass after the batched launch) next to
    FP64 general members, eager and as replayed hipGraphs."""
    st = hip.Stream()
    L = synth.laplacian2d(70)
    r, c, v = L.triplets()
    low = c <= r
    sym_lower = synth.from_triplets(L.m, L.n, r[low], c[low], v[low])
    rs = synth.random_sparse(900, 900, 0.01, seed=4)
    rr, cc, vv = rs.triplets()
    sl = cc < rr  # strictly lower: skew-symmetric storage has no diagonal
    skew_lower = synth.from_triplets(rs.m, rs.n, rr[sl], cc[sl], vv[sl])
    cases = [  # (matrix, s, precision, variant, symmetry)
        (synth.rmat(14, 16, seed=21), 64, "f32", "balanced", "general"),   # heavy strips, FP32 partials
        (synth.block_mix(1 << 12, 5), 32, "f64", "tiled", "general"),
        (synth.banded(5000, 20, 3), 16, "f32", "tiled", "general"),
        (sym_lower, 8, "f64", "balanced", "symmetric"),
        (synth.random_blocks(6000, 8, 4, 0.2, seed=2), 8, "f32", "strip", "general"),
        (skew_lower, 32, "f64", "balanced", "skew"),
    ]
    Ms, xs, ys, refs = [], [], [], []
    for i, (A, s, prec, var, sym) in enumerate(cases):
        print(f"case batch member {i}: s={s} {prec} {var} {sym}", flush=True)
        M = AbhsfMatrix(block_size=s, precision=prec, variant=var, stream=st)
        M.from_csr(A.m, A.n, A.rowptr, A.col, A.val, symmetry=sym)
        x = np.random.default_rng(i).uniform(-1, 1, A.n)
        vals = A.val if prec == "f64" else A.val.astype(np.float32).astype(np.float64)
        # the whole matrix: the stored triangle plus its mirror (negated for skew)
        rr_, cc_, _ = A.triplets()
        if sym == "general":
            F = synth.Csr(A.m, A.n, A.rowptr, A.col, vals)
        else:
            off = rr_ != cc_
            sgn = -1.0 if sym == "skew" else 1.0
            F = synth.from_triplets(A.m, A.n, np.concatenate([rr_, cc_[off]]), np.concatenate([cc_, rr_[off]]),
                                    np.concatenate([vals, sgn * vals[off]]))
        Ms.append(M)
        xs.append(DeviceArray.from_numpy(x))
        ys.append(DeviceArray.zeros(A.m))
        refs.append(pyoracle.csr_spmv(F.rowptr, F.col, F.val, x, 1.5, 0.0))
    B = AbhsfBatch(Ms, st)
    B.set_vectors(xs, ys)
    for graph in (False, True, True):
        for y in ys:
            y.fill_zero(st.handle)
        st.synchronize()
        B.spmv_async(1.5, 0.0, graph=graph)
        B.synchronize()
        for i, (y, ref) in enumerate(zip(ys, refs)):
            assert relerr(y.numpy(), ref) <= TOL, (i, graph)

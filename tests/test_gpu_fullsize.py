"""Parity at the benchmark configurations' full sizes (about two minutes on one
MI355X; the CPU builder and CSR products are the checks).

- Block-structured matrix, n = 2^22 (1.36e8 entries): every storage array of
  the GPU conversion equals the CPU builder's byte for byte at s = 16, 32, 64
  and 128 (the reference's block processor, preprocessing/test.cpp:36-108,
  237-368, restated by the oracle).  The tiled variant (s = 32, the tuner's
  pick; tiles reorder COO blocks, so its arrays are not exported) is checked
  by an SpMV on integer data, which is exact whatever the summation order.
- R-MAT scale 22 (6.5e7 entries): FP64 SpMV within 1e-12 relative max-norm
  of the CSR product at (256, tiled) and (64, balanced), and bitwise
  identical when repeated.
- 27-point stencil 160^3 (1.09e8 entries): SpMM with 32 right-hand sides in
  FP64 within 1e-12; with FP32 value storage and compensated accumulation,
  within 1e-12 of the FP64 product of the f32-rounded values, and within the
  FP32 rounding bound of the FP64 values.
"""
import numpy as np
import pytest

from abhsf_amd import AbhsfMatrix, synth
from abhsf_amd.hip import DeviceArray
from oracle import pyoracle

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
TOL = 1e-12


def relerr(got, ref):
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def block_mix():
    A = synth.block_mix(1 << 22, seed=0)
    assert A.nnz == 135855744
    return A


@pytest.mark.parametrize("s", [16, 32, 64, 128])
def test_block_mix_conversion_bit_exact(gpu, block_mix, s):
    A = block_mix
    print(f"case fullsize block_mix n=2^22 s={s} arrays", flush=True)
    d = [DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val)]
    M = AbhsfMatrix(block_size=s, variant="balanced")
    M.from_csr(A.m, A.n, *d, on_device=True, nnz=A.nnz)
    del d
    garr = M.export_all()
    st = M.stats()
    M.release()
    O = pyoracle.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, s.bit_length() - 1)
    assert st["blocks"] == dict(zip(("coo", "csr", "bitmap", "dense"), O.nb))
    assert st["smf_bits_preproc"] == O.smf_bits
    oarr = O.arrays()
    del O
    bad = [k for k in oarr if not np.array_equal(oarr[k], garr[k])]
    assert not bad, f"arrays differ from the CPU builder: {bad}"


def test_block_mix_tiled_integer_exact(gpu, block_mix):
    A = block_mix
    print("case fullsize block_mix n=2^22 s=32 tiled integer SpMV", flush=True)
    rng = np.random.default_rng(1)
    vi = (rng.integers(1, 9, A.nnz) * rng.choice([-1.0, 1.0], A.nnz)).astype(np.float64)
    xi = rng.integers(-8, 9, A.n).astype(np.float64)
    d = [DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, vi)]
    M = AbhsfMatrix(block_size=32, variant="tiled")
    M.from_csr(A.m, A.n, *d, on_device=True, nnz=A.nnz)
    del d
    dx, dy = DeviceArray.from_numpy(xi), DeviceArray.zeros(A.m)
    M.spmv_async(1.0, dx, 0.0, dy)
    M.synchronize()
    ref = pyoracle.csr_spmv(A.rowptr, A.col, vi, xi)
    assert np.array_equal(dy.numpy(), ref)
    M.release()


@pytest.fixture(scope="module")
def rmat22():
    A = synth.rmat(22, 16, seed=0)
    x = np.random.default_rng(0).uniform(-1.0, 1.0, A.n)
    return A, x, pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)


@pytest.mark.parametrize("s,variant", [(256, "tiled"), (64, "balanced")])
def test_rmat22_spmv(gpu, rmat22, s, variant):
    A, x, ref = rmat22
    print(f"case fullsize rmat22 s={s} {variant}", flush=True)
    d = [DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val)]
    M = AbhsfMatrix(block_size=s, variant=variant)
    M.from_csr(A.m, A.n, *d, on_device=True, nnz=A.nnz)
    del d
    dx = DeviceArray.from_numpy(x)
    outs = []
    for _ in range(2):
        dy = DeviceArray.zeros(A.m)
        M.spmv_async(1.0, dx, 0.0, dy)
        M.synchronize()
        outs.append(dy.numpy())
    M.release()
    assert relerr(outs[0], ref) <= TOL
    assert np.array_equal(outs[0], outs[1])


def test_stencil160_spmm_k32(gpu):
    A = synth.stencil27(160)
    assert A.nnz == 109215352
    k = 32
    print("case fullsize stencil27 160^3 SpMM k=32 f64/f32", flush=True)
    X = np.random.default_rng(32).uniform(-1.0, 1.0, (A.n, k))
    ref = np.empty((A.m, k))
    absref = np.empty((A.m, k))
    aval = np.abs(A.val)
    for j in range(k):
        xj = np.ascontiguousarray(X[:, j])
        ref[:, j] = pyoracle.csr_spmv(A.rowptr, A.col, A.val, xj)
        absref[:, j] = pyoracle.csr_spmv(A.rowptr, A.col, aval, np.abs(xj))
    d = [DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val)]
    dX = DeviceArray.from_numpy(X)
    for prec in ("f64", "f32"):
        M = AbhsfMatrix(block_size=64, precision=prec)
        M.from_csr(A.m, A.n, *d, on_device=True, nnz=A.nnz)
        dY = DeviceArray.zeros((A.m, k))
        M.spmm_async(k, 1.0, dX, k, 0.0, dY, k)
        M.synchronize()
        Y = dY.numpy()
        M.release()
        err = relerr(Y, ref)
        print(f"stencil160 k=32 {prec}: max-norm rel err vs FP64 CSR {err:.3e}", flush=True)
        if prec == "f64":
            assert err <= TOL
        else:
            # the stencil's values (26, -1) are exact in FP32, so the stored
            # matrix equals the FP64 one and the compensated sums keep FP64
            # accuracy; the FP32 rounding bound u * sum |a_ij x_j| holds a
            # fortiori
            assert err <= TOL
            assert np.all(np.abs(Y - ref) <= 2.0 ** -24 * absref + 1e-300)

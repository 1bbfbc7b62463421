"""Tiled variant: runs of consecutive light strips processed as one tile whose
COO blocks are stored in (block column, strip) order.

The tile order is a permutation of the oracle's COO blocks inside each tile
(export is refused for a tiled handle), so every entry is checked through an
exact SpMV: integer values and an integer x keep every product and partial
sum exact in FP64, so y must equal the oracle's CSR SpMV bit for bit whatever
the summation order -- an entry dropped, repeated or given a wrong row or
column changes y.  Floating-point inputs are then checked to the FP64
tolerance, with alpha/beta, and the block counts per scheme against the
oracle builder (tiling moves blocks, it does not change them).
"""
import numpy as np
import pytest

from abhsf_amd import AbhsfBatch, AbhsfError, AbhsfMatrix, capi, hip, synth
from abhsf_amd.hip import DeviceArray
from oracle import pyoracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(got, ref):
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (den if den > 0 else 1.0))


MATS = {
    "scatter_2^15": lambda: synth.block_mix(1 << 15, 2, 0.0, 0.0),
    "block_mix_2^15": lambda: synth.block_mix(1 << 15, 3),
    "rmat_14": lambda: synth.rmat(14, 8, seed=9),
    "laplacian_200": lambda: synth.laplacian2d(200),
    "banded_5000_40": lambda: synth.banded(5000, 40, 2),
    "random_300x257": lambda: synth.random_sparse(300, 257, 0.05, seed=3),
    "narrow_3000x20": lambda: synth.random_sparse(3000, 20, 0.2, seed=5),
    "wide_40x70000": lambda: synth.random_sparse(40, 70000, 0.002, seed=11),
    "one_row": lambda: synth.random_sparse(1, 5000, 0.01, seed=6),
    "empty_rows": lambda: synth.from_triplets(700, 900, [0, 5, 699], [3, 899, 0], [1.0, -2.0, 3.0]),
    "all_empty": lambda: synth.from_triplets(300, 300, [], [], []),
}


def integer_copy(A, seed):
    """A with integer values in [-8, 8] \\ {0}."""
    rng = np.random.default_rng(seed)
    v = rng.integers(1, 9, A.nnz) * rng.choice([-1.0, 1.0], A.nnz)
    return synth.Csr(A.m, A.n, A.rowptr, A.col, v.astype(np.float64), A.name + "_int")


def run(M, x, y0, alpha, beta):
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y0)
    M.spmv_async(alpha, dx, beta, dy)
    M.synchronize()
    return dy.numpy()


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("s", [2, 8, 16, 32, 64, 256])
def test_tiled_exact_and_fp64(gpu, name, s):
    A = MATS[name]()
    print(f"case tiled {name} s={s}", flush=True)
    k = s.bit_length() - 1
    O = pyoracle.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, k, pyoracle.PREPROC, 8)
    # exact: integer matrix and x
    Ai = integer_copy(A, 1)
    Mi = AbhsfMatrix(block_size=s, variant="tiled")
    Mi.from_csr(Ai.m, Ai.n, Ai.rowptr, Ai.col, Ai.val)
    st = Mi.stats()
    assert [st["blocks"][x] for x in capi.SCHEMES] == O.nb
    xi = np.random.default_rng(2).integers(-8, 9, A.n).astype(np.float64)
    got = run(Mi, xi, np.zeros(A.m), 1.0, 0.0)
    ref = pyoracle.csr_spmv(Ai.rowptr, Ai.col, Ai.val, xi)
    assert np.array_equal(got, ref), "tiled SpMV is not exact on integer data"
    Mi.release()
    # FP64 data, alpha/beta
    M = AbhsfMatrix(block_size=s, variant="tiled")
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    x = np.random.default_rng(8).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(9).uniform(-1, 1, A.m)
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0), (1.0, 1.0)):
        got = run(M, x, y0, alpha, beta)
        ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x, alpha, beta, y0)
        assert relerr(got, ref) <= TOL, (alpha, beta)
    # the storage bytes differ from the balanced variant's only by the tables
    Mb = AbhsfMatrix(block_size=s, variant="balanced")
    Mb.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    assert st["values"] == Mb.stats()["values"]


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("s", [8, 32, 128])
def test_tiled_heavy_strips(gpu, precision, s):
    """Power-law rows: heavy strips stay units + fix-up, tiles form between
    them; FP32 values use tiles of at most 256 rows."""
    A = synth.rmat(15, 16, seed=4)
    print(f"case tiled rmat15 s={s} {precision}", flush=True)
    M = AbhsfMatrix(block_size=s, precision=precision, variant="tiled")
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    x = np.random.default_rng(6).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(7).uniform(-1, 1, A.m)
    got = run(M, x, y0, 2.0, 0.5)
    vals = A.val if precision == "f64" else A.val.astype(np.float32).astype(np.float64)
    ref = pyoracle.csr_spmv(A.rowptr, A.col, vals, x, 2.0, 0.5, y0)
    assert relerr(got, ref) <= TOL


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("case", ["scatter_zs2", "dense_coo_zs1"])
def test_tiled_phase_a_shape(gpu, precision, case):
    """The tiles' COO decode takes two entries per block in its first phase,
    or one when blocks of 3+ entries are at least 1 in 64 of the tiles'
    blocks (stats()["tile_zs1"]).  One matrix on each side of that threshold,
    in both value precisions, checked exactly on integer data."""
    if case == "scatter_zs2":  # 12 uniform columns per row: tile blocks hold 1 or 2 entries
        A, s, want = synth.block_mix(1 << 18, 2, 0.0, 0.0), 32, False
    else:  # ~10 entries per 32 x 32 block, all COO, strips of ~1300 entries
        A, s, want = synth.random_sparse(4096, 4096, 0.01, seed=12), 32, True
    print(f"case tiled phase-A {case} {precision}", flush=True)
    Ai = integer_copy(A, 3)
    M = AbhsfMatrix(block_size=s, precision=precision, variant="tiled")
    M.from_csr(Ai.m, Ai.n, Ai.rowptr, Ai.col, Ai.val)
    st = M.stats()
    assert st["variant"] == "tiled"
    assert st["tile_zs1"] is want
    xi = np.random.default_rng(4).integers(-8, 9, A.n).astype(np.float64)
    got = run(M, xi, np.zeros(A.m), 1.0, 0.0)
    ref = pyoracle.csr_spmv(Ai.rowptr, Ai.col, Ai.val, xi)
    # integers up to 8 * 8 * nnz-per-row: exact in FP64 and in the f32-valued
    # storage (the values are small integers, exact in FP32)
    assert np.array_equal(got, ref), "tiled SpMV is not exact on integer data"


def test_tiled_in_batch(gpu):
    st = hip.Stream()
    mats = [(synth.block_mix(1 << 13, 5, 0.0, 0.0), 32, "tiled"), (synth.rmat(13, 8, seed=3), 64, "tiled"),
            (synth.laplacian2d(90), 16, "balanced"), (synth.banded(3000, 12, 1), 8, "tiled"),
            (synth.rmat(14, 16, seed=8), 256, "tiled")]
    Ms, xs, ys, refs = [], [], [], []
    for i, (A, s, v) in enumerate(mats):
        M = AbhsfMatrix(block_size=s, variant=v, stream=st)
        M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
        x = np.random.default_rng(i).uniform(-1, 1, A.n)
        Ms.append(M)
        xs.append(DeviceArray.from_numpy(x))
        ys.append(DeviceArray.zeros(A.m))
        refs.append(pyoracle.csr_spmv(A.rowptr, A.col, A.val, x))
    B = AbhsfBatch(Ms, st)
    B.set_vectors(xs, ys)
    for graph in (False, True):
        for y in ys:
            y.fill_zero(st.handle)
        B.spmv_async(1.0, 0.0, graph=graph)
        B.synchronize()
        for y, ref in zip(ys, refs):
            assert relerr(y.numpy(), ref) <= TOL, graph


@pytest.mark.parametrize("nrhs", [2, 5])
def test_tiled_spmm_column_by_column(gpu, nrhs):
    """SpMM on a tiled handle: each right-hand side through the tiled SpMV
    kernels, strided X and Y, alpha/beta, heavy strips included."""
    A = synth.rmat(14, 16, seed=6)
    print(f"case tiled spmm rmat14 nrhs={nrhs}", flush=True)
    M = AbhsfMatrix(block_size=64, variant="tiled")
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    rng = np.random.default_rng(nrhs)
    ldx, ldy = nrhs + 2, nrhs + 1
    X = rng.uniform(-1, 1, (A.n, ldx))
    Y0 = rng.uniform(-1, 1, (A.m, ldy))
    dX, dY = DeviceArray.from_numpy(X), DeviceArray.from_numpy(Y0)
    M.spmm_async(nrhs, 0.5, dX, ldx, -2.0, dY, ldy)
    M.synchronize()
    got = dY.numpy()
    for j in range(nrhs):
        ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, X[:, j].copy(), 0.5, -2.0, Y0[:, j].copy())
        assert relerr(got[:, j], ref) <= TOL, j
    assert np.array_equal(got[:, nrhs:], Y0[:, nrhs:])  # padding columns untouched


def test_tiled_refuses_non_spmv_uses(gpu):
    A = synth.block_mix(1 << 12, 1, 0.0, 0.0)
    M = AbhsfMatrix(block_size=32, variant="tiled")
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    with pytest.raises(AbhsfError, match="exported"):
        M.export_all()
    L = synth.laplacian2d(20)
    r, c, v = L.triplets()
    keep = c <= r
    lower = synth.from_triplets(L.m, L.n, r[keep], c[keep], v[keep])
    S = AbhsfMatrix(block_size=8, variant="tiled")
    with pytest.raises(AbhsfError, match="tiled"):
        S.from_csr(lower.m, lower.n, lower.rowptr, lower.col, lower.val, symmetry="symmetric")


def test_import_into_tiled_handle(gpu, tmp_path):
    """A blob saved from a strip-ordered handle loads into a tiled handle (the
    tile order is built on import) and gives the same SpMV."""
    A = synth.block_mix(1 << 13, 4)
    M = AbhsfMatrix(block_size=32, variant="balanced")
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    path = str(tmp_path / "m.npz")
    M.save(path)
    T = AbhsfMatrix.load(path, variant="tiled")
    x = np.random.default_rng(3).uniform(-1, 1, A.n)
    got = run(T, x, np.zeros(A.m), 1.0, 0.0)
    ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)
    assert relerr(got, ref) <= TOL
    with pytest.raises(AbhsfError, match="exported"):
        T.save(str(tmp_path / "t.npz"))

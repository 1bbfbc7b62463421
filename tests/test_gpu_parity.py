"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Conversion: every storage array must equal the oracle builder's byte for byte.
SpMV/SpMM FP64: relative max-norm error <= 1e-12 against the textbook CSR FP64
SpMV of the same inputs (TOL below).  FP32 value storage: error against the
FP64 oracle on the FP32-rounded values <= 1e-12 (compensated accumulation),
and the error against the FP64 values is reported.
"""
import functools
import os

import numpy as np
import pytest

from abhsf_amd import AbhsfMatrix, AbhsfError, capi, synth
from abhsf_amd.hip import DeviceArray
from oracle import pyoracle

from conftest import ROOT, read_mtx

TOL = 1e-12  # FP64 relative max-norm bound

pytestmark = pytest.mark.gpu


def relerr(got, ref):
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (den if den > 0 else 1.0))


def build(A, s, policy="preproc", precision="f64"):
    M = AbhsfMatrix(block_size=s, policy=policy, precision=precision)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    return M


def check_arrays(M, A, s, policy=pyoracle.PREPROC, val_bytes=8):
    k = s.bit_length() - 1
    O = pyoracle.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, k, policy, val_bytes)
    oa, ga = O.arrays(), M.export_all()
    for name in oa:
        assert oa[name].dtype == ga[name].dtype, name
        assert np.array_equal(oa[name], ga[name]), f"array {name} differs (s={s})"
    st = M.stats()
    assert [st["blocks"][x] for x in capi.SCHEMES] == O.nb
    return O, st


def spmv_check(M, A, x, y0=None, alpha=1.0, beta=0.0):
    dx = DeviceArray.from_numpy(x)
    dy = DeviceArray.from_numpy(y0 if y0 is not None else np.zeros(A.m))
    M.spmv_async(alpha, dx, beta, dy)
    M.synchronize()
    got = dy.numpy()
    ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x, alpha, beta, y0)
    return got, ref


MATS = {
    "laplacian_32": lambda: synth.laplacian2d(32),
    "random_300x257": lambda: synth.random_sparse(300, 257, 0.05, seed=3),
    "random_dense_ish": lambda: synth.random_sparse(130, 140, 0.6, seed=4),
    "block_mix_4096": lambda: synth.block_mix(4096, seed=1),
    "rmat_12": lambda: synth.rmat(12, 8, seed=2),
    "banded_1000": lambda: synth.banded(1000, 9, seed=5),
    "random_blocks_2000": lambda: synth.random_blocks(2000, 16, 4, 0.3, seed=6),
}


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("s", [2, 4, 8, 16, 32, 64, 128, 256])
def test_conversion_bit_exact_and_spmv(gpu, name, s):
    A = MATS[name]()
    print(f"case {name} s={s} m={A.m} n={A.n} nnz={A.nnz}", flush=True)
    M = build(A, s)
    check_arrays(M, A, s)
    x = np.random.default_rng(s).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(s + 1).uniform(-1, 1, A.m)
    got, ref = spmv_check(M, A, x, y0, alpha=1.0, beta=1.0)
    assert relerr(got, ref) <= TOL


def test_config1_laplacian256_s64(gpu):
    """The reference's CPU-path case: 256^2 5-point Laplacian, s=64, x ~ U[-1,1]
    seed 0, y = A x checked against CSR FP64; footprint = the reference's
    3,695,344 bits."""
    A = synth.laplacian2d(256)
    assert A.nnz == 326656
    M = build(A, 64)
    O, st = check_arrays(M, A, 64)
    assert st["smf_bits_preproc"] == 3695344
    assert st["blocks"]["csr"] == 1024 and st["blocks"]["coo"] == 3576
    x = np.random.default_rng(0).uniform(-1, 1, A.n)
    got, ref = spmv_check(M, A, x)
    assert relerr(got, ref) <= TOL
    # closed form: A * 1 = 4 - (number of neighbours)
    ones = np.ones(A.n)
    got1, _ = spmv_check(M, A, ones)
    g = 256
    i, j = np.divmod(np.arange(A.n), g)
    deg = (i > 0).astype(int) + (i < g - 1) + (j > 0) + (j < g - 1)
    assert np.array_equal(got1, 4.0 - deg)


def test_test_mtx_all_block_sizes(gpu):
    m, n, r, c, v = read_mtx(f"{ROOT}/tests/golden/test.mtx")
    A = synth.from_triplets(m, n, r, c, v)
    import json
    ka = json.load(open(f"{ROOT}/tests/golden/known_answers.json"))["test_mtx"]
    for s in (2, 4, 8, 16, 32, 64, 128, 256):
        M = build(A, s)
        check_arrays(M, A, s)
        assert M.stats()["smf_bits_preproc"] == ka["preproc_smf_bits"][str(s)]
        x = np.arange(1, n + 1, dtype=np.float64)
        got, ref = spmv_check(M, A, x)
        assert relerr(got, ref) <= TOL


def test_from_elements_accumulating_harness(gpu):
    """Reference harness semantics (spmv/unified/spmv.cpp:1176-1232): x = 0.5,
    y = 0, 5 warm-up + 20 timed accumulating spmv calls; result ||y||/25 equals
    the expected ||A*0.5||."""
    A = synth.laplacian2d(40)
    r, c, v = A.triplets()
    from abhsf_amd import matrix_properties
    perm = np.random.default_rng(9).permutation(r.size)  # any order is accepted
    M = AbhsfMatrix(block_size=8)
    M.from_elements((r[perm], c[perm], v[perm]), matrix_properties(A.m, A.n, A.nnz))
    check_arrays(M, A, 8)
    x = DeviceArray.from_numpy(np.full(A.n, 0.5))
    y = DeviceArray.zeros(A.m)
    for _ in range(25):
        M.spmv(x, y)
    expected = np.linalg.norm(pyoracle.csr_spmv(A.rowptr, A.col, A.val, np.full(A.n, 0.5)))
    assert abs(np.linalg.norm(y.numpy()) / 25 - expected) <= 1e-12 * expected


@pytest.mark.parametrize("s", [8, 32, 64, 256])
@pytest.mark.parametrize("nrhs", [1, 2, 3, 4, 8, 16, 32, 33, 64])
def test_spmm(gpu, s, nrhs):
    A = synth.block_mix(2048, seed=7)
    print(f"case spmm s={s} nrhs={nrhs}", flush=True)
    M = build(A, s)
    rng = np.random.default_rng(nrhs)
    ldx, ldy = nrhs + 3, nrhs + 1
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


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("s", [16, 64, 256])
@pytest.mark.parametrize("nrhs", [1, 3, 8, 13, 16, 20, 32])
def test_spmm_csr_blocks(gpu, precision, s, nrhs):
    """27-point stencil on 24^3: CSR blocks (s = 64, 256) whose rows each lane
    walks two at a time, plus bitmap blocks (s = 16); 13 right-hand sides
    make a column tile of 8 and one of 5.  Integer values and X keep every
    sum exact, so Y must equal the oracle bit for bit; then alpha/beta."""
    A = synth.stencil27(24)
    print(f"case spmm csr s={s} nrhs={nrhs} {precision}", flush=True)
    rng = np.random.default_rng(s + nrhs)
    vi = (rng.integers(1, 9, A.nnz) * rng.choice([-1.0, 1.0], A.nnz)).astype(np.float64)
    M = AbhsfMatrix(block_size=s, precision=precision)
    M.from_csr(A.m, A.n, A.rowptr, A.col, vi)
    X = rng.integers(-8, 9, (A.n, nrhs)).astype(np.float64)
    dX, dY = DeviceArray.from_numpy(X), DeviceArray.zeros((A.m, nrhs))
    M.spmm_async(nrhs, 1.0, dX, nrhs, 0.0, dY, nrhs)
    M.synchronize()
    got = dY.numpy()
    for j in range(nrhs):
        assert np.array_equal(got[:, j], pyoracle.csr_spmv(A.rowptr, A.col, vi, X[:, j].copy())), j
    Xf = rng.uniform(-1, 1, (A.n, nrhs))
    Y0 = rng.uniform(-1, 1, (A.m, nrhs))
    dX, dY = DeviceArray.from_numpy(Xf), DeviceArray.from_numpy(Y0)
    M.spmm_async(nrhs, -0.5, dX, nrhs, 2.0, dY, nrhs)
    M.synchronize()
    got = dY.numpy()
    for j in range(nrhs):
        ref = pyoracle.csr_spmv(A.rowptr, A.col, vi, Xf[:, j].copy(), -0.5, 2.0, Y0[:, j].copy())
        assert relerr(got[:, j], ref) <= TOL, j


@pytest.mark.parametrize("s", [8, 32, 64])
def test_msmf_policy(gpu, s):
    A = synth.block_mix(4096, seed=11)
    M = build(A, s, policy="msmf")
    check_arrays(M, A, s, policy=pyoracle.MSMF)
    x = np.random.default_rng(1).uniform(-1, 1, A.n)
    got, ref = spmv_check(M, A, x)
    assert relerr(got, ref) <= TOL


@pytest.mark.parametrize("s", [8, 32, 64, 128])
def test_fp32_storage_compensated(gpu, s):
    A = synth.block_mix(4096, seed=12)
    M = build(A, s, precision="f32")
    check_arrays(M, A, s, val_bytes=4)
    x = np.random.default_rng(2).uniform(-1, 1, A.n)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.zeros(A.m)
    M.spmv_async(1.0, dx, 0.0, dy)
    M.synchronize()
    got = dy.numpy()
    v32 = A.val.astype(np.float32).astype(np.float64)
    ref32 = pyoracle.csr_spmv(A.rowptr, A.col, v32, x)
    ref64 = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)
    e32, e64 = relerr(got, ref32), relerr(got, ref64)
    print(f"fp32 storage s={s}: err vs fp64 oracle on f32 values {e32:.2e}, vs fp64 values {e64:.2e}")
    assert e32 <= TOL
    assert e64 <= 1e-6


def test_beta_zero_ignores_y_and_alpha(gpu):
    A = synth.random_sparse(500, 400, 0.02, seed=8)
    M = build(A, 16)
    x = np.random.default_rng(3).uniform(-1, 1, A.n)
    y0 = np.full(A.m, np.nan)
    got, ref = spmv_check(M, A, x, y0, alpha=-3.0, beta=0.0)
    assert not np.any(np.isnan(got))
    assert relerr(got, ref) <= TOL


def full_matrix(m, n):
    r, c = np.divmod(np.arange(m * n), n)
    v = np.random.default_rng(0).standard_normal(m * n)
    return synth.from_triplets(m, n, r, c, v)


def test_edge_shapes(gpu):
    cases = [
        synth.from_triplets(1, 1, [0], [0], [2.5]),
        synth.from_triplets(5, 3, [], [], []),                          # empty
        synth.from_triplets(3, 1000, [0, 2, 2], [999, 0, 998], [1.0, 2.0, 3.0]),
        synth.from_triplets(1000, 3, [999, 0, 500], [2, 0, 1], [1.0, -1.0, 4.0]),
        full_matrix(257, 255),                                          # dense blocks at every s
    ]
    for A in cases:
        for s in (2, 8, 64, 256):
            M = build(A, s)
            check_arrays(M, A, s)
            x = np.random.default_rng(0).uniform(-1, 1, A.n)
            got, ref = spmv_check(M, A, x)
            assert relerr(got, ref) <= TOL, (A.m, A.n, s)


def test_invalid_input_rejected(gpu):
    M = AbhsfMatrix(block_size=8)
    # unsorted columns within a row
    with pytest.raises(AbhsfError) as e:
        M.from_csr(2, 4, np.array([0, 2, 3]), np.array([3, 1, 0]), np.ones(3))
    assert e.value.status == capi.ERR_INVALID
    # column out of range
    with pytest.raises(AbhsfError):
        M.from_csr(2, 4, np.array([0, 1, 2]), np.array([0, 4]), np.ones(2))
    # duplicate triplet
    from abhsf_amd import matrix_properties
    with pytest.raises(AbhsfError):
        M.from_elements((np.array([0, 0]), np.array([1, 1]), np.ones(2)), matrix_properties(2, 2, 2))
    # spmv before a successful build
    dx, dy = DeviceArray.zeros(4), DeviceArray.zeros(2)
    with pytest.raises(AbhsfError) as e:
        M.spmv(dx, dy)
    assert e.value.status == capi.ERR_STATE


def test_device_csr_input_and_determinism(gpu):
    A = synth.rmat(14, 16, seed=3)
    d_rp, d_c, d_v = (DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val))
    M = AbhsfMatrix(block_size=32)
    M.from_csr(A.m, A.n, d_rp, d_c, d_v, on_device=True)
    check_arrays(M, A, 32)
    x = np.random.default_rng(5).uniform(-1, 1, A.n)
    dx = DeviceArray.from_numpy(x)
    outs = []
    for _ in range(3):
        dy = DeviceArray.zeros(A.m)
        M.spmv_async(1.0, dx, 0.0, dy)
        M.synchronize()
        outs.append(dy.numpy())
    ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)
    assert relerr(outs[0], ref) <= TOL
    # FP64 results are deterministic: each strip accumulates in one
    # wavefront's LDS, and heavy strips' partials are summed in a fixed order
    same = all(np.array_equal(outs[0], o) for o in outs[1:])
    assert same, "repeated FP64 SpMV differs bit for bit"


@pytest.mark.parametrize("variant", ["balanced", "tiled"])
def test_repeated_spmv_bitwise_identical(gpu, variant):
    """Heavy strips (units + fix-up) and tiles: three runs, same bits."""
    A = synth.rmat(15, 16, seed=4)
    print(f"case determinism rmat15 s=64 {variant}", flush=True)
    M = AbhsfMatrix(block_size=64, variant=variant)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    dx = DeviceArray.from_numpy(np.random.default_rng(1).uniform(-1, 1, A.n))
    outs = []
    for _ in range(3):
        dy = DeviceArray.zeros(A.m)
        M.spmv_async(1.0, dx, 0.0, dy)
        M.synchronize()
        outs.append(dy.numpy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


@pytest.mark.parametrize("variant", ["strip", "balanced", "panel"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("s", [8, 64, 128, 256])
def test_heavy_strip_variants(gpu, variant, precision, s):
    """Power-law rows: heavy strips are split into units whose partial sums
    are combined by the fix-up pass (balanced) or processed whole (strip).
    The fix-up takes a strip in slices of at most 128 rows (64 with FP32
    values): one or two slices at s = 128, two or four at s = 256."""
    A = synth.rmat(15, 16, seed=4)
    print(f"case rmat15 s={s} {precision} {variant}", flush=True)
    M = AbhsfMatrix(block_size=s, precision=precision, variant=variant)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    check_arrays(M, A, s, val_bytes=8 if precision == "f64" else 4)
    x = np.random.default_rng(6).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(7).uniform(-1, 1, A.m)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y0)
    M.spmv_async(2.0, dx, 0.5, dy)
    M.synchronize()
    vals = A.val if precision == "f64" else A.val.astype(np.float32).astype(np.float64)
    ref = pyoracle.csr_spmv(A.rowptr, A.col, vals, x, 2.0, 0.5, y0)
    assert relerr(dy.numpy(), ref) <= TOL


def strips_of_many_units(seed):
    """Rows 0-63 full over 10,000 columns (about 312 units of 2,048 entries),
    rows 64-127 over 1,500 (47 units), rows 128-191 over 320 (10 units), the
    other rows tridiagonal: at s = 64 three heavy strips of very different
    unit counts, at s = 256 one of about 370."""
    n = 12000
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r0, w in ((0, 10000), (64, 1500), (128, 320)):
        rr = np.repeat(np.arange(r0, r0 + 64), w)
        rows.append(rr)
        cols.append((np.tile(np.arange(w), 64) * 7 + rr) % n)
    t = np.arange(192, n)
    for d in (-1, 0, 1):
        ok = (t + d >= 0) & (t + d < n)
        rows.append(t[ok])
        cols.append(t[ok] + d)
    r, c = np.concatenate(rows), np.concatenate(cols)
    return synth.from_triplets(n, n, r, c, rng.uniform(-1, 1, r.size), "many_units")


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("s", [64, 256])
def test_fixup_slices_by_unit_count(gpu, precision, s):
    """The fix-up narrows a heavy strip's row slices (down to 16 rows, more
    threads per row) when it has many units, and each thread ends with one
    round of all, half or a quarter of its chains: strips of about 370, 312,
    47 and 10 units cover the narrow slices and every last-round width."""
    A = strips_of_many_units(11)
    print(f"case many_units s={s} {precision}", flush=True)
    M = AbhsfMatrix(block_size=s, precision=precision, variant="balanced")
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    x = np.random.default_rng(12).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(13).uniform(-1, 1, A.m)
    dx = DeviceArray.from_numpy(x)
    vals = A.val if precision == "f64" else A.val.astype(np.float32).astype(np.float64)
    ref = pyoracle.csr_spmv(A.rowptr, A.col, vals, x, 1.5, -0.5, y0)
    outs = []
    for _ in range(2):
        dy = DeviceArray.from_numpy(y0)
        M.spmv_async(1.5, dx, -0.5, dy)
        M.synchronize()
        outs.append(dy.numpy())
    assert relerr(outs[0], ref) <= TOL
    assert np.array_equal(outs[0], outs[1]), "repeated SpMV differs bit for bit"


@pytest.mark.parametrize("variant", ["strip", "balanced"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("s", [64, 256])
def test_spmm_power_law_rows(gpu, variant, precision, s):
    """R-MAT: the first strip holds more than 1/4096 of the entries, so
    abhsf_spmm runs each right-hand side through the SpMV kernels (which
    split heavy strips) instead of giving the strip to one SpMM wave.
    Integer values and X: exact, for an odd nrhs with ldx > nrhs too."""
    A = synth.rmat(15, 16, seed=4)
    rng = np.random.default_rng(s)
    vi = (rng.integers(1, 9, A.nnz) * rng.choice([-1.0, 1.0], A.nnz)).astype(np.float64)
    print(f"case spmm power-law s={s} {variant} {precision}", flush=True)
    M = AbhsfMatrix(block_size=s, precision=precision, variant=variant)
    M.from_csr(A.m, A.n, A.rowptr, A.col, vi)
    st = M.stats()
    assert st["max_strip_work"] > max(4096, A.nnz // 4096)
    for nrhs, ldx, ldy in ((8, 8, 8), (13, 16, 15)):
        X = rng.integers(-8, 9, (A.n, ldx)).astype(np.float64)
        Y0 = rng.integers(-8, 9, (A.m, ldy)).astype(np.float64)
        dX, dY = DeviceArray.from_numpy(X), DeviceArray.from_numpy(Y0)
        M.spmm_async(nrhs, 1.0, dX, ldx, 2.0, dY, ldy)
        M.synchronize()
        got = dY.numpy()
        for j in range(nrhs):
            ref = pyoracle.csr_spmv(A.rowptr, A.col, vi, X[:, j].copy(), 1.0, 2.0, Y0[:, j].copy())
            assert np.array_equal(got[:, j], ref), (nrhs, j)
        assert np.array_equal(got[:, nrhs:], Y0[:, nrhs:])  # padding columns untouched


@functools.lru_cache(maxsize=1)
def two_full_rows(n=1 << 21, seed=0):
    """Rows 0 and 1 hold every column (2n entries: the strip that contains
    them is cut into ~2n/2048 units), the other rows up to 3 random entries;
    integer values in [-8, 8] \\ {0}."""
    rng = np.random.default_rng(seed)
    r = np.repeat(np.arange(2, n, dtype=np.int64), rng.integers(0, 4, n - 2))
    key = np.unique(r * n + rng.integers(0, n, r.size))  # sorted, duplicates merged
    rows = np.concatenate([np.zeros(n, np.int64), np.ones(n, np.int64), key // n])
    col = np.concatenate([np.arange(n), np.arange(n), key % n]).astype(np.int32)
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    val = (rng.integers(1, 9, col.size) * rng.choice([-1.0, 1.0], col.size)).astype(np.float64)
    return synth.Csr(n, n, rowptr, col, val, "two_full_rows")


@pytest.mark.parametrize("variant,precision", [("balanced", "f64"), ("tiled", "f64"), ("balanced", "f32")])
@pytest.mark.parametrize("s", [2, 8, 256])
def test_fixup_many_units_per_strip(gpu, variant, precision, s):
    """One strip of 2048+ units: the fix-up's threads of a row (256/s of
    them) each walk many units in their interleaved chains and then the
    remainder.  Integer data keeps every sum exact, so y must equal the
    oracle bit for bit whatever the summation order."""
    A = two_full_rows()
    print(f"case two_full_rows s={s} {variant} {precision}", flush=True)
    M = AbhsfMatrix(block_size=s, precision=precision, variant=variant)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    x = np.random.default_rng(1).integers(-8, 9, A.n).astype(np.float64)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.zeros(A.m)
    M.spmv_async(1.0, dx, 0.0, dy)
    M.synchronize()
    ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)
    assert np.array_equal(dy.numpy(), ref)
    y0 = np.random.default_rng(2).uniform(-1, 1, A.m)
    dy = DeviceArray.from_numpy(y0)
    M.spmv_async(-0.5, dx, 2.0, dy)
    M.synchronize()
    assert relerr(dy.numpy(), pyoracle.csr_spmv(A.rowptr, A.col, A.val, x, -0.5, 2.0, y0)) <= TOL


PANEL_MATS = {
    "scatter_2^15": lambda: synth.block_mix(1 << 15, 2, 0.0, 0.0),
    "block_mix_2^15": lambda: synth.block_mix(1 << 15, 3),
    "rmat_14": lambda: synth.rmat(14, 8, seed=9),
    "laplacian_200": lambda: synth.laplacian2d(200),
    "banded_5000_40": lambda: synth.banded(5000, 40, 2),
    "narrow_3000x20": lambda: synth.random_sparse(3000, 20, 0.2, seed=5),
    "one_row": lambda: synth.random_sparse(1, 5000, 0.01, seed=6),
    "empty_rows": lambda: synth.from_triplets(700, 900, [0, 5, 699], [3, 899, 0], [1.0, -2.0, 3.0]),
    "all_empty": lambda: synth.from_triplets(300, 300, [], [], []),
}


@pytest.mark.parametrize("name", list(PANEL_MATS))
@pytest.mark.parametrize("s", [16, 64, 256])
def test_panel_variant(gpu, name, s):
    """Column-panel variant: strips cut at the 8 panel boundaries, items
    placed per XCD by the slot table, split strips summed by the fix-up."""
    A = PANEL_MATS[name]()
    print(f"case panel {name} s={s}", flush=True)
    M = AbhsfMatrix(block_size=s, variant="panel")
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    check_arrays(M, A, s)
    x = np.random.default_rng(8).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(9).uniform(-1, 1, A.m)
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        got, ref = spmv_check(M, A, x, y0, alpha, beta)
        assert relerr(got, ref) <= TOL, (alpha, beta)
    # SpMM through the same handle (its kernel does not use the panel items)
    X = np.random.default_rng(10).uniform(-1, 1, (A.n, 3))
    dX, dY = DeviceArray.from_numpy(X), DeviceArray.zeros((A.m, 3))
    M.spmm_async(3, 1.0, dX, 3, 0.0, dY, 3)
    M.synchronize()
    Y = dY.numpy()
    for j in range(3):
        ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, X[:, j].copy())
        assert relerr(Y[:, j], ref) <= TOL, j


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("s", [16, 64, 256])
@pytest.mark.parametrize("case", ["stencil27_24", "block_mix_2048"])
def test_spmm_x_alignment(gpu, case, s, precision):
    """SpMM reads whole 16-byte-aligned column tiles of X as double2 and
    anything else one double at a time: X at a 16-byte boundary and X one
    double past it (same ldx) must both give the oracle's product exactly
    (integer values and X keep every sum exact, whatever the order)."""
    A = synth.stencil27(24) if case == "stencil27_24" else synth.block_mix(2048, seed=3)
    print(f"case spmm x alignment {case} s={s} {precision}", flush=True)
    rng = np.random.default_rng(s)
    vi = (rng.integers(1, 9, A.nnz) * rng.choice([-1.0, 1.0], A.nnz)).astype(np.float64)
    M = AbhsfMatrix(block_size=s, precision=precision)
    M.from_csr(A.m, A.n, A.rowptr, A.col, vi)
    for nrhs in (8, 16):
        X = rng.integers(-8, 9, (A.n, nrhs)).astype(np.float64)
        ref = np.stack([pyoracle.csr_spmv(A.rowptr, A.col, vi, X[:, j].copy()) for j in range(nrhs)], axis=1)
        flat = np.concatenate([[0.0], X.ravel()])
        buf = DeviceArray.from_numpy(flat)
        dX = DeviceArray.from_numpy(X)
        assert dX.address % 16 == 0 and (buf.address + 8) % 16 == 8
        for xaddr in (dX.address, buf.address + 8):
            dY = DeviceArray.zeros((A.m, nrhs))
            M.spmm_async(nrhs, 1.0, xaddr, nrhs, 0.0, dY, nrhs)
            M.synchronize()
            assert np.array_equal(dY.numpy(), ref), (nrhs, xaddr % 16)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("s", [16, 32, 64, 128])
def test_unaligned_x(gpu, precision, s):
    """x may start at any double (here x + 1 of a device array, 8 bytes past a
    16-byte boundary): dense-scheme blocks, whose paths load x in pairs, give
    the same result as from an aligned copy."""
    A = full_matrix(300, 301)  # dense-scheme blocks everywhere
    print(f"case unaligned x s={s} {precision}", flush=True)
    M = AbhsfMatrix(block_size=s, precision=precision)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    assert M.stats()["blocks"]["dense"] > 0
    x = np.random.default_rng(2).uniform(-1, 1, A.n)
    buf = DeviceArray.from_numpy(np.concatenate([[0.0], x]))  # x at byte offset 8
    assert (buf.address + 8) % 16 == 8
    dx = DeviceArray.from_numpy(x)
    outs = []
    for xaddr in (dx.address, buf.address + 8):
        dy = DeviceArray.zeros(A.m)
        M.spmv_async(1.0, xaddr, 0.0, dy)
        M.synchronize()
        outs.append(dy.numpy())
    assert np.array_equal(outs[0], outs[1])
    vals = A.val if precision == "f64" else A.val.astype(np.float32).astype(np.float64)
    assert relerr(outs[1], pyoracle.csr_spmv(A.rowptr, A.col, vals, x)) <= TOL


@pytest.mark.parametrize("case", ["test.mtx", "block_mix_2^13", "rmat_12", "laplacian_96"])
def test_smf_bits_msmf_pinned(gpu, case):
    """abhsf_stats.smf_bits_msmf is the structure footprint of
    block_statistics/mmf.cpp (top-level block CSR M ceil(log2(N+1)) +
    NNZB ceil(log2 N) plus, per nonzero block, the cheapest scheme + 2 bits):
    the oracle's mmf row (pinned to the reference's mmf binary) minus the
    value bits, at every s.  test.mtx at s = 256 is App. B's known answer
    (CSR-ABHSF 259 bits)."""
    if case == "test.mtx":
        m, n, r, c, v = read_mtx(os.path.join(ROOT, "tests", "golden", "test.mtx"))
        A = synth.from_triplets(m, n, r, c, v)
    else:
        A = {"block_mix_2^13": lambda: synth.block_mix(1 << 13, 7), "rmat_12": lambda: synth.rmat(12, 8, seed=2),
             "laplacian_96": lambda: synth.laplacian2d(96)}[case]()
    rows, cols, _ = A.triplets()
    print(f"case smf_bits_msmf {case}", flush=True)
    for k in range(1, 9):
        s = 1 << k
        M = AbhsfMatrix(block_size=s)
        M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
        got = M.stats()["smf_bits_msmf"]
        M.release()
        row = pyoracle.mmf_row(A.m, A.n, A.nnz, s, s, pyoracle.block_hist(rows, cols, k, k))
        assert got == row[6] - A.nnz * 64, (case, s)
        if case == "test.mtx" and s == 256:
            assert got == 259

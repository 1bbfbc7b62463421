"""Symmetric / skew-symmetric storage: one triangle stored, SpMV applies the
whole matrix (spmv/unified/spmv.cpp:1160-1166 counts nnz_all for it), against
the oracle's CSR SpMV of the expanded matrix."""
import numpy as np
import pytest

from abhsf_amd import AbhsfBatch, AbhsfError, AbhsfMatrix, capi, mmio, synth
from abhsf_amd.hip import DeviceArray
from abhsf_amd.matrix import matrix_properties, matrix_symmetry_t
from oracle import pyoracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(got, ref):
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (den if den > 0 else 1.0))


def triangle(A, lower=True, strict=False):
    r, c, v = A.triplets()
    keep = (r > c) if strict else (r >= c)
    if not lower:
        keep = (r < c) if strict else (r <= c)
    return synth.from_triplets(A.m, A.n, r[keep], c[keep], v[keep])


def expand(T, skew=False):
    r, c, v = T.triplets()
    off = r != c
    sign = -1.0 if skew else 1.0
    return synth.from_triplets(T.m, T.n, np.concatenate([r, c[off]]), np.concatenate([c, r[off]]),
                               np.concatenate([v, sign * v[off]]))


def symmetric_random(n, density, seed):
    A = synth.random_sparse(n, n, density, seed)
    r, c, v = A.triplets()
    L = synth.from_triplets(n, n, r[r >= c], c[r >= c], v[r >= c])
    return L


SYM_MATS = {
    "laplacian_100": lambda: triangle(synth.laplacian2d(100)),
    "random_700": lambda: symmetric_random(700, 0.02, 1),
    "rmat_12": lambda: triangle(synth.rmat(12, 8, seed=4)),
    "banded_3000_20_upper": lambda: triangle(synth.banded(3000, 20, 3), lower=False),
    "block_mix_4096": lambda: triangle(synth.block_mix(4096, seed=2)),
}


def run_spmv(M, n, x, y0, alpha, beta):
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y0)
    M.spmv_async(alpha, dx, beta, dy)
    M.synchronize()
    return dy.numpy()


@pytest.mark.parametrize("name", list(SYM_MATS))
@pytest.mark.parametrize("s", [8, 64, 256])
@pytest.mark.parametrize("variant", ["strip", "balanced", "panel"])
def test_symmetric_spmv(gpu, name, s, variant):
    T = SYM_MATS[name]()
    print(f"case symmetric {name} s={s} {variant}", flush=True)
    full = expand(T)
    M = AbhsfMatrix(block_size=s, variant=variant)
    M.from_csr(T.m, T.n, T.rowptr, T.col, T.val, symmetry="symmetric")
    assert M.stats()["nnz"] == T.nnz
    x = np.random.default_rng(1).uniform(-1, 1, T.n)
    y0 = np.random.default_rng(2).uniform(-1, 1, T.m)
    for alpha, beta in ((1.0, 0.0), (0.5, -1.5)):
        got = run_spmv(M, T.n, x, y0, alpha, beta)
        ref = pyoracle.csr_spmv(full.rowptr, full.col, full.val, x, alpha, beta, y0)
        assert relerr(got, ref) <= TOL, (alpha, beta)


@pytest.mark.parametrize("s", [16, 128])
def test_skew_symmetric_spmv_and_spmm(gpu, s):
    A = synth.random_sparse(900, 900, 0.03, 5)
    L = triangle(A, strict=True)
    full = expand(L, skew=True)
    M = AbhsfMatrix(block_size=s)
    M.from_csr(L.m, L.n, L.rowptr, L.col, L.val, symmetry="skew")
    x = np.random.default_rng(3).uniform(-1, 1, L.n)
    got = run_spmv(M, L.n, x, np.zeros(L.m), 1.0, 0.0)
    assert relerr(got, pyoracle.csr_spmv(full.rowptr, full.col, full.val, x)) <= TOL
    # x^T A x = 0 for a skew-symmetric A
    assert abs(float(x @ got)) <= 1e-10 * float(np.abs(x) @ np.abs(got))
    X = np.random.default_rng(4).uniform(-1, 1, (L.n, 3))
    dX, dY = DeviceArray.from_numpy(X), DeviceArray.zeros((L.m, 3))
    M.spmm_async(3, 2.0, dX, 3, 0.0, dY, 3)
    M.synchronize()
    Y = dY.numpy()
    for j in range(3):
        ref = pyoracle.csr_spmv(full.rowptr, full.col, full.val, X[:, j].copy(), 2.0)
        assert relerr(Y[:, j], ref) <= TOL, j


def test_symmetric_from_matrix_market(gpu, tmp_path):
    """The reader keeps the stored triangle (matrix_market_reader.h:191-197);
    from_elements with symmetric props applies both."""
    T = triangle(synth.laplacian2d(40))
    path = tmp_path / "lap_sym.mtx"
    r, c, v = T.triplets()
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real symmetric\n{T.m} {T.n} {T.nnz}\n")
        for a, b, w in zip(r, c, v):
            f.write(f"{a + 1} {b + 1} {float(w)!r}\n")
    A, props, nnz_all = mmio.read(str(path))
    assert props.symmetry == matrix_symmetry_t.SYMMETRIC and nnz_all == 2 * T.nnz - T.m
    M = AbhsfMatrix(block_size=32)
    M.from_elements(A.triplets(), props)
    x = np.full(T.n, 0.5)
    got = run_spmv(M, T.n, x, np.zeros(T.m), 1.0, 0.0)
    full = expand(T)
    assert relerr(got, pyoracle.csr_spmv(full.rowptr, full.col, full.val, x)) <= TOL
    # Laplacian rows sum to 4 - #neighbours: y = 0.5 * that
    deg = np.diff(full.rowptr) - 1
    assert np.allclose(got, 0.5 * (4 - deg))


def test_symmetric_rejects_both_triangles_and_skew_diagonal(gpu):
    A = synth.laplacian2d(20)  # both triangles
    M = AbhsfMatrix(block_size=16)
    with pytest.raises(AbhsfError, match="both L and U"):
        M.from_csr(A.m, A.n, A.rowptr, A.col, A.val, symmetry="symmetric")
    T = triangle(A)  # has the diagonal
    with pytest.raises(AbhsfError, match="diagonal"):
        M.from_csr(T.m, T.n, T.rowptr, T.col, T.val, symmetry="skew")
    # the handle stays usable
    M.from_csr(T.m, T.n, T.rowptr, T.col, T.val, symmetry="symmetric")
    x = np.random.default_rng(0).uniform(-1, 1, T.n)
    got = run_spmv(M, T.n, x, np.zeros(T.m), 1.0, 0.0)
    assert relerr(got, pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)) <= TOL
    # a batch member: the mirrored pass runs after the batched launch
    B = AbhsfBatch([M])
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.zeros(T.m)
    B.set_vectors([dx], [dy])
    B.spmv_async(1.0, 0.0)
    B.synchronize()
    assert relerr(dy.numpy(), pyoracle.csr_spmv(A.rowptr, A.col, A.val, x)) <= TOL
    B.release()


@pytest.mark.parametrize("skew", [False, True])
@pytest.mark.parametrize("s", [16, 256])
def test_symmetric_spmm_with_a_heavy_strip(gpu, skew, s):
    """An arrow matrix stored as its upper triangle: row 0 holds every column,
    so one strip holds more than 4096 entries and abhsf_spmm takes each
    right-hand side through the SpMV kernels; the mirrored triangle must be
    applied on that path too.  Integer data: exact."""
    n = 20000
    rng = np.random.default_rng(s)
    c = np.arange(n)
    r = np.concatenate([np.zeros(n, np.int64), np.arange(1, n)])
    cc = np.concatenate([c, np.arange(1, n)])
    v = (rng.integers(1, 9, r.size) * rng.choice([-1.0, 1.0], r.size)).astype(np.float64)
    if skew:  # no diagonal
        keep = r != cc
        r, cc, v = r[keep], cc[keep], v[keep]
    T = synth.from_triplets(n, n, r, cc, v)
    print(f"case symmetric arrow spmm s={s} skew={skew}", flush=True)
    M = AbhsfMatrix(block_size=s)
    M.from_csr(T.m, T.n, T.rowptr, T.col, T.val, symmetry="skew" if skew else "symmetric")
    st = M.stats()
    assert st["max_strip_work"] > max(4096, st["nnz"] // 4096)
    full = expand(T, skew)
    nrhs = 5
    X = rng.integers(-8, 9, (n, nrhs)).astype(np.float64)
    dX, dY = DeviceArray.from_numpy(X), DeviceArray.zeros((n, nrhs))
    M.spmm_async(nrhs, 1.0, dX, nrhs, 0.0, dY, nrhs)
    M.synchronize()
    got = dY.numpy()
    for j in range(nrhs):
        assert np.array_equal(got[:, j], pyoracle.csr_spmv(full.rowptr, full.col, full.val, X[:, j].copy())), j


@pytest.mark.parametrize("case", ["dense", "random-column", "mtx_general", "mtx_symmetric", "mtx_skew",
                                  "random-column:tiled", "mtx_general:tiled"])
def test_cpp_harness(gpu, case, tmp_path):
    """The C++ harness (abhsf_amd/csrc/spmv_unified.cpp, the reference's
    spmv/unified driver on this library): x = 0.5, 5 + 20 accumulating calls,
    printed ||y||/25 equals the printed expected ||A x||."""
    import os
    import re
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "abhsf_amd", "spmv_unified")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    case, _, variant = case.partition(":")
    if case.startswith("mtx"):
        kind = case[4:]
        A = synth.laplacian2d(30) if kind != "general" else synth.random_sparse(500, 500, 0.02, 7)
        T = A if kind == "general" else triangle(A, strict=(kind == "skew"))
        field = {"general": "general", "symmetric": "symmetric", "skew": "skew-symmetric"}[kind]
        path = tmp_path / f"{kind}.mtx"
        r, c, v = T.triplets()
        with open(path, "w") as f:
            f.write(f"%%MatrixMarket matrix coordinate real {field}\n{T.m} {T.n} {T.nnz}\n")
            for a, b, w in zip(r, c, v):
                f.write(f"{a + 1} {b + 1} {float(w)!r}\n")
        args = [exe, str(path), "0", "32"]
    else:
        args = [exe, case, "3000", "64"]
    if variant:
        args.append(variant)
    print("case harness", case, variant, flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    exp = float(re.search(r"Expected result = ([0-9.eE+-]+)", out.stdout).group(1))
    got = float(re.search(r"result = ([0-9.eE+-]+)\s*$", out.stdout.strip()).group(1))
    assert abs(got - exp) <= 1e-5 * max(exp, 1.0), out.stdout  # expected is printed with 6 digits

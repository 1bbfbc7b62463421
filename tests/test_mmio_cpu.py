"""Matrix Market input with the reference's semantics
(include/abhsf/utils/matrix_market_reader.h:41-200)."""
import os

import numpy as np
import pytest

from abhsf_amd import mmio, synth
from abhsf_amd.matrix import matrix_symmetry_t, matrix_type_t

from conftest import ROOT


def test_reads_reference_fixture():
    A, props, nnz_all = mmio.read(os.path.join(ROOT, "tests", "golden", "test.mtx"))
    # block_stats "props" for test.mtx: 6 7 15 15 15 2 3 (m n nnz_mtx nnz_stored nnz_all type symmetry)
    assert (props.m, props.n, props.nnz, nnz_all) == (6, 7, 15, 15)
    assert (int(props.type), int(props.symmetry)) == (2, 3)
    assert list(np.diff(A.rowptr)) == [3, 3, 1, 2, 3, 3]  # rows_nnz_stored


def test_drops_zeros_pattern_and_symmetric(tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% comment\n3 3 4\n1 1 2.0\n2 1 -1.0\n3 2 0.0\n3 3 5.0\n")
    A, props, nnz_all = mmio.read(str(p))
    assert props.symmetry == matrix_symmetry_t.SYMMETRIC and A.nnz == 3 and nnz_all == 4
    B, _, _ = mmio.read(str(p), expand_symmetric=True)
    D = np.zeros((3, 3))
    r, c, v = B.triplets()
    D[r, c] = v
    assert np.array_equal(D, [[2, -1, 0], [-1, 0, 0], [0, 0, 5]])
    q = tmp_path / "p.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 2\n2 1\n")
    P, pp, _ = mmio.read(str(q))
    assert pp.type == matrix_type_t.BINARY and list(P.val) == [1.0, 1.0]


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "coordinate"),
    ("not a banner\n", "banner"),
    ("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n2 1 1.0\n1 2 1.0\n", "both L and U"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n", "element"),
    ("%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 1 3\n", "non-real-non-binary"),
    ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 3 4\n", "non-real-non-binary"),
    # the triangle test sees stored zeros too (matrix_market_reader.h:175-177 before :179)
    ("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n2 1 1.0\n1 3 0.0\n", "both L and U"),
])
def test_malformed_inputs_raise(tmp_path, text, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(mmio.MatrixMarketError, match=msg):
        mmio.read(str(p))


def test_roundtrip(tmp_path):
    A = synth.random_sparse(50, 40, 0.1, seed=2)
    mmio.write(str(tmp_path / "a.mtx"), A)
    B, _, _ = mmio.read(str(tmp_path / "a.mtx"))
    assert np.array_equal(A.rowptr, B.rowptr) and np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)


@pytest.mark.parametrize("kind", ["general", "symmetric", "skew", "pattern"])
def test_binary_mm_roundtrip(tmp_path, kind):
    """utils/mtx2mm/mtx2mm.cpp:101-136 layout: 5 x u64 header, u32 rows, u32
    cols, f64 values (none for pattern); read with the .mtx semantics."""
    from abhsf_amd.matrix import matrix_properties
    A = synth.random_sparse(60, 60, 0.1, seed=3)
    r, c, v = A.triplets()
    if kind in ("symmetric", "skew"):
        keep = r > c if kind == "skew" else r >= c
        A = synth.from_triplets(60, 60, r[keep], c[keep], v[keep])
    sym = {"general": matrix_symmetry_t.UNSYMMETRIC, "symmetric": matrix_symmetry_t.SYMMETRIC,
           "skew": matrix_symmetry_t.SKEW_SYMMETRIC, "pattern": matrix_symmetry_t.UNSYMMETRIC}[kind]
    typ = matrix_type_t.BINARY if kind == "pattern" else matrix_type_t.REAL
    props = matrix_properties(A.m, A.n, A.nnz, typ, sym)
    path = str(tmp_path / "a.mm")
    mmio.write_mm(path, A, props)
    assert os.path.getsize(path) == 40 + 8 * A.nnz + (0 if kind == "pattern" else 8 * A.nnz)
    hdr = np.fromfile(path, dtype="<u8", count=5)
    assert list(hdr) == [A.m, A.n, A.nnz, int(typ), int(sym)]
    B, pb, nnz_all = mmio.read_mm(path)
    assert (pb.type, pb.symmetry) == (typ, sym)
    assert np.array_equal(A.rowptr, B.rowptr) and np.array_equal(A.col, B.col)
    assert np.array_equal(B.val, np.ones(A.nnz) if kind == "pattern" else A.val)
    off = int(np.sum(A.triplets()[0] != A.triplets()[1]))
    assert nnz_all == A.nnz + (off if sym != matrix_symmetry_t.UNSYMMETRIC else 0)
    # the same matrix through the text reader gives the same result
    mmio.write(str(tmp_path / "a.mtx"), A, "pattern" if kind == "pattern" else "real")
    if kind == "general":
        C, _, _ = mmio.read(str(tmp_path / "a.mtx"))
        assert np.array_equal(C.val, B.val) and np.array_equal(C.col, B.col)


def test_binary_mm_rejects_truncated_and_integer(tmp_path):
    p = tmp_path / "t.mm"
    np.array([2, 2, 3, 2, 3], dtype="<u8").tofile(str(p))
    with pytest.raises(mmio.MatrixMarketError, match="size"):
        mmio.read_mm(str(p))
    q = tmp_path / "i.mm"
    with open(q, "wb") as f:
        np.array([2, 2, 1, 1, 3], dtype="<u8").tofile(f)
        np.array([0], dtype="<u4").tofile(f)
        np.array([0], dtype="<u4").tofile(f)
        np.array([7], dtype="<i8").tofile(f)
    with pytest.raises(mmio.MatrixMarketError, match="non-real-non-binary"):
        mmio.read_mm(str(q))

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

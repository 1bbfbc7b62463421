"""CPU checks of the seeded synthetic generators (closed-form sizes,
determinism, sorted unique columns)."""
import numpy as np
import pytest

from abhsf_amd import synth


def well_formed(A):
    assert A.rowptr[0] == 0 and np.all(np.diff(A.rowptr) >= 0)
    assert A.col.min(initial=0) >= 0 and (A.col.size == 0 or A.col.max() < A.n)
    for r in range(0, A.m, max(1, A.m // 200)):
        c = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(c) > 0)


@pytest.mark.parametrize("g", [1, 2, 7, 64])
def test_laplacian_closed_form(g):
    A = synth.laplacian2d(g)
    assert A.nnz == 5 * g * g - 4 * g
    well_formed(A)
    assert np.all(A.val[A.col == np.repeat(np.arange(A.m), np.diff(A.rowptr))] == 4.0)


@pytest.mark.parametrize("g", [2, 5, 20])
def test_stencil27_closed_form(g):
    A = synth.stencil27(g)
    assert A.nnz == (3 * g - 2) ** 3
    well_formed(A)
    assert np.allclose(np.add.reduceat(A.val, A.rowptr[:-1]), [26 - (d - 1) for d in np.diff(A.rowptr)])


def test_block_mix_structure_and_determinism():
    A = synth.block_mix(1 << 16, seed=0)
    B = synth.block_mix(1 << 16, seed=0)
    C = synth.block_mix(1 << 16, seed=1)
    assert np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)
    assert not np.array_equal(A.rowptr, C.rowptr)
    well_formed(A)
    per_row = np.diff(A.rowptr)
    assert set(np.unique(per_row)) <= {12, 16, 24, 96}
    # strip types are constant over 32-row strips
    strips = per_row.reshape(-1, 32)
    kinds = np.where(strips.max(axis=1) == 96, 0, np.where(strips.max(axis=1) >= 16, 1, 2))
    frac = np.bincount(kinds, minlength=3) / kinds.size
    assert abs(frac[0] - 0.2) < 0.03 and abs(frac[1] - 0.3) < 0.03
    assert abs(np.mean(A.val)) < 0.02 and abs(np.std(A.val) - 1) < 0.02


def test_rmat_small():
    A = synth.rmat(12, 16, seed=0)
    B = synth.rmat(12, 16, seed=0)
    assert np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)
    well_formed(A)
    assert 0 < A.nnz <= 16 * 4096
    deg = np.diff(A.rowptr)
    assert deg[0] == deg.max()  # quadrant a = 0.57 concentrates on row 0
    assert A.val.min() >= -1 and A.val.max() < 1


def test_banded_and_random_blocks():
    A = synth.banded(100, 4, seed=0)
    assert A.nnz == 100 * 9 - 2 * (4 + 3 + 2 + 1)
    well_formed(A)
    B = synth.random_blocks(1000, 16, 4, 0.25, seed=0)
    well_formed(B)
    diag = np.repeat(np.arange(B.m), np.diff(B.rowptr)) == B.col
    assert diag.sum() == B.m  # diagonal always present

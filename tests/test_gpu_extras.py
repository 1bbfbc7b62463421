"""GPU checks of the reference's micro-benchmark closed form (spmv/orig/88.cpp)
per scheme, and of the serialised ABHSF blob (export -> import)."""
import numpy as np
import pytest

from abhsf_amd import AbhsfError, AbhsfMatrix, capi, synth
from abhsf_amd.hip import DeviceArray
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bnnz", [1, 2, 4, 8, 16, 32, 64])
def test_uniform_block_micro_benchmark_closed_form(gpu, bnnz):
    """88.cpp: 256 x 256 copies of one random 8x8 block with bnnz entries of
    value 1, x = 1, 5 warm-up + 20 accumulating calls: sum(y) =
    25 * bnnz * 256^2 (88.cpp:805-832); the block's scheme follows its z."""
    nb = 256
    A = synth.uniform_blocks(nb, 8, bnnz)
    M = AbhsfMatrix(block_size=8)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    want_scheme = pyoracle.choose(pyoracle.PREPROC, 8, bnnz)[0]
    st = M.stats()
    assert st["blocks"][capi.SCHEMES[want_scheme]] == nb * nb
    x, y = DeviceArray.from_numpy(np.ones(A.n)), DeviceArray.zeros(A.m)
    for _ in range(25):
        M.spmv_async(1.0, x, 1.0, y)
    M.synchronize()
    assert float(np.sum(y.numpy())) == 25.0 * bnnz * nb * nb


@pytest.mark.parametrize("s,prec", [(8, "f64"), (64, "f64"), (32, "f32"), (256, "f64")])
def test_blob_roundtrip(gpu, tmp_path, s, prec):
    A = synth.block_mix(4096, seed=3)
    M = AbhsfMatrix(block_size=s, precision=prec)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    path = str(tmp_path / "m.npz")
    M.save(path)
    L = AbhsfMatrix.load(path)
    a, b = M.export_all(), L.export_all()
    assert all(np.array_equal(a[k], b[k]) for k in a)
    assert M.stats()["smf_bits_preproc"] == L.stats()["smf_bits_preproc"]
    assert M.stats()["smf_bits_msmf"] == L.stats()["smf_bits_msmf"]
    x = DeviceArray.from_numpy(np.random.default_rng(1).uniform(-1, 1, A.n))
    y1, y2 = DeviceArray.zeros(A.m), DeviceArray.zeros(A.m)
    M.spmv_async(1.0, x, 0.0, y1)
    L.spmv_async(1.0, x, 0.0, y2)
    M.synchronize()
    L.synchronize()
    assert np.array_equal(y1.numpy(), y2.numpy())


def test_blob_rejects_inconsistent_arrays(gpu, tmp_path):
    A = synth.laplacian2d(40)
    M = AbhsfMatrix(block_size=16)
    M.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    path = str(tmp_path / "m.npz")
    M.save(path)
    with np.load(path) as f:
        arrays = {k: f[k].copy() for k in f.files}
    bad = dict(arrays)
    bad["coo_idx"] = bad["coo_idx"].copy()
    bad["coo_idx"][0] = 255  # lrow 15, lcol 15 in a block at the matrix edge or a wrong block
    bad["coo_bcol"] = bad["coo_bcol"].copy()
    bad["coo_bcol"][-1] = 60000  # block column beyond N
    np.savez(str(tmp_path / "bad.npz"), **bad)
    with pytest.raises(AbhsfError) as e:
        AbhsfMatrix.load(str(tmp_path / "bad.npz"))
    assert e.value.status == capi.ERR_INVALID

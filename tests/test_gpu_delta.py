"""Delta-coded coordinate storage (the reference's delta_matrix family,
spmv/unified/spmv.cpp:628-1033): the exported width streams, read back in
element order, are the oracle's restatement of the reference's varint
encoder byte for byte; SpMV against the oracle's CSR SpMV."""
import numpy as np
import pytest

from abhsf_amd import DeltaMatrix, synth
from abhsf_amd.hip import DeviceArray
from abhsf_amd.matrix import matrix_properties
from oracle import pyoracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
CHUNK = 512


def relerr(got, ref):
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / (den if den > 0 else 1.0))


def keys_of(A, K):
    r, c, _ = A.triplets()
    return (r.astype(np.uint64) << np.uint64(K)) + c.astype(np.uint64)


def stream_in_element_order(D, nnz):
    """Concatenate the tagged words of elements 1..nnz-1 in element order,
    little-endian bytes, from the exported width streams and tags."""
    tags = D.export("tags")
    words = [D.export(w) for w in ("w8", "w16", "w32", "w64")]
    t = ((tags[np.arange(nnz) >> 4] >> (2 * (np.arange(nnz) & 15)).astype(np.uint32)) & 3).astype(np.int64)
    idx = np.zeros(4, dtype=np.int64)
    out, word_of = [], []
    for e in range(nnz):
        w = int(words[t[e]][idx[t[e]]])
        idx[t[e]] += 1
        assert w & 3 == t[e], e  # the in-band tag equals the out-of-band one
        word_of.append(w)
        if e > 0:
            out += list(w.to_bytes(1 << t[e], "little"))
    assert list(idx) == [w.size for w in words]
    return np.array(out, dtype=np.uint8), word_of, t


MATS = {
    "laplacian_60": lambda: synth.laplacian2d(60),
    "rmat_13": lambda: synth.rmat(13, 8, seed=2),
    "random_900x3000": lambda: synth.random_sparse(900, 3000, 0.01, seed=3),
    "block_mix_8192": lambda: synth.block_mix(8192, seed=4),
    "one_row": lambda: synth.random_sparse(1, 5000, 0.02, seed=5),
    "single": lambda: synth.from_triplets(10, 10, [3], [7], [2.5]),
    "empty": lambda: synth.from_triplets(300, 200, [], [], []),
    # row gaps of 2^11 at K = 20: deltas of about 2^31 take the 8-byte words
    "wide_gaps": lambda: synth.from_triplets(1 << 21, 1 << 20, np.arange(0, 1 << 21, 2048),
                                             (np.arange(1024) * 997) % (1 << 20), np.linspace(-1, 1, 1024)),
}


@pytest.mark.parametrize("name", list(MATS))
def test_delta_layout_and_spmv(gpu, name):
    A = MATS[name]()
    print(f"case delta {name} nnz={A.nnz}", flush=True)
    D = DeltaMatrix()
    D.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    st = D.stats()
    K = max(1, (A.n - 1).bit_length())
    assert st["key_bits"] == K and st["nnz"] == A.nnz
    keys = keys_of(A, K)
    if A.nnz:
        stream, word_of, t = stream_in_element_order(D, A.nnz)
        assert np.array_equal(stream, pyoracle.varint_encode(keys))
        assert word_of[0] >> 2 == int(keys[0])  # element 0: delta from key 0
        # chunk bases and stream starts
        base, start = D.export("base"), D.export("start").reshape(-1, 4)
        nch = (A.nnz + CHUNK - 1) // CHUNK
        assert base.size == nch and start.shape == (nch, 4)
        for c in range(nch):
            assert base[c] == (0 if c == 0 else keys[c * CHUNK - 1])
            assert list(start[c]) == [int(np.sum(t[:c * CHUNK] == w)) for w in range(4)]
        assert np.array_equal(D.export("val"), A.val)
    if name == "wide_gaps":
        assert st["words"][8] > 0
    x = np.random.default_rng(1).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(2).uniform(-1, 1, A.m)
    for alpha, beta in ((1.0, 0.0), (-0.75, 2.0), (1.0, 1.0)):
        dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y0)
        D.spmv_async(alpha, dx, beta, dy)
        D.synchronize()
        ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x, alpha, beta, y0)
        assert relerr(dy.numpy(), ref) <= TOL, (alpha, beta)


def test_delta_from_elements_accumulates(gpu):
    """delta_matrix::spmv(x, y) accumulates (y += A x), the harness protocol
    of spmv/unified/spmv.cpp:1176-1232 with x = 0.5."""
    A = synth.laplacian2d(50)
    r, c, v = A.triplets()
    perm = np.random.default_rng(0).permutation(A.nnz)
    D = DeltaMatrix()
    D.from_elements((r[perm], c[perm], v[perm]), matrix_properties(A.m, A.n, A.nnz))
    x = DeviceArray.from_numpy(np.full(A.n, 0.5))
    y = DeviceArray.zeros(A.m)
    for _ in range(25):
        D.spmv(x, y)
    ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, np.full(A.n, 0.5))
    assert abs(np.linalg.norm(y.numpy()) / 25 - np.linalg.norm(ref)) <= 1e-12 * np.linalg.norm(ref)


U32_MATS = {
    "rmat_14": lambda: synth.rmat(14, 16, seed=2),             # empty rows: escapes in row-major order
    "laplacian_120": lambda: synth.laplacian2d(120),
    "block_mix_2^13": lambda: synth.block_mix(1 << 13, 4),
    "wide_30x200000": lambda: synth.random_sparse(30, 200000, 0.001, seed=7),
    "sparse_rows": lambda: synth.from_triplets(5000, 700, [0, 0, 1999, 4999], [0, 699, 5, 3], [1.0, 2.0, 3.0, 4.0]),
    # one entry per row: a 64-element step spans 8 row bands, so cache-block
    # keys take both the per-wave band bins and the global adds past them
    "diagonal_3000": lambda: synth.from_triplets(3000, 3000, np.arange(3000), np.arange(3000),
                                                 np.arange(1, 3001, dtype=np.float64)),
    "empty": lambda: synth.from_triplets(64, 64, [], [], []),
}


@pytest.mark.parametrize("codec", ["u32_row", "u32_block"])
@pytest.mark.parametrize("name", list(U32_MATS))
def test_u32_codecs_match_oracle_encoder(gpu, codec, name):
    """The u32 delta codecs (spmv/orig/delta.cpp:124-151 row << 31 | col;
    spmv/unified/spmv.cpp:636-689 8 x 8 cache blocks): every exported array
    equals the oracle encoder's, values in the codec's element order, and the
    SpMV matches the CSR product."""
    A = U32_MATS[name]()
    layout = pyoracle.DELTA32_ROWMAJOR if codec == "u32_row" else pyoracle.DELTA32_CACHEBLOCK
    print(f"case delta {codec} {name}", flush=True)
    D = DeltaMatrix(codec=codec)
    D.from_csr(A.m, A.n, A.rowptr, A.col, A.val)
    st = D.stats()
    assert st["codec"] == codec and st["key_bits"] == 31
    enc = pyoracle.delta32_encode(layout, A.m, A.rowptr, A.col, CHUNK)
    assert np.array_equal(D.export("da"), enc["da"])
    assert np.array_equal(D.export("base"), enc["base"])
    if A.nnz:
        assert np.array_equal(D.export("xstart"), enc["xstart"])
    assert np.array_equal(D.export("exc"), enc["exc"])
    assert np.array_equal(D.export("val"), A.val[enc["perm"].astype(np.int64)])
    assert st["words"][8] == enc["exc"].size and st["words"][4] == A.nnz - enc["exc"].size
    x = np.random.default_rng(3).uniform(-1, 1, A.n)
    y0 = np.random.default_rng(4).uniform(-1, 1, A.m)
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0), (1.0, 1.0)):
        dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y0)
        D.spmv_async(alpha, dx, beta, dy)
        D.synchronize()
        ref = pyoracle.csr_spmv(A.rowptr, A.col, A.val, x, alpha, beta, y0)
        assert relerr(dy.numpy(), ref) <= TOL, (alpha, beta)
    D.release()


def test_u32_block_rejects_wide_matrices(gpu):
    """The cache-block key holds column blocks in 25 bits (n <= 2^28)."""
    from abhsf_amd import AbhsfError
    D = DeltaMatrix(codec="u32_block")
    with pytest.raises(AbhsfError, match="2\\^28"):
        D.from_csr(1, (1 << 28) + 1, np.array([0, 1]), np.array([5]), np.array([1.0]))

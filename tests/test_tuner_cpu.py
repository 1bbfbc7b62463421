"""CPU checks of the tuner's oracle (oracle/mlp_ref.py) and of the tuner's
storage-cost model against the oracle builder."""
import numpy as np
import pytest

from abhsf_amd import synth, tuner
from oracle import mlp_ref, pyoracle


def test_mlp_ref_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    F = 7
    theta = mlp_ref.init(F, seed=1) * 0.5
    X = rng.standard_normal((13, F))
    y = rng.standard_normal(13)
    w = rng.uniform(0.5, 2.0, 13)
    g = mlp_ref.grad(theta, X, y, w)
    idx = rng.choice(theta.size, 80, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mlp_ref.loss(tp, X, y, w) - mlp_ref.loss(tm, X, y, w)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_mlp_ref_closed_form():
    """W1 = identity on the first F hidden units, W2 = identity, W3 = ones:
    out = sum(relu(x)) + b3."""
    F = 5
    L = mlp_ref.layout(F)
    t = np.zeros(L["P"])
    W1 = np.zeros((64, F))
    W1[:F, :F] = np.eye(F)
    t[L["W1"][0]:L["W1"][1]] = W1.ravel()
    t[L["W2"][0]:L["W2"][1]] = np.eye(64).ravel()
    t[L["W3"][0]:L["W3"][1]] = 1.0
    t[L["b3"][0]] = 0.25
    X = np.array([[1.0, -2.0, 3.0, -4.0, 5.0], [-1.0, -1.0, -1.0, -1.0, -1.0]])
    _, _, out = mlp_ref.forward(t, X)
    assert np.allclose(out, [9.25, 0.25])


@pytest.mark.parametrize("s", [2, 8, 16, 64, 256])
def test_tuner_choose_matches_oracle(s):
    for z in range(1, s * s + 1, max(1, s * s // 500)):
        assert tuner.choose(s, z) == pyoracle.choose(pyoracle.PREPROC, s, z)[0]


@pytest.mark.parametrize("s", [8, 32, 64, 128])
def test_layout_bytes_matches_oracle_storage(s):
    for A in (synth.block_mix(4096, seed=3), synth.rmat(11, 8, seed=2), synth.laplacian2d(40)):
        B = pyoracle.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, s.bit_length() - 1)
        arrays = B.arrays()
        stored = sum(a.nbytes for a in arrays.values())
        assert tuner.layout_bytes(s, A.m, A.n, B.nb, [int(v) for v in B.st.ne]) == stored
        # features from an exact histogram reproduce the oracle's block counts
        r, c, _ = A.triplets()
        h = pyoracle.block_hist(r, c, s.bit_length() - 1, s.bit_length() - 1)
        hist = np.zeros(s * s + 2, dtype=np.uint64)
        for z, cnt in h.items():
            hist[z] = cnt
        sf = tuner.size_features(A.rowptr, A.m, A.n, s, hist, sum(h.values()), A.nnz)
        assert sf["stored_bytes"] == pytest.approx(stored)
        tot = sum(B.nb)
        assert np.allclose(sf["blk_frac"], [v / tot for v in B.nb])


def test_feature_vector_length_matches_nfeat():
    from abhsf_amd import tuner
    g = {"log2_n": 1.0, "log2_nnz": 1.0, "log2_mean_row": 0.0, "row_cv": 0.0, "log2_max_row": 0.0,
         "empty_rows": 0.0}
    sf = {"log2_s": 3.0, "log2_blk_per_entry": 0.0, "blk_frac": [1, 0, 0, 0], "ent_frac": [1, 0, 0, 0],
          "log2_bytes_per_entry": 3.5, "strip_cv": 0.0, "log2_max_strip": 0.0, "heavy_frac": 0.0, "log2_M": 4.0,
          "tile_frac": 0.5}
    for v in tuner.VARIANTS + tuple(tuner.DELTA_CODECS):
        assert tuner.feature_vector(g, sf, v).shape == (tuner.NFEAT,)


@pytest.mark.parametrize("s", [2, 4, 8, 16, 32, 64, 128, 256])
def test_coo_blocks_hold_at_most_2s_entries(s):
    """COO is kept while 2kz + log s^2 <= kz + (s+1) log s^2 (the tie at
    z = 2s keeps COO: strict < in preprocessing/test.cpp:55-77), so COO
    blocks hold z <= 2s entries and z - 1 fits k + 1 bits: the u16 z1 of the
    tiled variant has 15 - k bits left for the strip index."""
    k = s.bit_length() - 1
    zc = tuner.coo_max_z(s)
    assert tuner.choose(s, zc) == 0 and tuner.choose(s, zc + 1) != 0
    assert all(tuner.choose(s, z) == 0 for z in range(1, zc + 1))
    assert zc <= 2 * s and zc - 1 < (1 << (k + 1))


@pytest.mark.parametrize("seed", range(4))
def test_tile_fraction_brute_force(seed):
    rng = np.random.default_rng(seed)
    A = synth.block_mix(1 << 11, seed, 0.2, 0.3) if seed % 2 else synth.rmat(11, 8, seed=seed)
    for s in (4, 16, 64):
        for stride in (1, 3):
            got = tuner.tile_fraction(A.rowptr, A.col, A.m, s, stride)
            M = -(-A.m // s)
            tot = pure = 0
            for I in range(0, M, stride):
                r0, r1 = I * s, min(I * s + s, A.m)
                cols = A.col[A.rowptr[r0]:A.rowptr[r1]]
                zs = {}
                for c in cols:
                    zs[int(c) // s] = zs.get(int(c) // s, 0) + 1
                w = len(cols)
                tot += w
                if w <= 2 * tuner.unit_entries(s) and all(tuner.choose(s, z) == 0 for z in zs.values()):
                    pure += w
            assert got == pytest.approx(pure / tot if tot else 0.0), (s, stride)


@pytest.mark.parametrize("name", ["rmat_13", "laplacian_60", "sparse_rows", "band_gaps", "one"])
def test_u32_delta_features_match_oracle_encoder(name):
    """The tuner's closed-form counts for the u32 delta codecs (escapes, stored
    bytes) equal what the oracle's encoder produces."""
    from oracle import pyoracle
    A = {"rmat_13": lambda: synth.rmat(13, 8, seed=5), "laplacian_60": lambda: synth.laplacian2d(60),
         "sparse_rows": lambda: synth.from_triplets(5000, 700, [0, 0, 1999, 4999], [0, 699, 5, 3], [1.0] * 4),
         "band_gaps": lambda: synth.from_triplets(100, 100, [0, 7, 16, 17, 40, 41, 99], [50, 1, 60, 2, 0, 99, 99],
                                                  [1.0] * 7),
         "one": lambda: synth.from_triplets(3, 3, [2], [2], [1.0])}[name]()
    for codec, layout in (("u32_row", pyoracle.DELTA32_ROWMAJOR), ("u32_block", pyoracle.DELTA32_CACHEBLOCK)):
        f = tuner.delta_features(A.rowptr, A.col, A.n, codec)
        enc = pyoracle.delta32_encode(layout, A.m, A.rowptr, A.col)
        nexc = enc["exc"].size
        assert round(f["ent_frac"][3] * A.nnz) == nexc, codec
        chunks = -(-A.nnz // 512)
        assert f["stored_bytes"] == 12 * A.nnz + 8 * nexc + 8 * chunks + 4 * (chunks + 1)


def test_sampled_delta_features_estimate_exact_counts():
    """Above SAMPLE_NNZ entries the delta features come from a weighted sample
    of row bands: on an R-MAT of 4e6 entries the estimated word classes and
    exception counts are close to the exact counts of the oracle encoders."""
    from oracle import pyoracle
    A = synth.rmat(18, 16, seed=3)
    assert A.nnz > tuner.SAMPLE_NNZ
    smp = tuner._band_sample(A.rowptr, A.col)
    assert smp[0].size < 0.8 * A.nnz  # a sample, not every element
    # varint word classes against the exact deltas of row-major keys
    K = max(1, tuner._ceil_log2(A.n))
    r, c, _ = A.triplets()
    keys = (r.astype(np.uint64) << np.uint64(K)) + c.astype(np.uint64)
    d = np.diff(keys, prepend=np.uint64(0))
    exact = np.array([np.mean(d < 64), np.mean((d >= 64) & (d < 1 << 14)), np.mean((d >= 1 << 14) & (d < 1 << 30))])
    est = np.array(tuner.delta_features(A.rowptr, A.col, A.n, "varint", smp)["ent_frac"][:3])
    assert np.all(np.abs(est - exact) < 0.05), (est, exact)
    for codec, layout in (("u32_row", pyoracle.DELTA32_ROWMAJOR), ("u32_block", pyoracle.DELTA32_CACHEBLOCK)):
        nexc = pyoracle.delta32_encode(layout, A.m, A.rowptr, A.col)["exc"].size
        f = tuner.delta_features(A.rowptr, A.col, A.n, codec, smp)
        assert abs(f["ent_frac"][3] * A.nnz - nexc) <= 0.35 * nexc + 50, (codec, f["ent_frac"][3] * A.nnz, nexc)

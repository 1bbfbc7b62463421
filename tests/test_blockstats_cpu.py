"""The block-statistics tables of abhsf_amd/blockstats.py (mmf rows, best
block size, per-size minima, block-size-subset regret) against the
reference's own programs (oracle/_ref/{block_stats,mmf,min-bss,min-schemes,
consistency}, built from block_statistics/*.cpp) on a set of matrices.
Histograms here come from the oracle (the GPU ones are compared with the
oracle in tests/test_tuner_gpu.py)."""
import os
import re
import subprocess

import numpy as np
import pytest

from abhsf_amd import blockstats, synth
from oracle import pyoracle

from conftest import ROOT, read_mtx

REF = os.path.join(ROOT, "oracle", "_ref")
ANSI = re.compile(r"\x1b\[[0-9;]*m")


def need_ref(*names):
    for n in names:
        if not os.path.exists(os.path.join(REF, n)):
            pytest.skip("reference binaries not built (make -C oracle ref needs /root/reference)")


def matrices():
    m, n, r, c, v = read_mtx(os.path.join(ROOT, "tests", "golden", "test.mtx"))
    return [("test", synth.from_triplets(m, n, r, c, v)), ("lap", synth.laplacian2d(30)),
            ("rnd", synth.random_sparse(200, 180, 0.05, seed=3)), ("rmat", synth.rmat(9, 8, seed=1))]


def oracle_hists(A):
    r, c, _ = A.triplets()
    return {(1 << a, 1 << b): pyoracle.block_hist(r, c, a, b) for a in range(1, 9) for b in range(1, 9)}


def write_mtx(path, A):
    r, c, v = A.triplets()
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{A.m} {A.n} {A.nnz}\n")
        for a, b, x in zip(r, c, v):
            f.write(f"{a + 1} {b + 1} {float(x)!r}\n")


@pytest.fixture(scope="module")
def refdir(tmp_path_factory):
    need_ref("block_stats", "mmf", "min-bss", "min-schemes", "consistency")
    d = tmp_path_factory.mktemp("bstats")
    for name, A in matrices():
        sub = d / name
        sub.mkdir()
        write_mtx(sub / "a.mtx", A)
        subprocess.run([os.path.join(REF, "block_stats"), "a.mtx"], cwd=sub, check=True, capture_output=True)
        subprocess.run([os.path.join(REF, "mmf")], cwd=sub, check=True, capture_output=True)
    return d


def read_mmf(path):
    t = {}
    for line in open(path):
        v = [int(x) for x in line.split()]
        t[(v[0], v[1])] = v[2:]
    return t


def test_mmf_tables_match_reference(refdir):
    for name, A in matrices():
        h = oracle_hists(A)
        for mult, fn in ((1, "mmf-single"), (2, "mmf-double")):
            assert blockstats.mmf_table(A, h, mult=mult) == read_mmf(refdir / name / fn), (name, fn)


def test_min_bss_and_min_schemes_match_reference(refdir):
    subprocess.run([os.path.join(REF, "min-bss")], cwd=refdir, check=True, capture_output=True)
    subprocess.run([os.path.join(REF, "min-schemes")], cwd=refdir, check=True, capture_output=True)
    lines = open(refdir / "mmf-bss-double").read().split("\n")
    header = lines[0].split()[2:]
    rows = {}
    for line in lines[1:]:
        if line.strip():
            v = line.split()
            rows[(int(v[0]), int(v[1]))] = [int(x) for x in v[2:]]
    tables = {name: blockstats.mmf_table(A, oracle_hists(A)) for name, A in matrices()}
    for j, name in enumerate(header):
        mins = blockstats.min_per_block_size(tables[name])
        for bs, vals in rows.items():
            assert mins[bs] == vals[j], (name, bs)
    for line in open(refdir / "mmf-schemes-double"):
        v = line.split()
        t = tables[v[0]]
        assert [min(r[k] for r in t.values()) for k in range(8)] == [int(x) for x in v[1:9]]
        _, (bs, _) = blockstats.block_opt(t)
        assert bs == (int(v[9]), int(v[10])), v[0]


@pytest.mark.parametrize("scheme", [0, 1])
def test_subset_regret_matches_reference(refdir, scheme):
    out = subprocess.run([os.path.join(REF, "consistency"), str(scheme), "1", "4"], cwd=refdir, check=True,
                         capture_output=True, text=True).stdout
    out = ANSI.sub("", out)
    want = {m.group(1): (float(m.group(2)), float(m.group(3)))
            for m in re.finditer(r"(B_\d+): Average = ([0-9.e+-]+), Maximum = ([0-9.e+-]+)", out)}
    tables = [blockstats.mmf_table(A, oracle_hists(A)) for _, A in matrices()]
    got = blockstats.subset_regret(tables, scheme_adaptive=bool(scheme))
    for k, (avg, mx) in want.items():
        assert got[k]["average"] == pytest.approx(avg, rel=1e-5, abs=1e-9), k
        assert got[k]["maximum"] == pytest.approx(mx, rel=1e-5, abs=1e-9), k


def test_rows_stats_match_reference(refdir):
    for name, A in matrices():
        rs = [int(x) for x in open(refdir / name / "rows_nnz_stored").read().split()]
        got = blockstats.rows_cols_nnz(A)
        assert list(got[0]) == rs
        assert abs(blockstats.rows_nnz_stdev(rs) - pyoracle.rows_stdev(rs)) < 1e-12


def test_morton_keys_and_delta_widths_brute_force():
    """morton.cpp:22-30 interleave and :119-147 gap histogram, restated by a
    plain per-bit loop and checked on hand cases."""
    from abhsf_amd import blockstats
    # hand cases: row bits go to even positions, column bits to odd ones
    assert list(blockstats.morton_keys([1, 0, 3, 2], [0, 1, 0, 3])) == [1, 2, 5, 14]
    A = synth.random_sparse(70, 90, 0.08, seed=4)
    r, c, _ = A.triplets()

    def morton(a, b):
        k = 0
        for i in range(32):
            k |= ((a >> i) & 1) << (2 * i) | ((b >> i) & 1) << (2 * i + 1)
        return k
    keys = sorted(morton(int(a), int(b)) for a, b in zip(r, c))
    assert list(blockstats.morton_keys(r, c)[np.argsort(blockstats.morton_keys(r, c))]) == keys
    hist = {}
    gaps = [b - a for a, b in zip(keys, keys[1:])]
    for g in gaps:
        k = -1 if g == 0 else (0 if g <= 1 else (g - 1).bit_length())
        hist[k] = hist.get(k, 0) + 1
    got = blockstats.delta_widths(A, "morton")
    assert got["hist"] == dict(sorted(hist.items()))
    assert got["largest"] == sorted(gaps, reverse=True)[:50]
    # row order: the delta format's keys; its gaps are the codec's deltas
    from oracle import pyoracle
    K = (A.n - 1).bit_length()
    rk = (r.astype(np.uint64) << np.uint64(K)) + c.astype(np.uint64)
    w = blockstats.delta_widths(A, "row")
    assert sum(w["hist"].values()) == A.nnz - 1 and -1 not in w["hist"]
    d = np.diff(rk)
    assert pyoracle.varint_encode(rk).size == sum(1 if g < 64 else 2 if g < 1 << 14 else 4 if g < 1 << 30 else 8
                                                  for g in d)

"""GPU index work bit-exact with the oracle: the stable counting sort
(partition labels -> permutation, and the SpMV row binning), CSR and dense
permutations (src/permute.c:120-285), read_partition (src/gooseberry.c:253-306).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,kmin,kmax", [
    (0, 0, 3), (1, 0, 0), (1023, 0, 2), (1024, 5, 9), (100001, 0, 255),
    (50000, 0, 256), (70000, 10, 70009), (3000, 0, 4294967294 // 4096),
])
def test_countingsort_matches_oracle(gb, oracle, n, kmin, kmax):
    rng = np.random.default_rng(n + kmax)
    keys = rng.integers(kmin, kmax + 1, n).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    out, dist = gb.countingsort_kv(keys, kmin, kmax, vals=vals)
    rout, rdist = oracle.countingsort_kv(keys, kmin, kmax, vals=vals)
    assert np.array_equal(out, rout) and np.array_equal(dist, rdist)


@pytest.mark.parametrize("case", ["tile_edges", "all_equal", "four_passes", "skewed_gaps",
                                  "no_values"])
def test_countingsort_tiles_and_distributions(gb, oracle, case):
    # tiles of 4096 keys: partial last tiles, a single digit everywhere,
    # full 32-bit keys, heavy and empty key ranges, the index payload
    rng = np.random.default_rng(len(case))
    vals = None
    if case == "tile_edges":
        for n in (4095, 4096, 4097, 3 * 4096 + 1):
            keys = rng.integers(0, 1000, n).astype(np.uint32)
            v = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
            out, dist = gb.countingsort_kv(keys, 0, 999, vals=v)
            rout, rdist = oracle.countingsort_kv(keys, 0, 999, vals=v)
            assert np.array_equal(out, rout) and np.array_equal(dist, rdist)
        return
    if case == "all_equal":
        keys, kmin, kmax = np.full(300000, 77, np.uint32), 0, 200
    elif case == "four_passes":
        keys = rng.integers(0, 1 << 32, 200000, dtype=np.uint64).astype(np.uint32)
        kmin, kmax = 0, 0xffffffff - 1
        keys = np.minimum(keys, kmax)
    elif case == "skewed_gaps":
        keys = (rng.zipf(1.5, 500000) % 70000).astype(np.uint32) * 3
        kmin, kmax = 0, 3 * 70000
    else:
        keys = rng.integers(5, 50000, 123457).astype(np.uint32)
        kmin, kmax = 5, 49999
    if case != "no_values":
        vals = rng.integers(0, 1 << 32, keys.size, dtype=np.uint64).astype(np.uint32)
    want_dist = kmax - kmin < 1 << 24
    out, dist = gb.countingsort_kv(keys, kmin, kmax, vals=vals, want_dist=want_dist)
    # a stable sort's order, and dist[k] = keys below kmin + k
    order = np.argsort(keys, kind="stable")
    rout = order.astype(np.uint32) if vals is None else vals[order]
    assert np.array_equal(out, rout)
    if want_dist:
        rdist = np.searchsorted(np.sort(keys), np.arange(kmin, kmax + 2, dtype=np.int64))
        assert np.array_equal(dist, rdist.astype(np.uint32))


def test_countingsort_rejects_out_of_range_keys(gb):
    with pytest.raises(gb.GooseberryError) as e:
        gb.countingsort_kv(np.array([0, 5, 9], np.uint32), 0, 8)
    assert e.value.name == "INVALIDINPUT"


def test_permute_sparse_matches_oracle(gb, oracle):
    from test_gpu_spmv import irregular
    n = 20000
    rp, ri, rv = irregular(n, 21, [(3, 5000), (4, 0)])
    rng = np.random.default_rng(22)
    rperm = rng.permutation(n).astype(np.uint32)
    cperm = rng.permutation(n).astype(np.uint32)
    for a, b in ((rperm, cperm), (rperm, None), (None, cperm)):
        got = gb.permute_sparse(rp, ri, rv, n, a, b)
        ref = oracle.permute_sparse(rp, ri, rv, n, a, b)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)


def test_permute_dense_matches_oracle(gb, oracle):
    rng = np.random.default_rng(23)
    v = rng.uniform(-1, 1, 5000 * 3)
    rp = rng.permutation(5000).astype(np.uint32)
    cp = rng.permutation(3).astype(np.uint32)
    for rev in (False, True):
        for a, b in ((rp, cp), (rp, None), (None, cp)):
            assert np.array_equal(gb.permute_dense(v, 5000, 3, a, b, rev),
                                  oracle.permute_dense(v, 5000, 3, a, b, rev))


def test_permute_rejects_non_permutation(gb):
    rp, ri, rv = gb.laplacian(2, 4)
    bad = np.arange(16, dtype=np.uint32)
    bad[3] = 2
    with pytest.raises(gb.GooseberryError):
        gb.permute_sparse(rp, ri, rv, 16, bad, None)


def test_read_partition_matches_oracle(gb, oracle, tmp_path):
    rng = np.random.default_rng(24)
    labels = rng.integers(0, 37, 10000)
    p = os.path.join(str(tmp_path), "x.part")
    with open(p, "w") as f:
        f.write("# partition\n" + "\n".join(map(str, labels)) + "\n")
    got = gb.read_partition(p)
    ref = oracle.read_partition(p)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)

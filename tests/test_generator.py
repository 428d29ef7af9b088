"""CPU checks of the generator restatement (tests/spd_restatement.py) that
the GPU generator is compared with: the hash against SplitMix64's published
output, the Zipf sampler against the exact Zipf(1.5) law, and the
symmetrised matrix's structure."""
import numpy as np
import scipy.sparse as sp
from scipy.special import zeta

import spd_restatement as R


def test_mix64_is_splitmix64_output():
    # SplitMix64 from state 0: first output mix64(0x9e3779b97f4a7c15)
    z = R.mix64(np.asarray([0x9e3779b97f4a7c15], dtype=np.uint64))
    assert int(z[0]) == 0xe220a8397b1dcdaf


def test_zipf_sampler_follows_the_law():
    n = 400000
    d = R.zipf_degrees(n, 11, 1.5, 1, 1 << 30)
    z = zeta(1.5)
    for k in range(1, 9):
        p = k ** -1.5 / z
        sd = np.sqrt(p * (1 - p) / n)
        assert abs(np.mean(d == k) - p) < 5 * sd, k
    # tail beyond 4095: sum_{k>=4096} k^-1.5 / zeta(1.5)
    tail = zeta(1.5, 4096) / z
    sd = np.sqrt(tail * (1 - tail) / n)
    assert abs(np.mean(d >= 4096) - tail) < 5 * sd


def test_zipf_clipped_mean():
    # E[min(max(Z, 2), 4096)] for Z ~ Zipf(1.5), from the exact pmf
    z = zeta(1.5)
    k = np.arange(1, 4096, dtype=np.float64)
    p = k ** -1.5 / z
    mean = float(np.sum(np.maximum(k, 2) * p) + 4096 * (1 - p.sum()))
    assert abs(mean - 97.82) < 0.01
    d = R.zipf_degrees(400000, 2, 1.5, 2, 4096)
    assert d.min() >= 2 and d.max() <= 4096
    assert abs(d.mean() - mean) < 5 * 450 / np.sqrt(400000)


def test_uniform_degrees_cover_the_range():
    d = R.uniform_degrees(100000, 5, 3, 27)
    assert d.min() == 3 and d.max() == 27
    counts = np.bincount(d - 3)
    assert counts.min() > 0.9 * 4000 and counts.max() < 1.1 * 4000


def test_symmetrised_matrix_structure():
    n = 4000
    rp, ri, rv = R.generate_uniform(n, 3, 3, 27)
    rp64 = rp.astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp64))
    cols = ri.astype(np.int64)
    o1, o2 = np.lexsort((rv, cols, rows)), np.lexsort((rv, rows, cols))
    assert np.array_equal(rows[o1], cols[o2]) and np.array_equal(cols[o1], rows[o2])
    assert np.array_equal(rv[o1], rv[o2])
    A = sp.csr_matrix((rv.copy(), ri.copy(), rp.copy()), shape=(n, n))
    last = rp64[1:] - 1
    assert np.all(ri[last] == np.arange(n))
    off = np.abs(A).sum(axis=1).A1 - rv[last]
    assert np.allclose(rv[last], 1.01 * off, rtol=1e-13)
    # every off-diagonal value in [-1, 0), no off-diagonal on the diagonal
    mask = np.ones(len(ri), dtype=bool)
    mask[last] = False
    assert np.all(rv[mask] < 0) and np.all(rv[mask] >= -1)
    assert np.all(ri[mask] != rows[mask])
    # strictly diagonally dominant with positive diagonal: SPD
    assert np.linalg.eigvalsh(A[:300, :300].toarray()).min() > 0
    d = R.uniform_degrees(n, 3, 3, 27)
    assert int(rp[-1]) == n + 2 * int((d - 1).sum())

"""Pinning of the oracle's sparse x sparse products (src/blas.c:94-234,
609-791) on the CPU.

The reference's tests hold no case for these functions, so the oracle is
pinned by what the mathematics fixes:
  * integer-valued entries make every sum exact whatever its order, so both
    products must equal scipy's A @ B exactly (a dropped, doubled or
    transposed term fails this);
  * the identity on either side returns the other operand entry for entry,
    in its own order (the sparse fill keeps a row's first-appearance order);
  * the sparse result's column order is checked against an independent
    formulation (numpy's first-occurrence indices over the flattened scan),
    and its values against the dense product element by element: the two
    sum the same products in the same order, one from +0.0, one from the
    first product.
"""
import numpy as np
import pytest
import scipy.sparse as sp


def random_csr(nrows, ncols, deg, seed, integer=False, empty_every=0, dups=False):
    """Rows of `deg` random columns (with repeats when dups), some empty."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(nrows):
        if empty_every and i % empty_every == 0:
            rows.append(np.zeros(0, dtype=np.int64))
            continue
        k = int(rng.integers(0, 2 * deg + 1))
        if dups:
            cols = rng.integers(0, ncols, size=k)
        else:
            cols = rng.choice(ncols, size=min(k, ncols), replace=False)
        rows.append(cols)
    rp = np.zeros(nrows + 1, dtype=np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ri = np.concatenate(rows).astype(np.uint32) if rp[-1] else np.zeros(0, np.uint32)
    if integer:
        rv = rng.integers(-4, 5, ri.size).astype(np.float64)
    else:
        rv = rng.uniform(-1.0, 1.0, ri.size)
    return rp, ri, rv


def scipy_of(m, nrows, ncols):
    rp, ri, rv = m
    # scipy sums duplicates itself; copies keep the inputs untouched
    return sp.csr_matrix((rv.copy(), ri.copy(), rp.copy()), shape=(nrows, ncols))


def scan_first_order(a, b):
    """Per row of C, the columns in order of their first product, from the
    flattened scan positions (an independent formulation)."""
    arp, ari, _ = a
    brp, bri, _ = b
    out = []
    for i in range(arp.size - 1):
        cols = [bri[brp[l]:brp[l + 1]] for l in ari[arp[i]:arp[i + 1]]]
        flat = np.concatenate(cols) if cols else np.zeros(0, np.uint32)
        uniq, first = np.unique(flat, return_index=True)
        out.append(uniq[np.argsort(first)])
    return out


@pytest.mark.parametrize("dups", [False, True])
def test_integer_products_equal_scipy(oracle, dups):
    a = random_csr(300, 200, 5, 1, integer=True, empty_every=7, dups=dups)
    b = random_csr(200, 250, 4, 2, integer=True, empty_every=5, dups=dups)
    ref = (scipy_of(a, 300, 200) @ scipy_of(b, 200, 250)).toarray()
    dense = oracle.spmultsp(a, b, 250).reshape(300, 250)
    assert np.array_equal(dense, ref)
    rp, ri, rv = oracle.spmultsp_sp(a, b, 250)
    got = sp.csr_matrix((rv, ri, rp), shape=(300, 250))
    assert np.array_equal(got.toarray(), ref)
    # every (row, col) reached by a product appears exactly once, zero sums
    # included (scipy's product would drop cancelled entries)
    order = scan_first_order(a, b)
    assert np.array_equal(np.diff(rp.astype(np.int64)), [len(o) for o in order])


def test_sparse_column_order_is_first_appearance(oracle):
    a = random_csr(400, 300, 6, 3, empty_every=9, dups=True)
    b = random_csr(300, 500, 5, 4, empty_every=11, dups=True)
    rp, ri, _ = oracle.spmultsp_sp(a, b, 500)
    order = scan_first_order(a, b)
    for i, o in enumerate(order):
        assert np.array_equal(ri[rp[i]:rp[i + 1]], o)


def test_sparse_values_match_dense_sums(oracle):
    a = random_csr(500, 400, 6, 5, empty_every=9, dups=True)
    b = random_csr(400, 300, 5, 6, dups=True)
    dense = oracle.spmultsp(a, b, 300).reshape(500, 300)
    rp, ri, rv = oracle.spmultsp_sp(a, b, 300)
    rows = np.repeat(np.arange(500), np.diff(rp.astype(np.int64)))
    d = dense[rows, ri]
    # equal to the bit, except a lone -0.0 product that the dense sum,
    # started at +0.0, turns into +0.0
    assert np.array_equal(d, rv)
    ref = (scipy_of(a, 500, 400) @ scipy_of(b, 400, 300)).toarray()
    assert np.max(np.abs(dense - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_identity_returns_the_other_operand(oracle):
    b = random_csr(250, 180, 5, 7, empty_every=4, dups=True)
    eye = (np.arange(251, dtype=np.uint32), np.arange(250, dtype=np.uint32),
           np.ones(250))
    rp, ri, rv = oracle.spmultsp_sp(eye, b, 180)
    # I B: row i is B's row i, duplicates merged at their first position
    order = scan_first_order(eye, b)
    for i in range(250):
        assert np.array_equal(ri[rp[i]:rp[i + 1]], order[i])
    assert np.array_equal(sp.csr_matrix((rv, ri, rp), shape=(250, 180)).toarray(),
                          scipy_of(b, 250, 180).toarray())
    a = random_csr(120, 250, 5, 8, empty_every=3)
    eye2 = (np.arange(251, dtype=np.uint32), np.arange(250, dtype=np.uint32),
            np.ones(250))
    rp, ri, rv = oracle.spmultsp_sp(a, eye2, 250)
    assert np.array_equal(rp, a[0]) and np.array_equal(ri, a[1])
    assert np.array_equal(rv, a[2])


def test_signed_zero_and_empty(oracle):
    # a single product -1 * 0.0 = -0.0: the sparse result keeps it, the
    # dense one (0.0 + -0.0) holds +0.0
    a = (np.array([0, 1, 1], np.uint32), np.array([0], np.uint32), np.array([-1.0]))
    b = (np.array([0, 1], np.uint32), np.array([1], np.uint32), np.array([0.0]))
    rp, ri, rv = oracle.spmultsp_sp(a, b, 2)
    assert list(rp) == [0, 1, 1] and list(ri) == [1]
    assert np.signbit(rv[0])
    dense = oracle.spmultsp(a, b, 2)
    assert not np.signbit(dense[1])
    # nothing to multiply
    z = (np.zeros(4, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    rp, ri, rv = oracle.spmultsp_sp(z, b, 2)
    assert list(rp) == [0, 0, 0, 0] and ri.size == 0


def test_sums_follow_scan_order(oracle):
    """Each element is the IEEE sum of its products in scan order; Python
    floats add in double precision, one product after another, so a sum in
    any other order (by column, reversed, pairwise) differs in the last
    bits for some element of this case."""
    a = random_csr(60, 40, 6, 9, dups=True)
    b = random_csr(40, 30, 6, 10, dups=True)
    arp, ari, arv = a
    brp, bri, brv = b
    sums = {}
    for i in range(60):
        for j in range(arp[i], arp[i + 1]):
            l = ari[j]
            for k in range(brp[l], brp[l + 1]):
                key = (i, int(bri[k]))
                p = float(arv[j]) * float(brv[k])
                sums[key] = sums[key] + p if key in sums else p
    dense = oracle.spmultsp(a, b, 30).reshape(60, 30)
    rp, ri, rv = oracle.spmultsp_sp(a, b, 30)
    got = {}
    for i in range(60):
        for q in range(rp[i], rp[i + 1]):
            got[(i, int(ri[q]))] = rv[q]
    assert got.keys() == sums.keys()
    for key, v in sums.items():
        assert got[key] == v and dense[key] == v
    # the check has teeth: some element's reversed-order sum differs
    rev = {}
    for i in range(60):
        prods = []
        for j in range(arp[i], arp[i + 1]):
            l = ari[j]
            for k in range(brp[l], brp[l + 1]):
                prods.append((int(bri[k]), float(arv[j]) * float(brv[k])))
        for c, p in reversed(prods):
            rev[(i, c)] = rev[(i, c)] + p if (i, c) in rev else p
    assert any(rev[k] != sums[k] for k in sums)

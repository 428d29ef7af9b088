"""GPU sparse x sparse products (src/blas.c:94-234, 609-791) against the
oracle, bit for bit: each C element sums its products in the reference's
scan order, and the sparse result lists a row's columns in the order of
their first product.  Values are compared as raw 64-bit patterns, so a
signed zero or a last-bit difference fails."""
import os
import subprocess

import numpy as np
import pytest

from test_oracle_spgemm import random_csr

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


CASES = [
    # (anrows, ancols, bncols, degA, degB, seed, dups, empty_every)
    (300, 200, 250, 5, 4, 1, False, 7),
    (2000, 1500, 1800, 8, 6, 2, True, 9),
    (30000, 30000, 30000, 8, 8, 3, True, 13),
    (500, 40000, 3000000, 6, 3, 4, True, 0),   # three radix passes on columns
    (1, 1, 1, 1, 1, 5, True, 0),
]


@pytest.mark.parametrize("stage", ["staged", "two_pass"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%dx%d" % c[:3])
def test_spgemm_bitwise_with_oracle(gb, oracle, case, stage, monkeypatch):
    # short rows staged in one pass (the default), or counted and then
    # written (GOOSEBERRY_SPGEMM_STAGE_MAX=0, the path past the staging limit)
    if stage == "two_pass":
        monkeypatch.setenv("GOOSEBERRY_SPGEMM_STAGE_MAX", "0")
    anr, anc, bnc, da, db, seed, dups, emp = case
    print("spgemm case", case, stage, flush=True)
    a = random_csr(anr, anc, da, seed, empty_every=emp, dups=dups)
    b = random_csr(anc, bnc, db, seed + 100, empty_every=emp, dups=dups)
    rp, ri, rv = gb.spmultsp_sp(anr, anc, bnc, a, b)
    orp, ori, orv = oracle.spmultsp_sp(a, b, bnc)
    assert np.array_equal(rp, orp) and np.array_equal(ri, ori)
    assert np.array_equal(bits(rv), bits(orv))
    if anr * bnc <= 1 << 26:
        dense = gb.spmultsp(anr, anc, bnc, a, b)
        assert np.array_equal(bits(dense), bits(oracle.spmultsp(a, b, bnc)))


def test_spgemm_signed_zero_empty_and_identity(gb, oracle):
    a = (np.array([0, 1, 1], np.uint32), np.array([0], np.uint32), np.array([-1.0]))
    b = (np.array([0, 1], np.uint32), np.array([1], np.uint32), np.array([0.0]))
    rp, ri, rv = gb.spmultsp_sp(2, 1, 2, a, b)
    assert list(rp) == [0, 1, 1] and list(ri) == [1] and np.signbit(rv[0])
    dense = gb.spmultsp(2, 1, 2, a, b)
    assert np.array_equal(bits(dense), bits(np.zeros(4)))
    z = (np.zeros(5, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    rp, ri, rv = gb.spmultsp_sp(4, 1, 2, z, b)
    assert list(rp) == [0] * 5 and ri.size == 0 and rv.size == 0
    assert np.array_equal(gb.spmultsp(4, 1, 2, z, b), np.zeros(8))
    # A I = A, entry for entry
    a = random_csr(700, 400, 6, 6, empty_every=5)
    eye = (np.arange(401, dtype=np.uint32), np.arange(400, dtype=np.uint32), np.ones(400))
    rp, ri, rv = gb.spmultsp_sp(700, 400, 400, a, eye)
    assert np.array_equal(rp, a[0]) and np.array_equal(ri, a[1])
    assert np.array_equal(bits(rv), bits(a[2]))


def test_spgemm_blocks(gb, oracle):
    a = random_csr(900, 700, 6, 7, dups=True)
    b = random_csr(700, 500, 5, 8, dups=True)
    full = oracle.spmultsp(a, b, 500).reshape(900, 500)
    # dense: only the block rows are written
    sentinel = np.full(900 * 500, 7.5)
    got = gb.spmultsp(900, 700, 500, a, b, blocks=[100, 250, 600],
                      crowval=sentinel.copy()).reshape(900, 500)
    assert np.array_equal(bits(got[100:600]), bits(full[100:600]))
    assert np.all(got[:100] == 7.5) and np.all(got[600:] == 7.5)
    # sparse: blocks that partition the rows give the whole product
    rp, ri, rv = gb.spmultsp_sp(900, 700, 500, a, b, blocks=[0, 300, 301, 900])
    orp, ori, orv = oracle.spmultsp_sp(a, b, 500)
    assert np.array_equal(rp, orp) and np.array_equal(ri, ori)
    assert np.array_equal(bits(rv), bits(orv))
    with pytest.raises(gb.GooseberryError) as e:
        gb.spmultsp_sp(900, 700, 500, a, b, blocks=[10, 900])
    assert e.value.name == "INVALIDINPUT"


def test_spgemm_invalid_inputs(gb):
    a = (np.array([0, 1], np.uint32), np.array([3], np.uint32), np.ones(1))
    b = (np.array([0, 1, 1], np.uint32), np.array([0], np.uint32), np.ones(1))
    with pytest.raises(gb.GooseberryError) as e:      # A column 3 >= ancols 2
        gb.spmultsp_sp(1, 2, 1, a, b)
    assert e.value.name == "INVALIDINPUT"
    a = (np.array([0, 1], np.uint32), np.array([1], np.uint32), np.ones(1))
    b = (np.array([0, 1, 2], np.uint32), np.array([0, 4], np.uint32), np.ones(2))
    with pytest.raises(gb.GooseberryError):           # B column 4 >= bncols 3
        gb.spmultsp(1, 2, 3, a, b)
    # 65537 x 65536 products exceed 32-bit positions: refused before any
    # product is formed
    m = 65537
    a = (np.array([0, m], np.uint32), np.zeros(m, np.uint32), np.ones(m))
    b = (np.array([0, 65536], np.uint32), np.arange(65536, dtype=np.uint32),
         np.ones(65536))
    with pytest.raises(gb.GooseberryError) as e:
        gb.spmultsp_sp(1, 1, 65536, a, b)
    assert e.value.name == "INVALIDINPUT"


def run(*args):
    from gooseberry_amd import CLI_PATH
    return subprocess.run([CLI_PATH] + [str(a) for a in args], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=300)


def test_cli_sparse_times_sparse(gb, oracle, tmp_path):
    a = random_csr(1200, 900, 6, 11, empty_every=10)
    b = random_csr(900, 1100, 5, 12, empty_every=8)
    # the text format infers the column count from the largest column used
    a[1][-1], b[1][-1] = 899, 1099
    fa, fb = os.path.join(str(tmp_path), "A.csr"), os.path.join(str(tmp_path), "B.csr")
    assert oracle.write_csr(fa, *a) == 1 and oracle.write_csr(fb, *b) == 1
    nra, nca, arp, ari, arv = oracle.read_sparse(fa)
    nrb, ncb, brp, bri, brv = oracle.read_sparse(fb)
    assert (nra, nca, nrb, ncb) == (1200, 900, 900, 1100)
    oa, ob = (arp, ari, arv), (brp, bri, brv)
    # sparse output file: blas_spmultsp_sp
    out, ref = os.path.join(str(tmp_path), "C.csr"), os.path.join(str(tmp_path), "R.csr")
    r = run("blas", "-x", "multiply", "-i", fa, "-i", fb, "-o", out, "-t")
    assert r.returncode == 0, r.stderr
    assert "Compute:" in r.stdout
    assert oracle.write_csr(ref, *oracle.spmultsp_sp(oa, ob, 1100)) == 1
    assert open(out).read() == open(ref).read()
    # dense output file: blas_spmultsp
    out, ref = os.path.join(str(tmp_path), "C.mat"), os.path.join(str(tmp_path), "R.mat")
    r = run("blas", "-x", "multiply", "-i", fa, "-i", fb, "-o", out)
    assert r.returncode == 0, r.stderr
    oracle.write_grid(ref, oracle.spmultsp(oa, ob, 1100), 1200, 1100)
    assert open(out).read() == open(ref).read()
    # row and column partitions: matrix_permute(A, rperm, cperm),
    # matrix_permute(B, cperm, NULL), C = A B, matrix_unpermute(C, rperm)
    rng = np.random.default_rng(13)
    rpf, cpf = os.path.join(str(tmp_path), "r.part"), os.path.join(str(tmp_path), "c.part")
    for p, lab in ((rpf, rng.integers(0, 5, 1200)), (cpf, rng.integers(0, 4, 900))):
        with open(p, "w") as f:
            f.write("\n".join(map(str, lab)) + "\n")
    out, ref = os.path.join(str(tmp_path), "P.csr"), os.path.join(str(tmp_path), "Q.csr")
    r = run("blas", "-x", "multiply", "-i", fa, "-i", fb, "-o", out, "-R", rpf, "-C", cpf)
    assert r.returncode == 0, r.stderr
    _, rperm, _ = oracle.read_partition(rpf, 1200)
    _, cperm, _ = oracle.read_partition(cpf, 900)
    pa = oracle.permute_sparse(arp, ari, arv, 900, rperm, cperm)
    pb = oracle.permute_sparse(brp, bri, brv, 1100, cperm, None)
    crp, cri, crv = oracle.spmultsp_sp(pa, pb, 1100)
    rren = np.empty(1200, np.uint32)
    rren[rperm] = np.arange(1200, dtype=np.uint32)
    crp, cri, crv = oracle.permute_sparse(crp, cri, crv, 1100, rren, None)
    assert oracle.write_csr(ref, crp, cri, crv) == 1
    assert open(out).read() == open(ref).read()
    # B B: 1100 columns against 900 rows do not chain and are refused
    r = run("blas", "-x", "multiply", "-i", fb, "-i", fb, "-o", out)
    assert r.returncode != 0 and "do not match" in r.stderr


def test_resident_spgemm(gb, oracle):
    a = random_csr(3000, 2500, 6, 21, empty_every=11, dups=True)
    b = random_csr(2600, 2000, 5, 22, dups=True)   # more rows than A has columns
    A = gb.Matrix(*a, ncols=2500)
    B = gb.Matrix(*b, ncols=2000)
    Cm = A.spmultsp(B)
    assert (Cm.nrows, Cm.ncols) == (3000, 2000)
    rp, ri, rv = Cm.download()
    bb = (b[0][:2501], b[1][:b[0][2500]], b[2][:b[0][2500]])
    orp, ori, orv = oracle.spmultsp_sp(a, bb, 2000)
    assert np.array_equal(rp, orp) and np.array_equal(ri, ori)
    assert np.array_equal(bits(rv), bits(orv))
    # the product is an ordinary resident matrix: its SpMV works
    x = np.random.default_rng(23).uniform(-1, 1, 2000)
    y = Cm.spmv(x)
    ref = oracle.spmult(3000, 2000, 1, orp, ori, orv, x)
    assert np.max(np.abs(y - ref)) <= 1e-14 * np.max(np.abs(ref))
    d = gb.DeviceArray(3000 * 2000)
    A.spmultsp_dense_device(B, d)
    assert np.array_equal(bits(d.numpy()), bits(oracle.spmultsp(a, bb, 2000)))
    S = gb.Matrix(*random_csr(100, 50, 3, 24), ncols=50)
    with pytest.raises(gb.GooseberryError):
        B.spmultsp(S)   # 2000 columns against 100 rows


def test_resident_product_plan_without_value_index(gb, oracle):
    # a product planned by default skips the value index (its build reads
    # C's values twice); a replan with the defaults adds it, and the SpMV's
    # bits stay those of the plan without it
    rp, ri, rv = gb.laplacian(2, 60)
    A = gb.Matrix(rp, ri, rv)
    C = A.spmultsp(A)
    assert C.info()["values8"] == 0
    x = np.random.default_rng(31).uniform(-1, 1, C.ncols)
    y = C.spmv(x)
    crp, cri, crv = C.download()
    assert np.array_equal(bits(y), bits(oracle.spmv(crp, cri, crv, x)))
    C.replan(None)
    assert C.info()["values8"] > 0   # the square of a stencil has few distinct values
    assert np.array_equal(bits(C.spmv(x)), bits(y))
    # a plan given to the product is taken as it is
    Cv = A.spmultsp(A, plan={"short_max": 64})
    assert Cv.info()["values8"] == C.info()["values8"]

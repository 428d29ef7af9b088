"""GPU Cuthill-McKee (gooseberry_amd/csrc/hip/gb_cm.hip) against the oracle's
restatement of src/permute.c:49-110,288-332: the same permutation, entry
for entry.  The graphs exercise both ways the GPU advances: long thin
graphs (one vertex at a time in one workgroup), wide ones (the parallel
step over a range of positions), a hub whose row is too long for the
one-workgroup kernel, isolated vertices and empty rows (placed in runs),
repeated entries, non-symmetric structure (reachability follows rows) and
diagonal blocks; then the CLI's -b flow."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def csr(A):
    A = sp.csr_matrix(A)
    return A.indptr.astype(np.uint32), A.indices.astype(np.uint32)


def grid(nx, ny, nz=1, seed=None):
    T = lambda m: sp.diags([-1, 2, -1], [-1, 0, 1], shape=(m, m))
    I = sp.eye
    if nz == 1:
        G = sp.kron(I(ny), T(nx)) + sp.kron(T(ny), I(nx))
    else:
        G = (sp.kron(I(nz), sp.kron(I(ny), T(nx))) + sp.kron(I(nz), sp.kron(T(ny), I(nx)))
             + sp.kron(T(nz), sp.kron(I(ny), I(nx))))
    G = G.tocsr()
    if seed is not None:
        q = np.random.default_rng(seed).permutation(G.shape[0])
        G = G[q][:, q].tocsr()
    return G


def random_graph(n, deg, seed, sym=True, comps=1, dup=False):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    bounds = np.linspace(0, n, comps + 1).astype(int)
    for c in range(comps):
        lo, hi = bounds[c], bounds[c + 1]
        m = (hi - lo) * deg
        rows.append(rng.integers(lo, hi, m))
        cols.append(rng.integers(lo, hi, m))
    r, c = np.concatenate(rows), np.concatenate(cols)
    if sym:
        r, c = np.concatenate([r, c]), np.concatenate([c, r])
    A = sp.coo_matrix((np.ones(r.size), (r, c)), shape=(n, n)).tocsr()
    if dup:
        # repeat every third row's entries (CSR with repeated columns)
        rp, ri = [0], []
        for i in range(n):
            e = list(A.indices[A.indptr[i]:A.indptr[i + 1]])
            if i % 3 == 0:
                e = e + e[: len(e) // 2]
            ri += e
            rp.append(len(ri))
        return np.array(rp, np.uint32), np.array(ri, np.uint32)
    return csr(A)


CASES = {
    "path": lambda: csr(sp.diags([-1, 2, -1], [-1, 0, 1], shape=(20000, 20000))),
    "grid2d_scrambled": lambda: csr(grid(300, 200, seed=1)),
    "grid3d": lambda: csr(grid(40, 40, 40)),
    "random_sym_4comps": lambda: random_graph(30000, 3, 2, comps=4),
    "random_directed": lambda: random_graph(20000, 2, 3, sym=False),
    "duplicates": lambda: random_graph(8000, 3, 4, dup=True),
    "sparse_many_components": lambda: random_graph(20000, 1, 5, comps=50),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_cm_matches_oracle(gb, oracle, name):
    print("cm case", name, flush=True)
    rp, ri = CASES[name]()
    got = gb.cuthillmckee(rp, ri)
    ref = oracle.cuthillmckee(rp, ri)
    assert np.array_equal(got, ref)


def test_hub_isolated_and_empty_rows(gb, oracle):
    # a hub joined to 3000 leaves (its row is too long for the one-workgroup
    # kernel), a chain hanging off one leaf, isolated and empty rows
    n = 5000
    rows, cols = [], []
    for k in range(1, 3001):
        rows += [0, k]
        cols += [k, 0]
    for k in range(3001, 3500):
        rows += [k, k + 1]
        cols += [k + 1, k]
    rows += [2, 3001]
    cols += [3001, 2]
    diag = np.arange(0, 4000)   # rows 4000.. have no diagonal; 4500.. are empty
    rows += list(diag)
    cols += list(diag)
    for k in range(4000, 4500, 2):
        rows += [k, k + 1]
        cols += [k + 1, k]
    A = sp.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n)).tocsr()
    rp, ri = csr(A)
    got = gb.cuthillmckee(rp, ri)
    assert np.array_equal(got, oracle.cuthillmckee(rp, ri))
    assert sorted(got.tolist()) == list(range(n))


def test_blocks_and_resident(gb, oracle):
    rp, ri = csr(grid(60, 50, seed=6))
    blocks = np.array([0, 700, 701, 1900, 3000], dtype=np.uint32)
    got = gb.cuthillmckee(rp, ri, blocks=blocks)
    assert np.array_equal(got, oracle.cuthillmckee(rp, ri, blocks=blocks))
    # positions outside the blocks keep the caller's values
    keep = np.arange(3000, dtype=np.uint32)[::-1].copy()
    part = gb.cuthillmckee(rp, ri, blocks=np.array([100, 900], np.uint32), perm=keep)
    assert np.array_equal(part[:100], keep[:100]) and np.array_equal(part[900:], keep[900:])
    assert sorted(part[100:900].tolist()) == list(range(100, 900))
    M = gb.Matrix(rp, ri, np.ones(ri.size))
    assert np.array_equal(M.cuthillmckee(), oracle.cuthillmckee(rp, ri))
    assert np.array_equal(M.cuthillmckee(blocks), got)
    with pytest.raises(gb.GooseberryError) as e:
        gb.cuthillmckee(rp, ri, ncols=3001)
    assert e.value.name == "INVALIDINPUT"


def run(*args):
    from gooseberry_amd import CLI_PATH
    return subprocess.run([CLI_PATH] + [str(a) for a in args], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=300)


def test_cli_bandwidth_reduction(gb, oracle, tmp_path):
    """gooseberry blas -b: A and x reordered by Cuthill-McKee (within the row
    partition's blocks when -R is given); the output is unpermuted by the
    row partition only, so without -R it stays in the new order, as in the
    reference (src/gooseberry_bin.c:1658-1675, 1805-1807)."""
    A = grid(40, 30, seed=7)
    A.data = np.random.default_rng(8).uniform(-1, 1, A.nnz)
    n = A.shape[0]
    a = os.path.join(str(tmp_path), "A.csr")
    v = os.path.join(str(tmp_path), "x.vec")
    rp, ri = csr(A)
    assert oracle.write_csr(a, rp, ri, A.data) == 1
    x = np.random.default_rng(9).uniform(-1, 1, n)
    oracle.write_grid(v, x, n, 1)
    nr, nc, orp, ori, orv = oracle.read_sparse(a)
    _, _, ox = oracle.read_grid(v)

    def expect(rperm=None, cperm=None, blocks=None):
        prp, pri, prv = orp, ori, orv
        px = ox
        if rperm is not None or cperm is not None:
            prp, pri, prv = oracle.permute_sparse(prp, pri, prv, nc, rperm, cperm)
            if cperm is not None:
                px = oracle.permute_dense(px, nc, 1, cperm)
        bperm = oracle.cuthillmckee(prp, pri, blocks=blocks)
        if rperm is not None and cperm is not None:
            rperm = bperm[rperm]
            cperm = bperm[cperm]
        prp, pri, prv = oracle.permute_sparse(prp, pri, prv, nc, bperm, bperm)
        px = oracle.permute_dense(px, nc, 1, bperm)
        y = oracle.spmult(nr, nc, 1, prp, pri, prv, px)
        if rperm is not None:
            y = oracle.permute_dense(y, nr, 1, rperm, rev=True)
        return np.array([float("%g" % t) for t in y])

    out = os.path.join(str(tmp_path), "y.vec")
    r = run("blas", "-x", "multiply", "-i", a, "-i", v, "-o", out, "-b")
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out)
    ref = expect()
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    # with row and column partitions
    rng = np.random.default_rng(10)
    rpf, cpf = os.path.join(str(tmp_path), "r.part"), os.path.join(str(tmp_path), "c.part")
    for p, lab in ((rpf, rng.integers(0, 4, n)), (cpf, rng.integers(0, 3, n))):
        with open(p, "w") as f:
            f.write("\n".join(map(str, lab)) + "\n")
    r = run("blas", "-x", "multiply", "-i", a, "-i", v, "-o", out, "-b", "-R", rpf, "-C", cpf)
    assert r.returncode == 0, r.stderr
    _, rperm, rdist = oracle.read_partition(rpf, nr)
    _, cperm, _ = oracle.read_partition(cpf, nc)
    ref = expect(rperm.copy(), cperm.copy(), blocks=rdist)
    got = np.loadtxt(out)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    # a non-square matrix is refused, as permute_cuthillmckee refuses it
    b = os.path.join(str(tmp_path), "B.csr")
    with open(b, "w") as f:
        f.write("1 1 3 2\n2 1\n")
    r = run("blas", "-x", "multiply", "-i", b, "-i", v, "-o", out, "-b")
    assert r.returncode != 0

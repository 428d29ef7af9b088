"""End-to-end `gooseberry blas` / `gooseberry cgd` on the GPU against the
oracle run through the reference's CLI flow (src/gooseberry_bin.c:1509-1857,
1860-2144): read -> (partition permutation) -> compute -> unpermute -> write.
"""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(*args):
    from gooseberry_amd import CLI_PATH
    return subprocess.run([CLI_PATH] + [str(a) for a in args], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=300)


def write_inputs(tmp_path, oracle, rp, ri, rv, vec):
    a = os.path.join(str(tmp_path), "A.csr")
    v = os.path.join(str(tmp_path), "x.vec")
    assert oracle.write_csr(a, rp, ri, rv) == 1
    with open(v, "w") as f:
        f.write("\n".join("%.17g" % t for t in vec) + "\n")
    return a, v


def test_blas_multiply_output_identical_to_oracle(gb, oracle, tmp_path):
    rp, ri, rv = gb.laplacian(2, 50)
    x = gb.uniform(0, 2500, stream=1)
    a, v = write_inputs(tmp_path, oracle, rp, ri, rv, x)
    out = os.path.join(str(tmp_path), "y.vec")
    r = run("blas", "-x", "multiply", "-i", a, "-i", v, "-o", out, "-r", "3", "-t")
    assert r.returncode == 0, r.stderr
    assert "Compute:" in r.stdout
    # oracle through the same text inputs
    nr, nc, orp, ori, orv = oracle.read_sparse(a)
    _, _, ox = oracle.read_grid(v)
    y = oracle.spmult(nr, nc, 1, orp, ori, orv, ox)
    ref = os.path.join(str(tmp_path), "ref.vec")
    oracle.write_grid(ref, y, nr, 1)
    assert open(out).read() == open(ref).read()


def test_blas_with_row_and_column_partitions(gb, oracle, tmp_path):
    from test_gpu_spmv import irregular
    n = 3000
    rp, ri, rv = irregular(n, 31, [(5, 900)], density=0.003)
    x = np.random.default_rng(32).uniform(-1, 1, n)
    a, v = write_inputs(tmp_path, oracle, rp, ri, rv, x)
    rng = np.random.default_rng(33)
    rl, cl = rng.integers(0, 7, n), rng.integers(0, 5, n)
    rpf, cpf = os.path.join(str(tmp_path), "r.part"), os.path.join(str(tmp_path), "c.part")
    for p, lab in ((rpf, rl), (cpf, cl)):
        with open(p, "w") as f:
            f.write("\n".join(map(str, lab)) + "\n")
    out = os.path.join(str(tmp_path), "y.vec")
    r = run("blas", "-x", "multiply", "-i", a, "-i", v, "-o", out, "-R", rpf, "-C", cpf)
    assert r.returncode == 0, r.stderr
    assert "Read 7-way row partition" in r.stdout
    # reference flow: matrix_permute(A, rperm, cperm); matrix_permute(x, cperm);
    # y = A x; matrix_unpermute(y, rperm)
    nr, nc, orp, ori, orv = oracle.read_sparse(a)
    _, _, ox = oracle.read_grid(v)
    _, rperm, _ = oracle.read_partition(rpf, nr)
    _, cperm, _ = oracle.read_partition(cpf, nc)
    prp, pri, prv = oracle.permute_sparse(orp, ori, orv, nc, rperm, cperm)
    px = oracle.permute_dense(ox, nc, 1, cperm)
    y = oracle.spmult(nr, nc, 1, prp, pri, prv, px)
    y = oracle.permute_dense(y, nr, 1, rperm, rev=True)
    got = np.loadtxt(out)
    ref = np.array([float("%g" % t) for t in y])
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_cgd_command(gb, oracle, tmp_path):
    rp, ri, rv = gb.laplacian(2, 40)
    b = gb.uniform(0, 1600)
    a, v = write_inputs(tmp_path, oracle, rp, ri, rv, b)
    out = os.path.join(str(tmp_path), "x.csr")  # written as GRID regardless
    nr, nc, orp, ori, orv = oracle.read_sparse(a)
    _, _, ob = oracle.read_grid(v)
    err = 1e-8 * np.linalg.norm(ob)
    r = run("cgd", "-i", a, "-i", v, "-o", out, "-e", "%.17g" % err, "-I", 2000, "-t")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "matrix = 1600x1600 vector = 1600x1"
    iters = [l for l in lines if l.startswith("ITER:")]
    assert all(re.match(r"^ITER: \d+ RMSE: \d+\.\d{6}$", l) for l in iters)
    assert [int(l.split()[1]) for l in iters] == list(range(len(iters)))
    xr, it, rmse, _ = oracle.cgd(orp, ori, orv, ob, err, niter=2000)
    assert abs(len(iters) - it) <= 2
    got = np.loadtxt(out)
    assert got.shape == (1600,)
    assert np.max(np.abs(got - xr)) <= 1e-4 * np.max(np.abs(xr))
    assert "Compute:" in r.stdout


def test_cgd_rejects_nonsquare(gb, oracle, tmp_path):
    a = os.path.join(str(tmp_path), "A.csr")
    with open(a, "w") as f:
        f.write("1 1.0 3 2.0\n2 1.0\n")
    v = os.path.join(str(tmp_path), "b.vec")
    with open(v, "w") as f:
        f.write("1\n2\n")
    r = run("cgd", "-i", a, "-i", v, "-o", os.path.join(str(tmp_path), "x.vec"))
    assert r.returncode == 2 and "not square" in r.stderr


@pytest.mark.parametrize("name,fmt", [("g.graph", 5), ("m.clu", 6), ("m.svm", 4),
                                      ("m.ij", 3)])
def test_blas_reads_the_other_sparse_formats(gb, oracle, tmp_path, name, fmt):
    """`gooseberry blas` picks the reader by extension (.graph/.metis/.chaco,
    .clu, .svm/.libfm, .ij/.point; src/gooseberry_bin.c:453-508); the output
    equals the oracle's SpMV of the oracle's reading of the same file."""
    from test_formats import generated_files
    text = generated_files(7)[(name, fmt)]
    a = os.path.join(str(tmp_path), name)
    with open(a, "w") as f:
        f.write(text)
    ref_read = {5: oracle.read_graph, 6: oracle.read_clu, 4: oracle.read_svm,
                3: oracle.read_point}[fmt]
    nr, nc, orp, ori, orv = ref_read(a)
    x = gb.uniform(3, nc, stream=1)
    v = os.path.join(str(tmp_path), "x.vec")
    with open(v, "w") as f:
        f.write("\n".join("%.17g" % t for t in x) + "\n")
    out = os.path.join(str(tmp_path), "y.vec")
    r = run("blas", "-x", "multiply", "-i", a, "-i", v, "-o", out)
    assert r.returncode == 0, r.stderr
    _, _, ox = oracle.read_grid(v)
    y = oracle.spmult(nr, nc, 1, orp, ori, orv, ox)
    ref = os.path.join(str(tmp_path), "ref.vec")
    oracle.write_grid(ref, y, nr, 1)
    assert open(out).read() == open(ref).read()

"""CPU tests of the product's host side: the C-ABI library loads and exports
every function include/gooseberry.h declares; CSR construction and vector
I/O are bit-exact with the oracle; the input generators; the CLI's exit
codes on paths that never reach the GPU; and that compute entry points fail
loudly (no CPU fallback) when no GPU is present.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAVE_GPU = os.path.exists("/dev/kfd")


def header_functions():
    text = open(os.path.join(ROOT, "include", "gooseberry.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gooseberry_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_entry_points():
    names = header_functions()
    for ref in ("gooseberry_spmult", "gooseberry_scale", "gooseberry_read_sparse_matrix",
                "gooseberry_read_dense_matrix", "gooseberry_write_dense_matrix",
                "gooseberry_write_sparse_matrix", "gooseberry_read_labels",
                "gooseberry_read_partition"):
        assert ref in names


def test_library_exports_every_declared_symbol(gb):
    out = subprocess.run(["nm", "-D", "--defined-only", gb.LIB_PATH],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing
    assert sorted(gb.EXPORTED) == sorted(header_functions())


def test_header_compiles_as_c_and_cxx(tmp_path):
    src = os.path.join(str(tmp_path), "t.c")
    with open(src, "w") as f:
        f.write('#include "gooseberry.h"\nint main(void){return GOOSEBERRY_SUCCESS==1?0:1;}\n')
    for cc in (["gcc", "-std=c99"], ["g++", "-x", "c++"]):
        subprocess.run(cc + ["-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                             "-o", os.path.join(str(tmp_path), "t")], check=True)


# -------------------------------------------------- CSR construction parity

def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w", newline="") as f:
        f.write(text)
    return p


CSR_CASES = [
    "# comment\n1 1.5 3 2.0\n\n2 -1 2 4 1 0.5\n% another\n5 7\n",
    "0 1.0 2 2.0 9\n1 3.0\n",
    "1 1e-3 2 -2.5e+10\r\n\r\n3 0.1\r\n",
    "",
    "\n\n\n",
    "4 1.0",
    "1 .5 2 abc 3 4\n' quote comment\n/ slash\n\"x\n2 2\n",
    "  3 1.0   1 2.0\n\t2 3.0\n",
]


@pytest.mark.parametrize("case", range(len(CSR_CASES)))
def test_read_csr_bit_exact_with_oracle(gb, oracle, tmp_path, case):
    p = _write(tmp_path, "m%d.csr" % case, CSR_CASES[case])
    ref = oracle.read_sparse(p)
    got = gb.read_sparse(p, gb.FORMAT_CSR)
    assert got[0] == ref[0] and got[1] == ref[1]
    for a, b in zip(got[2:], ref[2:]):
        assert a.dtype == b.dtype and np.array_equal(a, b)


def test_read_csr_random_files_bit_exact(gb, oracle, tmp_path):
    rng = np.random.default_rng(0)
    for t in range(5):
        n = int(rng.integers(1, 200))
        lines = []
        for i in range(n):
            k = int(rng.integers(0, 12))
            cols = rng.integers(1, 300, k)
            vals = rng.standard_normal(k) * 10.0 ** rng.integers(-5, 5, k)
            lines.append(" ".join("%d %.17g" % (c, v) for c, v in zip(cols, vals)))
            if rng.random() < 0.1:
                lines.append("# c")
        p = _write(tmp_path, "r%d.csr" % t, "\n".join(lines) + "\n")
        ref = oracle.read_sparse(p)
        got = gb.read_sparse(p)
        assert got[:2] == ref[:2]
        for a, b in zip(got[2:], ref[2:]):
            assert np.array_equal(a, b)


MTX_CASES = [
    "%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 3\n1 1 4.0\n2 1 -1.0\n3 3 2.0\n",
    "%%MatrixMarket matrix coordinate real general\n\n2 2 3\n1 2 1.5\n2 1 -3\n1 1 7\n",
    "2 3 2\n1 3\n2 1\n",
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n2 2\n",
]


@pytest.mark.parametrize("case", range(len(MTX_CASES)))
def test_read_mtx_bit_exact_with_oracle(gb, oracle, tmp_path, case):
    p = _write(tmp_path, "m%d.mtx" % case, MTX_CASES[case])
    ref = oracle.read_sparse(p, mtx=True)
    if isinstance(ref, int):
        with pytest.raises(gb.GooseberryError) as e:
            gb.read_sparse(p, gb.FORMAT_MATRIXMARKET)
        assert e.value.code == ref
        return
    got = gb.read_sparse(p, gb.FORMAT_MATRIXMARKET)
    assert got[:2] == ref[:2]
    for a, b in zip(got[2:], ref[2:]):
        assert np.array_equal(a, b)


GRID_CASES = ["1\n2.5\n-3e-7\n", "1 2 3\n4 5 6\n", "# c\n1 2\n3 4\n", "1 2 3\n4 5\n", ""]


@pytest.mark.parametrize("case", range(len(GRID_CASES)))
def test_read_grid_bit_exact_with_oracle(gb, oracle, tmp_path, case):
    p = _write(tmp_path, "g%d.vec" % case, GRID_CASES[case])
    ref = oracle.read_grid(p)
    if isinstance(ref, int):
        with pytest.raises(gb.GooseberryError) as e:
            gb.read_dense(p)
        assert e.value.code == ref
        return
    got = gb.read_dense(p)
    assert got[:2] == ref[:2] and np.array_equal(got[2], ref[2])


def test_writers_match_oracle_text(gb, oracle, tmp_path):
    rng = np.random.default_rng(1)
    v = rng.standard_normal(50) * 10.0 ** rng.integers(-8, 8, 50)
    a, b = os.path.join(str(tmp_path), "a.vec"), os.path.join(str(tmp_path), "b.vec")
    gb.write_dense(a, v, 25, 2)
    oracle.write_grid(b, v, 25, 2)
    assert open(a).read() == open(b).read()
    A = sp.random(30, 20, density=0.2, random_state=2, format="csr")
    gb.write_csr(a + ".csr", A.indptr, A.indices, A.data, 20)
    oracle.write_csr(b + ".csr", A.indptr, A.indices, A.data)
    assert open(a + ".csr").read() == open(b + ".csr").read()


def test_read_labels_matches_oracle(gb, oracle, tmp_path):
    p = _write(tmp_path, "l.part", "% c\n3\n1\n4\n1\n\n5\n")
    assert np.array_equal(gb.read_labels(p), oracle.read_labels(p))
    assert np.array_equal(gb.read_labels(p, 2), oracle.read_labels(p, 2))


def test_missing_file_codes(gb, tmp_path):
    with pytest.raises(gb.GooseberryError) as e:
        gb.read_sparse(os.path.join(str(tmp_path), "nope.csr"))
    assert e.value.name == "FILENOTFOUND"


# ---------------------------------------------------------------- generators

def test_laplacian_generators_match_kron_construction(gb):
    from test_oracle_pinning import lap2d, lap3d
    for dims, n, ref in ((2, 23, lap2d(23, 23)), (3, 9, lap3d(9))):
        rp, ri, rv = gb.laplacian(dims, n)
        A = sp.csr_matrix((rv, ri, rp), shape=ref.shape)
        assert np.array_equal(rp, ref.indptr) and np.array_equal(ri, ref.indices)
        # values 4/-1 (2D), 6/-1 (3D)
        assert np.array_equal(A.diagonal(), np.full(ref.shape[0], 2.0 * dims))
        assert (A - A.T).nnz == 0


def test_laplacian_sizes_of_bench_configs(gb):
    rp, ri, rv = gb.laplacian(2, 256)
    assert rp.size - 1 == 65536 and rp[-1] == 326656


def test_uniform_generator_matches_golden(gb):
    with open(os.path.join(ROOT, "tests", "golden", "rng_first16.json")) as f:
        gold = json.load(f)
    for key, rec in gold.items():
        seed, stream = map(int, key.split(":"))
        u = gb.uniform(seed, 16, lo=0.0, hi=1.0, stream=stream)
        assert np.array_equal(u, np.array(rec["u"])), key
        w = gb.uniform(seed, 16, stream=stream)
        assert np.array_equal(w, -1.0 + 2.0 * np.array(rec["u"])), key


# -------------------------------------------------------------------- CLI

def run_cli(*args):
    from gooseberry_amd import CLI_PATH
    return subprocess.run([CLI_PATH] + list(args), stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=120)


def test_cli_exit_codes_without_compute(tmp_path):
    assert run_cli().returncode == 1
    assert run_cli("nosuch").returncode == 1
    r = run_cli("help")
    assert r.returncode == 0 and "cgd" in r.stdout
    assert run_cli("sgd").returncode == 2           # not provided by this build
    assert run_cli("pagerank").returncode == 0      # usage only (nargs < 2)
    r = run_cli("pagerank", "-i", os.path.join(str(tmp_path), "no.csr"), "-t")
    assert r.returncode == 2 and "output vector file" in r.stderr
    assert run_cli("blas").returncode == 0          # usage only (nargs < 2)
    r = run_cli("cgd", "-i", os.path.join(str(tmp_path), "no.csr"), "-i",
                os.path.join(str(tmp_path), "no.vec"), "-o", os.path.join(str(tmp_path), "x.vec"))
    assert r.returncode == 2
    r = run_cli("cgd", "-i", "a.csr")
    assert r.returncode == 0 and "USAGE" in r.stderr
    r = run_cli("blas", "-x", "divide", "-i", "a.csr")
    assert r.returncode == 2


@pytest.mark.skipif(HAVE_GPU, reason="checks the no-GPU failure path")
def test_compute_fails_loudly_without_gpu(gb):
    rp, ri, rv = gb.laplacian(2, 4)
    with pytest.raises(gb.GooseberryError) as e:
        gb.spmult(16, 16, 1, rp, ri, rv, np.ones(16))
    assert e.value.name == "DEVICE"


def test_binary_csr_round_trip_and_checks(gb, oracle, tmp_path):
    """The binary CSR format: header, arrays as stored, exact size; the
    host reader gives back the written arrays and refuses a bad file."""
    from test_oracle_spgemm import random_csr
    rp, ri, rv = random_csr(3000, 2000, 6, 41, empty_every=7, dups=True)
    rv[::97] = -0.0
    path = os.path.join(str(tmp_path), "m.bcsr")
    gb.write_csr(path, rp, ri, rv, 2000, fmt=gb.FORMAT_BCSR)
    data = open(path, "rb").read()
    assert data[:8] == b"GBCSR\x00\x01\x00"
    assert len(data) == 24 + 4 * 3001 + 12 * int(rp[-1])
    nr, nc, grp, gri, grv = gb.read_sparse(path, gb.FORMAT_BCSR)
    assert (nr, nc) == (3000, 2000)
    assert np.array_equal(grp, rp) and np.array_equal(gri, ri)
    assert np.array_equal(grv.view(np.uint64), rv.view(np.uint64))
    for name, blob in (("short", data[:-1]), ("magic", b"X" + data[1:]),
                       ("col", data[:24 + 4 * 3001] + b"\xff\xff\xff\xff" +
                        data[24 + 4 * 3001 + 4:])):
        bad = os.path.join(str(tmp_path), name + ".bcsr")
        with open(bad, "wb") as f:
            f.write(blob)
        with pytest.raises(gb.GooseberryError) as e:
            gb.read_sparse(bad, gb.FORMAT_BCSR)
        assert e.value.name == "INVALIDINPUT", name


def test_device_reader_fails_loudly_without_gpu(gb, tmp_path):
    if HAVE_GPU:
        pytest.skip("a GPU is present")
    path = os.path.join(str(tmp_path), "a.csr")
    with open(path, "w") as f:
        f.write("1 2\n")
    with pytest.raises(gb.GooseberryError) as e:
        gb.read_sparse_device(path)
    assert e.value.name == "DEVICE"


def test_fault_hook_arguments(gb):
    # the hook only sets a countdown: no device is touched
    L = gb.lib()
    assert L.gooseberry_inject_fault(7, 0) == gb.ERROR_CODES["INVALIDINPUT"]
    assert L.gooseberry_inject_fault(-1, 0) == gb.ERROR_CODES["INVALIDINPUT"]
    for kind in (gb.FAULT_ALLOC, gb.FAULT_LAUNCH):
        assert L.gooseberry_inject_fault(kind, 3) == gb.SUCCESS
        assert L.gooseberry_inject_fault(kind, -1) == gb.SUCCESS


# ------------------------------------------------------ 64-bit row pointers

PROG64 = r'''
#include <stdio.h>
#include <stdlib.h>
#include "gooseberry.h"
/* built with GOOSEBERRY_64BIT_INDICES: the reference's names take 64-bit
 * row pointers */
int main(int argc, char **argv) {
  if (sizeof(ind_t) != 8) return 10;
  dim_t nr, nc;
  ind_t *rp;
  dim_t *ri;
  real_t *rv;
  int r = gooseberry_read_sparse_matrix(GOOSEBERRY_FORMAT_CSR, argv[1], &nr, &nc,
                                        &rp, &ri, &rv);
  if (r != GOOSEBERRY_SUCCESS) return 11;
  r = gooseberry_write_sparse_matrix(GOOSEBERRY_FORMAT_CSR, argv[2], nr, nc, rp, ri, rv);
  if (r != GOOSEBERRY_SUCCESS) return 12;
  printf("%u %u %llu\n", (unsigned)nr, (unsigned)nc, (unsigned long long)rp[nr]);
  if (argc > 3) {
    real_t *x = (real_t *)malloc((nc ? nc : 1) * sizeof(real_t));
    real_t *y = (real_t *)malloc((nr ? nr : 1) * sizeof(real_t));
    for (dim_t i = 0; i < nc; ++i) x[i] = 1.0 + (real_t)(i % 7) / 8.0;
    r = gooseberry_spmult(nr, nc, 1, rp, ri, rv, x, y, NULL, 0);
    if (r != GOOSEBERRY_SUCCESS) return 13;
    for (dim_t i = 0; i < nr; ++i) printf("%.17g\n", y[i]);
    free(x);
    free(y);
  }
  free(rp);
  free(ri);
  free(rv);
  return 0;
}
'''


def build_prog64(gb, tmp_path):
    """The program above, compiled as C99 and as C++ against the header with
    GOOSEBERRY_64BIT_INDICES and linked against the library."""
    src = os.path.join(str(tmp_path), "p64.c")
    with open(src, "w") as f:
        f.write(PROG64)
    libdir = os.path.dirname(gb.LIB_PATH)
    exe = os.path.join(str(tmp_path), "p64")
    for cc, out in ((["gcc", "-std=c99"], exe), (["g++", "-x", "c++"], exe + "xx")):
        subprocess.run(cc + ["-Wall", "-Werror", "-DGOOSEBERRY_64BIT_INDICES", "-I",
                             os.path.join(ROOT, "include"), src, "-L", libdir,
                             "-lgooseberry", "-Wl,-rpath," + libdir, "-o", out], check=True)
    return exe


def test_header_64bit_indices_maps_onto_i64(gb, tmp_path):
    """With GOOSEBERRY_64BIT_INDICES ind_t is 8 bytes and the reference's
    reader and writer names resolve to their _i64 forms: a .csr file goes
    through them byte for byte (host work, no GPU needed)."""
    exe = build_prog64(gb, tmp_path)
    from test_oracle_spgemm import random_csr
    rp, ri, rv = random_csr(300, 200, 5, 17, empty_every=9, dups=True)
    a = os.path.join(str(tmp_path), "a.csr")
    b = os.path.join(str(tmp_path), "b.csr")
    gb.write_csr(a, rp, ri, rv, 200)
    r = subprocess.run([exe, a, b], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.returncode
    nr, nc, nnz = map(int, r.stdout.split())
    assert (nr, nnz) == (300, int(rp[-1]))
    assert open(a, "rb").read() == open(b, "rb").read()
    # the library exports both widths; the 32-bit names stay 32-bit
    names = header_functions()
    for n in ("gooseberry_spmult", "gooseberry_cgd", "gooseberry_matrix_create"):
        assert n in names and n + "_i64" in names

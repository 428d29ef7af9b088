"""Calls from several host threads at once (include/gooseberry.h: one
library lock serialises them).  The reference's gooseberry_spmult keeps no
state between calls and may be called from any thread
(src/gooseberry.c:53-67, src/blas.c:552-606); here each thread's results
must be bitwise those of the same calls made from one thread.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _work(gb, k):
    rp, ri, rv = gb.laplacian(2 + (k % 2), 40 + 7 * k if k % 2 == 0 else 12 + k)
    n = rp.size - 1
    x = gb.uniform(100 + k, n)
    out = [gb.spmult(n, n, 1, rp, ri, rv, x) for _ in range(3)]
    M = gb.Matrix(rp, ri, rv)
    xs, info = M.cgd(x, 1e-9, gb.TOL_RELATIVE)
    out.append(xs)
    out.append(int(info.niter))
    M.free()
    return out


def test_concurrent_callers_match_serial(gb):
    nthreads = 6
    want = [_work(gb, k) for k in range(nthreads)]
    got = [None] * nthreads
    errors = []

    def run(k):
        try:
            for _ in range(2):
                got[k] = _work(gb, k)
        except Exception as e:  # reported below
            errors.append(e)

    ts = [threading.Thread(target=run, args=(k,)) for k in range(nthreads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(120)
    assert not errors, errors
    for k in range(nthreads):
        for a, b in zip(got[k], want[k]):
            if isinstance(a, int):
                assert a == b
            else:
                assert np.array_equal(a, b)

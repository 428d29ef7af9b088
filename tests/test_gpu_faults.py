"""Failure handling through the C ABI, every failure forced from the host
(no kernel faults or waits forever):

* injected device-allocation and launch failures map to error codes
  (NOMEMORY, DEVICE), the call that met them returns no object, and the
  library keeps working (the error model of include/gooseberry.h:85-94 in
  the reference, where failures are codes and the CLI exits 2);
* the server cancels queued and running requests, times out overdue ones
  and reports a request whose input its admission refuses, while the other
  requests finish with the same bits as alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def disarm(gb):
    yield
    gb.inject_fault(gb.FAULT_ALLOC, -1)
    gb.inject_fault(gb.FAULT_LAUNCH, -1)


def test_injected_allocation_failure(gb, oracle):
    rp, ri, rv = gb.laplacian(2, 80)
    x = gb.uniform(1, rp.size - 1)
    want = oracle.spmv(rp, ri, rv, x)
    # the first, then a later allocation of the matrix's construction
    for after in (0, 3, 6):
        gb.inject_fault(gb.FAULT_ALLOC, after)
        with pytest.raises(gb.GooseberryError) as e:
            gb.Matrix(rp, ri, rv)
        assert e.value.name == "NOMEMORY"
    gb.inject_fault(gb.FAULT_ALLOC, -1)
    M = gb.Matrix(rp, ri, rv)
    assert np.array_equal(M.spmv(x), want)


def test_injected_launch_failure_in_cg(gb):
    rp, ri, rv = gb.laplacian(3, 24)
    M = gb.Matrix(rp, ri, rv)
    b = gb.uniform(2, M.nrows)
    db = gb.DeviceArray.from_host(b)
    dx = gb.DeviceArray(M.nrows)
    ref = M.cgd_device(db, dx, 1e-10)
    x_ref = dx.numpy()
    # the solve's set-up launches (the iterations replay captured graphs,
    # which make no launch checks)
    for after in (0, 1):
        gb.inject_fault(gb.FAULT_LAUNCH, after)
        dy = gb.DeviceArray(M.nrows)
        rv_ = gb.lib().gooseberry_matrix_cgd(M.h, db.ptr, dy.ptr, 1e-10, gb.TOL_RELATIVE,
                                             0, None, None)
        assert rv_ == gb.ERROR_CODES["DEVICE"], after
        gb.inject_fault(gb.FAULT_LAUNCH, -1)
        # the interrupted solve starts again and gives the same answer
        info = M.cgd_device(db, dy, 1e-10)
        assert info.niter == ref.niter and np.array_equal(dy.numpy(), x_ref)


def test_injected_failure_in_batch_and_server(gb):
    # an allocation failure while a server admits a generated request: the
    # call returns the error, the request stays queued and is served later
    S = gb.Server(max_active=4, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
    for i in range(3):
        S.submit_sdd(i, 5000 + 1000 * i, 70 + i, 3, 27, 1e-10)
    gb.inject_fault(gb.FAULT_ALLOC, 2)
    with pytest.raises(gb.GooseberryError) as e:
        S.progress(2)
    assert e.value.name == "NOMEMORY"
    done = []
    for _ in range(500):
        done += S.progress(8)
        if len(done) == 3:
            break
    assert sorted(r["id"] for r in done) == [0, 1, 2]
    assert all(r["status"] == 1 and r["rel_residual"] <= 1e-10 for r in done)
    x1 = S.take_result(1, 6000)
    S1 = gb.Server(max_active=1, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
    S1.submit_sdd(1, 6000, 71, 3, 27, 1e-10)
    while not S1.progress(8):
        pass
    assert np.array_equal(S1.take_result(1, 6000), x1)


def _drain(S, want, steps=4):
    done = []
    for _ in range(2000):
        done += S.progress(steps)
        if len(done) >= want:
            break
    return {r["id"]: r for r in done}


def test_server_cancel_queued_and_running(gb):
    S = gb.Server(max_active=2, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
    sizes = {0: 200000, 1: 150000, 2: 8000, 3: 9000}
    for rid, n in sizes.items():
        S.submit_sdd(rid, n, 300 + rid, 3, 27, 1e-12)
    S.progress(1)  # 0 and 1 running, 2 and 3 queued
    a, q, _ = S.status()
    assert (a, q) == (2, 2)
    S.cancel(2)   # queued
    S.cancel(0)   # running
    with pytest.raises(gb.GooseberryError):
        S.cancel(99)
    a, q, _ = S.status()
    assert (a, q) == (1, 1)
    res = _drain(S, 4)
    assert res[0]["status"] == gb.REQUEST_CANCELLED
    assert res[2]["status"] == gb.REQUEST_CANCELLED
    assert res[1]["status"] == 1 and res[3]["status"] == 1
    # the requests that ran beside a cancelled one give the same bits alone
    x3 = S.take_result(3, 9000)
    S1 = gb.Server(max_active=1, precision=gb.PRECISION_MIXED_JACOBI, keep_results=True)
    S1.submit_sdd(3, 9000, 303, 3, 27, 1e-12)
    while not S1.progress(8):
        pass
    assert np.array_equal(S1.take_result(3, 9000), x3)
    a, q, m = S.status()
    assert a == 0 and q == 0


def test_server_timeout(gb):
    S = gb.Server(max_active=1, precision=gb.PRECISION_MIXED_JACOBI)
    S.set_timeout(1e-6)  # everything is overdue by the next call
    for rid in range(3):
        S.submit_sdd(rid, 6000, 400 + rid, 3, 27, 1e-10)
    res = _drain(S, 3)
    assert sorted(res) == [0, 1, 2]
    assert all(r["status"] == gb.REQUEST_TIMEDOUT for r in res.values())
    S.set_timeout(0)
    S.submit_sdd(7, 6000, 407, 3, 27, 1e-10)
    res = _drain(S, 1)
    assert res[7]["status"] == 1


def test_server_rejects_input_and_goes_on(gb):
    # no positive diagonal: Jacobi scaling refuses the matrix at admission
    rp = np.array([0, 2, 3], np.uint32)
    ri = np.array([0, 1, 0], np.uint32)
    rv = np.array([1.0, 2.0, 2.0])
    bad = gb.Matrix(rp, ri, rv)
    db, dx = gb.DeviceArray.from_host(np.ones(2)), gb.DeviceArray(2)
    S = gb.Server(max_active=2, precision=gb.PRECISION_MIXED_JACOBI)
    S.submit(0, bad, db, dx, 1e-10)
    S.submit_sdd(1, 7000, 501, 3, 27, 1e-10)
    res = _drain(S, 2)
    assert res[0]["status"] == gb.REQUEST_REJECTED
    assert res[1]["status"] == 1

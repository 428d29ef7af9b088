"""Host <-> device copies through the pinned staging buffers, and the
structure check of host CSR arrays after their upload.

Copies from pageable host memory to the device of 16 MiB and more, and
back to pageable memory of 4 MiB and more, go through two pinned buffers
filled and drained by several host threads; smaller copies, and copies
from or to memory the HIP runtime has pinned, go straight to the runtime.
Every byte must arrive, whatever the size's alignment, the number of
staging rounds and the direction.  The host entry points that take a CSR
(`gooseberry_matrix_create`, `gooseberry_spmult`, src/gooseberry.c:53-67)
check its structure on the device once it is uploaded, with the same
refusals as a host scan.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bytes(n, seed):
    return np.frombuffer(np.random.default_rng(seed).bytes(n), dtype=np.uint8).copy()


def _htod(gb, d_ptr, h_ptr, n):
    gb.check(gb.lib().gooseberry_memcpy_htod(d_ptr, h_ptr, n), "gooseberry_memcpy_htod")


def _dtoh(gb, h_ptr, d_ptr, n):
    gb.check(gb.lib().gooseberry_memcpy_dtoh(h_ptr, d_ptr, n), "gooseberry_memcpy_dtoh")


@pytest.mark.parametrize("nbytes", [
    (4 << 20) - 1,                 # the runtime's copies both ways
    4 << 20,                       # staged back to the host, in 2 MiB chunks
    (5 << 20) + 3,                 # parts that are not whole pages
    16 << 20,                      # staged both ways, four 4 MiB chunks
    (32 << 20) + 4097,             # 8 MiB chunks and a short last one
    (3 << 25) + 777,               # 16 MiB chunks, both buffers several times
])
def test_staged_round_trip(gb, nbytes):
    a = _bytes(nbytes, nbytes)
    d = gb.DeviceArray(nbytes, dtype=np.uint8)
    d.upload(a)
    # a device-side copy, so that the copy back reads what the device holds
    b = gb.DeviceArray(nbytes, dtype=np.uint8)
    b.copy_from(d)
    assert np.array_equal(b.numpy(), a)
    # from an odd device offset into a host slice at an odd offset
    off = 4099
    out = np.zeros(nbytes + 13, dtype=np.uint8)
    _dtoh(gb, out.ctypes.data + 13, d.ptr + off, nbytes - off)
    assert np.array_equal(out[13:13 + nbytes - off], a[off:])
    assert not out[:13].any() and not out[13 + nbytes - off:].any()
    # and up from a host slice at an odd offset
    _htod(gb, b.ptr, a.ctypes.data + off, nbytes - off)
    back = np.empty(nbytes - off, dtype=np.uint8)
    _dtoh(gb, back.ctypes.data, b.ptr, nbytes - off)
    assert np.array_equal(back, a[off:])
    d.free()
    b.free()


def test_staged_copy_of_float64_values(gb):
    v = np.random.default_rng(3).uniform(-1, 1, 5 << 20)  # 40 MiB
    d = gb.DeviceArray.from_host(v)
    assert np.array_equal(d.numpy(), v)
    d.free()


def test_staged_copy_from_a_caller_bound_to_one_cpu(gb):
    # the staging threads of a caller bound to one CPU (as an OpenMP
    # program's initial thread is under OMP_PROC_BIND) run on the CPUs the
    # process had at load time; the bytes are the same
    import os
    mask = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(mask)})
    try:
        a = _bytes((48 << 20) + 1, 21)
        d = gb.DeviceArray(a.size, dtype=np.uint8)
        d.upload(a)
        assert np.array_equal(d.numpy(), a)
        d.free()
    finally:
        os.sched_setaffinity(0, mask)


def test_pinned_host_memory_goes_straight(gb):
    # memory pinned by the HIP runtime the library itself runs on
    hip = C.CDLL("libamdhip64.so")
    hip.hipHostMalloc.restype = C.c_int
    hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    hip.hipHostFree.restype = C.c_int
    hip.hipHostFree.argtypes = [C.c_void_p]
    n = (9 << 20) + 5
    p = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(p), n, 0) == 0
    try:
        h = np.ctypeslib.as_array((C.c_uint8 * n).from_address(p.value))
        a = _bytes(n, 11)
        h[:] = a
        d = gb.DeviceArray(n, dtype=np.uint8)
        _htod(gb, d.ptr, p.value, n)
        assert np.array_equal(d.numpy(), a)
        h[:] = 0
        _dtoh(gb, p.value, d.ptr, n)
        assert np.array_equal(h, a)
        d.free()
    finally:
        hip.hipHostFree(p)


def test_host_spmult_staged_matches_oracle(gb, oracle):
    # a 2D 1000^2 Laplacian: its column indices (20 MB) and values (40 MB)
    # take several staging rounds; the products are the reference's bits
    rp, ri, rv = gb.laplacian(2, 1000)
    n = rp.size - 1
    x = gb.uniform(5, n)
    y = gb.spmult(n, n, 1, rp, ri, rv, x)
    assert np.array_equal(y, oracle.spmv(rp, ri, rv, x))
    # again, on staging buffers that are already there
    assert np.array_equal(gb.spmult(n, n, 1, rp, ri, rv, x), y)


def _random_csr(n, seed):
    rng = np.random.default_rng(seed)
    deg = rng.integers(1, 9, n)
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    ri = rng.integers(0, n, int(rp[-1])).astype(np.uint32)
    rv = rng.uniform(-1, 1, int(rp[-1]))
    return rp, ri, rv


@pytest.mark.parametrize("case", ["column", "decreasing"])
def test_host_csr_structure_refused_after_upload(gb, case):
    n = 1 << 20  # 4.5 M entries: the column indices go through the staging buffers
    rp, ri, rv = _random_csr(n, 8)
    if case == "column":
        ri[int(rp[-1]) // 2] = n            # a column index equal to ncols
    else:
        k = n // 2                          # rp[k] > rp[k + 1]
        rp[k] = rp[k + 1] + 1
    with pytest.raises(gb.GooseberryError) as e:
        gb.Matrix(rp, ri, rv, ncols=n)
    assert e.value.name == "INVALIDINPUT"
    with pytest.raises(gb.GooseberryError) as e:
        gb.spmult(n, n, 1, rp, ri, rv, np.ones(n))
    assert e.value.name == "INVALIDINPUT"
    # the library goes on working after the refusals
    rp, ri, rv = _random_csr(1000, 9)
    x = np.random.default_rng(1).uniform(-1, 1, 1000)
    M = gb.Matrix(rp, ri, rv)
    assert np.array_equal(M.spmv(x), gb.spmult(1000, 1000, 1, rp, ri, rv, x))


def test_host_csr_first_row_pointer(gb):
    # a resident matrix's row pointers start at 0; gooseberry_spmult reads
    # them from their first value (the reference's own offsets)
    rp, ri, rv = _random_csr(1000, 10)
    with pytest.raises(gb.GooseberryError) as e:
        gb.Matrix(np.concatenate([[1], rp[1:]]).astype(np.uint32), ri, rv)
    assert e.value.name == "INVALIDINPUT"
    h = C.c_void_p()
    three = np.array([3], np.uint32)  # no rows, row pointers not from 0
    rc = gb.lib().gooseberry_matrix_create(0, 5, three.ctypes.data_as(C.c_void_p),
                                           ri.ctypes.data_as(C.c_void_p),
                                           rv.ctypes.data_as(C.c_void_p), None, C.byref(h))
    assert rc == gb.ERROR_CODES["INVALIDINPUT"]

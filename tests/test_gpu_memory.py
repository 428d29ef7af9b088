"""Device memory: blocks of 4 MiB and up that the library frees are kept and
handed back to the next allocation of the same rounded size (2 MiB
granularity); the kept bytes count as free in gooseberry_device_memory; a
reused block serves a new computation correctly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIB = 1 << 20


def test_large_block_reused_for_same_rounded_size(gb):
    a = gb.DeviceArray(3 * MIB, np.uint8)          # 3 MiB: below the cache's floor
    b = gb.DeviceArray(9 * MIB + 5, np.uint8)      # rounds to 10 MiB
    pb = b.ptr
    b.free()
    c = gb.DeviceArray(10 * MIB - 7, np.uint8)     # the same rounded size
    assert c.ptr == pb
    d = gb.DeviceArray(12 * MIB, np.uint8)         # another size: a new block
    assert d.ptr != pb
    for x in (a, c, d):
        x.free()


def test_kept_bytes_count_as_free(gb):
    # blocks taken from the pool may not move the driver's free count, but
    # blocks the library keeps after a free must be reported as free
    arrs = [gb.DeviceArray(64 * MIB + 3 * i, np.uint8) for i in range(4)]
    gb.synchronize()
    f1, total = gb.device_memory()
    for x in arrs:
        x.free()
    gb.synchronize()
    f2, _ = gb.device_memory()
    assert f2 >= f1 + 4 * 64 * MIB - 8 * MIB and f2 <= total


def test_reused_blocks_give_the_same_results(gb, oracle):
    rp, ri, rv = gb.laplacian(3, 60)
    n = rp.size - 1
    x = gb.uniform(4, n)
    want = oracle.spmv(rp, ri, rv, x)
    for _ in range(3):
        M = gb.Matrix(rp, ri, rv)           # arrays of 1.5-12 MiB, reused each time
        junk = gb.DeviceArray.from_host(np.full(n, np.nan))
        y = M.spmv(x)
        assert np.max(np.abs(y - want)) <= 1e-14 * np.max(np.abs(want))
        del M
        junk.free()

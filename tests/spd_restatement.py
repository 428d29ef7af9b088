"""Numpy restatement of the random SPD generator of
gooseberry_amd/csrc/hip/gb_gen.hip (test infrastructure only).

Row i draws a degree d_i and emits d_i - 1 directed edges, edge e to a
column uniform over [0, n) \\ {i} with value -u_e, u_e uniform in (0, 1];
the pattern is symmetrised (out-edges in e order, then in-edges in e order,
then the diagonal, 1.01 x the row's absolute off-diagonal sum added left to
right).  For the uniform degree law the CSR arrays equal the device's bit
for bit; the Zipf draws use pow(), whose last bit may differ between libm
and the device at an acceptance boundary.
"""
import numpy as np

_K = [np.uint64(v) for v in (0xbf58476d1ce4e5b9, 0x94d049bb133111eb, 0x9e3779b97f4a7c15,
                             0x632be59bd9b4e019, 0xd1342543de82ef95, 0x7f4a7c159e3779b9)]
H_DEGREE, H_ZIPF_U, H_ZIPF_V, H_COLUMN, H_VALUE = 0xd0, 0xd1, 0xd2, 0xe0, 0xe1


def _u64(v):
    return np.asarray(v, dtype=np.uint64)


def mix64(z):
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _K[0]
        z = (z ^ (z >> np.uint64(27))) * _K[1]
        return z ^ (z >> np.uint64(31))


def hash3(a, b, c):
    a, b, c = _u64(a), _u64(b), _u64(c)
    with np.errstate(over="ignore"):
        return mix64(a * _K[2] ^ mix64(b + _K[3]) ^ mix64(c * _K[4] + _K[5]))


def u01(h):
    """Uniform in (0, 1]: ((h >> 11) + 1) 2^-53."""
    return ((h >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53


def uniform_degrees(n, seed, dmin, dmax):
    i = np.arange(n, dtype=np.uint64)
    span = np.uint64(dmax - dmin + 1)
    return (dmin + hash3(seed, H_DEGREE, i) % span).astype(np.int64)


def zipf_degrees(n, seed, s, dmin, dmax, tries=1024):
    """Zipf(s) by Devroye's rejection method, clipped to [dmin, dmax]."""
    am1 = s - 1.0
    b = 2.0 ** am1
    i = np.arange(n, dtype=np.uint64)
    x = np.ones(n)
    todo = np.arange(n)
    for t in range(tries):
        if not len(todo):
            break
        r = np.uint64(t << 8)
        U = u01(hash3(seed, np.uint64(H_ZIPF_U) + r, i[todo]))
        V = u01(hash3(seed, np.uint64(H_ZIPF_V) + r, i[todo]))
        X = np.floor(U ** (-1.0 / am1))
        T = (1.0 + 1.0 / X) ** am1
        acc = V * X * (T - 1.0) / (b - 1.0) <= T / b
        x[todo] = X
        todo = todo[~acc]
    d = np.where(~(x <= dmax), dmax, np.where(x < dmin, dmin, x))
    return d.astype(np.int64)


def symmetrise(n, seed, d):
    """CSR (rowptr u32, rowind u32, rowval f64) for degrees d."""
    o = d - 1
    m = int(o.sum())
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(o, out=off[1:])
    e = np.arange(m, dtype=np.uint64)
    src = np.repeat(np.arange(n, dtype=np.int64), o)
    dst = (hash3(seed, H_COLUMN, e) % np.uint64(n - 1)).astype(np.int64)
    dst += dst >= src
    val = -u01(hash3(seed, H_VALUE, e))
    indeg = np.bincount(dst, minlength=n)
    inptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(indeg, out=inptr[1:])
    ine = np.argsort(dst, kind="stable")
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(o + indeg + 1, out=rowptr[1:])
    nnz = int(rowptr[-1])
    rowind = np.empty(nnz, dtype=np.int64)
    rowval = np.empty(nnz, dtype=np.float64)
    ei = np.arange(m, dtype=np.int64)
    pos = rowptr[src] + (ei - off[src])
    rowind[pos] = dst
    rowval[pos] = val
    p = np.arange(m, dtype=np.int64)
    j = dst[ine]
    slot = rowptr[j] + o[j] + (p - inptr[j])
    rowind[slot] = src[ine]
    rowval[slot] = val[ine]
    last = rowptr[1:] - 1
    # sequential row sums: column k of the rows, added for every row at once
    lens = last - rowptr[:-1]
    acc = np.zeros(n)
    for k in range(int(lens.max()) if n else 0):
        live = lens > k
        acc[live] = acc[live] + np.abs(rowval[rowptr[:-1][live] + k])
    rowind[last] = np.arange(n)
    rowval[last] = np.where(lens > 0, 1.01 * acc, 1.0)
    return rowptr.astype(np.uint32), rowind.astype(np.uint32), rowval


def generate_uniform(n, seed, dmin, dmax):
    return symmetrise(n, seed, uniform_degrees(n, seed, dmin, dmax))

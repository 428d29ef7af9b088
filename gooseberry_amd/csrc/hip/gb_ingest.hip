// gb_ingest.hip -- CSR construction straight into device memory: the .csr
// text reader of src/io.c:369-498 run on the GPU, and the binary CSR
// format's loader.
//
// .csr text: one line per row, "col val col val ...".  A line is the text
// up to '\n' (a final line without one counts); a line whose first byte is
// one of # % ' " / is a comment (src/io.c:63-69, 406-410) and is skipped;
// any other line, blank ones included, is a row.  A row's entries come
// from the reference's loop: strtoull(.., 10) for a column, then strtod
// for a value, repeated while strtod converts something (src/io.c:415-446),
// so rows keep file order (no sort, no merge of repeated columns).
// Columns are shifted by mincol = min(1, smallest column) and ncols =
// maxcol - mincol + 1 (src/io.c:433-435, 464, 478).
//
// On the device:
//   1. newlines counted per 4 KiB chunk, scanned, and their positions
//      written: the start of every line
//   2. comment flags per line, scanned: the row of every line
//   3. count pass, one thread per line: entries per row (the conversions'
//      end positions only) and the smallest / largest column
//   4. row pointers = scan of the counts
//   5. fill pass: columns and values written at the row pointers; a value
//      whose rounding needs more than 19 digits' worth of precision is
//      listed and settled afterwards, one thread each, by exact
//      big-integer comparison.
// Every conversion is the C library's (gb_strtod.h), so the arrays are
// bitwise those of the host reader for the same text.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#define GB_HD __host__ __device__
#include "gb_strtod.h"
#include "gb_common.h"
#include "gb_stage.h"

#define GB_IN_CHUNK 4096                       // bytes per workgroup, line split
#define GB_IN_PER_THREAD (GB_IN_CHUNK / GB_BLOCK)  // 16
#define GB_IN_FIX0 4096                        // initial capacity of the hard list

struct gb_in_fix {
  uint64_t p;    // where strtod starts
  uint64_t end;  // the end of its line
  uint32_t out;  // entry index
  uint32_t pad;
};

__device__ __forceinline__ bool gb_in_is_comment(unsigned c) {
  return c == '#' || c == '%' || c == '\'' || c == '"' || c == '/';
}

// newlines per chunk
__global__ void __launch_bounds__(GB_BLOCK)
gb_in_nl_count(const char *__restrict__ txt, uint64_t len, uint32_t *__restrict__ cnt) {
  __shared__ uint32_t lds[GB_WAVES_PER_BLOCK];
  const uint64_t base = (uint64_t)blockIdx.x * GB_IN_CHUNK;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < GB_IN_PER_THREAD; ++i) {
    const uint64_t p = base + (uint64_t)i * GB_BLOCK + threadIdx.x;
    if (p < len && txt[p] == '\n') ++c;
  }
  c = gb_block_sum(c, lds);
  if (threadIdx.x == 0) cnt[blockIdx.x] = c;
}

// ls[0] = 0, ls[1 + k] = the position after the k-th newline; off = the
// exclusive scan of the chunk counts
__global__ void __launch_bounds__(GB_BLOCK)
gb_in_nl_pos(const char *__restrict__ txt, uint64_t len, const uint32_t *__restrict__ off,
             uint64_t *__restrict__ ls) {
  __shared__ uint32_t lds[GB_WAVES_PER_BLOCK];
  const uint64_t base = (uint64_t)blockIdx.x * GB_IN_CHUNK + (uint64_t)threadIdx.x *
                        GB_IN_PER_THREAD;
  uint32_t c = 0;
  for (int i = 0; i < GB_IN_PER_THREAD; ++i) {
    const uint64_t p = base + i;
    if (p < len && txt[p] == '\n') ++c;
  }
  uint32_t tot;
  uint32_t k = off[blockIdx.x] + gb_block_scan_excl(c, lds, &tot);
  for (int i = 0; i < GB_IN_PER_THREAD && c; ++i) {
    const uint64_t p = base + i;
    if (p < len && txt[p] == '\n') {
      ls[1 + k] = p + 1;
      ++k;
      --c;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ls[0] = 0;
}

// line k: [start, end) without its newline and trailing '\r'
__device__ __forceinline__ void gb_in_line(const char *txt, uint64_t len, const uint64_t *ls,
                                           uint32_t nl, uint32_t k, uint64_t *start,
                                           uint64_t *end) {
  const uint64_t a = ls[k];
  uint64_t b = k < nl ? ls[k + 1] - 1 : len;
  while (b > a && txt[b - 1] == '\r') --b;
  *start = a;
  *end = b;
}

// rowof[k] = 1 for a row line, 0 for a comment (scanned in place later)
__global__ void __launch_bounds__(GB_BLOCK)
gb_in_rowflag(const char *__restrict__ txt, uint64_t len, const uint64_t *__restrict__ ls,
              uint32_t nl, uint32_t nlines, uint32_t *__restrict__ rowof) {
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < nlines;
       k += gridDim.x * GB_BLOCK) {
    uint64_t a, b;
    gb_in_line(txt, len, ls, nl, k, &a, &b);
    rowof[k] = (b > a && gb_in_is_comment((unsigned char)txt[a])) ? 0u : 1u;
  }
}

// the reference's entry loop over one line; FILL false counts entries and
// tracks the column range, FILL true writes them from out0
template <bool FILL>
__device__ __forceinline__ uint32_t gb_in_parse_line(const char *txt, uint64_t a, uint64_t b,
                                                     uint32_t mincol, uint32_t out0,
                                                     uint32_t *rowind, double *rowval,
                                                     uint32_t *mn, uint32_t *mx,
                                                     gb_in_fix *fix, uint32_t fixcap,
                                                     uint32_t *nfix) {
  uint32_t nent = 0;
  uint64_t sp = a, c = 0;
  double v = 0.0;
  bool hard = false;
  sp = gbt_strtoull(txt, b, sp, 10, &c);
  uint64_t e = gbt_strtod(txt, b, sp, &v, &hard, FILL);
  while (e != sp) {
    const uint32_t col = (uint32_t)c;
    if (FILL) {
      const uint32_t o = out0 + nent;
      rowind[o] = col - mincol;
      rowval[o] = v;
      if (hard) {
        const uint32_t f = atomicAdd(nfix, 1u);
        if (f < fixcap) {
          fix[f].p = sp;
          fix[f].end = b;
          fix[f].out = o;
        }
      }
    } else {
      *mn = col < *mn ? col : *mn;
      *mx = col > *mx ? col : *mx;
    }
    ++nent;
    sp = e;
    sp = gbt_strtoull(txt, b, sp, 10, &c);
    e = gbt_strtod(txt, b, sp, &v, &hard, FILL);
  }
  return nent;
}

// count pass: cnt[row] = entries; minmax[0] = min column, minmax[1] = max
__global__ void __launch_bounds__(GB_BLOCK)
gb_in_count(const char *__restrict__ txt, uint64_t len, const uint64_t *__restrict__ ls,
            uint32_t nl, uint32_t nlines, const uint32_t *__restrict__ rowof,
            uint32_t *__restrict__ cnt, uint32_t *__restrict__ minmax) {
  __shared__ uint32_t smn, smx;
  if (threadIdx.x == 0) {
    smn = 0xffffffffu;
    smx = 0;
  }
  __syncthreads();
  uint32_t mn = 0xffffffffu, mx = 0;
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < nlines;
       k += gridDim.x * GB_BLOCK) {
    const uint32_t r = rowof[k];
    if (rowof[k + 1] == r) continue;  // comment
    uint64_t a, b;
    gb_in_line(txt, len, ls, nl, k, &a, &b);
    cnt[r] = gb_in_parse_line<false>(txt, a, b, 0, 0, nullptr, nullptr, &mn, &mx, nullptr, 0,
                                     nullptr);
  }
  if (mn != 0xffffffffu) atomicMin(&smn, mn);
  if (mx) atomicMax(&smx, mx);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (smn != 0xffffffffu) atomicMin(&minmax[0], smn);
    if (smx) atomicMax(&minmax[1], smx);
  }
}

// fill pass
__global__ void __launch_bounds__(GB_BLOCK)
gb_in_fill(const char *__restrict__ txt, uint64_t len, const uint64_t *__restrict__ ls,
           uint32_t nl, uint32_t nlines, const uint32_t *__restrict__ rowof,
           const uint32_t *__restrict__ rowptr, uint32_t mincol, uint32_t *__restrict__ rowind,
           double *__restrict__ rowval, gb_in_fix *__restrict__ fix, uint32_t fixcap,
           uint32_t *__restrict__ nfix) {
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < nlines;
       k += gridDim.x * GB_BLOCK) {
    const uint32_t r = rowof[k];
    if (rowof[k + 1] == r) continue;
    uint64_t a, b;
    gb_in_line(txt, len, ls, nl, k, &a, &b);
    gb_in_parse_line<true>(txt, a, b, mincol, rowptr[r], rowind, rowval, nullptr, nullptr,
                           fix, fixcap, nfix);
  }
}

// the listed values, exactly (one thread each; the big integers live in
// this kernel's private memory only)
__global__ void __launch_bounds__(64)
gb_in_fixup(const char *__restrict__ txt, const gb_in_fix *__restrict__ fix, uint32_t nfix,
            double *__restrict__ rowval) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= nfix) return;
  gbt_big L, R;
  const gb_in_fix f = fix[i];
  rowval[f.out] = gbt_strtod_exact(txt, f.end, f.p, &L, &R);
}

__global__ void gb_in_init_minmax(uint32_t *minmax) {
  minmax[0] = 0xffffffffu;
  minmax[1] = 0;
}

__global__ void __launch_bounds__(GB_BLOCK)
gb_in_sum64(const uint32_t *__restrict__ a, uint64_t n, unsigned long long *__restrict__ total) {
  __shared__ unsigned long long lds[GB_WAVES_PER_BLOCK];
  unsigned long long v = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * GB_BLOCK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * GB_BLOCK)
    v += a[i];
  v = gb_block_sum(v, lds);
  if (threadIdx.x == 0) atomicAdd(total, v);
}

// ---- MatrixMarket body (src/io.c:1060-1145): one coordinate per line

// flag[k] = 1 for an entry line (not blank after '\r' stripping, not a
// comment)
__global__ void __launch_bounds__(GB_BLOCK)
gb_mx_flag(const char *__restrict__ txt, uint64_t len, const uint64_t *__restrict__ ls,
           uint32_t nl, uint32_t nlines, uint32_t *__restrict__ flag) {
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < nlines;
       k += gridDim.x * GB_BLOCK) {
    uint64_t a, b;
    gb_in_line(txt, len, ls, nl, k, &a, &b);
    flag[k] = (b == a || gb_in_is_comment((unsigned char)txt[a])) ? 0u : 1u;
  }
}

// entry e (= rank of its line): row i, column j (1-based in the file),
// weight w when one converts (a weight the first 19 digits do not settle is
// listed for gb_in_fixup).  bad: 1 when a coordinate does not convert or is
// out of range; nwt counts the entries with a weight
__global__ void __launch_bounds__(GB_BLOCK)
gb_mx_parse(const char *__restrict__ txt, uint64_t len, const uint64_t *__restrict__ ls,
            uint32_t nl, uint32_t nlines, const uint32_t *__restrict__ ent,
            uint32_t nrows, uint32_t ncols, int sym, uint32_t *__restrict__ ei,
            uint32_t *__restrict__ ej, double *__restrict__ ew, uint32_t *__restrict__ bad,
            uint32_t *__restrict__ nwt, gb_in_fix *__restrict__ fix, uint32_t fixcap,
            uint32_t *__restrict__ nfix) {
  __shared__ uint32_t swt;
  if (threadIdx.x == 0) swt = 0;
  __syncthreads();
  uint32_t b_ = 0, wt = 0;
  for (uint32_t k = blockIdx.x * GB_BLOCK + threadIdx.x; k < nlines;
       k += gridDim.x * GB_BLOCK) {
    const uint32_t e = ent[k];
    if (ent[k + 1] == e) continue;
    uint64_t a, b;
    gb_in_line(txt, len, ls, nl, k, &a, &b);
    uint64_t v = 0;
    uint64_t q = gbt_strtoull(txt, b, a, 10, &v);
    if (q == a) { b_ = 1; continue; }
    const uint32_t i = (uint32_t)v - 1u;
    uint64_t q2 = gbt_strtoull(txt, b, q, 10, &v);
    if (q2 == q) { b_ = 1; continue; }
    const uint32_t j = (uint32_t)v - 1u;
    double w = 1.0;
    bool hard = false;
    const uint64_t q3 = gbt_strtod(txt, b, q2, &w, &hard, true);
    if (q3 != q2) {
      ++wt;
      if (hard) {  // settled by gb_in_fixup
        const uint32_t f = atomicAdd(nfix, 1u);
        if (f < fixcap) {
          fix[f].p = q2;
          fix[f].end = b;
          fix[f].out = e;
        }
      }
    } else {
      w = 1.0;
    }
    if (i >= nrows || j >= ncols || (sym && j >= nrows)) b_ = 1;
    ei[e] = i;
    ej[e] = j;
    ew[e] = w;
  }
  if (b_) atomicOr(bad, 1u);
  if (wt) atomicAdd(&swt, wt);
  __syncthreads();
  if (threadIdx.x == 0 && swt) atomicAdd(nwt, swt);
}

// record keys in file order, a symmetric file's mirrored entry first:
// record 2e = (j, i), 2e + 1 = (i, j); otherwise record e = (i, j)
__global__ void __launch_bounds__(GB_BLOCK)
gb_mx_keys(uint32_t nent, int sym, const uint32_t *__restrict__ ei,
           const uint32_t *__restrict__ ej, uint32_t *__restrict__ key) {
  for (uint32_t e = blockIdx.x * GB_BLOCK + threadIdx.x; e < nent; e += gridDim.x * GB_BLOCK) {
    if (sym) {
      key[2 * (uint64_t)e] = ej[e];
      key[2 * (uint64_t)e + 1] = ei[e];
    } else {
      key[e] = ei[e];
    }
  }
}

// CSR entries from the records in row order (ord = stable sort by row)
__global__ void __launch_bounds__(GB_BLOCK)
gb_mx_gather(uint32_t nnz, int sym, int wgts, const uint32_t *__restrict__ ord,
             const uint32_t *__restrict__ ei, const uint32_t *__restrict__ ej,
             const double *__restrict__ ew, uint32_t *__restrict__ rowind,
             double *__restrict__ rowval) {
  for (uint32_t p = blockIdx.x * GB_BLOCK + threadIdx.x; p < nnz; p += gridDim.x * GB_BLOCK) {
    const uint32_t r = ord[p];
    const uint32_t e = sym ? r >> 1 : r;
    const bool mirror = sym && !(r & 1u);
    rowind[p] = mirror ? ei[e] : ej[e];
    rowval[p] = wgts ? ew[e] : 1.0;
  }
}

static uint32_t gb_in_grid(uint64_t n) {
  uint64_t g = (n + GB_BLOCK - 1) / GB_BLOCK;
  if (g > 16384) g = 16384;
  return g ? (uint32_t)g : 1u;
}

static int gb_in_errno(int e) {
  switch (e) {
    case ENOENT: case ENOTDIR: return GOOSEBERRY_ERROR_FILENOTFOUND;
    case EACCES: case EPERM: return GOOSEBERRY_ERROR_PERMISSIONDENIED;
    case EINVAL: case ENAMETOOLONG: return GOOSEBERRY_ERROR_INVALIDINPUT;
    default: return GOOSEBERRY_ERROR_UNKNOWN;
  }
}

// A file range read into a staging buffer: pread by the staging threads,
// each over its own part of the range.
struct gb_in_file_src {
  int fd;
  uint64_t base;  // file offset of the range's first byte
};

static int gb_in_fill_file(void *ctx, uint64_t off, uint64_t len, char *buf) {
  const gb_in_file_src *f = (const gb_in_file_src *)ctx;
  uint64_t done = 0;
  while (done < len) {
    const ssize_t r = pread(f->fd, buf + done, (size_t)(len - done),
                            (off_t)(f->base + off + done));
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return GOOSEBERRY_ERROR_READ;  // an error, or the file ended early
    done += (uint64_t)r;
  }
  return GOOSEBERRY_SUCCESS;
}

// bytes [offset, offset + bytes) of an open file into device memory: the
// staging threads read into one pinned buffer while the other one's copy to
// the device runs
static int gb_in_file_to_device(FILE *f, uint64_t offset, uint64_t bytes, char *d_dst) {
  gb_in_file_src src = {fileno(f), offset};
  return gb_stage_htod(d_dst, bytes, gb_in_fill_file, &src);
}

extern "C" {

int gbd_load_file(const char *file, char **d_txt, uint64_t *r_len) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (!file || !d_txt || !r_len) return GOOSEBERRY_ERROR_INVALIDINPUT;
  FILE *f = fopen(file, "rb");
  if (!f) return gb_in_errno(errno);
  struct stat st;
  if (fstat(fileno(f), &st) != 0) {
    fclose(f);
    return GOOSEBERRY_ERROR_READ;
  }
  const uint64_t len = (uint64_t)st.st_size;
  char *d = nullptr;
  hipError_t he = gb_dev_alloc((void **)&d, len + 16);
  if (he != hipSuccess) {
    fclose(f);
    return gb_hip_fail(he, "gb_dev_alloc", __FILE__, __LINE__);
  }
  rv = gb_in_file_to_device(f, 0, len, d);
  fclose(f);
  if (rv != GOOSEBERRY_SUCCESS) {
    (void)gb_dev_free(d);
    return rv;
  }
  *d_txt = d;
  *r_len = len;
  return GOOSEBERRY_SUCCESS;
}

int gbd_load_file_range(const char *file, uint64_t offset, uint64_t bytes, void *d_dst) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  FILE *f = fopen(file, "rb");
  if (!f) return gb_in_errno(errno);
  rv = gb_in_file_to_device(f, offset, bytes, (char *)d_dst);
  fclose(f);
  return rv;
}

int gbd_parse_csr(const char *d_txt, uint64_t len, uint32_t *r_nrows, uint32_t *r_ncols,
                  uint32_t *r_nnz, uint32_t **d_rowptr, uint32_t **d_rowind,
                  double **d_rowval) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  hipStream_t s = gb_stream();
  hipError_t he = hipSuccess;
  const uint64_t nchunks = (len + GB_IN_CHUNK - 1) / GB_IN_CHUNK;
  if (nchunks >= 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t *ccnt = nullptr, *rowof = nullptr, *cnt = nullptr, *minmax = nullptr,
           *nfix = nullptr, *rowptr = nullptr, *rowind = nullptr;
  uint64_t *ls = nullptr;
  double *rowval = nullptr;
  gb_in_fix *fix = nullptr;
  unsigned long long *tot = nullptr;
  uint32_t nl = 0, nlines = 0, nrows = 0, nnz = 0, hmm[2] = {0xffffffffu, 0}, hnfix = 0;
  uint32_t fixcap = GB_IN_FIX0;
  uint64_t last = 0;
  unsigned long long nnz64 = 0;
#define GB_IN_TRY(call)                                         \
  do {                                                          \
    he = (call);                                                \
    if (he != hipSuccess) {                                     \
      rv = gb_hip_fail(he, #call, __FILE__, __LINE__);          \
      goto done;                                                \
    }                                                           \
  } while (0)
#define GB_IN_RV(call)                                          \
  do {                                                          \
    rv = (call);                                                \
    if (rv != GOOSEBERRY_SUCCESS) goto done;                    \
  } while (0)

  // 1. line starts
  GB_IN_TRY(gb_dev_alloc((void **)&ccnt, (nchunks + 1) * 4));
  if (nchunks) {
    gb_in_nl_count<<<(uint32_t)nchunks, GB_BLOCK, 0, s>>>(d_txt, len, ccnt);
    GB_IN_TRY(hipGetLastError());
  }
  GB_IN_RV(gbd_scan_u32(ccnt, ccnt, nchunks, ccnt + nchunks));
  GB_IN_TRY(hipMemcpyAsync(&nl, ccnt + nchunks, 4, hipMemcpyDeviceToHost, s));
  GB_IN_TRY(hipStreamSynchronize(s));
  if (nl == 0xffffffffu) {
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto done;
  }
  GB_IN_TRY(gb_dev_alloc((void **)&ls, ((uint64_t)nl + 2) * 8));
  if (nchunks) {
    gb_in_nl_pos<<<(uint32_t)nchunks, GB_BLOCK, 0, s>>>(d_txt, len, ccnt, ls);
    GB_IN_TRY(hipGetLastError());
    GB_IN_TRY(hipMemcpyAsync(&last, ls + nl, 8, hipMemcpyDeviceToHost, s));
    GB_IN_TRY(hipStreamSynchronize(s));
  }
  // a last line without a newline counts; text ending in one has no
  // line after it
  nlines = nl + (len > last ? 1u : 0u);
  // 2. rows
  GB_IN_TRY(gb_dev_alloc((void **)&rowof, ((uint64_t)nlines + 1) * 4));
  if (nlines) {
    gb_in_rowflag<<<gb_in_grid(nlines), GB_BLOCK, 0, s>>>(d_txt, len, ls, nl, nlines, rowof);
    GB_IN_TRY(hipGetLastError());
  }
  GB_IN_RV(gbd_scan_u32(rowof, rowof, nlines, rowof + nlines));
  GB_IN_TRY(hipMemcpyAsync(&nrows, rowof + nlines, 4, hipMemcpyDeviceToHost, s));
  GB_IN_TRY(hipStreamSynchronize(s));
  // 3. count pass
  GB_IN_TRY(gb_dev_alloc((void **)&cnt, ((uint64_t)nrows + 1) * 4));
  GB_IN_TRY(gb_dev_alloc((void **)&minmax, 8));
  GB_IN_TRY(gb_dev_alloc((void **)&tot, 8));
  gb_in_init_minmax<<<1, 1, 0, s>>>(minmax);
  GB_IN_TRY(hipGetLastError());
  GB_IN_RV(gbd_zero(tot, 8));
  if (nlines) {
    gb_in_count<<<gb_in_grid(nlines), GB_BLOCK, 0, s>>>(d_txt, len, ls, nl, nlines, rowof, cnt,
                                                       minmax);
    GB_IN_TRY(hipGetLastError());
  }
  if (nrows) {
    gb_in_sum64<<<gb_in_grid(nrows) > 1024u ? 1024u : gb_in_grid(nrows), GB_BLOCK, 0, s>>>(
        cnt, nrows, tot);
    GB_IN_TRY(hipGetLastError());
  }
  GB_IN_TRY(hipMemcpyAsync(hmm, minmax, 8, hipMemcpyDeviceToHost, s));
  GB_IN_TRY(hipMemcpyAsync(&nnz64, tot, 8, hipMemcpyDeviceToHost, s));
  GB_IN_TRY(hipStreamSynchronize(s));
  if (nnz64 > 0xffffffffull) {
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;  // beyond 32-bit row pointers
    goto done;
  }
  nnz = (uint32_t)nnz64;
  {
    // src/io.c:401-402,433-435: mincol starts at 1, maxcol at 0
    const uint32_t mincol = hmm[0] < 1u ? hmm[0] : 1u;
    const uint32_t maxcol = hmm[1];
    // 4. row pointers
    GB_IN_TRY(gb_dev_alloc((void **)&rowptr, ((uint64_t)nrows + 1) * 4));
    GB_IN_RV(gbd_scan_u32(cnt, rowptr, nrows, rowptr + nrows));
    GB_IN_TRY(gb_dev_alloc((void **)&rowind, ((uint64_t)nnz + 1) * 4));
    GB_IN_TRY(gb_dev_alloc((void **)&rowval, ((uint64_t)nnz + 1) * 8));
    // 5. fill pass, again with a larger list if the first one overflowed
    for (int pass = 0; pass < 2; ++pass) {
      GB_IN_TRY(gb_dev_alloc((void **)&fix, (uint64_t)fixcap * sizeof(gb_in_fix)));
      if (!nfix) GB_IN_TRY(gb_dev_alloc((void **)&nfix, 4));
      GB_IN_RV(gbd_zero(nfix, 4));
      if (nlines) {
        gb_in_fill<<<gb_in_grid(nlines), GB_BLOCK, 0, s>>>(d_txt, len, ls, nl, nlines, rowof,
                                                          rowptr, mincol, rowind, rowval, fix,
                                                          fixcap, nfix);
        GB_IN_TRY(hipGetLastError());
      }
      GB_IN_TRY(hipMemcpyAsync(&hnfix, nfix, 4, hipMemcpyDeviceToHost, s));
      GB_IN_TRY(hipStreamSynchronize(s));
      if (hnfix <= fixcap) break;
      (void)gb_dev_free(fix);
      fix = nullptr;
      fixcap = hnfix;
    }
    if (hnfix > fixcap) {
      rv = GOOSEBERRY_ERROR_UNKNOWN;
      goto done;
    }
    if (hnfix) {
      gb_in_fixup<<<gb_div_up(hnfix, 64), 64, 0, s>>>(d_txt, fix, hnfix, rowval);
      GB_IN_TRY(hipGetLastError());
    }
    GB_IN_TRY(hipStreamSynchronize(s));
    *r_nrows = nrows;
    *r_ncols = maxcol - mincol + 1u;
    *r_nnz = nnz;
    *d_rowptr = rowptr;
    *d_rowind = rowind;
    *d_rowval = rowval;
    rowptr = rowind = nullptr;
    rowval = nullptr;
  }
  rv = GOOSEBERRY_SUCCESS;
done:
  (void)gb_dev_free(ccnt); (void)gb_dev_free(ls); (void)gb_dev_free(rowof); (void)gb_dev_free(cnt);
  (void)gb_dev_free(minmax); (void)gb_dev_free(tot); (void)gb_dev_free(fix); (void)gb_dev_free(nfix);
  (void)gb_dev_free(rowptr); (void)gb_dev_free(rowind); (void)gb_dev_free(rowval);
  return rv;
#undef GB_IN_TRY
#undef GB_IN_RV
}

// The body of a MatrixMarket file (text at d_txt[off, len)): the lines
// after the size line.  nrows, ncols and sym come from the header, parsed
// on the host.  Rows are filled by a stable sort of the entries' records by
// row, a symmetric file's mirrored entry first (src/io.c:1103-1144).
int gbd_parse_mtx_body(const char *d_txt, uint64_t len, uint64_t off, uint32_t nrows,
                       uint32_t ncols, int sym, uint32_t *r_nnz, uint32_t **d_rowptr,
                       uint32_t **d_rowind, double **d_rowval) {
  int rv = gb_ensure_init();
  if (rv != GOOSEBERRY_SUCCESS) return rv;
  if (off > len) return GOOSEBERRY_ERROR_INVALIDINPUT;
  hipStream_t s = gb_stream();
  hipError_t he = hipSuccess;
  const char *txt = d_txt + off;
  const uint64_t blen = len - off;
  const uint64_t nchunks = (blen + GB_IN_CHUNK - 1) / GB_IN_CHUNK;
  if (nchunks >= 0xffffffffull) return GOOSEBERRY_ERROR_INVALIDINPUT;
  uint32_t *ccnt = nullptr, *ent = nullptr, *ei = nullptr, *ej = nullptr, *flags = nullptr,
           *key = nullptr, *ord = nullptr, *rowptr = nullptr, *rowind = nullptr,
           *nfix = nullptr;
  uint64_t *ls = nullptr;
  double *ew = nullptr, *rowval = nullptr;
  gb_in_fix *fix = nullptr;
  uint32_t nl = 0, nlines = 0, nent = 0, hflags[3] = {0, 0, 0}, hnfix = 0, fixcap = GB_IN_FIX0;
  uint64_t last = 0, nnz64 = 0;
  uint32_t nnz = 0;
#define GB_MX_TRY(call)                                         \
  do {                                                          \
    he = (call);                                                \
    if (he != hipSuccess) {                                     \
      rv = gb_hip_fail(he, #call, __FILE__, __LINE__);          \
      goto done;                                                \
    }                                                           \
  } while (0)
#define GB_MX_RV(call)                                          \
  do {                                                          \
    rv = (call);                                                \
    if (rv != GOOSEBERRY_SUCCESS) goto done;                    \
  } while (0)
  // lines of the body
  GB_MX_TRY(gb_dev_alloc((void **)&ccnt, (nchunks + 1) * 4));
  if (nchunks) {
    gb_in_nl_count<<<(uint32_t)nchunks, GB_BLOCK, 0, s>>>(txt, blen, ccnt);
    GB_MX_TRY(hipGetLastError());
  }
  GB_MX_RV(gbd_scan_u32(ccnt, ccnt, nchunks, ccnt + nchunks));
  GB_MX_TRY(hipMemcpyAsync(&nl, ccnt + nchunks, 4, hipMemcpyDeviceToHost, s));
  GB_MX_TRY(hipStreamSynchronize(s));
  if (nl == 0xffffffffu) {
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto done;
  }
  GB_MX_TRY(gb_dev_alloc((void **)&ls, ((uint64_t)nl + 2) * 8));
  if (nchunks) {
    gb_in_nl_pos<<<(uint32_t)nchunks, GB_BLOCK, 0, s>>>(txt, blen, ccnt, ls);
    GB_MX_TRY(hipGetLastError());
    GB_MX_TRY(hipMemcpyAsync(&last, ls + nl, 8, hipMemcpyDeviceToHost, s));
    GB_MX_TRY(hipStreamSynchronize(s));
  }
  nlines = nl + (blen > last ? 1u : 0u);
  // entry lines and their ranks
  GB_MX_TRY(gb_dev_alloc((void **)&ent, ((uint64_t)nlines + 1) * 4));
  if (nlines) {
    gb_mx_flag<<<gb_in_grid(nlines), GB_BLOCK, 0, s>>>(txt, blen, ls, nl, nlines, ent);
    GB_MX_TRY(hipGetLastError());
  }
  GB_MX_RV(gbd_scan_u32(ent, ent, nlines, ent + nlines));
  GB_MX_TRY(hipMemcpyAsync(&nent, ent + nlines, 4, hipMemcpyDeviceToHost, s));
  GB_MX_TRY(hipStreamSynchronize(s));
  nnz64 = (uint64_t)nent * (sym ? 2u : 1u);
  if (nnz64 > 0xffffffffull) {
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto done;
  }
  nnz = (uint32_t)nnz64;
  GB_MX_TRY(gb_dev_alloc((void **)&ei, ((uint64_t)nent + 1) * 4));
  GB_MX_TRY(gb_dev_alloc((void **)&ej, ((uint64_t)nent + 1) * 4));
  GB_MX_TRY(gb_dev_alloc((void **)&ew, ((uint64_t)nent + 1) * 8));
  GB_MX_TRY(gb_dev_alloc((void **)&flags, 8));
  GB_MX_TRY(gb_dev_alloc((void **)&nfix, 4));
  for (int pass = 0; pass < 2; ++pass) {
    GB_MX_TRY(gb_dev_alloc((void **)&fix, (uint64_t)fixcap * sizeof(gb_in_fix)));
    GB_MX_RV(gbd_zero(flags, 8));
    GB_MX_RV(gbd_zero(nfix, 4));
    if (nlines) {
      gb_mx_parse<<<gb_in_grid(nlines), GB_BLOCK, 0, s>>>(txt, blen, ls, nl, nlines, ent, nrows,
                                                         ncols, sym, ei, ej, ew, flags,
                                                         flags + 1, fix, fixcap, nfix);
      GB_MX_TRY(hipGetLastError());
    }
    GB_MX_TRY(hipMemcpyAsync(hflags, flags, 8, hipMemcpyDeviceToHost, s));
    GB_MX_TRY(hipMemcpyAsync(&hnfix, nfix, 4, hipMemcpyDeviceToHost, s));
    GB_MX_TRY(hipStreamSynchronize(s));
    if (hnfix <= fixcap) break;
    (void)gb_dev_free(fix);
    fix = nullptr;
    fixcap = hnfix;
  }
  // a coordinate missing or out of range; weights on some lines only
  if (hflags[0] || (hflags[1] != 0 && hflags[1] != nent) || hnfix > fixcap) {
    rv = GOOSEBERRY_ERROR_INVALIDINPUT;
    goto done;
  }
  if (hnfix) {
    gb_in_fixup<<<gb_div_up(hnfix, 64), 64, 0, s>>>(txt, fix, hnfix, ew);
    GB_MX_TRY(hipGetLastError());
  }
  // rows: a stable sort of the records by row gives the fill order and,
  // through its run starts, the row pointers
  GB_MX_TRY(gb_dev_alloc((void **)&rowptr, ((uint64_t)nrows + 2) * 4));
  GB_MX_TRY(gb_dev_alloc((void **)&rowind, ((uint64_t)nnz + 1) * 4));
  GB_MX_TRY(gb_dev_alloc((void **)&rowval, ((uint64_t)nnz + 1) * 8));
  if (nnz) {
    GB_MX_TRY(gb_dev_alloc((void **)&key, (uint64_t)nnz * 4));
    GB_MX_TRY(gb_dev_alloc((void **)&ord, (uint64_t)nnz * 4));
    gb_mx_keys<<<gb_in_grid(nent), GB_BLOCK, 0, s>>>(nent, sym, ei, ej, key);
    GB_MX_TRY(hipGetLastError());
    GB_MX_RV(gbd_countingsort_kv_unchecked(key, nullptr, 0, nrows - 1, nnz, ord, rowptr));
    gb_mx_gather<<<gb_in_grid(nnz), GB_BLOCK, 0, s>>>(nnz, sym, hflags[1] != 0, ord, ei, ej, ew,
                                                     rowind, rowval);
    GB_MX_TRY(hipGetLastError());
  } else {
    GB_MX_RV(gbd_zero(rowptr, ((uint64_t)nrows + 1) * 4));
  }
  GB_MX_TRY(hipStreamSynchronize(s));
  *r_nnz = nnz;
  *d_rowptr = rowptr;
  *d_rowind = rowind;
  *d_rowval = rowval;
  rowptr = rowind = nullptr;
  rowval = nullptr;
  rv = GOOSEBERRY_SUCCESS;
done:
  (void)gb_dev_free(ccnt); (void)gb_dev_free(ls); (void)gb_dev_free(ent); (void)gb_dev_free(ei); (void)gb_dev_free(ej);
  (void)gb_dev_free(ew); (void)gb_dev_free(flags); (void)gb_dev_free(nfix); (void)gb_dev_free(fix); (void)gb_dev_free(key);
  (void)gb_dev_free(ord); (void)gb_dev_free(rowptr); (void)gb_dev_free(rowind); (void)gb_dev_free(rowval);
  return rv;
#undef GB_MX_TRY
#undef GB_MX_RV
}

}  // extern "C"

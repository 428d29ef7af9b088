/* gb_host.h -- internal declarations shared by the host C files. */
#ifndef GB_HOST_H
#define GB_HOST_H

#include <stdint.h>
#include <stdio.h>

#include "gooseberry.h"
#include "../gb_device.h"

int gb_upload(void **d, const void *h, size_t bytes);
int gb_check_csr(dim_t nrows, dim_t ncols, const ind_t *rowptr,
    const dim_t *rowind);
gbd_matrix *gb_matrix_dev(gooseberry_matrix_t mat);
/* GOOSEBERRY_ERROR_UNIMPLEMENTED, with its error text, for operations that
 * take no matrix held as row parts */
int gb_parts_unsupported(void);
int gb_matrix_adopt(dim_t nrows, dim_t ncols, ind_t *d_rowptr,
    dim_t *d_rowind, real_t *d_rowval, ind_t nnz,
    const gooseberry_plan_t *plan, gooseberry_matrix_t *r_mat);

/* reads and checks a binary CSR header (the file must be exactly the size
 * it announces); leaves f at the row pointers */
int gb_bcsr_header(FILE *f, dim_t *nrows, dim_t *ncols, uint64_t *nnz);

/* One recursive library lock serialises the public entry points: the
 * library stream, its cached graphs and workspaces and the block cache are
 * shared by every caller.  GB_API_GUARD() takes it for the rest of the
 * enclosing function and releases it on every return path. */
void gb_api_lock(void);
void gb_api_unlock(int *scope);
#define GB_API_GUARD()                                                    \
  int gb_api_scope_ __attribute__((cleanup(gb_api_unlock), unused)) =     \
      (gb_api_lock(), 0)

#endif

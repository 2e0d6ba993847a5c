/*
 * merbit_host.h -- host-side graph ingest with the reference's semantics (C ABI).
 *
 * These replace the input side of the reference's SpMV driver: the config
 * registry lookup and MTX reader of GRAPH::GRAPH (src/class.cu:60-141), the
 * PageRank value rewrite Fun_Values (src/class.cu:22-34), the (col,row) sort
 * GRAPH::sort(false) (src/class.cu:143-177) and cusparseXcoo2csr
 * (include/tile.cuh:571-574).  All arrays are host memory owned by the
 * caller.  Functions return 0 or a negative MBX_* status (merbit_amd.h).
 */
#ifndef MERBIT_HOST_H
#define MERBIT_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Config registry lookup (src/class.cu:71-77): lines "ID path [nrow ncol nnz]";
 * copies the path of `gid` into path_out.  A missing ID is MBX_EINVAL (the
 * reference loops forever). */
int mbh_config_lookup(const char* config_file, int32_t gid, char* path_out, int32_t cap);

/* MTX reader (src/class.cu:79-123).  A first line containing "symmetric"
 * mirrors every entry, the diagonal included (nnz doubles, duplicates kept);
 * "basezero" means 0-based indices; '%' lines are comments.  Square only
 * (nrow = the declared row count).  Call with row_ind == NULL to get nrow and
 * nnz, then with arrays of nnz entries (val may be NULL; absent values read
 * as 1). */
int mbh_mtx_read(const char* path, int32_t* nrow, int64_t* nnz, int32_t* row_ind,
                 int32_t* col_ind, double* val);

/* CSR of A^T with values 1/out-degree(row) from a loaded COO: the matrix
 * Fun_Merbit operates on.  row_ptr_t[nrow+1], col_t[nnz], val_t[nnz];
 * columns sorted within rows. */
int mbh_ref_ingest(int32_t nrow, int64_t nnz, const int32_t* row_ind,
                   const int32_t* col_ind, int32_t* row_ptr_t, int32_t* col_t,
                   double* val_t);

#ifdef __cplusplus
}
#endif
#endif /* MERBIT_HOST_H */

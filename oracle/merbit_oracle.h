/*
 * merbit_oracle.h -- CPU restatement of MERBIT's merge-path tile SpMV.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (merbit_amd/, include/)
 * links, loads or calls this library.  It is used by tests/, by the smoke
 * check in __graft_entry__.py and by the cpu_baseline leg of bench.py, and
 * there only as the checker or the reported CPU baseline.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  Pinning of each
 * function (what independent check holds it) is listed in merbit_oracle.c.
 */
#ifndef MERBIT_ORACLE_H
#define MERBIT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/utils.cuh:5-29 merge_search: coordinate (coord_x = nnz consumed,
 * coord_y = rows consumed) where merge diagonal `diag` crosses the path that
 * merges row ends row_ptr[1..nrow] with nnz indices 0..nnz-1. */
void orc_merge_search(const int32_t* row_ptr, int32_t nrow, int32_t nnz,
                      int64_t diag, int32_t* coord_x, int32_t* coord_y);

/* include/tile.cuh:28-122 TILE/Tile_Kernel: the reference's 32-lane plan.
 * tile_x/tile_y: uint32[tile_num+1] (MSB of tile_y = "fast" tile), lane_desc:
 * uint64[lane_num] holding the reference bit image (u32 layout
 * flag<<18|dy<<9|dx for sigma<15, u64 layout flag<<22|dy<<11|dx otherwise).
 * Returns lane_num; tile_num = (lane_num+31)/32.  Returns -1 on bad sigma. */
int32_t orc_ref_lane_count(int32_t nrow, int32_t nnz, int32_t sigma);
int32_t orc_ref_plan(const int32_t* row_ptr, int32_t nrow, int32_t nnz,
                     int32_t sigma, uint32_t* tile_x, uint32_t* tile_y,
                     uint64_t* lane_desc);

/* Plain row-sequential CSR y = alpha*A*x + beta*y in FP64 (beta == 0: y is
 * not read).  The textbook CSR product the reference's comparator rows
 * compute (src/spmv.cu:6-64 via cuSPARSE). */
void orc_spmv_csr_f64(int32_t nrow, const int32_t* row_ptr, const int32_t* col,
                      const double* val, double alpha, const double* x,
                      double beta, double* y);

/* Accurate reference: y_i = sum_k val_k*x_col_k evaluated with error-free
 * TwoProduct (fma) and TwoSum (Knuth), i.e. Ogita-Rump-Oishi Dot2 in FP64,
 * plus absr_i = sum_k |val_k*x_col_k| (the condition denominator).  absr may
 * be NULL.  Values/x may be given as double (ty 0), float (ty 1) or bf16
 * bit patterns (ty 2, uint16); x as double (tx 0) or float (tx 1). */
void orc_spmv_accurate(int32_t nrow, const int32_t* row_ptr, const int32_t* col,
                       const void* val, int32_t val_type, const void* x,
                       int32_t x_type, double* y, double* absr);

/* Reference-order emulator: reproduces the summation grouping of one
 * merbit_kernel launch (include/tile.cuh:476-561 with fast_track 267-290,
 * load_smem 293-305, normal_track 308-352, load_mem 354-375 and
 * warp_segmented_sum include/utils.cuh:58-83), on a zeroed output
 * (tile.cuh:597), with shared/global atomics applied in warp order then block
 * order.  FP64, BLOCK = 256, 32-lane tiles.  Returns 0 or -1 on bad sigma. */
int32_t orc_merbit_emulate(int32_t nrow, int32_t nnz, const int32_t* row_ptr,
                           const int32_t* col, const double* val, double alpha,
                           const double* x, int32_t sigma, double* y);

/* Reference ingest (src/class.cu:22-34 values := 1/out-degree(row),
 * src/class.cu:143-177 sort by (col,row), include/tile.cuh:571-574 coo2csr):
 * from a loaded COO (row_ind, col_ind) build the CSR of A^T that Fun_Merbit
 * operates on.  row_ptr_t[nrow+1], col_t[nnz], val_t[nnz] (double). */
void orc_ref_ingest(int32_t nrow, int32_t nnz, const int32_t* row_ind,
                    const int32_t* col_ind, int32_t* row_ptr_t, int32_t* col_t,
                    double* val_t);

/* Transpose of a CSR (counting sort by column, stable): CSC as CSR of A^T. */
void orc_csr_transpose(int32_t nrow, int32_t ncol, int32_t nnz,
                       const int32_t* row_ptr, const int32_t* col,
                       int32_t* col_ptr, int32_t* row_idx, int32_t* perm);

/* CPU baseline: FP64 CSR SpMV parallelised with OpenMP over nnz-balanced
 * static row chunks.  Runs `warmup` untimed then `reps` timed products and
 * returns the median seconds per product; *threads_used receives the
 * thread count. */
double orc_spmv_csr_omp_time(int32_t nrow, const int32_t* row_ptr,
                             const int32_t* col, const double* val,
                             const double* x, double* y, int32_t nthreads,
                             int32_t warmup, int32_t reps,
                             int32_t* threads_used);

#ifdef __cplusplus
}
#endif
#endif

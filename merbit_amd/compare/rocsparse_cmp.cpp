// rocsparse_cmp.cpp -- rocSPARSE CSR SpMV as a measured comparator
// (libmerbit_cmp.so).  Not part of the product path: the benchmark times it
// beside the merge-path kernel on the same device arrays, in place of the
// reference's vendor-library rows (src/spmv.cu:6-64 Fun_CSRMV, 124-157
// Fun_MERGEMV).  The analysis / preprocessing stage runs once at create, as
// the reference does before its timed loop.  The COO algorithms (the
// reference's Fun_COOMV row, spmv.cu:69-118, which runs on its sorted COO
// arrays) take their row indices from rocsparse_csr2coo, run once at create.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cstdint>
#include <new>

#include "merbit_amd.h"

extern "C" {

// Algorithm ids accepted by mbc_rocsparse_create.
enum {
  MBC_ADAPTIVE = 0,
  MBC_LRB = 1,
  MBC_ROWSPLIT = 2,
  MBC_NNZSPLIT = 3,
  MBC_COO = 4,         // segmented COO (deterministic)
  MBC_COO_ATOMIC = 5,  // atomic COO (summation order not fixed)
};

struct mbc_rocsparse {
  rocsparse_handle handle = nullptr;
  rocsparse_spmat_descr mat = nullptr;
  rocsparse_dnvec_descr x = nullptr, y = nullptr;
  rocsparse_datatype type = rocsparse_datatype_f64_r;
  rocsparse_spmv_alg alg = rocsparse_spmv_alg_csr_adaptive;
  void* buffer = nullptr;
  size_t buffer_bytes = 0;
  int32_t* coo_row = nullptr;  // COO algorithms: row index of every entry
};

void mbc_rocsparse_destroy(mbc_rocsparse* h);

static int fail(mbc_rocsparse* h, int st) {
  mbc_rocsparse_destroy(h);
  return st;
}

static int spmv_stage(mbc_rocsparse* h, rocsparse_spmv_stage stage) {
  const double one64 = 1.0, zero64 = 0.0;
  const float one32 = 1.0f, zero32 = 0.0f;
  const bool f64 = h->type == rocsparse_datatype_f64_r;
  const void* alpha = f64 ? (const void*)&one64 : (const void*)&one32;
  const void* beta = f64 ? (const void*)&zero64 : (const void*)&zero32;
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdeprecated-declarations"
  const rocsparse_status st =
      rocsparse_spmv(h->handle, rocsparse_operation_none, alpha, h->mat, h->x, beta, h->y,
                     h->type, h->alg, stage, &h->buffer_bytes, h->buffer);
#pragma GCC diagnostic pop
  return st == rocsparse_status_success ? MBX_OK : MBX_EHIP;
}

// y = A*x for CSR (int32 indices) with FP64 (value_kind MBX_VALUE_F64) or FP32
// (MBX_VALUE_F32) values, x and y; all pointers device memory.  Allocates the
// algorithm's buffer and runs its preprocessing on `stream` (synchronised).
int mbc_rocsparse_create(int32_t value_kind, int32_t alg, int32_t nrow, int32_t ncol,
                         int32_t nnz, const int32_t* row_ptr, const int32_t* col,
                         const void* val, const void* x, void* y, void* stream,
                         mbc_rocsparse** out) {
  if (!out || nrow < 0 || ncol < 0 || nnz < 0) return MBX_EINVAL;
  *out = nullptr;
  mbc_rocsparse* h = new (std::nothrow) mbc_rocsparse();
  if (!h) return MBX_ENOMEM;
  if (value_kind == MBX_VALUE_F64)
    h->type = rocsparse_datatype_f64_r;
  else if (value_kind == MBX_VALUE_F32)
    h->type = rocsparse_datatype_f32_r;
  else
    return fail(h, MBX_EUNSUPPORTED);
  switch (alg) {
    case MBC_ADAPTIVE: h->alg = rocsparse_spmv_alg_csr_adaptive; break;
    case MBC_LRB: h->alg = rocsparse_spmv_alg_csr_lrb; break;
    case MBC_ROWSPLIT: h->alg = rocsparse_spmv_alg_csr_rowsplit; break;
    case MBC_NNZSPLIT: h->alg = rocsparse_spmv_alg_csr_nnzsplit; break;
    case MBC_COO: h->alg = rocsparse_spmv_alg_coo; break;
    case MBC_COO_ATOMIC: h->alg = rocsparse_spmv_alg_coo_atomic; break;
    default: return fail(h, MBX_EUNSUPPORTED);
  }
  const bool coo = alg == MBC_COO || alg == MBC_COO_ATOMIC;
  if (rocsparse_create_handle(&h->handle) != rocsparse_status_success) return fail(h, MBX_EHIP);
  if (rocsparse_set_stream(h->handle, (hipStream_t)stream) != rocsparse_status_success)
    return fail(h, MBX_EHIP);
  if (coo) {
    if (hipMalloc((void**)&h->coo_row, sizeof(int32_t) * (nnz > 0 ? (size_t)nnz : 1)) !=
        hipSuccess) {
      h->coo_row = nullptr;
      return fail(h, MBX_ENOMEM);
    }
    if (nnz > 0 && rocsparse_csr2coo(h->handle, row_ptr, nnz, nrow, h->coo_row,
                                     rocsparse_index_base_zero) != rocsparse_status_success)
      return fail(h, MBX_EHIP);
  }
  const rocsparse_status mst =
      coo ? rocsparse_create_coo_descr(&h->mat, nrow, ncol, nnz, h->coo_row,
                                       const_cast<int32_t*>(col), const_cast<void*>(val),
                                       rocsparse_indextype_i32, rocsparse_index_base_zero,
                                       h->type)
          : rocsparse_create_csr_descr(&h->mat, nrow, ncol, nnz, const_cast<int32_t*>(row_ptr),
                                       const_cast<int32_t*>(col), const_cast<void*>(val),
                                       rocsparse_indextype_i32, rocsparse_indextype_i32,
                                       rocsparse_index_base_zero, h->type);
  if (mst != rocsparse_status_success ||
      rocsparse_create_dnvec_descr(&h->x, ncol, const_cast<void*>(x), h->type) !=
          rocsparse_status_success ||
      rocsparse_create_dnvec_descr(&h->y, nrow, y, h->type) != rocsparse_status_success)
    return fail(h, MBX_EHIP);
  int st = spmv_stage(h, rocsparse_spmv_stage_buffer_size);
  if (st != MBX_OK) return fail(h, st);
  if (hipMalloc(&h->buffer, h->buffer_bytes > 0 ? h->buffer_bytes : 16) != hipSuccess) {
    h->buffer = nullptr;
    return fail(h, MBX_ENOMEM);
  }
  st = spmv_stage(h, rocsparse_spmv_stage_preprocess);
  if (st != MBX_OK) return fail(h, st);
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return fail(h, MBX_EHIP);
  *out = h;
  return MBX_OK;
}

// One product on the stream given at create; no synchronisation.
int mbc_rocsparse_spmv(mbc_rocsparse* h) {
  if (!h) return MBX_EINVAL;
  return spmv_stage(h, rocsparse_spmv_stage_compute);
}

// Bytes of the algorithm's analysis buffer (reported beside the timings).
int64_t mbc_rocsparse_buffer_bytes(const mbc_rocsparse* h) {
  return h ? (int64_t)h->buffer_bytes : -1;
}

void mbc_rocsparse_destroy(mbc_rocsparse* h) {
  if (!h) return;
  if (h->x) rocsparse_destroy_dnvec_descr(h->x);
  if (h->y) rocsparse_destroy_dnvec_descr(h->y);
  if (h->mat) rocsparse_destroy_spmat_descr(h->mat);
  if (h->handle) rocsparse_destroy_handle(h->handle);
  if (h->buffer) (void)hipFree(h->buffer);
  if (h->coo_row) (void)hipFree(h->coo_row);
  delete h;
}

}  // extern "C"

// merbit_tune.hip -- timing autotuner for the tile shape.
//
// For a matrix (or batch) already in HBM, builds a plan for every
// instantiated items-per-thread value, times `reps` products of each with
// HIP events after two warm-up calls, and returns the fastest.  The tile size
// is 256*items_per_thread merge items.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "merbit_amd.h"

namespace {

const int kSpmvCandidates[] = {2, 4, 6, 8, 12, 16};
const int kSpmmCandidates[] = {2, 4};

template <class Call>
int time_calls(hipStream_t s, int reps, Call call, float* us) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    return MBX_EHIP;
  }
  int st = MBX_OK;
  for (int i = 0; i < 2 && st == MBX_OK; ++i) st = call();
  if (st == MBX_OK && hipEventRecord(a, s) != hipSuccess) st = MBX_EHIP;
  for (int i = 0; i < reps && st == MBX_OK; ++i) st = call();
  float ms = 0.f;
  if (st == MBX_OK && (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                       hipEventElapsedTime(&ms, a, b) != hipSuccess))
    st = MBX_EHIP;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *us = ms * 1000.f / (float)reps;
  return st;
}

int spmv_call(const mbx_plan* p, int kind, const int32_t* rp, const int32_t* col,
              const void* val, const void* x, void* y, hipStream_t s) {
  switch (kind) {
    case MBX_VALUE_F64:
      return mbx_spmv_f64(p, rp, col, (const double*)val, 1.0, (const double*)x, 0.0,
                          (double*)y, s);
    case MBX_VALUE_F32:
      return mbx_spmv_f32(p, rp, col, (const float*)val, 1.f, (const float*)x, 0.f, (float*)y,
                          s);
    case MBX_VALUE_BF16:
      return mbx_spmv_bf16(p, rp, col, (const uint16_t*)val, 1.f, (const float*)x, 0.f,
                           (float*)y, s);
    default:
      return MBX_EINVAL;
  }
}

}  // namespace

extern "C" {

int mbx_autotune_spmv(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                      const int32_t* d_col, const void* d_val, int32_t value_kind,
                      const void* d_x, void* d_y, int32_t reps, void* stream, int32_t* best_ipt,
                      float* best_us) {
  if (!best_ipt || !best_us || reps < 1) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  *best_ipt = 0;
  *best_us = 0.f;
  for (int ipt : kSpmvCandidates) {
    mbx_plan* p = nullptr;
    int st = mbx_plan_create(d_row_ptr, nrow, ncol, nnz, ipt, 256, stream, &p);
    if (st == MBX_EOVERFLOW) continue;
    if (st != MBX_OK) return st;
    float us = 0.f;
    st = time_calls(s, reps,
                    [&] { return spmv_call(p, value_kind, d_row_ptr, d_col, d_val, d_x, d_y, s); },
                    &us);
    mbx_plan_destroy(p);
    if (st != MBX_OK) return st;
    if (*best_ipt == 0 || us < *best_us) {
      *best_ipt = ipt;
      *best_us = us;
    }
  }
  return *best_ipt ? MBX_OK : MBX_EOVERFLOW;
}

int mbx_autotune_spmm(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                      const int32_t* d_col, const float* d_val, int32_t k, const float* d_x,
                      int64_t ldx, int32_t x_layout, float* d_y, int64_t ldy, int32_t y_layout,
                      void* d_workspace, size_t workspace_bytes, int32_t reps, void* stream,
                      int32_t* best_ipt, float* best_us) {
  if (!best_ipt || !best_us || reps < 1) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  *best_ipt = 0;
  *best_us = 0.f;
  for (int ipt : kSpmmCandidates) {
    mbx_plan* p = nullptr;
    int st = mbx_plan_create(d_row_ptr, nrow, ncol, nnz, ipt, 256, stream, &p);
    if (st == MBX_EOVERFLOW) continue;
    if (st != MBX_OK) return st;
    float us = 0.f;
    st = time_calls(s, reps,
                    [&] {
                      return mbx_spmm_f32(p, d_row_ptr, d_col, d_val, k, 1.f, d_x, ldx, x_layout,
                                          0.f, d_y, ldy, y_layout, d_workspace, workspace_bytes,
                                          s);
                    },
                    &us);
    mbx_plan_destroy(p);
    if (st != MBX_OK) return st;
    if (*best_ipt == 0 || us < *best_us) {
      *best_ipt = ipt;
      *best_us = us;
    }
  }
  return *best_ipt ? MBX_OK : MBX_EOVERFLOW;
}

int mbx_autotune_batch(const mbx_batch_item* items, int32_t count, int32_t value_kind,
                       int32_t reps, void* stream, int32_t* best_ipt, float* best_us) {
  if (!best_ipt || !best_us || reps < 1) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  *best_ipt = 0;
  *best_us = 0.f;
  for (int ipt : kSpmvCandidates) {
    mbx_batch* b = nullptr;
    int st = mbx_batch_create(items, count, value_kind, ipt, stream, &b);
    if (st == MBX_EOVERFLOW) continue;
    if (st != MBX_OK) return st;
    float us = 0.f;
    st = time_calls(s, reps, [&] { return mbx_batch_spmv(b, 1.0, 0.0, s); }, &us);
    mbx_batch_destroy(b);
    if (st != MBX_OK) return st;
    if (*best_ipt == 0 || us < *best_us) {
      *best_ipt = ipt;
      *best_us = us;
    }
  }
  return *best_ipt ? MBX_OK : MBX_EOVERFLOW;
}

}  // extern "C"

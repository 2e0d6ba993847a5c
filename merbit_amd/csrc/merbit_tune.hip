// merbit_tune.hip -- timing autotuner for the tile shape, and the pattern
// statistics a caller (or a tuning model) can use as features.
//
// For a matrix (or batch) already in HBM, builds a plan for every
// instantiated items-per-thread value, times `reps` products of each with
// HIP events after two warm-up calls, and returns the fastest.  The tile size
// is 256*items_per_thread merge items.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

namespace {

const int kSpmvCandidates[] = {2, 4, 6, 8, 10, 12, 14, 16};
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

namespace {

// Items per thread x stream policy x tiles per CU, and with `sorted` the
// same shapes with tile-sorted gathers where a plan takes them.
int autotune_launch(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                    const int32_t* d_col, const void* d_val, int32_t value_kind,
                    const void* d_x, void* d_y, int32_t reps, hipStream_t s, bool sorted,
                    int32_t* best_ipt, int32_t* best_stream, int32_t* best_wg_per_cu,
                    int32_t* best_sorted, float* best_us) {
  const int kModes[] = {MBX_STREAM_CACHED, MBX_STREAM_NONTEMPORAL};
  const int kResidency[] = {0, 3, 4, 5};
  *best_ipt = 0;
  *best_us = 0.f;
  if (best_sorted) *best_sorted = 0;
  for (int ipt : kSpmvCandidates) {
    mbx_plan* p = nullptr;
    int st = mbx_plan_create(d_row_ptr, nrow, ncol, nnz, ipt, 256, s, &p);
    if (st == MBX_EOVERFLOW) continue;
    if (st != MBX_OK) return st;
    for (int form = 0; form < (sorted ? 2 : 1); ++form) {
      if (form == 1) {
        st = mbx_plan_sort_cols(p, d_col, s);
        if (st == MBX_EUNSUPPORTED || st == MBX_EINVAL || nnz == 0) break;  // plain only
        if (st != MBX_OK) {
          mbx_plan_destroy(p);
          return st;
        }
      }
      for (int mode : kModes) {
        for (int wg : kResidency) {
          p->stream_mode = mode;
          p->wg_per_cu = wg;
          float us = 0.f;
          st = time_calls(
              s, reps,
              [&] { return spmv_call(p, value_kind, d_row_ptr, d_col, d_val, d_x, d_y, s); },
              &us);
          if (st != MBX_OK) {
            mbx_plan_destroy(p);
            return st;
          }
          if (*best_ipt == 0 || us < *best_us) {
            *best_ipt = ipt;
            *best_stream = mode;
            *best_wg_per_cu = wg;
            if (best_sorted) *best_sorted = form;
            *best_us = us;
          }
        }
      }
    }
    mbx_plan_destroy(p);
  }
  return *best_ipt ? MBX_OK : MBX_EOVERFLOW;
}

}  // namespace

extern "C" {

int mbx_plan_stats(const mbx_plan* p, const int32_t* d_row_ptr, mbx_pattern_stats* out) {
  if (!p || !out || (p->nrow > 0 && !d_row_ptr)) return MBX_EINVAL;
  std::vector<int32_t> rp((size_t)p->nrow + 1);
  std::vector<int4> coord((size_t)p->ntiles + 1);
  if (hipMemcpy(rp.data(), d_row_ptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost) !=
          hipSuccess ||
      (p->ntiles > 0 && hipMemcpy(coord.data(), p->d_coord, sizeof(int4) * coord.size(),
                                  hipMemcpyDeviceToHost) != hipSuccess))
    return MBX_EHIP;
  mbx_pattern_stats st{};
  st.nrow = p->nrow;
  st.ncol = p->ncol;
  st.nnz = p->nnz;
  st.ntiles = p->ntiles;
  st.tile_items = p->tile_items;
  // row-length moments about the mean (population), GRAPH::desc
  // (src/class.cu:255-320); empty rows as GRAPH::dangling (class.cu:181-216)
  const double n = (double)std::max(p->nrow, 1);
  const double mean = (double)p->nnz / n;
  double m2 = 0.0, m3 = 0.0, m4 = 0.0;
  for (int32_t r = 0; r < p->nrow; ++r) {
    const int32_t len = rp[(size_t)r + 1] - rp[(size_t)r];
    if (len == 0) ++st.empty_rows;
    st.max_row = std::max(st.max_row, len);
    const double d = (double)len - mean;
    m2 += d * d;
    m3 += d * d * d;
    m4 += d * d * d * d;
  }
  st.mean = mean;
  st.var = m2 / n;
  st.skew = st.var > 0.0 ? m3 / n / std::pow(st.var, 1.5) : 0.0;
  st.kurt = st.var > 0.0 ? m4 / n / (st.var * st.var) : 0.0;
  // tile statistics from the plan's coordinates (TILE::desc,
  // include/tile.cuh:177-233: tiles inside one row are its fast tiles)
  for (int32_t t = 0; t < p->ntiles; ++t) {
    const int4 c0 = coord[(size_t)t], c1 = coord[(size_t)t + 1];
    const int32_t rows = c1.x - c0.x;
    st.max_rows_per_tile = std::max(st.max_rows_per_tile, rows);
    if (rows == 0) ++st.single_row_tiles;
    if (t > 0 && rows > 0 && c0.z < c0.y) {
      ++st.split_tiles;
      const long long ts = ((long long)c0.x + c0.z) / p->tile_items;
      if (t - ts > 64) ++st.long_rows;
    }
  }
  *out = st;
  return MBX_OK;
}

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

int mbx_autotune_spmv_launch(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                             const int32_t* d_col, const void* d_val, int32_t value_kind,
                             const void* d_x, void* d_y, int32_t reps, void* stream,
                             int32_t* best_ipt, int32_t* best_stream, int32_t* best_wg_per_cu,
                             float* best_us) {
  if (!best_ipt || !best_stream || !best_wg_per_cu || !best_us || reps < 1) return MBX_EINVAL;
  return autotune_launch(d_row_ptr, nrow, ncol, nnz, d_col, d_val, value_kind, d_x, d_y, reps,
                         (hipStream_t)stream, false, best_ipt, best_stream, best_wg_per_cu,
                         nullptr, best_us);
}

int mbx_autotune_spmv_form(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                           const int32_t* d_col, const void* d_val, int32_t value_kind,
                           const void* d_x, void* d_y, int32_t reps, void* stream,
                           int32_t* best_ipt, int32_t* best_stream, int32_t* best_wg_per_cu,
                           int32_t* best_sorted, float* best_us) {
  if (!best_ipt || !best_stream || !best_wg_per_cu || !best_sorted || !best_us || reps < 1)
    return MBX_EINVAL;
  return autotune_launch(d_row_ptr, nrow, ncol, nnz, d_col, d_val, value_kind, d_x, d_y, reps,
                         (hipStream_t)stream, true, best_ipt, best_stream, best_wg_per_cu,
                         best_sorted, best_us);
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

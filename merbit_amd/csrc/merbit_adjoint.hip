// merbit_adjoint.hip -- the adjoint of y = A*x with respect to x and to the
// values of A (no reference counterpart: the reference computes y only).
//
//   dx   = A^T g         by the merge-path tile SpMV on the transposed pattern
//                        (CSC of A, built once per pattern on the device with a
//                        stable radix sort, so every column keeps its rows in
//                        ascending order and the sums are deterministic); the
//                        values are gathered into CSC order per call.
//   dval = g[row]*x[col] (SDDMM over the pattern) walking the forward plan's
//                        merge-path tiles, which give every nonzero its row.

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>
#include <new>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

using mbx_detail::lds_merge_search;
using mbx_detail::xcd_tile;

namespace {

// Row of every nonzero: the last r with row_ptr[r] <= k.
__global__ void row_of_nnz_kernel(const int32_t* __restrict__ row_ptr, int nrow, int nnz,
                                  int32_t* __restrict__ row, int32_t* __restrict__ iota) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  int lo = 0, hi = nrow - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (row_ptr[mid] <= k) lo = mid; else hi = mid - 1;
  }
  row[k] = lo;
  iota[k] = k;
}

// Column pointer of the sorted keys: col_ptr[c] = first p with key[p] >= c.
__global__ void col_ptr_kernel(const uint32_t* __restrict__ key, int nnz, int ncol,
                               int32_t* __restrict__ col_ptr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > nnz) return;
  const int lo = (p == 0) ? -1 : (int)key[p - 1];
  const int hi = (p == nnz) ? ncol : (int)key[p];
  for (int c = lo + 1; c <= hi; ++c) col_ptr[c] = p;
}

__global__ void transpose_rows_kernel(const int32_t* __restrict__ perm,
                                      const int32_t* __restrict__ row, int nnz,
                                      int32_t* __restrict__ row_idx) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nnz) row_idx[p] = row[perm[p]];
}

template <class T>
__global__ void gather_kernel(const T* __restrict__ val, const int32_t* __restrict__ perm,
                              int nnz, T* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nnz) out[p] = val[perm[p]];
}

// dval[k] = g[row(k)] * x[col(k)] for the nonzeros of one forward tile.
template <int WG, int IPT, class T>
__global__ __launch_bounds__(WG) void sddmm_tile_kernel(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const T* __restrict__ g, const T* __restrict__ x, T* __restrict__ dval,
    const int4* __restrict__ coord, int ntiles) {
  constexpr int TI = WG * IPT;
  __shared__ int s_re[TI + 1];
  __shared__ int s_row[TI];
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int tid = threadIdx.x;
  const int4 c0 = coord[tile];
  const int4 c1 = coord[tile + 1];
  const int r0 = c0.x, k0 = c0.y;
  const int rows_t = c1.x - r0;
  const int nnz_t = c1.y - k0;
  int ci[IPT];
  {
    int re[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      re[j] = INT_MAX;
      ci[j] = 0;
      if (i < rows_t) re[j] = row_ptr[r0 + 1 + i];
      if (i < nnz_t) ci[j] = col[k0 + i];
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      if (i < rows_t) s_re[i] = re[j] - k0;
    }
    if (tid == 0) s_re[rows_t] = INT_MAX;
  }
  __syncthreads();
  const int d0 = tid * IPT;
  const int items = min(max(nnz_t + rows_t - d0, 0), IPT);
  if (items > 0) {
    int r = lds_merge_search(s_re, d0, nnz_t, rows_t);
    int k = d0 - r;
    int re = s_re[r];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      if (i < items) {
        if (k < re) {
          s_row[k] = r;
          ++k;
        } else {
          ++r;
          re = s_re[r];
        }
      }
    }
  }
  __syncthreads();
  T xv[IPT], gv[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = tid + j * WG;
    xv[j] = T(0);
    gv[j] = T(0);
    if (i < nnz_t) {
      xv[j] = x[ci[j]];
      gv[j] = g[r0 + s_row[i]];
    }
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = tid + j * WG;
    if (i < nnz_t) dval[k0 + i] = gv[j] * xv[j];
  }
}

constexpr int kBlock = 256;

inline int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

struct mbx_adjoint {
  int32_t nrow, ncol, nnz;
  int32_t* col_ptr;   // ncol + 1: row pointer of A^T
  int32_t* row_idx;   // nnz: column indices of A^T
  int32_t* perm;      // nnz: CSR position of every CSC entry
  void* vt;           // nnz values in CSC order (8 bytes each)
  mbx_plan* plan_t;   // merge-path plan of A^T
};

extern "C" {

int mbx_adjoint_create(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                       int32_t ncol, int32_t nnz, int32_t items_per_thread, void* stream,
                       mbx_adjoint** out) {
  if (!out) return MBX_EINVAL;
  *out = nullptr;
  if (nrow < 0 || ncol < 0 || nnz < 0 || (nrow > 0 && !d_row_ptr) || (nnz > 0 && !d_col))
    return MBX_EINVAL;
  if (nnz > 0 && (nrow == 0 || ncol == 0)) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  mbx_adjoint* a = new (std::nothrow) mbx_adjoint();
  if (!a) return MBX_ENOMEM;
  a->nrow = nrow;
  a->ncol = ncol;
  a->nnz = nnz;
  const size_t m = (size_t)(nnz > 0 ? nnz : 1);
  int32_t *row = nullptr, *iota = nullptr;
  uint32_t* keys = nullptr;
  void* tmp = nullptr;
  int st = MBX_OK;
  if (hipMalloc((void**)&a->col_ptr, sizeof(int32_t) * ((size_t)ncol + 1)) != hipSuccess ||
      hipMalloc((void**)&a->row_idx, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc((void**)&a->perm, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc(&a->vt, 8 * m) != hipSuccess || hipMalloc((void**)&row, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc((void**)&iota, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc((void**)&keys, sizeof(uint32_t) * m) != hipSuccess) {
    st = MBX_ENOMEM;
  }
  if (st == MBX_OK && nnz > 0) {
    row_of_nnz_kernel<<<blocks(nnz), kBlock, 0, s>>>(d_row_ptr, nrow, nnz, row, iota);
    int bits = 1;
    while (bits < 31 && (1ll << bits) < (long long)ncol) ++bits;
    size_t tmp_bytes = 0;
    // column indices are non-negative: sort them as unsigned keys (stable)
    const uint32_t* ukeys = reinterpret_cast<const uint32_t*>(d_col);
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, ukeys, keys, iota, a->perm, nnz,
                                           0, bits, s) != hipSuccess ||
        hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, ukeys, keys, iota, a->perm, nnz, 0,
                                           bits, s) != hipSuccess)
      st = MBX_EHIP;
    if (st == MBX_OK) {
      col_ptr_kernel<<<blocks((long long)nnz + 1), kBlock, 0, s>>>(keys, nnz, ncol, a->col_ptr);
      transpose_rows_kernel<<<blocks(nnz), kBlock, 0, s>>>(a->perm, row, nnz, a->row_idx);
    }
  } else if (st == MBX_OK) {
    if (hipMemsetAsync(a->col_ptr, 0, sizeof(int32_t) * ((size_t)ncol + 1), s) != hipSuccess)
      st = MBX_EHIP;
  }
  if (st == MBX_OK && hipGetLastError() != hipSuccess) st = MBX_EHIP;
  if (st == MBX_OK)
    st = mbx_plan_create(a->col_ptr, ncol, nrow, nnz, items_per_thread, 256, stream, &a->plan_t);
  // the temporaries die here: wait for the kernels that use them
  if (hipStreamSynchronize(s) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  if (row) (void)hipFree(row);
  if (iota) (void)hipFree(iota);
  if (keys) (void)hipFree(keys);
  if (tmp) (void)hipFree(tmp);
  if (st != MBX_OK) {
    mbx_adjoint_destroy(a);
    return st;
  }
  *out = a;
  return MBX_OK;
}

void mbx_adjoint_destroy(mbx_adjoint* a) {
  if (!a) return;
  if (a->col_ptr) (void)hipFree(a->col_ptr);
  if (a->row_idx) (void)hipFree(a->row_idx);
  if (a->perm) (void)hipFree(a->perm);
  if (a->vt) (void)hipFree(a->vt);
  mbx_plan_destroy(a->plan_t);
  delete a;
}

int mbx_adjoint_transpose(const mbx_adjoint* a, int32_t* h_col_ptr, int32_t* h_row_idx,
                          int32_t* h_perm) {
  if (!a) return MBX_EINVAL;
  const size_t m = (size_t)a->nnz;
  if ((h_col_ptr && hipMemcpy(h_col_ptr, a->col_ptr, sizeof(int32_t) * ((size_t)a->ncol + 1),
                              hipMemcpyDeviceToHost) != hipSuccess) ||
      (h_row_idx && m && hipMemcpy(h_row_idx, a->row_idx, sizeof(int32_t) * m,
                                   hipMemcpyDeviceToHost) != hipSuccess) ||
      (h_perm && m && hipMemcpy(h_perm, a->perm, sizeof(int32_t) * m, hipMemcpyDeviceToHost) !=
                          hipSuccess))
    return MBX_EHIP;
  return MBX_OK;
}

int mbx_adjoint_dx(const mbx_adjoint* a, int32_t value_kind, const void* d_val, double alpha,
                   const void* d_g, double beta, void* d_dx, void* stream) {
  if (!a || (a->nnz > 0 && (!d_val || !d_g)) || (a->ncol > 0 && !d_dx)) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int nnz = a->nnz;
  switch (value_kind) {
    case MBX_VALUE_F64:
      if (nnz > 0)
        gather_kernel<double><<<blocks(nnz), kBlock, 0, s>>>((const double*)d_val, a->perm, nnz,
                                                             (double*)a->vt);
      return mbx_spmv_f64(a->plan_t, a->col_ptr, a->row_idx, (const double*)a->vt, alpha,
                          (const double*)d_g, beta, (double*)d_dx, stream);
    case MBX_VALUE_F32:
      if (nnz > 0)
        gather_kernel<float><<<blocks(nnz), kBlock, 0, s>>>((const float*)d_val, a->perm, nnz,
                                                           (float*)a->vt);
      return mbx_spmv_f32(a->plan_t, a->col_ptr, a->row_idx, (const float*)a->vt, (float)alpha,
                          (const float*)d_g, (float)beta, (float*)d_dx, stream);
    case MBX_VALUE_BF16:
      if (nnz > 0)
        gather_kernel<uint16_t><<<blocks(nnz), kBlock, 0, s>>>((const uint16_t*)d_val, a->perm,
                                                              nnz, (uint16_t*)a->vt);
      return mbx_spmv_bf16(a->plan_t, a->col_ptr, a->row_idx, (const uint16_t*)a->vt,
                           (float)alpha, (const float*)d_g, (float)beta, (float*)d_dx, stream);
    default:
      return MBX_EINVAL;
  }
}

int mbx_sddmm(const mbx_plan* p, int32_t value_kind, const int32_t* d_row_ptr,
              const int32_t* d_col, const void* d_g, const void* d_x, void* d_dval,
              void* stream) {
  if (!mbx_detail::plan_csr_args_ok(p, d_row_ptr, d_col, d_dval)) return MBX_EINVAL;
  if (p->nnz > 0 && (!d_g || !d_x)) return MBX_EINVAL;
  if (value_kind != MBX_VALUE_F64 && value_kind != MBX_VALUE_F32) return MBX_EINVAL;
  if (p->wg != 256) return MBX_EUNSUPPORTED;
  if (p->ntiles == 0) return MBX_OK;
  hipStream_t s = (hipStream_t)stream;
  const bool f64 = value_kind == MBX_VALUE_F64;
  switch (p->ipt) {
#define MBX_CASE(I)                                                                         \
  case I:                                                                                   \
    if (f64)                                                                                \
      sddmm_tile_kernel<256, I, double><<<p->ntiles, 256, 0, s>>>(                          \
          d_row_ptr, d_col, (const double*)d_g, (const double*)d_x, (double*)d_dval,        \
          p->d_coord, p->ntiles);                                                           \
    else                                                                                    \
      sddmm_tile_kernel<256, I, float><<<p->ntiles, 256, 0, s>>>(                           \
          d_row_ptr, d_col, (const float*)d_g, (const float*)d_x, (float*)d_dval,           \
          p->d_coord, p->ntiles);                                                           \
    break;
    MBX_CASE(2) MBX_CASE(4) MBX_CASE(6) MBX_CASE(8) MBX_CASE(12) MBX_CASE(16)
#undef MBX_CASE
    default:
      return MBX_EUNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

}  // extern "C"

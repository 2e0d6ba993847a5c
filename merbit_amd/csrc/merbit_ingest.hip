// merbit_ingest.hip -- the reference's input preparation on the device: from a
// loaded COO graph to the CSR of A^T with values 1/out-degree(row), the matrix
// its SpMV driver operates on.  Replaces Fun_Values (src/class.cu:9-34),
// GRAPH::sort(false) (src/class.cu:143-177, a Thrust sort by (col, row)) and
// cusparseXcoo2csr (include/tile.cuh:571-574).
//
// Out-degrees are an integer histogram (LDS-free global atomics on counts:
// exact and order independent).  The (col, row) order comes from one radix
// sort of 64-bit keys col<<32 | row; equal keys are identical entries, so the
// result does not depend on the sort's treatment of ties.  Row pointers of
// A^T are the positions where the sorted column changes.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>
#include <new>

#include "merbit_amd.h"

namespace {

constexpr int kBlock = 256;
inline int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// Out-degree histogram and sort keys; flags entries outside [0, n).
__global__ void ingest_keys_kernel(const int32_t* __restrict__ row_ind,
                                   const int32_t* __restrict__ col_ind, int nnz, int n,
                                   int32_t* __restrict__ deg, uint64_t* __restrict__ keys,
                                   int32_t* __restrict__ bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int r = row_ind[k], c = col_ind[k];
  if (r < 0 || r >= n || c < 0 || c >= n) {
    atomicOr(bad, 1);
    keys[k] = 0;
    return;
  }
  atomicAdd(deg + r, 1);
  keys[k] = ((uint64_t)(uint32_t)c << 32) | (uint32_t)r;
}

// Entry p of A^T: column index = the original row, value = 1/out-degree of
// that row (computed in FP64, stored as V, as the reference's double literal
// 1.0 / count does); row pointers where the original column changes.
template <class V>
__global__ void ingest_emit_kernel(const uint64_t* __restrict__ keys, int nnz, int n,
                                   const int32_t* __restrict__ deg, int32_t* __restrict__ row_ptr_t,
                                   int32_t* __restrict__ col_t, V* __restrict__ val_t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > nnz) return;
  const int lo = (p == 0) ? -1 : (int)(keys[p - 1] >> 32);
  const int hi = (p == nnz) ? n : (int)(keys[p] >> 32);
  for (int c = lo + 1; c <= hi; ++c) row_ptr_t[c] = p;
  if (p < nnz) {
    const int r = (int)(uint32_t)keys[p];
    col_t[p] = r;
    val_t[p] = (V)(1.0 / (double)deg[r]);
  }
}

}  // namespace

extern "C" {

int mbx_ingest_ref(int32_t n, int32_t nnz, const int32_t* d_row_ind, const int32_t* d_col_ind,
                   int32_t value_kind, int32_t* d_row_ptr_t, int32_t* d_col_t, void* d_val_t,
                   void* stream) {
  if (n < 0 || nnz < 0 || !d_row_ptr_t) return MBX_EINVAL;
  if (nnz > 0 && (!d_row_ind || !d_col_ind || !d_col_t || !d_val_t)) return MBX_EINVAL;
  if (value_kind != MBX_VALUE_F64 && value_kind != MBX_VALUE_F32) return MBX_EINVAL;
  if (nnz > 0 && n == 0) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const size_t m = (size_t)(nnz > 0 ? nnz : 1);
  int32_t *deg = nullptr, *bad = nullptr;
  uint64_t *keys = nullptr, *sorted = nullptr;
  void* tmp = nullptr;
  int st = MBX_OK;
  if (hipMalloc((void**)&deg, sizeof(int32_t) * ((size_t)n + 1)) != hipSuccess ||
      hipMalloc((void**)&bad, sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&keys, sizeof(uint64_t) * m) != hipSuccess ||
      hipMalloc((void**)&sorted, sizeof(uint64_t) * m) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK && (hipMemsetAsync(deg, 0, sizeof(int32_t) * ((size_t)n + 1), s) != hipSuccess ||
                       hipMemsetAsync(bad, 0, sizeof(int32_t), s) != hipSuccess))
    st = MBX_EHIP;
  if (st == MBX_OK && nnz > 0) {
    ingest_keys_kernel<<<blocks(nnz), kBlock, 0, s>>>(d_row_ind, d_col_ind, nnz, n, deg, keys,
                                                      bad);
    int bits = 1;
    while (bits < 31 && (1ll << bits) < (long long)n) ++bits;
    size_t tmp_bytes = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys, sorted, nnz, 0, 32 + bits,
                                          s) != hipSuccess ||
        hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1) != hipSuccess ||
        hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, keys, sorted, nnz, 0, 32 + bits, s) !=
            hipSuccess)
      st = MBX_EHIP;
  }
  if (st == MBX_OK) {
    if (value_kind == MBX_VALUE_F64)
      ingest_emit_kernel<double><<<blocks((long long)nnz + 1), kBlock, 0, s>>>(
          sorted, nnz, n, deg, d_row_ptr_t, d_col_t, (double*)d_val_t);
    else
      ingest_emit_kernel<float><<<blocks((long long)nnz + 1), kBlock, 0, s>>>(
          sorted, nnz, n, deg, d_row_ptr_t, d_col_t, (float*)d_val_t);
    if (hipGetLastError() != hipSuccess) st = MBX_EHIP;
  }
  int32_t h_bad = 0;
  if (st == MBX_OK &&
      (hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    st = MBX_EHIP;
  else if (st != MBX_OK)
    (void)hipStreamSynchronize(s);  // the temporaries die here
  if (st == MBX_OK && h_bad) st = MBX_EINVAL;
  if (deg) (void)hipFree(deg);
  if (bad) (void)hipFree(bad);
  if (keys) (void)hipFree(keys);
  if (sorted) (void)hipFree(sorted);
  if (tmp) (void)hipFree(tmp);
  return st;
}

}  // extern "C"

// merbit_validate.hip -- structural check of a device-resident CSR pattern
// (mbx_csr_validate).  The products read x through a bounded buffer resource,
// so a column index outside [0, ncol) reads 0 there instead of faulting; this
// entry point reports such input instead of letting it pass silently, and the
// adjoint (whose transposition indexes by column) refuses it through the same
// check.  The reference trusts its input (GRAPH builds the CSR itself,
// src/class.cu:143-177); there is no counterpart to cite.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <new>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

namespace {

constexpr int kBlock = 256;

// Counters, device side: [0] row_ptr errors, [1] column errors, [2] unsorted
// pairs, [3] first bad column position (min), [4] row_ptr[nrow].
struct Counters {
  unsigned long long rp_err, col_err, unsorted;
  int first_bad, last_rp;
};

__device__ __forceinline__ void wave_count(bool bad, unsigned long long* dst) {
  const unsigned long long m = __ballot(bad);
  if (m && (threadIdx.x & 63) == __ffsll(m) - 1) atomicAdd(dst, (unsigned long long)__popcll(m));
}

// One thread per row r < nrow: row_ptr[r] <= row_ptr[r+1], both in [0, nnz],
// row_ptr[0] == 0; marks the first entry of every non-empty row (starts has
// nnz + 1 bytes, zeroed) for the sortedness test.
__global__ void csr_rows_kernel(const int32_t* __restrict__ rp, int nrow, int nnz,
                                uint8_t* __restrict__ starts, Counters* __restrict__ c) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  bool bad = false;
  if (r < nrow) {
    const int a = rp[r], b = rp[r + 1];
    bad = a > b || a < 0 || b > nnz || (r == 0 && a != 0);
    if (!bad && a < b) starts[a] = 1;
    if (r == nrow - 1) c->last_rp = b;
  }
  wave_count(bad, &c->rp_err);
}

// One thread per entry k < nnz: column in [0, ncol); a column below its
// predecessor's in the same row counts as an unsorted pair.
__global__ void csr_cols_kernel(const int32_t* __restrict__ col, int nnz, int ncol,
                                const uint8_t* __restrict__ starts, Counters* __restrict__ c) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  bool bad = false, desc = false;
  if (k < nnz) {
    const int v = col[k];
    bad = v < 0 || v >= ncol;
    if (bad) atomicMin(&c->first_bad, k);
    desc = k > 0 && !starts[k] && v < col[k - 1];
  }
  wave_count(bad, &c->col_err);
  wave_count(desc, &c->unsorted);
}

}  // namespace

namespace mbx_detail {

int csr_check(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow, int32_t ncol,
              int32_t nnz, hipStream_t s, mbx_csr_report* out) {
  mbx_csr_report rep = {};
  rep.first_bad_col = -1;
  if (nrow < 0 || ncol < 0 || nnz < 0 || (nrow > 0 && !d_row_ptr) || (nnz > 0 && !d_col))
    return MBX_EINVAL;
  if (nrow == 0) {
    rep.nnz_mismatch = nnz != 0;
    if (out) *out = rep;
    return rep.nnz_mismatch ? MBX_EINVAL : MBX_OK;
  }
  Counters* d_c = nullptr;
  uint8_t* starts = nullptr;
  if (hipMalloc((void**)&d_c, sizeof(Counters)) != hipSuccess ||
      hipMalloc((void**)&starts, (size_t)nnz + 1) != hipSuccess) {
    (void)hipGetLastError();
    if (d_c) (void)hipFree(d_c);
    return MBX_ENOMEM;
  }
  Counters h = {0, 0, 0, INT32_MAX, 0};
  int st = MBX_OK;
  if (hipMemcpyAsync(d_c, &h, sizeof(Counters), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(starts, 0, (size_t)nnz + 1, s) != hipSuccess)
    st = MBX_EHIP;
  if (st == MBX_OK) {
    csr_rows_kernel<<<(nrow + kBlock - 1) / kBlock, kBlock, 0, s>>>(d_row_ptr, nrow, nnz, starts,
                                                                      d_c);
    if (nnz > 0)
      csr_cols_kernel<<<(int)(((long long)nnz + kBlock - 1) / kBlock), kBlock, 0, s>>>(
          d_col, nnz, ncol, starts, d_c);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&h, d_c, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = MBX_EHIP;
  }
  (void)hipFree(d_c);
  (void)hipFree(starts);
  if (st != MBX_OK) return st;
  rep.row_ptr_errors = (int64_t)h.rp_err;
  rep.col_errors = (int64_t)h.col_err;
  rep.unsorted_pairs = (int64_t)h.unsorted;
  rep.nnz_mismatch = h.last_rp != nnz;
  rep.first_bad_col = h.col_err ? h.first_bad : -1;
  if (out) *out = rep;
  return (rep.row_ptr_errors || rep.col_errors || rep.nnz_mismatch) ? MBX_EINVAL : MBX_OK;
}

}  // namespace mbx_detail

extern "C" int mbx_csr_validate(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                                int32_t ncol, int32_t nnz, void* stream, mbx_csr_report* out) {
  return mbx_detail::csr_check(d_row_ptr, d_col, nrow, ncol, nnz, (hipStream_t)stream, out);
}

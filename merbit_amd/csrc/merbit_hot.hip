// merbit_hot.hip -- hot-column split of a plan (mbx_plan_split_hot).
//
// On power-law patterns a few columns take a large share of the nonzeros
// (the 2^20-row R-MAT: the 12288 most frequent columns hold 54 % of them).
// Every gather of x goes through the vector L1, whose outstanding-miss
// slots bound the product; hot columns repeat across every tile yet their
// x lines are spread over all of x.  The split ranks columns by nonzero
// count and keeps, in the plan,
//   hot[h]   the h-th most frequent column (h < nhot, ties by column index),
//   colh[k]  the column array with every hot entry replaced by -1 - h,
// so that a product gathers x[hot[.]] once into a compact array (xh), a
// persistent tile kernel stages xh into LDS once per workgroup, and each
// nonzero reads x from LDS (hot) or through the bounded buffer resource
// (cold; a negative index reads 0 there without a memory access).  The
// products and their summation order are unchanged, so results are
// bit-identical to the plain product.  No reference counterpart: the
// reference gathers x[col] for every nonzero (include/tile.cuh:293-305).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <new>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

namespace {

constexpr int kBlock = 256;

inline int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// Column counts by sorting: the column indices (those outside [0, ncol)
// mapped to the key ncol) are radix-sorted on the bits of ncol, the sorted
// keys run-length encoded, and each run's length stored at its column.  No
// atomics: one atomic increment per nonzero took 2.2 ms on the 2^20-row
// R-MAT, bound by the atomic rate and, on its hottest columns, serialised
// on one counter.
__global__ void col_keys_kernel(const int32_t* __restrict__ col, int nnz, int ncol,
                                int32_t* __restrict__ keys) {
  const long long k = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (k < nnz) {
    const int c = col[k];
    keys[k] = (c >= 0 && c < ncol) ? c : ncol;
  }
}

__global__ void run_store_kernel(const int32_t* __restrict__ uniq,
                                 const int32_t* __restrict__ len, const int32_t* __restrict__ nruns,
                                 int max_runs, int ncol, int32_t* __restrict__ cnt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < max_runs && i < *nruns && uniq[i] < ncol) cnt[uniq[i]] = len[i];
}

__global__ void iota_kernel(int n, int32_t* __restrict__ iota) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < n) iota[k] = k;
}

// The nhot most frequent columns: slots, the hot list and the entries they
// cover (summed in 64 bits).
__global__ void hot_pick_kernel(const int32_t* __restrict__ sorted_col,
                                const int32_t* __restrict__ sorted_cnt, int nhot,
                                int32_t* __restrict__ slot, int32_t* __restrict__ hot,
                                unsigned long long* __restrict__ covered) {
  const int h = blockIdx.x * kBlock + threadIdx.x;
  if (h >= nhot) return;
  const int c = sorted_col[h];
  slot[c] = h;
  hot[h] = c;
  atomicAdd(covered, (unsigned long long)sorted_cnt[h]);
}

// colh[k] = -1 - slot of a hot column, the column itself otherwise; a
// negative (invalid) column becomes ncol, which the bounded gathers read as
// 0 as the plain product reads it, rather than a hot slot; the entries past
// nnz up to the next multiple of 4 (read by the 16-byte column groups of the
// staging) are 0.
__global__ void colh_kernel(const int32_t* __restrict__ col, int nnz, int padded, int ncol,
                            const int32_t* __restrict__ slot, int32_t* __restrict__ colh) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= padded) return;
  if (k >= nnz) {
    colh[k] = 0;
    return;
  }
  const int c = col[k];
  const int s = (c >= 0 && c < ncol) ? slot[c] : -1;
  colh[k] = s >= 0 ? -1 - s : (c >= 0 ? c : ncol);
}

void drop_hot(mbx_plan* p) {
  if (p->d_colh || p->d_hot || p->d_xh) (void)hipDeviceSynchronize();  // products may read them
  if (p->d_colh) (void)hipFree(p->d_colh);
  if (p->d_hot) (void)hipFree(p->d_hot);
  if (p->d_xh) (void)hipFree(p->d_xh);
  p->d_colh = nullptr;
  p->d_hot = nullptr;
  p->d_xh = nullptr;
  p->nhot = 0;
}

}  // namespace

namespace mbx_detail {
void plan_drop_hot(mbx_plan* p) { drop_hot(p); }

int count_columns(const int32_t* d_col, int nnz, int ncol, int32_t* cnt, hipStream_t s) {
  if (ncol <= 0) return MBX_OK;
  if (hipMemsetAsync(cnt, 0, 4 * (size_t)ncol, s) != hipSuccess) return MBX_EHIP;
  if (nnz <= 0) return MBX_OK;
  const int end_bit = 32 - __builtin_clz((unsigned)ncol);  // keys 0 .. ncol
  const int max_runs = (int)(nnz < (long long)ncol + 1 ? nnz : (long long)ncol + 1);
  int32_t *k0 = nullptr, *k1 = nullptr, *uniq = nullptr, *len = nullptr, *nruns = nullptr;
  void* tmp = nullptr;
  size_t sort_bytes = 0, rle_bytes = 0;
  int st = MBX_OK;
  hipcub::DoubleBuffer<int32_t> keys;
  if (hipMallocAsync((void**)&k0, 4 * (size_t)nnz, s) != hipSuccess ||
      hipMallocAsync((void**)&k1, 4 * (size_t)nnz, s) != hipSuccess ||
      hipMallocAsync((void**)&uniq, 4 * (size_t)max_runs, s) != hipSuccess ||
      hipMallocAsync((void**)&len, 4 * (size_t)max_runs, s) != hipSuccess ||
      hipMallocAsync((void**)&nruns, sizeof(int32_t), s) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK) {
    keys = hipcub::DoubleBuffer<int32_t>(k0, k1);
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, keys, nnz, 0, end_bit, s) !=
            hipSuccess ||
        hipcub::DeviceRunLengthEncode::Encode(nullptr, rle_bytes, k0, uniq, len, nruns, nnz,
                                              s) != hipSuccess ||
        hipMallocAsync(&tmp, sort_bytes > rle_bytes ? sort_bytes : rle_bytes, s) != hipSuccess)
      st = MBX_ENOMEM;
  }
  if (st == MBX_OK) {
    col_keys_kernel<<<blocks(nnz), kBlock, 0, s>>>(d_col, nnz, ncol, k0);
    if (hipGetLastError() != hipSuccess ||
        hipcub::DeviceRadixSort::SortKeys(tmp, sort_bytes, keys, nnz, 0, end_bit, s) !=
            hipSuccess ||
        hipcub::DeviceRunLengthEncode::Encode(tmp, rle_bytes, keys.Current(), uniq, len, nruns,
                                              nnz, s) != hipSuccess)
      st = MBX_EHIP;
  }
  if (st == MBX_OK) {
    run_store_kernel<<<blocks(max_runs), kBlock, 0, s>>>(uniq, len, nruns, max_runs, ncol, cnt);
    if (hipGetLastError() != hipSuccess) st = MBX_EHIP;
  }
  // stream-ordered: released once the stream has used them
  for (void* q : {(void*)k0, (void*)k1, (void*)uniq, (void*)len, (void*)nruns, tmp})
    if (q) (void)hipFreeAsync(q, s);
  if (st != MBX_OK) (void)hipGetLastError();
  return st;
}
}  // namespace mbx_detail

namespace mbx_detail {

int split_hot_counted(mbx_plan* p, const int32_t* d_col, int32_t max_hot,
                      const int32_t* cnt_given, hipStream_t s, double* covered) {
  if (!p || max_hot < 0) return MBX_EINVAL;
  if (covered) *covered = 0.0;
  drop_hot(p);
  if (!d_col || max_hot == 0 || p->nnz == 0 || p->ncol == 0) return MBX_OK;
  if (max_hot > kMaxHotColumns) return MBX_EUNSUPPORTED;
  if (p->d_scol) return MBX_EUNSUPPORTED;  // the sorted columns exclude the split
  const int ncol = p->ncol, nnz = p->nnz;
  const int nhot = max_hot < ncol ? max_hot : ncol;
  const int padded = (nnz + 3) / 4 * 4;
  int32_t *cnt = nullptr, *iota = nullptr, *scnt = nullptr, *scol = nullptr, *slot = nullptr;
  unsigned long long* d_cov = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int st = MBX_OK;
  if ((!cnt_given && hipMallocAsync((void**)&cnt, 4 * (size_t)ncol, s) != hipSuccess) ||
      hipMallocAsync((void**)&iota, 4 * (size_t)ncol, s) != hipSuccess ||
      hipMallocAsync((void**)&scnt, 4 * (size_t)ncol, s) != hipSuccess ||
      hipMallocAsync((void**)&scol, 4 * (size_t)ncol, s) != hipSuccess ||
      hipMallocAsync((void**)&slot, 4 * (size_t)ncol, s) != hipSuccess ||
      hipMallocAsync((void**)&d_cov, sizeof(unsigned long long), s) != hipSuccess ||
      hipMalloc((void**)&p->d_colh, 4 * (size_t)padded) != hipSuccess ||
      hipMalloc((void**)&p->d_hot, 4 * (size_t)nhot) != hipSuccess ||
      hipMalloc(&p->d_xh, ((size_t)nhot * 8 + 15) / 16 * 16) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK &&
      (hipMemsetAsync(slot, 0xFF, 4 * (size_t)ncol, s) != hipSuccess ||
       hipMemsetAsync(d_cov, 0, sizeof(unsigned long long), s) != hipSuccess ||
       hipMemsetAsync(p->d_xh, 0, ((size_t)nhot * 8 + 15) / 16 * 16, s) != hipSuccess))
    st = MBX_EHIP;
  if (st == MBX_OK) {
    iota_kernel<<<blocks(ncol), kBlock, 0, s>>>(ncol, iota);
    if (!cnt_given) st = count_columns(d_col, nnz, ncol, cnt, s);
  }
  const int32_t* counts = cnt_given ? cnt_given : cnt;
  if (st == MBX_OK) {
    // most frequent first; equal counts keep ascending column order (the
    // sort is stable); counts are at most nnz, so cbits bits order them
    const int cbits = 32 - __builtin_clz((unsigned)nnz);
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp_bytes, counts, scnt, iota, scol,
                                                     ncol, 0, cbits, s) != hipSuccess ||
        hipMallocAsync(&tmp, tmp_bytes ? tmp_bytes : 1, s) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairsDescending(tmp, tmp_bytes, counts, scnt, iota, scol, ncol,
                                                     0, cbits, s) != hipSuccess)
      st = MBX_EHIP;
  }
  unsigned long long h_cov = 0;
  if (st == MBX_OK) {
    hot_pick_kernel<<<blocks(nhot), kBlock, 0, s>>>(scol, scnt, nhot, slot, p->d_hot, d_cov);
    colh_kernel<<<blocks(padded), kBlock, 0, s>>>(d_col, nnz, padded, ncol, slot, p->d_colh);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&h_cov, d_cov, sizeof(h_cov), hipMemcpyDeviceToHost, s) != hipSuccess)
      st = MBX_EHIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  for (void* q : {(void*)cnt, (void*)iota, (void*)scnt, (void*)scol, (void*)slot,
                  (void*)d_cov, tmp})
    if (q) (void)hipFreeAsync(q, s);
  if (st != MBX_OK) {
    (void)hipGetLastError();
    drop_hot(p);
    return st;
  }
  p->nhot = nhot;
  if (covered) *covered = (double)h_cov / (double)nnz;
  return MBX_OK;
}

}  // namespace mbx_detail

extern "C" int mbx_plan_split_hot(mbx_plan* p, const int32_t* d_col, int32_t max_hot,
                                  void* stream, double* covered) {
  return mbx_detail::split_hot_counted(p, d_col, max_hot, nullptr, (hipStream_t)stream,
                                       covered);
}

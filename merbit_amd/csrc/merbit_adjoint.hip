// merbit_adjoint.hip -- the adjoint of y = A*x with respect to x and to the
// values of A (no reference counterpart: the reference computes y only).
//
//   dx   = A^T g         by the merge-path tile SpMV on the transposed pattern
//                        (CSC of A, built once per pattern on the device with a
//                        stable radix sort, so every column keeps its rows in
//                        ascending order and the sums are deterministic); the
//                        values are moved into CSC order per call by a
//                        two-pass bucketed permutation (below).
//
// Bucketed permutation.  Moving nnz values from CSR to CSC order is a random
// permutation; done as one gather it reads a whole memory line per value from
// an array larger than the caches.  Instead, CSC positions are cut into
// buckets of kBucket consecutive positions, and CSR positions into chunks of
// kChunk.  Pass 1 (one workgroup per chunk) reads the chunk's values
// coalesced, regroups them in LDS by bucket and writes each bucket's run
// contiguously into a staging array laid out bucket after bucket; pass 2
// (one workgroup per bucket) loads the bucket's staging slice into LDS and
// writes the bucket's CSC positions in order, each value read from LDS.
// Every global access is coalesced or a short contiguous run.  The schedule
// (LDS slot per CSR entry, staging index per pass-1 slot, staging index per
// CSC position) is computed once at create with two stable radix sorts; the
// values moved are bit-for-bit those of the plain gather val[perm[p]].
//   dval = g[row]*x[col] (SDDMM over the pattern) walking the forward plan's
//                        merge-path tiles, which give every nonzero its row.

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>
#include <new>
#include <vector>

#include "merbit_amd.h"
#include "merbit_internal.hpp"
#include "merbit_tile.hpp"

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

// Both sizes are limited by LDS: a chunk's or a bucket's values must fit a
// workgroup's LDS at 8 bytes each (128 KB of the 160 KB per CU).
constexpr int kChunk = 16384;   // CSR entries per pass-1 workgroup (LDS slots, < 65536)
constexpr int kBucket = 16384;  // CSC positions per bucket
constexpr int kPassWg = 256;

// Four consecutive elements as one vector access (16 bytes for FP32, 8 for
// bfloat16, two 16-byte halves for FP64); the pointer must be aligned to the
// group size.
template <class T>
using vec4_t = T __attribute__((ext_vector_type(4)));

// Pass 1: the chunk's values into LDS at their bucket-grouped slots, then
// out to the staging array, slot order (consecutive lanes write consecutive
// staging entries inside each bucket run).  Consecutive chunks run on one
// XCD, so the runs they add to the same staging lines meet in one L2.  VEC:
// values and slots read four at a time (val aligned to 4 elements).
template <class T, bool VEC>
__global__ __launch_bounds__(kPassWg) void perm_pass1_kernel(const T* __restrict__ val,
                                                             const uint16_t* __restrict__ lslot,
                                                             const int32_t* __restrict__ spos,
                                                             int nnz, int nsrc, int nchunks,
                                                             T* __restrict__ stage) {
  extern __shared__ __align__(16) char smem[];
  T* s_v = reinterpret_cast<T*>(smem);  // kChunk values
  const long long base = (long long)xcd_tile(blockIdx.x, nchunks) * kChunk;
  const int n = (int)min((long long)kChunk, (long long)nnz - base);
  // sources at or past nsrc are virtual (value 0): val holds nsrc values
  const int nv = (int)max(0ll, min((long long)n, (long long)nsrc - base));
  const int n4 = VEC ? (nv & ~3) : 0;
  if (VEC) {
#pragma unroll 4
    for (int j = 4 * threadIdx.x; j < n4; j += 4 * kPassWg) {
      const vec4_t<T> v =
          __builtin_nontemporal_load(reinterpret_cast<const vec4_t<T>*>(val + base + j));
      const vec4_t<uint16_t> l =
          __builtin_nontemporal_load(reinterpret_cast<const vec4_t<uint16_t>*>(lslot + base + j));
      s_v[l.x] = v.x;
      s_v[l.y] = v.y;
      s_v[l.z] = v.z;
      s_v[l.w] = v.w;
    }
  }
  for (int j = n4 + threadIdx.x; j < n; j += kPassWg)
    s_v[lslot[base + j]] = j < nv ? val[base + j] : T(0);
  __syncthreads();
  // one slot per lane: a store instruction covers 64 consecutive slots, so
  // it touches as few bucket runs (staging lines) as possible
#pragma unroll 8
  for (int j = threadIdx.x; j < n; j += kPassWg)
    stage[__builtin_nontemporal_load(spos + base + j)] = s_v[j];
}

// Pass 2: one workgroup per bucket loads the bucket's staging slice into LDS
// (coalesced, four values at a time), then writes the bucket's CSC positions
// in order, each value read from LDS at its staging index.  Buckets start at
// multiples of kBucket in library-allocated arrays, so every group of four
// is aligned.
template <class T>
__global__ __launch_bounds__(kPassWg) void perm_pass2_kernel(const T* __restrict__ stage,
                                                             const int32_t* __restrict__ lidx,
                                                             int nnz, T* __restrict__ out) {
  extern __shared__ __align__(16) char smem[];
  T* s_v = reinterpret_cast<T*>(smem);
  const long long base = (long long)blockIdx.x * kBucket;
  const int n = (int)min((long long)kBucket, (long long)nnz - base);
  const int n4 = n & ~3;
#pragma unroll 4
  for (int j = 4 * threadIdx.x; j < n4; j += 4 * kPassWg)
    *reinterpret_cast<vec4_t<T>*>(s_v + j) =
        *reinterpret_cast<const vec4_t<T>*>(stage + base + j);
  for (int j = n4 + threadIdx.x; j < n; j += kPassWg) s_v[j] = stage[base + j];
  __syncthreads();
  const int b = (int)base;  // staging indices of this bucket are base + [0, n)
#pragma unroll 4
  for (int j = 4 * threadIdx.x; j < n4; j += 4 * kPassWg) {
    const vec4_t<int32_t> i =
        __builtin_nontemporal_load(reinterpret_cast<const vec4_t<int32_t>*>(lidx + base + j));
    vec4_t<T> v;
    v.x = s_v[i.x - b];
    v.y = s_v[i.y - b];
    v.z = s_v[i.z - b];
    v.w = s_v[i.w - b];
    *reinterpret_cast<vec4_t<T>*>(out + base + j) = v;
  }
  for (int j = n4 + threadIdx.x; j < n; j += kPassWg) out[base + j] = s_v[lidx[base + j] - b];
}

// inv[perm[p]] = p: the CSC position of every CSR entry.
__global__ void perm_inv_kernel(const int32_t* __restrict__ perm, int nnz,
                                int32_t* __restrict__ inv) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nnz) inv[perm[p]] = p;
}

// Sort keys of CSR entry k: pass-1 order is chunk(k) above bucket(inv[k]);
// staging order is bucket(inv[k]) (both sorts are stable, so k breaks ties).
__global__ void perm_key1_kernel(const int32_t* __restrict__ inv, int nnz, int bucket_bits,
                                 uint32_t* __restrict__ key1, uint32_t* __restrict__ key2,
                                 int32_t* __restrict__ iota) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t b = (uint32_t)(inv[k] / kBucket);
  key1[k] = ((uint32_t)(k / kChunk) << bucket_bits) | b;
  key2[k] = b;
  iota[k] = k;
}

// lslot[k] = slot of CSR entry k in its chunk, from the pass-1 order.
__global__ void perm_lslot_kernel(const int32_t* __restrict__ order1, int nnz,
                                  uint16_t* __restrict__ lslot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) lslot[order1[i]] = (uint16_t)(i % kChunk);
}

// sidx[k] = staging index of CSR entry k (its rank in (bucket, k) order).
__global__ void perm_sidx_kernel(const int32_t* __restrict__ order2, int nnz,
                                 int32_t* __restrict__ sidx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nnz) sidx[order2[r]] = r;
}

// spos[i] = sidx[order1[i]] and lidx[p] = sidx[perm[p]].
__global__ void perm_final_kernel(const int32_t* __restrict__ order1,
                                  const int32_t* __restrict__ perm,
                                  const int32_t* __restrict__ sidx, int nnz,
                                  int32_t* __restrict__ spos, int32_t* __restrict__ lidx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  spos[i] = sidx[order1[i]];
  lidx[i] = sidx[perm[i]];
}

// dval[k] = g[row(k)] * x[col(k)] for the nonzeros of one forward tile.  x
// is gathered through a buffer resource over its ncol entries when it spans
// at most 4 GiB (x_rsrc), so a column index outside [0, ncol) reads 0, as in
// the forward tile kernel, instead of faulting.  RSRC (x within the buffer
// resource's 4 GiB) is a template parameter: with both gather paths in one
// kernel the FP64 tile at 8 items per thread took 118 VGPRs (4 waves per
// SIMD; 252 at 12 items), with one path 41 (62), and the SDDMM 13 % less
// time on the R-MAT.
template <int WG, int IPT, class T, bool RSRC>
__device__ __forceinline__ void sddmm_tile(int tile, int* s_re, int* s_row,
                                           const int32_t* __restrict__ row_ptr,
                                           const int32_t* __restrict__ col,
                                           const T* __restrict__ g, const T* __restrict__ x,
                                           T* __restrict__ dval, const int4* __restrict__ coord,
                                           int ncol) {
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
      // the pattern is read once per call and dval written once: non-temporal
      // accesses keep them from displacing the x and g lines the gathers reuse
      if (i < rows_t) re[j] = __builtin_nontemporal_load(row_ptr + r0 + 1 + i);
      if (i < nnz_t) ci[j] = __builtin_nontemporal_load(col + k0 + i);
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
  if constexpr (RSRC) {
    const __amdgpu_buffer_rsrc_t rx =
        mbx_detail::buffer_rsrc(x, (uint32_t)ncol * (uint32_t)sizeof(T));
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      xv[j] = T(0);
      gv[j] = T(0);
      if (i < nnz_t) {
        mbx_detail::buffer_gather(rx, ci[j], xv[j]);
        gv[j] = g[r0 + s_row[i]];
      }
    }
  } else {
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
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = tid + j * WG;
    if (i < nnz_t) __builtin_nontemporal_store(gv[j] * xv[j], dval + k0 + i);
  }
}

template <int WG, int IPT, class T, bool RSRC>
__global__ __launch_bounds__(WG) void sddmm_tile_kernel(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const T* __restrict__ g, const T* __restrict__ x, T* __restrict__ dval,
    const int4* __restrict__ coord, int ntiles, int ncol) {
  constexpr int TI = WG * IPT;
  __shared__ int s_re[TI + 1];
  __shared__ int s_row[TI];
  sddmm_tile<WG, IPT, T, RSRC>(xcd_tile(blockIdx.x, ntiles), s_re, s_row, row_ptr, col, g, x,
                               dval, coord, ncol);
}

// The plain SDDMM, one workgroup per tile, on the bounded or the direct
// gather path.
template <int IPT, class T>
void launch_sddmm(const mbx_plan* p, bool x_rsrc, const int32_t* rp, const int32_t* col,
                  const T* g, const T* x, T* dval, hipStream_t s) {
  if (x_rsrc)
    sddmm_tile_kernel<256, IPT, T, true><<<p->ntiles, 256, 0, s>>>(rp, col, g, x, dval,
                                                                   p->d_coord, p->ntiles, p->ncol);
  else
    sddmm_tile_kernel<256, IPT, T, false><<<p->ntiles, 256, 0, s>>>(
        rp, col, g, x, dval, p->d_coord, p->ntiles, p->ncol);
}

constexpr int kBlock = 256;

inline int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace


namespace {

// The bucketed-permutation schedule of an adjoint whose perm is built
// (stream-ordered; temporaries freed after a synchronisation).
int build_perm_schedule(mbx_adjoint* a, hipStream_t s) {
  const int nnz = a->nnz;
  if (nnz == 0) return MBX_OK;
  const size_t m = (size_t)nnz;
  int bucket_bits = 1, chunk_bits = 1;
  while (bucket_bits < 31 && (1ll << bucket_bits) * kBucket < (long long)nnz) ++bucket_bits;
  while (chunk_bits < 31 && (1ll << chunk_bits) * kChunk < (long long)nnz) ++chunk_bits;
  if (bucket_bits + chunk_bits > 32) return MBX_EOVERFLOW;
  int32_t *inv = nullptr, *iota = nullptr, *order1 = nullptr, *order2 = nullptr;
  uint32_t *key1 = nullptr, *key2 = nullptr, *skey = nullptr;
  void* tmp = nullptr;
  int st = MBX_OK;
  if (hipMalloc((void**)&inv, 4 * m) != hipSuccess || hipMalloc((void**)&iota, 4 * m) != hipSuccess ||
      hipMalloc((void**)&order1, 4 * m) != hipSuccess ||
      hipMalloc((void**)&order2, 4 * m) != hipSuccess ||
      hipMalloc((void**)&key1, 4 * m) != hipSuccess || hipMalloc((void**)&key2, 4 * m) != hipSuccess ||
      hipMalloc((void**)&skey, 4 * m) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK) {
    perm_inv_kernel<<<blocks(nnz), kBlock, 0, s>>>(a->perm, nnz, inv);
    perm_key1_kernel<<<blocks(nnz), kBlock, 0, s>>>(inv, nnz, bucket_bits, key1, key2, iota);
    size_t b1 = 0, b2 = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, b1, key1, skey, iota, order1, nnz, 0,
                                           bucket_bits + chunk_bits, s) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(nullptr, b2, key2, skey, iota, order2, nnz, 0,
                                           bucket_bits, s) != hipSuccess)
      st = MBX_EHIP;
    const size_t tb = b1 > b2 ? b1 : b2;
    if (st == MBX_OK && hipMalloc(&tmp, tb ? tb : 1) != hipSuccess) st = MBX_ENOMEM;
    if (st == MBX_OK &&
        (hipcub::DeviceRadixSort::SortPairs(tmp, b1, key1, skey, iota, order1, nnz, 0,
                                            bucket_bits + chunk_bits, s) != hipSuccess ||
         hipcub::DeviceRadixSort::SortPairs(tmp, b2, key2, skey, iota, order2, nnz, 0,
                                            bucket_bits, s) != hipSuccess))
      st = MBX_EHIP;
    if (st == MBX_OK) {
      int32_t* sidx = inv;  // inv is dead once the keys are built
      perm_lslot_kernel<<<blocks(nnz), kBlock, 0, s>>>(order1, nnz, a->lslot);
      perm_sidx_kernel<<<blocks(nnz), kBlock, 0, s>>>(order2, nnz, sidx);
      perm_final_kernel<<<blocks(nnz), kBlock, 0, s>>>(order1, a->perm, sidx, nnz, a->spos,
                                                       a->lidx);
      if (hipGetLastError() != hipSuccess) st = MBX_EHIP;
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  if (inv) (void)hipFree(inv);
  if (iota) (void)hipFree(iota);
  if (order1) (void)hipFree(order1);
  if (order2) (void)hipFree(order2);
  if (key1) (void)hipFree(key1);
  if (key2) (void)hipFree(key2);
  if (skey) (void)hipFree(skey);
  if (tmp) (void)hipFree(tmp);
  return st;
}

// Values of `val` (CSR order) into CSC order in vt, by the two passes.
template <class T>
void permute_values(const mbx_adjoint* a, const T* val, hipStream_t s) {
  const int nnz = a->nnz;
  const int g1 = (int)(((long long)nnz + kChunk - 1) / kChunk);
  const size_t lds1 = sizeof(T) * kChunk;
  if ((reinterpret_cast<uintptr_t>(val) & (4 * sizeof(T) - 1)) == 0) {
    (void)hipFuncSetAttribute((const void*)perm_pass1_kernel<T, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
    perm_pass1_kernel<T, true><<<g1, kPassWg, lds1, s>>>(val, a->lslot, a->spos, nnz, a->nsrc,
                                                          g1, (T*)a->stage);
  } else {
    (void)hipFuncSetAttribute((const void*)perm_pass1_kernel<T, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
    perm_pass1_kernel<T, false><<<g1, kPassWg, lds1, s>>>(val, a->lslot, a->spos, nnz,
                                                           a->nsrc, g1, (T*)a->stage);
  }
  const int g2 = (int)(((long long)nnz + kBucket - 1) / kBucket);
  const size_t lds2 = sizeof(T) * kBucket;
  (void)hipFuncSetAttribute((const void*)perm_pass2_kernel<T>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
  perm_pass2_kernel<T><<<g2, kPassWg, lds2, s>>>((const T*)a->stage, a->lidx, nnz, (T*)a->vt);
}

// Slot q of the blocked A^T holds CSC entry tperm[q], CSR entry perm[.];
// its padding slots (tperm -1) are given by tb_pad_kernel.
__global__ void tb_real_kernel(const int32_t* __restrict__ tperm, int n,
                               const int32_t* __restrict__ perm, int32_t* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && tperm[q] >= 0) out[q] = perm[tperm[q]];
}

// The i-th padding slot reads virtual source nnz + i.
__global__ void tb_pad_kernel(const int32_t* __restrict__ dst, int n, int nnz,
                              int32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[dst[i]] = nnz + i;
}

void drop_blocks(mbx_adjoint* a) {
  if (a->tb) mbx_colblock_destroy(a->tb);
  mbx_detail::perm_free(a->tbp);
  a->tb = nullptr;
  a->tbp = nullptr;
}

}  // namespace

namespace mbx_detail {

int perm_schedule_build(mbx_adjoint* a, hipStream_t s) { return build_perm_schedule(a, s); }

void perm_free(mbx_adjoint* a) {
  if (!a) return;
  if (a->perm) (void)hipFree(a->perm);
  if (a->lslot) (void)hipFree(a->lslot);
  if (a->spos) (void)hipFree(a->spos);
  if (a->lidx) (void)hipFree(a->lidx);
  if (a->stage) (void)hipFree(a->stage);
  if (a->vt) (void)hipFree(a->vt);
  delete a;
}

void perm_apply(const mbx_adjoint* a, int value_bytes, const void* val, hipStream_t s) {
  if (a->nnz == 0) return;
  if (value_bytes == 8)
    permute_values<double>(a, (const double*)val, s);
  else if (value_bytes == 4)
    permute_values<float>(a, (const float*)val, s);
  else
    permute_values<uint16_t>(a, (const uint16_t*)val, s);
}

}  // namespace mbx_detail

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
  // The transposition indexes by column (radix-sort key width, column
  // pointers): refuse columns outside [0, ncol) and a malformed row_ptr
  // before any of it runs.
  {
    const int cst = mbx_detail::csr_check(d_row_ptr, d_col, nrow, ncol, nnz, s, nullptr);
    if (cst != MBX_OK) return cst;
  }
  mbx_adjoint* a = new (std::nothrow) mbx_adjoint();
  if (!a) return MBX_ENOMEM;
  a->nrow = nrow;
  a->ncol = ncol;
  a->nnz = nnz;
  a->nsrc = nnz;
  const size_t m = (size_t)(nnz > 0 ? nnz : 1);
  int32_t *row = nullptr, *iota = nullptr;
  uint32_t* keys = nullptr;
  void* tmp = nullptr;
  int st = MBX_OK;
  if (hipMalloc((void**)&a->col_ptr, sizeof(int32_t) * ((size_t)ncol + 1)) != hipSuccess ||
      hipMalloc((void**)&a->row_idx, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc((void**)&a->perm, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc(&a->vt, 8 * m) != hipSuccess || hipMalloc(&a->stage, 8 * m) != hipSuccess ||
      hipMalloc((void**)&a->lslot, sizeof(uint16_t) * m) != hipSuccess ||
      hipMalloc((void**)&a->spos, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc((void**)&a->lidx, sizeof(int32_t) * m) != hipSuccess ||
      hipMalloc((void**)&row, sizeof(int32_t) * m) != hipSuccess ||
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
  if (st == MBX_OK) st = build_perm_schedule(a, s);
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
  if (a->stage) (void)hipFree(a->stage);
  if (a->lslot) (void)hipFree(a->lslot);
  if (a->spos) (void)hipFree(a->spos);
  if (a->lidx) (void)hipFree(a->lidx);
  drop_blocks(a);
  mbx_plan_destroy(a->plan_t);
  delete a;
}

int mbx_adjoint_set_blocks(mbx_adjoint* a, int32_t nblocks, void* stream) {
  if (!a || nblocks < -1 || nblocks > 16) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (a->tb) {  // a dx on any stream may still read them
    (void)hipDeviceSynchronize();
    drop_blocks(a);
  }
  if (nblocks == 0) return MBX_OK;
  int st = mbx_colblock_create(a->col_ptr, a->row_idx, a->ncol, a->nrow, a->nnz,
                               nblocks < 0 ? 0 : nblocks, a->plan_t->ipt, stream, &a->tb);
  if (st != MBX_OK) return st;
  // destinations: the blocked form's slots up to its last block; sources:
  // the nnz CSR values, then one virtual (zero) source per padding slot
  int32_t nb = 0;
  int32_t bounds[17];
  (void)mbx_colblock_info(a->tb, &nb, nullptr, bounds);
  std::vector<int32_t> pad = mbx_detail::colblock_padding(a->tb);
  const long long n = (long long)a->nnz + (long long)pad.size();
  mbx_adjoint* p = new (std::nothrow) mbx_adjoint();
  if (!p) {
    drop_blocks(a);
    return MBX_ENOMEM;
  }
  a->tbp = p;
  p->nnz = (int32_t)n;
  p->nsrc = a->nnz;
  const size_t m = (size_t)(n > 0 ? n : 1);
  int32_t* d_pad = nullptr;
  if (hipMalloc((void**)&p->perm, 4 * m) != hipSuccess ||
      hipMalloc((void**)&p->lslot, 2 * m) != hipSuccess ||
      hipMalloc((void**)&p->spos, 4 * m) != hipSuccess ||
      hipMalloc((void**)&p->lidx, 4 * m) != hipSuccess || hipMalloc(&p->stage, 8 * m) != hipSuccess ||
      hipMalloc(&p->vt, 8 * m) != hipSuccess ||
      hipMalloc((void**)&d_pad, 4 * (pad.size() + 1)) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK && !pad.empty() &&
      hipMemcpyAsync(d_pad, pad.data(), 4 * pad.size(), hipMemcpyHostToDevice, s) != hipSuccess)
    st = MBX_EHIP;
  if (st == MBX_OK && n > 0) {
    tb_real_kernel<<<blocks(n), kBlock, 0, s>>>(mbx_detail::colblock_slot_map(a->tb), (int)n,
                                                a->perm, p->perm);
    if (!pad.empty())
      tb_pad_kernel<<<blocks((long long)pad.size()), kBlock, 0, s>>>(d_pad, (int)pad.size(),
                                                                     a->nnz, p->perm);
    if (hipGetLastError() != hipSuccess) st = MBX_EHIP;
  }
  if (st == MBX_OK && n > 0) st = build_perm_schedule(p, s);
  if (hipStreamSynchronize(s) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  if (d_pad) (void)hipFree(d_pad);
  if (st != MBX_OK) drop_blocks(a);
  return st;
}

int mbx_adjoint_auto(mbx_adjoint* a, int32_t value_kind, void* stream, mbx_launch_choice* out) {
  if (!a) return MBX_EINVAL;
  mbx_launch_choice c = {};
  mbx_plan* p = nullptr;
  // A^T is an ncol x nrow CSR: row pointer col_ptr, column indices row_idx
  int st = mbx_plan_create_auto(a->col_ptr, a->row_idx, a->ncol, a->nrow, a->nnz, value_kind,
                                stream, &p, &c);
  if (st != MBX_OK) return st;
  if (a->tb) (void)hipDeviceSynchronize();  // a dx on any stream may still use the old plan
  mbx_plan_destroy(a->plan_t);
  a->plan_t = p;
  // the rule's column-blocked form for spread columns; the single product
  // on A^T otherwise
  st = mbx_adjoint_set_blocks(a, c.rule == MBX_RULE_SPREAD_COLUMNS ? c.column_blocks : 0,
                              stream);
  if (st == MBX_OK && out) *out = c;
  return st;
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
  if (a->tb) {  // one permutation from CSR order into the blocked A^T, then its products
    const int vb = value_kind == MBX_VALUE_F64 ? 8 : value_kind == MBX_VALUE_F32 ? 4 : 2;
    if (value_kind != MBX_VALUE_F64 && value_kind != MBX_VALUE_F32 &&
        value_kind != MBX_VALUE_BF16)
      return MBX_EINVAL;
    if (a->tbp->nnz > 0) mbx_detail::perm_apply(a->tbp, vb, d_val, s);
    return mbx_colblock_spmv(a->tb, value_kind, a->tbp->vt, alpha, d_g, beta, d_dx, stream);
  }
  switch (value_kind) {
    case MBX_VALUE_F64:
      if (nnz > 0) permute_values<double>(a, (const double*)d_val, s);
      return mbx_spmv_f64(a->plan_t, a->col_ptr, a->row_idx, (const double*)a->vt, alpha,
                          (const double*)d_g, beta, (double*)d_dx, stream);
    case MBX_VALUE_F32:
      if (nnz > 0) permute_values<float>(a, (const float*)d_val, s);
      return mbx_spmv_f32(a->plan_t, a->col_ptr, a->row_idx, (const float*)a->vt, (float)alpha,
                          (const float*)d_g, (float)beta, (float*)d_dx, stream);
    case MBX_VALUE_BF16:
      if (nnz > 0) permute_values<uint16_t>(a, (const uint16_t*)d_val, s);
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
  const int x_rsrc = (uint64_t)p->ncol * (f64 ? 8u : 4u) <= 0xFFFFFFFFull;
  switch (p->ipt) {
#define MBX_CASE(I)                                                                         \
  case I:                                                                                   \
    if (f64)                                                                                \
      launch_sddmm<I>(p, x_rsrc, d_row_ptr, d_col, (const double*)d_g, (const double*)d_x,  \
                      (double*)d_dval, s);                                                  \
    else                                                                                    \
      launch_sddmm<I>(p, x_rsrc, d_row_ptr, d_col, (const float*)d_g, (const float*)d_x,    \
                      (float*)d_dval, s);                                                   \
    break;
    MBX_CASE(2) MBX_CASE(4) MBX_CASE(6) MBX_CASE(8) MBX_CASE(10) MBX_CASE(12) MBX_CASE(14)
    MBX_CASE(16)
#undef MBX_CASE
    default:
      return MBX_EUNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

}  // extern "C"

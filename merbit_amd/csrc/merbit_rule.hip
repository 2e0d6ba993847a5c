// merbit_rule.hip -- the launch rule: pattern statistics -> launch shape
// (mbx_launch_rule, mbx_plan_create_auto).
//
// The reference computes pattern statistics (degree moments, empty rows,
// src/class.cu:223-381; fast and normal tiles per block, include/tile.cuh:
// 177-233) but compiles one SIGMA and BLOCK for every matrix.  Here the
// statistics choose the launch, without a timing sweep, from what bounds
// the product (measured on MI355X, warm and rotated copies):
//   * few tiles (under 16 per CU at 8 items per thread): the launch is
//     short and unevenly filled, 4 items per thread give twice the tiles
//     (2 when even that leaves CUs without a tile), unless tile-sorted
//     gathers (below) still have 2 tiles per CU (a uniform random batch
//     member of 95160 rows: 14.9 against 16.0 us);
//   * local columns (every tile of 2048 merge items spans < 65536
//     columns, e.g. banded): the gathers hit L2 and the stream is the bound:
//     every tile resident, 16-bit column offsets, and for FP32 / bfloat16
//     values FP32 products with FP64 sums (the compensated sums are
//     VALU-bound once the stream is cut; 4-byte products in LDS let more
//     tiles' loads share a CU: banded BF16 155 -> 137 us, FP32 171 -> 164);
//     8 items per thread for FP64 values, 12 for FP32 and bfloat16 values
//     (their shorter values leave a tile's loads short; with 4-byte products
//     the larger tile still fits: FP32 kernel 155.5 -> 151.9 us);
//   * x of at most 2 MiB: the gathers stay in L2; tile-sorted gathers
//     (below), else 8 items per thread, at most 4 tiles per CU;
//   * larger x with a hot column set (the kHot* most frequent columns hold
//     at least a quarter of the entries, power-law patterns): the product is
//     bound by the vector L1's miss slots; tile-sorted gathers, else the hot
//     values in LDS (mbx_plan_split_hot) with the tiles per CU following
//     from that LDS; FP32 / bfloat16 values with FP32 products, whose
//     smaller tile LDS leaves room for more hot values at 3 tiles per CU;
//   where a pattern is not local and ncol fits the sorted words (2^20 at 16
//   items per thread), tile-sorted gathers (mbx_plan_sort_cols) at 16 items
//   per thread: each gather instruction reads 64 consecutive sorted columns
//   of its tile, so lines shared within the tile are requested once (warm
//   R-MAT FP64 85.8 -> 76.6 us against the hot split, FP32 67.4 -> 62.1;
//   a uniform random batch member of 194898 rows 57.9 -> 50.0);
//   * larger x without one (uniform random graphs): 3 tiles per CU, and the
//     column-blocked form (mbx_colblock_*) is recommended with slices of at
//     most 3.5 MiB of x.
// Whatever the rule, a pattern whose longest row is at most 1/16 of the
// chosen tile gets row-aligned tiles (mbx_plan_create_rows) at up to 8
// items per thread: no row crosses a tile boundary, so the carry fixup
// launch goes (the 256^2 Laplacian 6.1 -> 4.2 us per product in a graph,
// banded FP32 161 -> 156 us) for at most 1/16 more tiles.
// The stream policy stays MBX_STREAM_AUTO (by the product's bytes).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <limits.h>
#include <stdint.h>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kHotF64 = 4096;   // 32 KB of FP64 x: 3 tiles per CU with it in LDS
// 31 KB of FP32 x: with 4-byte products (MBX_ACC_F32_PRODUCTS) 3 tiles per
// CU at 8 items per thread (R-MAT FP32 kernel 64.2 us against 66.4 for the
// compensated sums with 5800 hot values, 4 per CU; BF16 62.1 against 63.7)
constexpr int kHotF32 = 7900;
constexpr double kHotShare = 0.25;
constexpr long long kXFitsBytes = 2ll << 20;
constexpr long long kSliceBytes = (7ll << 20) / 2;  // column-blocked form's x slice
constexpr int kFewTilesPerCu = 16;
// Tile-sorted gathers: 16 items per thread (the slot bits leave 20 for the
// column), or for uniform patterns of at most 2^19 columns 32, the tiles of
// 8192 items run by 512 threads (13 slot bits): twice the entries per x line
// (batch members of 194898 rows 52.9 -> 46.8 us, of 100399 rows 18.2 ->
// 16.2, of 95160 rows 15.7 either way).  The few-tiles rule takes them from
// 1 tile per CU on (a member of 33513 rows: 10.7 us at 4 items per thread
// unsorted, 10.2 sorted at 16).
constexpr int kSortIpt = 16;
constexpr int kSortWideIpt = 32;
constexpr int kSortFewHotIpt = 12;  // few tiles, power law
constexpr long long kSortMaxCols = 1ll << 20;
constexpr long long kSortWideMaxCols = 1ll << 19;
constexpr int kSortMinTilesPerCu = 1;
// Row-aligned tiles at 12 items per thread measured slower (banded BF16 /
// FP32 tile kernel +8 %, against +-1 % at 8): the rule keeps them to <= 8.
constexpr int kRowTilesMaxIpt = 8;

inline int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// Entries in the k most frequent columns (sorted counts, descending).
__global__ void rule_top_kernel(const int32_t* __restrict__ scnt, int k,
                                unsigned long long* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < k) atomicAdd(out, (unsigned long long)scnt[i]);
}

// Longest row (atomicMax of the per-wave maxima).
__global__ void rule_max_row_kernel(const int32_t* __restrict__ row_ptr, int nrow,
                                    int* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int len = i < nrow ? row_ptr[i + 1] - row_ptr[i] : 0;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) len = max(len, __shfl_xor(len, d));
  if ((threadIdx.x & 63) == 0 && len > 0) atomicMax(out, len);
}

// Widest column range of one tile (one workgroup per tile).
__global__ __launch_bounds__(256) void rule_span_kernel(const int32_t* __restrict__ col,
                                                        const int4* __restrict__ coord,
                                                        int* __restrict__ span) {
  __shared__ int s_lo[4], s_hi[4];
  const int t = blockIdx.x;
  const int k0 = coord[t].y, k1 = coord[t + 1].y;
  int lo = INT_MAX, hi = INT_MIN;
  for (int k = k0 + (int)threadIdx.x; k < k1; k += 256) {
    const int c = col[k];
    lo = min(lo, c);
    hi = max(hi, c);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    lo = min(lo, __shfl_xor(lo, d));
    hi = max(hi, __shfl_xor(hi, d));
  }
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0 && k1 > k0) {
    lo = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
    hi = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
    const long long w = (long long)hi - lo;
    atomicMax(span, w > INT_MAX ? INT_MAX : (int)w);
  }
}

// The column features: the widest tile span at 8 items per thread and the
// share of the entries in the `hot` most frequent columns.  With cnt_keep
// (ncol counters, the caller's), the column counts are left there.
int column_features(const int32_t* rp, const int32_t* col, int nrow, int ncol, int nnz, int hot,
                    hipStream_t s, int* span_out, double* share_out, int* max_row_out,
                    int32_t* cnt_keep) {
  *span_out = 0;
  *share_out = 0.0;
  *max_row_out = 0;
  if (nnz == 0 || ncol == 0) return MBX_OK;
  mbx_plan* p8 = nullptr;
  int st = mbx_plan_create(rp, nrow, ncol, nnz, 8, 256, s, &p8);
  if (st != MBX_OK) return st;
  int32_t *cnt = nullptr, *scnt = nullptr;
  int* d_span = nullptr;
  unsigned long long* d_top = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  if (cnt_keep) cnt = cnt_keep;
  if ((!cnt_keep && hipMallocAsync((void**)&cnt, 4 * (size_t)ncol, s) != hipSuccess) ||
      hipMallocAsync((void**)&scnt, 4 * (size_t)ncol, s) != hipSuccess ||
      hipMallocAsync((void**)&d_span, 2 * sizeof(int), s) != hipSuccess ||
      hipMallocAsync((void**)&d_top, sizeof(unsigned long long), s) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK && (hipMemsetAsync(d_span, 0, 2 * sizeof(int), s) != hipSuccess ||
                       hipMemsetAsync(d_top, 0, sizeof(unsigned long long), s) != hipSuccess))
    st = MBX_EHIP;
  if (st == MBX_OK) {
    rule_span_kernel<<<p8->ntiles, 256, 0, s>>>(col, p8->d_coord, d_span);
    rule_max_row_kernel<<<blocks(nrow), kBlock, 0, s>>>(rp, nrow, d_span + 1);
    st = mbx_detail::count_columns(col, nnz, ncol, cnt, s);
  }
  const int cbits = 32 - __builtin_clz((unsigned)nnz);  // counts are at most nnz
  if (st == MBX_OK) {
    if (hipcub::DeviceRadixSort::SortKeysDescending(nullptr, tmp_bytes, cnt, scnt, ncol, 0, cbits,
                                                    s) != hipSuccess ||
        hipMallocAsync(&tmp, tmp_bytes ? tmp_bytes : 1, s) != hipSuccess ||
        hipcub::DeviceRadixSort::SortKeysDescending(tmp, tmp_bytes, cnt, scnt, ncol, 0, cbits, s) !=
            hipSuccess)
      st = MBX_EHIP;
  }
  const int k = hot < ncol ? hot : ncol;
  unsigned long long top = 0;
  int span[2] = {0, 0};  // widest tile span, longest row
  if (st == MBX_OK) {
    rule_top_kernel<<<blocks(k), kBlock, 0, s>>>(scnt, k, d_top);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&top, d_top, sizeof(top), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(span, d_span, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
      st = MBX_EHIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  for (void* q : {(void*)(cnt_keep ? nullptr : cnt), (void*)scnt, (void*)d_span, (void*)d_top, tmp})
    if (q) (void)hipFreeAsync(q, s);
  mbx_plan_destroy(p8);
  if (st != MBX_OK) {
    (void)hipGetLastError();
    return st;
  }
  *span_out = span[0];
  *share_out = (double)top / (double)nnz;
  *max_row_out = span[1];
  return MBX_OK;
}

// mbx_launch_rule, leaving the column counts in cnt_keep when it is given.
int launch_rule(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow, int32_t ncol,
                int32_t nnz, int32_t value_kind, void* stream, mbx_launch_choice* out,
                int32_t* cnt_keep) {
  if (!out) return MBX_EINVAL;
  if (nrow < 0 || ncol < 0 || nnz < 0 || (nrow > 0 && !d_row_ptr) || (nnz > 0 && !d_col))
    return MBX_EINVAL;
  if (value_kind != MBX_VALUE_F64 && value_kind != MBX_VALUE_F32 &&
      value_kind != MBX_VALUE_BF16)
    return MBX_EINVAL;
  mbx_launch_choice c = {};
  c.stream_mode = MBX_STREAM_AUTO;
  c.accumulation = MBX_ACC_COMPENSATED_F32;
  const bool f64 = value_kind == MBX_VALUE_F64;
  const int sx = f64 ? 8 : 4;
  const int hot = f64 ? kHotF64 : kHotF32;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return MBX_EHIP;
  }
  c.cus = cus;
  c.x_bytes = (double)ncol * sx;
  c.ntiles8 = (int32_t)(((long long)nrow + nnz + 2047) / 2048);
  if (nnz > 0) {
    int span = 0, max_row = 0;
    double share = 0.0;
    const int st = column_features(d_row_ptr, d_col, nrow, ncol, nnz, hot, (hipStream_t)stream,
                                   &span, &share, &max_row, cnt_keep);
    if (st != MBX_OK) return st;
    c.max_tile_span = span;
    c.hot_share = share;
    c.max_row = max_row;
  }
  const bool local = nnz > 0 && c.max_tile_span < 65536;
  const bool sortable = nnz > 0 && !local && ncol <= kSortMaxCols;
  const bool wide_ok = ncol <= kSortWideMaxCols;
  // sorted gathers at 16 items per thread, or (uniform patterns) the wide
  // tiles when there is a tile of them per CU (their tiles are 1/4 of those
  // at 8 items per thread)
  auto sorted = [&](bool uniform) {
    const bool wide = uniform && wide_ok &&
                      (long long)c.ntiles8 >= 4ll * kSortMinTilesPerCu * cus;
    c.items_per_thread = wide ? kSortWideIpt : kSortIpt;
    c.sort_cols = 1;
  };
  if (c.ntiles8 < kFewTilesPerCu * cus) {
    // tiles of 16 items per thread are half those of 8.  A power-law pattern
    // takes 12 items per thread: the hot lines are shared within smaller
    // tiles too, and more tiles fill the chip (R-MAT of 2^18 rows and 2^22
    // entries: 23.6 us at 12 against 26.2 at 16 and 29.4 at 32)
    if (sortable && (long long)c.ntiles8 >= 2ll * kSortMinTilesPerCu * cus) {
      if (c.hot_share >= kHotShare) {
        c.items_per_thread = kSortFewHotIpt;
        c.sort_cols = 1;
      } else {
        sorted(true);
      }
    } else  // twice the tiles at 4 items per thread; 2 when even those leave CUs idle
      c.items_per_thread = 2ll * c.ntiles8 < cus ? 2 : 4;
    c.rule = MBX_RULE_FEW_TILES;
  } else if (local) {
    c.items_per_thread = f64 ? 8 : 12;
    c.compress_cols = 1;
    if (!f64) c.accumulation = MBX_ACC_F32_PRODUCTS;
    c.rule = MBX_RULE_LOCAL;
  } else if ((long long)ncol * sx <= kXFitsBytes) {
    if (sortable) {
      sorted(true);
    } else {
      c.items_per_thread = 8;
      c.wg_per_cu = 4;
    }
    c.rule = MBX_RULE_X_FITS;
  } else if (c.hot_share >= kHotShare) {
    if (sortable) {
      sorted(false);
    } else {
      c.items_per_thread = 8;
      c.hot_columns = hot;
    }
    if (!f64) c.accumulation = MBX_ACC_F32_PRODUCTS;
    c.rule = MBX_RULE_HOT_COLUMNS;
  } else {
    c.items_per_thread = 8;
    c.wg_per_cu = 3;
    const long long b = ((long long)ncol * sx + kSliceBytes - 1) / kSliceBytes;
    c.column_blocks = (int32_t)(b > 16 ? 16 : b);
    c.rule = MBX_RULE_SPREAD_COLUMNS;
  }
  c.row_tiles = nnz > 0 && c.items_per_thread <= kRowTilesMaxIpt &&
                (long long)c.max_row * mbx_detail::kRowTileDiv <= 256ll * c.items_per_thread;
  *out = c;
  return MBX_OK;
}

}  // namespace

extern "C" {

int mbx_launch_rule(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow, int32_t ncol,
                    int32_t nnz, int32_t value_kind, void* stream, mbx_launch_choice* out) {
  return launch_rule(d_row_ptr, d_col, nrow, ncol, nnz, value_kind, stream, out, nullptr);
}

int mbx_plan_create_auto(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                         int32_t ncol, int32_t nnz, int32_t value_kind, void* stream,
                         mbx_plan** out, mbx_launch_choice* choice) {
  if (!out) return MBX_EINVAL;
  *out = nullptr;
  // the column counts are made once, for the rule and the hot split
  int32_t* cnt = nullptr;
  hipStream_t s = (hipStream_t)stream;
  if (nnz > 0 && ncol > 0 && hipMallocAsync((void**)&cnt, 4 * (size_t)ncol, s) != hipSuccess) {
    (void)hipGetLastError();
    return MBX_ENOMEM;
  }
  mbx_launch_choice c;
  mbx_plan* p = nullptr;
  int st = launch_rule(d_row_ptr, d_col, nrow, ncol, nnz, value_kind, stream, &c, cnt);
  if (st == MBX_OK && c.row_tiles) {
    st = mbx_plan_create_rows(d_row_ptr, nrow, ncol, nnz, c.items_per_thread, 256, stream, &p);
    if (st == MBX_EUNSUPPORTED) {  // a row longer than the features saw (no columns)
      c.row_tiles = 0;
      st = MBX_OK;
    }
  }
  if (st == MBX_OK && !c.row_tiles)
    st = mbx_plan_create(d_row_ptr, nrow, ncol, nnz, c.items_per_thread, 256, stream, &p);
  if (st != MBX_OK) {
    if (cnt) (void)hipFreeAsync(cnt, s);
    return st;
  }
  (void)mbx_plan_set_stream(p, c.stream_mode);
  (void)mbx_plan_set_residency(p, c.wg_per_cu);
  (void)mbx_plan_set_accumulation(p, c.accumulation);
  if (c.compress_cols) {
    st = mbx_plan_compress_cols(p, d_col, stream);
    if (st == MBX_EUNSUPPORTED) {  // cannot happen for a local pattern; keep 32-bit columns
      c.compress_cols = 0;
      st = MBX_OK;
    }
  }
  if (st == MBX_OK && c.sort_cols) {
    st = mbx_plan_sort_cols(p, d_col, stream);
    if (st == MBX_EUNSUPPORTED || st == MBX_EINVAL) {  // columns out of range: plain gathers
      c.sort_cols = 0;
      st = MBX_OK;
    }
  }
  if (st == MBX_OK && c.hot_columns > 0)
    st = mbx_detail::split_hot_counted(p, d_col, c.hot_columns, cnt, (hipStream_t)stream,
                                       nullptr);
  if (cnt) (void)hipFreeAsync(cnt, s);
  if (st != MBX_OK) {
    mbx_plan_destroy(p);
    return st;
  }
  if (choice) *choice = c;
  *out = p;
  return MBX_OK;
}

}  // extern "C"

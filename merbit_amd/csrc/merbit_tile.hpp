// merbit_tile.hpp -- the merge-path tile computation shared by the SpMV
// kernels of one matrix (merbit_spmv.hip) and of a batch of matrices in one
// persistent launch (merbit_batch.hip).  Internal to libmerbit_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "merbit_internal.hpp"

namespace mbx_detail {

// ---------------------------------------------------------------------------
// Arithmetic policies
// ---------------------------------------------------------------------------

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

struct F64 {
  using val_t = double;
  using x_t = double;
  using y_t = double;
  using scal_t = double;
  using lds_t = double;
  using acc_t = double;
  __device__ static lds_t product(val_t v, x_t x) { return v * x; }
  __device__ static acc_t zero() { return 0.0; }
  __device__ static void add(acc_t& a, lds_t p) { a += p; }
  // a precedes b in merge order
  __device__ static acc_t combine(acc_t a, acc_t b) { return a + b; }
  __device__ static acc_t shfl_up(acc_t a, int d) { return __shfl_up(a, d); }
  __device__ static lds_t to_lds(acc_t a) { return a; }
  __device__ static double to_f64(lds_t a) { return a; }
  __device__ static void store(y_t* y, scal_t alpha, lds_t v, scal_t beta, bool beta_zero) {
    *y = beta_zero ? alpha * v : alpha * v + beta * *y;
  }
  __device__ static void store_f64(y_t* y, scal_t alpha, double v, scal_t beta, bool beta_zero) {
    *y = beta_zero ? alpha * v : alpha * v + beta * *y;
  }
};

// Compensated FP32: products split by TwoProduct (p + e == v*x exactly),
// running sums kept as (s, c) with s updated by TwoSum and c collecting the
// rounding errors.
template <class V>
struct CompF32 {
  using val_t = V;
  using x_t = float;
  using y_t = float;
  using scal_t = float;
  using lds_t = float2;
  using acc_t = float2;
  __device__ static float widen(float v) { return v; }
  __device__ static float widen(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
  __device__ static lds_t product(val_t v, x_t x) {
    const float a = widen(v);
    const float p = a * x;
    return make_float2(p, __builtin_fmaf(a, x, -p));
  }
  __device__ static acc_t zero() { return make_float2(0.f, 0.f); }
  __device__ static void add(acc_t& a, lds_t p) {
    float s, e;
    two_sum(a.x, p.x, s, e);
    a.x = s;
    a.y += e + p.y;
  }
  __device__ static acc_t combine(acc_t a, acc_t b) {
    float s, e;
    two_sum(a.x, b.x, s, e);
    return make_float2(s, (a.y + b.y) + e);
  }
  __device__ static acc_t shfl_up(acc_t a, int d) {
    return make_float2(__shfl_up(a.x, d), __shfl_up(a.y, d));
  }
  __device__ static lds_t to_lds(acc_t a) { return a; }
  __device__ static double to_f64(lds_t a) { return (double)a.x + (double)a.y; }
  __device__ static void store(y_t* y, scal_t alpha, lds_t v, scal_t beta, bool beta_zero) {
    const float r = v.x + v.y;
    *y = beta_zero ? alpha * r : alpha * r + beta * *y;
  }
  __device__ static void store_f64(y_t* y, scal_t alpha, double v, scal_t beta, bool beta_zero) {
    *y = beta_zero ? (float)((double)alpha * v)
                   : (float)((double)alpha * v + (double)beta * (double)*y);
  }
};

using F32C = CompF32<float>;
using BF16C = CompF32<uint16_t>;

// merge_search over the global row_ptr (include/utils.cuh:5-29): the
// (rows consumed, nnz consumed) coordinate of tile boundary t.
__device__ __forceinline__ int2 tile_start_coord(const int32_t* __restrict__ row_ptr, int nrow,
                                                 int nnz, int tile_items, int t) {
  const long long path = (long long)nrow + nnz;
  const long long diag = min((long long)t * tile_items, path);
  long long lo = diag - nnz > 0 ? diag - nnz : 0;
  long long hi = diag < nrow ? diag : nrow;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if ((long long)row_ptr[mid + 1] <= diag - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  return make_int2((int)lo, (int)(diag - lo));
}


// LDS of one tile: local row ends, then the product / row-sum area (products
// of the tile's nonzeros first, sums of the rows ending in the tile after
// them: together at most WG*IPT entries, the merge-path identity).
template <int WG, int IPT, class P>
struct TileSmem {
  int re[WG * IPT + 1];                 // row_ptr[r0+1+i] - k0
  typename P::lds_t v[WG * IPT];
  typename P::acc_t wtail[WG / 64];
  int wflag[WG / 64];
  double tail;
};

// One merge-path tile of y = alpha*A*x + beta*y (the work of one workgroup).
// Rows ending in the tile are written to y; carry[2*tile] receives the
// partial of the row open at the tile end and carry[2*tile+1] the partial of
// a first row that began in an earlier tile (completed by fixup_tile).
template <int WG, int IPT, class P>
__device__ __forceinline__ void spmv_tile(
    TileSmem<WG, IPT, P>& sm, int tile, const int32_t* __restrict__ row_ptr,
    const int32_t* __restrict__ col, const typename P::val_t* __restrict__ val,
    const typename P::x_t* __restrict__ x, typename P::y_t* __restrict__ y,
    const int2* __restrict__ coord, double* __restrict__ carry, typename P::scal_t alpha,
    typename P::scal_t beta, int beta_zero) {
  constexpr int NW = WG / 64;
  using lds_t = typename P::lds_t;
  using acc_t = typename P::acc_t;
  int* s_re = sm.re;
  lds_t* s_v = sm.v;
  const int tid = threadIdx.x;
  const int2 c0 = coord[tile];
  const int2 c1 = coord[tile + 1];
  const int r0 = c0.x, k0 = c0.y;
  const int rows_t = c1.x - r0;
  const int nnz_t = c1.y - k0;

  // 1. Stage: row ends of the rows ending here, then products.  The open
  //    row at the tile end never ends inside the tile: its end is +inf.
  //    All global loads of the tile are issued before the first wait.
  {
    int re[IPT];
    int ci[IPT];
    typename P::val_t vv[IPT];
    typename P::x_t xv[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      re[j] = INT_MAX;
      ci[j] = 0;
      vv[j] = typename P::val_t(0);
      if (i < rows_t) re[j] = row_ptr[r0 + 1 + i];
      if (i < nnz_t) {
        ci[j] = col[k0 + i];
        vv[j] = val[k0 + i];
      }
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      if (i < rows_t) s_re[i] = re[j] - k0;
    }
    if (tid == 0) s_re[rows_t] = INT_MAX;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      xv[j] = typename P::x_t(0);
      if (i < nnz_t) xv[j] = x[ci[j]];
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      if (i < nnz_t) s_v[i] = P::product(vv[j], xv[j]);
    }
  }
  __syncthreads();

  // 2. Per-thread walk over merge items [tid*IPT, tid*IPT+IPT) of the tile,
  //    starting from a merge-path search in the LDS row ends.
  const int d0 = tid * IPT;
  const int items = min(max(nnz_t + rows_t - d0, 0), IPT);
  lds_t* s_y = s_v + nnz_t;
  acc_t sum = P::zero(), head = P::zero();
  int head_row = 0;
  bool flagged = false;
  if (items > 0) {
    int r = lds_merge_search(s_re, d0, nnz_t, rows_t);
    int k = d0 - r;
    int re = s_re[r];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      if (i < items) {
        if (k < re) {
          P::add(sum, s_v[k]);
          ++k;
        } else {
          if (flagged) {
            s_y[r] = P::to_lds(sum);  // row began and ended in this thread
          } else {
            head = sum;
            head_row = r;
            flagged = true;
          }
          sum = P::zero();
          ++r;
          re = s_re[r];
        }
      }
    }
  }

  // 3. Segmented inclusive scan of the threads' open-row partials.  A
  //    thread with a row end starts a new segment; seg is the nearest such
  //    lane at or below this one.
  const int lane = tid & 63, wave = tid >> 6;
  const unsigned long long fmask = __ballot(flagged);
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const unsigned long long fb = fmask & upto;
  const int seg = fb ? 63 - __clzll(fb) : -1;
  acc_t s = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const acc_t o = P::shfl_up(s, d);
    if (lane >= d && lane - d >= seg) s = P::combine(o, s);
  }
  const acc_t cin = P::shfl_up(s, 1);
  if (lane == 63) {
    sm.wtail[wave] = s;
    sm.wflag[wave] = fmask != 0ull;
  }
  __syncthreads();
  acc_t win = P::zero();
#pragma unroll
  for (int w = 0; w < NW - 1; ++w)
    if (w < wave) win = sm.wflag[w] ? sm.wtail[w] : P::combine(win, sm.wtail[w]);
  const bool ext = (fmask & ((1ull << lane) - 1ull)) == 0ull;
  if (flagged) {
    acc_t c = (lane == 0) ? P::zero() : cin;
    if (ext) c = P::combine(win, c);
    s_y[head_row] = P::to_lds(P::combine(c, head));
  }
  if (tid == WG - 1) {
    acc_t t = s;
    if (fmask == 0ull) t = P::combine(win, t);
    sm.tail = P::to_f64(t);
  }
  __syncthreads();

  // 4. Coalesced write-out of the rows that end in this tile, and the carry
  //    slots: tail = partial of the row open at the tile end, head = partial
  //    of a row that began in an earlier tile (completed by the fixup).
  const bool head_multi = rows_t > 0 && row_ptr[r0] < k0;
  for (int i = tid; i < rows_t; i += WG) {
    const lds_t v = s_y[i];
    if (i == 0 && head_multi)
      carry[2 * tile + 1] = P::to_f64(v);
    else
      P::store(y + r0 + i, alpha, v, beta, beta_zero != 0);
  }
  if (tid == 0) carry[2 * tile] = sm.tail;
}

// Carry fixup, one wave per tile t >= 1: when t's first row began in an
// earlier tile ts and ends in t, y[row] = alpha*(tail[ts] + ... + tail[t-1]
// + head[t]) + beta*y[row].  Lane l sums tails ts+l, ts+l+64, ... in order,
// then a fixed xor butterfly adds the lanes: the result does not depend on
// scheduling, and rows spanning thousands of tiles stay 64-way parallel.
template <class P>
__device__ __forceinline__ void fixup_tile(int t, int lane, const int32_t* __restrict__ row_ptr,
                                           const int2* __restrict__ coord,
                                           const double* __restrict__ carry,
                                           typename P::y_t* __restrict__ y, int tile_items,
                                           typename P::scal_t alpha, typename P::scal_t beta,
                                           int beta_zero) {
  const int2 c0 = coord[t];
  const int2 c1 = coord[t + 1];
  const int r0 = c0.x, k0 = c0.y;
  if (c1.x <= r0) return;  // the row does not end in this tile
  const int rs = row_ptr[r0];
  if (rs >= k0) return;    // the row began in this tile
  const int ts = (int)(((long long)r0 + rs) / tile_items);
  double acc = 0.0;
  for (int s = ts + lane; s < t; s += 64) acc += carry[2 * s];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d);
  if (lane == 0) P::store_f64(y + r0, alpha, acc + carry[2 * t + 1], beta, beta_zero != 0);
}


}  // namespace mbx_detail

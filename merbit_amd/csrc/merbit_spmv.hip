// merbit_spmv.hip -- merge-path tile SpMV for MI355X (gfx950, wave64).
//
// One SpMV is two stream-ordered kernels over a plan made once per matrix:
//
//   plan      partition_kernel: merge-path search of every tile boundary over
//             (row ends row_ptr[1..nrow], nnz index 0..nnz-1); stores the
//             (rows consumed, nnz consumed) coordinate of each tile start.
//             Reference: merge_search include/utils.cuh:5-29 and the TILE plan
//             include/tile.cuh:28-122 (which stores a per-32-lane descriptor;
//             here only per-tile coordinates are kept and each thread finds
//             its own start by a search in LDS).
//   tile      tile_spmv_kernel: one workgroup per tile of WG*IPT merge items.
//             Stages the tile's row ends and its products val*x[col] in LDS,
//             walks IPT merge items per thread, then a 64-lane segmented scan
//             (shuffles + ballot) and a cross-wave scan finish the rows that
//             end in the tile.  Rows wholly inside the tile are written to y
//             with coalesced stores; the partial of the row open at the tile
//             end (tail) and of a row that began in an earlier tile (head) go
//             to per-tile carry slots.
//             Reference: merbit_kernel include/tile.cuh:476-561 (fast_track,
//             load_smem, normal_track, warp_segmented_sum, load_mem).
//   fixup     carry_fixup_kernel: for every row that spans tiles, sums the
//             tail slots of the tiles it covers and the head slot of the tile
//             where it ends, in a fixed order, and writes y.  Replaces the
//             reference's block-boundary global atomics and its ping-pong
//             zeroing (tile.cuh:354-375, 597-603), so the result is
//             deterministic and y needs no pre-zeroing.
//
// Value policies: F64 (plain FP64 sums, the parity path) and F32C / BF16C
// (FP32 or bfloat16 values, FP32 x/y, TwoProduct + TwoSum compensated FP32
// sums, FP64 carries).

#include <hip/hip_runtime.h>
#include <limits.h>
#include <new>
#include <stdint.h>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

using mbx_detail::lds_merge_search;
using mbx_detail::xcd_tile;

namespace {

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

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// merge_search over the global row_ptr (include/utils.cuh:5-29).
__global__ void partition_kernel(const int32_t* __restrict__ row_ptr, int nrow,
                                 int nnz, int tile_items, int ntiles,
                                 int2* __restrict__ coord) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
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
  coord[t] = make_int2((int)lo, (int)(diag - lo));
}

template <int WG, int IPT, class P>
__global__ __launch_bounds__(WG) void tile_spmv_kernel(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const typename P::val_t* __restrict__ val, const typename P::x_t* __restrict__ x,
    typename P::y_t* __restrict__ y, const int2* __restrict__ coord,
    double* __restrict__ carry, int ntiles, typename P::scal_t alpha,
    typename P::scal_t beta, int beta_zero) {
  constexpr int T = WG * IPT;
  constexpr int NW = WG / 64;
  using lds_t = typename P::lds_t;
  using acc_t = typename P::acc_t;
  __shared__ int s_re[T + 1];   // local row ends: row_ptr[r0+1+i] - k0
  __shared__ lds_t s_v[T];      // products [0,nnz_t), row sums after them
  __shared__ acc_t s_wtail[NW];
  __shared__ int s_wflag[NW];
  __shared__ double s_tail;

  const int tile = xcd_tile(blockIdx.x, ntiles);
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
    s_wtail[wave] = s;
    s_wflag[wave] = fmask != 0ull;
  }
  __syncthreads();
  acc_t win = P::zero();
#pragma unroll
  for (int w = 0; w < NW - 1; ++w)
    if (w < wave) win = s_wflag[w] ? s_wtail[w] : P::combine(win, s_wtail[w]);
  const bool ext = (fmask & ((1ull << lane) - 1ull)) == 0ull;
  if (flagged) {
    acc_t c = (lane == 0) ? P::zero() : cin;
    if (ext) c = P::combine(win, c);
    s_y[head_row] = P::to_lds(P::combine(c, head));
  }
  if (tid == WG - 1) {
    acc_t t = s;
    if (fmask == 0ull) t = P::combine(win, t);
    s_tail = P::to_f64(t);
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
  if (tid == 0) carry[2 * tile] = s_tail;
}

// Carry fixup, one wave per tile t >= 1: when t's first row began in an
// earlier tile ts and ends in t, y[row] = alpha*(tail[ts] + ... + tail[t-1]
// + head[t]) + beta*y[row].  Lane l sums tails ts+l, ts+l+64, ... in order,
// then a fixed xor butterfly adds the lanes: the result does not depend on
// scheduling, and rows spanning thousands of tiles stay 64-way parallel.
template <class P>
__global__ __launch_bounds__(256) void carry_fixup_kernel(
    const int32_t* __restrict__ row_ptr, const int2* __restrict__ coord,
    const double* __restrict__ carry, typename P::y_t* __restrict__ y, int ntiles,
    int tile_items, typename P::scal_t alpha, typename P::scal_t beta, int beta_zero) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t < 1 || t >= ntiles) return;  // uniform per wave
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

}  // namespace

// ---------------------------------------------------------------------------
// Plan and C ABI
// ---------------------------------------------------------------------------

namespace {

bool supported_shape(int wg, int ipt) {
  if (wg != 256) return false;
  switch (ipt) {
    case 2: case 4: case 6: case 8: case 12: case 16: return true;
    default: return false;
  }
}

template <class P, int IPT>
hipError_t launch_tile(const mbx_plan* p, hipStream_t s, const int32_t* rp,
                       const int32_t* col, const typename P::val_t* val,
                       const typename P::x_t* x, typename P::y_t* y,
                       typename P::scal_t alpha, typename P::scal_t beta) {
  const int bz = beta == typename P::scal_t(0);
  if (p->ev_begin) (void)hipEventRecord(p->ev_begin, s);
  tile_spmv_kernel<256, IPT, P><<<p->ntiles, 256, 0, s>>>(
      rp, col, val, x, y, p->d_coord, p->d_carry, p->ntiles, alpha, beta, bz);
  if (p->ev_end) (void)hipEventRecord(p->ev_end, s);
  p->ev_begin = p->ev_end = nullptr;
  if (p->ntiles > 1) {
    const int grid = (p->ntiles + 3) / 4;
    carry_fixup_kernel<P><<<grid, 256, 0, s>>>(rp, p->d_coord, p->d_carry, y, p->ntiles,
                                               p->tile_items, alpha, beta, bz);
  }
  return hipGetLastError();
}

template <class P>
int run_spmv(const mbx_plan* p, const int32_t* rp, const int32_t* col,
             const typename P::val_t* val, typename P::scal_t alpha,
             const typename P::x_t* x, typename P::scal_t beta, typename P::y_t* y,
             void* stream) {
  if (!p) return MBX_EINVAL;
  if (p->nrow > 0 && (!rp || !y)) return MBX_EINVAL;
  if (p->nnz > 0 && (!col || !val || !x)) return MBX_EINVAL;
  if (p->ntiles == 0) return MBX_OK;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  switch (p->ipt) {
    case 2: e = launch_tile<P, 2>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 4: e = launch_tile<P, 4>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 6: e = launch_tile<P, 6>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 8: e = launch_tile<P, 8>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 12: e = launch_tile<P, 12>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 16: e = launch_tile<P, 16>(p, s, rp, col, val, x, y, alpha, beta); break;
    default: return MBX_EUNSUPPORTED;
  }
  return e == hipSuccess ? MBX_OK : MBX_EHIP;
}

}  // namespace

extern "C" {

int mbx_plan_create(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol,
                    int32_t nnz, int32_t items_per_thread, int32_t wg_size,
                    void* stream, mbx_plan** out) {
  if (!out) return MBX_EINVAL;
  *out = nullptr;
  if (nrow < 0 || ncol < 0 || nnz < 0 || (nrow > 0 && !d_row_ptr)) return MBX_EINVAL;
  const int ipt = items_per_thread ? items_per_thread : 8;
  const int wg = wg_size ? wg_size : 256;
  if (!supported_shape(wg, ipt)) return MBX_EUNSUPPORTED;
  const long long T = (long long)wg * ipt;
  const long long path = (long long)nrow + nnz;
  if (path + T >= (long long)INT_MAX) return MBX_EOVERFLOW;
  mbx_plan* p = new (std::nothrow) mbx_plan();
  if (!p) return MBX_ENOMEM;
  p->nrow = nrow;
  p->ncol = ncol;
  p->nnz = nnz;
  p->wg = wg;
  p->ipt = ipt;
  p->tile_items = (int)T;
  p->ntiles = (int)((path + T - 1) / T);
  p->d_coord = nullptr;
  p->d_carry = nullptr;
  p->ev_begin = p->ev_end = nullptr;
  const size_t nt1 = (size_t)p->ntiles + 1;
  if (hipMalloc((void**)&p->d_coord, sizeof(int2) * nt1) != hipSuccess ||
      hipMalloc((void**)&p->d_carry, sizeof(double) * 2 * nt1) != hipSuccess) {
    (void)hipGetLastError();
    mbx_plan_destroy(p);
    return MBX_ENOMEM;
  }
  const int grid = (p->ntiles + 1 + 255) / 256;
  partition_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(d_row_ptr, nrow, nnz, p->tile_items,
                                                          p->ntiles, p->d_coord);
  if (hipGetLastError() != hipSuccess) {
    mbx_plan_destroy(p);
    return MBX_EHIP;
  }
  *out = p;
  return MBX_OK;
}

void mbx_plan_destroy(mbx_plan* p) {
  if (!p) return;
  if (p->d_coord) (void)hipFree(p->d_coord);
  if (p->d_carry) (void)hipFree(p->d_carry);
  delete p;
}

int mbx_plan_info(const mbx_plan* p, int32_t* ntiles, int32_t* tile_items, int32_t* wg_size) {
  if (!p) return MBX_EINVAL;
  if (ntiles) *ntiles = p->ntiles;
  if (tile_items) *tile_items = p->tile_items;
  if (wg_size) *wg_size = p->wg;
  return MBX_OK;
}

int mbx_plan_partition(const mbx_plan* p, int32_t* h_rows, int32_t* h_nnz, int32_t count) {
  if (!p || !h_rows || !h_nnz || count < 0 || count > p->ntiles + 1) return MBX_EINVAL;
  if (count == 0) return MBX_OK;
  int2* tmp = new (std::nothrow) int2[count];
  if (!tmp) return MBX_ENOMEM;
  if (hipMemcpy(tmp, p->d_coord, sizeof(int2) * (size_t)count, hipMemcpyDeviceToHost) !=
      hipSuccess) {
    delete[] tmp;
    return MBX_EHIP;
  }
  for (int i = 0; i < count; ++i) {
    h_rows[i] = tmp[i].x;
    h_nnz[i] = tmp[i].y;
  }
  delete[] tmp;
  return MBX_OK;
}

int mbx_plan_set_events(mbx_plan* p, void* ev_begin, void* ev_end) {
  if (!p) return MBX_EINVAL;
  p->ev_begin = (hipEvent_t)ev_begin;
  p->ev_end = (hipEvent_t)ev_end;
  return MBX_OK;
}

int mbx_spmv_f64(const mbx_plan* p, const int32_t* rp, const int32_t* col,
                 const double* val, double alpha, const double* x, double beta,
                 double* y, void* stream) {
  return run_spmv<F64>(p, rp, col, val, alpha, x, beta, y, stream);
}

int mbx_spmv_f32(const mbx_plan* p, const int32_t* rp, const int32_t* col,
                 const float* val, float alpha, const float* x, float beta, float* y,
                 void* stream) {
  return run_spmv<F32C>(p, rp, col, val, alpha, x, beta, y, stream);
}

int mbx_spmv_bf16(const mbx_plan* p, const int32_t* rp, const int32_t* col,
                  const uint16_t* val, float alpha, const float* x, float beta, float* y,
                  void* stream) {
  return run_spmv<BF16C>(p, rp, col, val, alpha, x, beta, y, stream);
}

const char* mbx_status_str(int st) {
  switch (st) {
    case MBX_OK: return "ok";
    case MBX_EINVAL: return "invalid argument";
    case MBX_EOVERFLOW: return "nrow + nnz exceeds the int32 merge-path range";
    case MBX_EUNSUPPORTED: return "unsupported items_per_thread / wg_size / k";
    case MBX_ENOMEM: return "device allocation failed";
    case MBX_EHIP: return "HIP runtime or launch error";
    default: return "unknown status";
  }
}

}  // extern "C"

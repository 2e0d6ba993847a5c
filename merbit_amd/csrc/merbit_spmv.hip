// merbit_spmv.hip -- merge-path tile SpMV for MI355X (gfx950, wave64).
//
// One SpMV is two stream-ordered kernels over a plan made once per matrix:
//
//   plan      partition_kernel: merge-path search of every tile boundary over
//             (row ends row_ptr[1..nrow], nnz index 0..nnz-1); stores the
//             (rows consumed, nnz consumed, row_ptr[rows]) coordinate of each
//             tile start.  Reference: merge_search include/utils.cuh:5-29 and
//             the TILE plan include/tile.cuh:28-122 (which stores a per-32-lane
//             descriptor; here only per-tile coordinates are kept and every
//             thread finds its start from row-end marks inside the tile).
//   tile      tile_spmv_kernel: one workgroup per tile of WG*IPT merge items.
//             Loads the tile's row ends, columns and values in 16-byte
//             groups, gathers x, stores the products val*x[col] in LDS and
//             marks the merge position of every row end; each thread walks
//             its IPT items, then DPP wave scans and a cross-wave step finish
//             the rows that end in the tile.  Rows wholly inside the tile are
//             written to y with coalesced stores; the partial of the row open
//             at the tile end (tail) and of a row that began in an earlier
//             tile (head) go to per-tile carry slots.
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
// sums per thread, FP64 partials across threads and tiles).

#include <hip/hip_runtime.h>
#include <limits.h>
#include <new>
#include <stdint.h>

#include <mutex>

#include "merbit_amd.h"
#include "merbit_internal.hpp"
#include "merbit_tile.hpp"

using mbx_detail::timed_launch;
using mbx_detail::xcd_tile;

namespace {

using mbx_detail::BF16C;
using mbx_detail::F32C;
using mbx_detail::F64;
using mbx_detail::TileSmem;

__global__ void plan_stats_kernel(const int4* __restrict__ coord, int ntiles,
                                  int* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) mbx_detail::plan_stats_update(coord, t, out);
}

// The ends of row_ptr (out[2] = row_ptr[0], out[3] = row_ptr[nrow]), read
// back by plan creation to refuse a row_ptr that does not run from 0 to nnz.
__global__ void row_ptr_ends_kernel(const int32_t* __restrict__ row_ptr, int nrow,
                                    int* __restrict__ out) {
  if (threadIdx.x == 0) {
    out[2] = row_ptr[0];
    out[3] = row_ptr[nrow];
  }
}

__global__ void partition_kernel(const int32_t* __restrict__ row_ptr, int nrow, int nnz,
                                 int tile_items, int ntiles, int4* __restrict__ coord) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  coord[t] = mbx_detail::tile_start_coord(row_ptr, nrow, nnz, tile_items, t);
}

// Row-aligned tiles: boundary t is the start of the row open at merge
// diagonal t * spacing, i.e. the path point (i, row_ptr[i]) at or before it.
// With spacing = tile_items - (longest row), every tile holds at most
// tile_items merge items and none starts or ends inside a row.
__global__ void partition_rows_kernel(const int32_t* __restrict__ row_ptr, int nrow, int nnz,
                                      int spacing, int ntiles, int4* __restrict__ coord) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  int4 c = mbx_detail::tile_start_coord(row_ptr, nrow, nnz, spacing, t);
  c.y = c.z;
  coord[t] = c;
}

// Longest row (out[0], atomicMax) and the ends of row_ptr (out[2], out[3]).
__global__ void max_row_kernel(const int32_t* __restrict__ row_ptr, int nrow,
                               int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int len = 0;
  if (i < nrow) len = row_ptr[i + 1] - row_ptr[i];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) len = max(len, __shfl_xor(len, d));
  if ((threadIdx.x & 63) == 0 && len > 0) atomicMax(out, len);
  if (i == 0) {
    out[2] = row_ptr[0];
    out[3] = row_ptr[nrow];
  }
}

// Column span of one tile (one workgroup per tile): base = smallest column,
// written to coord[t].w; *wide is set when the span needs more than 16 bits.
__global__ __launch_bounds__(256) void col_span_kernel(const int32_t* __restrict__ col,
                                                       int4* __restrict__ coord,
                                                       int* __restrict__ wide) {
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
  if (threadIdx.x == 0) {
    lo = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
    hi = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
    const bool any = k1 > k0;
    coord[t].w = any ? lo : 0;
    if (any && (long long)hi - lo > 0xFFFF) atomicOr(wide, 1);
  }
}

// 16-bit offsets of every nonzero of a tile from the tile's base column.
__global__ __launch_bounds__(256) void col16_kernel(const int32_t* __restrict__ col,
                                                    const int4* __restrict__ coord,
                                                    uint16_t* __restrict__ col16) {
  const int t = blockIdx.x;
  const int k0 = coord[t].y, k1 = coord[t + 1].y, base = coord[t].w;
  for (int k = k0 + (int)threadIdx.x; k < k1; k += 256)
    col16[k] = (uint16_t)(col[k] - base);
}

template <int WG, int IPT, bool VEC, class P>
__global__ __launch_bounds__(WG) void tile_spmv_kernel(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const typename P::val_t* __restrict__ val, const typename P::x_t* __restrict__ x,
    typename P::y_t* __restrict__ y, const int4* __restrict__ coord,
    double* __restrict__ carry, int ntiles, int ncol, typename P::scal_t alpha,
    typename P::scal_t beta, int beta_zero) {
  extern __shared__ __align__(16) char smem[];
  TileSmem<WG, IPT, P> sm(smem);
  mbx_detail::spmv_tile<WG, IPT, VEC, P>(sm, xcd_tile(blockIdx.x, ntiles), ncol, row_ptr, col,
                                         val, x, y, coord, carry, alpha, beta, beta_zero);
}

// Persistent tile kernel of a hot-column plan (P = HotX<...>): each
// workgroup copies the hot x values into LDS once (16-byte loads), then runs
// the tiles of its share: each XCD takes a contiguous range of tiles, its
// workgroups take every stride-th tile of it.
template <int WG, int IPT, class P>
__global__ __launch_bounds__(WG) void tile_spmv_hot_kernel(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ colh,
    const typename P::val_t* __restrict__ val, const typename P::x_t* __restrict__ x,
    typename P::y_t* __restrict__ y, const int4* __restrict__ coord,
    double* __restrict__ carry, int ntiles, int ncol, typename P::scal_t alpha,
    typename P::scal_t beta, int beta_zero, const typename P::x_t* __restrict__ xh, int nhot) {
  extern __shared__ __align__(16) char smem[];
  TileSmem<WG, IPT, P> sm(smem);
  using X = typename P::x_t;
  X* s_xh = reinterpret_cast<X*>(smem + TileSmem<WG, IPT, P>::bytes());
  const int nv = (nhot * (int)sizeof(X) + 15) / 16;
  for (int i = threadIdx.x; i < nv; i += WG)
    reinterpret_cast<int4*>(s_xh)[i] = reinterpret_cast<const int4*>(xh)[i];
  sm.xh = s_xh;
  __syncthreads();
  const mbx_detail::PersistentShare sh = mbx_detail::persistent_share(blockIdx.x, gridDim.x,
                                                                      ntiles);
  for (int i = sh.w; i < sh.len; i += sh.stride) {
    mbx_detail::spmv_tile<WG, IPT, true, P>(sm, sh.first + i, ncol, row_ptr, colh, val, x, y,
                                            coord, carry, alpha, beta, beta_zero);
    __syncthreads();  // the next tile reuses the LDS
  }
}

// Tile-sorted column words of a plan (mbx_plan_sort_cols), one workgroup
// per tile (see sort_tile_cols).
template <int WG, int IPT>
__global__ __launch_bounds__(WG) void sort_cols_kernel(const int32_t* __restrict__ col,
                                                       int ncol,
                                                       const int4* __restrict__ coord,
                                                       int32_t* __restrict__ scol,
                                                       int* __restrict__ bad) {
  __shared__ uint32_t s_key[1 << mbx_detail::sort_slot_bits<WG, IPT>()];
  __shared__ int s_wsum[WG / 64];
  const int t = blockIdx.x;
  mbx_detail::sort_tile_cols<WG, IPT>(col, ncol, coord[t].y, coord[t + 1].y, scol, bad, s_key,
                                      s_wsum);
}

// Carry fixup, one thread per tile (see fixup_tile).
template <class P>
__global__ __launch_bounds__(256) void carry_fixup_kernel(
    const int4* __restrict__ coord, const double* __restrict__ carry,
    typename P::y_t* __restrict__ y, int ntiles, int tile_items, typename P::scal_t alpha,
    typename P::scal_t beta, int beta_zero) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  mbx_detail::fixup_tile<P>(t >= 1 && t < ntiles, t, coord, carry, y, tile_items, alpha, beta,
                            beta_zero);
}

}  // namespace

// ---------------------------------------------------------------------------
// Plan and C ABI
// ---------------------------------------------------------------------------

namespace {

bool supported_shape(int wg, int ipt) {
  if (wg != 256) return false;
  switch (ipt) {
    case 2: case 4: case 6: case 8: case 10: case 12: case 14: case 16: case 32: return true;
    default: return false;
  }
}

// The hot-column launch (mbx_plan_split_hot) of P at IPT: the hot x gather,
// then a persistent grid of as many workgroups as fit with the hot values in
// LDS (at most wg_per_cu per CU when the plan caps residency).  False when
// the hot values do not fit a workgroup's LDS for this value kind: the
// caller then runs the plain product on the caller's columns.
template <class P, int IPT>
bool launch_hot(const mbx_plan* p, hipStream_t s, const int32_t* rp,
                const typename P::val_t* val, const typename P::x_t* x, typename P::y_t* y,
                typename P::scal_t alpha, typename P::scal_t beta, int bz) {
  using Q = mbx_detail::HotX<P>;
  using X = typename P::x_t;
  const size_t xh_bytes = ((size_t)p->nhot * sizeof(X) + 15) / 16 * 16;
  const size_t lds =
      mbx_detail::residency_lds(TileSmem<256, IPT, Q>::bytes() + xh_bytes, p->wg_per_cu);
  if (lds > mbx_detail::kLdsPerCu || p->cus <= 0) return false;
  auto* kernel = tile_spmv_hot_kernel<256, IPT, Q>;
  mbx_detail::allow_lds(kernel, lds);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    return false;
  }
  if (p->wg_per_cu > 0 && per_cu > p->wg_per_cu) per_cu = p->wg_per_cu;
  const long long g = (long long)p->cus * per_cu;
  const int grid = (int)(g < p->ntiles ? g : p->ntiles);
  mbx_detail::hot_gather_kernel<X><<<(p->nhot + 255) / 256, 256, 0, s>>>(x, p->d_hot, p->nhot,
                                                               static_cast<X*>(p->d_xh));
  timed_launch(p, kernel, grid, 256, lds, s, rp, p->d_colh, val, x, y, p->d_coord, p->d_carry,
               p->ntiles, p->ncol, alpha, beta, bz, static_cast<const X*>(p->d_xh), p->nhot);
  return true;
}

template <class P, int IPT>
hipError_t launch_tile(const mbx_plan* p, hipStream_t s, const int32_t* rp,
                       const int32_t* col, const typename P::val_t* val,
                       const typename P::x_t* x, typename P::y_t* y,
                       typename P::scal_t alpha, typename P::scal_t beta) {
  // 32 items per thread: tiles of 8192 merge items run by 512 threads of 16
  constexpr int WG = IPT == 32 ? 512 : 256, KI = IPT == 32 ? 16 : IPT;
  const int bz = beta == typename P::scal_t(0);
  const bool vec = mbx_detail::vec_path_ok(rp, col, val, p->ncol);
  const long long footprint = 4ll * (p->nrow + 1) +
                              (long long)p->nnz * (4 + sizeof(typename P::val_t)) +
                              (long long)p->ncol * sizeof(typename P::x_t) +
                              (long long)p->nrow * sizeof(typename P::y_t);
  const bool nt = mbx_detail::stream_nontemporal(p->stream_mode, footprint);
  const size_t lds = mbx_detail::residency_lds(TileSmem<WG, KI, P>::bytes(), p->wg_per_cu);
  // the plan's 16-bit columns, instantiated for the tile shapes the banded
  // and graph products use
  constexpr bool kC16 = IPT == 8 || IPT == 12;
  bool launched = false;
  if constexpr (kC16) {
    if (vec && p->d_col16) {
      const int32_t* c16 = reinterpret_cast<const int32_t*>(p->d_col16);
      if (nt) {
        using Q = mbx_detail::Col16<mbx_detail::Streamed<P>>;
        mbx_detail::allow_lds(tile_spmv_kernel<WG, KI, true, Q>, lds);
        timed_launch(p, tile_spmv_kernel<WG, KI, true, Q>, p->ntiles, WG, lds, s, rp, c16, val,
                     x, y, p->d_coord, p->d_carry, p->ntiles, p->ncol, alpha, beta, bz);
      } else {
        using Q = mbx_detail::Col16<P>;
        mbx_detail::allow_lds(tile_spmv_kernel<WG, KI, true, Q>, lds);
        timed_launch(p, tile_spmv_kernel<WG, KI, true, Q>, p->ntiles, WG, lds, s, rp, c16, val,
                     x, y, p->d_coord, p->d_carry, p->ntiles, p->ncol, alpha, beta, bz);
      }
      launched = true;
    }
  }
  // the plan's tile-sorted columns (read in place of the caller's)
  constexpr bool kSortShape = KI == 8 || KI == 12 || KI == 16;
  if constexpr (kSortShape) {
    if (!launched && p->d_scol && mbx_detail::vec_path_ok(rp, p->d_scol, val, p->ncol)) {
      if (nt) {
        using Q = mbx_detail::SortX<mbx_detail::Streamed<P>>;
        mbx_detail::allow_lds(tile_spmv_kernel<WG, KI, true, Q>, lds);
        timed_launch(p, tile_spmv_kernel<WG, KI, true, Q>, p->ntiles, WG, lds, s, rp,
                     p->d_scol, val, x, y, p->d_coord, p->d_carry, p->ntiles, p->ncol, alpha,
                     beta, bz);
      } else {
        using Q = mbx_detail::SortX<P>;
        mbx_detail::allow_lds(tile_spmv_kernel<WG, KI, true, Q>, lds);
        timed_launch(p, tile_spmv_kernel<WG, KI, true, Q>, p->ntiles, WG, lds, s, rp,
                     p->d_scol, val, x, y, p->d_coord, p->d_carry, p->ntiles, p->ncol, alpha,
                     beta, bz);
      }
      launched = true;
    }
  }
  if constexpr (kC16) {  // hot-column plans: the tile shapes of the power-law products
    if (!launched && vec && p->d_colh && p->nhot > 0) {
      launched = nt ? launch_hot<mbx_detail::Streamed<P>, IPT>(p, s, rp, val, x, y, alpha, beta,
                                                               bz)
                    : launch_hot<P, IPT>(p, s, rp, val, x, y, alpha, beta, bz);
    }
  }
  if (launched) {
  } else if (vec && nt) {
    mbx_detail::allow_lds(tile_spmv_kernel<WG, KI, true, mbx_detail::Streamed<P>>, lds);
    timed_launch(p, tile_spmv_kernel<WG, KI, true, mbx_detail::Streamed<P>>, p->ntiles, WG,
                 lds, s, rp, col, val, x, y, p->d_coord, p->d_carry, p->ntiles, p->ncol, alpha,
                 beta, bz);
  } else if (vec) {
    mbx_detail::allow_lds(tile_spmv_kernel<WG, KI, true, P>, lds);
    timed_launch(p, tile_spmv_kernel<WG, KI, true, P>, p->ntiles, WG, lds, s, rp, col, val, x,
                 y, p->d_coord, p->d_carry, p->ntiles, p->ncol, alpha, beta, bz);
  } else {
    mbx_detail::allow_lds(tile_spmv_kernel<WG, KI, false, P>, lds);
    timed_launch(p, tile_spmv_kernel<WG, KI, false, P>, p->ntiles, WG, lds, s, rp, col, val,
                 x, y, p->d_coord, p->d_carry, p->ntiles, p->ncol, alpha, beta, bz);
  }
  p->ev_begin = p->ev_end = nullptr;
  if (p->split_rows > 0) {  // some row crosses a tile boundary
    const int grid = (p->ntiles + 255) / 256;
    carry_fixup_kernel<P><<<grid, 256, 0, s>>>(p->d_coord, p->d_carry, y, p->ntiles,
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
    case 10: e = launch_tile<P, 10>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 12: e = launch_tile<P, 12>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 14: e = launch_tile<P, 14>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 16: e = launch_tile<P, 16>(p, s, rp, col, val, x, y, alpha, beta); break;
    case 32: e = launch_tile<P, 32>(p, s, rp, col, val, x, y, alpha, beta); break;
    default: return MBX_EUNSUPPORTED;
  }
  return e == hipSuccess ? MBX_OK : MBX_EHIP;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point (mbx_spmv_f64_noalloc): the tile
// coordinates, then the FP64 carry slots, of at most kNoAllocTiles tiles of
// the default shape.  One copy per GPU, held until the process ends.  Host
// code reaches a GPU's copy only through the address hipGetSymbolAddress
// returns with that GPU current (noalloc_scratch); the array's name is not a
// device address on the host.
// ---------------------------------------------------------------------------

constexpr int kNoAllocIpt = 8;
constexpr int kNoAllocTileItems = 256 * kNoAllocIpt;
constexpr size_t kNoAllocBytes = size_t(4) << 20;
// coordinate (int4) + two carry slots (FP64) per tile boundary
constexpr int kNoAllocTiles = (int)(kNoAllocBytes / (sizeof(int4) + 2 * sizeof(double))) - 1;
constexpr size_t kNoAllocCarryOffset = sizeof(int4) * ((size_t)kNoAllocTiles + 1);

__device__ __align__(256) char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current GPU's scratch, looked up once per device.
int noalloc_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return MBX_EHIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return MBX_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return MBX_EHIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return MBX_OK;
}

// mbx_plan_create and mbx_plan_create_rows.
int plan_create(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                int32_t items_per_thread, int32_t wg_size, void* stream, bool row_tiles,
                mbx_plan** out) {
  if (!out) return MBX_EINVAL;
  *out = nullptr;
  if (nrow < 0 || ncol < 0 || nnz < 0 || (nrow > 0 && !d_row_ptr)) return MBX_EINVAL;
  if (nrow == 0 && nnz > 0) return MBX_EINVAL;
  const int ipt = items_per_thread ? items_per_thread : 8;
  const int wg = wg_size ? wg_size : 256;
  if (!supported_shape(wg, ipt)) return MBX_EUNSUPPORTED;
  const long long T = (long long)wg * ipt;
  const long long path = (long long)nrow + nnz;
  if (path + T >= (long long)INT_MAX) return MBX_EOVERFLOW;
  // Row-aligned tiles are spaced by T minus the longest row, which must be
  // at most T / kRowTileDiv (read back before the partition is sized).
  long long spacing = T;
  if (row_tiles && nrow > 0) {
    hipStream_t st = (hipStream_t)stream;
    int* d_m = nullptr;
    int h_m[4] = {0, 0, 0, 0};
    int rc = MBX_OK;
    if (hipMallocAsync((void**)&d_m, 4 * sizeof(int), st) != hipSuccess) {
      (void)hipGetLastError();
      return MBX_ENOMEM;
    }
    if (hipMemsetAsync(d_m, 0, 4 * sizeof(int), st) != hipSuccess) rc = MBX_EHIP;
    if (rc == MBX_OK) {
      max_row_kernel<<<(nrow + 255) / 256, 256, 0, st>>>(d_row_ptr, nrow, d_m);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(h_m, d_m, 4 * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = MBX_EHIP;
    }
    (void)hipFreeAsync(d_m, st);
    if (hipStreamSynchronize(st) != hipSuccess && rc == MBX_OK) rc = MBX_EHIP;
    if (rc != MBX_OK) {
      (void)hipGetLastError();
      return rc;
    }
    if (h_m[2] != 0 || h_m[3] != nnz) return MBX_EINVAL;
    if ((long long)h_m[0] * mbx_detail::kRowTileDiv > T) return MBX_EUNSUPPORTED;
    spacing = T - h_m[0];
  }
  mbx_plan* p = new (std::nothrow) mbx_plan();
  if (!p) return MBX_ENOMEM;
  p->nrow = nrow;
  p->ncol = ncol;
  p->nnz = nnz;
  p->wg = wg;
  p->ipt = ipt;
  p->tile_items = (int)T;
  p->tile_spacing = (int)spacing;
  p->max_rows = (int)T;
  p->split_rows = 0;
  p->ntiles = (int)((path + spacing - 1) / spacing);
  p->d_coord = nullptr;
  p->d_carry = nullptr;
  p->coord_refs = new (std::nothrow) std::atomic<int>(1);
  p->stream_mode = MBX_STREAM_AUTO;
  p->wg_per_cu = 0;
  p->ev_begin = p->ev_end = nullptr;
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      p->cus = 0;
    (void)hipGetLastError();
  }
  if (!p->coord_refs) {
    mbx_plan_destroy(p);
    return MBX_ENOMEM;
  }
  const size_t nt1 = (size_t)p->ntiles + 1;
  if (hipMalloc((void**)&p->d_coord, sizeof(int4) * nt1) != hipSuccess ||
      hipMalloc((void**)&p->d_carry, sizeof(double) * 2 * nt1) != hipSuccess) {
    (void)hipGetLastError();
    mbx_plan_destroy(p);
    return MBX_ENOMEM;
  }
  hipStream_t st = (hipStream_t)stream;
  const int grid = (p->ntiles + 1 + 255) / 256;
  if (spacing != T)
    partition_rows_kernel<<<grid, 256, 0, st>>>(d_row_ptr, nrow, nnz, (int)spacing, p->ntiles,
                                                p->d_coord);
  else
    partition_kernel<<<grid, 256, 0, st>>>(d_row_ptr, nrow, nnz, p->tile_items, p->ntiles,
                                           p->d_coord);
  // Read back once (plan creation synchronises the stream): whether any row
  // crosses a tile boundary (products skip the fixup launch when none does),
  // the most rows ending in one tile, and the ends of row_ptr.  The carry
  // slots (at least 4 ints) are unused until a product.
  int* d_stats = reinterpret_cast<int*>(p->d_carry);
  if (hipMemsetAsync(d_stats, 0, 4 * sizeof(int), st) != hipSuccess) {
    mbx_plan_destroy(p);
    return MBX_EHIP;
  }
  if (p->ntiles > 0)
    plan_stats_kernel<<<(p->ntiles + 255) / 256, 256, 0, st>>>(p->d_coord, p->ntiles, d_stats);
  if (nrow > 0) row_ptr_ends_kernel<<<1, 64, 0, st>>>(d_row_ptr, nrow, d_stats);
  int h_stats[4] = {0, 0, 0, 0};
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h_stats, d_stats, 4 * sizeof(int), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    mbx_plan_destroy(p);
    return MBX_EHIP;
  }
  if (nrow > 0 && (h_stats[2] != 0 || h_stats[3] != nnz)) {
    mbx_plan_destroy(p);
    return MBX_EINVAL;
  }
  p->max_rows = h_stats[0];
  p->split_rows = h_stats[1];
  *out = p;
  return MBX_OK;
}

}  // namespace

extern "C" {

int mbx_plan_create(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol,
                    int32_t nnz, int32_t items_per_thread, int32_t wg_size,
                    void* stream, mbx_plan** out) {
  return plan_create(d_row_ptr, nrow, ncol, nnz, items_per_thread, wg_size, stream, false, out);
}

int mbx_plan_create_rows(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol,
                         int32_t nnz, int32_t items_per_thread, int32_t wg_size,
                         void* stream, mbx_plan** out) {
  return plan_create(d_row_ptr, nrow, ncol, nnz, items_per_thread, wg_size, stream, true, out);
}

void mbx_plan_destroy(mbx_plan* p) {
  if (!p) return;
  // the coordinates go with the last plan object that shares them
  if (!p->coord_refs || p->coord_refs->fetch_sub(1) == 1) {
    if (p->d_coord) (void)hipFree(p->d_coord);
    delete p->coord_refs;
  }
  if (p->d_carry) (void)hipFree(p->d_carry);
  if (p->d_col16) (void)hipFree(p->d_col16);
  if (p->d_scol) (void)hipFree(p->d_scol);
  if (p->d_colh || p->d_hot || p->d_xh) mbx_detail::plan_drop_hot(p);
  delete p;
}

int mbx_plan_clone(const mbx_plan* p, mbx_plan** out) {
  if (!p || !out || !p->coord_refs) return MBX_EINVAL;
  *out = nullptr;
  mbx_plan* c = new (std::nothrow) mbx_plan(*p);
  if (!c) return MBX_ENOMEM;
  c->d_col16 = nullptr;  // the clone reads the caller's columns
  c->d_scol = nullptr;
  c->d_colh = nullptr;
  c->d_hot = nullptr;
  c->d_xh = nullptr;
  c->nhot = 0;
  c->ev_begin = c->ev_end = nullptr;
  c->d_carry = nullptr;
  if (hipMalloc((void**)&c->d_carry, sizeof(double) * 2 * ((size_t)p->ntiles + 1)) !=
      hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return MBX_ENOMEM;
  }
  p->coord_refs->fetch_add(1);
  *out = c;
  return MBX_OK;
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
  int4* tmp = new (std::nothrow) int4[count];
  if (!tmp) return MBX_ENOMEM;
  if (hipMemcpy(tmp, p->d_coord, sizeof(int4) * (size_t)count, hipMemcpyDeviceToHost) !=
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

int mbx_plan_set_stream(mbx_plan* p, int32_t mode) {
  if (!p || mode < MBX_STREAM_AUTO || mode > MBX_STREAM_NONTEMPORAL) return MBX_EINVAL;
  p->stream_mode = mode;
  return MBX_OK;
}

int mbx_plan_set_residency(mbx_plan* p, int32_t wg_per_cu) {
  if (!p || wg_per_cu < 0 || wg_per_cu > 16) return MBX_EINVAL;
  p->wg_per_cu = wg_per_cu;
  return MBX_OK;
}

int mbx_plan_set_accumulation(mbx_plan* p, int32_t mode) {
  if (!p || (mode != MBX_ACC_COMPENSATED_F32 && mode != MBX_ACC_F64 &&
             mode != MBX_ACC_F32_PRODUCTS))
    return MBX_EINVAL;
  p->accumulation = mode;
  return MBX_OK;
}

int mbx_plan_compress_cols(mbx_plan* p, const int32_t* d_col, void* stream) {
  if (!p) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (p->d_col16) {  // products on any stream may still read them
    (void)hipDeviceSynchronize();
    (void)hipFree(p->d_col16);
    p->d_col16 = nullptr;
  }
  if (!d_col) return MBX_OK;  // back to the caller's 32-bit columns
  if (p->ipt != 8 && p->ipt != 12) return MBX_EUNSUPPORTED;
  if (p->d_scol) return MBX_EUNSUPPORTED;  // the sorted columns exclude them
  if (p->ntiles == 0 || p->nnz == 0) return MBX_OK;
  int* d_wide = nullptr;
  uint16_t* c16 = nullptr;
  const size_t n16 = ((size_t)p->nnz + 7) / 8 * 8;  // whole 8-byte groups
  if (hipMalloc((void**)&d_wide, sizeof(int)) != hipSuccess) return MBX_ENOMEM;
  if (hipMalloc((void**)&c16, sizeof(uint16_t) * n16) != hipSuccess) {
    (void)hipFree(d_wide);
    return MBX_ENOMEM;
  }
  int st = MBX_OK, wide = 0;
  if (hipMemsetAsync(d_wide, 0, sizeof(int), s) != hipSuccess ||
      hipMemsetAsync(c16, 0, sizeof(uint16_t) * n16, s) != hipSuccess)
    st = MBX_EHIP;
  if (st == MBX_OK) {
    col_span_kernel<<<p->ntiles, 256, 0, s>>>(d_col, p->d_coord, d_wide);
    col16_kernel<<<p->ntiles, 256, 0, s>>>(d_col, p->d_coord, c16);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&wide, d_wide, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = MBX_EHIP;
  }
  (void)hipFree(d_wide);
  if (st == MBX_OK && wide) st = MBX_EUNSUPPORTED;
  if (st != MBX_OK) {
    (void)hipFree(c16);
    return st;
  }
  p->d_col16 = c16;
  return MBX_OK;
}

int mbx_plan_sort_cols(mbx_plan* p, const int32_t* d_col, void* stream) {
  if (!p) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (p->d_scol) {  // products on any stream may still read them
    (void)hipDeviceSynchronize();
    (void)hipFree(p->d_scol);
    p->d_scol = nullptr;
  }
  if (!d_col) return MBX_OK;  // back to the caller's columns
  if (p->ipt != 8 && p->ipt != 12 && p->ipt != 16 && p->ipt != 32) return MBX_EUNSUPPORTED;
  if (p->d_col16 || (p->d_colh && p->nhot > 0)) return MBX_EUNSUPPORTED;
  const int sb = p->ipt == 8    ? mbx_detail::sort_slot_bits<256, 8>()
                 : p->ipt == 32 ? mbx_detail::sort_slot_bits<512, 16>()
                                : mbx_detail::sort_slot_bits<256, 16>();  // 12 for IPT 12 too
  if ((long long)p->ncol > (1ll << (32 - sb))) return MBX_EUNSUPPORTED;
  if (p->ntiles == 0 || p->nnz == 0) return MBX_OK;
  const size_t padded = ((size_t)p->nnz + 3) / 4 * 4;  // whole 16-byte groups
  int* d_bad = nullptr;
  int32_t* sc = nullptr;
  if (hipMalloc((void**)&d_bad, sizeof(int)) != hipSuccess) return MBX_ENOMEM;
  if (hipMalloc((void**)&sc, sizeof(int32_t) * padded) != hipSuccess) {
    (void)hipFree(d_bad);
    return MBX_ENOMEM;
  }
  int st = MBX_OK, bad = 0;
  if (hipMemsetAsync(d_bad, 0, sizeof(int), s) != hipSuccess ||
      hipMemsetAsync(sc, 0, sizeof(int32_t) * padded, s) != hipSuccess)
    st = MBX_EHIP;
  if (st == MBX_OK) {
    switch (p->ipt) {
      case 8:
        sort_cols_kernel<256, 8><<<p->ntiles, 256, 0, s>>>(d_col, p->ncol, p->d_coord, sc, d_bad);
        break;
      case 12:
        sort_cols_kernel<256, 12><<<p->ntiles, 256, 0, s>>>(d_col, p->ncol, p->d_coord, sc,
                                                             d_bad);
        break;
      case 16:
        sort_cols_kernel<256, 16><<<p->ntiles, 256, 0, s>>>(d_col, p->ncol, p->d_coord, sc,
                                                             d_bad);
        break;
      default:  // 32: tiles of 512 threads x 16 items
        sort_cols_kernel<512, 16><<<p->ntiles, 512, 0, s>>>(d_col, p->ncol, p->d_coord, sc,
                                                             d_bad);
        break;
    }
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = MBX_EHIP;
  }
  (void)hipFree(d_bad);
  if (st == MBX_OK && bad) st = MBX_EINVAL;  // a column outside [0, ncol)
  if (st != MBX_OK) {
    (void)hipFree(sc);
    return st;
  }
  p->d_scol = sc;
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
  if (p && p->accumulation == MBX_ACC_F64)
    return run_spmv<mbx_detail::F32A>(p, rp, col, val, alpha, x, beta, y, stream);
  if (p && p->accumulation == MBX_ACC_F32_PRODUCTS)
    return run_spmv<mbx_detail::F32R>(p, rp, col, val, alpha, x, beta, y, stream);
  return run_spmv<F32C>(p, rp, col, val, alpha, x, beta, y, stream);
}

int mbx_spmv_bf16(const mbx_plan* p, const int32_t* rp, const int32_t* col,
                  const uint16_t* val, float alpha, const float* x, float beta, float* y,
                  void* stream) {
  if (p && p->accumulation == MBX_ACC_F64)
    return run_spmv<mbx_detail::BF16A>(p, rp, col, val, alpha, x, beta, y, stream);
  if (p && p->accumulation == MBX_ACC_F32_PRODUCTS)
    return run_spmv<mbx_detail::BF16R>(p, rp, col, val, alpha, x, beta, y, stream);
  return run_spmv<BF16C>(p, rp, col, val, alpha, x, beta, y, stream);
}

int mbx_spmv_f64_noalloc(const int32_t* d_row_ptr, const int32_t* d_col, const double* d_val,
                         int32_t nrow, int32_t ncol, int32_t nnz, double alpha,
                         const double* d_x, double beta, double* d_y, void* stream) {
  if (nrow < 0 || ncol < 0 || nnz < 0 || (nrow == 0 && nnz > 0)) return MBX_EINVAL;
  const long long path = (long long)nrow + nnz;
  const long long ntiles = (path + kNoAllocTileItems - 1) / kNoAllocTileItems;
  if (ntiles > kNoAllocTiles) return MBX_ESCRATCH;
  if (nrow > 0 && (!d_row_ptr || !d_y)) return MBX_EINVAL;
  if (nnz > 0 && (!d_col || !d_val || !d_x)) return MBX_EINVAL;
  if (ntiles == 0) return MBX_OK;
  void* base = nullptr;
  const int sst = noalloc_scratch(&base);
  if (sst != MBX_OK) return sst;
  // A plan of the default shape over the scratch: every tile coordinate is
  // written by the partition kernel and every carry slot the fixup reads is
  // written by its tile first, so nothing needs clearing between calls.  The
  // fixup launch is not skipped (whether a row crosses a tile boundary is not
  // read back: that would synchronise).
  mbx_plan p{};
  p.nrow = nrow;
  p.ncol = ncol;
  p.nnz = nnz;
  p.wg = 256;
  p.ipt = kNoAllocIpt;
  p.tile_items = kNoAllocTileItems;
  p.tile_spacing = kNoAllocTileItems;
  p.ntiles = (int)ntiles;
  p.max_rows = kNoAllocTileItems;
  p.split_rows = 1;
  p.d_coord = static_cast<int4*>(base);
  p.d_carry = reinterpret_cast<double*>(static_cast<char*>(base) + kNoAllocCarryOffset);
  p.stream_mode = MBX_STREAM_AUTO;
  hipStream_t s = (hipStream_t)stream;
  partition_kernel<<<(p.ntiles + 1 + 255) / 256, 256, 0, s>>>(d_row_ptr, nrow, nnz,
                                                               p.tile_items, p.ntiles, p.d_coord);
  if (hipGetLastError() != hipSuccess) return MBX_EHIP;
  const hipError_t e =
      launch_tile<F64, kNoAllocIpt>(&p, s, d_row_ptr, d_col, d_val, d_x, d_y, alpha, beta);
  return e == hipSuccess ? MBX_OK : MBX_EHIP;
}

const char* mbx_status_str(int st) {
  switch (st) {
    case MBX_OK: return "ok";
    case MBX_EINVAL: return "invalid argument";
    case MBX_EOVERFLOW: return "nrow + nnz exceeds the int32 merge-path range";
    case MBX_EUNSUPPORTED: return "unsupported items_per_thread / wg_size / k";
    case MBX_ENOMEM: return "device allocation failed";
    case MBX_EHIP: return "HIP runtime or launch error";
    case MBX_ESCRATCH:
      return "nrow + nnz exceeds the no-allocation scratch (131071 tiles of 2048 merge items)";
    default: return "unknown status";
  }
}

}  // extern "C"

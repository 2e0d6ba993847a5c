// merbit_batch.hip -- many uneven CSR matrices in one persistent launch,
// HIP-graph capture of repeated products, and the timing autotuner.
//
// A batch concatenates the merge-path partitions of its matrices into one
// global tile space.  A persistent kernel (one workgroup per resident slot)
// walks that space, each XCD taking a contiguous range of global tiles; every
// tile runs exactly the per-tile code of the single-matrix kernel, so a batch
// product is bit-identical to calling mbx_spmv_* once per matrix.  One fixup
// launch completes the rows that cross tiles in every matrix.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <new>
#include <vector>

#include "merbit_amd.h"
#include "merbit_internal.hpp"
#include "merbit_tile.hpp"

namespace {

using mbx_detail::BF16C;
using mbx_detail::F32C;
using mbx_detail::F64;
using mbx_detail::TileSmem;

struct BatchMat {
  const int32_t* row_ptr;
  const int32_t* col;
  const void* val;
  const void* x;
  void* y;
  int32_t nrow, ncol, nnz;
  int32_t tile_base;   // first global tile of this matrix
  int32_t coord_base;  // first coordinate (tile_base + matrix index)
};

__global__ void batch_partition_kernel(const BatchMat* __restrict__ mats, int count,
                                       int tile_items, int ncoord, int4* __restrict__ coord) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncoord) return;
  int lo = 0, hi = count - 1;  // last matrix whose coord_base <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mats[mid].coord_base <= g) lo = mid; else hi = mid - 1;
  }
  const BatchMat m = mats[lo];
  coord[g] = mbx_detail::tile_start_coord(m.row_ptr, m.nrow, m.nnz, tile_items,
                                          g - m.coord_base);
}

__device__ __forceinline__ int batch_matrix_of_(const int* __restrict__ tile_base, int count,
                                                int g);

__global__ void batch_max_rows_kernel(const BatchMat* __restrict__ mats,
                                      const int* __restrict__ tile_base, int count,
                                      int total_tiles, const int4* __restrict__ coord,
                                      int* __restrict__ out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_tiles) return;
  const int m = batch_matrix_of_(tile_base, count, g);
  mbx_detail::plan_stats_update(coord + mats[m].coord_base, g - mats[m].tile_base, out);
}

__device__ __forceinline__ int batch_matrix_of(const int* __restrict__ tile_base, int count,
                                               int g) {
  int lo = 0, hi = count - 1;  // last matrix whose tile_base <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_base[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int batch_matrix_of_(const int* __restrict__ tile_base, int count,
                                                int g) {
  return batch_matrix_of(tile_base, count, g);
}

template <int WG, int IPT, bool VEC, class P>
__global__ __launch_bounds__(WG) void batch_spmv_kernel(
    const BatchMat* __restrict__ mats, const int* __restrict__ tile_base, int count,
    int total_tiles, const int4* __restrict__ coord, double* __restrict__ carry,
    typename P::scal_t alpha, typename P::scal_t beta, int beta_zero) {
  extern __shared__ __align__(16) char smem[];
  TileSmem<WG, IPT, P> sm(smem);
  constexpr int X = mbx_detail::kXcds;
  const int b = blockIdx.x, nwg = gridDim.x;
  int first = 0, len = total_tiles, w = b, stride = nwg;
  if (nwg >= X) {  // every XCD has workgroups: each takes a contiguous range
    const int xcd = b % X;
    w = b / X;
    stride = (nwg - xcd + X - 1) / X;  // workgroups on this XCD
    const int q = total_tiles / X, r = total_tiles % X;
    first = xcd * q + min(xcd, r);
    len = q + (xcd < r ? 1 : 0);
  }
  for (int i = w; i < len; i += stride) {
    const int g = first + i;
    const int m = batch_matrix_of(tile_base, count, g);
    const BatchMat M = mats[m];
    mbx_detail::spmv_tile<WG, IPT, VEC, P>(
        sm, g - M.tile_base, M.ncol, M.row_ptr, M.col, static_cast<const typename P::val_t*>(M.val),
        static_cast<const typename P::x_t*>(M.x), static_cast<typename P::y_t*>(M.y),
        coord + M.coord_base, carry + 2ll * M.coord_base, alpha, beta, beta_zero);
    __syncthreads();  // the next tile reuses the LDS
  }
}

// Kernel shape of a batch's items_per_thread: 32 is a tile of 8192 merge
// items run by 512 threads of 16 items (the larger tile shares more x lines
// between its sorted gathers), every other value 256 threads.
template <int IPT>
struct BatchShape {
  static constexpr int WG = IPT == 32 ? 512 : 256;
  static constexpr int KI = IPT == 32 ? 16 : IPT;
  static constexpr bool kSort = KI == 8 || KI == 12 || KI == 16;
};

// Tile-sorted column words of every member (mbx_batch_sort_cols): one
// workgroup per global tile reads its member's columns from `mats` and writes
// the words to the member's array in `sorted` (see sort_tile_cols).
template <int WG, int IPT>
__global__ __launch_bounds__(WG) void batch_sort_cols_kernel(
    const BatchMat* __restrict__ mats, const BatchMat* __restrict__ sorted,
    const int* __restrict__ tile_base, int count, const int4* __restrict__ coord,
    int* __restrict__ bad) {
  __shared__ uint32_t s_key[1 << mbx_detail::sort_slot_bits<WG, IPT>()];
  __shared__ int s_wsum[WG / 64];
  const int g = blockIdx.x;
  const int m = batch_matrix_of(tile_base, count, g);
  const BatchMat M = mats[m];
  const int4* c = coord + M.coord_base + (g - M.tile_base);
  mbx_detail::sort_tile_cols<WG, IPT>(M.col, M.ncol, c[0].y, c[1].y,
                                      const_cast<int32_t*>(sorted[m].col), bad, s_key, s_wsum);
}

template <class P>
__global__ __launch_bounds__(256) void batch_fixup_kernel(
    const BatchMat* __restrict__ mats, const int* __restrict__ tile_base, int count,
    int total_tiles, int tile_items, const int4* __restrict__ coord,
    const double* __restrict__ carry, typename P::scal_t alpha, typename P::scal_t beta,
    int beta_zero) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const bool in = g < total_tiles;
  const int m = in ? batch_matrix_of(tile_base, count, g) : 0;
  const BatchMat M = mats[m];
  const int t = in ? g - M.tile_base : 0;
  // every lane calls fixup_tile (it cooperates across the wave); lanes of
  // different matrices pass their own matrix's pointers
  mbx_detail::fixup_tile<P>(in && t >= 1, t, coord + M.coord_base, carry + 2ll * M.coord_base,
                            static_cast<typename P::y_t*>(M.y), tile_items, alpha, beta,
                            beta_zero);
}

}  // namespace

struct mbx_batch {
  int32_t count, kind, ipt, tile_items, total_tiles, ncoord, grid;
  bool vec;            // every matrix suits the vector staging path
  int32_t max_rows;    // most rows ending in one tile over the batch
  int32_t split_rows;  // tiles whose first row began earlier (fixup needed)
  BatchMat* d_mats;
  int* d_tile_base;
  int4* d_coord;  // per tile boundary: rows, nnz consumed, row_ptr[rows]
  double* d_carry;
  long long footprint;  // bytes one batch product reads and writes
  int32_t stream_mode;  // MBX_STREAM_*
  int32_t wg_per_cu;    // resident tiles per CU; 0 = as many as fit
  int32_t cus;          // compute units of the device
  // mbx_batch_sort_cols: the member table with every column array replaced
  // by the member's tile-sorted words (in d_scol), and the resident grid of
  // the kernel that reads them
  BatchMat* d_mats_sorted;
  int32_t* d_scol;
  int32_t grid_sorted;
};

namespace {

// Resident workgroups of the kernel variant the batch will launch.
template <class P, int IPT>
int batch_grid(int device, bool vec, bool sorted) {
  using S = BatchShape<IPT>;
  if constexpr (!P::kSorted && S::kSort) {
    if (sorted) return batch_grid<mbx_detail::SortX<P>, IPT>(device, true, false);
  }
  int per_cu = 0, cus = 0;
  const size_t lds = TileSmem<S::WG, S::KI, P>::bytes();
  hipError_t e;
  if constexpr (P::kSorted)  // vector staging only
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, batch_spmv_kernel<S::WG, S::KI, true, P>, S::WG, lds);
  else
    e = vec ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, batch_spmv_kernel<S::WG, S::KI, true, P>, S::WG, lds)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, batch_spmv_kernel<S::WG, S::KI, false, P>, S::WG, lds);
  if (e != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return std::max(1, per_cu) * std::max(1, cus);
}

template <class P, int IPT>
void batch_launch(const mbx_batch* b, hipStream_t s, typename P::scal_t alpha,
                  typename P::scal_t beta) {
  using S = BatchShape<IPT>;
  constexpr int WG = S::WG, KI = S::KI;
  const int bz = beta == typename P::scal_t(0);
  int grid = std::min(b->grid, b->total_tiles);
  if (b->wg_per_cu > 0) grid = std::min(grid, b->wg_per_cu * b->cus);  // persistent
  const size_t lds = mbx_detail::residency_lds(TileSmem<WG, KI, P>::bytes(), b->wg_per_cu);
  const bool nt = mbx_detail::stream_nontemporal(b->stream_mode, b->footprint);
  bool launched = false;
  if constexpr (S::kSort) {  // the members' tile-sorted columns
    if (b->vec && b->d_mats_sorted) {
      int sgrid = std::min(b->grid_sorted, b->total_tiles);
      if (b->wg_per_cu > 0) sgrid = std::min(sgrid, b->wg_per_cu * b->cus);
      if (nt) {
        using Q = mbx_detail::SortX<mbx_detail::Streamed<P>>;
        mbx_detail::allow_lds(batch_spmv_kernel<WG, KI, true, Q>, lds);
        batch_spmv_kernel<WG, KI, true, Q><<<sgrid, WG, lds, s>>>(
            b->d_mats_sorted, b->d_tile_base, b->count, b->total_tiles, b->d_coord, b->d_carry,
            alpha, beta, bz);
      } else {
        using Q = mbx_detail::SortX<P>;
        mbx_detail::allow_lds(batch_spmv_kernel<WG, KI, true, Q>, lds);
        batch_spmv_kernel<WG, KI, true, Q><<<sgrid, WG, lds, s>>>(
            b->d_mats_sorted, b->d_tile_base, b->count, b->total_tiles, b->d_coord, b->d_carry,
            alpha, beta, bz);
      }
      launched = true;
    }
  }
  if (launched) {
  } else if (b->vec && nt) {
    mbx_detail::allow_lds(batch_spmv_kernel<WG, KI, true, mbx_detail::Streamed<P>>, lds);
    batch_spmv_kernel<WG, KI, true, mbx_detail::Streamed<P>><<<grid, WG, lds, s>>>(
        b->d_mats, b->d_tile_base, b->count, b->total_tiles, b->d_coord, b->d_carry, alpha, beta,
        bz);
  } else if (b->vec) {
    mbx_detail::allow_lds(batch_spmv_kernel<WG, KI, true, P>, lds);
    batch_spmv_kernel<WG, KI, true, P><<<grid, WG, lds, s>>>(
        b->d_mats, b->d_tile_base, b->count, b->total_tiles, b->d_coord, b->d_carry, alpha, beta,
        bz);
  } else {
    mbx_detail::allow_lds(batch_spmv_kernel<WG, KI, false, P>, lds);
    batch_spmv_kernel<WG, KI, false, P><<<grid, WG, lds, s>>>(
        b->d_mats, b->d_tile_base, b->count, b->total_tiles, b->d_coord, b->d_carry, alpha, beta,
        bz);
  }
  if (b->split_rows > 0)
    batch_fixup_kernel<P><<<(b->total_tiles + 255) / 256, 256, 0, s>>>(
      b->d_mats, b->d_tile_base, b->count, b->total_tiles, b->tile_items, b->d_coord,
      b->d_carry, alpha, beta, bz);
}

template <class P>
int batch_dispatch(const mbx_batch* b, hipStream_t s, double alpha, double beta, bool grid_only,
                   int device, bool sorted) {
  using S = typename P::scal_t;
  switch (b->ipt) {
#define MBX_CASE(I)                                                    \
  case I:                                                             \
    if (grid_only) return batch_grid<P, I>(device, b->vec, sorted);   \
    batch_launch<P, I>(b, s, (S)alpha, (S)beta);                      \
    return 0;
    MBX_CASE(2) MBX_CASE(4) MBX_CASE(6) MBX_CASE(8) MBX_CASE(10) MBX_CASE(12) MBX_CASE(14)
    MBX_CASE(16) MBX_CASE(32)
#undef MBX_CASE
    default: return -1;
  }
}

int batch_run(const mbx_batch* b, hipStream_t s, double alpha, double beta, bool grid_only,
              int device, bool sorted = false) {
  switch (b->kind) {
    case MBX_VALUE_F64: return batch_dispatch<F64>(b, s, alpha, beta, grid_only, device, sorted);
    case MBX_VALUE_F32:
      return batch_dispatch<F32C>(b, s, alpha, beta, grid_only, device, sorted);
    case MBX_VALUE_BF16:
      return batch_dispatch<BF16C>(b, s, alpha, beta, grid_only, device, sorted);
    default: return -1;
  }
}

bool vec_item_ok(const mbx_batch_item& it, int kind) {
  const int32_t* rp = it.row_ptr;
  const int32_t* col = it.col;
  switch (kind) {
    case MBX_VALUE_F64: return mbx_detail::vec_path_ok(rp, col, (const double*)it.val, it.ncol);
    case MBX_VALUE_F32: return mbx_detail::vec_path_ok(rp, col, (const float*)it.val, it.ncol);
    default: return mbx_detail::vec_path_ok(rp, col, (const uint16_t*)it.val, it.ncol);
  }
}

bool ipt_ok(int ipt) {
  return ipt == 2 || ipt == 4 || ipt == 6 || ipt == 8 || ipt == 10 || ipt == 12 || ipt == 14 ||
         ipt == 16 || ipt == 32;
}

}  // namespace

extern "C" {

int mbx_batch_create(const mbx_batch_item* items, int32_t count, int32_t value_kind,
                     int32_t items_per_thread, void* stream, mbx_batch** out) {
  if (!out) return MBX_EINVAL;
  *out = nullptr;
  if (!items || count <= 0) return MBX_EINVAL;
  if (value_kind != MBX_VALUE_F64 && value_kind != MBX_VALUE_F32 && value_kind != MBX_VALUE_BF16)
    return MBX_EINVAL;
  const int ipt = items_per_thread ? items_per_thread : 8;
  if (!ipt_ok(ipt)) return MBX_EUNSUPPORTED;
  const long long T = 256ll * ipt;
  std::vector<BatchMat> h((size_t)count);
  std::vector<int> tb((size_t)count);
  long long tiles = 0, footprint = 0;
  bool vec = true;
  const long long s_val = value_kind == MBX_VALUE_F64 ? 8 : value_kind == MBX_VALUE_F32 ? 4 : 2;
  const long long s_vec = value_kind == MBX_VALUE_F64 ? 8 : 4;
  for (int m = 0; m < count; ++m) {
    const mbx_batch_item& it = items[m];
    if (it.nrow < 0 || it.ncol < 0 || it.nnz < 0) return MBX_EINVAL;
    if (it.nrow > 0 && (!it.row_ptr || !it.y)) return MBX_EINVAL;
    if (it.nnz > 0 && (!it.col || !it.val || !it.x)) return MBX_EINVAL;
    const long long path = (long long)it.nrow + it.nnz;
    if (path + T >= (long long)INT32_MAX) return MBX_EOVERFLOW;
    const long long nt = (path + T - 1) / T;
    h[(size_t)m] = BatchMat{it.row_ptr, it.col, it.val, it.x, it.y, it.nrow, it.ncol, it.nnz,
                            (int32_t)tiles, (int32_t)(tiles + m)};
    vec = vec && vec_item_ok(it, value_kind);
    footprint += 4ll * (it.nrow + 1) + (long long)it.nnz * (4 + s_val) +
                 s_vec * ((long long)it.ncol + it.nrow);
    tb[(size_t)m] = (int)tiles;
    tiles += nt;
    if (tiles + count >= (long long)INT32_MAX / 2) return MBX_EOVERFLOW;
  }
  mbx_batch* b = new (std::nothrow) mbx_batch();
  if (!b) return MBX_ENOMEM;
  b->count = count;
  b->kind = value_kind;
  b->vec = vec;
  b->footprint = footprint;
  b->stream_mode = MBX_STREAM_AUTO;
  b->wg_per_cu = 0;
  b->cus = 1;
  b->ipt = ipt;
  b->tile_items = (int)T;
  b->total_tiles = (int)tiles;
  b->ncoord = (int)tiles + count;
  b->max_rows = b->tile_items;
  b->split_rows = 0;
  b->d_mats = nullptr;
  b->d_tile_base = nullptr;
  b->d_coord = nullptr;
  b->d_carry = nullptr;
  b->d_mats_sorted = nullptr;
  b->d_scol = nullptr;
  b->grid_sorted = 0;
  hipStream_t s = (hipStream_t)stream;
  if (hipMalloc((void**)&b->d_mats, sizeof(BatchMat) * (size_t)count) != hipSuccess ||
      hipMalloc((void**)&b->d_tile_base, sizeof(int) * (size_t)count) != hipSuccess ||
      hipMalloc((void**)&b->d_coord, sizeof(int4) * (size_t)b->ncoord) != hipSuccess ||
      hipMalloc((void**)&b->d_carry, sizeof(double) * 2 * (size_t)b->ncoord) != hipSuccess) {
    (void)hipGetLastError();
    mbx_batch_destroy(b);
    return MBX_ENOMEM;
  }
  int device = 0;
  if (hipMemcpyAsync(b->d_mats, h.data(), sizeof(BatchMat) * (size_t)count,
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b->d_tile_base, tb.data(), sizeof(int) * (size_t)count,
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipGetDevice(&device) != hipSuccess) {
    mbx_batch_destroy(b);
    return MBX_EHIP;
  }
  batch_partition_kernel<<<(b->ncoord + 255) / 256, 256, 0, s>>>(b->d_mats, count, b->tile_items,
                                                                  b->ncoord, b->d_coord);
  int* d_stats = reinterpret_cast<int*>(b->d_carry);  // carries are unused until a product
  int h_stats[2] = {0, 0};
  if (hipMemsetAsync(d_stats, 0, 2 * sizeof(int), s) != hipSuccess) {
    mbx_batch_destroy(b);
    return MBX_EHIP;
  }
  if (b->total_tiles > 0)
    batch_max_rows_kernel<<<(b->total_tiles + 255) / 256, 256, 0, s>>>(
        b->d_mats, b->d_tile_base, count, b->total_tiles, b->d_coord, d_stats);
  if (hipMemcpyAsync(h_stats, d_stats, 2 * sizeof(int), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    mbx_batch_destroy(b);
    return MBX_EHIP;
  }
  b->max_rows = h_stats[0];
  b->split_rows = h_stats[1];
  b->grid = batch_run(b, s, 0.0, 0.0, true, device);
  if (hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device) !=
          hipSuccess ||
      b->cus < 1)
    b->cus = 1;
  // the host vectors die here: wait for the copies that read them
  if (hipGetLastError() != hipSuccess || b->grid <= 0 || hipStreamSynchronize(s) != hipSuccess) {
    mbx_batch_destroy(b);
    return MBX_EHIP;
  }
  *out = b;
  return MBX_OK;
}

void mbx_batch_destroy(mbx_batch* b) {
  if (!b) return;
  if (b->d_mats) (void)hipFree(b->d_mats);
  if (b->d_tile_base) (void)hipFree(b->d_tile_base);
  if (b->d_coord) (void)hipFree(b->d_coord);
  if (b->d_carry) (void)hipFree(b->d_carry);
  if (b->d_mats_sorted) (void)hipFree(b->d_mats_sorted);
  if (b->d_scol) (void)hipFree(b->d_scol);
  delete b;
}

int mbx_batch_sort_cols(mbx_batch* b, int32_t enable, void* stream) {
  if (!b) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (b->d_mats_sorted || b->d_scol) {  // products on any stream may still read them
    (void)hipDeviceSynchronize();
    if (b->d_mats_sorted) (void)hipFree(b->d_mats_sorted);
    if (b->d_scol) (void)hipFree(b->d_scol);
    b->d_mats_sorted = nullptr;
    b->d_scol = nullptr;
  }
  if (!enable) return MBX_OK;
  if (b->ipt != 8 && b->ipt != 12 && b->ipt != 16 && b->ipt != 32) return MBX_EUNSUPPORTED;
  if (!b->vec) return MBX_EUNSUPPORTED;  // the scalar staging path reads the columns as given
  const int sb = b->ipt == 8    ? mbx_detail::sort_slot_bits<256, 8>()
                 : b->ipt == 32 ? mbx_detail::sort_slot_bits<512, 16>()
                                : mbx_detail::sort_slot_bits<256, 16>();  // 12 for IPT 12 too
  std::vector<BatchMat> h((size_t)b->count);
  if (hipMemcpy(h.data(), b->d_mats, sizeof(BatchMat) * (size_t)b->count,
                hipMemcpyDeviceToHost) != hipSuccess)
    return MBX_EHIP;
  size_t total = 0;
  std::vector<size_t> off((size_t)b->count);
  for (int m = 0; m < b->count; ++m) {
    if ((long long)h[(size_t)m].ncol > (1ll << (32 - sb))) return MBX_EUNSUPPORTED;
    off[(size_t)m] = total;
    total += ((size_t)h[(size_t)m].nnz + 3) / 4 * 4;  // whole 16-byte groups
  }
  if (b->total_tiles == 0 || total == 0) return MBX_OK;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return MBX_EHIP;
  const int grid = batch_run(b, s, 0.0, 0.0, true, device, true);
  if (grid <= 0) return MBX_EHIP;
  BatchMat* d_sorted = nullptr;
  int32_t* sc = nullptr;
  int* d_bad = nullptr;
  if (hipMalloc((void**)&d_sorted, sizeof(BatchMat) * (size_t)b->count) != hipSuccess ||
      hipMalloc((void**)&sc, sizeof(int32_t) * total) != hipSuccess ||
      hipMalloc((void**)&d_bad, sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    if (d_sorted) (void)hipFree(d_sorted);
    if (sc) (void)hipFree(sc);
    return MBX_ENOMEM;
  }
  std::vector<BatchMat> hs = h;
  for (int m = 0; m < b->count; ++m) hs[(size_t)m].col = sc + off[(size_t)m];
  int st = MBX_OK, bad = 0;
  if (hipMemcpyAsync(d_sorted, hs.data(), sizeof(BatchMat) * (size_t)b->count,
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(sc, 0, sizeof(int32_t) * total, s) != hipSuccess ||
      hipMemsetAsync(d_bad, 0, sizeof(int), s) != hipSuccess)
    st = MBX_EHIP;
  if (st == MBX_OK) {
    switch (b->ipt) {
      case 8:
        batch_sort_cols_kernel<256, 8><<<b->total_tiles, 256, 0, s>>>(
            b->d_mats, d_sorted, b->d_tile_base, b->count, b->d_coord, d_bad);
        break;
      case 12:
        batch_sort_cols_kernel<256, 12><<<b->total_tiles, 256, 0, s>>>(
            b->d_mats, d_sorted, b->d_tile_base, b->count, b->d_coord, d_bad);
        break;
      case 16:
        batch_sort_cols_kernel<256, 16><<<b->total_tiles, 256, 0, s>>>(
            b->d_mats, d_sorted, b->d_tile_base, b->count, b->d_coord, d_bad);
        break;
      default:  // 32: tiles of 512 threads x 16 items
        batch_sort_cols_kernel<512, 16><<<b->total_tiles, 512, 0, s>>>(
            b->d_mats, d_sorted, b->d_tile_base, b->count, b->d_coord, d_bad);
        break;
    }
    // the host table dies here: wait for the copy that reads it
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = MBX_EHIP;
  }
  (void)hipFree(d_bad);
  if (st == MBX_OK && bad) st = MBX_EINVAL;  // a column outside [0, ncol)
  if (st != MBX_OK) {
    (void)hipFree(d_sorted);
    (void)hipFree(sc);
    return st;
  }
  b->d_mats_sorted = d_sorted;
  b->d_scol = sc;
  b->grid_sorted = grid;
  return MBX_OK;
}

int mbx_batch_info(const mbx_batch* b, int32_t* total_tiles, int32_t* tile_items,
                   int32_t* grid) {
  if (!b) return MBX_EINVAL;
  if (total_tiles) *total_tiles = b->total_tiles;
  if (tile_items) *tile_items = b->tile_items;
  if (grid) *grid = std::min(b->grid, b->total_tiles);
  return MBX_OK;
}

int mbx_batch_set_residency(mbx_batch* b, int32_t wg_per_cu) {
  if (!b || wg_per_cu < 0 || wg_per_cu > 16) return MBX_EINVAL;
  b->wg_per_cu = wg_per_cu;
  return MBX_OK;
}

int mbx_batch_set_stream(mbx_batch* b, int32_t mode) {
  if (!b || mode < MBX_STREAM_AUTO || mode > MBX_STREAM_NONTEMPORAL) return MBX_EINVAL;
  b->stream_mode = mode;
  return MBX_OK;
}

int mbx_batch_spmv(const mbx_batch* b, double alpha, double beta, void* stream) {
  if (!b) return MBX_EINVAL;
  if (b->total_tiles == 0) return MBX_OK;
  if (batch_run(b, (hipStream_t)stream, alpha, beta, false, 0) != 0) return MBX_EUNSUPPORTED;
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

int mbx_batch_graph_create(const mbx_batch* b, double alpha, double beta, int32_t repeats,
                           void** graph_exec) {
  if (!b) return MBX_EINVAL;
  return mbx_detail::capture_graph(
      repeats, [&](hipStream_t s) { return mbx_batch_spmv(b, alpha, beta, s); }, graph_exec);
}

int mbx_spmv_graph_create(const mbx_plan* plan, int32_t value_kind, const int32_t* d_row_ptr,
                          const int32_t* d_col, const void* d_val, double alpha,
                          const void* d_x, double beta, void* d_y, int32_t repeats,
                          void** graph_exec) {
  if (!plan) return MBX_EINVAL;
  auto fn = [&](hipStream_t s) -> int {
    switch (value_kind) {
      case MBX_VALUE_F64:
        return mbx_spmv_f64(plan, d_row_ptr, d_col, (const double*)d_val, alpha,
                            (const double*)d_x, beta, (double*)d_y, s);
      case MBX_VALUE_F32:
        return mbx_spmv_f32(plan, d_row_ptr, d_col, (const float*)d_val, (float)alpha,
                            (const float*)d_x, (float)beta, (float*)d_y, s);
      case MBX_VALUE_BF16:
        return mbx_spmv_bf16(plan, d_row_ptr, d_col, (const uint16_t*)d_val, (float)alpha,
                             (const float*)d_x, (float)beta, (float*)d_y, s);
      default:
        return MBX_EINVAL;
    }
  };
  return mbx_detail::capture_graph(repeats, fn, graph_exec);
}

int mbx_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return MBX_EINVAL;
  return hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream) == hipSuccess
             ? MBX_OK
             : MBX_EHIP;
}

void mbx_graph_destroy(void* graph_exec) {
  if (graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec);
}

}  // extern "C"

// probes.hip -- bandwidth probes for the headline product (libmerbit_probe.so).
// Not part of the product path: the benchmark times them beside the FP64
// tile kernel, at its launch shape (one workgroup of 256 threads per tile of
// a plan's 2048 or 3072 merge items, XCD-contiguous tile order, the plan's
// residency cap and non-temporal row_ptr/col loads), to show what the two
// halves of the tile's memory work cost on their own:
//   stream probe  the tile's row_ptr, col and value groups, loaded exactly as
//                 the tile kernel stages them, summed per workgroup; the
//                 rows ending in the tile are written (the y stream);
//   gather probe  the tile's column groups and the x gathers of its
//                 nonzeros through the same buffer resource, summed per
//                 workgroup and written once per tile; on a plan with
//                 tile-sorted columns (mbx_plan_sort_cols) the plan's sorted
//                 words, gathered in their order as the product does.
// Each writes what it reads into a result so no load is dead; the results
// themselves are not meaningful.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "merbit_amd.h"
#include "merbit_internal.hpp"
#include "merbit_tile.hpp"

namespace {

using P = mbx_detail::Streamed<mbx_detail::F64>;
using mbx_detail::TileRegs;

__device__ __forceinline__ double block_sum(double v, double* s_w) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  return s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

template <int IPT>
__global__ __launch_bounds__(256) void stream_probe_kernel(const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ col,
                                                           const double* __restrict__ val,
                                                           const int4* __restrict__ coord,
                                                           int ntiles, double* __restrict__ y) {
  __shared__ double s_w[4];
  const int tile = mbx_detail::xcd_tile(blockIdx.x, ntiles);
  TileRegs<256, IPT, P> R{};  // groups no thread loads stay 0
  mbx_detail::tile_coords(R, coord, tile);
  mbx_detail::tile_issue_loads(R, rp, col, val);
  constexpr int NG = TileRegs<256, IPT, P>::NG;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < NG; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) s += R.vg[j][v] + (double)(R.cg[j][v] + R.rg[j][v]);
  const double t = block_sum(s, s_w);
  for (int i = threadIdx.x; i < R.rows(); i += 256) y[R.r0() + i] = t;
}

template <int IPT, class Q>
__global__ __launch_bounds__(256) void gather_probe_kernel(const int32_t* __restrict__ col,
                                                           const double* __restrict__ x, int ncol,
                                                           const int4* __restrict__ coord,
                                                           int ntiles, double* __restrict__ y) {
  __shared__ double s_w[4];
  const int tile = mbx_detail::xcd_tile(blockIdx.x, ntiles);
  constexpr int NG = TileRegs<256, IPT, Q>::NG;
  TileRegs<256, IPT, Q> R;
  mbx_detail::tile_coords(R, coord, tile);
  const int k0 = R.k0(), nnz_t = R.nnz();
  double s = 0.0;
  if (nnz_t > 0) {
    const int kb = k0 >> 2, klast = (k0 + nnz_t - 1) >> 2, nkg = klast - kb + 1;
#pragma unroll
    for (int j = 0; j < NG; ++j)
      if (j == 0 || j * 256 < nkg)
        mbx_detail::load_group<Q::kIndexNT>(col, min(kb + (int)threadIdx.x + j * 256, klast),
                                            R.cg[j]);
    mbx_detail::tile_issue_gathers(R, x, ncol);
#pragma unroll
    for (int j = 0; j < NG; ++j)
      if (j == 0 || j * 256 < nkg)
#pragma unroll
        for (int v = 0; v < 4; ++v) s += R.xv[j][v];
  }
  const double t = block_sum(s, s_w);
  if (threadIdx.x == 0) y[tile] = t;
}

template <int IPT>
int launch(const mbx_plan* p, int which, const int32_t* rp, const int32_t* col,
           const double* val, const double* x, double* y, hipStream_t s) {
  const size_t lds = mbx_detail::residency_lds(64, p->wg_per_cu);
  if (which == 0) {
    mbx_detail::allow_lds(stream_probe_kernel<IPT>, lds);
    stream_probe_kernel<IPT><<<p->ntiles, 256, lds, s>>>(rp, col, val, p->d_coord, p->ntiles, y);
  } else if (p->d_scol) {
    using Q = mbx_detail::SortX<P>;
    mbx_detail::allow_lds(gather_probe_kernel<IPT, Q>, lds);
    gather_probe_kernel<IPT, Q><<<p->ntiles, 256, lds, s>>>(p->d_scol, x, p->ncol, p->d_coord,
                                                            p->ntiles, y);
  } else {
    mbx_detail::allow_lds(gather_probe_kernel<IPT, P>, lds);
    gather_probe_kernel<IPT, P><<<p->ntiles, 256, lds, s>>>(col, x, p->ncol, p->d_coord,
                                                            p->ntiles, y);
  }
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

}  // namespace

extern "C" {

// which: 0 = stream probe (y: nrow doubles), 1 = gather probe (y: ntiles
// doubles).  The plan must have 8, 12 or 16 items per thread; its residency
// cap applies.  row_ptr, col and val 16-byte aligned.
int mbp_probe(const mbx_plan* p, int32_t which, const int32_t* rp, const int32_t* col,
              const double* val, const double* x, double* y, void* stream) {
  if (!p || !y || (which != 0 && which != 1)) return MBX_EINVAL;
  if (p->ntiles == 0) return MBX_OK;
  if (!rp || (p->nnz > 0 && (!col || (which == 0 && !val) || (which == 1 && !x))))
    return MBX_EINVAL;
  if (!mbx_detail::vec_path_ok(rp, col, val ? val : reinterpret_cast<const double*>(col),
                               p->ncol))
    return MBX_EUNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  if (p->ipt == 8) return launch<8>(p, which, rp, col, val, x, y, s);
  if (p->ipt == 12) return launch<12>(p, which, rp, col, val, x, y, s);
  if (p->ipt == 16) return launch<16>(p, which, rp, col, val, x, y, s);
  return MBX_EUNSUPPORTED;
}

}  // extern "C"

// merbit_pagerank.hip -- the reference's iterative driver with its damping
// factor put to use: power iterations of PageRank on the matrix the reference
// operates on (CSR of P^T, P = row-stochastic transition matrix of the loaded
// graph, values 1/out-degree; src/class.cu:22-34, include/tile.cuh:571-574).
// The reference carries the factor in PARAMETER::damp (include/class.cuh:
// 119-146) but its loop repeats y = A*x with x fixed (tile.cuh:596-603).
//
// One iteration, for x of total mass s = sum(x):
//   y  = damp * A x                     (merge-path tile SpMV, alpha = damp)
//   x' = y + (sum(x) - sum(y)) / n
// sum(y) = damp * (mass on nodes with out-degree > 0), so the uniform term is
// the teleport share (1-damp)*s/n plus the damped mass of dangling nodes
// spread evenly; x' keeps the mass s.
//
// Both sums are deterministic: every block sums a fixed grid-stride set in a
// fixed order, and every block of the update kernel reduces the same
// partials in the same order, so all of them add the same constant.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxParts = 1024;  // partial-sum blocks (a multiple of 64)

int part_blocks(int n) {
  const long long b = ((long long)n + 4 * kBlock - 1) / (4 * kBlock);
  return (int)(b < 1 ? 1 : (b > kMaxParts ? kMaxParts : b));
}

// part[b] = sum of x over block b's grid-stride set, part[kMaxParts+b] the
// same of y; fixed order within each thread, then a fixed LDS tree.
__global__ __launch_bounds__(kBlock) void pr_partial_kernel(const double* __restrict__ x,
                                                            const double* __restrict__ y, int n,
                                                            double* __restrict__ part) {
  __shared__ double sx[kBlock], sy[kBlock];
  const int t = threadIdx.x;
  double ax = 0.0, ay = 0.0;
  for (long long i = (long long)blockIdx.x * kBlock + t; i < n; i += (long long)gridDim.x * kBlock) {
    ax += x[i];
    ay += y[i];
  }
  sx[t] = ax;
  sy[t] = ay;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      sx[t] += sx[t + w];
      sy[t] += sy[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x] = sx[0];
    part[kMaxParts + blockIdx.x] = sy[0];
  }
}

// y[i] += (sum(x) - sum(y)) / n, the sums rebuilt from `nparts` partials by
// the first wave of every block (same order everywhere).
__global__ __launch_bounds__(kBlock) void pr_update_kernel(double* __restrict__ y, int n,
                                                           const double* __restrict__ part,
                                                           int nparts) {
  __shared__ double s_c;
  const int t = threadIdx.x;
  if (t < 64) {
    double d = 0.0;
    for (int i = t; i < nparts; i += 64) d += part[i] - part[kMaxParts + i];
    for (int m = 32; m > 0; m >>= 1) d += __shfl_xor(d, m);
    if (t == 0) s_c = d / (double)n;
  }
  __syncthreads();
  const double c = s_c;
  for (long long i = (long long)blockIdx.x * kBlock + t; i < n; i += (long long)gridDim.x * kBlock)
    y[i] += c;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

size_t mbx_pagerank_workspace_bytes(int32_t nrow) {
  if (nrow < 0) return 0;
  return align256(sizeof(double) * (size_t)(nrow > 0 ? nrow : 1)) +
         align256(sizeof(double) * 2 * kMaxParts);
}

int mbx_pagerank_f64(const mbx_plan* plan, const int32_t* d_row_ptr, const int32_t* d_col,
                     const double* d_val, double damp, int32_t iters, const double* d_x0,
                     double* d_y, void* d_work, size_t work_bytes, void* stream) {
  if (!mbx_detail::plan_csr_args_ok(plan, d_row_ptr, d_col, d_val)) return MBX_EINVAL;
  const int n = plan->nrow;
  if (plan->ncol != n || iters < 1 || !(damp >= 0.0 && damp <= 1.0)) return MBX_EINVAL;
  if (n == 0) return MBX_OK;
  if (!d_x0 || !d_y || !d_work || work_bytes < mbx_pagerank_workspace_bytes(n))
    return MBX_EINVAL;
  if ((reinterpret_cast<uintptr_t>(d_work) & 255u) != 0) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  double* other = (double*)d_work;
  double* part = (double*)((char*)d_work + align256(sizeof(double) * (size_t)n));
  const int nb = part_blocks(n);
  const int ub = (int)(((long long)n + kBlock - 1) / kBlock < 4096
                           ? ((long long)n + kBlock - 1) / kBlock
                           : 4096);
  const double* src = d_x0;
  for (int t = 0; t < iters; ++t) {
    // the last iteration writes d_y; earlier ones alternate so that holds
    double* dst = ((iters - 1 - t) % 2 == 0) ? d_y : other;
    int st = mbx_spmv_f64(plan, d_row_ptr, d_col, d_val, damp, src, 0.0, dst, stream);
    if (st != MBX_OK) return st;
    pr_partial_kernel<<<nb, kBlock, 0, s>>>(src, dst, n, part);
    pr_update_kernel<<<ub, kBlock, 0, s>>>(dst, n, part, nb);
    src = dst;
  }
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

int mbx_pagerank_graph_create(const mbx_plan* plan, const int32_t* d_row_ptr,
                              const int32_t* d_col, const double* d_val, double damp,
                              int32_t iters, const double* d_x0, double* d_y, void* d_work,
                              size_t work_bytes, void** graph_exec) {
  if (!plan) return MBX_EINVAL;
  return mbx_detail::capture_graph(
      1,
      [&](hipStream_t s) {
        return mbx_pagerank_f64(plan, d_row_ptr, d_col, d_val, damp, iters, d_x0, d_y, d_work,
                                work_bytes, s);
      },
      graph_exec);
}

}  // extern "C"

// merbit_colblock.hip -- column-blocked form of a CSR pattern for products
// whose x gathers are spread evenly over a long x (random graphs).
//
// The merge-path product gathers x[col] once per nonzero.  When x is larger
// than one XCD's 4 MB L2 and the columns have no hot set (a random graph:
// 2.4M FP32 x values, 9.6 MB), most gathers miss L2 and the product is bound
// by line fills from the MALL, not by HBM.  Splitting the columns into B
// ranges and the matrix into B CSR matrices over all rows,
//   A = A_0 + A_1 + ... + A_{B-1},  A_j keeping the columns [lo_j, lo_{j+1}),
// lets each product A_j x_j gather from a slice of x that stays in L2.  The
// products are chained on one stream (y = alpha*A_0 x_0 + beta*y, then
// y = alpha*A_j x_j + y), each a plain merge-path SpMV over its own plan.
//
// The B matrices live in one block-major nonzero order: block after block,
// rows in order within a block, entries of a row in CSR order.  Block starts
// are padded to 4 entries so every block's col and values keep the 16-byte
// alignment of the vector staging path.  The pattern is converted once
// (mbx_colblock_create); values are moved into block-major order with
// mbx_colblock_values whenever they change, so the form pays when one set of
// values serves many products.  A row's sum is split over the blocks and
// added in block order: results are deterministic, but rounded differently
// from the single CSR product (FP64 agrees to a few units in the last place).
//
// The adjoint dx = A^T g gathers g by A's rows, so it is blocked the other
// way: mbx_colblock_adjoint_prepare builds the CSC of A (the adjoint's
// radix sort), a column-blocked form of A^T over it (blocks of A's rows) and
// the adjoint's two-pass bucketed permutation from this form's slots to the
// A^T form's slots: a bijection of [0, slots), because this form keeps spare
// slots after its last block that, with the padding of both forms, pair up
// in order.  dx is that one permutation and the A^T form's chained product.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <new>
#include <vector>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

constexpr int kMaxBlocks = 16;
// Spare slots after the last block: the A^T form of a pattern pads at most
// 3 slots per block, so this form's slot count always covers it and one
// bijection of [0, slots) maps this form's order onto the A^T form's.
constexpr int kReserve = 4 * kMaxBlocks;
constexpr int kDefaultTilesPerCu = 4;

struct mbx_colblock {
  int32_t nrow = 0, ncol = 0, nnz = 0, nb = 0;
  int32_t stride = 0;                 // row_ptr entries per block (nrow+1 rounded to 4)
  int32_t stream_mode = MBX_STREAM_AUTO;
  int32_t lo[kMaxBlocks + 1] = {};    // column ranges
  int64_t base[kMaxBlocks + 1] = {};  // padded first slot of each block
  int64_t slots = 0;                  // value slots: base[nb] plus kReserve spare ones
  int32_t* row_ptr = nullptr;         // nb * stride: block j's row pointer at j * stride
  int32_t* col = nullptr;             // slots: column minus lo_j, block-major
  int32_t* perm = nullptr;            // slots: CSR position of each slot, -1 for padding
  mbx_plan* plan[kMaxBlocks] = {};
  // adjoint (mbx_colblock_adjoint_prepare): the column-blocked form of A^T
  // and the permutation that moves this form's values into its order
  mbx_colblock* t = nullptr;
  mbx_adjoint* tp = nullptr;  // bucketed permutation: this form's slots -> the A^T form's
  int32_t nnz_b[kMaxBlocks] = {};
  bool sorted = false;  // mbx_colblock_sort_cols: the block plans gather in sorted order
};

namespace {

struct Bounds {
  int32_t v[kMaxBlocks + 1];
};

__device__ __forceinline__ int block_of(int c, int nb, int ncol) {
  const long long j = ((long long)c * nb) / ncol;
  return (int)(j < 0 ? 0 : (j >= nb ? nb - 1 : j));
}

// One thread per row: entries of the row in each column block.
__global__ void cb_count_kernel(const int32_t* __restrict__ row_ptr,
                                const int32_t* __restrict__ col, int nrow, int ncol, int nb,
                                int stride, int32_t* __restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  int c16[kMaxBlocks];
#pragma unroll
  for (int j = 0; j < kMaxBlocks; ++j) c16[j] = 0;
  for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
    const int j = block_of(col[k], nb, ncol);
#pragma unroll
    for (int q = 0; q < kMaxBlocks; ++q) c16[q] += (q == j);
  }
  for (int j = 0; j < nb; ++j) cnt[(long long)j * stride + r] = c16[j];
}

// Exclusive-scanned counts -> each block's own row pointer (starting at 0).
__global__ void cb_rebase_kernel(const int32_t* __restrict__ scanned, int nrow, int nb,
                                 int stride, Bounds start, int32_t* __restrict__ row_ptr) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nb * stride) return;
  const int j = (int)(i / stride), r = (int)(i - (long long)j * stride);
  row_ptr[i] = r <= nrow ? scanned[i] - start.v[j] : 0;
}

// One thread per row: place the row's entries in its blocks' rows.
__global__ void cb_scatter_kernel(const int32_t* __restrict__ row_ptr,
                                  const int32_t* __restrict__ col, int nrow, int ncol, int nb,
                                  int stride, Bounds lo, Bounds base,
                                  const int32_t* __restrict__ brp, int32_t* __restrict__ bcol,
                                  int32_t* __restrict__ perm) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  int seen[kMaxBlocks];
#pragma unroll
  for (int j = 0; j < kMaxBlocks; ++j) seen[j] = 0;
  for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
    const int c = col[k];
    const int j = block_of(c, nb, ncol);
    int s = 0;
#pragma unroll
    for (int q = 0; q < kMaxBlocks; ++q)
      if (q == j) s = seen[q]++;
    const long long p = (long long)base.v[j] + brp[(long long)j * stride + r] + s;
    bcol[p] = c - lo.v[j];
    perm[p] = k;
  }
}

template <class T>
__global__ void cb_values_kernel(const int32_t* __restrict__ perm, long long slots,
                                 const T* __restrict__ val, T* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots) return;
  const int k = perm[i];
  out[i] = k >= 0 ? val[k] : T(0);
}

// Block-major values back to CSR order: out[perm[i]] = in[i] on real slots.
template <class T>
__global__ void cb_to_csr_kernel(const int32_t* __restrict__ perm, long long slots,
                                 const T* __restrict__ in, T* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots) return;
  const int k = perm[i];
  if (k >= 0) out[k] = in[i];
}

// Inverse of the slot -> CSR position map over the real slots.
__global__ void cb_inverse_kernel(const int32_t* __restrict__ perm, long long slots,
                                  int32_t* __restrict__ inv) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots) return;
  const int k = perm[i];
  if (k >= 0) inv[k] = (int32_t)i;
}

// Slot q of the A^T form holds CSC entry tperm[q] (or padding, -1), that is
// CSR entry csc_perm[.], which sits in slot inv[.] of this form; *bad is set
// if some entry has no slot (a pattern other than the one the form was built
// from).  Padding slots are filled by cb_pair_kernel.
__global__ void cb_real_slot_kernel(const int32_t* __restrict__ tperm, long long tslots,
                                    const int32_t* __restrict__ csc_perm,
                                    const int32_t* __restrict__ inv, int32_t* __restrict__ out,
                                    int* __restrict__ bad) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= tslots) return;
  const int p = tperm[q];
  if (p < 0) return;
  const int v = inv[csc_perm[p]];
  out[q] = v < 0 ? 0 : v;
  if (v < 0) atomicOr(bad, 1);
}

// Whether (rp, col) is the pattern the form was built from: every row has
// the length of its blocks' rows together (with rp[0] == 0, so the row
// pointers agree) ...
__global__ void cb_same_rows_kernel(const int32_t* __restrict__ rp,
                                    const int32_t* __restrict__ brp, int nrow, int nb,
                                    int stride, int* __restrict__ bad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  int len = 0;
  for (int j = 0; j < nb; ++j)
    len += brp[(long long)j * stride + r + 1] - brp[(long long)j * stride + r];
  if (rp[r + 1] - rp[r] != len || (r == 0 && rp[0] != 0)) atomicOr(bad, 1);
}

// ... and every real slot i of block j holds the column col[perm[i]] - lo_j.
__global__ void cb_same_cols_kernel(const int32_t* __restrict__ col,
                                    const int32_t* __restrict__ perm,
                                    const int32_t* __restrict__ bcol, long long slots, int nb,
                                    Bounds lo, Bounds base, int* __restrict__ bad) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots) return;
  const int k = perm[i];
  if (k < 0) return;
  int j = 0;
  while (j + 1 < nb && i >= base.v[j + 1]) ++j;
  if (col[k] != bcol[i] + lo.v[j]) atomicOr(bad, 1);
}

// out[dst[i]] = src[i]: the spare and padding slots of both forms, paired in
// order.
__global__ void cb_pair_kernel(const int32_t* __restrict__ dst, const int32_t* __restrict__ src,
                               int n, int32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[dst[i]] = src[i];
}

void destroy(mbx_colblock* cb);

// MBX_EINVAL unless (rp, col) is the pattern cb was built from (the same
// nnz is the caller's check); synchronises s.
int same_pattern(const mbx_colblock* cb, const int32_t* rp, const int32_t* col, hipStream_t s) {
  int* d_bad = nullptr;
  if (hipMalloc((void**)&d_bad, sizeof(int)) != hipSuccess) return MBX_ENOMEM;
  int st = MBX_OK, bad = 0;
  if (hipMemsetAsync(d_bad, 0, sizeof(int), s) != hipSuccess) st = MBX_EHIP;
  if (st == MBX_OK && cb->nrow > 0)
    cb_same_rows_kernel<<<(cb->nrow + 255) / 256, 256, 0, s>>>(rp, cb->row_ptr, cb->nrow, cb->nb,
                                                               cb->stride, d_bad);
  if (st == MBX_OK && cb->nnz > 0) {
    Bounds lo, base;
    for (int j = 0; j <= kMaxBlocks; ++j) {
      lo.v[j] = j <= cb->nb ? cb->lo[j] : 0;
      base.v[j] = j <= cb->nb ? (int32_t)cb->base[j] : 0;
    }
    cb_same_cols_kernel<<<(int)((cb->slots + 255) / 256), 256, 0, s>>>(
        col, cb->perm, cb->col, cb->slots, cb->nb, lo, base, d_bad);
  }
  if (st == MBX_OK &&
      (hipGetLastError() != hipSuccess ||
       hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    st = MBX_EHIP;
  (void)hipFree(d_bad);
  if (st == MBX_OK && bad) st = MBX_EINVAL;
  return st;
}

void drop_adjoint(mbx_colblock* cb) {
  if (cb->t) destroy(cb->t);
  mbx_detail::perm_free(cb->tp);
  cb->t = nullptr;
  cb->tp = nullptr;
}

void destroy(mbx_colblock* cb) {
  if (!cb) return;
  (void)hipDeviceSynchronize();
  drop_adjoint(cb);
  for (int j = 0; j < kMaxBlocks; ++j)
    if (cb->plan[j]) mbx_plan_destroy(cb->plan[j]);
  if (cb->row_ptr) (void)hipFree(cb->row_ptr);
  if (cb->col) (void)hipFree(cb->col);
  if (cb->perm) (void)hipFree(cb->perm);
  delete cb;
}

int value_bytes(int32_t kind) {
  switch (kind) {
    case MBX_VALUE_F64: return 8;
    case MBX_VALUE_F32: return 4;
    case MBX_VALUE_BF16: return 2;
    default: return 0;
  }
}

int build(mbx_colblock* cb, const int32_t* rp, const int32_t* col, int ipt, hipStream_t s) {
  const int nrow = cb->nrow, nb = cb->nb, stride = cb->stride;
  const long long n = (long long)nb * stride;
  int32_t* cnt = nullptr;
  int32_t* scanned = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int st = MBX_OK;
  if (hipMalloc(&cnt, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc(&scanned, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc(&cb->row_ptr, sizeof(int32_t) * n) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK && hipMemsetAsync(cnt, 0, sizeof(int32_t) * n, s) != hipSuccess) st = MBX_EHIP;
  if (st == MBX_OK && nrow > 0)
    cb_count_kernel<<<(nrow + 255) / 256, 256, 0, s>>>(rp, col, nrow, cb->ncol, nb, stride, cnt);
  if (st == MBX_OK &&
      hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, scanned, (int)n, s) != hipSuccess)
    st = MBX_EHIP;
  if (st == MBX_OK && hipMalloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1) != hipSuccess)
    st = MBX_ENOMEM;
  if (st == MBX_OK &&
      hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, scanned, (int)n, s) != hipSuccess)
    st = MBX_EHIP;
  // block starts in the unpadded order, read back to size and pad the blocks,
  // and the total, which must be the nnz the caller gave (row_ptr[nrow])
  int32_t h_start[kMaxBlocks + 1] = {};
  for (int j = 0; j <= nb && st == MBX_OK; ++j) {
    const long long at = j < nb ? (long long)j * stride : (long long)(nb - 1) * stride + nrow;
    if (hipMemcpyAsync(&h_start[j], scanned + at, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
        hipSuccess)
      st = MBX_EHIP;
  }
  if (st == MBX_OK && hipStreamSynchronize(s) != hipSuccess) st = MBX_EHIP;
  if (st == MBX_OK && h_start[nb] != cb->nnz) st = MBX_EINVAL;
  Bounds start{}, lo{}, base{};
  int32_t nnz_b[kMaxBlocks] = {};
  if (st == MBX_OK) {
    for (int j = 0; j < nb; ++j) {
      start.v[j] = h_start[j];
      nnz_b[j] = (j + 1 < nb ? h_start[j + 1] : cb->nnz) - h_start[j];
      lo.v[j] = cb->lo[j];
    }
    cb->base[0] = 0;
    for (int j = 0; j < nb; ++j) cb->base[j + 1] = cb->base[j] + (nnz_b[j] + 3) / 4 * 4;
    for (int j = 0; j < nb; ++j) cb->nnz_b[j] = nnz_b[j];
    cb->slots = cb->base[nb] + kReserve;
    if (cb->slots > INT32_MAX) st = MBX_EOVERFLOW;
    for (int j = 0; j <= nb; ++j) base.v[j] = (int32_t)cb->base[j];
  }
  const long long slots = cb->slots;
  if (st == MBX_OK &&
      (hipMalloc(&cb->col, sizeof(int32_t) * (slots > 0 ? slots : 1)) != hipSuccess ||
       hipMalloc(&cb->perm, sizeof(int32_t) * (slots > 0 ? slots : 1)) != hipSuccess))
    st = MBX_ENOMEM;
  if (st == MBX_OK && slots > 0 &&
      (hipMemsetAsync(cb->col, 0, sizeof(int32_t) * slots, s) != hipSuccess ||
       hipMemsetAsync(cb->perm, 0xFF, sizeof(int32_t) * slots, s) != hipSuccess))
    st = MBX_EHIP;
  if (st == MBX_OK) {
    cb_rebase_kernel<<<(int)((n + 255) / 256), 256, 0, s>>>(scanned, nrow, nb, stride, start,
                                                             cb->row_ptr);
    if (nrow > 0)
      cb_scatter_kernel<<<(nrow + 255) / 256, 256, 0, s>>>(rp, col, nrow, cb->ncol, nb, stride,
                                                           lo, base, cb->row_ptr, cb->col,
                                                           cb->perm);
    if (hipGetLastError() != hipSuccess) st = MBX_EHIP;
  }
  for (int j = 0; j < nb && st == MBX_OK; ++j)
    st = mbx_plan_create(cb->row_ptr + (long long)j * stride, nrow, cb->lo[j + 1] - cb->lo[j],
                         nnz_b[j], ipt, 256, s, &cb->plan[j]);
  if (hipStreamSynchronize(s) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  if (tmp) (void)hipFree(tmp);
  if (scanned) (void)hipFree(scanned);
  if (cnt) (void)hipFree(cnt);
  return st;
}

}  // namespace

namespace mbx_detail {

std::vector<int32_t> colblock_padding(const mbx_colblock* cb) {
  std::vector<int32_t> pad;
  for (int j = 0; j < cb->nb; ++j)
    for (long long q = cb->base[j] + cb->nnz_b[j]; q < cb->base[j + 1]; ++q)
      pad.push_back((int32_t)q);
  return pad;
}

const int32_t* colblock_slot_map(const mbx_colblock* cb) { return cb->perm; }

}  // namespace mbx_detail

extern "C" {

int mbx_colblock_create(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                        int32_t ncol, int32_t nnz, int32_t nblocks, int32_t items_per_thread,
                        void* stream, mbx_colblock** out) {
  if (!out) return MBX_EINVAL;
  *out = nullptr;
  if (nrow < 0 || ncol < 0 || nnz < 0 || nblocks < 0 || nblocks > kMaxBlocks) return MBX_EINVAL;
  if (!d_row_ptr || (nnz > 0 && !d_col)) return MBX_EINVAL;
  if (nnz > 0 && ncol == 0) return MBX_EINVAL;
  int nb = nblocks;
  if (nb == 0) {  // slices of at most 3.5 MiB of 4-byte x (one XCD's L2 is 4 MB)
    const long long slice = (7ll << 20) / 2 / 4;
    nb = (int)((ncol + slice - 1) / slice);
    nb = nb < 1 ? 1 : (nb > kMaxBlocks ? kMaxBlocks : nb);
  }
  if (ncol > 0 && nb > ncol) nb = ncol;
  if (nb < 1) nb = 1;
  const long long stride = ((long long)nrow + 1 + 3) / 4 * 4;
  if (stride * nb > INT32_MAX) return MBX_EOVERFLOW;
  mbx_colblock* cb = new (std::nothrow) mbx_colblock;
  if (!cb) return MBX_ENOMEM;
  cb->nrow = nrow;
  cb->ncol = ncol;
  cb->nnz = nnz;
  cb->nb = nb;
  cb->stride = (int32_t)stride;
  for (int j = 0; j <= nb; ++j)  // lo_j = ceil(j * ncol / nb): c is in block floor(c * nb / ncol)
    cb->lo[j] = (int32_t)(((long long)j * ncol + nb - 1) / nb);
  const int st = build(cb, d_row_ptr, d_col, items_per_thread, (hipStream_t)stream);
  if (st != MBX_OK) {
    destroy(cb);
    return st;
  }
  // the block products are gather-bound: 4 tiles per CU measured fastest for
  // 3 to 6 blocks of a random graph (mbx_colblock_set_residency changes it)
  (void)mbx_colblock_set_residency(cb, kDefaultTilesPerCu);
  *out = cb;
  return MBX_OK;
}

void mbx_colblock_destroy(mbx_colblock* cb) { destroy(cb); }

int mbx_colblock_info(const mbx_colblock* cb, int32_t* nblocks, int64_t* value_slots,
                      int32_t* col_bounds) {
  if (!cb) return MBX_EINVAL;
  if (nblocks) *nblocks = cb->nb;
  if (value_slots) *value_slots = cb->slots;
  if (col_bounds)
    for (int j = 0; j <= cb->nb; ++j) col_bounds[j] = cb->lo[j];
  return MBX_OK;
}

int mbx_colblock_set_stream(mbx_colblock* cb, int32_t mode) {
  if (!cb || mode < MBX_STREAM_AUTO || mode > MBX_STREAM_NONTEMPORAL) return MBX_EINVAL;
  cb->stream_mode = mode;
  if (cb->t) cb->t->stream_mode = mode;
  return MBX_OK;
}

int mbx_colblock_set_residency(mbx_colblock* cb, int32_t wg_per_cu) {
  if (!cb || wg_per_cu < 0 || wg_per_cu > 16) return MBX_EINVAL;
  for (int j = 0; j < cb->nb; ++j) mbx_plan_set_residency(cb->plan[j], wg_per_cu);
  if (cb->t) mbx_colblock_set_residency(cb->t, wg_per_cu);
  return MBX_OK;
}

int mbx_colblock_sort_cols(mbx_colblock* cb, int32_t enable, void* stream) {
  if (!cb) return MBX_EINVAL;
  for (int j = 0; j < cb->nb; ++j) (void)mbx_plan_sort_cols(cb->plan[j], nullptr, stream);
  if (cb->t) (void)mbx_colblock_sort_cols(cb->t, 0, stream);
  cb->sorted = false;
  if (!enable) return MBX_OK;
  int st = MBX_OK;
  for (int j = 0; j < cb->nb && st == MBX_OK; ++j)
    st = mbx_plan_sort_cols(cb->plan[j], cb->col + cb->base[j], stream);
  if (st == MBX_OK && cb->t) {
    st = mbx_colblock_sort_cols(cb->t, 1, stream);
    if (st == MBX_EUNSUPPORTED) st = MBX_OK;  // the A^T form keeps its plain gathers
  }
  if (st != MBX_OK) {
    (void)mbx_colblock_sort_cols(cb, 0, stream);
    return st;
  }
  cb->sorted = true;
  return MBX_OK;
}

int mbx_colblock_adjoint_prepare(mbx_colblock* cb, const int32_t* d_row_ptr,
                                 const int32_t* d_col, int32_t nblocks,
                                 int32_t items_per_thread, void* stream, int32_t* nblocks_out) {
  if (!cb || nblocks < 0 || nblocks > kMaxBlocks) return MBX_EINVAL;
  if (!d_row_ptr || (cb->nnz > 0 && !d_col)) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (cb->t) (void)hipDeviceSynchronize();  // a dx on any stream may still read them
  drop_adjoint(cb);
  mbx_adjoint* adj = nullptr;
  // validates the pattern itself (row_ptr from 0 to nnz, columns in range)
  int st = mbx_adjoint_create(d_row_ptr, d_col, cb->nrow, cb->ncol, cb->nnz, items_per_thread,
                              stream, &adj);
  if (st == MBX_OK) st = same_pattern(cb, d_row_ptr, d_col, s);
  if (st == MBX_OK)
    st = mbx_colblock_create(adj->col_ptr, adj->row_idx, cb->ncol, cb->nrow, cb->nnz, nblocks,
                             items_per_thread, stream, &cb->t);
  if (st == MBX_OK && cb->sorted) {  // the A^T form gathers as this one does, where it can
    const int ss = mbx_colblock_sort_cols(cb->t, 1, stream);
    if (ss != MBX_EUNSUPPORTED) st = ss;
  }
  // The permutation over [0, slots): destination q (a slot of the A^T form,
  // or one past its last block) reads source slot perm[q] of this form.
  // Real entries pair through the CSC position; the A^T form's padding and
  // the slots past its last block take this form's padding and spare slots
  // in order (equal counts: slots - nnz on each side).
  int32_t* inv = nullptr;
  int* d_bad = nullptr;
  int32_t *d_pd = nullptr, *d_ps = nullptr;
  const int nnz = cb->nnz;
  const long long slots = cb->slots;
  mbx_adjoint* tp = nullptr;
  std::vector<int32_t> pd, ps;
  if (st == MBX_OK) {
    const mbx_colblock* t = cb->t;
    for (int j = 0; j < t->nb; ++j)
      for (long long q = t->base[j] + t->nnz_b[j]; q < t->base[j + 1]; ++q) pd.push_back((int32_t)q);
    for (long long q = t->base[t->nb]; q < slots; ++q) pd.push_back((int32_t)q);
    for (int j = 0; j < cb->nb; ++j)
      for (long long q = cb->base[j] + cb->nnz_b[j]; q < cb->base[j + 1]; ++q)
        ps.push_back((int32_t)q);
    for (long long q = cb->base[cb->nb]; q < slots; ++q) ps.push_back((int32_t)q);
    if (t->base[t->nb] > slots || pd.size() != ps.size()) st = MBX_EOVERFLOW;  // cannot happen
    tp = new (std::nothrow) mbx_adjoint();
    if (!tp && st == MBX_OK) st = MBX_ENOMEM;
  }
  const size_t m = (size_t)slots;
  if (st == MBX_OK) {
    tp->nnz = (int32_t)slots;
    tp->nsrc = (int32_t)slots;
    cb->tp = tp;
    if (hipMalloc((void**)&inv, sizeof(int32_t) * (nnz > 0 ? nnz : 1)) != hipSuccess ||
        hipMalloc((void**)&d_bad, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&d_pd, sizeof(int32_t) * (pd.size() + 1)) != hipSuccess ||
        hipMalloc((void**)&d_ps, sizeof(int32_t) * (ps.size() + 1)) != hipSuccess ||
        hipMalloc((void**)&tp->perm, sizeof(int32_t) * m) != hipSuccess ||
        hipMalloc((void**)&tp->lslot, sizeof(uint16_t) * m) != hipSuccess ||
        hipMalloc((void**)&tp->spos, sizeof(int32_t) * m) != hipSuccess ||
        hipMalloc((void**)&tp->lidx, sizeof(int32_t) * m) != hipSuccess ||
        hipMalloc(&tp->stage, 8 * m) != hipSuccess || hipMalloc(&tp->vt, 8 * m) != hipSuccess)
      st = MBX_ENOMEM;
  } else if (tp) {
    delete tp;
    tp = nullptr;
  }
  if (st == MBX_OK &&
      (hipMemsetAsync(inv, 0xFF, sizeof(int32_t) * (nnz > 0 ? nnz : 1), s) != hipSuccess ||
       hipMemsetAsync(d_bad, 0, sizeof(int), s) != hipSuccess ||
       hipMemsetAsync(tp->perm, 0, sizeof(int32_t) * m, s) != hipSuccess ||
       hipMemcpyAsync(d_pd, pd.data(), sizeof(int32_t) * pd.size(), hipMemcpyHostToDevice, s) !=
           hipSuccess ||
       hipMemcpyAsync(d_ps, ps.data(), sizeof(int32_t) * ps.size(), hipMemcpyHostToDevice, s) !=
           hipSuccess))
    st = MBX_EHIP;
  if (st == MBX_OK) {
    const long long treal = cb->t->base[cb->t->nb];
    cb_inverse_kernel<<<(int)((slots + 255) / 256), 256, 0, s>>>(cb->perm, slots, inv);
    if (treal > 0)
      cb_real_slot_kernel<<<(int)((treal + 255) / 256), 256, 0, s>>>(cb->t->perm, treal,
                                                                     adj->perm, inv, tp->perm,
                                                                     d_bad);
    const int np = (int)pd.size();
    cb_pair_kernel<<<(np + 255) / 256, 256, 0, s>>>(d_pd, d_ps, np, tp->perm);
  }
  int bad = 0;
  if (st == MBX_OK &&
      (hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    st = MBX_EHIP;
  if (st == MBX_OK && bad) st = MBX_EINVAL;
  if (st == MBX_OK) st = mbx_detail::perm_schedule_build(tp, s);
  if (d_bad) (void)hipFree(d_bad);
  if (d_pd) (void)hipFree(d_pd);
  if (d_ps) (void)hipFree(d_ps);
  if (st == MBX_OK && hipGetLastError() != hipSuccess) st = MBX_EHIP;
  if (hipStreamSynchronize(s) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  if (inv) (void)hipFree(inv);
  if (adj) mbx_adjoint_destroy(adj);
  if (st != MBX_OK) {
    drop_adjoint(cb);
    return st;
  }
  cb->t->stream_mode = cb->stream_mode;
  if (cb->nb > 0 && cb->plan[0]) mbx_colblock_set_residency(cb->t, cb->plan[0]->wg_per_cu);
  if (nblocks_out) *nblocks_out = cb->t->nb;
  return MBX_OK;
}

int mbx_colblock_dx(const mbx_colblock* cb, int32_t value_kind, const void* d_val_blocked,
                    double alpha, const void* d_g, double beta, void* d_dx, void* stream) {
  if (!cb || !cb->t) return MBX_EINVAL;
  const int vb = value_bytes(value_kind);
  if (vb == 0) return MBX_EINVAL;
  if (!d_val_blocked || (cb->ncol > 0 && !d_dx) || (cb->nnz > 0 && !d_g)) return MBX_EINVAL;
  // one two-pass bucketed permutation into the A^T form's order
  mbx_detail::perm_apply(cb->tp, vb, d_val_blocked, (hipStream_t)stream);
  return mbx_colblock_spmv(cb->t, value_kind, cb->tp->vt, alpha, d_g, beta, d_dx, stream);
}

int mbx_colblock_values(const mbx_colblock* cb, int32_t value_kind, const void* d_val,
                        void* d_val_blocked, void* stream) {
  if (!cb) return MBX_EINVAL;
  const int vb = value_bytes(value_kind);
  if (vb == 0) return MBX_EINVAL;
  const long long slots = cb->slots;
  if (!d_val_blocked || (cb->nnz > 0 && !d_val)) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int grid = (int)((slots + 255) / 256);
  if (vb == 8)
    cb_values_kernel<<<grid, 256, 0, s>>>(cb->perm, slots, (const uint64_t*)d_val,
                                          (uint64_t*)d_val_blocked);
  else if (vb == 4)
    cb_values_kernel<<<grid, 256, 0, s>>>(cb->perm, slots, (const uint32_t*)d_val,
                                          (uint32_t*)d_val_blocked);
  else
    cb_values_kernel<<<grid, 256, 0, s>>>(cb->perm, slots, (const uint16_t*)d_val,
                                          (uint16_t*)d_val_blocked);
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

int mbx_colblock_values_to_csr(const mbx_colblock* cb, int32_t value_kind,
                               const void* d_val_blocked, void* d_val, void* stream) {
  if (!cb) return MBX_EINVAL;
  const int vb = value_bytes(value_kind);
  if (vb == 0 || !d_val_blocked || (cb->nnz > 0 && !d_val)) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const long long slots = cb->slots;
  const int grid = (int)((slots + 255) / 256);
  if (vb == 8)
    cb_to_csr_kernel<<<grid, 256, 0, s>>>(cb->perm, slots, (const uint64_t*)d_val_blocked,
                                          (uint64_t*)d_val);
  else if (vb == 4)
    cb_to_csr_kernel<<<grid, 256, 0, s>>>(cb->perm, slots, (const uint32_t*)d_val_blocked,
                                          (uint32_t*)d_val);
  else
    cb_to_csr_kernel<<<grid, 256, 0, s>>>(cb->perm, slots, (const uint16_t*)d_val_blocked,
                                          (uint16_t*)d_val);
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

int mbx_colblock_spmv(const mbx_colblock* cb, int32_t value_kind, const void* d_val_blocked,
                      double alpha, const void* d_x, double beta, void* d_y, void* stream) {
  if (!cb) return MBX_EINVAL;
  const int vb = value_bytes(value_kind);
  if (vb == 0) return MBX_EINVAL;
  if (cb->nrow > 0 && !d_y) return MBX_EINVAL;
  if (cb->nnz > 0 && (!d_val_blocked || !d_x)) return MBX_EINVAL;
  const int xb = value_kind == MBX_VALUE_F64 ? 8 : 4;
  // AUTO judges the whole chained product: its stream is read once per
  // product even when each block's part would stay in the MALL
  int mode = cb->stream_mode;
  if (mode == MBX_STREAM_AUTO) {
    const long long bytes = (long long)cb->nb * 4 * (cb->nrow + 1) + (4ll + vb) * cb->nnz;
    mode = mbx_detail::stream_nontemporal(MBX_STREAM_AUTO, bytes) ? MBX_STREAM_NONTEMPORAL
                                                                   : MBX_STREAM_CACHED;
  }
  for (int j = 0; j < cb->nb; ++j) {
    cb->plan[j]->stream_mode = mode;
    const int32_t* rp = cb->row_ptr + (long long)j * cb->stride;
    const int32_t* col = cb->col + cb->base[j];
    const char* val = static_cast<const char*>(d_val_blocked) + cb->base[j] * vb;
    const char* x = d_x ? static_cast<const char*>(d_x) + (long long)cb->lo[j] * xb : nullptr;
    const double bj = j == 0 ? beta : 1.0;
    int st;
    if (value_kind == MBX_VALUE_F64)
      st = mbx_spmv_f64(cb->plan[j], rp, col, (const double*)val, alpha, (const double*)x, bj,
                        (double*)d_y, stream);
    else if (value_kind == MBX_VALUE_F32)
      st = mbx_spmv_f32(cb->plan[j], rp, col, (const float*)val, (float)alpha,
                        (const float*)x, (float)bj, (float*)d_y, stream);
    else
      st = mbx_spmv_bf16(cb->plan[j], rp, col, (const uint16_t*)val, (float)alpha,
                         (const float*)x, (float)bj, (float*)d_y, stream);
    if (st != MBX_OK) return st;
  }
  return MBX_OK;
}

int mbx_colblock_sddmm(const mbx_colblock* cb, int32_t value_kind, const void* d_g,
                       const void* d_x, void* d_dval_blocked, void* stream) {
  if (!cb) return MBX_EINVAL;
  if (value_kind != MBX_VALUE_F64 && value_kind != MBX_VALUE_F32) return MBX_EINVAL;
  if (cb->nnz > 0 && (!d_g || !d_x || !d_dval_blocked)) return MBX_EINVAL;
  const int vb = value_kind == MBX_VALUE_F64 ? 8 : 4;
  for (int j = 0; j < cb->nb; ++j) {
    const int32_t* rp = cb->row_ptr + (long long)j * cb->stride;
    const char* x = d_x ? static_cast<const char*>(d_x) + (long long)cb->lo[j] * vb : nullptr;
    char* dval = d_dval_blocked ? static_cast<char*>(d_dval_blocked) + cb->base[j] * vb : nullptr;
    const int st = mbx_sddmm(cb->plan[j], value_kind, rp, cb->col + cb->base[j], d_g, x, dval,
                             stream);
    if (st != MBX_OK) return st;
  }
  return MBX_OK;
}

}  // extern "C"

// merbit_internal.hpp -- pieces shared by the HIP translation units of
// libmerbit_amd.so (not part of the C ABI).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <vector>

#include "merbit_amd.h"

// The merge-path plan: one (rows consumed, nnz consumed, row_ptr[rows])
// coordinate per tile boundary of wg*ipt merge items, plus FP64 carry slots
// (tail, head) per tile used by the SpMV carry fixup.  Clones share the
// coordinates (counted by coord_refs) and own their carry slots.
struct mbx_plan {
  int32_t nrow, ncol, nnz;
  int32_t wg, ipt, tile_items, ntiles;
  // merge items between tile starts: tile_items, or for row-aligned tiles
  // (mbx_plan_create_rows) tile_items minus the longest row
  int32_t tile_spacing;
  int32_t max_rows;    // most rows ending in one tile (pattern statistic)
  int32_t split_rows;  // tiles whose first row began earlier (fixup needed)
  int4* d_coord;  // per tile boundary: rows, nnz consumed, row_ptr[rows]
  double* d_carry;
  std::atomic<int>* coord_refs;  // plan objects sharing d_coord
  int32_t stream_mode;           // MBX_STREAM_*
  int32_t wg_per_cu;             // resident tiles per CU; 0 = as many as fit
  mutable hipEvent_t ev_begin, ev_end;
  // mbx_plan_compress_cols: column index minus its tile's base (d_coord[t].w)
  uint16_t* d_col16;
  int32_t accumulation;  // MBX_ACC_*: sums of the FP32 / bfloat16 value forms
  // mbx_plan_split_hot: the nhot most frequent columns (hot), the column
  // array with hot entries as -1 - slot (colh, padded to a multiple of 4) and
  // the per-product x values of the hot columns (xh, nhot * 8 bytes)
  int32_t* d_colh;
  int32_t* d_hot;
  void* d_xh;
  int32_t nhot;
  int32_t cus;  // compute units of the plan's device (persistent grids)
  // mbx_plan_sort_cols: every tile's columns sorted, in gather order, with
  // the element's tile position in the low bits (nnz padded to whole groups)
  int32_t* d_scol;
};

// Most hot columns a plan keeps (mbx_plan_split_hot): 128 KB of FP32 x.
constexpr int kMaxHotColumns = 32768;

namespace mbx_detail {
// Row-aligned tiles (mbx_plan_create_rows) need the longest row to be at
// most tile_items / kRowTileDiv, so tiles are at most 1/kRowTileDiv fewer
// items apart than plain ones.
constexpr int kRowTileDiv = 16;
}  // namespace mbx_detail

struct mbx_colblock;

// Adjoint object (mbx_adjoint_*): the CSC of A and the schedule that moves
// values into CSC order; also read by the column-blocked form's adjoint.
struct mbx_adjoint {
  int32_t nrow, ncol, nnz;
  int32_t* col_ptr;   // ncol + 1: row pointer of A^T
  int32_t* row_idx;   // nnz: column indices of A^T
  int32_t* perm;      // nnz: CSR position of every CSC entry
  void* vt;           // nnz values in CSC order (8 bytes each)
  mbx_plan* plan_t;   // merge-path plan of A^T
  // bucketed permutation schedule (built and described in merbit_adjoint.hip)
  uint16_t* lslot;    // nnz: LDS slot of CSR entry k within its chunk
  int32_t* spos;      // nnz: staging index of pass-1 slot i (chunk-major)
  int32_t* lidx;      // nnz: staging index of CSC position p
  void* stage;        // nnz values, bucket after bucket (8 bytes each)
  int32_t nsrc;       // sources held by the value array; those past it read 0
  // optional (mbx_adjoint_set_blocks): a column-blocked form of A^T and the
  // permutation from CSR order straight into its order
  mbx_colblock* tb;
  mbx_adjoint* tbp;
};

namespace mbx_detail {

// The adjoint's two-pass bucketed permutation on any bijection: for an
// object with nnz, perm (destination p reads source perm[p]) and the lslot,
// spos, lidx, stage and vt buffers allocated (nnz entries; 8 bytes for stage
// and vt), perm_schedule_build computes the schedule (synchronises s) and
// perm_apply moves values: vt[p] = val[perm[p]].
int perm_schedule_build(mbx_adjoint* a, hipStream_t s);
void perm_apply(const mbx_adjoint* a, int value_bytes, const void* val, hipStream_t s);
// Frees a permutation object's buffers (perm, lslot, spos, lidx, stage, vt)
// and the object.
void perm_free(mbx_adjoint* a);

// Column-blocked form internals read by the adjoint: the padding slots
// before the end of its last block, and its slot -> source position map
// (device array, -1 on padding).
std::vector<int32_t> colblock_padding(const mbx_colblock* cb);
const int32_t* colblock_slot_map(const mbx_colblock* cb);

// mbx_csr_validate's check (merbit_validate.hip), for entry points that must
// refuse a malformed pattern: MBX_EINVAL when row_ptr is not a valid
// non-decreasing pointer from 0 to nnz or a column lies outside [0, ncol).
// Allocates temporaries and synchronises s.
int csr_check(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow, int32_t ncol,
              int32_t nnz, hipStream_t s, mbx_csr_report* out);

// Frees a plan's hot-column split (merbit_hot.hip), waiting for the device.
void plan_drop_hot(mbx_plan* p);
// cnt[c] = entries of d_col equal to c, for c < ncol (entries outside
// [0, ncol) are not counted); cnt is cleared first.  Stream-ordered.
int count_columns(const int32_t* d_col, int nnz, int ncol, int32_t* cnt, hipStream_t s);
// mbx_plan_split_hot with the column counts of d_col given (cnt_given, as
// count_columns makes them) or, when it is null, made here.
int split_hot_counted(mbx_plan* p, const int32_t* d_col, int32_t max_hot,
                      const int32_t* cnt_given, hipStream_t s, double* covered);

constexpr int kXcds = 8;

// Workgroups are dispatched round-robin over the 8 XCDs; give XCD j the
// contiguous tile range starting at j*q+min(j,r) so neighbouring tiles (which
// share x lines and the cache lines at tile boundaries) meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n / kXcds, r = n % kXcds;
  const int j = b % kXcds, i = b / kXcds;
  return j * q + min(j, r) + i;
}

// The tiles of persistent workgroup b of nwg (a grid sized to stay resident):
// each XCD takes a contiguous range of the tiles, its workgroups every
// stride-th tile of it: first + w, first + w + stride, ... below first + len.
struct PersistentShare {
  int first, len, w, stride;
};
__device__ __forceinline__ PersistentShare persistent_share(int b, int nwg, int ntiles) {
  PersistentShare s{0, ntiles, b, nwg};
  if (nwg >= kXcds) {
    const int xcd = b % kXcds;
    s.w = b / kXcds;
    s.stride = (nwg - xcd + kXcds - 1) / kXcds;
    const int q = ntiles / kXcds, r = ntiles % kXcds;
    s.first = xcd * q + min(xcd, r);
    s.len = q + (xcd < r ? 1 : 0);
  }
  return s;
}

// x values of the hot columns for one product of a hot-column plan:
// xh[h] = x[hot[h]] (mbx_plan_split_hot).
template <class X>
__global__ __launch_bounds__(256) void hot_gather_kernel(const X* __restrict__ x,
                                                         const int32_t* __restrict__ hot,
                                                         int nhot, X* __restrict__ xh) {
  const int h = blockIdx.x * 256 + threadIdx.x;
  if (h < nhot) xh[h] = x[hot[h]];
}

// Merge-path search in LDS (the SpMM and SDDMM tiles): first row r in
// [max(0,d-nnz_t), min(d,rows_t)] whose local end s_re[r] exceeds d-r-1 (an
// nnz index is consumed before an equal row end), include/utils.cuh:5-29.
__device__ __forceinline__ int lds_merge_search(const int* s_re, int d, int nnz_t, int rows_t) {
  int lo = max(0, d - nnz_t), hi = min(d, rows_t);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s_re[mid] <= d - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The vector staging path reads row_ptr and col in 16-byte groups, FP64 and
// FP32 values in 16-byte and bfloat16 values in 8-byte groups, and gathers x
// by 32-bit byte offsets (x of at most 4 GiB: FP64 x is FP64, else FP32).
template <class V>
inline bool vec_path_ok(const int32_t* rp, const int32_t* col, const V* val, int32_t ncol) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(rp) | reinterpret_cast<uintptr_t>(col);
  const uintptr_t v = reinterpret_cast<uintptr_t>(val);
  const uint64_t x_bytes = (uint64_t)ncol * (sizeof(V) == 8 ? 8u : 4u);
  return (a & 15u) == 0 && (v & (sizeof(V) == 2 ? 7u : 15u)) == 0 && x_bytes <= 0xFFFFFFFFull;
}

// Whether a product reads its row_ptr/col/value stream non-temporally:
// asked for, or (MBX_STREAM_AUTO) when the product's bytes exceed what stays
// resident in the 256 MB MALL between products.
constexpr long long kResidentBytes = 240ll << 20;
inline bool stream_nontemporal(int32_t mode, long long footprint) {
  return mode == MBX_STREAM_NONTEMPORAL || (mode == MBX_STREAM_AUTO && footprint > kResidentBytes);
}

// LDS a tile launch reserves: what the tile needs, or more when the caller
// limits the tiles resident on a CU to wg_per_cu (the LDS then admits no
// more; fewer tiles per CU can mean fewer L1 misses for their gathers).
constexpr size_t kLdsPerCu = 160 * 1024;
inline size_t residency_lds(size_t need, int wg_per_cu) {
  if (wg_per_cu <= 0) return need;
  const size_t cap = kLdsPerCu / (size_t)wg_per_cu - 1024;
  return need > cap ? need : cap;
}

// Launch a tile kernel with `lds` bytes of dynamic LDS (the attribute allows
// more than 64 KB when asked).
template <class K>
inline void allow_lds(K* kernel, size_t lds) {
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// Launch-side validation shared by every entry point taking a plan.
inline bool plan_csr_args_ok(const mbx_plan* p, const void* rp, const void* col,
                             const void* val) {
  if (!p) return false;
  if (p->nrow > 0 && !rp) return false;
  if (p->nnz > 0 && (!col || !val)) return false;
  return true;
}

// Capture `repeats` calls of fn(stream) into an executable graph.
template <class F>
inline int capture_graph(int repeats, F fn, void** exec_out) {
  if (!exec_out || repeats < 1) return MBX_EINVAL;
  *exec_out = nullptr;
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return MBX_EHIP;
  int st = MBX_OK;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    st = MBX_EHIP;
  } else {
    for (int i = 0; i < repeats && st == MBX_OK; ++i) st = fn(cs);
    if (hipStreamEndCapture(cs, &g) != hipSuccess && st == MBX_OK) st = MBX_EHIP;
  }
  hipGraphExec_t exec = nullptr;
  if (st == MBX_OK && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess)
    st = MBX_EHIP;
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  if (st == MBX_OK) *exec_out = exec;
  (void)hipGetLastError();
  return st;
}

// kernel<<<grid, block, lds, s>>>(args...).  When the plan carries timing
// events (mbx_plan_set_events), they are attached to this dispatch instead
// (hipExtLaunchKernel) and then cleared: their interval is the kernel's own
// execution, the interval a kernel trace reports, without the marker packets
// and dispatch latency that event records around the launch would add.  A
// plan with only one of the two events set launches plainly; either way the
// plan's events are cleared.
template <class... KA, class... A>
inline void timed_launch(const mbx_plan* p, void (*kernel)(KA...), unsigned grid,
                         unsigned block, size_t lds, hipStream_t s, A... args) {
  if (p->ev_begin && p->ev_end)
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (uint32_t)lds, s, p->ev_begin,
                          p->ev_end, 0u, args...);
  else
    kernel<<<grid, block, lds, s>>>(args...);
  p->ev_begin = p->ev_end = nullptr;
}

}  // namespace mbx_detail

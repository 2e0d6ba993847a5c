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
  // Stream loads (row_ptr, col, values) are plain.  Streamed<F64> reads
  // row_ptr and col non-temporally (2 % faster on a cold R-MAT, 1 % on the
  // banded matrix) and keeps the values plain (non-temporal FP64 values
  // cost the banded product 9 % and the R-MAT hot-column product 2-5 %).
  static constexpr bool kIndexNT = false;
  static constexpr bool kValueNT = false;
  static constexpr bool kStreamedIndexNT = true;
  static constexpr bool kStreamedValueNT = false;
  static constexpr bool kCol16 = false;
  static constexpr bool kHotX = false;
  static constexpr bool kSorted = false;
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
  __device__ static acc_t shfl_xor(acc_t a, int d) { return __shfl_xor(a, d); }
  __device__ static lds_t to_lds(acc_t a) { return a; }
  __device__ static double to_f64(lds_t a) { return a; }
  // a row completed from an incoming carry c (earlier lanes) and this
  // lane's partial h
  __device__ static lds_t finish(double c, acc_t h) { return c + h; }
  __device__ static void store(y_t* y, scal_t alpha, lds_t v, scal_t beta, bool beta_zero) {
    *y = beta_zero ? alpha * v : alpha * v + beta * *y;
  }
  __device__ static void store_f64(y_t* y, scal_t alpha, double v, scal_t beta, bool beta_zero) {
    *y = beta_zero ? alpha * v : alpha * v + beta * *y;
  }
};

// Compensated FP32: products split by TwoProduct (p + e == v*x exactly),
// each thread's running sums kept as (s, c) with s updated by TwoSum and c
// collecting the rounding errors.  Partials that cross threads (the
// wavefront scan, the cross-wave carries and the tile carries) are FP64.
template <class V>
struct CompF32 {
  // Stream loads are plain; Streamed<CompF32> reads row_ptr, col and the
  // values non-temporally (4-8 % faster on the cold banded and R-MAT
  // matrices).
  static constexpr bool kIndexNT = false;
  static constexpr bool kValueNT = false;
  static constexpr bool kStreamedIndexNT = true;
  static constexpr bool kStreamedValueNT = true;
  static constexpr bool kCol16 = false;
  static constexpr bool kHotX = false;
  static constexpr bool kSorted = false;
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
  __device__ static acc_t shfl_xor(acc_t a, int d) {
    return make_float2(__shfl_xor(a.x, d), __shfl_xor(a.y, d));
  }
  __device__ static lds_t to_lds(acc_t a) { return a; }
  __device__ static double to_f64(lds_t a) { return (double)a.x + (double)a.y; }
  // the FP64 value split back into an unevaluated float pair
  __device__ static lds_t finish(double c, acc_t h) {
    const double d = c + ((double)h.x + (double)h.y);
    const float hi = (float)d;
    return make_float2(hi, (float)(d - (double)hi));
  }
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

// FP32 or bfloat16 values and FP32 x, y with FP64 products and sums: the
// product of two FP32 values (24-bit significands) is exact in FP64, so each
// nonzero costs one FP64 add in the walk instead of a TwoSum.  With packed
// FP32 off, an FP64 add issues at the FP32 rate, so this form needs far
// fewer VALU cycles than CompF32 for the same bytes; its rows carry FP64
// rounding until the final rounding of y to FP32.
template <class V>
struct Acc64 {
  static constexpr bool kIndexNT = false;
  static constexpr bool kValueNT = false;
  static constexpr bool kStreamedIndexNT = true;
  static constexpr bool kStreamedValueNT = true;
  static constexpr bool kCol16 = false;
  static constexpr bool kHotX = false;
  static constexpr bool kSorted = false;
  using val_t = V;
  using x_t = float;
  using y_t = float;
  using scal_t = float;
  using lds_t = double;
  using acc_t = double;
  __device__ static lds_t product(val_t v, x_t x) {
    return (double)CompF32<V>::widen(v) * (double)x;
  }
  __device__ static acc_t zero() { return 0.0; }
  __device__ static void add(acc_t& a, lds_t p) { a += p; }
  __device__ static acc_t combine(acc_t a, acc_t b) { return a + b; }
  __device__ static acc_t shfl_up(acc_t a, int d) { return __shfl_up(a, d); }
  __device__ static acc_t shfl_xor(acc_t a, int d) { return __shfl_xor(a, d); }
  __device__ static lds_t to_lds(acc_t a) { return a; }
  __device__ static double to_f64(lds_t a) { return a; }
  __device__ static lds_t finish(double c, acc_t h) { return c + h; }
  __device__ static void store(y_t* y, scal_t alpha, lds_t v, scal_t beta, bool beta_zero) {
    store_f64(y, alpha, v, beta, beta_zero);
  }
  __device__ static void store_f64(y_t* y, scal_t alpha, double v, scal_t beta, bool beta_zero) {
    *y = beta_zero ? (float)((double)alpha * v)
                   : (float)((double)alpha * v + (double)beta * (double)*y);
  }
};

using F32A = Acc64<float>;
using BF16A = Acc64<uint16_t>;

// FP32 or bfloat16 values and FP32 x, y with each product rounded once to
// FP32 and the rows summed in FP64: the products take 4 bytes of LDS
// instead of Acc64's 8, so a tile needs half the product area and more tiles
// (and their loads) fit a CU.  Each product's rounding adds at most 2^-24 of
// |val*x| to a row's error, and a row's partial is rounded to FP32 where it
// is kept in LDS and where a carry completes it: at most 3 * 2^-24 (about
// 1.8e-7) of sum |val*x| in all, the FP32 rounding of y included.
template <class V>
struct Acc64R : Acc64<V> {
  using typename Acc64<V>::val_t;
  using typename Acc64<V>::x_t;
  using typename Acc64<V>::y_t;
  using typename Acc64<V>::scal_t;
  using typename Acc64<V>::acc_t;
  using lds_t = float;
  __device__ static lds_t product(val_t v, x_t x) { return CompF32<V>::widen(v) * x; }
  __device__ static void add(acc_t& a, lds_t p) { a += (double)p; }
  __device__ static lds_t to_lds(acc_t a) { return (float)a; }
  __device__ static double to_f64(lds_t a) { return (double)a; }
  __device__ static lds_t finish(double c, acc_t h) { return (float)(c + h); }
  __device__ static void store(y_t* y, scal_t alpha, lds_t v, scal_t beta, bool beta_zero) {
    Acc64<V>::store_f64(y, alpha, (double)v, beta, beta_zero);
  }
};

using F32R = Acc64R<float>;
using BF16R = Acc64R<uint16_t>;

// P in its streamed mode: the stream loads P marks are non-temporal, so that
// data read once per product does not displace the x lines the gathers
// reuse.  For products whose data does not stay cache-resident from one
// product to the next; when it does, the plain loads of P keep it there
// (non-temporal loads cost a warm R-MAT product 6 %).
template <class P>
struct Streamed : P {
  static constexpr bool kIndexNT = P::kStreamedIndexNT;
  static constexpr bool kValueNT = P::kStreamedValueNT;
};

// P reading 16-bit column offsets from its tile's base column (the plan's
// compressed columns, mbx_plan_compress_cols) instead of 32-bit columns:
// 2 bytes less per nonzero in the stream.  Vector staging path only.
template <class P>
struct Col16 : P {
  static constexpr bool kCol16 = true;
};

// P reading x of the plan's hot columns from LDS (mbx_plan_split_hot): the
// column array holds -1 - slot for a hot column, whose x value sits in the
// workgroup's copy of the hot values (TileSmem::xh); the buffer gather of a
// negative index reads 0 without touching memory.  Vector staging path only.
template <class P>
struct HotX : P {
  static constexpr bool kHotX = true;
};

// P reading the plan's tile-sorted columns (mbx_plan_sort_cols) instead of
// the caller's: within each tile the nonzeros' columns are sorted and laid
// out so that each gather instruction's 64 lanes read 64 consecutive sorted
// columns, which share x lines and so need fewer vector-L1 requests.  Each
// 32-bit word holds the column in its high 32 - kSortSlotBits bits and the
// element's position in the tile (its product slot) in the low bits; the
// gathered x values are written to their slots in LDS and read back there by
// the threads that hold the matching values, so every product, and the
// result, is the same as the plain product's.  Vector staging path only.
template <class P>
struct SortX : P {
  static constexpr bool kSorted = true;
};

// Bits of the tile position in a tile-sorted column word: enough for the
// WG * IPT merge items of a tile.
template <int WG, int IPT>
constexpr int sort_slot_bits() {
  int b = 0;
  while ((1 << b) < WG * IPT) ++b;
  return b;
}

// out[0] = most rows ending in one tile, out[1] = tiles whose first row
// began in an earlier tile (they need the carry fixup kernel).
__device__ __forceinline__ void plan_stats_update(const int4* __restrict__ coord, int t,
                                                  int* __restrict__ out) {
  const int4 c0 = coord[t];
  const int4 c1 = coord[t + 1];
  atomicMax(out, c1.x - c0.x);
  if (t > 0 && c1.x > c0.x && c0.z < c0.y) atomicAdd(out + 1, 1);
}

// merge_search over the global row_ptr (include/utils.cuh:5-29): the
// (rows consumed, nnz consumed) coordinate of tile boundary t, and where the
// row open there starts (row_ptr[rows], so tiles need not load it).
__device__ __forceinline__ int4 tile_start_coord(const int32_t* __restrict__ row_ptr, int nrow,
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
  return make_int4((int)lo, (int)(diag - lo), nrow > 0 ? row_ptr[lo] : 0, 0);
}


// One aligned block of a staged stream.  NT: a non-temporal load (the
// block is read once; a non-temporal hint keeps it from displacing the x
// lines that the gathers reuse in the caches).
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));
template <bool NT, class E>
__device__ __forceinline__ E block_load(const E* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// 16-byte aligned groups of 4 consecutive elements.  A group is loaded only
// when it holds at least one element of the wanted range, so a load never
// leaves the 16-byte block of an in-range element (and never faults).
template <bool NT = false>
__device__ __forceinline__ void load_group(const int32_t* p, int g, int (&o)[4]) {
  const i32x4_t a = block_load<NT>(reinterpret_cast<const i32x4_t*>(p) + g);
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
template <bool NT = false>
__device__ __forceinline__ void load_group(const float* p, int g, float (&o)[4]) {
  const f32x4_t a = block_load<NT>(reinterpret_cast<const f32x4_t*>(p) + g);
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
// Four 16-bit column offsets (an 8-byte aligned group) plus the tile's base.
template <bool NT = false>
__device__ __forceinline__ void load_col16(const uint16_t* p, int g, int base, int (&o)[4]) {
  const u32x2_t a = block_load<NT>(reinterpret_cast<const u32x2_t*>(p) + g);
  o[0] = base + (int)(a.x & 0xFFFFu);
  o[1] = base + (int)(a.x >> 16);
  o[2] = base + (int)(a.y & 0xFFFFu);
  o[3] = base + (int)(a.y >> 16);
}
// bfloat16 groups stay packed (two values per 32-bit word) until their use.
struct Bf16x4 {
  uint32_t w[2];
  __device__ uint16_t operator[](int v) const {
    return (uint16_t)((v & 1) ? (w[v >> 1] >> 16) : (w[v >> 1] & 0xFFFFu));
  }
};
template <bool NT = false>
__device__ __forceinline__ void load_group(const uint16_t* p, int g, Bf16x4& o) {
  const u32x2_t a = block_load<NT>(reinterpret_cast<const u32x2_t*>(p) + g);
  o.w[0] = a.x;
  o.w[1] = a.y;
}
// FP64 groups are two 16-byte halves, each loaded only if it holds an element
// of [lo, hi).
template <bool NT = false>
__device__ __forceinline__ void load_group(const double* p, int g, double (&o)[4], int lo,
                                           int hi) {
  o[0] = o[1] = o[2] = o[3] = 0.0;
  const int e = 4 * g;
  if (e + 1 >= lo && e < hi) {
    const f64x2_t a = block_load<NT>(reinterpret_cast<const f64x2_t*>(p) + 2 * g);
    o[0] = a.x; o[1] = a.y;
  }
  if (e + 3 >= lo && e + 2 < hi) {
    const f64x2_t b = block_load<NT>(reinterpret_cast<const f64x2_t*>(p) + 2 * g + 1);
    o[2] = b.x; o[3] = b.y;
  }
}
template <bool NT = false, class T, class G>
__device__ __forceinline__ void load_group(const T* p, int g, G& o, int, int) {
  load_group<NT>(p, g, o);
}

// Value group g, every 16-byte block clamped to the blocks that hold elements
// of [lo, hi) (the caller guarantees lo < hi): always in bounds.
template <bool NT = false>
__device__ __forceinline__ void load_group_clamped(const double* p, int g, double (&o)[4], int lo,
                                                   int hi) {
  const int hb = lo >> 1, he = (hi - 1) >> 1;  // 16-byte blocks (2 doubles)
  const f64x2_t* q = reinterpret_cast<const f64x2_t*>(p);
  const f64x2_t a = block_load<NT>(q + min(max(2 * g, hb), he));
  const f64x2_t b = block_load<NT>(q + min(max(2 * g + 1, hb), he));
  o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}

// Register image of 4 consecutive values of type T.
template <class T>
struct ValGroup {
  T v[4];
  __device__ T operator[](int i) const { return v[i]; }
};
template <>
struct ValGroup<uint16_t> : Bf16x4 {};

template <bool NT = false>
__device__ __forceinline__ void load_group(const double* p, int g, ValGroup<double>& o, int lo,
                                           int hi) {
  load_group<NT>(p, g, o.v, lo, hi);
}
template <bool NT = false>
__device__ __forceinline__ void load_group(const float* p, int g, ValGroup<float>& o, int, int) {
  load_group<NT>(p, g, o.v);
}
template <bool NT = false>
__device__ __forceinline__ void load_group_clamped(const double* p, int g, ValGroup<double>& o,
                                                   int lo, int hi) {
  load_group_clamped<NT>(p, g, o.v, lo, hi);
}
template <bool NT = false, class T>
__device__ __forceinline__ void load_group_clamped(const T* p, int g, ValGroup<T>& o, int lo,
                                                   int hi) {
  load_group<NT>(p, g, o, lo, hi);
}

// Cross-lane moves through DPP (GFX9 controls): row_shr:n shifts within
// rows of 16 lanes, row_bcast:15 / row_bcast:31 pass lane 15 of a row to the
// next row / lane 31 to rows 2 and 3, wave_shr:1 shifts the whole wave by
// one.  Lanes without a source (or in rows left out by ROW_MASK) read 0.
constexpr int kDppRowShr = 0x110;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;
constexpr int kDppWaveShr1 = 0x138;
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, true);
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = dpp_move<CTRL, ROW_MASK>(__double2loint(v));
  const int hi = dpp_move<CTRL, ROW_MASK>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ int wave_prefix_sum(int v) {
  v += dpp_move<kDppRowShr | 1>(v);
  v += dpp_move<kDppRowShr | 2>(v);
  v += dpp_move<kDppRowShr | 4>(v);
  v += dpp_move<kDppRowShr | 8>(v);
  v += dpp_move<kDppRowBcast15, 0xa>(v);
  v += dpp_move<kDppRowBcast31, 0xc>(v);
  return v;
}

// Segmented inclusive sum over the 64 lanes of a wave: lane l adds the
// values of lanes seg..l-1, where seg (< l, or -1) is the last lane at or
// below l that starts a segment.  Earlier lanes' sums are added on the left.
__device__ __forceinline__ double wave_segmented_sum(double s, int lane, int seg) {
  const int reach = min(lane & 15, lane - seg);  // lanes back within row and segment
  double o = dpp_move<kDppRowShr | 1>(s);
  if (reach >= 1) s = o + s;
  o = dpp_move<kDppRowShr | 2>(s);
  if (reach >= 2) s = o + s;
  o = dpp_move<kDppRowShr | 4>(s);
  if (reach >= 4) s = o + s;
  o = dpp_move<kDppRowShr | 8>(s);
  if (reach >= 8) s = o + s;
  o = dpp_move<kDppRowBcast15, 0xa>(s);
  if ((lane & 16) && seg < (lane & ~15)) s = o + s;
  o = dpp_move<kDppRowBcast31, 0xc>(s);
  if (lane >= 32 && seg < 32) s = o + s;
  return s;
}

// Raw buffer resource over `bytes` bytes at p: loads through it take a
// 32-bit byte offset, and offsets past the end read 0 instead of faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           0x00020000 /* 32-bit data format */);
}
// The byte offset is formed in unsigned arithmetic (a column near 2^29 gives
// an offset past INT_MAX; the hardware reads it as unsigned).
__device__ __forceinline__ void buffer_gather(__amdgpu_buffer_rsrc_t r, int c, float& out) {
  out = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)((unsigned)c << 2), 0, 0));
}
__device__ __forceinline__ void buffer_gather(__amdgpu_buffer_rsrc_t r, int c, double& out) {
  typedef unsigned v2u __attribute__((ext_vector_type(2)));
  const v2u w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)((unsigned)c << 3), 0, 0);
  out = __hiloint2double((int)w.y, (int)w.x);
}

// Row-end marks.  The end item of local row i sits at merge position
// p = (row_ptr[r0+1+i] - k0) + i of the tile; bit p % IPT of s_mask[p / IPT]
// marks it for the thread owning that item and s_wcnt[wave of p] counts it.
template <int IPT>
__device__ __forceinline__ void mark_row_end(unsigned* s_mask, int* s_wcnt, int i, int re) {
  const int p = re + i;
  atomicOr(s_mask + p / IPT, 1u << (p % IPT));
  atomicAdd(s_wcnt + p / (64 * IPT), 1);
}

// Scalar staging: element i of the tile is handled by thread i % WG.  The
// marks must be zero on entry; the function contains a workgroup barrier
// between issuing the loads and setting the marks.  x is gathered through a
// buffer resource when it spans at most 4 GiB (a column index outside
// [0, ncol) then reads 0, as on the vector path), directly otherwise.
template <int WG, int IPT, class P>
__device__ __forceinline__ void stage_scalar(
    unsigned* s_mask, int* s_wcnt, typename P::lds_t* s_v, const int32_t* __restrict__ row_ptr,
    const int32_t* __restrict__ col, const typename P::val_t* __restrict__ val,
    const typename P::x_t* __restrict__ x, int ncol, int r0, int k0, int rows_t, int nnz_t) {
  const int tid = threadIdx.x;
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
  const uint64_t x_bytes = (uint64_t)ncol * sizeof(typename P::x_t);
  if (x_bytes <= 0xFFFFFFFFull) {  // uniform
    const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(x, (uint32_t)x_bytes);
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      xv[j] = typename P::x_t(0);
      if (i < nnz_t) buffer_gather(rx, ci[j], xv[j]);
    }
  } else {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      xv[j] = typename P::x_t(0);
      if (i < nnz_t) xv[j] = x[ci[j]];
    }
  }
  __syncthreads();  // marks zeroed by every thread
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = tid + j * WG;
    if (i < rows_t) mark_row_end<IPT>(s_mask, s_wcnt, i, re[j] - k0);
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = tid + j * WG;
    if (i < nnz_t) s_v[i] = P::product(vv[j], xv[j]);
  }
}

// Vector staging: 16-byte groups of 4 elements, group g of the tile by thread
// g % WG (row_ptr, col and FP32 values 16-byte aligned, FP64 values 16-byte,
// bfloat16 values 8-byte aligned; x of at most 4 GiB, gathered by 32-bit
// byte offsets).  Group indices are clamped into the tile's range, so every
// load is in bounds and issued unconditionally (no branch makes the compiler
// wait on a load before issuing the next); elements outside the tile are
// masked where they are used.
__device__ __forceinline__ void lds_store4(double* dst, const double (&v)[4]) {
  reinterpret_cast<double2*>(dst)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2*>(dst)[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void lds_store4(float* dst, const float (&v)[4]) {
  reinterpret_cast<float4*>(dst)[0] = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void lds_store4(float2* dst, const float2 (&v)[4]) {
  reinterpret_cast<float4*>(dst)[0] = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
  reinterpret_cast<float4*>(dst)[1] = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
}

// Registers of one tile on the vector staging path: its coordinates, the
// 16-byte row-end, column and value groups this thread loads, and the x
// values gathered for its column groups.  Filled in three steps so that a
// persistent workgroup can keep the next tile's loads in flight while it
// works on the current one: tile_coords, tile_issue_loads (needs the
// coordinates), tile_issue_gathers (needs the columns).
template <int WG, int IPT, class P>
struct TileRegs {
  static constexpr int NG = (WG * IPT / 4 + 2 + WG - 1) / WG;  // groups per thread
  int4 c0, c1;
  int rg[NG][4];
  int cg[NG][4];
  ValGroup<typename P::val_t> vg[NG];
  typename P::x_t xv[NG][4];
  __device__ int r0() const { return c0.x; }
  __device__ int k0() const { return c0.y; }
  __device__ int rows() const { return c1.x - c0.x; }
  __device__ int nnz() const { return c1.y - c0.y; }
};

template <int WG, int IPT, class P>
__device__ __forceinline__ void tile_coords(TileRegs<WG, IPT, P>& R,
                                            const int4* __restrict__ coord, int tile) {
  R.c0 = coord[tile];
  R.c1 = coord[tile + 1];
}

// Row groups first (clamped into valid entries of row_ptr), then the nonzero
// groups: no load waits for another.  Group rounds that no thread needs are
// skipped (uniform branches).
template <int WG, int IPT, class P>
__device__ __forceinline__ void tile_issue_loads(TileRegs<WG, IPT, P>& R,
                                                 const int32_t* __restrict__ row_ptr,
                                                 const int32_t* __restrict__ col,
                                                 const typename P::val_t* __restrict__ val) {
  constexpr int NG = TileRegs<WG, IPT, P>::NG;
  const int tid = threadIdx.x;
  const int e0 = R.r0() + 1, rows_t = R.rows(), k0 = R.k0(), nnz_t = R.nnz();
  const int rb = e0 >> 2, rlast = (e0 + rows_t - 1) >> 2;
  const int k1 = k0 + nnz_t;
  const int kb = k0 >> 2, klast = (k1 - 1) >> 2;
  const int nrg = rlast - rb + 1, nkg = klast - kb + 1;
  const int rl = max(rlast, rb - 1 < 0 ? 0 : rb - 1);
#pragma unroll
  for (int j = 0; j < NG; ++j)
    if (j == 0 || j * WG < nrg)
      load_group<P::kIndexNT>(row_ptr, min(rb + tid + j * WG, rl), R.rg[j]);
  if (nnz_t > 0) {  // uniform per tile
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j == 0 || j * WG < nkg) {
        const int g = min(kb + tid + j * WG, klast);
        if constexpr (P::kCol16)
          load_col16<P::kIndexNT>(reinterpret_cast<const uint16_t*>(col), g, R.c0.w, R.cg[j]);
        else
          load_group<P::kIndexNT>(col, g, R.cg[j]);
        load_group_clamped<P::kValueNT>(val, g, R.vg[j], k0, k1);
      }
    }
  }
}

// x gathers through a buffer resource over x: elements of a group outside
// the tile (nonzeros of the neighbouring tiles, or past the end of the column
// array) read some in-bounds value or 0, and their products are dropped by
// tile_write_stage; no load can fault.
template <int WG, int IPT, class P>
__device__ __forceinline__ void tile_issue_gathers(TileRegs<WG, IPT, P>& R,
                                                   const typename P::x_t* __restrict__ x,
                                                   int ncol) {
  constexpr int NG = TileRegs<WG, IPT, P>::NG;
  const int nnz_t = R.nnz();
  if (nnz_t > 0) {
    const int kb = R.k0() >> 2, nkg = ((R.k0() + nnz_t - 1) >> 2) - kb + 1;
    const __amdgpu_buffer_rsrc_t rx =
        buffer_rsrc(x, (uint32_t)ncol * (uint32_t)sizeof(typename P::x_t));
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j == 0 || j * WG < nkg) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if constexpr (P::kSorted)
            buffer_gather(rx, (int)((uint32_t)R.cg[j][v] >> sort_slot_bits<WG, IPT>()),
                          R.xv[j][v]);
          else
            buffer_gather(rx, R.cg[j][v], R.xv[j][v]);
        }
      }
    }
  }
}

// Tile-sorted columns (P::kSorted): each gathered x value goes to the LDS
// slot of its element (the low bits of its column word, offset as the
// products are), as an x_t in the first bytes of that product slot.  Only
// elements of the tile are written; the words of the clamped groups at the
// tile's ends belong to other tiles.  Needs a barrier before the slots are
// read (tile_write_stage).
template <int WG, int IPT, class P>
__device__ __forceinline__ void sorted_scatter_x(const TileRegs<WG, IPT, P>& R,
                                                 typename P::lds_t* s_v) {
  constexpr int NG = TileRegs<WG, IPT, P>::NG;
  constexpr uint32_t kMask = (1u << sort_slot_bits<WG, IPT>()) - 1u;
  using x_t = typename P::x_t;
  const int tid = threadIdx.x;
  const int k0 = R.k0(), nnz_t = R.nnz();
  if (nnz_t > 0) {
    const int kb = k0 >> 2, nkg = ((k0 + nnz_t - 1) >> 2) - kb + 1;
    const int off = k0 & 3;
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j == 0 || j * WG < nkg) {
        const int q = tid + j * WG;  // group kb + q
        const int lo = k0 - 4 * (kb + q), hi = lo + nnz_t;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (v >= lo && v < hi)
            *reinterpret_cast<x_t*>(s_v + ((uint32_t)R.cg[j][v] & kMask) + off) = R.xv[j][v];
      }
    }
  }
}

// Row-end marks (the marks must be zero and visible to the workgroup) and
// products to LDS: the product of tile element i goes to s_v[i + (k0 & 3)],
// so every group starts on a 4-slot boundary and groups wholly inside the
// tile are written with vector LDS stores.
template <int WG, int IPT, class P>
__device__ __forceinline__ void tile_write_stage(const TileRegs<WG, IPT, P>& R, unsigned* s_mask,
                                                 int* s_wcnt, typename P::lds_t* s_v) {
  constexpr int NG = TileRegs<WG, IPT, P>::NG;
  using lds_t = typename P::lds_t;
  const int tid = threadIdx.x;
  const int e0 = R.r0() + 1, rows_t = R.rows(), k0 = R.k0(), nnz_t = R.nnz();
  const int rb = e0 >> 2, nrg = ((e0 + rows_t - 1) >> 2) - rb + 1;
  const int kb = k0 >> 2, nkg = ((k0 + nnz_t - 1) >> 2) - kb + 1;
  // element v of row group rb + q is local row v - lo
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    if (j == 0 || j * WG < nrg) {
      const int lo = e0 - 4 * (rb + tid + j * WG), hi = lo + rows_t;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (v >= lo && v < hi) mark_row_end<IPT>(s_mask, s_wcnt, v - lo, R.rg[j][v] - k0);
    }
  }
  if (nnz_t > 0) {
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j == 0 || j * WG < nkg) {
        const int q = tid + j * WG;  // group kb + q
        const int lo = k0 - 4 * (kb + q), hi = lo + nnz_t;
        typename P::x_t xr[4];
        if constexpr (P::kSorted) {  // x of these elements, scattered to their slots
#pragma unroll
          for (int v = 0; v < 4; ++v)
            xr[v] = (v >= lo && v < hi)
                        ? *reinterpret_cast<const typename P::x_t*>(s_v + 4 * q + v)
                        : typename P::x_t(0);
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v) xr[v] = R.xv[j][v];
        }
        lds_t pr[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) pr[v] = P::product(R.vg[j][v], xr[v]);
        if (lo <= 0 && hi >= 4) {
          lds_store4(s_v + 4 * q, pr);
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (v >= lo && v < hi) s_v[4 * q + v] = pr[v];
        }
      }
    }
  }
}

// Hot columns (P::kHotX): x of every nonzero whose column index is -1 - h is
// the LDS value s_xh[h].  Run after the gathers' results are needed anyway
// (before the products): an LDS read per element, the gathered value kept
// for cold columns.
template <int WG, int IPT, class P>
__device__ __forceinline__ void hot_substitute(TileRegs<WG, IPT, P>& R,
                                               const typename P::x_t* s_xh) {
  constexpr int NG = TileRegs<WG, IPT, P>::NG;
  const int nnz_t = R.nnz();
  if (nnz_t > 0) {
    const int kb = R.k0() >> 2, nkg = ((R.k0() + nnz_t - 1) >> 2) - kb + 1;
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j == 0 || j * WG < nkg) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = R.cg[j][v];
          const typename P::x_t h = s_xh[c < 0 ? -1 - c : 0];
          if (c < 0) R.xv[j][v] = h;
        }
      }
    }
  }
}

// LDS of one tile, carved from dynamic shared memory: the product / row-sum
// area (products of the tile's nonzeros first, sums of the rows ending in the
// tile after them: together at most WG*IPT entries, the merge-path identity,
// plus up to 3 slots of vector-path offset), the cross-wave scan scratch and
// the row-end marks (one word per thread, one count per wave).
template <int WG, int IPT, class P>
struct TileSmem {
  static constexpr int NW = WG / 64;
  static constexpr size_t kV = sizeof(typename P::lds_t) * (WG * IPT + 4);
  static constexpr size_t kW = (sizeof(double) * NW + 15) / 16 * 16;
  static constexpr size_t kF = (sizeof(int) * NW + 15) / 16 * 16;  // one int per wave
  static constexpr size_t kM = sizeof(unsigned) * WG;
  typename P::lds_t* v;
  double* wtail;
  int* wflag;   // NW
  double* tail;
  unsigned* mask;
  int* wcnt;    // NW
  const typename P::x_t* xh = nullptr;  // hot x values (P::kHotX), set by the kernel
  __device__ TileSmem(char* base)
      : v(reinterpret_cast<typename P::lds_t*>(base)),
        wtail(reinterpret_cast<double*>(base + kV)),
        wflag(reinterpret_cast<int*>(base + kV + kW)),
        tail(reinterpret_cast<double*>(base + kV + kW + kF)),
        mask(reinterpret_cast<unsigned*>(base + kV + kW + kF + 16)),
        wcnt(reinterpret_cast<int*>(base + kV + kW + kF + 16 + kM)) {}
  static constexpr size_t bytes() { return kV + kW + kF + 16 + kM + kF; }
};

// The walk of one thread over `items` merge items from local coordinate
// (r, k); bit i of m marks item i as a row end.  A nonzero item adds its
// product to the open row; a row-end item writes the open row's sum to
// s_y[r] (for the thread's first row end this is only the part from this
// thread: the scan adds the earlier threads' part) and opens the next row.
// What remains in `sum` is the partial of the row open after the last item.
template <int IPT, bool FULL, class P>
__device__ __forceinline__ void walk_items(unsigned m, const typename P::lds_t* s_p,
                                           typename P::lds_t* s_y, int r, int k, int items,
                                           typename P::acc_t& sum) {
  using acc_t = typename P::acc_t;
  using lds_t = typename P::lds_t;
  // Item i is the (i - ends before it)-th nonzero of the thread: all product
  // reads are issued before the first use (a row-end item reads a slot it
  // does not use, at most the first row-sum slot).
  lds_t pv[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) pv[i] = s_p[k + i - __popc(m & ((1u << i) - 1u))];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const bool end = (m >> i) & 1u;
    if (end) s_y[r] = P::to_lds(sum);
    acc_t ns = sum;
    P::add(ns, pv[i]);
    const bool nz = !end && (FULL || i < items);
    sum = end ? P::zero() : (nz ? ns : sum);
    r += end ? 1 : 0;
  }
}

// The rest of one tile once its products and row-end marks are in LDS (and
// a barrier has made them visible): per-thread walk, wave and cross-wave
// carry scans, write-out.  Rows ending in the tile are written to y;
// carry[2*tile] receives the partial of the row open at the tile end and
// carry[2*tile+1] the partial of a first row that began in an earlier tile
// (completed by fixup_tile).  s_p: the tile's products.
template <int WG, int IPT, class P>
__device__ __forceinline__ void finish_tile(
    TileSmem<WG, IPT, P>& sm, typename P::lds_t* s_p, int tile, int4 c0, int4 c1,
    typename P::y_t* __restrict__ y, double* __restrict__ carry, typename P::scal_t alpha,
    typename P::scal_t beta, int beta_zero) {
  constexpr int NW = WG / 64;
  using lds_t = typename P::lds_t;
  using acc_t = typename P::acc_t;
  const int tid = threadIdx.x;
  const int r0 = c0.x, k0 = c0.y;
  const int rows_t = c1.x - r0;
  const int nnz_t = c1.y - k0;

  // Per-thread walk over merge items [tid*IPT, tid*IPT+IPT) of the tile.
  // Its start row is the number of row ends before its first item: the
  // counts of the earlier waves plus a wave scan of the marks.  Every tile
  // but a matrix's last has WG*IPT items, so the walk without item checks
  // is the common (wave-uniform) case.
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const unsigned m = sm.mask[tid];
  const int nm = __popc(m);
  int r = wave_prefix_sum(nm) - nm;
#pragma unroll
  for (int w = 0; w < NW - 1; ++w)
    if (w < wave) r += sm.wcnt[w];
  const int d0 = tid * IPT;
  const int items = min(max(nnz_t + rows_t - d0, 0), IPT);
  lds_t* s_y = s_p + nnz_t;
  acc_t sum = P::zero();
  if (items == IPT)
    walk_items<IPT, true, P>(m, s_p, s_y, r, d0 - r, IPT, sum);
  else if (items > 0)
    walk_items<IPT, false, P>(m, s_p, s_y, r, d0 - r, items, sum);
  const bool flagged = m != 0u;  // the thread closes row r (its head)

  // Segmented inclusive scan (FP64) of the threads' open-row partials.  A
  // thread with a row end starts a new segment; seg is the nearest such lane
  // at or below this one.
  const unsigned long long fmask = __ballot(flagged);
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const unsigned long long fb = fmask & upto;
  const int seg = fb ? 63 - __clzll(fb) : -1;
  const double s = wave_segmented_sum(P::to_f64(P::to_lds(sum)), lane, seg);
  const double cin = dpp_move<kDppWaveShr1>(s);
  if (lane == 63) {
    sm.wtail[wave] = s;
    sm.wflag[wave] = fmask != 0ull;
  }
  __syncthreads();
  double win = 0.0;
#pragma unroll
  for (int w = 0; w < NW - 1; ++w)
    if (w < wave) win = sm.wflag[w] ? sm.wtail[w] : win + sm.wtail[w];
  const bool ext = (fmask & ((1ull << lane) - 1ull)) == 0ull;
  if (flagged) {
    double c = (lane == 0) ? 0.0 : cin;
    if (ext) c = win + c;
    s_y[r] = P::finish(c, s_y[r]);
  }
  if (tid == WG - 1) *sm.tail = (fmask == 0ull) ? win + s : s;
  __syncthreads();

  // Coalesced write-out of the rows that end in this tile, and the carry
  // slots: tail = partial of the row open at the tile end, head = partial of
  // a row that began in an earlier tile (completed by the fixup).
  const bool head_multi = rows_t > 0 && c0.z < k0;  // row r0 began earlier
  for (int i = tid; i < rows_t; i += WG) {
    const lds_t v = s_y[i];
    if (i == 0 && head_multi)
      carry[2 * tile + 1] = P::to_f64(v);  // completed by the fixup kernel
    else
      P::store(y + r0 + i, alpha, v, beta, beta_zero != 0);
  }
  if (tid == 0) carry[2 * tile] = *sm.tail;
}

// Zero the row-end marks: after the previous tile's mark reads (which precede
// its carry-scan barrier) and before the barrier that precedes new marks.
template <int WG, int IPT, class P>
__device__ __forceinline__ void clear_marks(TileSmem<WG, IPT, P>& sm) {
  sm.mask[threadIdx.x] = 0u;
  if (threadIdx.x < WG / 64) sm.wcnt[threadIdx.x] = 0;
}

// One merge-path tile of y = alpha*A*x + beta*y (the work of one workgroup).
template <int WG, int IPT, bool VEC, class P>
__device__ __forceinline__ void spmv_tile(
    TileSmem<WG, IPT, P>& sm, int tile, int ncol, const int32_t* __restrict__ row_ptr,
    const int32_t* __restrict__ col, const typename P::val_t* __restrict__ val,
    const typename P::x_t* __restrict__ x, typename P::y_t* __restrict__ y,
    const int4* __restrict__ coord, double* __restrict__ carry, typename P::scal_t alpha,
    typename P::scal_t beta, int beta_zero) {
  static_assert(WG % 64 == 0 && WG / 64 <= 8, "tiles of up to 8 waves");
  static_assert(VEC || !P::kCol16, "16-bit columns are read by the vector staging only");
  static_assert(VEC || !P::kHotX, "hot columns are read by the vector staging only");
  static_assert(VEC || !P::kSorted, "sorted columns are read by the vector staging only");
  static_assert(!(P::kSorted && (P::kHotX || P::kCol16)),
                "sorted columns exclude the hot split and 16-bit columns");
  typename P::lds_t* s_p = sm.v;
  int4 c0, c1;
  // Stage: products to LDS and the row-end marks, every global load of the
  // tile issued before the first wait.
  if (VEC) {
    TileRegs<WG, IPT, P> R;
    tile_coords(R, coord, tile);
    tile_issue_loads(R, row_ptr, col, val);
    tile_issue_gathers(R, x, ncol);
    clear_marks(sm);
    // the previous tile of a persistent workgroup is done with the LDS (its
    // loop's barrier), so the x values can go to their slots before this
    // barrier, which then makes them visible
    if constexpr (P::kSorted) sorted_scatter_x(R, sm.v);
    __syncthreads();
    if constexpr (P::kHotX) hot_substitute(R, sm.xh);
    tile_write_stage(R, sm.mask, sm.wcnt, sm.v);
    s_p += R.k0() & 3;
    c0 = R.c0;
    c1 = R.c1;
  } else {
    c0 = coord[tile];
    c1 = coord[tile + 1];
    clear_marks(sm);
    stage_scalar<WG, IPT, P>(sm.mask, sm.wcnt, sm.v, row_ptr, col, val, x, ncol, c0.x, c0.y,
                             c1.x - c0.x, c1.y - c0.y);
  }
  __syncthreads();
  finish_tile(sm, s_p, tile, c0, c1, y, carry, alpha, beta, beta_zero);
}

// Plan-time build of one tile's sorted column words (mbx_plan_sort_cols,
// mbx_batch_sort_cols; read through SortX): the words (column << SB) | i of
// the tile's nonzeros i = 0..n-1 are sorted in LDS (bitonic; the slot bits
// make every key distinct), then written to the tile's positions in the
// order the vector staging gathers them -- round j, wave w, element v, lane
// l of the 16-byte groups -- so each gather instruction's 64 lanes take 64
// consecutive sorted columns.  *bad is set when a column lies outside
// [0, ncol).  One workgroup of WG threads per tile; s_key holds 2^SB words,
// s_wsum one int per wave.
template <int WG, int IPT>
__device__ __forceinline__ void sort_tile_cols(const int32_t* __restrict__ col, int ncol, int k0,
                                               int k1, int32_t* __restrict__ scol,
                                               int* __restrict__ bad, uint32_t* s_key,
                                               int* s_wsum) {
  constexpr int SB = sort_slot_bits<WG, IPT>();
  constexpr int N2 = 1 << SB;
  constexpr int NW = WG / 64;
  constexpr int NG = (WG * IPT / 4 + 2 + WG - 1) / WG;  // as TileRegs
  constexpr int C = 4 * NG;  // gather-order positions per thread
  const int tid = threadIdx.x;
  const int n = k1 - k0;
  int out_of_range = 0;
  for (int i = tid; i < N2; i += WG) {
    uint32_t key = 0xFFFFFFFFu;
    if (i < n) {
      const int c = col[k0 + i];
      out_of_range |= (c < 0 || c >= ncol) ? 1 : 0;
      key = ((uint32_t)c << SB) | (uint32_t)i;
    }
    s_key[i] = key;
  }
  if (out_of_range) atomicOr(bad, 1);
  __syncthreads();
  for (int k = 2; k <= N2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < N2; i += WG) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint32_t a = s_key[i], b = s_key[ixj];
          if ((a > b) == ((i & k) == 0)) {
            s_key[i] = b;
            s_key[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // gather-order position kk = ((j * NW + w) * 4 + v) * 64 + l holds element
  // v of group kb + j * WG + w * 64 + l; thread tid ranks kk in [tid*C, tid*C+C)
  const int kb = k0 >> 2;
  int cnt = 0;
#pragma unroll
  for (int u = 0; u < C; ++u) {
    const int kk = tid * C + u;
    const int q = ((kk >> 8) / NW) * WG + ((kk >> 8) % NW) * 64 + (kk & 63);
    const int p = 4 * (kb + q) + ((kk >> 6) & 3);
    cnt += (p >= k0 && p < k1) ? 1 : 0;
  }
  const int incl = wave_prefix_sum(cnt);
  if ((tid & 63) == 63) s_wsum[tid >> 6] = incl;
  __syncthreads();
  int rank = incl - cnt;
  for (int w = 0; w < NW; ++w)
    if (w < (tid >> 6)) rank += s_wsum[w];
#pragma unroll
  for (int u = 0; u < C; ++u) {
    const int kk = tid * C + u;
    const int q = ((kk >> 8) / NW) * WG + ((kk >> 8) % NW) * 64 + (kk & 63);
    const int p = 4 * (kb + q) + ((kk >> 6) & 3);
    if (p >= k0 && p < k1) scol[p] = (int32_t)s_key[rank++];
  }
}

// Carry fixup for tile t >= 1 whose first row began in an earlier tile ts and
// ends in t: y[row] = alpha*(tail[ts] + ... + tail[t-1] + head[t]) +
// beta*y[row].  One thread per tile: runs of at most 64 tails are summed by
// their own thread in order; longer runs are summed by the whole wave (lane l
// takes tails ts+l, ts+l+64, ... in order, then a fixed xor butterfly).  The
// order depends only on the run length, so results are reproducible.  Every
// lane of the wave must call this (inactive lanes pass valid = false).
template <class P>
__device__ __forceinline__ void fixup_tile(bool valid, int t, const int4* __restrict__ coord,
                                           const double* __restrict__ carry,
                                           typename P::y_t* __restrict__ y, int tile_items,
                                           typename P::scal_t alpha, typename P::scal_t beta,
                                           int beta_zero) {
  constexpr int kShort = 64;
  const int lane = threadIdx.x & 63;
  int r0 = 0, ts = 0;
  if (valid) {
    const int4 c0 = coord[t];
    const int4 c1 = coord[t + 1];
    r0 = c0.x;
    valid = c1.x > r0 && c0.z < c0.y;  // the row ends here and began earlier
    ts = (int)(((long long)r0 + c0.z) / tile_items);
  }
  if (valid && t - ts <= kShort) {
    // in order; loads issued 8 at a time so a run is not one round trip per tile
    double acc = 0.0;
    int s = ts;
    for (; s + 8 <= t; s += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = carry[2 * (s + u)];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; s < t; ++s) acc += carry[2 * s];
    P::store_f64(y + r0, alpha, acc + carry[2 * t + 1], beta, beta_zero != 0);
  }
  unsigned long long longs = __ballot(valid && t - ts > kShort);
  while (longs) {
    const int l = __ffsll(longs) - 1;
    longs &= longs - 1;
    const int lt = __shfl(t, l), lts = __shfl(ts, l), lr = __shfl(r0, l);
    // lanes of a batch wave may belong to different matrices: take lane l's
    const double* lc = reinterpret_cast<const double*>(__shfl((long long)carry, l));
    typename P::y_t* ly = reinterpret_cast<typename P::y_t*>(__shfl((long long)y, l));
    double acc = 0.0;
    for (int s = lts + lane; s < lt; s += 64) acc += lc[2 * s];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d);
    if (lane == 0)
      P::store_f64(ly + lr, alpha, acc + lc[2 * lt + 1], beta, beta_zero != 0);
  }
}

}  // namespace mbx_detail

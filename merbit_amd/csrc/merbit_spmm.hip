// merbit_spmm.hip -- merge-path tile SpMM Y = alpha*A*X + beta*Y (FP32).
//
// The SpMV tile kernel generalised to KC-wide vectors: a tile's col/val stay
// in registers, and for each chunk of KC right-hand sides the products
// val*X[col, chunk] are staged in LDS as KC-vectors, walked per thread along
// the merge path, combined by a per-component segmented scan, and rows ending
// in the tile are written with coalesced stores.  X is read as rows (the KC
// values of one column index are contiguous), so each nonzero gathers one
// KC*4-byte segment instead of KC scattered words; a column-major X is
// repacked once per call into the caller's workspace.  Carries of rows that
// cross tiles (FP64, per right-hand side) are completed by a fixup kernel in a
// fixed order, as for SpMV.
//
// Two tile kernels: for k <= 4 one lane per path thread (tile_spmm_kernel,
// KC = 4); for wider X, G lanes per path thread (tile_spmm_rows_kernel), so
// that G lanes read one nonzero's X segment together and each lane carries a
// single float4 through the walk and the scan.

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "merbit_amd.h"
#include "merbit_internal.hpp"

using mbx_detail::lds_merge_search;
using mbx_detail::xcd_tile;

namespace {

constexpr int kTransRows = 128;

// Column-major X (ncol x k, leading dimension ldx) -> row-major Xr (ld ldr):
// coalesced reads down each column, coalesced writes along each row.
__global__ __launch_bounds__(256) void repack_rows_kernel(const float* __restrict__ X,
                                                          long long ldx, int x_row_major,
                                                          int ncol, int k,
                                                          float* __restrict__ Xr, int ldr) {
  __shared__ float t[64][kTransRows + 1];
  const int c0 = blockIdx.x * kTransRows;
  const int j0 = blockIdx.y * 64;
  const int kk = min(64, k - j0);
  for (int e = threadIdx.x; e < kk * kTransRows; e += 256) {
    int j, cc;
    if (x_row_major) {
      cc = e / kk;
      j = e % kk;
    } else {
      j = e / kTransRows;
      cc = e % kTransRows;
    }
    const int c = c0 + cc;
    if (c < ncol)
      t[j][cc] = x_row_major ? X[(long long)c * ldx + j0 + j] : X[c + (long long)(j0 + j) * ldx];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kk * kTransRows; e += 256) {
    const int cc = e / kk, j = e % kk;
    const int c = c0 + cc;
    if (c < ncol) Xr[(long long)c * ldr + j0 + j] = t[j][cc];
  }
}

// Column-major X with 16-byte aligned columns (ldx % 4 == 0) -> row-major Xr:
// 16-byte loads down each column, 16-byte stores along each row, transposed
// through LDS (128 rows x 64 right-hand sides per workgroup).
__global__ __launch_bounds__(256) void repack_cols_vec_kernel(const float* __restrict__ X,
                                                              long long ldx, int ncol, int k,
                                                              float* __restrict__ Xr,
                                                              int ldr) {
  __shared__ float t[64][kTransRows + 1];
  const int c0 = blockIdx.x * kTransRows;
  const int j0 = blockIdx.y * 64;
  const int kk = min(64, k - j0);
  constexpr int C4 = kTransRows / 4;  // float4 per column segment
  for (int e = threadIdx.x; e < kk * C4; e += 256) {
    const int j = e / C4, cc = 4 * (e - j * C4);
    const int c = c0 + cc;
    const float* src = X + (long long)(j0 + j) * ldx + c;
    float4 v;
    if (c + 3 < ncol) {
      v = *reinterpret_cast<const float4*>(src);
    } else {
      v.x = c < ncol ? src[0] : 0.f;
      v.y = c + 1 < ncol ? src[1] : 0.f;
      v.z = c + 2 < ncol ? src[2] : 0.f;
      v.w = 0.f;
    }
    t[j][cc] = v.x;
    t[j][cc + 1] = v.y;
    t[j][cc + 2] = v.z;
    t[j][cc + 3] = v.w;
  }
  __syncthreads();
  const int k4 = (kk + 3) / 4;  // float4 groups per row segment (the last one padded with 0)
  for (int e = threadIdx.x; e < kTransRows * k4; e += 256) {
    const int cc = e / k4, g = e - cc * k4;
    const int c = c0 + cc;
    if (c >= ncol) continue;
    const int j = 4 * g;
    const float4 v = make_float4(t[j][cc], j + 1 < kk ? t[j + 1][cc] : 0.f,
                                 j + 2 < kk ? t[j + 2][cc] : 0.f, j + 3 < kk ? t[j + 3][cc] : 0.f);
    *reinterpret_cast<float4*>(Xr + (long long)c * ldr + j0 + j) = v;
  }
}

template <int KC>
struct Vec {
  float v[KC];
};

template <int KC>
__device__ __forceinline__ void vzero(Vec<KC>& a) {
#pragma unroll
  for (int q = 0; q < KC; ++q) a.v[q] = 0.f;
}

template <int WG, int IPT, int KC>
__global__ __launch_bounds__(WG) void tile_spmm_kernel(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ Xr, int ldr, int k,
    float* __restrict__ Y, long long ldy, int y_row_major, const int4* __restrict__ coord,
    double* __restrict__ carry, int ntiles, float alpha, float beta, int beta_zero) {
  constexpr int T = WG * IPT;
  constexpr int NW = WG / 64;
  constexpr int KG = KC / 4;  // float4 groups per chunk
  __shared__ int s_re[T + 1];
  __shared__ float4 s_v[T * KG];  // product vectors [0,nnz_t), row sums after
  __shared__ float s_wtail[NW][KC];
  __shared__ int s_wflag[NW];
  __shared__ float s_tail[KC];

  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int tid = threadIdx.x;
  const int4 c0 = coord[tile];
  const int4 c1 = coord[tile + 1];
  const int r0 = c0.x, k0 = c0.y;
  const int rows_t = c1.x - r0;
  const int nnz_t = c1.y - k0;

  int ci[IPT];
  float vv[IPT];
  {
    int re[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = tid + j * WG;
      re[j] = INT_MAX;
      ci[j] = 0;
      vv[j] = 0.f;
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
  }
  __syncthreads();

  const int d0 = tid * IPT;
  const int items = min(max(nnz_t + rows_t - d0, 0), IPT);
  const int r_start = items > 0 ? lds_merge_search(s_re, d0, nnz_t, rows_t) : 0;
  const int k_start = d0 - r_start;
  const bool head_multi = rows_t > 0 && c0.z < k0;
  const int lane = tid & 63, wave = tid >> 6;
  float4* s_y = s_v + nnz_t * KG;

  for (int cb = 0; cb < k; cb += KC) {
    const int kc = min(KC, k - cb);
    // 1. products for this chunk of right-hand sides
    {
      float4 xv[IPT][KG];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const int i = tid + j * WG;
#pragma unroll
        for (int g = 0; g < KG; ++g) {
          xv[j][g] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (i < nnz_t && 4 * g < kc)
            xv[j][g] = *reinterpret_cast<const float4*>(Xr + (long long)ci[j] * ldr + cb + 4 * g);
        }
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const int i = tid + j * WG;
        if (i < nnz_t) {
#pragma unroll
          for (int g = 0; g < KG; ++g) {
            const float4 a = xv[j][g];
            s_v[i * KG + g] = make_float4(vv[j] * a.x, vv[j] * a.y, vv[j] * a.z, vv[j] * a.w);
          }
        }
      }
    }
    __syncthreads();

    // 2. walk
    Vec<KC> sum, head;
    vzero(sum);
    vzero(head);
    int head_row = 0;
    bool flagged = false;
    if (items > 0) {
      int r = r_start, kk = k_start;
      int re = s_re[r];
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        if (i < items) {
          if (kk < re) {
#pragma unroll
            for (int g = 0; g < KG; ++g) {
              const float4 p = s_v[kk * KG + g];
              sum.v[4 * g] += p.x;
              sum.v[4 * g + 1] += p.y;
              sum.v[4 * g + 2] += p.z;
              sum.v[4 * g + 3] += p.w;
            }
            ++kk;
          } else {
            if (flagged) {
#pragma unroll
              for (int g = 0; g < KG; ++g)
                s_y[r * KG + g] = make_float4(sum.v[4 * g], sum.v[4 * g + 1],
                                              sum.v[4 * g + 2], sum.v[4 * g + 3]);
            } else {
              head = sum;
              head_row = r;
              flagged = true;
            }
            vzero(sum);
            ++r;
            re = s_re[r];
          }
        }
      }
    }

    // 3. per-component segmented scan across the tile
    const unsigned long long fmask = __ballot(flagged);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long fb = fmask & upto;
    const int seg = fb ? 63 - __clzll(fb) : -1;
    Vec<KC> s = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const bool take = lane >= d && lane - d >= seg;
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        const float o = __shfl_up(s.v[q], d);
        if (take) s.v[q] = o + s.v[q];
      }
    }
    Vec<KC> cin;
#pragma unroll
    for (int q = 0; q < KC; ++q) cin.v[q] = __shfl_up(s.v[q], 1);
    if (lane == 63) {
#pragma unroll
      for (int q = 0; q < KC; ++q) s_wtail[wave][q] = s.v[q];
      s_wflag[wave] = fmask != 0ull;
    }
    __syncthreads();
    Vec<KC> win;
    vzero(win);
#pragma unroll
    for (int w = 0; w < NW - 1; ++w)
      if (w < wave) {
        const bool f = s_wflag[w];
#pragma unroll
        for (int q = 0; q < KC; ++q) win.v[q] = f ? s_wtail[w][q] : win.v[q] + s_wtail[w][q];
      }
    const bool ext = (fmask & ((1ull << lane) - 1ull)) == 0ull;
    if (flagged) {
      Vec<KC> c;
#pragma unroll
      for (int q = 0; q < KC; ++q) {
        c.v[q] = (lane == 0) ? 0.f : cin.v[q];
        if (ext) c.v[q] = win.v[q] + c.v[q];
      }
#pragma unroll
      for (int g = 0; g < KG; ++g)
        s_y[head_row * KG + g] =
            make_float4(c.v[4 * g] + head.v[4 * g], c.v[4 * g + 1] + head.v[4 * g + 1],
                        c.v[4 * g + 2] + head.v[4 * g + 2], c.v[4 * g + 3] + head.v[4 * g + 3]);
    }
    if (tid == WG - 1) {
#pragma unroll
      for (int q = 0; q < KC; ++q) s_tail[q] = (fmask == 0ull) ? win.v[q] + s.v[q] : s.v[q];
    }
    __syncthreads();

    // 4. write-out: coalesced along rows (column-major Y) or along columns
    const float* sy = reinterpret_cast<const float*>(s_y);
    const int n_out = rows_t * kc;
    for (int e = tid; e < n_out; e += WG) {
      int i, q;
      if (y_row_major) {
        i = e / kc;
        q = e - i * kc;
      } else {
        q = e / rows_t;
        i = e - q * rows_t;
      }
      const float v = sy[i * KC + q];
      const int jcol = cb + q;
      if (i == 0 && head_multi) {
        carry[(2ll * tile + 1) * k + jcol] = (double)v;
      } else {
        float* yp = y_row_major ? Y + (long long)(r0 + i) * ldy + jcol
                                : Y + (r0 + i) + (long long)jcol * ldy;
        *yp = beta_zero ? alpha * v : alpha * v + beta * *yp;
      }
    }
    if (tid < kc) carry[(2ll * tile) * k + cb + tid] = (double)s_tail[tid];
    __syncthreads();
  }
}

// The same tile SpMM with G lanes per merge-path thread: lane g of a group
// owns right-hand sides [4g, 4g+4) of the chunk (KC = 4G), so a chunk covers
// a whole X row segment of 16G bytes.  The product stage maps G consecutive
// lanes onto one nonzero's segment: one wave instruction reads 64/G whole
// segments (one 128-byte line each at G = 8) instead of 64 scattered 16-byte
// pieces, and each nonzero's row is fetched once per chunk of 4G columns.
// The walk and the segmented scan then move one float4 per lane.  The tile is
// the plan's (TI = 256 * items per thread path steps); each of the WG / G
// path threads walks TI * G / WG of them.  Products live in dynamic LDS
// (TI * G float4).
template <int WG, int TI, int G>
__global__ __launch_bounds__(WG) void tile_spmm_rows_kernel(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ Xr, int ldr, int k,
    float* __restrict__ Y, long long ldy, int y_row_major, const int4* __restrict__ coord,
    double* __restrict__ carry, int ntiles, float alpha, float beta, int beta_zero) {
  constexpr int KC = 4 * G;
  constexpr int NW = WG / 64;
  constexpr int PW = 64 / G;         // path threads per wave
  constexpr int IPP = TI * G / WG;   // path steps per path thread
  // product-stage loads in flight per lane: 2 at G = 2 (k <= 8, fewer
  // registers, twice the resident workgroups: k = 8 162 against 177 us on
  // the R-MAT), 8 at G = 4 (k = 32 586 against 597 at 4, 646 at 2)
  constexpr int U = G <= 2 ? 2 : 8;
  static_assert((TI * G) % WG == 0 && 64 % G == 0, "tile shape");
  __shared__ int s_re[TI + 1];
  __shared__ float s_wtail[NW][KC];
  __shared__ int s_wflag[NW];
  __shared__ float s_tail[KC];
  extern __shared__ float4 s_v[];  // [TI * G]: products [0, nnz_t), row sums after

  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int tid = threadIdx.x;
  const int4 c0 = coord[tile];
  const int4 c1 = coord[tile + 1];
  const int r0 = c0.x, k0 = c0.y;
  const int rows_t = c1.x - r0;
  const int nnz_t = c1.y - k0;

  for (int i = tid; i < rows_t; i += WG) s_re[i] = row_ptr[r0 + 1 + i] - k0;
  if (tid == 0) s_re[rows_t] = INT_MAX;
  __syncthreads();

  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane % G, pl = lane / G;
  const int d0 = (tid / G) * IPP;
  const int items = min(max(nnz_t + rows_t - d0, 0), IPP);
  const int r_start = items > 0 ? lds_merge_search(s_re, d0, nnz_t, rows_t) : 0;
  const int k_start = d0 - r_start;
  const bool head_multi = rows_t > 0 && c0.z < k0;
  float4* s_y = s_v + nnz_t * G;
  const int n_prod = nnz_t * G;

  for (int cb = 0; cb < k; cb += KC) {
    const int kc = min(KC, k - cb);
    // 1. products: G consecutive lanes per nonzero, one float4 each
    for (int base = 0; base < n_prod; base += U * WG) {
      int ci[U];
      float vv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = base + u * WG + tid;
        ci[u] = 0;
        vv[u] = 0.f;
        if (e < n_prod) {
          ci[u] = col[k0 + e / G];
          vv[u] = val[k0 + e / G];
        }
      }
      float4 xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = base + u * WG + tid;
        xv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (e < n_prod && 4 * g < kc)
          xv[u] = *reinterpret_cast<const float4*>(Xr + (long long)ci[u] * ldr + cb + 4 * g);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = base + u * WG + tid;
        if (e < n_prod)
          s_v[e] = make_float4(vv[u] * xv[u].x, vv[u] * xv[u].y, vv[u] * xv[u].z,
                               vv[u] * xv[u].w);
      }
    }
    __syncthreads();

    // 2. walk: every lane of a path thread follows the same path, on its float4
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f), head = sum;
    int head_row = 0;
    bool flagged = false;
    if (items > 0) {
      int r = r_start, kk = k_start;
      int re = s_re[r];
#pragma unroll
      for (int i = 0; i < IPP; ++i) {
        if (i < items) {
          if (kk < re) {
            const float4 p = s_v[kk * G + g];
            sum.x += p.x;
            sum.y += p.y;
            sum.z += p.z;
            sum.w += p.w;
            ++kk;
          } else {
            if (flagged) {
              s_y[r * G + g] = sum;
            } else {
              head = sum;
              head_row = r;
              flagged = true;
            }
            sum = make_float4(0.f, 0.f, 0.f, 0.f);
            ++r;
            re = s_re[r];
          }
        }
      }
    }

    // 3. segmented scan over the path threads, per component
    const unsigned long long fmask = __ballot(flagged);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long fb = fmask & upto;
    const int seg = fb ? (63 - __clzll(fb)) / G : -1;
    float4 s = sum;
#pragma unroll
    for (int d = 1; d < PW; d <<= 1) {
      const bool take = pl >= d && pl - d >= seg;
      const float ox = __shfl_up(s.x, d * G), oy = __shfl_up(s.y, d * G);
      const float oz = __shfl_up(s.z, d * G), ow = __shfl_up(s.w, d * G);
      if (take) s = make_float4(ox + s.x, oy + s.y, oz + s.z, ow + s.w);
    }
    const float4 cin = make_float4(__shfl_up(s.x, G), __shfl_up(s.y, G), __shfl_up(s.z, G),
                                   __shfl_up(s.w, G));
    if (pl == PW - 1) {
      s_wtail[wave][4 * g] = s.x;
      s_wtail[wave][4 * g + 1] = s.y;
      s_wtail[wave][4 * g + 2] = s.z;
      s_wtail[wave][4 * g + 3] = s.w;
      if (g == 0) s_wflag[wave] = fmask != 0ull;
    }
    __syncthreads();
    float4 win = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int w = 0; w < NW - 1; ++w)
      if (w < wave) {
        const bool f = s_wflag[w];
        const float* t = s_wtail[w] + 4 * g;
        win = f ? make_float4(t[0], t[1], t[2], t[3])
                : make_float4(win.x + t[0], win.y + t[1], win.z + t[2], win.w + t[3]);
      }
    const bool ext = (fmask & ((1ull << (pl * G)) - 1ull)) == 0ull;
    if (flagged) {
      float4 c = (pl == 0) ? make_float4(0.f, 0.f, 0.f, 0.f) : cin;
      if (ext) c = make_float4(win.x + c.x, win.y + c.y, win.z + c.z, win.w + c.w);
      s_y[head_row * G + g] = make_float4(c.x + head.x, c.y + head.y, c.z + head.z, c.w + head.w);
    }
    if (tid >= WG - G) {
      const float4 t = (fmask == 0ull)
                           ? make_float4(win.x + s.x, win.y + s.y, win.z + s.z, win.w + s.w)
                           : s;
      s_tail[4 * g] = t.x;
      s_tail[4 * g + 1] = t.y;
      s_tail[4 * g + 2] = t.z;
      s_tail[4 * g + 3] = t.w;
    }
    __syncthreads();

    // 4. write-out, as the per-lane kernel
    const float* sy = reinterpret_cast<const float*>(s_y);
    const int n_out = rows_t * kc;
    for (int e = tid; e < n_out; e += WG) {
      int i, q;
      if (y_row_major) {
        i = e / kc;
        q = e - i * kc;
      } else {
        q = e / rows_t;
        i = e - q * rows_t;
      }
      const float v = sy[i * KC + q];
      const int jcol = cb + q;
      if (i == 0 && head_multi) {
        carry[(2ll * tile + 1) * k + jcol] = (double)v;
      } else {
        float* yp = y_row_major ? Y + (long long)(r0 + i) * ldy + jcol
                                : Y + (r0 + i) + (long long)jcol * ldy;
        *yp = beta_zero ? alpha * v : alpha * v + beta * *yp;
      }
    }
    if (tid < kc) carry[(2ll * tile) * k + cb + tid] = (double)s_tail[tid];
    __syncthreads();
  }
}

// One thread per (tile t >= 1, right-hand side q): a row that began in an
// earlier tile ts and ends in t is completed by summing tail[ts..t-1] then
// head[t] in order (FP64).
__global__ __launch_bounds__(256) void spmm_fixup_kernel(
    const int4* __restrict__ coord, const double* __restrict__ carry, int k,
    float* __restrict__ Y, long long ldy, int y_row_major, int ntiles, int tile_items,
    float alpha, float beta, int beta_zero) {
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  if (g >= (long long)ntiles * k) return;
  const int t = (int)(g / k), q = (int)(g - (long long)t * k);
  if (t < 1) return;
  const int4 c0 = coord[t];
  const int4 c1 = coord[t + 1];
  const int r0 = c0.x, k0 = c0.y;
  if (c1.x <= r0) return;
  const int rs = c0.z;
  if (rs >= k0) return;
  const int ts = (int)(((long long)r0 + rs) / tile_items);
  // in order; loads issued 8 at a time so a long run is not one round trip
  // per tile
  double acc = 0.0;
  int s = ts;
  for (; s + 8 <= t; s += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = carry[(2ll * (s + u)) * k + q];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; s < t; ++s) acc += carry[(2ll * s) * k + q];
  acc += carry[(2ll * t + 1) * k + q];
  float* yp = y_row_major ? Y + (long long)r0 * ldy + q : Y + r0 + (long long)q * ldy;
  *yp = beta_zero ? (float)((double)alpha * acc)
                  : (float)((double)alpha * acc + (double)beta * (double)*yp);
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Row-major X usable in place: 16-byte aligned rows holding whole float4
// groups (the kernel reads KC-wide chunks as float4).
bool x_direct(const float* X, long long ldx, int x_layout, int k) {
  return x_layout == MBX_ROW_MAJOR && k % 4 == 0 && ldx % 4 == 0 &&
         (reinterpret_cast<uintptr_t>(X) & 15u) == 0;
}

size_t workspace_bytes(const mbx_plan* p, int k, bool repack) {
  const int ldr = (k + 3) / 4 * 4;
  size_t b = align_up(sizeof(double) * 2 * ((size_t)p->ntiles + 1) * (size_t)k, 256);
  if (repack) b += align_up(sizeof(float) * (size_t)p->ncol * (size_t)ldr, 256);
  return b;
}

template <int IPT, int KC>
void launch_spmm(const mbx_plan* p, hipStream_t s, const int32_t* rp, const int32_t* col,
                 const float* val, const float* Xr, int ldr, int k, float* Y, long long ldy,
                 int yrm, double* carry, float alpha, float beta) {
  const int bz = beta == 0.f;
  mbx_detail::timed_launch(p, tile_spmm_kernel<256, IPT, KC>, p->ntiles, 256, 0, s, rp, col, val,
                           Xr, ldr, k, Y, ldy, yrm, p->d_coord, carry, p->ntiles, alpha, beta,
                           bz);
  if (p->ntiles > 1)
    spmm_fixup_kernel<<<(int)(((long long)p->ntiles * k + 255) / 256), 256, 0, s>>>(
        p->d_coord, carry, k, Y, ldy, yrm, p->ntiles, p->tile_items, alpha, beta, bz);
}

template <int WG, int TI, int G>
void launch_spmm_rows(const mbx_plan* p, hipStream_t s, const int32_t* rp, const int32_t* col,
                      const float* val, const float* Xr, int ldr, int k, float* Y,
                      long long ldy, int yrm, double* carry, float alpha, float beta) {
  const int bz = beta == 0.f;
  const size_t lds = sizeof(float4) * (size_t)TI * G;
  auto kern = tile_spmm_rows_kernel<WG, TI, G>;
  mbx_detail::allow_lds(kern, lds + sizeof(int) * (TI + 1) + 4096);
  mbx_detail::timed_launch(p, kern, p->ntiles, WG, lds, s, rp, col, val, Xr, ldr, k, Y, ldy, yrm,
                           p->d_coord, carry, p->ntiles, alpha, beta, bz);
  if (p->ntiles > 1)
    spmm_fixup_kernel<<<(int)(((long long)p->ntiles * k + 255) / 256), 256, 0, s>>>(
        p->d_coord, carry, k, Y, ldy, yrm, p->ntiles, p->tile_items, alpha, beta, bz);
}

}  // namespace

extern "C" {

size_t mbx_spmm_workspace_bytes(const mbx_plan* p, int32_t k, const float* d_x, int64_t ldx,
                                int32_t x_layout) {
  if (!p || k <= 0) return 0;
  return workspace_bytes(p, k, !x_direct(d_x, ldx, x_layout, k));
}

int mbx_spmm_f32(const mbx_plan* p, const int32_t* rp, const int32_t* col, const float* val,
                 int32_t k, float alpha, const float* X, int64_t ldx, int32_t x_layout,
                 float beta, float* Y, int64_t ldy, int32_t y_layout, void* workspace,
                 size_t workspace_size, void* stream) {
  if (!mbx_detail::plan_csr_args_ok(p, rp, col, val) || k <= 0) return MBX_EINVAL;
  if (x_layout != MBX_COL_MAJOR && x_layout != MBX_ROW_MAJOR) return MBX_EINVAL;
  if (y_layout != MBX_COL_MAJOR && y_layout != MBX_ROW_MAJOR) return MBX_EINVAL;
  if ((p->ncol > 0 && !X) || (p->nrow > 0 && !Y)) return MBX_EINVAL;
  if (x_layout == MBX_COL_MAJOR ? ldx < p->ncol : ldx < k) return MBX_EINVAL;
  if (y_layout == MBX_COL_MAJOR ? ldy < p->nrow : ldy < k) return MBX_EINVAL;
  if (p->wg != 256 || (p->ipt != 2 && p->ipt != 4)) return MBX_EUNSUPPORTED;
  if (p->ntiles == 0) return MBX_OK;
  const bool repack = !x_direct(X, ldx, x_layout, k);
  const size_t need = workspace_bytes(p, k, repack);
  if (!workspace || workspace_size < need) return MBX_EINVAL;
  if ((reinterpret_cast<uintptr_t>(workspace) & 255u) != 0) return MBX_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  char* ws = static_cast<char*>(workspace);
  double* carry = reinterpret_cast<double*>(ws);
  const float* Xr = X;
  int ldr = (int)ldx;
  if (repack) {
    ldr = (k + 3) / 4 * 4;
    float* xr = reinterpret_cast<float*>(
        ws + align_up(sizeof(double) * 2 * ((size_t)p->ntiles + 1) * (size_t)k, 256));
    if (p->ncol > 0) {
      dim3 grid((p->ncol + kTransRows - 1) / kTransRows, (k + 63) / 64);
      if (x_layout == MBX_COL_MAJOR && ldx % 4 == 0 &&
          (reinterpret_cast<uintptr_t>(X) & 15u) == 0)
        repack_cols_vec_kernel<<<grid, 256, 0, s>>>(X, ldx, p->ncol, k, xr, ldr);
      else
        repack_rows_kernel<<<grid, 256, 0, s>>>(X, ldx, x_layout == MBX_ROW_MAJOR, p->ncol, k,
                                                xr, ldr);
    }
    Xr = xr;
  }
  const int yrm = y_layout == MBX_ROW_MAJOR;
  // k <= 4: one lane per path thread with KC = 4.  Wider: G lanes per path
  // thread, G chosen so each path thread walks 8 steps (G * items per thread)
  // and no group of a chunk is idle for k = 8.
  if (k <= 4) {
    if (p->ipt == 2)
      launch_spmm<2, 4>(p, s, rp, col, val, Xr, ldr, k, Y, ldy, yrm, carry, alpha, beta);
    else
      launch_spmm<4, 4>(p, s, rp, col, val, Xr, ldr, k, Y, ldy, yrm, carry, alpha, beta);
  } else if (p->ipt == 2) {
    if (k <= 8)
      launch_spmm_rows<256, 512, 2>(p, s, rp, col, val, Xr, ldr, k, Y, ldy, yrm, carry, alpha,
                                    beta);
    else
      launch_spmm_rows<256, 512, 4>(p, s, rp, col, val, Xr, ldr, k, Y, ldy, yrm, carry, alpha,
                                    beta);
  } else {
    launch_spmm_rows<256, 1024, 2>(p, s, rp, col, val, Xr, ldr, k, Y, ldy, yrm, carry, alpha,
                                   beta);
  }
  return hipGetLastError() == hipSuccess ? MBX_OK : MBX_EHIP;
}

}  // extern "C"

/*
 * merbit_oracle.c -- CPU restatement of MERBIT's merge-path tile SpMV.
 *
 * TEST INFRASTRUCTURE ONLY (see merbit_oracle.h).  The product never links
 * or calls this file.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_merge_search   brute-force linear merge on small inputs, the
 *                      coordinate invariants (x+y = diag, row_ptr[y] <= x <=
 *                      row_ptr[y+1], monotone in diag), end points.
 *   orc_ref_plan       brute-force path enumeration; sum of popcounts of the
 *                      flag fields = nrow; hand-written bit images of a 4x4
 *                      matrix (tests/golden/ref_plan_*.json).
 *   orc_spmv_csr_f64   scipy.sparse CSR matvec and numpy dense matvec; the
 *                      5-point Laplacian closed form with x = 1.
 *   orc_spmv_accurate  exact rational sums (Python fractions) on random and
 *                      cancellation-heavy rows: result within 1 ulp.
 *   orc_merbit_emulate hand-derived results on tiny matrices whose values
 *                      make the summation grouping visible; agreement with
 *                      orc_spmv_accurate within 1e-12 condition-scaled.
 *   orc_ref_ingest     hand-written COO -> CSR(A^T) example and scipy
 *                      transpose; column sums of the result with x = 1.
 *   orc_csr_transpose  scipy.sparse transpose.
 */
#include "merbit_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int64_t i64min(int64_t a, int64_t b) { return a < b ? a : b; }
static inline int64_t i64max(int64_t a, int64_t b) { return a > b ? a : b; }

/* include/utils.cuh:5-29.  Smallest y in [max(diag-nnz,0), min(diag,nrow)]
 * with row_ptr[y+1] > diag-y-1; x = diag - y.  Tie rule: an nnz index is
 * consumed before the row end that equals it. */
void orc_merge_search(const int32_t* row_ptr, int32_t nrow, int32_t nnz,
                      int64_t diag, int32_t* coord_x, int32_t* coord_y) {
  int64_t lo = i64max(diag - nnz, 0);
  int64_t hi = i64min(diag, nrow);
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)row_ptr[mid + 1] <= diag - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  *coord_x = (int32_t)(diag - lo);
  *coord_y = (int32_t)i64min(lo, nrow);
}

int32_t orc_ref_lane_count(int32_t nrow, int32_t nnz, int32_t sigma) {
  if (sigma < 1 || sigma > 42) return -1;
  return (int32_t)(((int64_t)nrow + nnz + sigma - 1) / sigma);
}

/* include/tile.cuh:28-104 (Tile_Kernel) and 106-122 (TILE ctor). */
int32_t orc_ref_plan(const int32_t* row_ptr, int32_t nrow, int32_t nnz,
                     int32_t sigma, uint32_t* tile_x, uint32_t* tile_y,
                     uint64_t* lane_desc) {
  const int32_t lane_num = orc_ref_lane_count(nrow, nnz, sigma);
  if (lane_num < 0) return -1;
  const int32_t tile_num = (lane_num + 31) / 32;
  const int wide = sigma >= 15; /* DescTypeT: u32 for SIGMA < 15 (tile.cuh:4) */
  const int d1 = wide ? 22 : 18, d2 = wide ? 11 : 9; /* tile.cuh:51-52 */
  const int64_t path = (int64_t)nrow + nnz;
  for (int32_t t = 0; t < tile_num; ++t) {
    int32_t wx, wy;
    orc_merge_search(row_ptr, nrow, nnz, (int64_t)t * 32 * sigma, &wx, &wy);
    const int32_t lanes = ((t + 1) * 32 < lane_num) ? 32 : lane_num - t * 32;
    int all_same = 1;
    for (int32_t l = 0; l < lanes; ++l) {
      const int64_t gid = (int64_t)t * 32 + l;
      const int64_t diag = gid * sigma;
      const int32_t lane_nnz = ((gid + 1) * sigma < path) ? sigma : (int32_t)(path - diag);
      int32_t cx, cy;
      orc_merge_search(row_ptr, nrow, nnz, diag, &cx, &cy);
      const uint64_t dx = (uint64_t)(cx - wx), dy = (uint64_t)(cy - wy);
      uint64_t flags = 0;
      for (int32_t i = 0; i < lane_nnz; ++i) { /* tile.cuh:71-82 */
        if (cx < row_ptr[cy + 1]) {
          ++cx;
        } else {
          flags |= (uint64_t)1 << i;
          ++cy;
        }
      }
      uint64_t d = (flags << d1) | (dy << d2) | dx;
      if (!wide) d &= 0xFFFFFFFFull;
      lane_desc[gid] = d;
      if (cy != wy) all_same = 0; /* tile.cuh:87-90 */
    }
    tile_x[t] = (uint32_t)wx;
    tile_y[t] = (uint32_t)wy | (all_same ? 0x80000000u : 0u);
  }
  tile_x[tile_num] = (uint32_t)nnz; /* tile.cuh:99-103 */
  tile_y[tile_num] = (uint32_t)nrow;
  return lane_num;
}

void orc_spmv_csr_f64(int32_t nrow, const int32_t* row_ptr, const int32_t* col,
                      const double* val, double alpha, const double* x,
                      double beta, double* y) {
  for (int32_t r = 0; r < nrow; ++r) {
    double s = 0.0;
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += val[k] * x[col[k]];
    y[r] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[r];
  }
}

/* bf16 bit pattern -> float (exact: bf16 is the top half of an IEEE float). */
static inline double bf16_to_double(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, sizeof f);
  return (double)f;
}

static inline double load_val(const void* p, int32_t ty, int64_t k) {
  switch (ty) {
    case 0: return ((const double*)p)[k];
    case 1: return (double)((const float*)p)[k];
    default: return bf16_to_double(((const uint16_t*)p)[k]);
  }
}

/* Knuth TwoSum and fma-based TwoProduct (error-free transformations). */
static inline void two_sum(double a, double b, double* s, double* e) {
  double t = a + b;
  double bb = t - a;
  *e = (a - (t - bb)) + (b - bb);
  *s = t;
}

void orc_spmv_accurate(int32_t nrow, const int32_t* row_ptr, const int32_t* col,
                       const void* val, int32_t val_type, const void* x,
                       int32_t x_type, double* y, double* absr) {
  for (int32_t r = 0; r < nrow; ++r) {
    double s = 0.0, c = 0.0, a = 0.0;
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
      const double v = load_val(val, val_type, k);
      const double xv = load_val(x, x_type, col[k]);
      const double p = v * xv;
      const double pe = fma(v, xv, -p);
      double t, e;
      two_sum(s, p, &t, &e);
      s = t;
      c += e + pe;
      a += fabs(p);
    }
    y[r] = s + c;
    if (absr) absr[r] = a;
  }
}

/* ---- reference-order emulator ------------------------------------------ */

/* tile.cuh:308-352 normal_track + utils.cuh:58-83 warp_segmented_sum for one
 * 32-lane tile whose products are already in s_data (block-local base). */
static void emu_normal_tile(int32_t t, int32_t nrow, int32_t nnz, int32_t sigma,
                            int32_t lane_num, const uint64_t* lane_desc,
                            const double* s_data, double* s_y) {
  const int wide = sigma >= 15;
  const int dshift = wide ? 11 : 9, fshift = wide ? 22 : 18;
  const uint64_t dmask = ((uint64_t)1 << dshift) - 1;
  const int64_t path = (int64_t)nrow + nnz;
  const int32_t len = ((t + 1) * 32 < lane_num) ? 32 : lane_num - t * 32;
  double lsum[32];
  int lflag[32], liy[32];
  for (int32_t l = 0; l < len; ++l) {
    const int64_t gid = (int64_t)t * 32 + l;
    const int32_t lane_nnz = ((gid + 1) * sigma < path) ? sigma : (int32_t)(path - gid * sigma);
    const uint64_t d = lane_desc[gid];
    int32_t ix = (int32_t)(d & dmask);
    int32_t iy = (int32_t)((d >> dshift) & dmask);
    const uint64_t bits = d >> fshift;
    int first = 1, lane_flag = (l == 0);
    double sum = 0.0;
    for (int32_t i = 0; i < lane_nnz; ++i) {
      const int flag = (int)((bits >> i) & 1u);
      /* tile.cuh:330: sum += (1-flag)*s_data[ind_x] */
      sum += flag ? 0.0 : s_data[ix];
      if (flag) {
        if (first) {
          s_y[iy] += sum; /* LDS atomicAdd, tile.cuh:336 */
          first = 0;
        } else {
          s_y[iy] = sum; /* tile.cuh:341 */
        }
        sum = 0.0;
      }
      ix += 1 - flag;
      iy += flag;
      lane_flag |= flag;
    }
    lsum[l] = sum;
    lflag[l] = lane_flag;
    liy[l] = iy;
  }
  /* utils.cuh:61-74: Hillis-Steele segmented scan with shfl_up */
  double ls[32];
  int lf[32];
  for (int32_t l = 0; l < len; ++l) {
    ls[l] = lsum[l];
    lf[l] = lflag[l];
  }
  for (int32_t i = 1;; i <<= 1) {
    int any = 0;
    for (int32_t l = 0; l < len; ++l) any |= !lf[l];
    if (!any) break;
    double ts[32];
    int tf[32];
    for (int32_t l = 0; l < len; ++l) {
      ts[l] = (l >= i) ? ls[l - i] : ls[l];
      tf[l] = (l >= i) ? lf[l - i] : lf[l];
    }
    for (int32_t l = 0; l < len; ++l) {
      ls[l] += (!lf[l] && l >= i) ? ts[l] : 0.0;
      lf[l] = lf[l] ? lf[l] : tf[l];
    }
  }
  /* utils.cuh:76-82: flagged lane L adds lane L-1's carry (lane 0 the last
   * lane's) into that lane's open row. */
  for (int32_t l = 0; l < len; ++l) {
    const int32_t loc = (l == 0) ? len - 1 : l - 1;
    if (lflag[l]) s_y[liy[loc]] += ls[loc];
  }
}

int32_t orc_merbit_emulate(int32_t nrow, int32_t nnz, const int32_t* row_ptr,
                           const int32_t* col, const double* val, double alpha,
                           const double* x, int32_t sigma, double* y) {
  const int32_t lane_num = orc_ref_lane_count(nrow, nnz, sigma);
  if (lane_num < 0) return -1;
  const int32_t tile_num = (lane_num + 31) / 32;
  uint32_t* tile_x = (uint32_t*)malloc(sizeof(uint32_t) * (tile_num + 1));
  uint32_t* tile_y = (uint32_t*)malloc(sizeof(uint32_t) * (tile_num + 1));
  uint64_t* desc = (uint64_t*)malloc(sizeof(uint64_t) * (lane_num > 0 ? lane_num : 1));
  const int32_t BLOCK = 256;
  double* shared = (double*)malloc(sizeof(double) * (BLOCK * sigma + 1));
  orc_ref_plan(row_ptr, nrow, nnz, sigma, tile_x, tile_y, desc);
  memset(y, 0, sizeof(double) * nrow); /* tile.cuh:597 */
  const int32_t grid = (lane_num + BLOCK - 1) / BLOCK; /* tile.cuh:596 */
  for (int32_t b = 0; b < grid; ++b) {
    /* tile.cuh:500-507 */
    const int32_t bts = b * BLOCK / 32;
    const int32_t bte = ((b + 1) * BLOCK / 32 < tile_num + 1) ? (b + 1) * BLOCK / 32 : tile_num;
    const int32_t bxs = (int32_t)tile_x[bts], bxe = (int32_t)tile_x[bte];
    const int32_t bxn = bxe - bxs;
    const int32_t bys = (int32_t)(tile_y[bts] & 0x7FFFFFFFu);
    const int32_t bye = (int32_t)(tile_y[bte] & 0x7FFFFFFFu);
    const int32_t byn = (bye == nrow) ? bye - bys : bye - bys + 1;
    memset(shared, 0, sizeof(double) * (BLOCK * sigma + 1)); /* 511-516 */
    double* s_data = shared;
    double* s_y = shared + bxn;
    for (int32_t w = 0; w < BLOCK / 32; ++w) { /* 520-556 */
      const int32_t t = bts + w;
      if (t >= tile_num) continue;
      const int32_t wxs = (int32_t)tile_x[t], wxe = (int32_t)tile_x[t + 1];
      if (wxs >= wxe) continue;
      const uint32_t ty = tile_y[t];
      const int fast = (int)(ty >> 31);
      const int32_t wys = (int32_t)(ty & 0x7FFFFFFFu);
      const int32_t wbys = wys - bys, wxn = wxe - wxs;
      if (fast) { /* fast_track, tile.cuh:267-290 */
        double v[32];
        for (int l = 0; l < 32; ++l) {
          v[l] = 0.0;
          for (int32_t idx = l; idx < wxn; idx += 32)
            v[l] = fma(val[wxs + idx], x[col[wxs + idx]], v[l]); /* contracted */
        }
        for (int dl = 16; dl > 0; dl >>= 1) { /* warp_reduce, utils.cuh:32-42 */
          double nv[32];
          for (int l = 0; l < 32; ++l) nv[l] = v[l] + v[l ^ dl];
          memcpy(v, nv, sizeof v);
        }
        s_y[wbys] += v[0];
      } else { /* load_smem then normal_track, tile.cuh:293-305, 308-352 */
        const int32_t wbxs = wxs - bxs;
        for (int32_t idx = 0; idx < wxn; ++idx)
          s_data[wbxs + idx] = val[wxs + idx] * x[col[wxs + idx]];
        emu_normal_tile(t, nrow, nnz, sigma, lane_num, desc, s_data + wbxs, s_y + wbys);
      }
    }
    /* load_mem, tile.cuh:354-375 (first and last row via global atomics) */
    for (int32_t idx = 0; idx < byn; ++idx) {
      const int32_t r = bys + idx;
      if (idx == 0 || idx == byn - 1)
        y[r] += alpha * s_y[idx];
      else
        y[r] = alpha * s_y[idx];
    }
  }
  free(tile_x);
  free(tile_y);
  free(desc);
  free(shared);
  return 0;
}

/* ---- ingest and transpose ---------------------------------------------- */

void orc_ref_ingest(int32_t nrow, int32_t nnz, const int32_t* row_ind,
                    const int32_t* col_ind, int32_t* row_ptr_t, int32_t* col_t,
                    double* val_t) {
  int32_t* cnt = (int32_t*)calloc((size_t)nrow, sizeof(int32_t));
  for (int32_t k = 0; k < nnz; ++k) cnt[row_ind[k]]++; /* class.cu:2-9 */
  /* Sort by (col,row) (class.cu:163-175), then row pointer from the sorted
   * col_ind (tile.cuh:574): a stable counting sort by row inside a counting
   * sort by column gives the same order. */
  int32_t* byrow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  int32_t* rstart = (int32_t*)calloc((size_t)nrow + 1, sizeof(int32_t));
  for (int32_t k = 0; k < nnz; ++k) rstart[row_ind[k] + 1]++;
  for (int32_t r = 0; r < nrow; ++r) rstart[r + 1] += rstart[r];
  for (int32_t k = 0; k < nnz; ++k) byrow[rstart[row_ind[k]]++] = k;
  memset(row_ptr_t, 0, sizeof(int32_t) * ((size_t)nrow + 1));
  for (int32_t k = 0; k < nnz; ++k) row_ptr_t[col_ind[k] + 1]++;
  for (int32_t c = 0; c < nrow; ++c) row_ptr_t[c + 1] += row_ptr_t[c];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nrow + 1));
  memcpy(fill, row_ptr_t, sizeof(int32_t) * ((size_t)nrow + 1));
  for (int32_t i = 0; i < nnz; ++i) {
    const int32_t k = byrow[i];
    const int32_t p = fill[col_ind[k]]++;
    col_t[p] = row_ind[k];
    val_t[p] = 1.0 / (double)cnt[row_ind[k]]; /* class.cu:11-19 */
  }
  free(cnt);
  free(byrow);
  free(rstart);
  free(fill);
}

void orc_csr_transpose(int32_t nrow, int32_t ncol, int32_t nnz,
                       const int32_t* row_ptr, const int32_t* col,
                       int32_t* col_ptr, int32_t* row_idx, int32_t* perm) {
  memset(col_ptr, 0, sizeof(int32_t) * ((size_t)ncol + 1));
  for (int32_t k = 0; k < nnz; ++k) col_ptr[col[k] + 1]++;
  for (int32_t c = 0; c < ncol; ++c) col_ptr[c + 1] += col_ptr[c];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ncol + 1));
  memcpy(fill, col_ptr, sizeof(int32_t) * ((size_t)ncol + 1));
  for (int32_t r = 0; r < nrow; ++r)
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
      const int32_t p = fill[col[k]]++;
      row_idx[p] = r;
      perm[p] = k;
    }
  free(fill);
}

/* ---- CPU baseline -------------------------------------------------------- */

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int cmp_double(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

double orc_spmv_csr_omp_time(int32_t nrow, const int32_t* row_ptr,
                             const int32_t* col, const double* val,
                             const double* x, double* y, int32_t nthreads,
                             int32_t warmup, int32_t reps,
                             int32_t* threads_used) {
  int nt = nthreads > 0 ? nthreads : 1;
#ifndef _OPENMP
  nt = 1;
#endif
  /* nnz-balanced static chunks: thread t owns rows [b[t], b[t+1]) */
  int32_t* bnd = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nt + 1));
  const int64_t nnz = row_ptr[nrow];
  for (int t = 0; t <= nt; ++t) {
    const int64_t target = nnz * t / nt;
    int32_t lo = 0, hi = nrow;
    while (lo < hi) {
      int32_t mid = (lo + hi) / 2;
      if (row_ptr[mid] < target) lo = mid + 1; else hi = mid;
    }
    bnd[t] = (t == nt) ? nrow : lo;
  }
  double* times = (double*)malloc(sizeof(double) * (size_t)(reps > 0 ? reps : 1));
  for (int it = 0; it < warmup + reps; ++it) {
    const double t0 = now_s();
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      for (int32_t r = bnd[tid]; r < bnd[tid + 1]; ++r) {
        double s = 0.0;
        for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += val[k] * x[col[k]];
        y[r] = s;
      }
    }
    const double t1 = now_s();
    if (it >= warmup) times[it - warmup] = t1 - t0;
  }
  qsort(times, (size_t)reps, sizeof(double), cmp_double);
  const double med = reps > 0 ? times[reps / 2] : 0.0;
  if (threads_used) *threads_used = nt;
  free(bnd);
  free(times);
  return med;
}

/*
 * merbit_amd.h -- C ABI of the MI355X (gfx950) merge-path tile SpMV.
 *
 * Plain C: device pointers, sizes and an opaque plan; `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  Products (spmv,
 * spmm, batch, adjoint, sddmm) are stream-ordered, never synchronise the host
 * and allocate nothing; creating or destroying a plan, batch or adjoint
 * allocates and synchronises.
 * Every function returns MBX_OK (0) or a negative status (mbx_status_str).
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository root).  The reference is a C++ template API
 * without an FFI; the C++ host mirror in merbit_amd/host/merbit.hpp keeps its
 * names (Fun_Merbit, TILE, GRAPH, PARAMETER, ENV, Timer, Fun_Check) on top of
 * these entry points.
 *
 * Matrix layout (as the reference kernel consumes it): 0-based CSR with
 * int32 row_ptr[nrow+1], int32 col[nnz] and values[nnz]; nrow + nnz must stay
 * below 2^31 minus one tile.  The caller owns the CSR arrays, x and y.
 */
#ifndef MERBIT_AMD_H
#define MERBIT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MBX_OK = 0,
  MBX_EINVAL = -1,       /* bad argument (null pointer, negative size, ...) */
  MBX_EOVERFLOW = -2,    /* nrow + nnz (+ one tile) does not fit int32 */
  MBX_EUNSUPPORTED = -3, /* items_per_thread / wg_size / k not instantiated */
  MBX_ENOMEM = -4,       /* device allocation failed */
  MBX_EHIP = -5,         /* a HIP runtime call or launch failed */
  MBX_ESCRATCH = -6      /* larger than mbx_spmv_f64_noalloc's static scratch */
};

typedef struct mbx_plan mbx_plan;

/* Replaces TILE<BLOCK,SIGMA>::TILE(offset, nrow, nnz), include/tile.cuh:106-122
 * (with Tile_Kernel, tile.cuh:28-104, and merge_search, utils.cuh:5-29).
 * Runs the merge-path partition search over (row_ptr, nnz index) on `stream`
 * and keeps one (row, nnz) coordinate per tile of wg_size*items_per_thread
 * merge items plus the per-tile carry slots.  items_per_thread plays the
 * reference's SIGMA (instantiated: 2, 4, 6, 8, 10, 12, 14, 16 and 32, whose
 * tiles of 256*32 merge items run as 512 threads of 16 items) and wg_size
 * its BLOCK (256); 0 selects the defaults (8, 256).
 * Synchronises `stream` once: it reads back whether any row crosses a tile
 * boundary (the carry fixup launch is skipped when none does) and the ends
 * row_ptr[0], row_ptr[nrow]; MBX_EINVAL unless they are 0 and nnz. */
int mbx_plan_create(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol,
                    int32_t nnz, int32_t items_per_thread, int32_t wg_size,
                    void* stream, mbx_plan** out);

/* The plan of mbx_plan_create with row-aligned tiles (no reference
 * counterpart: the reference's merge path always splits rows at tile
 * boundaries, tile.cuh:28-104).  Tile t starts at the start of the row open
 * at merge diagonal t*(T - L), T = wg_size*items_per_thread and L the
 * longest row, so no tile holds more than T merge items and no row crosses
 * a tile boundary: products need no carry fixup launch.  For patterns of
 * short rows (banded, stencils, bounded degree): MBX_EUNSUPPORTED when
 * L > T/16.  Rows are summed within one tile, so FP64 results may round
 * differently from mbx_plan_create's (both within the parity bar).
 * Synchronises `stream` twice (the longest row, then the plan statistics);
 * products, clones and the plan options work on it as on any plan. */
int mbx_plan_create_rows(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol,
                         int32_t nnz, int32_t items_per_thread, int32_t wg_size,
                         void* stream, mbx_plan** out);

/* Replaces ~TILE, include/tile.cuh:19-24.  Waits for the device. */
void mbx_plan_destroy(mbx_plan* plan);

/* A second plan object on the same partition with carry slots of its own
 * (no reference counterpart: the reference's plan serves one stream).
 * Products on a plan use its carry slots, so two products on one plan must
 * not overlap in time; a plan and its clones share the read-only tile
 * coordinates and may run products concurrently on different streams when
 * their outputs differ.  Allocates 16*(ntiles+1) bytes; no kernel runs.
 * Every plan object is destroyed on its own; the coordinates are freed with
 * the last one. */
int mbx_plan_clone(const mbx_plan* plan, mbx_plan** out);

/* Plan geometry: number of tiles, merge items per tile, workgroup size. */
int mbx_plan_info(const mbx_plan* plan, int32_t* ntiles, int32_t* tile_items,
                  int32_t* wg_size);

/* Copies the ntiles+1 tile start coordinates (rows consumed, nnz consumed)
 * to host arrays (synchronous; for tests and diagnostics, like TILE::show,
 * include/tile.cuh:124-175). */
int mbx_plan_partition(const mbx_plan* plan, int32_t* h_rows, int32_t* h_nnz,
                       int32_t count);

/* Cache policy of a product's row_ptr/col/value stream (no reference
 * counterpart).  MBX_STREAM_CACHED: plain loads, so a matrix that fits the
 * 256 MB MALL stays there for the next product (repeated products of one
 * matrix).  MBX_STREAM_NONTEMPORAL: non-temporal loads for the parts the
 * value kind gains from (row_ptr and col; FP32/BF16 values too), so data
 * read once does not displace the x lines the gathers reuse (matrices that
 * do not stay resident).  MBX_STREAM_AUTO (the default): non-temporal when
 * the product's bytes exceed 240 MiB.  Results are identical in every mode. */
enum { MBX_STREAM_AUTO = 0, MBX_STREAM_CACHED = 1, MBX_STREAM_NONTEMPORAL = 2 };
int mbx_plan_set_stream(mbx_plan* plan, int32_t mode);

/* Tiles resident per CU for the plan's SpMV launches (no reference
 * counterpart): 0 (the default) lets as many run as registers and LDS
 * allow; 1..16 reserves extra LDS per tile so that at most that many share
 * a CU.  Fewer resident tiles can mean fewer L1 misses for their x gathers
 * (a timing knob for the autotuner; results are identical). */
int mbx_plan_set_residency(mbx_plan* plan, int32_t wg_per_cu);

/* Accumulation of the FP32 / bfloat16 value forms (no reference
 * counterpart).  MBX_ACC_COMPENSATED_F32 (the default): TwoProduct + TwoSum
 * compensated FP32 sums per thread, FP64 partials across threads and tiles.
 * MBX_ACC_F64: each product formed exactly in FP64 (FP32 x FP32 fits its
 * significand) and the rows summed in FP64; x and y stay FP32 and y is
 * rounded once.  With packed FP32 off an FP64 add issues at the FP32 rate,
 * so MBX_ACC_F64 spends one VALU add per nonzero where TwoSum spends about
 * six; the compensated form is the VALU-bound one on HBM-bound matrices.
 * MBX_ACC_F32_PRODUCTS: each product rounded once to FP32 and the rows
 * summed in FP64, a row's partial rounded to FP32 where it is kept in LDS
 * (error at most 3 * 2^-24, about 1.8e-7, of sum |val*x|, y's rounding
 * included); the products take half the LDS of MBX_ACC_F64, so more tiles
 * and their loads fit a CU (the launch rule's choice for local patterns). */
enum { MBX_ACC_COMPENSATED_F32 = 0, MBX_ACC_F64 = 1, MBX_ACC_F32_PRODUCTS = 2 };
int mbx_plan_set_accumulation(mbx_plan* plan, int32_t mode);

/* Compressed column indices (no reference counterpart).  Stores every
 * nonzero's column as a 16-bit offset from its tile's smallest column, so
 * products on this plan read 2 bytes of column index per nonzero instead of
 * 4 (banded and other local patterns: a tile of 2048 merge items spans a few
 * hundred columns).  Applies to plans with items_per_thread 8 or 12 and to
 * products on the vector staging path (16-byte aligned row_ptr, col and
 * values); other products read d_col as before, so d_col must still be the
 * pattern's column array.  Results are bit-identical either way.
 * MBX_EUNSUPPORTED (nothing kept) when a tile spans more than 65536 columns
 * or the plan's items_per_thread is not 8 or 12; d_col NULL drops the
 * compressed columns.  Plan clones read the caller's columns.  Synchronises
 * `stream` (and the device, when it replaces compressed columns). */
int mbx_plan_compress_cols(mbx_plan* plan, const int32_t* d_col, void* stream);

/* Hot-column split (no reference counterpart; the reference gathers
 * x[col] for every nonzero, include/tile.cuh:293-305).  Ranks the columns of
 * d_col by nonzero count and keeps the max_hot most frequent (ties by column
 * index; at most 32768) with a copy of the column array in which their
 * entries name a slot instead.  Products on this plan with items_per_thread
 * 8 or 12 on the vector staging path then gather the hot columns' x values
 * once per product into the plan (nhot values) and run a persistent tile
 * kernel that keeps them in each workgroup's LDS; every other x value is
 * gathered as before.  For power-law patterns, whose hot columns take a
 * large share of the gathers.  Results are bit-identical to the plain
 * product; the hot values must fit one workgroup's LDS beside the tile
 * (FP64 x: about 16000 columns), else the product runs the plain kernel.
 * *covered (may be NULL) receives the share of the nonzeros in the kept
 * columns.  max_hot 0 or d_col NULL drops the split.  The plan's products
 * then read its own column copy, so d_col must stay the pattern's column
 * array.  Allocates 4*nnz + 12*max_hot bytes and synchronises `stream`
 * (and the device, when it replaces a split); clones do not inherit it. */
int mbx_plan_split_hot(mbx_plan* plan, const int32_t* d_col, int32_t max_hot, void* stream,
                       double* covered);

/* Tile-sorted gathers (no reference counterpart; the reference gathers
 * x[col] in the order of the nonzeros, include/tile.cuh:293-305).  Keeps a
 * copy of the column array in which every tile's columns are sorted and
 * laid out in the order the tile kernel gathers them, each word carrying the
 * element's position in its tile: each gather instruction then reads 64
 * consecutive sorted columns, which share x lines, so the vector L1 sees
 * fewer requests (uniform random and power-law columns, whose tiles touch
 * some x lines many times).  The gathered values go back to their elements'
 * product slots in LDS, so the products and results are bit-identical to the
 * plain product's.  Applies to plans with items_per_thread 8, 12 or 16 and
 * to products on the vector staging path; other products read d_col as
 * before, so d_col must stay the pattern's column array.  MBX_EUNSUPPORTED
 * (nothing kept) for other shapes, ncol above 2^21 (items_per_thread 8) or
 * 2^20 (12, 16), or a plan with 16-bit columns or a hot split (which in turn
 * refuse a sorted plan); MBX_EINVAL when a column lies outside [0, ncol).
 * d_col NULL drops the copy.  Allocates 4*nnz bytes and synchronises
 * `stream` (and the device, when it replaces a copy); clones do not inherit
 * it. */
int mbx_plan_sort_cols(mbx_plan* plan, const int32_t* d_col, void* stream);

/* Optional timing hook: when ev_begin and ev_end (hipEvent_t) are both
 * non-NULL, the next spmv or spmm call on this plan attaches them to its tile
 * kernel's dispatch (hipExtLaunchKernel): their interval is that kernel's
 * own execution, as a kernel trace reports it.  The next call clears the
 * plan's events whether or not both were set.  For eager calls; do not set
 * them on a plan whose next call is captured into a graph. */
int mbx_plan_set_events(mbx_plan* plan, void* ev_begin, void* ev_end);

/* Replaces merbit_kernel<double,BLOCK,SIGMA>, include/tile.cuh:476-561, and the
 * block-boundary atomics of load_mem, tile.cuh:354-375: y = alpha*A*x +
 * beta*y in FP64 (beta == 0: y is not read, so it need not be initialised).
 * Deterministic: the cross-tile carry fixup sums in a fixed order. */
int mbx_spmv_f64(const mbx_plan* plan, const int32_t* d_row_ptr,
                 const int32_t* d_col, const double* d_val, double alpha,
                 const double* d_x, double beta, double* d_y, void* stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  y = alpha*A*x + beta*y in FP64 without a plan: the partition
 * search, the tile kernel and the carry fixup of mbx_plan_create(items
 * per thread 8, wg 256) + mbx_spmv_f64 run back to back on `stream`, with the
 * tile coordinates and carry slots in a static 4 MiB device array (one per
 * GPU, held until the process ends), so the result is bit-identical to that
 * pair.  Allocates nothing, never synchronises (capturable); row_ptr must run
 * from 0 to nnz (not checked here: mbx_csr_validate checks it).  Limit: at
 * most 131071 tiles, nrow + nnz <= 268,433,408; larger inputs return
 * MBX_ESCRATCH.  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (other streams, other host threads) race and
 * must be serialised by the caller. */
int mbx_spmv_f64_noalloc(const int32_t* d_row_ptr, const int32_t* d_col, const double* d_val,
                         int32_t nrow, int32_t ncol, int32_t nnz, double alpha,
                         const double* d_x, double beta, double* d_y, void* stream);

/* Mixed precision forms of mbx_spmv_f64 (no reference counterpart: the
 * reference instantiates float and double only, src/class.cu:427-430).
 * Values FP32 (f32) or bfloat16 bit patterns (bf16), x and y FP32; products
 * are split error-free (TwoProduct) and rows summed with TwoSum compensation
 * in FP32; carries of rows that cross tiles are fixed up in FP64. */
int mbx_spmv_f32(const mbx_plan* plan, const int32_t* d_row_ptr,
                 const int32_t* d_col, const float* d_val, float alpha,
                 const float* d_x, float beta, float* d_y, void* stream);
int mbx_spmv_bf16(const mbx_plan* plan, const int32_t* d_row_ptr,
                  const int32_t* d_col, const uint16_t* d_val, float alpha,
                  const float* d_x, float beta, float* d_y, void* stream);

/* Dense operand layouts for SpMM: element (i, j) of an n x k matrix with
 * leading dimension ld is at [i + j*ld] (column-major) or [i*ld + j]. */
enum { MBX_COL_MAJOR = 0, MBX_ROW_MAJOR = 1 };

/* Batched right-hand sides (no reference counterpart; the reference has SpMV
 * only): Y = alpha*A*X + beta*Y with X ncol x k and Y nrow x k in FP32, FP32
 * values, FP32 accumulation, FP64 carries for rows that cross tiles.  Uses a
 * plan made with items_per_thread 2 or 4 (MBX_EUNSUPPORTED otherwise).  The
 * caller provides a 256-byte aligned device workspace of at least
 * mbx_spmm_workspace_bytes(plan, k, X, ldx, x_layout) bytes; a column-major
 * (or unaligned row-major) X is repacked into it as rows once per call.
 * Items per thread 2 suits wide X (k > 8: 16 right-hand sides per pass),
 * 4 suits k <= 8; mbx_autotune_spmm times both. */
size_t mbx_spmm_workspace_bytes(const mbx_plan* plan, int32_t k, const float* d_x,
                                int64_t ldx, int32_t x_layout);
int mbx_spmm_f32(const mbx_plan* plan, const int32_t* d_row_ptr, const int32_t* d_col,
                 const float* d_val, int32_t k, float alpha, const float* d_x, int64_t ldx,
                 int32_t x_layout, float beta, float* d_y, int64_t ldy, int32_t y_layout,
                 void* d_workspace, size_t workspace_bytes, void* stream);

/* Value kinds for the kind-generic entry points below. */
enum { MBX_VALUE_F64 = 0, MBX_VALUE_F32 = 1, MBX_VALUE_BF16 = 2 };

/* Many CSR matrices in one persistent launch (no reference counterpart; the
 * reference times one graph per process, src/main.cu:25-65).  Each item
 * names device arrays already in HBM: values of the batch's value kind
 * (double, float or bfloat16 bits), x and y double for MBX_VALUE_F64 and
 * float otherwise.  mbx_batch_spmv computes y_m = alpha*A_m*x_m + beta*y_m
 * for every item with the same per-tile arithmetic as mbx_spmv_*, so the
 * results are bit-identical to one call per matrix.  items_per_thread takes
 * the plan's values and also 32: tiles of 8192 merge items run by 512
 * threads of 16 items (for mbx_batch_sort_cols, whose sorted gathers share
 * more x lines in a larger tile; members up to 2^19 columns), whose results
 * no single-matrix plan reproduces bit for bit.  mbx_batch_create
 * synchronises `stream` once (it uploads the descriptors). */
typedef struct {
  const int32_t* row_ptr;
  const int32_t* col;
  const void* val;
  const void* x;
  void* y;
  int32_t nrow, ncol, nnz;
} mbx_batch_item;
typedef struct mbx_batch mbx_batch;
int mbx_batch_create(const mbx_batch_item* items, int32_t count, int32_t value_kind,
                     int32_t items_per_thread, void* stream, mbx_batch** out);
void mbx_batch_destroy(mbx_batch* batch);
int mbx_batch_info(const mbx_batch* batch, int32_t* total_tiles, int32_t* tile_items,
                   int32_t* grid);
int mbx_batch_spmv(const mbx_batch* batch, double alpha, double beta, void* stream);
/* The stream cache policy (MBX_STREAM_*) of the batch's products; AUTO
 * decides on the bytes of the whole batch. */
int mbx_batch_set_stream(mbx_batch* batch, int32_t mode);
/* Tiles resident per CU for the batch's launches (mbx_plan_set_residency);
 * the persistent grid shrinks to wg_per_cu per CU. */
int mbx_batch_set_residency(mbx_batch* batch, int32_t wg_per_cu);
/* Tile-sorted gathers for every member (as mbx_plan_sort_cols; no reference
 * counterpart): enable 1 builds the members' sorted column copies from the
 * items' col arrays, 0 drops them.  Batches with items_per_thread 8, 12, 16
 * or 32 whose members all suit the vector staging path; MBX_EUNSUPPORTED
 * otherwise or when a member's ncol exceeds the limit of mbx_plan_sort_cols,
 * MBX_EINVAL when a column lies outside [0, ncol).  Results stay
 * bit-identical.  Allocates 4 bytes per nonzero and synchronises `stream`. */
int mbx_batch_sort_cols(mbx_batch* batch, int32_t enable, void* stream);

/* HIP graphs of repeated products (the reference's timed loop of
 * maxiter*perrnd launches, include/tile.cuh:596-603): `repeats` back-to-back
 * products captured once, replayed by mbx_graph_launch.  The graph holds the
 * pointers given here.  graph_exec is a hipGraphExec_t. */
int mbx_batch_graph_create(const mbx_batch* batch, double alpha, double beta, int32_t repeats,
                           void** graph_exec);
int mbx_spmv_graph_create(const mbx_plan* plan, int32_t value_kind, const int32_t* d_row_ptr,
                          const int32_t* d_col, const void* d_val, double alpha,
                          const void* d_x, double beta, void* d_y, int32_t repeats,
                          void** graph_exec);
int mbx_graph_launch(void* graph_exec, void* stream);
void mbx_graph_destroy(void* graph_exec);

/* PageRank power iterations on the reference's operated matrix (CSR of P^T,
 * values 1/out-degree: mbx_ingest_ref's output), the use of PARAMETER::damp
 * (include/class.cuh:119-146) that the reference's loop (include/tile.cuh:
 * 596-603, y = A*x repeated with x fixed) leaves out.  Each iteration is
 *   y = damp*A*x;  x' = y + (sum(x) - sum(y))/n
 * (teleport plus the damped mass of dangling nodes, spread evenly; the mass
 * sum(x) is kept).  The square plan's product runs `iters` times from d_x0
 * (not modified) and leaves the last iterate in d_y.  d_work: 256-byte
 * aligned device memory of mbx_pagerank_workspace_bytes(nrow) bytes.  FP64;
 * the sums are deterministic.  The graph variant captures all iterations as
 * one HIP graph for mbx_graph_launch. */
size_t mbx_pagerank_workspace_bytes(int32_t nrow);
int mbx_pagerank_f64(const mbx_plan* plan, const int32_t* d_row_ptr, const int32_t* d_col,
                     const double* d_val, double damp, int32_t iters, const double* d_x0,
                     double* d_y, void* d_work, size_t work_bytes, void* stream);
int mbx_pagerank_graph_create(const mbx_plan* plan, const int32_t* d_row_ptr,
                              const int32_t* d_col, const double* d_val, double damp,
                              int32_t iters, const double* d_x0, double* d_y, void* d_work,
                              size_t work_bytes, void** graph_exec);

/* Pattern statistics of a planned matrix (features for choosing the tile
 * shape): row-length moments about the mean and the empty-row count as the
 * reference's GRAPH::desc / GRAPH::dangling report them (src/class.cu:181-320),
 * and tile statistics from the plan (TILE::desc, include/tile.cuh:177-233:
 * single-row tiles are its fast tiles).  Copies row_ptr and the plan's
 * coordinates to the host (synchronous). */
typedef struct mbx_pattern_stats {
  int32_t nrow, ncol, nnz;
  double mean, var, skew, kurt; /* row lengths */
  int32_t empty_rows, max_row;
  int32_t ntiles, tile_items;
  int32_t single_row_tiles;     /* tiles that end no row */
  int32_t split_tiles;          /* tiles whose first row began in an earlier tile */
  int32_t max_rows_per_tile;
  int32_t long_rows;            /* split rows spanning more than 64 tiles */
} mbx_pattern_stats;
int mbx_plan_stats(const mbx_plan* plan, const int32_t* d_row_ptr, mbx_pattern_stats* out);

/* Launch rule: the statistics of a pattern choose its launch without a
 * timing sweep.  The reference computes degree and tile statistics
 * (GRAPH::desc, src/class.cu:223-381; TILE::desc, include/tile.cuh:177-233)
 * but compiles one SIGMA and BLOCK (CMakeLists.txt:4-6) for every matrix.
 * Features: tiles at 8 items per thread, the widest column range of one such
 * tile, x bytes, and the share of the entries in the 4096 (FP64 x) or 7900
 * (FP32 x) most frequent columns.  "Sorted" below: tile-sorted gathers
 * (sort_cols, mbx_plan_sort_cols) at 16 items per thread, taken when the
 * pattern is not local and ncol is at most 2^20; for the few-tiles and
 * x-fits rules at 32 (tiles of 8192 items, 512 threads) when ncol is at
 * most 2^19 and that leaves a tile per CU.  Rules, first match:
 *   MBX_RULE_FEW_TILES       fewer than 16 tiles per CU: sorted when that
 *                            still gives a tile per CU (at 12 items per
 *                            thread when the hot-column share is at least
 *                            1/4, a power law), else 4 items per
 *                            thread (2 when that gives fewer tiles than CUs);
 *   MBX_RULE_LOCAL           every tile spans < 65536 columns (banded):
 *                            8 (FP64) or 12 (FP32, BF16) items per thread, all tiles
 *                            resident, 16-bit columns, FP32 products and
 *                            FP64 sums for FP32/BF16 values;
 *   MBX_RULE_X_FITS          x of at most 2 MiB: sorted, else 8 items per
 *                            thread and at most 4 tiles per CU;
 *   MBX_RULE_HOT_COLUMNS     the most frequent columns hold >= 1/4 of the
 *                            entries: sorted, else 8 and the hot-column
 *                            split; FP32/BF16 values with FP32 products and
 *                            FP64 sums;
 *   MBX_RULE_SPREAD_COLUMNS  otherwise: 8, 3 tiles per CU, and the
 *                            column-blocked form recommended (column_blocks
 *                            for mbx_colblock_create).
 * Whatever the rule, tiles are row-aligned (row_tiles, mbx_plan_create_rows:
 * no carry fixup launch) when the longest row (feature max_row) is at most
 * 1/16 of the chosen tile and items per thread at most 8.  The stream policy stays MBX_STREAM_AUTO.
 * Synchronises `stream` (reads the features back).  mbx_plan_create_auto
 * makes the plan with the choice applied (row-aligned tiles, items per
 * thread, residency, accumulation, 16-bit columns, sorted gathers, hot
 * split; a pattern whose columns the sort refuses keeps the plain gathers);
 * column_blocks is a recommendation for the caller (a plan is a CSR
 * product). */
enum {
  MBX_RULE_FEW_TILES = 1,
  MBX_RULE_LOCAL = 2,
  MBX_RULE_X_FITS = 3,
  MBX_RULE_HOT_COLUMNS = 4,
  MBX_RULE_SPREAD_COLUMNS = 5
};
typedef struct mbx_launch_choice {
  int32_t rule;              /* MBX_RULE_* that decided */
  int32_t items_per_thread;
  int32_t stream_mode;       /* MBX_STREAM_* */
  int32_t wg_per_cu;         /* tiles per CU (0: as many as fit) */
  int32_t accumulation;      /* MBX_ACC_* (FP32 / bfloat16 values) */
  int32_t compress_cols;     /* 1: mbx_plan_compress_cols */
  int32_t hot_columns;       /* mbx_plan_split_hot max_hot (0: none) */
  int32_t column_blocks;     /* > 0: the column-blocked form with this many blocks */
  int32_t row_tiles;         /* 1: row-aligned tiles (mbx_plan_create_rows) */
  int32_t sort_cols;         /* 1: tile-sorted gathers (mbx_plan_sort_cols) */
  /* the features */
  int32_t cus, ntiles8, max_tile_span, max_row;
  double x_bytes, hot_share;
} mbx_launch_choice;
int mbx_launch_rule(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow, int32_t ncol,
                    int32_t nnz, int32_t value_kind, void* stream, mbx_launch_choice* out);
int mbx_plan_create_auto(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                         int32_t ncol, int32_t nnz, int32_t value_kind, void* stream,
                         mbx_plan** out, mbx_launch_choice* choice);

/* Timing autotuner (no reference counterpart: SIGMA and BLOCK are
 * compile-time in the reference, CMakeLists.txt:4-6): times every
 * instantiated items_per_thread on the given operands (which it overwrites in
 * y) and returns the fastest with its time per product in microseconds. */
/* As mbx_autotune_spmv, over the whole launch shape: every items_per_thread
 * with both stream policies (MBX_STREAM_CACHED, MBX_STREAM_NONTEMPORAL) and
 * tile residencies 0, 3, 4, 5 (mbx_plan_set_stream / _set_residency).
 * Returns the fastest triple; apply it to a plan of that items_per_thread
 * with the two setters. */
int mbx_autotune_spmv_launch(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                             const int32_t* d_col, const void* d_val, int32_t value_kind,
                             const void* d_x, void* d_y, int32_t reps, void* stream,
                             int32_t* best_ipt, int32_t* best_stream, int32_t* best_wg_per_cu,
                             float* best_us);
/* As mbx_autotune_spmv_launch, with tile-sorted gathers (mbx_plan_sort_cols)
 * as a further candidate at every items_per_thread that takes them;
 * *best_sorted is 1 when the fastest launch sorts (apply it with
 * mbx_plan_sort_cols on the plan of that shape).  No reference counterpart. */
int mbx_autotune_spmv_form(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                           const int32_t* d_col, const void* d_val, int32_t value_kind,
                           const void* d_x, void* d_y, int32_t reps, void* stream,
                           int32_t* best_ipt, int32_t* best_stream, int32_t* best_wg_per_cu,
                           int32_t* best_sorted, float* best_us);
int mbx_autotune_spmv(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                      const int32_t* d_col, const void* d_val, int32_t value_kind,
                      const void* d_x, void* d_y, int32_t reps, void* stream, int32_t* best_ipt,
                      float* best_us);
int mbx_autotune_spmm(const int32_t* d_row_ptr, int32_t nrow, int32_t ncol, int32_t nnz,
                      const int32_t* d_col, const float* d_val, int32_t k, const float* d_x,
                      int64_t ldx, int32_t x_layout, float* d_y, int64_t ldy, int32_t y_layout,
                      void* d_workspace, size_t workspace_bytes, int32_t reps, void* stream,
                      int32_t* best_ipt, float* best_us);
int mbx_autotune_batch(const mbx_batch_item* items, int32_t count, int32_t value_kind,
                       int32_t reps, void* stream, int32_t* best_ipt, float* best_us);

/* Column-blocked form of a pattern (no reference counterpart).  For products
 * whose x gathers spread evenly over an x larger than one XCD's 4 MB L2 (a
 * random graph), the columns are split into nblocks ranges and A into the
 * CSR matrices A_j over all rows keeping the columns of range j; a product
 * chains y = alpha*A_0*x_0 + beta*y, then y = alpha*A_j*x_j + y, each a
 * merge-path SpMV whose x slice stays in L2.  Row sums are split over the
 * blocks and added in block order: deterministic, rounded differently from
 * mbx_spmv_* on the CSR.
 *   mbx_colblock_create: builds the block-major pattern (block after block,
 *       rows in order, a row's entries in CSR order; block starts padded to
 *       4 slots) and one plan per block; nblocks 0..16 (0: slices of at most
 *       3.5 MiB of 4-byte x, the FP32/BF16 value kinds; FP64 x is twice as
 *       long, so FP64 callers pass twice that count), clamped to ncol.
 *       Synchronises `stream`.
 *   mbx_colblock_info: block count, value slots (the length of a blocked
 *       value array: nnz, the block padding and 64 spare slots after the
 *       last block) and the nblocks+1 column bounds.
 *   mbx_colblock_values: moves values of the given kind from CSR order into
 *       a blocked array of value_slots entries (padding slots set to 0);
 *       mbx_colblock_values_to_csr moves them back (nnz entries).
 *   mbx_colblock_spmv: the chained product on blocked values; x and y as
 *       for mbx_spmv_f64 (MBX_VALUE_F64) or mbx_spmv_f32/bf16 (FP32 x, y).
 *   mbx_colblock_sddmm: dval = g[row]*x[col] in block-major order (FP64 or
 *       FP32), as mbx_sddmm on each block with its x slice; padding slots
 *       are not written.
 *   mbx_colblock_adjoint_prepare: builds the column-blocked form of A^T
 *       (blocks over A's rows, nblocks 0..16 as above, reported in
 *       *nblocks_out when not NULL) from the same CSR pattern the form was
 *       made from (checked: MBX_EINVAL when the row lengths or the column
 *       of any entry differ, or the pattern is malformed), and a two-pass bucketed
 *       permutation from this form's slots to that form's.  Synchronises
 *       `stream`.
 *   mbx_colblock_dx: dx = alpha*A^T*g + beta*dx from values in this form's
 *       block-major order: the permutation (into buffers the form owns),
 *       then the A^T form's chained product, whose g slices stay in L2.
 *       MBX_EINVAL before mbx_colblock_adjoint_prepare.
 *   mbx_colblock_set_stream / _set_residency: the stream cache policy and
 *       tiles per CU of every block's product (and of the A^T form's), as
 *       mbx_plan_set_stream / mbx_plan_set_residency; MBX_STREAM_AUTO (the
 *       default) judges the bytes of the whole chained product, and the
 *       residency defaults to 4 tiles per CU (the gather-bound block
 *       products' best).  Calls on one
 *       column-blocked form must not overlap (they share its plans' carry
 *       slots and its dx value buffer). */
typedef struct mbx_colblock mbx_colblock;
int mbx_colblock_create(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                        int32_t ncol, int32_t nnz, int32_t nblocks, int32_t items_per_thread,
                        void* stream, mbx_colblock** out);
void mbx_colblock_destroy(mbx_colblock* cb);
int mbx_colblock_info(const mbx_colblock* cb, int32_t* nblocks, int64_t* value_slots,
                      int32_t* col_bounds);
int mbx_colblock_values(const mbx_colblock* cb, int32_t value_kind, const void* d_val,
                        void* d_val_blocked, void* stream);
int mbx_colblock_values_to_csr(const mbx_colblock* cb, int32_t value_kind,
                               const void* d_val_blocked, void* d_val, void* stream);
int mbx_colblock_spmv(const mbx_colblock* cb, int32_t value_kind, const void* d_val_blocked,
                      double alpha, const void* d_x, double beta, void* d_y, void* stream);
int mbx_colblock_sddmm(const mbx_colblock* cb, int32_t value_kind, const void* d_g,
                       const void* d_x, void* d_dval_blocked, void* stream);
int mbx_colblock_set_stream(mbx_colblock* cb, int32_t mode);
int mbx_colblock_set_residency(mbx_colblock* cb, int32_t wg_per_cu);
/* Tile-sorted gathers in every block product (mbx_plan_sort_cols on each
 * block's plan, and on the A^T form's once prepared; no reference
 * counterpart): enable 1 sorts, 0 drops.  MBX_EUNSUPPORTED (nothing kept)
 * when the blocks' items_per_thread is not 8, 12 or 16 or a block spans
 * more columns than the sorted words hold; results stay bit-identical.
 * Synchronises `stream`. */
int mbx_colblock_sort_cols(mbx_colblock* cb, int32_t enable, void* stream);
int mbx_colblock_adjoint_prepare(mbx_colblock* cb, const int32_t* d_row_ptr,
                                 const int32_t* d_col, int32_t nblocks,
                                 int32_t items_per_thread, void* stream, int32_t* nblocks_out);
int mbx_colblock_dx(const mbx_colblock* cb, int32_t value_kind, const void* d_val_blocked,
                    double alpha, const void* d_g, double beta, void* d_dx, void* stream);

/* Adjoint of y = A*x (no reference counterpart).  An adjoint object holds the
 * transposed pattern (CSC of A built on the device by a stable radix sort:
 * rows stay ascending within every column, so dx sums deterministically) and
 * its merge-path plan.  mbx_adjoint_create synchronises `stream` once.
 *   mbx_adjoint_dx:  dx = alpha*A^T*g + beta*dx, values of the given kind
 *                    gathered into CSC order, then the tile SpMV on A^T.
 *   mbx_sddmm:       dval[k] = g[row(k)] * x[col(k)] over the pattern of a
 *                    forward plan (FP64 or FP32 g, x and dval).
 *   mbx_adjoint_transpose: host copies of col_ptr[ncol+1], row_idx[nnz] and
 *                    perm[nnz] (CSR position of each CSC entry); any may be NULL. */
typedef struct mbx_adjoint mbx_adjoint;
int mbx_adjoint_create(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                       int32_t ncol, int32_t nnz, int32_t items_per_thread, void* stream,
                       mbx_adjoint** out);
void mbx_adjoint_destroy(mbx_adjoint* adj);
int mbx_adjoint_transpose(const mbx_adjoint* adj, int32_t* h_col_ptr, int32_t* h_row_idx,
                          int32_t* h_perm);
int mbx_adjoint_dx(const mbx_adjoint* adj, int32_t value_kind, const void* d_val, double alpha,
                   const void* d_g, double beta, void* d_dx, void* stream);
/* Optional: dx through a column-blocked form of A^T (mbx_colblock_*, blocks
 * of A's rows), for graphs whose g outgrows one XCD's L2.  One two-pass
 * bucketed permutation moves the CSR-order values straight into that
 * form's order (its padding slots read virtual zero sources), then its
 * chained product sums dx.  nblocks: 1..16, -1 for slices of at most
 * 3.5 MiB of 4-byte g, 0 to return to the single merge-path product on
 * A^T.  Synchronises `stream` (and the device when it replaces a blocked
 * form). */
int mbx_adjoint_set_blocks(mbx_adjoint* a, int32_t nblocks, void* stream);
/* The launch rule (mbx_launch_rule) for the transposed pattern, for values
 * of value_kind: the plan of A^T is rebuilt with the rule's shape and
 * options (mbx_plan_create_auto on col_ptr / row_idx: items per thread,
 * tiles per CU, 16-bit columns, FP64 sums, the hot-column split of A^T's
 * most frequent columns, i.e. A's longest rows), and dx goes through the
 * column-blocked A^T when the rule finds spread columns (its
 * column_blocks), through the single product otherwise.  *choice (may be
 * NULL) receives the rule's choice.  Synchronises `stream` (and the device
 * when it replaces a blocked form). */
int mbx_adjoint_auto(mbx_adjoint* a, int32_t value_kind, void* stream, mbx_launch_choice* choice);

int mbx_sddmm(const mbx_plan* plan, int32_t value_kind, const int32_t* d_row_ptr,
              const int32_t* d_col, const void* d_g, const void* d_x, void* d_dval,
              void* stream);

/* Reference-semantics input preparation on the device: the CSR of A^T of a
 * square COO graph (n x n, int32 indices, duplicates kept) with values
 * 1/out-degree(row) -- Fun_Values (src/class.cu:9-34), GRAPH::sort(false)
 * (src/class.cu:143-177) and cusparseXcoo2csr (include/tile.cuh:571-574).
 * Device arrays: d_row_ind, d_col_ind [nnz] in; d_row_ptr_t [n+1],
 * d_col_t [nnz] (ascending within rows), d_val_t [nnz] (FP64 or FP32 by
 * value_kind) out.  An index outside [0, n) is MBX_EINVAL.  Allocates
 * temporaries and synchronises `stream`. */
int mbx_ingest_ref(int32_t n, int32_t nnz, const int32_t* d_row_ind, const int32_t* d_col_ind,
                   int32_t value_kind, int32_t* d_row_ptr_t, int32_t* d_col_t, void* d_val_t,
                   void* stream);

/* Structural check of a device-resident CSR pattern (no reference
 * counterpart: the reference builds its CSR itself and trusts it,
 * src/class.cu:143-177).  The products read x through a bounded buffer
 * resource, so a column index outside [0, ncol) reads 0 there instead of
 * faulting; call this once on untrusted input to have such a pattern
 * reported instead.  Counts:
 *   row_ptr_errors  rows r with row_ptr[r] > row_ptr[r+1], or an end outside
 *                   [0, nnz], plus row_ptr[0] != 0;
 *   col_errors      entries whose column is outside [0, ncol)
 *                   (first_bad_col: the first such position, -1 if none);
 *   unsorted_pairs  adjacent entries of one row with descending columns (the
 *                   reference's layout has ascending columns; the products
 *                   do not need them, so this alone is not an error);
 *   nnz_mismatch    1 when row_ptr[nrow] != nnz.
 * Returns MBX_EINVAL when any count but unsorted_pairs is non-zero (the
 * report is filled), MBX_OK otherwise.  Allocates nnz + 1 bytes of
 * temporaries and synchronises `stream`.  mbx_adjoint_create and
 * mbx_colblock_adjoint_prepare run the same check and refuse such input with
 * MBX_EINVAL; mbx_plan_create refuses row_ptr[0] != 0 and
 * row_ptr[nrow] != nnz. */
typedef struct mbx_csr_report {
  int64_t row_ptr_errors;
  int64_t col_errors;
  int64_t unsorted_pairs;
  int32_t nnz_mismatch;
  int32_t first_bad_col;
} mbx_csr_report;
int mbx_csr_validate(const int32_t* d_row_ptr, const int32_t* d_col, int32_t nrow,
                     int32_t ncol, int32_t nnz, void* stream, mbx_csr_report* out);

/* Status text, like the CSR5 error codes (include/csr5/detail/common.h:13-22). */
const char* mbx_status_str(int status);

#ifdef __cplusplus
}
#endif
#endif /* MERBIT_AMD_H */

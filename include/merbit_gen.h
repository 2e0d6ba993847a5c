/*
 * merbit_gen.h -- host-side workload generators and CSR utilities (C ABI).
 *
 * The reference loads its matrices from MTX files named in a config registry
 * (src/class.cu:60-141, config/config.txt); these generators make the
 * synthetic matrices the benchmarks and tests use, bit-reproducibly, from a
 * project-defined PRNG (splitmix64 seeding xoshiro256**).  All arrays are
 * host memory owned by the caller.  Functions return 0 or a negative
 * MBX_* status from merbit_amd.h.
 */
#ifndef MERBIT_GEN_H
#define MERBIT_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Uniform [lo, hi) and standard normal (Box-Muller) fills. */
int mbg_fill_uniform_f64(double* out, int64_t n, double lo, double hi, uint64_t seed);
int mbg_fill_uniform_f32(float* out, int64_t n, float lo, float hi, uint64_t seed);
int mbg_fill_normal_f64(double* out, int64_t n, uint64_t seed);
int mbg_fill_normal_f32(float* out, int64_t n, uint64_t seed);

/* FP32 -> bfloat16 bit patterns, round to nearest even (NaN kept quiet). */
int mbg_f32_to_bf16(const float* in, uint16_t* out, int64_t n);

/* 5-point Laplacian on a g x g grid, natural order i = r*g + c, diagonal 4,
 * neighbours -1, columns sorted.  nnz = 5*g*g - 4*g. */
int64_t mbg_laplacian5_nnz(int32_t g);
int mbg_laplacian5(int32_t g, int32_t* row_ptr, int32_t* col, double* val);

/* R-MAT (Chakrabarti et al.) with quadrant probabilities a, b, c (d = 1-a-b-c)
 * on 2^scale vertices, `edges` samples, duplicates kept as separate entries,
 * columns sorted within rows.  row_ptr[2^scale + 1], col[edges]. */
int mbg_rmat(int32_t scale, int64_t edges, double a, double b, double c, uint64_t seed,
             int32_t* row_ptr, int32_t* col);

/* Banded n x n matrix with diagonals at offsets -h..h (bandwidth 2h+1).
 * nnz = (2h+1)*n - h*(h+1).  Values are left to the caller. */
int64_t mbg_banded_nnz(int32_t n, int32_t h);
int mbg_banded(int32_t n, int32_t h, int32_t* row_ptr, int32_t* col);

/* Random n x ncol pattern: row lengths ~ Poisson(lambda) capped at ncol,
 * distinct uniformly drawn columns, sorted.  Two calls: first with col == NULL
 * to get row_ptr (and so nnz = row_ptr[n]), then with col to fill it; the
 * same seed gives the same matrix. */
int mbg_poisson_rows(int32_t n, int32_t ncol, double lambda, uint64_t seed,
                     int32_t* row_ptr, int32_t* col);

#ifdef __cplusplus
}
#endif
#endif /* MERBIT_GEN_H */

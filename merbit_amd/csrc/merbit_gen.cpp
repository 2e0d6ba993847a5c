// merbit_gen.cpp -- synthetic workloads (R-MAT, Laplacian, banded, Poisson
// random patterns) and value fills for the merge-path SpMV benchmarks.
//
// PRNG: splitmix64 expands the 64-bit seed into the xoshiro256** state
// (Blackman & Vigna); uniform doubles take the top 53 bits.  Everything is
// sequential and bit-reproducible across compilers and platforms.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "merbit_amd.h"
#include "merbit_gen.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t v = z;
      v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
      v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
      s[i] = v ^ (v >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  // Box-Muller, one normal per pair of uniforms (1-u keeps the log finite).
  double normal() {
    const double u1 = 1.0 - uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // Knuth's multiplicative Poisson sampler (fine for lambda <= ~100).
  int poisson(double lambda) {
    const double L = std::exp(-lambda);
    int k = 0;
    double p = 1.0;
    do {
      ++k;
      p *= uniform();
    } while (p > L);
    return k - 1;
  }
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

}  // namespace

extern "C" {

int mbg_fill_uniform_f64(double* out, int64_t n, double lo, double hi, uint64_t seed) {
  if (n < 0 || (n > 0 && !out)) return MBX_EINVAL;
  Rng g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * g.uniform();
  return MBX_OK;
}

int mbg_fill_uniform_f32(float* out, int64_t n, float lo, float hi, uint64_t seed) {
  if (n < 0 || (n > 0 && !out)) return MBX_EINVAL;
  Rng g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = (float)(lo + (double)(hi - lo) * g.uniform());
  return MBX_OK;
}

int mbg_fill_normal_f64(double* out, int64_t n, uint64_t seed) {
  if (n < 0 || (n > 0 && !out)) return MBX_EINVAL;
  Rng g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = g.normal();
  return MBX_OK;
}

int mbg_fill_normal_f32(float* out, int64_t n, uint64_t seed) {
  if (n < 0 || (n > 0 && !out)) return MBX_EINVAL;
  Rng g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = (float)g.normal();
  return MBX_OK;
}

int mbg_f32_to_bf16(const float* in, uint16_t* out, int64_t n) {
  if (n < 0 || (n > 0 && (!in || !out))) return MBX_EINVAL;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u;
    std::memcpy(&u, &in[i], 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
      out[i] = (uint16_t)((u >> 16) | 0x0040u);  // quiet NaN
    } else {
      const uint32_t lsb = (u >> 16) & 1u;
      out[i] = (uint16_t)((u + 0x7FFFu + lsb) >> 16);
    }
  }
  return MBX_OK;
}

int64_t mbg_laplacian5_nnz(int32_t g) {
  if (g <= 0) return 0;
  return 5ll * g * g - 4ll * g;
}

int mbg_laplacian5(int32_t g, int32_t* row_ptr, int32_t* col, double* val) {
  if (g <= 0 || !row_ptr || !col || !val) return MBX_EINVAL;
  int64_t k = 0;
  row_ptr[0] = 0;
  for (int32_t r = 0; r < g; ++r)
    for (int32_t c = 0; c < g; ++c) {
      const int32_t i = r * g + c;
      auto put = [&](int32_t j, double v) {
        col[k] = j;
        val[k] = v;
        ++k;
      };
      if (r > 0) put(i - g, -1.0);
      if (c > 0) put(i - 1, -1.0);
      put(i, 4.0);
      if (c < g - 1) put(i + 1, -1.0);
      if (r < g - 1) put(i + g, -1.0);
      row_ptr[i + 1] = (int32_t)k;
    }
  return MBX_OK;
}

int mbg_rmat(int32_t scale, int64_t edges, double a, double b, double c, uint64_t seed,
             int32_t* row_ptr, int32_t* col) {
  if (scale < 1 || scale > 30 || edges < 0 || edges >= INT32_MAX || !row_ptr ||
      (edges > 0 && !col))
    return MBX_EINVAL;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return MBX_EINVAL;
  const int64_t n = 1ll << scale;
  std::vector<uint32_t> er((size_t)edges), ec((size_t)edges);
  Rng g(seed);
  const double ab = a + b, abc = a + b + c;
  for (int64_t e = 0; e < edges; ++e) {
    uint32_t r = 0, cc = 0;
    for (int l = 0; l < scale; ++l) {
      const double u = g.uniform();
      const uint32_t rb = u >= ab;                      // lower half for c, d
      const uint32_t cb = (u >= a && u < ab) || u >= abc;  // right half for b, d
      r = (r << 1) | rb;
      cc = (cc << 1) | cb;
    }
    er[(size_t)e] = r;
    ec[(size_t)e] = cc;
  }
  std::memset(row_ptr, 0, sizeof(int32_t) * (size_t)(n + 1));
  for (int64_t e = 0; e < edges; ++e) row_ptr[er[(size_t)e] + 1]++;
  for (int64_t i = 0; i < n; ++i) row_ptr[i + 1] += row_ptr[i];
  std::vector<int32_t> fill(row_ptr, row_ptr + n);
  for (int64_t e = 0; e < edges; ++e) col[fill[er[(size_t)e]]++] = (int32_t)ec[(size_t)e];
  for (int64_t i = 0; i < n; ++i) std::sort(col + row_ptr[i], col + row_ptr[i + 1]);
  return MBX_OK;
}

int64_t mbg_banded_nnz(int32_t n, int32_t h) {
  if (n <= 0 || h < 0) return 0;
  int64_t nnz = 0;
  for (int32_t d = -h; d <= h; ++d) nnz += std::max<int64_t>(0, (int64_t)n - std::abs(d));
  return nnz;
}

int mbg_banded(int32_t n, int32_t h, int32_t* row_ptr, int32_t* col) {
  if (n <= 0 || h < 0 || !row_ptr || !col) return MBX_EINVAL;
  int64_t k = 0;
  row_ptr[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t lo = std::max(0, i - h), hi = std::min(n - 1, i + h);
    for (int32_t j = lo; j <= hi; ++j) col[k++] = j;
    row_ptr[i + 1] = (int32_t)k;
  }
  return MBX_OK;
}

int mbg_poisson_rows(int32_t n, int32_t ncol, double lambda, uint64_t seed, int32_t* row_ptr,
                     int32_t* col) {
  if (n < 0 || ncol <= 0 || lambda < 0 || lambda > 200 || !row_ptr) return MBX_EINVAL;
  Rng lens(seed);
  int64_t k = 0;
  row_ptr[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int l = std::min(lens.poisson(lambda), ncol);
    k += l;
    if (k >= INT32_MAX) return MBX_EOVERFLOW;
    row_ptr[i + 1] = (int32_t)k;
  }
  if (!col) return MBX_OK;
  Rng cols(seed ^ 0x5DEECE66Dull);
  std::vector<int32_t> tmp;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t want = row_ptr[i + 1] - row_ptr[i];
    tmp.clear();
    while ((int32_t)tmp.size() < want) {
      while ((int32_t)tmp.size() < want) tmp.push_back((int32_t)cols.below((uint32_t)ncol));
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    }
    std::copy(tmp.begin(), tmp.end(), col + row_ptr[i]);
  }
  return MBX_OK;
}

}  // extern "C"

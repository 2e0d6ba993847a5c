// merbit_host.cpp -- reference-semantics graph ingest on the host.
//
// Reads the config registry and MTX files the way GRAPH::GRAPH does
// (src/class.cu:60-141) and builds the CSR of A^T with 1/out-degree values
// that Fun_Merbit operates on (src/class.cu:22-34,143-177;
// include/tile.cuh:571-574), with counting sorts instead of a device sort:
// the result is the same exact integer arrangement.

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "merbit_amd.h"
#include "merbit_host.h"

namespace {

struct File {
  FILE* f;
  explicit File(const char* p) : f(std::fopen(p, "r")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

bool read_line(FILE* f, std::string& line) {
  line.clear();
  char buf[4096];
  while (std::fgets(buf, sizeof buf, f)) {
    line += buf;
    if (!line.empty() && line.back() == '\n') break;
  }
  return !line.empty();
}

bool blank(const std::string& s) {
  for (char c : s)
    if (c != ' ' && c != '\t' && c != '\r' && c != '\n') return false;
  return true;
}

}  // namespace

extern "C" {

int mbh_config_lookup(const char* config_file, int32_t gid, char* path_out, int32_t cap) {
  if (!config_file || !path_out || cap <= 0) return MBX_EINVAL;
  File fh(config_file);
  if (!fh.f) return MBX_EINVAL;
  std::string line;
  while (read_line(fh.f, line)) {
    char* end = nullptr;
    const char* s = line.c_str();
    errno = 0;
    const long id = std::strtol(s, &end, 10);
    if (end == s || errno) continue;
    if (id != gid) continue;
    while (*end == ' ' || *end == '\t') ++end;
    const char* p = end;
    while (*end && *end != ' ' && *end != '\t' && *end != '\r' && *end != '\n') ++end;
    const size_t n = (size_t)(end - p);
    if (n == 0 || n + 1 > (size_t)cap) return MBX_EINVAL;
    std::memcpy(path_out, p, n);
    path_out[n] = 0;
    return MBX_OK;
  }
  return MBX_EINVAL;
}

int mbh_mtx_read(const char* path, int32_t* nrow, int64_t* nnz, int32_t* row_ind,
                 int32_t* col_ind, double* val) {
  if (!path || !nrow || !nnz) return MBX_EINVAL;
  File fh(path);
  if (!fh.f) return MBX_EINVAL;
  std::string line;
  bool first = true, symmetric = false, zero_base = false;
  // header and comments (class.cu:84-95)
  for (;;) {
    if (!read_line(fh.f, line)) return MBX_EINVAL;
    if (first) {
      symmetric = line.find("symmetric") != std::string::npos;
      zero_base = line.find("basezero") != std::string::npos;
      first = false;
    }
    if (line[0] != '%' && !blank(line)) break;
  }
  long long R = 0, C = 0, N = 0;
  if (std::sscanf(line.c_str(), "%lld %lld %lld", &R, &C, &N) != 3) return MBX_EINVAL;
  if (R < 0 || C < 0 || N < 0 || R != C || R > INT32_MAX) return MBX_EINVAL;
  const long long total = symmetric ? 2 * N : N;
  if (total > INT32_MAX) return MBX_EOVERFLOW;
  *nrow = (int32_t)R;
  *nnz = total;
  if (!row_ind) return MBX_OK;
  if (total > 0 && !col_ind) return MBX_EINVAL;
  const long long base = zero_base ? 0 : 1;
  long long i = 0;
  while (read_line(fh.f, line)) {
    if (blank(line) || line[0] == '%') continue;
    if (i >= N) return MBX_EINVAL;  // more entries than declared
    long long r, c;
    double v = 1.0;
    const int got = std::sscanf(line.c_str(), "%lld %lld %lf", &r, &c, &v);
    if (got < 2) return MBX_EINVAL;
    r -= base;
    c -= base;
    if (r < 0 || r >= R || c < 0 || c >= R) return MBX_EINVAL;
    row_ind[i] = (int32_t)r;
    col_ind[i] = (int32_t)c;
    if (val) val[i] = v;
    if (symmetric) {  // mirror, diagonal included (class.cu:116-121)
      row_ind[N + i] = (int32_t)c;
      col_ind[N + i] = (int32_t)r;
      if (val) val[N + i] = v;
    }
    ++i;
  }
  return i == N ? MBX_OK : MBX_EINVAL;
}

int mbh_ref_ingest(int32_t nrow, int64_t nnz, const int32_t* row_ind, const int32_t* col_ind,
                   int32_t* row_ptr_t, int32_t* col_t, double* val_t) {
  if (nrow < 0 || nnz < 0 || nnz > INT32_MAX || !row_ptr_t) return MBX_EINVAL;
  if (nnz > 0 && (!row_ind || !col_ind || !col_t || !val_t)) return MBX_EINVAL;
  const size_t n = (size_t)nrow, m = (size_t)nnz;
  std::vector<int32_t> deg(n, 0);
  for (size_t k = 0; k < m; ++k) {
    if (row_ind[k] < 0 || row_ind[k] >= nrow || col_ind[k] < 0 || col_ind[k] >= nrow)
      return MBX_EINVAL;
    deg[(size_t)row_ind[k]]++;
  }
  // stable counting sort by row, then by column: order (col, row)
  std::vector<int32_t> start(n + 1, 0), byrow(m);
  for (size_t r = 0; r < n; ++r) start[r + 1] = start[r] + deg[r];
  {
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (size_t k = 0; k < m; ++k) byrow[(size_t)fill[(size_t)row_ind[k]]++] = (int32_t)k;
  }
  std::memset(row_ptr_t, 0, sizeof(int32_t) * (n + 1));
  for (size_t k = 0; k < m; ++k) row_ptr_t[(size_t)col_ind[k] + 1]++;
  for (size_t c = 0; c < n; ++c) row_ptr_t[c + 1] += row_ptr_t[c];
  std::vector<int32_t> fill(row_ptr_t, row_ptr_t + n);
  for (size_t i = 0; i < m; ++i) {
    const size_t k = (size_t)byrow[i];
    const size_t p = (size_t)fill[(size_t)col_ind[k]]++;
    col_t[p] = row_ind[k];
    val_t[p] = 1.0 / (double)deg[(size_t)row_ind[k]];
  }
  return MBX_OK;
}

}  // extern "C"

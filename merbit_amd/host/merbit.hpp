// merbit.hpp -- C++ host mirror of the reference's SpMV host API.
//
// Same names and roles as the reference (include/class.cuh, include/tile.cuh,
// include/utils.cuh), rebuilt over the C ABI of libmerbit_amd.so:
//
//   Timer       include/class.cuh:5-31    HIP events; start() syncs the device
//   ENV         include/class.cuh:36-59   device + stream (no cuSPARSE/cuBLAS)
//   GRAPH<V>    include/class.cuh:62-81   the loaded graph as COO (row_ind,
//               src/class.cu:60-177       col_ind, values 1/out-degree; sort)
//                                         and as the operated CSR (A^T), on device
//   PARAMETER   include/class.cuh:119-146 x (= 1) and y on device
//   TILE        include/tile.cuh:6-25     the merge-path plan (mbx_plan)
//   Fun_Merbit  include/tile.cuh:563-615  plan + maxiter*perrnd timed SpMVs
//   Fun_Check   include/utils.cuh:116-132 per-row relative check
//   FSTR        include/inc.cuh:70        the global result log stream
//
// Errors throw std::runtime_error (the reference's TRYCATCH, inc.cuh:124-132,
// catches them in the driver).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "merbit_amd.h"
#include "merbit_host.h"

inline std::fstream FSTR;

namespace merbit {

inline void check(int st, const char* what) {
  if (st != MBX_OK)
    throw std::runtime_error(std::string(what) + ": " + mbx_status_str(st));
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* device_copy(const std::vector<T>& h) {
  T* d = nullptr;
  const size_t bytes = sizeof(T) * (h.empty() ? 1 : h.size());
  hip_check(hipMalloc((void**)&d, bytes), "hipMalloc");
  if (!h.empty()) hip_check(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice),
                            "hipMemcpy");
  return d;
}

template <class T>
void device_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace merbit

struct Timer {
  hipEvent_t start_event, stop_event;
  Timer() {
    merbit::hip_check(hipEventCreate(&start_event), "hipEventCreate");
    merbit::hip_check(hipEventCreate(&stop_event), "hipEventCreate");
  }
  void start(hipStream_t stream) {
    merbit::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    merbit::hip_check(hipEventRecord(start_event, stream), "hipEventRecord");
  }
  float stop(hipStream_t stream) {
    merbit::hip_check(hipEventRecord(stop_event, stream), "hipEventRecord");
    merbit::hip_check(hipEventSynchronize(stop_event), "hipEventSynchronize");
    float ms = 0.f;
    merbit::hip_check(hipEventElapsedTime(&ms, start_event, stop_event), "hipEventElapsedTime");
    return ms;
  }
  ~Timer() {
    (void)hipEventDestroy(start_event);
    (void)hipEventDestroy(stop_event);
  }
  Timer(const Timer&) = delete;
  Timer& operator=(const Timer&) = delete;
};

struct ENV {
  const int device_id = 0;
  const std::string config;
  const std::string log;
  hipStream_t stream = nullptr;
  ENV(int id, const std::string& config, const std::string& log)
      : device_id(id), config(config), log(log) {
    merbit::hip_check(hipSetDevice(device_id), "hipSetDevice");
    merbit::hip_check(hipStreamCreate(&stream), "hipStreamCreate");
  }
  ~ENV() { (void)hipStreamDestroy(stream); }
  ENV(const ENV&) = delete;
  ENV& operator=(const ENV&) = delete;
};

// The loaded graph in HBM, in the reference's two views:
//  * COO (include/class.cuh:62-81): entry k is (row_ind[k], col_ind[k],
//    values[k]) -- an edge of the loaded graph with value 1/out-degree of its
//    row (Fun_Values, src/class.cu:22-34); sort(false) orders the entries by
//    (col_ind, row_ind), sort(true) by (row_ind, col_ind) (class.cu:143-177);
//  * the operated CSR (offset, col, values): A^T after sort(false) and
//    cusparseXcoo2csr over col_ind (include/tile.cuh:571-574), so col is
//    row_ind (the same device array) and offset the row pointer of col_ind.
// Entries stay in the sort(false) order unless sort(true) is called; the
// CSR view is valid in that order only (csr_valid).  Host copies are kept
// for verification.
template <typename V>
struct GRAPH {
  int graph_id = -1;
  int nrow = 0;
  int nnz = 0;
  bool owner = true;
  int* row_ind = nullptr;  // COO row (source) of each entry; the CSR's col
  int* col_ind = nullptr;  // COO column (target) of each entry; the CSR's row
  V* values = nullptr;
  int* offset = nullptr;   // CSR row pointer over col_ind (sort(false) order)
  int* col = nullptr;      // == row_ind
  bool csr_valid = true;
  std::vector<int32_t> h_offset, h_col, h_col_ind;
  std::vector<V> h_values;

  GRAPH() = default;

  // src/class.cu:60-141: config ID -> MTX path -> COO -> reference ingest
  GRAPH(const std::string& config, int gid) : graph_id(gid) {
    char path[4096];
    merbit::check(mbh_config_lookup(config.c_str(), gid, path, sizeof path),
                  ("config ID " + std::to_string(gid) + " in " + config).c_str());
    load_mtx(path);
  }

  static GRAPH from_mtx(const std::string& path, int gid = -1) {
    GRAPH g;
    g.graph_id = gid;
    g.load_mtx(path);
    return g;
  }

  // A matrix given directly as host CSR (synthetic workloads).
  GRAPH(int gid, std::vector<int32_t> offset_h, std::vector<int32_t> col_h, std::vector<V> val_h)
      : graph_id(gid), h_offset(std::move(offset_h)), h_col(std::move(col_h)),
        h_values(std::move(val_h)) {
    nrow = (int)h_offset.size() - 1;
    nnz = (int)h_col.size();
    upload();
  }

  GRAPH(GRAPH&& o) noexcept { *this = std::move(o); }
  GRAPH& operator=(GRAPH&& o) noexcept {
    std::swap(graph_id, o.graph_id);
    std::swap(nrow, o.nrow);
    std::swap(nnz, o.nnz);
    std::swap(owner, o.owner);
    std::swap(row_ind, o.row_ind);
    std::swap(col_ind, o.col_ind);
    std::swap(values, o.values);
    std::swap(offset, o.offset);
    std::swap(col, o.col);
    std::swap(csr_valid, o.csr_valid);
    h_offset.swap(o.h_offset);
    h_col.swap(o.h_col);
    h_col_ind.swap(o.h_col_ind);
    h_values.swap(o.h_values);
    return *this;
  }
  GRAPH(const GRAPH&) = delete;
  GRAPH& operator=(const GRAPH&) = delete;

  ~GRAPH() {
    if (owner) {
      merbit::device_free(offset);
      merbit::device_free(row_ind);  // col is the same array
      merbit::device_free(col_ind);
      merbit::device_free(values);
    }
    col = nullptr;
  }

  // src/class.cu:143-177: dir false sorts the entries by (col_ind, row_ind)
  // -- the order the CSR view needs --, dir true by (row_ind, col_ind).
  // Stable, so duplicate entries (equal index pairs and values) keep their
  // order.  A set-up step on the host copies, uploaded again.
  void sort(const bool dir) {
    std::vector<int32_t> perm((size_t)nnz);
    for (int k = 0; k < nnz; ++k) perm[(size_t)k] = k;
    const std::vector<int32_t>& a = dir ? h_col : h_col_ind;  // primary key
    const std::vector<int32_t>& b = dir ? h_col_ind : h_col;  // secondary key
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t p, int32_t q) {
      return a[(size_t)p] != a[(size_t)q] ? a[(size_t)p] < a[(size_t)q] : b[(size_t)p] < b[(size_t)q];
    });
    std::vector<int32_t> r((size_t)nnz), c((size_t)nnz);
    std::vector<V> v((size_t)nnz);
    for (int k = 0; k < nnz; ++k) {
      r[(size_t)k] = h_col[(size_t)perm[(size_t)k]];
      c[(size_t)k] = h_col_ind[(size_t)perm[(size_t)k]];
      v[(size_t)k] = h_values[(size_t)perm[(size_t)k]];
    }
    h_col.swap(r);
    h_col_ind.swap(c);
    h_values.swap(v);
    if (nnz > 0) {
      merbit::hip_check(hipMemcpy(row_ind, h_col.data(), sizeof(int32_t) * nnz,
                                  hipMemcpyHostToDevice), "hipMemcpy");
      merbit::hip_check(hipMemcpy(col_ind, h_col_ind.data(), sizeof(int32_t) * nnz,
                                  hipMemcpyHostToDevice), "hipMemcpy");
      merbit::hip_check(hipMemcpy(values, h_values.data(), sizeof(V) * nnz,
                                  hipMemcpyHostToDevice), "hipMemcpy");
    }
    csr_valid = !dir;
  }

 private:
  void load_mtx(const std::string& path) {
    int32_t n = 0;
    int64_t m = 0;
    merbit::check(mbh_mtx_read(path.c_str(), &n, &m, nullptr, nullptr, nullptr),
                  ("read " + path).c_str());
    std::vector<int32_t> ri((size_t)m), ci((size_t)m);
    merbit::check(mbh_mtx_read(path.c_str(), &n, &m, ri.data(), ci.data(), nullptr),
                  ("read " + path).c_str());
    nrow = n;
    nnz = (int)m;
    h_offset.assign((size_t)n + 1, 0);
    h_col.assign((size_t)m, 0);
    std::vector<double> v((size_t)m);
    merbit::check(mbh_ref_ingest(n, m, ri.data(), ci.data(), h_offset.data(), h_col.data(),
                                 v.data()),
                  "ingest");
    h_values.assign(v.begin(), v.end());
    upload();
  }

  void upload() {
    // the COO column of every entry: the CSR row it sits in
    h_col_ind.assign((size_t)nnz, 0);
    for (int r = 0; r < nrow; ++r)
      for (int k = h_offset[(size_t)r]; k < h_offset[(size_t)r + 1]; ++k) h_col_ind[(size_t)k] = r;
    offset = merbit::device_copy(h_offset);
    row_ind = merbit::device_copy(h_col);
    col_ind = merbit::device_copy(h_col_ind);
    values = merbit::device_copy(h_values);
    col = row_ind;
    csr_valid = true;
  }
};

template <typename V>
struct PARAMETER {
  GRAPH<V>* graph;
  double damp;
  int perrnd;
  int maxiter;
  int iter = 0;
  V* person = nullptr;         // x, filled with 1 (class.cuh:137-139)
  V* pagerank_real = nullptr;  // y
  PARAMETER(GRAPH<V>* g, double damp, int perrnd, int maxiter)
      : graph(g), damp(damp), perrnd(perrnd), maxiter(maxiter) {
    std::vector<V> ones((size_t)g->nrow, V(1));
    person = merbit::device_copy(ones);
    merbit::hip_check(hipMalloc((void**)&pagerank_real, sizeof(V) * (g->nrow > 0 ? g->nrow : 1)),
                      "hipMalloc");
  }
  ~PARAMETER() {
    merbit::device_free(person);
    merbit::device_free(pagerank_real);
  }
  PARAMETER(const PARAMETER&) = delete;
  PARAMETER& operator=(const PARAMETER&) = delete;
};

// include/tile.cuh:6-25: the plan.  SIGMA = items per thread, BLOCK = threads.
template <unsigned BLOCK, unsigned SIGMA>
struct TILE {
  mbx_plan* plan = nullptr;
  int tile_num = 0;
  int tile_items = 0;
  TILE(const int* offset, int nrow, int nnz, hipStream_t stream = nullptr) {
    merbit::check(mbx_plan_create(offset, nrow, nrow, nnz, (int32_t)SIGMA, (int32_t)BLOCK,
                                  stream, &plan),
                  "mbx_plan_create");
    int32_t nt = 0, ti = 0, wg = 0;
    merbit::check(mbx_plan_info(plan, &nt, &ti, &wg), "mbx_plan_info");
    tile_num = nt;
    tile_items = ti;
  }
  ~TILE() { mbx_plan_destroy(plan); }
  TILE(const TILE&) = delete;
  TILE& operator=(const TILE&) = delete;
};

// The launch rule's plan (mbx_plan_create_auto) in place of a compile-time
// SIGMA and BLOCK: the pattern's statistics choose the shape and options
// (no reference counterpart; the reference compiles one SIGMA for every
// matrix, CMakeLists.txt:4-6).
struct AUTO_TILE {
  mbx_plan* plan = nullptr;
  mbx_launch_choice choice{};
  AUTO_TILE(const int* offset, const int* col, int nrow, int nnz, int32_t value_kind,
            hipStream_t stream = nullptr) {
    merbit::check(mbx_plan_create_auto(offset, col, nrow, nrow, nnz, value_kind, stream, &plan,
                                       &choice),
                  "mbx_plan_create_auto");
  }
  ~AUTO_TILE() { mbx_plan_destroy(plan); }
  AUTO_TILE(const AUTO_TILE&) = delete;
  AUTO_TILE& operator=(const AUTO_TILE&) = delete;
};

namespace merbit {

template <typename V>
constexpr int32_t value_kind() {
  return std::is_same<V, double>::value ? MBX_VALUE_F64 : MBX_VALUE_F32;
}

inline int spmv(const mbx_plan* p, const GRAPH<double>& g, double alpha, const double* x,
                double beta, double* y, hipStream_t s) {
  return mbx_spmv_f64(p, g.offset, g.col, g.values, alpha, x, beta, y, s);
}

inline int spmv(const mbx_plan* p, const GRAPH<float>& g, float alpha, const float* x,
                float beta, float* y, hipStream_t s) {
  return mbx_spmv_f32(p, g.offset, g.col, g.values, alpha, x, beta, y, s);
}

}  // namespace merbit

struct MerbitTimes {
  float plan_ms = 0.f;  // tim1
  float spmv_ms = 0.f;  // tim2 / (maxiter * perrnd)
};

namespace merbit {
// The times of the calling thread's last Fun_Merbit / Fun_PageRank (the
// reference writes them to its log only).
inline MerbitTimes& last_times() {
  static thread_local MerbitTimes t;
  return t;
}
}  // namespace merbit

// include/tile.cuh:563-615.  Sorts the graph's entries by (col_ind, row_ind)
// and takes the CSR view as the reference does (tile.cuh:571-574), builds
// the plan (tim1), runs maxiter*perrnd SpMVs y = A*x on env.stream (tim2),
// leaves y in para.pagerank_real and appends " gid&MERBIT&SIGMA&tim1&
// tim2/iters" to FSTR when it is open; merbit::last_times() holds the times.
namespace merbit {
// The timed loop of products (include/tile.cuh:596-605) on a built plan,
// captured once as a HIP graph and replayed in one launch; y lands in
// para.pagerank_real and the time per product in t.spmv_ms.
template <typename V>
void timed_products(const ENV& env, const PARAMETER<V>& para, const mbx_plan* plan,
                    MerbitTimes& t) {
  const GRAPH<V>& g = *para.graph;
  Timer timer;
  const int iters = para.maxiter * para.perrnd;
  void* graph = nullptr;
  if (iters > 0)
    merbit::check(mbx_spmv_graph_create(plan, merbit::value_kind<V>(), g.offset, g.col, g.values,
                                        1.0, para.person, 0.0, para.pagerank_real, iters, &graph),
                  "graph capture");
  timer.start(env.stream);
  if (graph) merbit::check(mbx_graph_launch(graph, env.stream), "graph launch");
  const float tim2 = timer.stop(env.stream);
  mbx_graph_destroy(graph);
  t.spmv_ms = iters > 0 ? tim2 / (float)iters : 0.f;
}
}  // namespace merbit

template <typename V, unsigned BLOCK, unsigned SIGMA>
void Fun_Merbit(const ENV& env, const PARAMETER<V>& para) {
  static_assert(std::is_same<V, double>::value || std::is_same<V, float>::value,
                "Fun_Merbit is instantiated for float and double");
  Timer timer;
  MerbitTimes& t = merbit::last_times();
  t = MerbitTimes();
  if (!para.graph->csr_valid) para.graph->sort(false);
  const GRAPH<V>& g = *para.graph;
  timer.start(env.stream);
  TILE<BLOCK, SIGMA> tile(g.offset, g.nrow, g.nnz, env.stream);
  t.plan_ms = timer.stop(env.stream);
  merbit::timed_products(env, para, tile.plan, t);
  if (FSTR.is_open())
    FSTR << " " << g.graph_id << "&MERBIT&" << SIGMA << "&" << t.plan_ms << "&" << t.spmv_ms
         << "\n";
}

// Fun_Merbit with the launch rule's plan (AUTO_TILE) instead of a fixed
// SIGMA: tim1 includes the rule's statistics; the log's sigma field reads
// "auto:<items per thread>".  Returns the rule's choice.
template <typename V>
mbx_launch_choice Fun_Merbit_Auto(const ENV& env, const PARAMETER<V>& para) {
  static_assert(std::is_same<V, double>::value || std::is_same<V, float>::value,
                "Fun_Merbit_Auto is instantiated for float and double");
  Timer timer;
  MerbitTimes& t = merbit::last_times();
  t = MerbitTimes();
  if (!para.graph->csr_valid) para.graph->sort(false);
  const GRAPH<V>& g = *para.graph;
  timer.start(env.stream);
  AUTO_TILE tile(g.offset, g.col, g.nrow, g.nnz, merbit::value_kind<V>(), env.stream);
  t.plan_ms = timer.stop(env.stream);
  merbit::timed_products(env, para, tile.plan, t);
  if (FSTR.is_open())
    FSTR << " " << g.graph_id << "&MERBIT&auto:" << tile.choice.items_per_thread << "&"
         << t.plan_ms << "&" << t.spmv_ms << "\n";
  return tile.choice;
}

// PageRank with the damping factor the reference carries in PARAMETER::damp
// but does not apply (include/class.cuh:119-146): para.maxiter power
// iterations y = damp*A*x, x' = y + (sum(x) - sum(y))/n from x0 = para.person,
// captured as one HIP graph; the last iterate lands in para.pagerank_real.
// FP64 only.  Returns the plan time and the time per iteration.
template <unsigned BLOCK, unsigned SIGMA>
MerbitTimes Fun_PageRank(const ENV& env, const PARAMETER<double>& para) {
  Timer timer;
  MerbitTimes t;
  if (!para.graph->csr_valid) para.graph->sort(false);
  const GRAPH<double>& g = *para.graph;
  timer.start(env.stream);
  TILE<BLOCK, SIGMA> tile(g.offset, g.nrow, g.nnz, env.stream);
  t.plan_ms = timer.stop(env.stream);
  if (para.maxiter < 1 || g.nrow == 0) return t;
  const size_t wbytes = mbx_pagerank_workspace_bytes(g.nrow);
  void* work = nullptr;
  merbit::hip_check(hipMalloc(&work, wbytes), "hipMalloc");
  void* graph = nullptr;
  const int st = mbx_pagerank_graph_create(tile.plan, g.offset, g.col, g.values, para.damp,
                                           para.maxiter, para.person, para.pagerank_real, work,
                                           wbytes, &graph);
  if (st != MBX_OK) {
    merbit::device_free(work);
    merbit::check(st, "PageRank graph capture");
  }
  timer.start(env.stream);
  const int lst = mbx_graph_launch(graph, env.stream);
  const float ms = timer.stop(env.stream);
  mbx_graph_destroy(graph);
  merbit::device_free(work);
  merbit::check(lst, "graph launch");
  t.spmv_ms = ms / (float)para.maxiter;
  if (FSTR.is_open())
    FSTR << " " << g.graph_id << "&PAGERANK&" << SIGMA << "&" << t.plan_ms << "&" << t.spmv_ms
         << "\n";
  return t;
}

// include/utils.cuh:116-132: flags rows with |X-Y| > rtol*|X|, prints the
// first few and the total flagged error; returns the number of flagged rows.
template <typename V>
int Fun_Check(const V* X, const V* Y, int nrow, double rtol = 1e-4, std::ostream& os = std::cout) {
  double err = 0.0;
  int bad = 0;
  for (int i = 0; i < nrow; ++i) {
    const double d = std::fabs((double)X[i] - (double)Y[i]);
    if (d > std::fabs((double)X[i]) * rtol || std::isnan(d)) {
      err += d;
      if (bad++ < 10) os << "X[" << i << "] = " << X[i] << ", Y[" << i << "] = " << Y[i] << "\n";
    }
  }
  os << ", total error = " << err << std::endl;
  return bad;
}

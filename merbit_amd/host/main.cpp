// main.cpp -- config-ID driver: ./merbit_main <ID> (src/main.cu:3-84).
//
//   merbit_main <ID> [--config FILE] [--log FILE] [--device N] [--type f64|f32]
//               [--sigma 2|4|6|8|10|12|14|16|auto] [--perrnd N] [--maxiter N] [--check]
//               [--pagerank [--damp D]] [--sort-rows]
//   merbit_main --mtx FILE [...]
//
// Loads graph <ID> from the registry (lines "ID path [nrow ncol nnz]"), builds
// the operated matrix with the reference's semantics, runs Fun_Merbit
// (maxiter*perrnd SpMVs, default 20*20 = 400 as src/main.cu:32) and appends
// " gid&MERBIT&sigma&tim1&tim2" to the log.  --check verifies y against a
// sequential FP64 CSR product on the host (Fun_Check).  --pagerank runs
// Fun_PageRank instead: maxiter damped power iterations (FP64), checked
// against the same iterations on the host.  --sigma auto builds the plan with
// the launch rule (Fun_Merbit_Auto: shape and options from the pattern's
// statistics) instead of a fixed SIGMA.

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "merbit.hpp"

namespace {

struct Args {
  int gid = -1;
  std::string config = "config/config.txt";
  std::string log;
  std::string mtx;
  std::string type = "f64";
  int device = 0, sigma = 8, perrnd = 20, maxiter = 20;
  double damp = 0.85;
  bool check = false, pagerank = false, sort_rows = false;
};

int usage() {
  std::cerr << "usage: merbit_main <ID> [--config FILE] [--log FILE] [--device N] "
               "[--type f64|f32] [--sigma 2|4|...|16|auto] [--perrnd N] [--maxiter N] "
               "[--check]\n"
               "                   [--pagerank [--damp D]]\n"
               "       merbit_main --mtx FILE [same options]\n";
  return 2;
}

// PageRank iterations on the host, in the order the device computes them
// mathematically (y = damp*A*x, then the uniform correction); the check's
// tolerance absorbs the different summation order.
std::vector<double> host_pagerank(const GRAPH<double>& g, const std::vector<double>& x0,
                                  double damp, int iters) {
  std::vector<double> x = x0, y((size_t)g.nrow);
  for (int t = 0; t < iters; ++t) {
    double sx = 0.0, sy = 0.0;
    for (int r = 0; r < g.nrow; ++r) {
      double s = 0.0;
      for (int k = g.h_offset[r]; k < g.h_offset[r + 1]; ++k)
        s += g.h_values[k] * x[(size_t)g.h_col[k]];
      y[(size_t)r] = damp * s;
      sx += x[(size_t)r];
      sy += y[(size_t)r];
    }
    const double c = (sx - sy) / (double)g.nrow;
    for (int r = 0; r < g.nrow; ++r) x[(size_t)r] = y[(size_t)r] + c;
  }
  return x;
}

template <unsigned S>
int run_pagerank(const ENV& env, GRAPH<double>& graph, const Args& a) {
  PARAMETER<double> para(&graph, a.damp, a.perrnd, a.maxiter);
  const MerbitTimes t = Fun_PageRank<256, S>(env, para);
  std::cout << "graph = " << graph.graph_id << ", PageRank damp = " << a.damp
            << ", iterations = " << a.maxiter << ", nrow = " << graph.nrow
            << ", nnz = " << graph.nnz << "\n"
            << "tim1 = " << t.plan_ms << " ms, " << t.spmv_ms << " ms/iteration\n";
  if (!a.check) return 0;
  std::vector<double> y((size_t)graph.nrow);
  merbit::hip_check(hipMemcpy(y.data(), para.pagerank_real, sizeof(double) * y.size(),
                              hipMemcpyDeviceToHost),
                    "hipMemcpy");
  const std::vector<double> ref =
      host_pagerank(graph, std::vector<double>((size_t)graph.nrow, 1.0), a.damp, a.maxiter);
  const int bad = Fun_Check<double>(ref.data(), y.data(), graph.nrow, 1e-12);
  std::cout << (bad ? "CHECK FAILED" : "check ok") << " (" << bad << " rows)\n";
  return bad ? 1 : 0;
}

// The report line and --check of one Fun_Merbit / Fun_Merbit_Auto run.
template <typename V>
int report(const GRAPH<V>& graph, const PARAMETER<V>& para, const Args& a,
           const std::string& sigma) {
  const MerbitTimes t = merbit::last_times();
  const double bytes = 4.0 * (graph.nrow + 1) + graph.nnz * (4.0 + sizeof(V)) +
                       2.0 * graph.nrow * sizeof(V);
  std::cout << "graph = " << graph.graph_id << ", block = 256, sigma = " << sigma
            << ", nrow = " << graph.nrow << ", nnz = " << graph.nnz << "\n"
            << "tim1 = " << t.plan_ms << " ms, tim2 = " << t.spmv_ms << " ms/SpMV, "
            << bytes / (t.spmv_ms * 1e-3) / 1e9 << " GB/s\n";
  if (!a.check) return 0;
  std::vector<V> y((size_t)graph.nrow), ref((size_t)graph.nrow);
  merbit::hip_check(hipMemcpy(y.data(), para.pagerank_real, sizeof(V) * y.size(),
                              hipMemcpyDeviceToHost),
                    "hipMemcpy");
  for (int r = 0; r < graph.nrow; ++r) {  // x = 1
    double s = 0.0;
    for (int k = graph.h_offset[r]; k < graph.h_offset[r + 1]; ++k) s += (double)graph.h_values[k];
    ref[(size_t)r] = (V)s;
  }
  const double rtol = std::is_same<V, double>::value ? 1e-12 : 1e-5;
  const int bad = Fun_Check<V>(ref.data(), y.data(), graph.nrow, rtol);
  std::cout << (bad ? "CHECK FAILED" : "check ok") << " (" << bad << " rows)\n";
  return bad ? 1 : 0;
}

template <typename V, unsigned S>
int run_sigma(const ENV& env, GRAPH<V>& graph, const Args& a) {
  if constexpr (std::is_same<V, double>::value) {
    if (a.pagerank) return run_pagerank<S>(env, graph, a);
  } else if (a.pagerank) {
    std::cerr << "--pagerank is FP64 only\n";
    return 2;
  }
  PARAMETER<V> para(&graph, a.damp, a.perrnd, a.maxiter);
  // --sort-rows leaves the entries in (row, col) order, as after the
  // reference's GRAPH::sort(true); Fun_Merbit sorts them back to (col, row)
  // before it takes the CSR view, as the reference's does (tile.cuh:571)
  if (a.sort_rows) graph.sort(true);
  Fun_Merbit<V, 256, S>(env, para);
  return report(graph, para, a, std::to_string(S));
}

// --sigma auto: the launch rule's plan.
template <typename V>
int run_auto(const ENV& env, GRAPH<V>& graph, const Args& a) {
  if (a.pagerank) {
    std::cerr << "--pagerank takes a fixed --sigma\n";
    return 2;
  }
  PARAMETER<V> para(&graph, a.damp, a.perrnd, a.maxiter);
  if (a.sort_rows) graph.sort(true);
  const mbx_launch_choice c = Fun_Merbit_Auto<V>(env, para);
  std::cout << "launch rule " << c.rule << ": items per thread " << c.items_per_thread
            << ", row tiles " << c.row_tiles << ", hot columns " << c.hot_columns
            << ", 16-bit columns " << c.compress_cols << "\n";
  return report(graph, para, a, "auto:" + std::to_string(c.items_per_thread));
}

template <typename V>
int run(const Args& a) {
  ENV env(a.device, a.config, a.log);
  GRAPH<V> graph = a.mtx.empty() ? GRAPH<V>(a.config, a.gid) : GRAPH<V>::from_mtx(a.mtx, a.gid);
  switch (a.sigma) {
    case -1: return run_auto<V>(env, graph, a);
    case 2: return run_sigma<V, 2>(env, graph, a);
    case 4: return run_sigma<V, 4>(env, graph, a);
    case 6: return run_sigma<V, 6>(env, graph, a);
    case 8: return run_sigma<V, 8>(env, graph, a);
    case 10: return run_sigma<V, 10>(env, graph, a);
    case 12: return run_sigma<V, 12>(env, graph, a);
    case 14: return run_sigma<V, 14>(env, graph, a);
    case 16: return run_sigma<V, 16>(env, graph, a);
    default: std::cerr << "unsupported sigma " << a.sigma << "\n"; return 2;
  }
}

}  // namespace

int main(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string s = argv[i];
    auto next = [&](void) -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    if (s == "--config" && (v = next())) a.config = v;
    else if (s == "--log" && (v = next())) a.log = v;
    else if (s == "--mtx" && (v = next())) a.mtx = v;
    else if (s == "--type" && (v = next())) a.type = v;
    else if (s == "--device" && (v = next())) a.device = std::atoi(v);
    else if (s == "--sigma" && (v = next())) a.sigma = std::strcmp(v, "auto") ? std::atoi(v) : -1;
    else if (s == "--perrnd" && (v = next())) a.perrnd = std::atoi(v);
    else if (s == "--maxiter" && (v = next())) a.maxiter = std::atoi(v);
    else if (s == "--check") a.check = true;
    else if (s == "--pagerank") a.pagerank = true;
    else if (s == "--sort-rows") a.sort_rows = true;
    else if (s == "--damp" && (v = next())) a.damp = std::atof(v);
    else if (!s.empty() && s[0] != '-' && a.gid < 0) a.gid = std::atoi(s.c_str());
    else return usage();
  }
  if (a.gid < 0 && a.mtx.empty()) return usage();
  if (!a.log.empty()) FSTR.open(a.log, std::ios_base::app);
  int rc = 1;
  try {  // TRYCATCH, include/inc.cuh:124-132
    rc = a.type == "f32" ? run<float>(a) : run<double>(a);
  } catch (const std::exception& e) {
    std::cerr << "Exception caught: " << e.what() << std::endl;
    rc = 1;
  }
  if (FSTR.is_open()) FSTR.close();
  return rc;
}

// f2m_solver -- command-line F2M solver over the C ABI (counterpart of
// RUN.cpp:12-31, `F2M.solver.exe <tsplib-file> <topK>`).
//
//   f2m_solver <tsplib-file> <topK> [--early-exit] [--quiet] [--input-order]
//                                  [--reference-recovery] [--adaptive]
//
// Nodes are numbered along the Morton curve of the points (f2m_spatial_order;
// --input-order keeps the input order), which only changes speed and, among
// equal-cost candidates at the K-th rank, which one the kNN rule keeps.  The
// exact recovery runs where the reference's recovery is not exact
// (--reference-recovery reproduces the reference's output there instead).
// --adaptive: topK is the first K; while some node's K-th nearest cost is below
// its dual bound (f2m_knn_certify), K doubles (at most 80) and the solve restarts
// from the previous potentials; the report gains a line "kNN K: <K> (certified)".
//
// Writes next to the input, as the reference does (README.md:32-45):
//   <file>.top<K>.F2M_solution  "label1 label2 x" per edge with x != 0, label1 <
//                               label2, x printed with 4 significant digits
//                               (graph_tour.hpp:186-214)
//   <file>.top<K>.F2M_report    Optimal Found / Optimal Value (15 digits) / Time /
//                               MAX edge rank, or "ERROR: Optimal Solution not
//                               Found" (graph_tour.hpp:218-241)
//   <file>.top<K>.DM            kept edges "v1 v2 d", 0-based point indices
//                               (FXY.hpp:146-165)
//   <file>_duplicates           if the input repeats a point (FXY.hpp:168-183)
// Time is the solve only (graph::RUN scope, graph_cuda.cu:51,131).
// Exit codes follow the reference: 0 after usage and format messages,
// 1 on a device error.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <tuple>
#include <vector>

#include "f2m.h"

static int die(f2m_ctx* ctx, f2m_status s, const char* what) {
  std::cerr << "f2m_solver: " << what << ": " << f2m_strerror(s);
  if (ctx && *f2m_last_error(ctx)) std::cerr << " (" << f2m_last_error(ctx) << ")";
  std::cerr << "\n";
  return 1;
}

int main(int argc, char** argv) {
  bool early = false, quiet = false, input_order = false, ref_recovery = false, adaptive = false;
  std::vector<std::string> pos;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--early-exit")) early = true;
    else if (!std::strcmp(argv[i], "--quiet")) quiet = true;
    else if (!std::strcmp(argv[i], "--input-order")) input_order = true;
    else if (!std::strcmp(argv[i], "--reference-recovery")) ref_recovery = true;
    else if (!std::strcmp(argv[i], "--adaptive")) adaptive = true;
    else pos.push_back(argv[i]);
  }
  if (pos.size() != 2) {
    std::cout << "usage: f2m_solver fileXY topK [--early-exit] [--quiet] [--input-order] [--reference-recovery]"
                 " [--adaptive]\n"
                 "fileXY == file in TSPLIB format\n"
                 "topK == the number of top K closest neighbors for each node\n";
    return 0;
  }
  const std::string file = pos[0];
  const int K = std::atoi(pos[1].c_str());
  if (K < 1) {
    std::cout << "topK must be a positive integer\n";
    return 0;
  }
  f2m_points* pts = nullptr;
  f2m_status s = f2m_tsplib_read(file.c_str(), &pts);
  if (s == F2M_ERR_FORMAT || s == F2M_ERR_IO) {
    std::cout << "ERROR MESSAGE:\nYour file : " << file << " seems to be in the wrong format\n"
              << "You must provide file in TSPLIB format\n";
    return 0;
  }
  if (s) return die(nullptr, s, "reading input");
  const int64_t n = f2m_points_count(pts);
  std::vector<int64_t> label(n);
  std::vector<double> x(n), y(n);
  f2m_points_copy(pts, label.data(), x.data(), y.data());
  if (f2m_points_duplicates(pts) > 0) f2m_points_write_duplicates(pts, (file + "_duplicates").c_str());
  f2m_points_free(pts);
  if (n < 4) {
    std::cout << "need at least 4 distinct points\n";
    return 0;
  }

  f2m_ctx* ctx = nullptr;
  if ((s = f2m_ctx_create(0, nullptr, &ctx))) return die(nullptr, s, "device");
  // node v of the graph is input point perm[v]
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
  if (!input_order) {
    if ((s = f2m_spatial_order(ctx, (int32_t)n, x.data(), y.data(), perm.data()))) return die(ctx, s, "order");
    std::vector<int64_t> l2(n);
    std::vector<double> x2(n), y2(n);
    for (int64_t i = 0; i < n; ++i) {
      l2[i] = label[perm[i]];
      x2[i] = x[perm[i]];
      y2[i] = y[perm[i]];
    }
    label.swap(l2);
    x.swap(x2);
    y.swap(y2);
  }
  f2m_opts o;
  f2m_default_opts(&o);
  o.early_exit = early ? 1 : 0;
  if (ref_recovery) o.exact = 0;
  // solve with K; with --adaptive, while the kNN restriction is not certified
  // (f2m_knn_certify > 0) double K (at most 80) and re-solve from the previous
  // potentials (f2m_solve_warm)
  f2m_graph g;
  f2m_stats st;
  int Kc = K;
  int64_t viol = -1;
  std::vector<double> Fw;
  double secs = 0;
  for (;;) {
    if ((s = f2m_knn_build(ctx, (int32_t)n, x.data(), y.data(), Kc, F2M_COST_RAW_FP64, &g)))
      return die(ctx, s, "kNN graph");
    auto t0 = std::chrono::steady_clock::now();
    s = Fw.empty() ? f2m_solve(ctx, &g, &o, &st) : f2m_solve_warm(ctx, &g, &o, Fw.data(), &st);
    if (s) return die(ctx, s, "solve");
    secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!adaptive) break;
    if ((s = f2m_knn_certify(ctx, &g, &viol))) return die(ctx, s, "certificate");
    if (!quiet) std::cout << "K " << Kc << ": objective " << std::setprecision(15) << st.objective
                          << ", nodes failing the kNN certificate " << viol << "\n";
    const int next = std::min<int64_t>(std::min(2 * Kc, 80), n - 1);
    if (viol == 0 || next <= Kc) break;
    Fw.resize(n);
    if ((s = f2m_result_copy(ctx, Fw.data(), nullptr, nullptr))) return die(ctx, s, "results");
    Kc = next;
  }
  std::vector<int32_t> rowptr(n + 1), nbr(g.nslots);
  std::vector<double> cost(g.nslots);
  if ((s = f2m_graph_copy(ctx, &g, rowptr.data(), nbr.data(), cost.data()))) return die(ctx, s, "graph copy");

  const std::string stem = file + ".top" + std::to_string(K);
  {  // .DM: kept edges in increasing (cost, v1, v2), input point indices
    std::vector<std::tuple<double, int32_t, int32_t>> e;
    for (int64_t v = 0; v < n; ++v)
      for (int32_t q = rowptr[v]; q < rowptr[v + 1]; ++q)
        if (nbr[q] > v) {
          const int32_t a = perm[v], b = perm[nbr[q]];
          e.emplace_back(cost[q], std::min(a, b), std::max(a, b));
        }
    std::sort(e.begin(), e.end());
    std::ofstream f(stem + ".DM");
    for (auto& t : e) f << std::get<1>(t) << " " << std::get<2>(t) << " " << std::get<0>(t) << "\n";
  }
  std::vector<double> xs(g.nslots);
  if ((s = f2m_result_copy(ctx, nullptr, xs.data(), nullptr))) return die(ctx, s, "results");
  f2m_ctx_destroy(ctx);

  std::ofstream rep(stem + ".F2M_report");
  if (st.status == 0) {
    std::ofstream sol(stem + ".F2M_solution");
    std::vector<std::tuple<int64_t, int64_t, double>> out;
    for (int64_t v = 0; v < n; ++v)
      for (int32_t q = rowptr[v]; q < rowptr[v + 1]; ++q) {
        const int64_t w = nbr[q];
        if (w <= v || xs[q] == 0.0) continue;
        int64_t a = label[v], b = label[w];
        if (a > b) std::swap(a, b);
        out.emplace_back(a, b, xs[q]);
      }
    std::sort(out.begin(), out.end());
    for (auto& t : out) sol << std::get<0>(t) << " " << std::get<1>(t) << " " << std::setprecision(4) << std::get<2>(t) << "\n";
    rep << "Optimal Found: Yes\n";
    rep << "Optimal Value: " << std::setprecision(15) << st.objective << "\n";
    rep << "Time: " << secs << "\n";
    rep << "MAX edge rank: " << st.max_rank << "\n";
    if (adaptive) rep << "kNN K: " << Kc << (viol == 0 ? " (certified)" : " (not certified)") << "\n";
    if (!quiet) {
      std::cout << "Optimal Found: Yes\n";
      std::cout << "Optimal Value: " << std::setprecision(15) << st.objective << "\n";
      std::cout << "RunTime: " << secs << "\n";
      std::cout << "MAX edge rank: " << st.max_rank << "\n";
      std::cout << "Solution File: " << stem << ".F2M_solution\n";
    }
  } else {
    rep << "ERROR: Optimal Solution not Found\n";
    if (!quiet) std::cout << "ERROR: Optimal Solution not Found\n";
  }
  if (!quiet)
    std::cout << "Nodes: " << n << " Edges: " << g.nslots / 2 << " sweeps: " << st.sweeps
              << " combs: " << st.n_combs << " irregular: " << st.n_irregular_a << "/" << st.n_irregular_b
              << " dual bound: " << std::setprecision(15) << st.dual << "\n";
  return 0;
}

// Host-code driver for the sanitizer build (tests/test_sanitize.py): the
// TSPLIB reader (f2m_amd/csrc/f2m_tsplib.cpp) and the CPU oracle
// (oracle/f2m_oracle.cpp, oracle/f2m_exact_oracle.cpp), built together with
// -fsanitize=address,undefined and run on the files named on the command
// line: read, write duplicates, kNN graph, solve (exact and reference
// recovery, 1 and 4 threads), warm start, exact stage, certificate.  Prints
// one line per file; any sanitizer report makes the process fail.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "f2m.h"
#include "f2m_oracle.h"

static int run(const char* path, const char* dup_out) {
  f2m_points* p = nullptr;
  const f2m_status s = f2m_tsplib_read(path, &p);
  if (s) {
    std::printf("%s: status %d\n", path, (int)s);
    return s == F2M_ERR_FORMAT || s == F2M_ERR_IO ? 0 : 1;
  }
  const int64_t n = f2m_points_count(p);
  std::vector<int64_t> lab(n);
  std::vector<double> x(n), y(n);
  f2m_points_copy(p, lab.data(), x.data(), y.data());
  if (f2m_points_duplicates(p) > 0 && f2m_points_write_duplicates(p, dup_out) != F2M_OK) return 1;
  f2m_points_free(p);
  if (n < 12) {
    std::printf("%s: n %lld (too small to solve)\n", path, (long long)n);
    return 0;
  }
  const int K = 10;
  void* g = orc_graph_build((int)n, x.data(), y.data(), K, 0, 0, 4);
  if (!g) return 1;
  const int64_t m = orc_graph_nslots(g);
  std::vector<int32_t> rp(n + 1), nb(m), root(n), hr(n);
  std::vector<double> c(m), F(n), xs(m), F2(n), x2(m);
  orc_graph_export(g, rp.data(), nb.data(), c.data());
  orc_graph_free(g);
  orc_opts o;
  orc_default_opts(&o);
  orc_stats st, st0, stw, stx;
  if (orc_solve((int)n, rp.data(), nb.data(), c.data(), &o, 4, F.data(), xs.data(), root.data(), &st)) return 1;
  orc_opts o0 = o;
  o0.exact = 0;
  if (orc_solve((int)n, rp.data(), nb.data(), c.data(), &o0, 1, F2.data(), x2.data(), root.data(), &st0)) return 1;
  if (orc_solve_warm((int)n, rp.data(), nb.data(), c.data(), &o, 2, F.data(), F2.data(), x2.data(), root.data(),
                     &stw))
    return 1;
  const int rc = orc_exact((int)n, rp.data(), nb.data(), c.data(), st.d_min, F.data(), &o, F2.data(), x2.data(),
                           hr.data(), &stx);
  const int64_t viol = orc_knn_certify((int)n, rp.data(), c.data(), K, st.d_min, F.data());
  std::printf("%s: n %lld slots %lld objective %.10g ref-recovery %.10g warm %.10g exact rc %d %.10g cert %lld\n",
              path, (long long)n, (long long)m, st.objective, st0.objective, stw.objective, rc, stx.objective,
              (long long)viol);
  return std::fabs(st.objective - stw.objective) <= 1e-9 * std::fabs(st.objective) ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: san_driver <dup-out> <tsplib>...\n");
    return 2;
  }
  int bad = 0;
  for (int i = 2; i < argc; ++i) bad |= run(argv[i], argv[1]);
  return bad;
}

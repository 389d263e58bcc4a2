// ref_axb.cpp -- C entry point over the reference's own least-squares solver
// (Ax_b.cpp / Ax_b.h, compiled from the reference tree by oracle/ref.mk into
// oracle/_ref/).  TEST INFRASTRUCTURE ONLY: used by tests/test_oracle_ref.py
// to check orc_axb (oracle/f2m_oracle.cpp) against the reference bit for bit.
#include <string>
#include <vector>

#include "Ax_b.h"

extern "C" int ref_axb(int m, const int* row_ptr, const int* col, const int* b, int nv, double* x_out,
                       double* max_err) {
  // Ax_b::parse_Ax_b input: per equation, (var, coef) pairs then (-1, rhs)
  std::vector<int> A;
  for (int j = 0; j < m; ++j) {
    for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k) {
      A.push_back(col[k]);
      A.push_back(1);
    }
    A.push_back(-1);
    A.push_back(b[j]);
  }
  Ax_b S(A);
  S.solve();
  for (int i = 0; i < nv && i < (int)S.x.size(); ++i) x_out[i] = S.x[i];
  if (max_err) *max_err = S.d_max_error;
  return S.go_max;
}

// linalg.cpp -- small dense symmetric eigensolver for the PCA initialisation.
//
// Cyclic Jacobi rotations (Rutishauser's formulation, Golub & Van Loan
// section 8.5): each rotation zeroes one off-diagonal entry; sweeps repeat
// until the off-diagonal mass is negligible relative to the diagonal.  Used
// on matrices of at most a few hundred rows (the covariance when d is small,
// the Rayleigh-Ritz projection of a subspace iteration otherwise).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "host_common.h"

namespace pm {

void jacobi_eigh(int m, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V) {
  V.assign(static_cast<size_t>(m) * m, 0.0);
  for (int i = 0; i < m; ++i) V[static_cast<size_t>(i) * m + i] = 1.0;
  auto a = [&](int i, int j) -> double & { return A[static_cast<size_t>(i) * m + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < m; ++i) {
      diag += a(i, i) * a(i, i);
      for (int j = i + 1; j < m; ++j) off += a(i, j) * a(i, j);
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p) {
      for (int q = p + 1; q < m; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        const double app = a(p, p), aqq = a(q, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        // A <- J^T A J on rows/columns p and q
        for (int k = 0; k < m; ++k) {
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        a(p, q) = 0.0;
        a(q, p) = 0.0;
        for (int k = 0; k < m; ++k) {
          double &vkp = V[static_cast<size_t>(k) * m + p], &vkq = V[static_cast<size_t>(k) * m + q];
          const double x = vkp, y = vkq;
          vkp = c * x - s * y;
          vkq = s * x + c * y;
        }
      }
    }
  }
  // sort descending by eigenvalue; sign: largest-magnitude entry positive
  std::vector<int> order(m);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return a(x, x) > a(y, y); });
  std::vector<double> Vs(static_cast<size_t>(m) * m);
  w.assign(m, 0.0);
  for (int c2 = 0; c2 < m; ++c2) {
    const int src = order[c2];
    w[c2] = a(src, src);
    int best = 0;
    double bv = -1.0;
    for (int k = 0; k < m; ++k) {
      const double v = std::fabs(V[static_cast<size_t>(k) * m + src]);
      if (v > bv) {
        bv = v;
        best = k;
      }
    }
    const double sg = V[static_cast<size_t>(best) * m + src] < 0 ? -1.0 : 1.0;
    for (int k = 0; k < m; ++k) Vs[static_cast<size_t>(k) * m + c2] = sg * V[static_cast<size_t>(k) * m + src];
  }
  V.swap(Vs);
}

}  // namespace pm

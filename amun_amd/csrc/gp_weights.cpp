// gp_weights.cpp -- host-side weights of the multi-dimensional Gaussian
// process reconstruction in quad precision (built with g++ and libquadmath:
// the covariance matrix of the ngp^nd stencil is ill-conditioned for kernel
// widths of several cells, so the inverse needs the reference's widest real
// kind to give the weights to double precision).
#include <quadmath.h>
#include <cmath>
#include <vector>

// prepare_mgp (interpolations.F90:558-711) with invert (algebra.F90:
// 1198-1305, Gauss-Jordan with the reference's zero-pivot row addition):
// out[(d * 2 + s) * n + m] = (xgp(:, s, d) A^-1)(m), A the covariance of the
// cell averages over the stencil (i fastest), xgp that of the face value at
// the lower (s = 0, offset - 1/2 from the stencil cell) or upper side along
// d; no symmetrisation.  out holds 2 nd ngp^nd values.  Returns false when A
// is singular.
extern "C" bool amun_mgp_weights_quad(int nd, int ngp, double sgp, double* out) {
  typedef __float128 Q;
  const int mg = (ngp - 1) / 2, n = nd == 3 ? ngp * ngp * ngp : ngp * ngp;
  const Q sig = (Q)(std::sqrt(2.0) * sgp);
  const Q fc = (Q)(0.5 * std::sqrt(3.14159265358979323846264338327950)) * sig;
  auto erfd = [&](int t) {
    return erfq(((Q)(1.0 * t) + 0.5Q) / sig) - erfq(((Q)(1.0 * t) - 0.5Q) / sig);
  };
  std::vector<int> o[3] = {std::vector<int>(n), std::vector<int>(n), std::vector<int>(n)};
  for (int m = 0, c = (nd == 3 ? -mg : 0); c <= (nd == 3 ? mg : 0); c++)
    for (int b = -mg; b <= mg; b++)
      for (int a = -mg; a <= mg; a++, m++) { o[0][m] = a; o[1][m] = b; o[2][m] = c; }
  std::vector<Q> g((size_t)n * 2 * n);
  auto G = [&](int i, int j) -> Q& { return g[(size_t)i * 2 * n + j]; };
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) {
      if (j < n) {
        Q v = fc * erfd(o[0][i] - o[0][j]) * erfd(o[1][i] - o[1][j]);
        if (nd == 3) v = v * erfd(o[2][i] - o[2][j]);
        G(i, j) = v;
      } else {
        G(i, j) = (i + n == j) ? 1.0Q : 0.0Q;
      }
    }
  for (int k = 0; k < n - 1; k++) {
    if (G(k, k) == 0.0Q) {
      bool found = false;
      for (int i = k + 1; i < n && !found; i++)
        if (G(i, k) != 0.0Q) {
          for (int j = 0; j < 2 * n; j++) G(k, j) = G(k, j) + G(i, j);
          found = true;
        }
      if (!found) return false;
    }
    for (int j = k + 1; j < n; j++) {
      const Q t = G(j, k) / G(k, k);
      for (int i = k; i < 2 * n; i++) G(j, i) = G(j, i) - t * G(k, i);
    }
  }
  for (int i = 0; i < n; i++)
    if (G(i, i) == 0.0Q) return false;
  for (int i = 0; i < n; i++) {
    const Q t = G(i, i);
    for (int j = i; j < 2 * n; j++) G(i, j) = G(i, j) / t;
  }
  for (int k = n - 2; k >= 0; k--)
    for (int i = 0; i <= k; i++) {
      const Q t = G(i, k + 1);
      for (int j = k; j < 2 * n; j++) G(i, j) = G(i, j) - G(k + 1, j) * t;
    }
  std::vector<Q> x(n);
  for (int d = 0; d < nd; d++)
    for (int s = 0; s < 2; s++) {
      for (int m = 0; m < n; m++) {
        const Q z = ((Q)(1.0 * o[d][m]) + (s ? 0.5Q : -0.5Q)) / sig;
        Q e = expq(-(z * z));
        for (int t = 0; t < nd; t++)
          if (t != d) e = e * erfd(o[t][m]);
        x[m] = e;
      }
      for (int m = 0; m < n; m++) {
        Q acc = 0.0Q;
        for (int i = 0; i < n; i++) acc = acc + x[i] * G(i, m + n);
        out[(size_t)(d * 2 + s) * n + m] = (double)acc;
      }
    }
  return true;
}

// Dense Hermitian eigensolver, FP64, written from the textbook method:
// complex Householder reduction to Hermitian tridiagonal form, a diagonal
// unitary rescaling that makes the sub-diagonal real, implicit-shift QL
// iteration on the real symmetric tridiagonal matrix, back-transformation.
// Semantics of src/eigensolver/dense.rs:50-98 (faer self_adjoint_eigen,
// ascending eigenvalues, unit eigenvectors as columns). Eigenvectors are
// unique only up to a phase (and rotation within degenerate subspaces).
// Test infrastructure (oracle).
#include <algorithm>
#include <numeric>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// Implicit QL on a real symmetric tridiagonal matrix: d (diagonal), e
// (sub-diagonal, e[i] couples i and i+1, e[n-1] unused); z is the n×n
// accumulated eigenvector matrix stored column-major (z[col*n + row]).
static void tridiag_ql(std::vector<double>& d, std::vector<double>& e, std::vector<double>& z, size_t n) {
  if (n == 0) return;
  e[n - 1] = 0.0;
  for (size_t l = 0; l < n; ++l) {
    int iter = 0;
    size_t m;
    do {
      for (m = l; m + 1 < n; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 || std::fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m != l) {
        if (++iter > 300) throw std::runtime_error("tridiagonal QL did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        size_t i;
        bool early = false;
        for (i = m; i-- > l;) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            early = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          double* zi = &z[i * n];
          double* zi1 = &z[(i + 1) * n];
          for (size_t k = 0; k < n; ++k) {
            f = zi1[k];
            zi1[k] = s * zi[k] + c * f;
            zi[k] = c * zi[k] - s * f;
          }
        }
        if (early) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

void eigh(size_t n, const cplx* a_in, double* w, cplx* v_out) { eigh_lowest(n, a_in, n, w, v_out); }

void eigh_lowest(size_t n, const cplx* a_in, size_t n_keep, double* w, cplx* v_out) {
  if (n == 0) return;
  // working copy, column-major a[j*n + i] = A(i, j)
  std::vector<cplx> a(a_in, a_in + n * n);
  std::vector<cplx> q(n * n, cplx(0.0, 0.0));
  for (size_t i = 0; i < n; ++i) q[i * n + i] = 1.0;
  std::vector<cplx> v(n), p(n), wv(n);
  for (size_t k = 0; k + 2 < n; ++k) {
    size_t m = n - k - 1;  // length of the column below the diagonal
    cplx* col = &a[k * n + k + 1];
    double xn2 = 0.0;
    for (size_t i = 0; i < m; ++i) xn2 += std::norm(col[i]);
    double xn = std::sqrt(xn2);
    if (xn == 0.0) continue;
    cplx x0 = col[0];
    cplx ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cplx(1.0, 0.0);
    cplx alpha = -ph * xn;
    for (size_t i = 0; i < m; ++i) v[i] = col[i];
    v[0] -= alpha;
    double vn2 = 0.0;
    for (size_t i = 0; i < m; ++i) vn2 += std::norm(v[i]);
    if (vn2 == 0.0) continue;
    double vn = std::sqrt(vn2);
    for (size_t i = 0; i < m; ++i) v[i] /= vn;
    // p = A_sub v, sub-block rows/cols k+1..n-1
    for (size_t i = 0; i < m; ++i) p[i] = 0.0;
    for (size_t j = 0; j < m; ++j) {
      cplx vj = v[j];
      const cplx* aj = &a[(k + 1 + j) * n + k + 1];
      for (size_t i = 0; i < m; ++i) p[i] += aj[i] * vj;
    }
    cplx K(0.0, 0.0);
    for (size_t i = 0; i < m; ++i) K += std::conj(v[i]) * p[i];
    double Kr = K.real();
    for (size_t i = 0; i < m; ++i) wv[i] = p[i] - Kr * v[i];
    // A_sub -= 2 (v w^H + w v^H)
    for (size_t j = 0; j < m; ++j) {
      cplx cw = std::conj(wv[j]), cv = std::conj(v[j]);
      cplx* aj = &a[(k + 1 + j) * n + k + 1];
      for (size_t i = 0; i < m; ++i) aj[i] -= 2.0 * (v[i] * cw + wv[i] * cv);
    }
    // column k and row k below/right of the diagonal
    col[0] = alpha;
    for (size_t i = 1; i < m; ++i) col[i] = 0.0;
    a[(k + 1) * n + k] = std::conj(alpha);
    for (size_t i = 1; i < m; ++i) a[(k + 1 + i) * n + k] = 0.0;
    // Q <- Q H  (columns k+1..n-1): s = 2 Q_sub v, Q_sub -= s v^H
    std::vector<cplx>& sv = p;  // reuse p as scratch of length >= n
    if (sv.size() < n) sv.resize(n);
    for (size_t r = 0; r < n; ++r) sv[r] = 0.0;
    for (size_t j = 0; j < m; ++j) {
      cplx vj = v[j];
      const cplx* qj = &q[(k + 1 + j) * n];
      for (size_t r = 0; r < n; ++r) sv[r] += qj[r] * vj;
    }
    for (size_t j = 0; j < m; ++j) {
      cplx cvj = 2.0 * std::conj(v[j]);
      cplx* qj = &q[(k + 1 + j) * n];
      for (size_t r = 0; r < n; ++r) qj[r] -= sv[r] * cvj;
    }
  }
  // Hermitian tridiagonal: diag real, sub-diag e_k = a(k+1, k)
  std::vector<double> d(n), e(n, 0.0);
  std::vector<cplx> dph(n);
  dph[0] = 1.0;
  for (size_t k = 0; k < n; ++k) d[k] = a[k * n + k].real();
  for (size_t k = 0; k + 1 < n; ++k) {
    cplx ek = a[k * n + k + 1];
    double ab = std::abs(ek);
    e[k] = ab;
    dph[k + 1] = ab > 0 ? dph[k] * (ek / ab) : dph[k];
  }
  std::vector<double> z(n * n, 0.0);
  for (size_t i = 0; i < n; ++i) z[i * n + i] = 1.0;
  tridiag_ql(d, e, z, n);
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return d[x] < d[y]; });
  // V = Q · diag(dph) · Z  (lowest n_keep columns)
  for (size_t jj = 0; jj < n; ++jj) w[jj] = d[order[jj]];
  for (size_t jj = 0; jj < n_keep; ++jj) {
    size_t j = order[jj];
    cplx* out = &v_out[jj * n];
    for (size_t r = 0; r < n; ++r) out[r] = 0.0;
    for (size_t t = 0; t < n; ++t) {
      cplx c = dph[t] * z[j * n + t];
      if (c == cplx(0.0, 0.0)) continue;
      const cplx* qt = &q[t * n];
      for (size_t r = 0; r < n; ++r) out[r] += qt[r] * c;
    }
  }
}

}  // namespace orc

// Small dense complex linear algebra on the host (FP64), for the recycled
// Krylov subspace of the GMRES solver: eigenvalues of an upper Hessenberg
// matrix, selected eigenvectors, harmonic Ritz vectors of an Arnoldi
// Hessenberg and a thin QR.  Sizes are at most kMaxRestart + 1, so plain
// loops are enough; the arithmetic is written out on (re, im) pairs so it
// carries no C99 Annex G special-value handling.
//
// Eigenvalues: single-shift QR iteration with Wilkinson shifts and Givens
// rotations on the active window (Francis 1961/62; Wilkinson, "The Algebraic
// Eigenvalue Problem", 1965, ch. 8).  Eigenvectors: inverse iteration with the
// Hessenberg LU of H - lambda I (Wilkinson ch. 9).  Harmonic Ritz vectors:
// Morgan, "GMRES with deflated restarting", SIAM J. Sci. Comput. 24 (2002):
// eigenvectors of H_m + |h_{m+1,m}|^2 H_m^{-H} e_m e_m^T.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace vie {
namespace dense {

struct zd {
  double re = 0, im = 0;
  zd() = default;
  zd(double r, double i = 0) : re(r), im(i) {}
};
inline zd operator+(zd a, zd b) { return {a.re + b.re, a.im + b.im}; }
inline zd operator-(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
inline zd operator-(zd a) { return {-a.re, -a.im}; }
inline zd operator*(zd a, zd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline zd operator*(double s, zd a) { return {s * a.re, s * a.im}; }
inline zd conj(zd a) { return {a.re, -a.im}; }
inline double norm2(zd a) { return a.re * a.re + a.im * a.im; }
inline double abs(zd a) { return std::sqrt(norm2(a)); }  // entries are far from overflow
inline zd operator/(zd a, zd b) {
  // Smith's scaling, so |b| near the range limits does not overflow
  if (std::fabs(b.re) >= std::fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
inline zd csqrt(zd a) {
  const double m = abs(a);
  if (m == 0) return {0, 0};
  const double re = std::sqrt(0.5 * (m + std::fabs(a.re)));
  const double im = a.im / (2 * re);
  return a.re >= 0 ? zd(re, im) : zd(std::fabs(im), a.im >= 0 ? re : -re);
}

// Column-major n x m matrix.
struct Mat {
  int n = 0, m = 0;
  std::vector<zd> a;
  Mat() = default;
  Mat(int rows, int cols) : n(rows), m(cols), a((size_t)rows * cols) {}
  zd& operator()(int i, int j) { return a[(size_t)j * n + i]; }
  const zd& operator()(int i, int j) const { return a[(size_t)j * n + i]; }
};

// Rotation [c s; -conj(s) c], real c, that maps (x, y) to (r, 0).
inline void givens(zd x, zd y, double* c, zd* s) {
  const double ax = abs(x), ay = abs(y);
  if (ay == 0) {
    *c = 1;
    *s = {0, 0};
    return;
  }
  if (ax == 0) {
    *c = 0;
    *s = (1.0 / ay) * conj(y);
    return;
  }
  const double r = std::sqrt(ax * ax + ay * ay);
  *c = ax / r;
  *s = (1.0 / (ax * r)) * (x * conj(y));
}

// Solves U x = b for an upper Hessenberg U (entries below the subdiagonal
// ignored), Gaussian elimination with adjacent-row pivoting: O(n^2).  False
// when U is singular.
inline bool hessenberg_solve(Mat U, std::vector<zd>& b) {
  const int n = U.n;
  for (int k = 0; k + 1 < n; k++) {
    if (norm2(U(k + 1, k)) > norm2(U(k, k))) {
      for (int j = k; j < n; j++) std::swap(U(k, j), U(k + 1, j));
      std::swap(b[k], b[k + 1]);
    }
    if (norm2(U(k, k)) == 0) return false;
    const zd f = U(k + 1, k) / U(k, k);
    for (int j = k + 1; j < n; j++) U(k + 1, j) = U(k + 1, j) - f * U(k, j);
    b[k + 1] = b[k + 1] - f * b[k];
  }
  for (int i = n - 1; i >= 0; i--) {
    if (norm2(U(i, i)) == 0) return false;
    zd t = b[i];
    for (int j = i + 1; j < n; j++) t = t - U(i, j) * b[j];
    b[i] = t / U(i, i);
  }
  return true;
}

// Eigenvalues of an upper Hessenberg H (entries below the subdiagonal are
// ignored).  False when the iteration does not settle.
inline bool hessenberg_eigvals(Mat H, std::vector<zd>& w) {
  const int n = H.n;
  const double eps = 2.220446049250313e-16;
  w.assign(n, zd(0, 0));
  int hi = n - 1, iter = 0, total = 0;
  while (hi >= 0) {
    int l = hi;
    for (; l > 0; l--) {
      const double scale = abs(H(l, l)) + abs(H(l - 1, l - 1));
      if (abs(H(l, l - 1)) <= eps * (scale > 0 ? scale : 1.0)) {
        H(l, l - 1) = {0, 0};
        break;
      }
    }
    if (l == hi) {
      w[hi] = H(hi, hi);
      hi--;
      iter = 0;
      continue;
    }
    if (++total > 60 * n) return false;
    zd shift;
    if (iter > 0 && iter % 11 == 0) {
      shift = H(hi, hi) + zd(abs(H(hi, hi - 1)), 0);  // exceptional shift
    } else {
      const zd a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
      const zd half = 0.5 * (a - d), disc = csqrt(half * half + b * c);
      const zd m1 = 0.5 * (a + d) + disc, m2 = 0.5 * (a + d) - disc;
      shift = norm2(m1 - d) < norm2(m2 - d) ? m1 : m2;
    }
    iter++;
    // one implicit sweep over the active window l..hi
    zd x = H(l, l) - shift, y = H(l + 1, l);
    for (int k = l; k < hi; k++) {
      double c;
      zd s;
      givens(x, y, &c, &s);
      for (int j = std::max(l, k - 1); j <= hi; j++) {
        const zd u = H(k, j), v = H(k + 1, j);
        H(k, j) = c * u + s * v;
        H(k + 1, j) = c * v - conj(s) * u;
      }
      const int rmax = std::min(k + 2, hi);
      for (int i = l; i <= rmax; i++) {
        const zd u = H(i, k), v = H(i, k + 1);
        H(i, k) = c * u + conj(s) * v;
        H(i, k + 1) = c * v - s * u;
      }
      if (k < hi - 1) {
        x = H(k + 1, k);
        y = H(k + 2, k);
      }
    }
  }
  return true;
}

// Eigenvector of an upper Hessenberg H for the eigenvalue lambda (unit
// 2-norm), by two steps of inverse iteration with the pivoted Hessenberg LU
// of H - lambda I (a pivot that vanishes is replaced by eps |H|).
inline std::vector<zd> hessenberg_eigvec(const Mat& H, zd lambda) {
  const int n = H.n;
  double hn = 0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i <= std::min(j + 1, n - 1); i++) hn = std::max(hn, abs(H(i, j)));
  const double tiny = 2.220446049250313e-16 * (hn > 0 ? hn : 1.0);
  // U: upper triangular after eliminating the subdiagonal; row swaps only
  // between neighbours, recorded in piv / mult.
  Mat U(n, n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i <= std::min(j + 1, n - 1); i++) U(i, j) = H(i, j);
  for (int i = 0; i < n; i++) U(i, i) = U(i, i) - lambda;
  std::vector<zd> mult(n);
  std::vector<char> swp(n, 0);
  for (int k = 0; k + 1 < n; k++) {
    if (norm2(U(k + 1, k)) > norm2(U(k, k))) {
      swp[k] = 1;
      for (int j = k; j < n; j++) std::swap(U(k, j), U(k + 1, j));
    }
    if (norm2(U(k, k)) == 0) U(k, k) = zd(tiny, 0);
    mult[k] = U(k + 1, k) / U(k, k);
    U(k + 1, k) = {0, 0};
    for (int j = k + 1; j < n; j++) U(k + 1, j) = U(k + 1, j) - mult[k] * U(k, j);
  }
  if (norm2(U(n - 1, n - 1)) == 0) U(n - 1, n - 1) = zd(tiny, 0);
  std::vector<zd> v(n, zd(1, 0));
  for (int it = 0; it < 2; it++) {
    if (it > 0)  // apply the row operations of the LU to the right-hand side
      for (int k = 0; k + 1 < n; k++) {
        if (swp[k]) std::swap(v[k], v[k + 1]);
        v[k + 1] = v[k + 1] - mult[k] * v[k];
      }
    for (int i = n - 1; i >= 0; i--) {
      zd s = v[i];
      for (int j = i + 1; j < n; j++) s = s - U(i, j) * v[j];
      v[i] = s / U(i, i);
    }
    double nrm = 0;
    for (int i = 0; i < n; i++) nrm += norm2(v[i]);
    nrm = std::sqrt(nrm);
    for (int i = 0; i < n; i++) v[i] = (1.0 / nrm) * v[i];
  }
  return v;
}

// k harmonic Ritz vectors (columns of P, m x k) of the Arnoldi Hessenberg
// Hbar ((m+1) x m): eigenvectors of H_m + |h_{m+1,m}|^2 H_m^{-H} e_m e_m^T
// for the k eigenvalues of smallest modulus.
inline bool harmonic_ritz(const Mat& Hbar, int k, Mat& P) {
  const int m = Hbar.m;
  if (k < 1 || k > m || Hbar.n != m + 1) return false;
  // f = H_m^{-H} e_m.  H_m^H is lower Hessenberg; reversing its row and
  // column order makes it upper Hessenberg, solved in O(m^2).
  Mat R(m, m);
  for (int j = 0; j < m; j++)
    for (int i = 0; i < m; i++) R(i, j) = conj(Hbar(m - 1 - j, m - 1 - i));
  std::vector<zd> f(m, zd(0, 0));
  f[0] = {1, 0};
  if (!hessenberg_solve(R, f)) return false;
  std::reverse(f.begin(), f.end());
  Mat M(m, m);
  for (int j = 0; j < m; j++)
    for (int i = 0; i < m; i++) M(i, j) = Hbar(i, j);
  const double h2 = norm2(Hbar(m, m - 1));
  for (int i = 0; i < m; i++) M(i, m - 1) = M(i, m - 1) + h2 * f[i];
  std::vector<zd> w;
  if (!hessenberg_eigvals(M, w)) return false;
  std::vector<int> idx(m);
  for (int i = 0; i < m; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return norm2(w[a]) < norm2(w[b]); });
  P = Mat(m, k);
  for (int c = 0; c < k; c++) {
    const std::vector<zd> v = hessenberg_eigvec(M, w[idx[c]]);
    for (int i = 0; i < m; i++) P(i, c) = v[i];
  }
  return true;
}

// Thin QR of A (n x k, n >= k) by Gram-Schmidt with one reorthogonalisation:
// A = Q R, Q n x k with orthonormal columns, R k x k upper triangular.
// False when a column is (numerically) dependent on the ones before it.
inline bool thin_qr(const Mat& A, Mat& Q, Mat& R) {
  const int n = A.n, k = A.m;
  Q = A;
  R = Mat(k, k);
  for (int j = 0; j < k; j++) {
    double n0 = 0;
    for (int r = 0; r < n; r++) n0 += norm2(Q(r, j));
    for (int pass = 0; pass < 2; pass++)
      for (int i = 0; i < j; i++) {
        zd d = {0, 0};
        for (int r = 0; r < n; r++) d = d + conj(Q(r, i)) * Q(r, j);
        for (int r = 0; r < n; r++) Q(r, j) = Q(r, j) - d * Q(r, i);
        R(i, j) = R(i, j) + d;
      }
    double nrm = 0;
    for (int r = 0; r < n; r++) nrm += norm2(Q(r, j));
    if (!(nrm > 1e-24 * n0) || !(nrm > 0)) return false;
    nrm = std::sqrt(nrm);
    R(j, j) = {nrm, 0};
    for (int r = 0; r < n; r++) Q(r, j) = (1.0 / nrm) * Q(r, j);
  }
  return true;
}

// Inverse of an upper triangular k x k R.
inline bool upper_inverse(const Mat& R, Mat& Ri) {
  const int k = R.n;
  Ri = Mat(k, k);
  for (int j = 0; j < k; j++) {
    if (norm2(R(j, j)) == 0) return false;
    Ri(j, j) = zd(1, 0) / R(j, j);
    for (int i = j - 1; i >= 0; i--) {
      zd s = {0, 0};
      for (int l = i + 1; l <= j; l++) s = s + R(i, l) * Ri(l, j);
      Ri(i, j) = (zd(0, 0) - s) / R(i, i);
    }
  }
  return true;
}

}  // namespace dense
}  // namespace vie

// Host symmetric eigen-solvers used on small matrices: the Lanczos projected
// tridiagonal and rho blocks.  These take the place of SciFortran/LAPACK
// eigh (src/MATRIX/MATRIX_BLOCKS.f90:623, src/DMRG_SUPERBLOCK.f90:449).
//
// Method: Householder reduction to tridiagonal form, then the implicit QL
// iteration with Wilkinson-type shift (Golub & Van Loan, "Matrix
// Computations", sec. 8.3; Wilkinson & Reinsch, Handbook vol. II, II/2-II/3).
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

#include "nss_internal.h"

namespace nss {

// Implicit QL on the symmetric tridiagonal (d, e) with e[i] coupling i,i+1.
// Applies every plane rotation to the columns of the nr x n row-major z.
static void ql_implicit(int n, double* d, double* e, int nr, double* z) {
  if (n <= 1) return;
  e[n - 1] = 0.0;
  // Deflation is tested against the norm scale seen so far (running max of
  // |d|+|e|), so clusters of eigenvalues far below ||T|| (the null space of a
  // rank-deficient rho) converge to absolute accuracy eps*||T||.
  double tst = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    tst = std::max(tst, std::fabs(d[l]) + std::fabs(e[l]));
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 2.220446049250313e-16 * std::max(dd, tst)) break;
      }
      if (m == l) break;
      if (++iter > 80) throw Error(NSS_E_NOCONV, "tridiagonal QL did not converge");
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool deflated = false;
      for (; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < nr; ++k) {
          double* zk = z + (size_t)k * n;
          f = zk[i + 1];
          zk[i + 1] = s * zk[i] + c * f;
          zk[i] = c * zk[i] - s * f;
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

// Sort eigenvalues ascending and permute the columns of z accordingly.
static void sort_pairs(int n, double* w, int nr, double* z) {
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] < w[b]; });
  std::vector<double> tw(n), tz((size_t)nr * n);
  for (int j = 0; j < n; ++j) {
    tw[j] = w[idx[j]];
    for (int k = 0; k < nr; ++k) tz[(size_t)k * n + j] = z[(size_t)k * n + idx[j]];
  }
  std::copy(tw.begin(), tw.end(), w);
  std::copy(tz.begin(), tz.end(), z);
}

void tridiag_eig(int n, const double* d, const double* e, double* w, int nr, const int* rows,
                 double* zrows) {
  std::vector<double> ee(n, 0.0);
  for (int i = 0; i < n; ++i) w[i] = d[i];
  for (int i = 0; i + 1 < n; ++i) ee[i] = e[i];
  for (int k = 0; k < nr; ++k) {
    double* zk = zrows + (size_t)k * n;
    std::fill(zk, zk + n, 0.0);
    zk[rows[k]] = 1.0;
  }
  ql_implicit(n, w, ee.data(), nr, zrows);
  sort_pairs(n, w, nr, zrows);
}

// Lowest eigenvalue by bisection on Sturm counts (Wilkinson, The Algebraic
// Eigenvalue Problem, 5.37-5.39: the negative pivots of the LDL^T of T - x I
// count the eigenvalues below x), from the Gershgorin interval down to a few
// ulps of its width.  The last component of the unit eigenvector by inverse
// iteration: two solves with T - theta I, factored by Gaussian elimination
// with partial pivoting (the tridiagonal LU of LAPACK's dgttrf, U with two
// superdiagonals), zero pivots replaced by eps ||T||.
void tridiag_lowest(int n, const double* d, const double* e, double* theta, double* zlast, double* zfull) {
  if (n <= 0) return;
  if (n == 1) {
    *theta = d[0];
    *zlast = 1.0;
    if (zfull) zfull[0] = 1.0;
    return;
  }
  const double eps = std::numeric_limits<double>::epsilon();
  double lo = HUGE_VAL, hi = -HUGE_VAL, emax2 = 0.0, tnorm = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
    tnorm = std::max(tnorm, std::fabs(d[i]) + r);
    if (i + 1 < n) emax2 = std::max(emax2, e[i] * e[i]);
  }
  const double pivmin = std::numeric_limits<double>::min() * std::max(1.0, emax2);
  auto below = [&](double x) {  // number of eigenvalues < x (at least one?)
    double q = d[0] - x;
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) return true;
    for (int i = 1; i < n; ++i) {
      q = d[i] - x - e[i - 1] * e[i - 1] / q;
      if (std::fabs(q) < pivmin) q = -pivmin;
      if (q < 0.0) return true;
    }
    return false;
  };
  const double tol = 2.0 * eps * std::max(tnorm, std::numeric_limits<double>::min());
  for (int it = 0; it < 256 && hi - lo > tol; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (below(mid)) hi = mid;
    else lo = mid;
  }
  const double th = 0.5 * (lo + hi);
  *theta = th;
  // LU of T - th I with row interchanges: dl (multipliers), dd, du, du2, piv
  std::vector<double> dl(n - 1), dd(n), du(n - 1), du2(n > 2 ? n - 2 : 1, 0.0), x(n);
  std::vector<char> piv(n - 1, 0);
  for (int i = 0; i < n; ++i) dd[i] = d[i] - th;
  for (int i = 0; i + 1 < n; ++i) dl[i] = du[i] = e[i];
  const double tiny = eps * std::max(tnorm, std::numeric_limits<double>::min());
  for (int i = 0; i + 1 < n; ++i) {
    if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
      if (dd[i] == 0.0) dd[i] = tiny;
      const double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
      if (i + 2 < n) du2[i] = 0.0;
    } else {
      const double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = t - f * dd[i + 1];
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      piv[i] = 1;
    }
  }
  if (dd[n - 1] == 0.0) dd[n - 1] = tiny;
  // start: a fixed sign pattern of unit entries (not orthogonal to any
  // eigenvector of a generic T)
  for (int i = 0; i < n; ++i) x[i] = (i % 3 == 1) ? -1.0 : 1.0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i + 1 < n; ++i) {  // L: apply the interchanges and multipliers
      if (piv[i]) std::swap(x[i], x[i + 1]);
      x[i + 1] -= dl[i] * x[i];
    }
    x[n - 1] /= dd[n - 1];  // U
    x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dd[n - 2];
    for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dd[i];
    double nrm = 0.0, mx = 0.0;
    for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(x[i]));
    for (int i = 0; i < n; ++i) nrm += (x[i] / mx) * (x[i] / mx);
    nrm = mx * std::sqrt(nrm);
    for (int i = 0; i < n; ++i) x[i] /= nrm;
  }
  *zlast = x[n - 1];
  if (zfull) std::copy(x.begin(), x.end(), zfull);
}

void tridiag_eig_full(int n, const double* d, const double* e, double* w, double* z) {
  std::vector<int> rows(n);
  std::iota(rows.begin(), rows.end(), 0);
  tridiag_eig(n, d, e, w, n, rows.data(), z);
}

// a (n x n row-major, symmetric) -> eigenvectors as columns, w ascending.
void sym_eig(int n, double* a, double* w) {
  if (n <= 0) return;
  if (n == 1) {
    w[0] = a[0];
    a[0] = 1.0;
    return;
  }
  std::vector<double> A(a, a + (size_t)n * n);
  std::vector<double> d(n), e(n, 0.0);
  std::vector<std::vector<double>> vs;  // Householder vectors (acting on k+1..n-1)
  std::vector<double> vnorm2;
  std::vector<double> p(n), q(n);
  auto at = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double alpha = 0.0;
    for (int i = 0; i < m; ++i) alpha += at(k + 1 + i, k) * at(k + 1 + i, k);
    alpha = std::sqrt(alpha);
    std::vector<double> v(m);
    for (int i = 0; i < m; ++i) v[i] = at(k + 1 + i, k);
    if (alpha == 0.0) {
      vs.emplace_back();
      vnorm2.push_back(0.0);
      continue;
    }
    if (v[0] > 0.0) alpha = -alpha;
    v[0] -= alpha;
    double vn = 0.0;
    for (double x : v) vn += x * x;
    if (vn == 0.0) {
      vs.emplace_back();
      vnorm2.push_back(0.0);
      continue;
    }
    // A22 <- H A22 H, H = I - 2 v v^T / vn
    const double beta = 2.0 / vn;
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      const double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
      for (int j = 0; j < m; ++j) s += row[j] * v[j];
      p[i] = beta * s;
    }
    double K = 0.0;
    for (int i = 0; i < m; ++i) K += v[i] * p[i];
    K *= 0.5 * beta;
    for (int i = 0; i < m; ++i) q[i] = p[i] - K * v[i];
    for (int i = 0; i < m; ++i) {
      double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
      for (int j = 0; j < m; ++j) row[j] -= v[i] * q[j] + q[i] * v[j];
    }
    at(k + 1, k) = alpha;
    at(k, k + 1) = alpha;
    for (int i = 1; i < m; ++i) at(k + 1 + i, k) = at(k, k + 1 + i) = 0.0;
    vs.push_back(v);
    vnorm2.push_back(vn);
  }
  for (int i = 0; i < n; ++i) d[i] = at(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = at(i + 1, i);
  // Q = H_0 H_1 ... accumulated backwards into z (row-major n x n)
  std::vector<double> Z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
  for (int k = (int)vs.size() - 1; k >= 0; --k) {
    if (vnorm2[k] == 0.0) continue;
    const std::vector<double>& v = vs[k];
    const int m = (int)v.size();
    const double beta = 2.0 / vnorm2[k];
    // Z[k+1:, k+1:] <- (I - beta v v^T) Z[k+1:, k+1:]
    for (int j = k + 1; j < n; ++j) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += v[i] * Z[(size_t)(k + 1 + i) * n + j];
      s *= beta;
      if (s == 0.0) continue;
      for (int i = 0; i < m; ++i) Z[(size_t)(k + 1 + i) * n + j] -= s * v[i];
    }
  }
  ql_implicit(n, d.data(), e.data(), n, Z.data());
  for (int i = 0; i < n; ++i) w[i] = d[i];
  sort_pairs(n, w, n, Z.data());
  std::copy(Z.begin(), Z.end(), a);
}

}  // namespace nss

// Host-only test hooks (no device needed): expose the eigen-solvers so the
// CPU test-suite can check them against LAPACK.
extern "C" int nss_host_sym_eig(int n, double* a, double* w) {
  try {
    nss::sym_eig(n, a, w);
  } catch (const nss::Error& e) {
    nss::set_last_error(e.what());
    return e.code;
  }
  return NSS_OK;
}

extern "C" int nss_host_tridiag_lowest(int n, const double* d, const double* e, double* theta, double* zlast) {
  try {
    nss::tridiag_lowest(n, d, e, theta, zlast);
  } catch (const nss::Error& er) {
    nss::set_last_error(er.what());
    return er.code;
  }
  return NSS_OK;
}

extern "C" int nss_host_tridiag_eig(int n, const double* d, const double* e, double* w, double* z) {
  try {
    nss::tridiag_eig_full(n, d, e, w, z);
  } catch (const nss::Error& er) {
    nss::set_last_error(er.what());
    return er.code;
  }
  return NSS_OK;
}

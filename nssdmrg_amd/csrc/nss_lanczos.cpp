// GPU Lanczos eigen-solver for the superblock Hamiltonian: the role of
// SciFortran sp_eigh in sb_diag (src/DMRG_SUPERBLOCK.f90:376-437).
//
// Thick-restart Lanczos (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000))
// with full re-orthogonalisation (two classical Gram-Schmidt passes).  The
// Krylov basis lives in HBM; H*v, projections and updates run as kernels on
// the context stream; only the projection coefficients (<= kmax doubles per
// iteration) come back to the host, which keeps the small projected matrix
// and solves it in FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "nss_internal.h"

namespace nss {

namespace {

struct Lz {
  nss_ctx* ctx;
  SbDevice* sb;
  int64_t n;
  int kmax;
  DevArr<double> V;     // (kmax+1) x n
  DevArr<double> w;     // n
  DevArr<double> coef;  // 2*(kmax+1) + 2
  DevArr<double> part;  // reduction scratch
  std::vector<double> T;  // (kmax+1)^2, host projected matrix
  int numop = 0, numop32 = 0;

  double* vec(int j) { return V.p + (size_t)j * n; }
  double& t(int i, int j) { return T[(size_t)i * (kmax + 1) + j]; }

  void apply(const double* x, double* y, bool f32) {
    if (f32) {
      sb_hv_fp32(ctx, sb, x, y);
      ++numop32;
    } else {
      sb_hv(ctx, sb, x, y);
      ++numop;
    }
  }
  // w <- w - V[0..m) h (twice), returns h (length m) and beta = ||w|| on host
  void orthogonalize(int m, std::vector<double>& h, double& beta) {
    hipStream_t s = ctx->stream;
    double* h1 = coef.p;
    double* h2 = coef.p + (kmax + 1);
    double* nb = coef.p + 2 * (kmax + 1);
    launch_gemv_t(V.p, n, m, w.p, n, part.p, h1, s);
    launch_gemv_n(V.p, n, m, h1, w.p, n, 0, s);
    launch_gemv_t(V.p, n, m, w.p, n, part.p, h2, s);
    launch_gemv_n(V.p, n, m, h2, w.p, n, 0, s);
    launch_norm(w.p, n, part.p, nb, s);
    std::vector<double> hc(2 * (kmax + 1) + 1);
    NSS_HIP(hipMemcpyAsync(hc.data(), coef.p, hc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    NSS_HIP(hipStreamSynchronize(s));
    h.assign(m, 0.0);
    for (int i = 0; i < m; ++i) h[i] = hc[i] + hc[(kmax + 1) + i];
    beta = hc[2 * (kmax + 1)];
  }
};

}  // namespace

// One thick-restart Lanczos run from the (unnormalised) device vector v0.
// Returns Ritz values; Ritz vectors written to x (neigen x n).
// soft: stop (returning the current Ritz vectors) instead of failing when
// maxiter restart cycles are exhausted (used by the FP32 pre-solve).
static std::vector<double> trlan(Lz& L, const double* v0, int neigen, int maxiter, double tol, bool f32,
                                 double* x, bool soft = false) {
  const int64_t n = L.n;
  const int kmax = L.kmax;
  hipStream_t s = L.ctx->stream;
  L.T.assign((size_t)(kmax + 1) * (kmax + 1), 0.0);
  // v_0 = v0 / ||v0||
  double* nb = L.coef.p + 2 * (kmax + 1);
  launch_norm(v0, n, L.part.p, nb, s);
  launch_div(v0, L.vec(0), n, nb, s);
  int k = 0;       // number of kept Ritz vectors (arrow size) after a restart
  int j = 0;       // current basis vector index being expanded
  int cycles = 0;
  std::vector<double> th;
  std::vector<double> Y;  // m x m eigenvectors (row-major, columns = vectors)
  const int nkeep = std::min(std::max(2 * neigen, neigen + 8), std::max(neigen, kmax / 2));
  for (;;) {
    // expand basis from j to kmax
    bool converged = false, breakdown = false;
    int m = 0;
    double beta = 0.0;
    for (; j < kmax; ++j) {
      L.apply(L.vec(j), L.w.p, f32);
      std::vector<double> h;
      L.orthogonalize(j + 1, h, beta);
      // projected matrix: full column j (arrow after a restart, tridiagonal otherwise)
      for (int i = 0; i <= j; ++i) {
        const bool keep = (i == j) || (i == j - 1) || (j == k && i < k);
        const double v = keep ? h[i] : 0.0;
        L.t(i, j) = v;
        L.t(j, i) = v;
      }
      m = j + 1;
      L.t(m, j) = L.t(j, m) = beta;
      const bool can_check = m >= neigen;
      const bool cheap = (k == 0);
      const bool at_end = (m == kmax) || (m == n);
      double hn = 0.0;
      for (int i = 0; i < m; ++i) hn = std::max(hn, std::fabs(L.t(i, i)));
      breakdown = beta <= 1e-14 * std::max(1.0, hn);
      if (can_check && (cheap || at_end || breakdown || (m - k) % 5 == 0)) {
        th.assign(m, 0.0);
        Y.assign((size_t)m * m, 0.0);
        if (k == 0) {
          std::vector<double> d(m), e(m, 0.0);
          for (int i = 0; i < m; ++i) d[i] = L.t(i, i);
          for (int i = 0; i + 1 < m; ++i) e[i] = L.t(i + 1, i);
          if (at_end || breakdown) {
            tridiag_eig_full(m, d.data(), e.data(), th.data(), Y.data());
          } else {
            // only the last row of the eigenvector matrix for residuals
            std::vector<double> zl(m);
            int row = m - 1;
            tridiag_eig(m, d.data(), e.data(), th.data(), 1, &row, zl.data());
            for (int i = 0; i < m; ++i) Y[(size_t)(m - 1) * m + i] = zl[i];
          }
        } else {
          for (int a = 0; a < m; ++a)
            for (int b = 0; b < m; ++b) Y[(size_t)a * m + b] = L.t(a, b);
          sym_eig(m, Y.data(), th.data());
        }
        converged = true;
        for (int i = 0; i < neigen; ++i) {
          const double res = std::fabs(beta * Y[(size_t)(m - 1) * m + i]);
          if (!(res <= tol * std::max(1.0, std::fabs(th[i])))) converged = false;
        }
        if (breakdown) converged = true;
        if (converged || at_end) {
          if (k == 0 && !(at_end || breakdown)) {
            std::vector<double> d(m), e(m, 0.0);
            for (int i = 0; i < m; ++i) d[i] = L.t(i, i);
            for (int i = 0; i + 1 < m; ++i) e[i] = L.t(i + 1, i);
            tridiag_eig_full(m, d.data(), e.data(), th.data(), Y.data());
          }
          break;
        }
      }
      if (breakdown) break;
      launch_div(L.w.p, L.vec(j + 1), n, L.coef.p + 2 * (kmax + 1), s);
    }
    if (m == 0) throw Error(NSS_E_STATE, "lanczos: empty Krylov space");
    if (m < neigen) throw Error(NSS_E_NOCONV, "lanczos: invariant Krylov space smaller than neigen");
    const bool give_up = soft && cycles >= maxiter;
    if (converged || breakdown || m == n || give_up) {
      // Ritz vectors x_i = V[0..m) y_i
      DevArr<double> yd;
      yd.resize(L.ctx->pool, (size_t)m * neigen);
      std::vector<double> yc((size_t)m * neigen);
      for (int i = 0; i < neigen; ++i)
        for (int a = 0; a < m; ++a) yc[(size_t)i * m + a] = Y[(size_t)a * m + i];
      NSS_HIP(hipMemcpyAsync(yd.p, yc.data(), yc.size() * sizeof(double), hipMemcpyHostToDevice, s));
      for (int i = 0; i < neigen; ++i) {
        launch_gemv_n(L.V.p, n, m, yd.p + (size_t)i * m, x + (size_t)i * n, n, 1, s);
        launch_norm(x + (size_t)i * n, n, L.part.p, nb, s);
        launch_div(x + (size_t)i * n, x + (size_t)i * n, n, nb, s);
      }
      NSS_HIP(hipStreamSynchronize(s));
      return std::vector<double>(th.begin(), th.begin() + neigen);
    }
    // thick restart: keep the nkeep lowest Ritz vectors + the residual direction
    if (++cycles > maxiter) throw Error(NSS_E_NOCONV, "lanczos: no convergence within LANC_NITER restarts");
    const int kk = std::min(nkeep, m - 1);
    DevArr<double> U, yd;
    U.resize(L.ctx->pool, (size_t)kk * n);
    yd.resize(L.ctx->pool, (size_t)m * kk);
    std::vector<double> yc((size_t)m * kk);
    for (int i = 0; i < kk; ++i)
      for (int a = 0; a < m; ++a) yc[(size_t)i * m + a] = Y[(size_t)a * m + i];
    NSS_HIP(hipMemcpyAsync(yd.p, yc.data(), yc.size() * sizeof(double), hipMemcpyHostToDevice, s));
    for (int i = 0; i < kk; ++i) launch_gemv_n(L.V.p, n, m, yd.p + (size_t)i * m, U.p + (size_t)i * n, n, 1, s);
    // residual direction: w / beta is v_m (not yet stored when m == kmax)
    launch_div(L.w.p, L.vec(kk), n, L.coef.p + 2 * (kmax + 1), s);
    NSS_HIP(hipMemcpyAsync(L.V.p, U.p, (size_t)kk * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    NSS_HIP(hipStreamSynchronize(s));
    std::fill(L.T.begin(), L.T.end(), 0.0);
    for (int i = 0; i < kk; ++i) {
      L.t(i, i) = th[i];
      const double si = beta * Y[(size_t)(m - 1) * m + i];
      L.t(i, kk) = L.t(kk, i) = si;
    }
    k = kk;
    j = kk;
  }
}

LanczosOut lanczos_device(nss_ctx* ctx, SbDevice* sb, int neigen, int kmax, int maxiter, double tol,
                          uint64_t seed, int v0_mode, int precision, double* xdev) {
  const int64_t n = sb_n(sb);
  NSS_REQUIRE(n > 0, NSS_E_STATE, "lanczos: empty superblock");
  NSS_REQUIRE(neigen >= 1 && neigen <= 4 && neigen <= n, NSS_E_ARG, "lanczos: 1 <= neigen <= min(4,n)");
  if (kmax <= 0) kmax = 150;
  kmax = (int)std::min<int64_t>(kmax, n);
  kmax = std::max(kmax, std::min<int>((int)n, neigen + 2));
  if (maxiter <= 0) maxiter = 512;
  Lz L;
  L.ctx = ctx;
  L.sb = sb;
  L.n = n;
  L.kmax = kmax;
  L.V.resize(ctx->pool, (size_t)(kmax + 1) * n);
  L.w.resize(ctx->pool, n);
  L.coef.resize(ctx->pool, 2 * (size_t)(kmax + 1) + 2);
  L.part.resize(ctx->pool, (size_t)gemv_t_chunks(n) * (kmax + 1) + 1);
  DevArr<double> v0;
  v0.resize(ctx->pool, n);
  if (v0_mode == 1) launch_fill_const(v0.p, n, 1.0 / std::sqrt((double)n), ctx->stream);
  else launch_fill_splitmix(v0.p, n, seed, ctx->stream);
  LanczosOut out;
  if (precision == NSS_FP32_MIXED) {
    // FP32-MFMA H*v to an FP32-attainable residual, then the FP64 restart
    std::vector<double> e32 = trlan(L, v0.p, neigen, 2, std::max(tol, 1e-5), true, xdev, true);
    // start vector of the FP64 run: sum of the FP32 Ritz vectors
    NSS_HIP(hipMemcpyAsync(v0.p, xdev, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if (neigen > 1) {
      // v0 += x_i: gemv_n computes w -= V h, so h = -1
      const double m1 = -1.0;
      DevArr<double> h1;
      h1.resize(ctx->pool, 1);
      NSS_HIP(hipMemcpyAsync(h1.p, &m1, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      for (int i = 1; i < neigen; ++i) launch_gemv_n(xdev + (size_t)i * n, n, 1, h1.p, v0.p, n, 0, ctx->stream);
      NSS_HIP(hipStreamSynchronize(ctx->stream));
    }
    out.evals = trlan(L, v0.p, neigen, maxiter, tol, false, xdev);
  } else {
    out.evals = trlan(L, v0.p, neigen, maxiter, tol, false, xdev);
  }
  out.numop = L.numop;
  out.numop_fp32 = L.numop32;
  return out;
}

// Dense path of sb_diag (m_sb <= LANC_DIM_THRESHOLD, src/DMRG_SUPERBLOCK.f90:438-467):
// H_sb is assembled on the device column by column (H e_j) and diagonalised.
LanczosOut dense_diag(nss_ctx* ctx, SbDevice* sb, int neigen, double* xdev) {
  const int64_t n = sb_n(sb);
  NSS_REQUIRE(n > 0 && n <= 8192, NSS_E_ARG, "dense_diag: dimension out of range");
  hipStream_t s = ctx->stream;
  DevArr<double> E, H;
  E.resize(ctx->pool, (size_t)n);
  H.resize(ctx->pool, (size_t)n * n);
  std::vector<double> hh((size_t)n * n);
  std::vector<double> unit(n, 0.0);
  for (int64_t j = 0; j < n; ++j) {
    unit.assign(n, 0.0);
    unit[j] = 1.0;
    NSS_HIP(hipMemcpyAsync(E.p, unit.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    sb_hv(ctx, sb, E.p, H.p + (size_t)j * n);
    NSS_HIP(hipStreamSynchronize(s));
  }
  NSS_HIP(hipMemcpyAsync(hh.data(), H.p, hh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  NSS_HIP(hipStreamSynchronize(s));
  // hh holds columns H e_j contiguously: hh[j*n + i] = H[i][j]; symmetrise
  std::vector<double> a((size_t)n * n);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) a[(size_t)i * n + j] = 0.5 * (hh[(size_t)j * n + i] + hh[(size_t)i * n + j]);
  std::vector<double> w(n);
  sym_eig((int)n, a.data(), w.data());
  std::vector<double> x((size_t)neigen * n);
  for (int i = 0; i < neigen; ++i)
    for (int64_t r = 0; r < n; ++r) x[(size_t)i * n + r] = a[(size_t)r * n + i];
  NSS_HIP(hipMemcpyAsync(xdev, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice, s));
  NSS_HIP(hipStreamSynchronize(s));
  LanczosOut out;
  out.evals.assign(w.begin(), w.begin() + neigen);
  out.numop = (int)n;
  return out;
}

}  // namespace nss

// GPU Lanczos eigen-solver for the superblock Hamiltonian: the role of
// SciFortran sp_eigh in sb_diag (src/DMRG_SUPERBLOCK.f90:376-437).
//
// Thick-restart Lanczos (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000))
// with full re-orthogonalisation (a classical Gram-Schmidt pass against the
// whole basis every iteration).  The
// Krylov basis lives in HBM.  An iteration is a fixed sequence of launches
// on the context stream (W and Y of H*v, Y also leaving the partial sums of
// alpha_j = v_j^T H v_j per wave of each output tile; one Gram-Schmidt pass that
// subtracts alpha_j v_j and beta_{j-1} v_{j-1} on the fly; one
// normalisation) that writes its projection coefficients and beta into a
// per-iteration slot in device memory; the host does not wait for it.  The
// host reads the slots back only at convergence checks, whose spacing
// follows the observed residual decay, and solves the small projected
// matrix in FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nss_internal.h"

namespace nss {

namespace {

struct Lz {
  nss_ctx* ctx;
  SbDevice* sb;
  int64_t n;
  int kmax;
  int S;                 // slot stride: h1[kmax+1], h2[kmax+1], beta
  DevArr<double> V;      // (kmax+1) x n
  DevArr<double> w;      // n
  DevArr<double> slots;  // kmax x S
  DevArr<double> part;   // Gram-Schmidt partials
  DevArr<double> sq;     // per-block sums of squares
  DevArr<double> scal;   // scratch scalars
  DevArr<double> adot;   // FP64 H*v: the Y launch's partials of alpha_j, one per wave of each tile
  int64_t nyt = 0;       // their count (4 x the Y plan's tiles)
  std::vector<double> T;  // (kmax+1)^2 host projected matrix
  int numop = 0, numop32 = 0;

  double* vec(int j) { return V.p + (size_t)j * n; }
  double& t(int i, int j) { return T[(size_t)i * (kmax + 1) + j]; }

  int restart_at = 0;  // first iteration after a thick restart (no three-term predecessor)

  // enqueue iteration j: w = H v_j; w -= beta_{j-1} v_{j-1} + alpha_j v_j;
  // one classical Gram-Schmidt pass against V[0..j]; v_{j+1} = w / ||w||.
  // Slot j: [0..j] Gram-Schmidt coefficients h, [kmax+1+j] alpha_j,
  // [kmax+1+j-1] beta_{j-1}, [2(kmax+1)] beta_j, so that the projected
  // matrix entries are h[i] + slot[kmax+1+i].
  void iterate(int j, bool f32) {
    hipStream_t s = ctx->stream;
    double* slot = slots.p + (size_t)j * S;
    const bool has_prev = j > restart_at;
    const double* bprev = has_prev ? slots.p + (size_t)(j - 1) * S + 2 * (kmax + 1) : nullptr;
    double* bout = has_prev ? slot + (kmax + 1) + (j - 1) : nullptr;
    if (!f32 && nyt > 0) {
      // alpha_j = v_j^T H v_j from the Y launch's per-tile partials; the
      // Gram-Schmidt pass subtracts alpha_j v_j and beta_{j-1} v_{j-1} on the fly
      sb_hv_dot(ctx, sb, vec(j), w.p, adot.p);
      ++numop;
      launch_orth_pass(V.p, n, j + 1, w.p, n, part.p, slot, sq.p, s, adot.p, vec(j), slot + (kmax + 1) + j, nyt,
                       has_prev ? vec(j - 1) : nullptr, bprev, bout);
    } else {
      if (f32) {
        sb_hv_fp32(ctx, sb, vec(j), w.p);
        ++numop32;
      } else {
        sb_hv(ctx, sb, vec(j), w.p);
        ++numop;
      }
      double* apart = part.p + (size_t)(kmax + 1) * lz_nch(n);  // alpha partials, past the pass's partials
      launch_lz_recurrence(w.p, has_prev ? vec(j - 1) : nullptr, vec(j), n, bprev, bout, apart, s);
      launch_orth_pass(V.p, n, j + 1, w.p, n, part.p, slot, sq.p, s, apart, vec(j), slot + (kmax + 1) + j);
    }
    launch_normalize(w.p, vec(j + 1), n, sq.p, slot + 2 * (kmax + 1), s);
  }
};

// Host-side split of a solve (NSS_PROFILE=1), without stream synchronisations
// of its own: enqueueing the iterations, waiting at the convergence checks,
// the host eigen-solves of the checks, the Ritz vectors at the end.
struct LzClock {
  nss_ctx* ctx;
  double t0 = 0.0;
  static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
  explicit LzClock(nss_ctx* c) : ctx(c), t0(c->prof ? now() : 0.0) {}
  void lap(const char* name) {
    if (!ctx->prof) return;
    const double t = now();
    ctx->timers[name] += t - t0;
    t0 = t;
  }
  void count(const char* name) {
    if (ctx->prof) ctx->timers[name] += 1;
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
  const int S = L.S;
  hipStream_t s = L.ctx->stream;
  L.T.assign((size_t)(kmax + 1) * (kmax + 1), 0.0);
  launch_norm(v0, n, L.part.p, L.scal.p, s);
  launch_div(v0, L.vec(0), n, L.scal.p, s);
  int k = 0;  // kept Ritz vectors (arrow size) after a restart
  int j = 0;  // next basis vector to expand
  L.restart_at = 0;
  NSS_HIP(hipMemsetAsync(L.slots.p, 0, (size_t)kmax * S * sizeof(double), s));
  int cycles = 0;
  std::vector<double> th, Y;
  const int nkeep = std::min(std::max(2 * neigen, neigen + 8), std::max(neigen, kmax / 2));
  nss_ctx* ctx = L.ctx;
  if (ctx->lz_pin_n < (size_t)kmax * S) {
    if (ctx->lz_pin) NSS_HIP(hipHostFree(ctx->lz_pin));
    ctx->lz_pin = nullptr;
    ctx->lz_pin_n = 0;
    NSS_HIP(hipHostMalloc((void**)&ctx->lz_pin, (size_t)kmax * S * sizeof(double), hipHostMallocDefault));
    ctx->lz_pin_n = (size_t)kmax * S;
  }
  if (!ctx->lz_ev) NSS_HIP(hipEventCreateWithFlags(&ctx->lz_ev, hipEventDisableTiming));
  LzClock clk(L.ctx);
  for (;;) {
    bool converged = false, breakdown = false;
    int m = 0;
    double beta = 0.0;
    int read_upto = j;     // slots [read_upto, j) not yet on the host
    int next_check = j + 1;
    double last_res = -1.0;
    int last_res_m = 0;
    bool ahead = false;  // iteration j already enqueued behind the last check's copy
    while (j < kmax) {
      if (!ahead) L.iterate(j, f32);
      ahead = false;
      ++j;
      const bool at_end = (j == kmax) || (j == n);
      if (j < next_check && !at_end) continue;
      // copy the slots of iterations read_upto .. j-1 to pinned memory, then
      // enqueue iteration j before waiting on the copy, so that the device
      // keeps working while the host runs the check (if the check ends the
      // run, that iteration's output is left unused: it writes only slot j,
      // w and basis vector j+1)
      const int cnt = j - read_upto;
      clk.lap("lz.enqueue");
      double* hs = ctx->lz_pin;
      NSS_HIP(hipMemcpyAsync(hs, L.slots.p + (size_t)read_upto * S, (size_t)cnt * S * sizeof(double),
                             hipMemcpyDeviceToHost, s));
      NSS_HIP(hipEventRecord(ctx->lz_ev, s));
      if (!at_end && lz_check_ahead()) {
        L.iterate(j, f32);
        ahead = true;
      }
      clk.lap("lz.enqueue");
      NSS_HIP(hipEventSynchronize(ctx->lz_ev));
      clk.lap("lz.wait");
      clk.count("lz.checks");
      double hn = 0.0;
      for (int it = read_upto; it < j; ++it) {
        const double* sl = hs + (size_t)(it - read_upto) * S;
        for (int i = 0; i <= it; ++i) {
          const bool keep = (i == it) || (i == it - 1) || (it == k && i < k);
          const double v = keep ? sl[i] + sl[(kmax + 1) + i] : 0.0;
          L.t(i, it) = v;
          L.t(it, i) = v;
        }
        const double b = sl[2 * (kmax + 1)];
        L.t(it + 1, it) = L.t(it, it + 1) = b;
        for (int i = 0; i <= it; ++i) hn = std::max(hn, std::fabs(L.t(i, i)));
        m = it + 1;
        beta = b;
        if (b <= 1e-14 * std::max(1.0, hn) || !std::isfinite(b)) {  // invariant subspace
          breakdown = true;
          break;
        }
      }
      read_upto = j;
      if (breakdown) {  // later iterations used a null vector: discard them
        j = m;
        ahead = false;
      }
      if (m < neigen) {
        if (breakdown) throw Error(NSS_E_NOCONV, "lanczos: invariant Krylov space smaller than neigen");
        next_check = std::max(neigen, j + 1);
        continue;
      }
      th.assign(m, 0.0);
      Y.assign((size_t)m * m, 0.0);
      if (k == 0) {
        std::vector<double> d(m), e(m, 0.0);
        for (int i = 0; i < m; ++i) d[i] = L.t(i, i);
        for (int i = 0; i + 1 < m; ++i) e[i] = L.t(i + 1, i);
        if (neigen == 1) {  // the test needs only the lowest pair's last component
          tridiag_lowest(m, d.data(), e.data(), &th[0], &Y[(size_t)(m - 1) * m]);
        } else {
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
      double worst = 0.0;  // max over pairs of residual / tolerance
      for (int i = 0; i < neigen; ++i) {
        const double res = std::fabs(beta * Y[(size_t)(m - 1) * m + i]);
        const double lim = tol * std::max(1.0, std::fabs(th[i]));
        worst = std::max(worst, res / lim);
        if (!(res <= lim)) converged = false;
      }
      if (breakdown) converged = true;
      if (converged || at_end || m == n) {
        if (k == 0) {
          std::vector<double> d(m), e(m, 0.0);
          for (int i = 0; i < m; ++i) d[i] = L.t(i, i);
          for (int i = 0; i + 1 < m; ++i) e[i] = L.t(i + 1, i);
          if (neigen == 1 && (converged || m == n || (soft && cycles >= maxiter))) {
            // only the lowest Ritz vector is needed: its inverse iteration
            std::vector<double> z(m);
            double zl = 0.0;
            tridiag_lowest(m, d.data(), e.data(), &th[0], &zl, z.data());
            for (int a = 0; a < m; ++a) Y[(size_t)a * m] = z[a];
          } else {  // full eigenvectors of the tridiagonal (also for a thick restart)
            tridiag_eig_full(m, d.data(), e.data(), th.data(), Y.data());
          }
        }
        break;
      }
      // schedule the next check from the observed decay of the residual
      int step = 1;
      if (last_res > 0.0 && worst < last_res && m > last_res_m) {
        const double rate = std::pow(worst / last_res, 1.0 / (m - last_res_m));
        if (rate < 1.0) {
          const double need = std::log(worst) / -std::log(rate);
          step = (int)std::clamp(need * 0.8, 1.0, 8.0);
        }
      } else if (m < 8) {
        step = 1;
      } else {
        step = 4;
      }
      last_res = worst;
      last_res_m = m;
      next_check = std::min(kmax, j + step);
      clk.lap("lz.host");
    }
    clk.lap("lz.host");
    if (m == 0) throw Error(NSS_E_STATE, "lanczos: empty Krylov space");
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
        launch_norm(x + (size_t)i * n, n, L.part.p, L.scal.p, s);
        launch_div(x + (size_t)i * n, x + (size_t)i * n, n, L.scal.p, s);
      }
      NSS_HIP(hipStreamSynchronize(s));
      clk.lap("lz.ritz");
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
    // residual direction v_m (already normalised in slot m) moves to position kk
    NSS_HIP(hipMemcpyAsync(L.vec(kk), L.vec(m), n * sizeof(double), hipMemcpyDeviceToDevice, s));
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
    L.restart_at = kk;
    // slots of the new iterations must not carry stale recurrence terms
    NSS_HIP(hipMemsetAsync(L.slots.p + (size_t)kk * S, 0, (size_t)(kmax - kk) * S * sizeof(double), s));
    clk.lap("lz.restart");
  }
}

double lz_mixed_inner() {
  static const double v = [] {
    const char* e = std::getenv("NSS_MIXED_INNER");
    const double x = e ? std::atof(e) : 1e-5;
    return (x > 0.0 && x < 1.0) ? x : 1e-5;
  }();
  return v;
}

// Ritz values within this fraction of |theta_1| of the lowest join the FP64
// block of the next cycle (NSS_MIXED_BLOCK_WIN, default 1e-4).
static double lz_mixed_window() {
  static const double v = [] {
    const char* e = std::getenv("NSS_MIXED_BLOCK_WIN");
    const double x = e ? std::atof(e) : 1e-4;
    return x >= 0.0 ? x : 1e-4;
  }();
  return v;
}

// Rounding level of the FP32-MFMA H*v for the orthogonality estimates of
// the mixed solver's correction space (NSS_MIXED_PRO_EPS, default the FP32
// machine epsilon 2^-23); its full Gram-Schmidt pass runs above sqrt of it.
static double lz_mixed_eps() {
  static const double v = [] {
    const char* e = std::getenv("NSS_MIXED_PRO_EPS");
    const double x = e ? std::atof(e) : 0.0;
    return (x > 0.0 && x < 1e-2) ? x : 1.1920928955078125e-07;
  }();
  return v;
}

// Largest correction space of one cycle of the mixed solver
// (NSS_MIXED_MAXM, read per solve; default 64): a cycle that reaches it ends
// with the Rayleigh-Ritz of what it has, and the next cycle starts from that
// (longer FP32 spaces grown from random starts degraded into Ritz values far
// outside the spectrum and forced FP64 hand-overs).
static int lz_mixed_maxm() {
  const char* e = std::getenv("NSS_MIXED_MAXM");
  const int v = e ? std::atoi(e) : 0;
  return v >= 8 ? v : 64;
}

// NSS_MIXED_FULL_REORTH=1: the correction space's Gram-Schmidt pass in every
// iteration instead of partial reorthogonalisation.
static int lz_mixed_full_reorth() {
  static const int v = [] {
    const char* e = std::getenv("NSS_MIXED_FULL_REORTH");
    return e && e[0] == '1' ? 1 : 0;
  }();
  return v;
}

// Lowest eigenpair in mixed precision: FP32-MFMA H*v for the bulk of the
// Krylov work, FP64 for everything that decides the answer.  Each cycle:
//  1. FP64 block step: W = H X for the block X of p <= 3 orthonormal vectors
//     (V[0..p)), the exact Rayleigh-Ritz G = X^T W = Q Theta Q^T, X <- X Q,
//     W <- W Q and the exact residuals R = W - X Theta.  The run ends here
//     when |R_1| <= tol max(1, |theta_1|), the FP64 criterion of the
//     thick-restart solver.
//  2. A Lanczos space V[p..) built with FP32-MFMA H*v from R_1 (FP64
//     vectors and recurrence; every H v projected out of X, then the
//     partial-reorthogonalisation iteration of the batched solver, its
//     estimates at the FP32 rounding level lz_mixed_eps(), so that the
//     pass against V runs only when semi-orthogonality at sqrt of that level
//     is at risk), together with the exact couplings C = R^T V (X^T H v =
//     R^T v for v orthogonal to X).
//  3. Rayleigh-Ritz of H over [X V]: M = [[Theta, C], [C^T, T]], with the
//     FP32 rounding confined to the tridiagonal T of the correction space.
//     The cycle ends when the lowest pair's residual estimate has fallen by
//     lz_mixed_inner() below |R_1|, or at lz_mixed_maxm() vectors; the Ritz
//     vectors of M whose values lie within lz_mixed_window() |theta_1| of the
//     lowest become the next block, orthonormalised in FP64.
// A near-degenerate partner (a state closer to the ground state than the
// FP32 rounding of H*v, which an FP32 Krylov space mixes with it at O(1))
// thereby enters the block and is separated exactly by the FP64 Rayleigh-Ritz.
// A cycle whose |R_1| does not fall below a quarter of the previous one and
// whose theta_1 does not fall by more than 1% of the previous |R_1| (with an
// unchanged block size) hands over to the FP64 thick-restart solver from X_1.
// NSS_MIXED_TRACE=1: one line per cycle of the mixed solver on stderr.
static bool lz_mixed_trace() {
  static const bool v = [] {
    const char* e = std::getenv("NSS_MIXED_TRACE");
    return e && e[0] == '1';
  }();
  return v;
}

static double lanczos_mixed_block(Lz& L, const double* v0, int maxiter, double tol, double* x) {
  nss_ctx* ctx = L.ctx;
  SbDevice* sb = L.sb;
  const int64_t n = L.n;
  const int kmax = L.kmax, S = L.S;
  hipStream_t s = ctx->stream;
  constexpr int PMAX = 3;
  if (kmax < PMAX + 4) return trlan(L, v0, 1, maxiter, tol, false, x)[0];
  const int64_t gch = gemv_t_chunks(n);
  DevArr<double> W, R, Tmp, Cd, Gd, hd, gpart;
  W.resize(ctx->pool, (size_t)PMAX * n);
  R.resize(ctx->pool, (size_t)PMAX * n);
  Tmp.resize(ctx->pool, (size_t)PMAX * n);
  Cd.resize(ctx->pool, (size_t)PMAX * (kmax + 1));
  Gd.resize(ctx->pool, (size_t)PMAX * PMAX);
  hd.resize(ctx->pool, (size_t)(kmax + 1) * PMAX);
  gpart.resize(ctx->pool, (size_t)PMAX * gch + 1);
  launch_norm(v0, n, L.part.p, L.scal.p, s);
  launch_div(v0, L.vec(0), n, L.scal.p, s);
  const double inner = lz_mixed_inner(), window = lz_mixed_window();
  // the correction space's chain: V[p..), orthogonality estimates at the
  // FP32-MFMA rounding level
  const int kb1 = kmax + 1;
  const int64_t nchk = lz_nch(n);
  const int full = lz_mixed_full_reorth();
  const int maxm = lz_mixed_maxm();
  DevArr<double> cpart, om, cbp, hx;
  DevArr<int32_t> cst, dmap;
  DevArr<LzChainDev> dch;
  cpart.resize(ctx->pool, (size_t)(kb1 + 2) * nchk);
  om.resize(ctx->pool, (size_t)3 * (kb1 + 1));
  cbp.resize(ctx->pool, 1);
  hx.resize(ctx->pool, PMAX);
  cst.resize(ctx->pool, 2);
  dmap.resize(ctx->pool, nchk);
  dch.resize(ctx->pool, 1);
  std::vector<double> om0((size_t)3 * (kb1 + 1), 0.0);
  om0[0] = 1.0;
  LzChainDev ch{};
  ch.w = L.w.p;
  ch.x = nullptr;
  ch.slots = L.slots.p;
  ch.part = cpart.p;
  ch.om = om.p;
  ch.st = cst.p;
  ch.bpre = cbp.p;
  ch.n = n;
  ch.sstride = S;
  ch.eps = lz_mixed_eps();
  // the new vector's loss against the previous one: an eighth of the
  // semi-orthogonality threshold (NSS_MIXED_PSI overrides).  alpha comes from
  // the same w it is subtracted from, so the FP64 path's eps sqrt(n) (the
  // rounding of length-n dot products) does not model it: at 12.8M states it
  // exceeds sqrt(eps) and forced a full pass every iteration; one FP32
  // rounding unit let spaces grown from random starts degrade (measured on
  // the spin-1 m=1024 and periodic-ring workloads)
  ch.psi = std::sqrt(ch.eps) / 8.0;
  if (const char* e = std::getenv("NSS_MIXED_PSI")) {
    const double v = std::atof(e);
    if (v > 0.0 && v < 1.0) ch.psi = v;
  }
  ch.nch = (int32_t)nchk;
  ch.blk0 = 0;
  ch.active = 1;
  NSS_HIP(hipMemsetAsync(dmap.p, 0, (size_t)nchk * sizeof(int32_t), s));
  int p = 1, prev_p = 0;
  double prev_res = -1.0, prev_theta = 0.0;
  std::vector<double> G, th, nrm(PMAX), sl, cc, M, ev;
  for (int cycle = 0; cycle < std::max(8, maxiter); ++cycle) {
    // 1. FP64 block step
    for (int i = 0; i < p; ++i) {
      sb_hv(ctx, sb, L.vec(i), W.p + (size_t)i * n);
      ++L.numop;
    }
    for (int j = 0; j < p; ++j) launch_gemv_t(L.V.p, n, p, W.p + (size_t)j * n, n, gpart.p, Gd.p + j * p, s);
    G.resize((size_t)p * p);
    NSS_HIP(hipMemcpyAsync(G.data(), Gd.p, G.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    NSS_HIP(hipStreamSynchronize(s));
    for (int a = 0; a < p; ++a)  // G[j*p + i] = X_i . W_j; symmetrise
      for (int b = 0; b < a; ++b) G[(size_t)a * p + b] = G[(size_t)b * p + a] = 0.5 * (G[(size_t)a * p + b] + G[(size_t)b * p + a]);
    th.assign(p, 0.0);
    sym_eig(p, G.data(), th.data());  // eigenvectors as columns
    if (p > 1) {
      std::vector<double> q((size_t)p * p);
      for (int i = 0; i < p; ++i)
        for (int k = 0; k < p; ++k) q[(size_t)i * p + k] = G[(size_t)k * p + i];  // column i contiguous
      NSS_HIP(hipMemcpyAsync(hd.p, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice, s));
      for (int pass = 0; pass < 2; ++pass) {  // X <- X Q, then W <- W Q
        double* base = pass == 0 ? L.V.p : W.p;
        for (int i = 0; i < p; ++i) launch_gemv_n(base, n, p, hd.p + (size_t)i * p, Tmp.p + (size_t)i * n, n, 1, s);
        NSS_HIP(hipMemcpyAsync(base, Tmp.p, (size_t)p * n * sizeof(double), hipMemcpyDeviceToDevice, s));
      }
    }
    // R_i = W_i - theta_i X_i and |R_i|
    NSS_HIP(hipMemcpyAsync(R.p, W.p, (size_t)p * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    NSS_HIP(hipMemcpyAsync(hd.p + (size_t)PMAX * PMAX, th.data(), p * sizeof(double), hipMemcpyHostToDevice, s));
    for (int i = 0; i < p; ++i)
      launch_gemv_n(L.vec(i), n, 1, hd.p + (size_t)PMAX * PMAX + i, R.p + (size_t)i * n, n, 0, s);
    launch_norm(R.p, n, L.part.p, L.scal.p, s);
    double res1 = 0.0;
    NSS_HIP(hipMemcpyAsync(&res1, L.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    NSS_HIP(hipStreamSynchronize(s));
    const double lim = tol * std::max(1.0, std::fabs(th[0]));
    if (lz_mixed_trace())
      std::fprintf(stderr, "mixed n %lld cycle %d p %d theta %.15g res %.3e\n", (long long)n, cycle, p, th[0], res1);
    if (res1 <= lim) {
      NSS_HIP(hipMemcpyAsync(x, L.vec(0), n * sizeof(double), hipMemcpyDeviceToDevice, s));
      NSS_HIP(hipStreamSynchronize(s));
      return th[0];
    }
    // no progress from the FP32 space: FP64 from here.  Progress is a
    // residual down by 4, or (early cycles from a random start, whose
    // residual need not fall yet) theta down by more than 1% of the previous
    // residual; a cycle that grew the block has only just separated a
    // near-degenerate pair, and is exempt.
    if (prev_res > 0.0 && p == prev_p && res1 > 0.25 * prev_res && !(prev_theta - th[0] > 0.01 * prev_res)) break;
    prev_res = res1;
    prev_theta = th[0];
    prev_p = p;
    // 2. FP32 Lanczos space from R_1, orthogonal to X
    NSS_HIP(hipMemcpyAsync(L.w.p, R.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    launch_orth_pass(L.V.p, n, p, L.w.p, n, L.part.p, hd.p, L.sq.p, s);
    launch_normalize(L.w.p, L.vec(p), n, L.sq.p, L.scal.p + 1, s);
    launch_gemv_t(R.p, n, p, L.vec(p), n, gpart.p, Cd.p, s);  // C[k*p + i] = R_i . v_{p+k}
    NSS_HIP(hipMemsetAsync(L.slots.p, 0, (size_t)kmax * S * sizeof(double), s));
    // the correction space V[p..) as one chain of the partial-reorthogonalisation
    // kernels (inner iteration k in slot k), its estimates restarted
    ch.V = L.vec(p);
    ch.kmax = kmax - p;
    NSS_HIP(hipMemcpyAsync(dch.p, &ch, sizeof(LzChainDev), hipMemcpyHostToDevice, s));
    NSS_HIP(hipMemcpyAsync(om.p, om0.data(), om0.size() * sizeof(double), hipMemcpyHostToDevice, s));
    NSS_HIP(hipMemsetAsync(cst.p, 0, 2 * sizeof(int32_t), s));
    const double target = std::max(0.5 * lim, inner * res1);
    std::vector<double> dV(kmax + 1, 0.0), eV(kmax + 1, 0.0), bV(kmax + 1, 0.0);
    // checks (host round trips) at jj = 4, then where the residual estimate's
    // geometric decay since the previous check predicts the target
    int jj = 0, read_upto = 0, next_check = std::min(4, kmax - p), m = 0, m_prev = 0;
    double est_prev = 0.0;
    bool breakdown = false;
    for (;;) {
      sb_hv_fp32(ctx, sb, L.vec(p + jj), L.w.p);
      ++L.numop32;
      launch_orth_pass(L.V.p, n, p, L.w.p, n, L.part.p, hx.p, nullptr, s);  // w orthogonal to X
      launch_lzb_iteration(dch.p, dmap.p, (int)nchk, 1, jj, kb1, full, s);
      ++jj;
      const int j = p + jj;
      launch_gemv_t(R.p, n, p, L.vec(j), n, gpart.p, Cd.p + (size_t)jj * p, s);
      const bool at_end = j == kmax || j == n || jj >= maxm;
      if (jj < next_check && !at_end) continue;
      const int cnt = jj - read_upto;
      sl.resize((size_t)cnt * S);
      NSS_HIP(hipMemcpyAsync(sl.data(), L.slots.p + (size_t)read_upto * S, sl.size() * sizeof(double),
                             hipMemcpyDeviceToHost, s));
      cc.resize((size_t)(jj + 1) * p);
      NSS_HIP(hipMemcpyAsync(cc.data(), Cd.p, cc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      NSS_HIP(hipStreamSynchronize(s));
      double hn = std::fabs(th[0]);
      for (int k = read_upto; k < jj; ++k) {
        const double* q = sl.data() + (size_t)(k - read_upto) * S;
        dV[k] = q[k] + q[kb1 + k];
        if (k > 0) eV[k - 1] = q[k - 1] + q[kb1 + k - 1];
        bV[k] = q[2 * kb1];
        hn = std::max(hn, std::fabs(dV[k]));
        m = k + 1;
        if (bV[k] <= 1e-14 * std::max(1.0, hn) || !std::isfinite(bV[k])) {
          breakdown = true;
          break;
        }
      }
      read_upto = jj;
      // 3. Rayleigh-Ritz over [X, V[p..p+m)]
      const int D = p + m;
      M.assign((size_t)D * D, 0.0);
      for (int i = 0; i < p; ++i) M[(size_t)i * D + i] = th[i];
      for (int k = 0; k < m; ++k) {
        for (int i = 0; i < p; ++i) M[(size_t)i * D + p + k] = M[(size_t)(p + k) * D + i] = cc[(size_t)k * p + i];
        M[(size_t)(p + k) * D + p + k] = dV[k];
        if (k + 1 < m) M[(size_t)(p + k) * D + p + k + 1] = M[(size_t)(p + k + 1) * D + p + k] = eV[k];
      }
      ev.assign(D, 0.0);
      sym_eig(D, M.data(), ev.data());
      const double est = std::fabs(bV[m - 1] * M[(size_t)(D - 1) * D + 0]);
      if (!(est <= target || at_end || breakdown)) {
        int step = 4;
        if (m_prev > 0 && est > 0.0 && est < est_prev) {
          const double rate = std::log(est / est_prev) / (m - m_prev);  // < 0
          step = (int)std::clamp(std::ceil(1.1 * std::log(target / est) / rate), 1.0, 16.0);
        }
        m_prev = m;
        est_prev = est;
        next_check = std::min(kmax - p, jj + step);
        continue;
      }
      // the next block: the lowest Ritz vector and those within the window
      int np = 1;
      while (np < PMAX && np < D && ev[np] - ev[0] <= window * std::max(1.0, std::fabs(ev[0]))) ++np;
      if (lz_mixed_trace())
        std::fprintf(stderr, "mixed   fp32 m %d est %.3e target %.3e%s%s gaps %.3e %.3e next p %d\n", m, est, target,
                     at_end ? " at_end" : "", breakdown ? " breakdown" : "", D > 1 ? ev[1] - ev[0] : 0.0,
                     D > 2 ? ev[2] - ev[0] : 0.0, np);
      std::vector<double> y((size_t)np * D);
      for (int i = 0; i < np; ++i)
        for (int a = 0; a < D; ++a) y[(size_t)i * D + a] = M[(size_t)a * D + i];
      NSS_HIP(hipMemcpyAsync(hd.p, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, s));
      for (int i = 0; i < np; ++i) launch_gemv_n(L.V.p, n, D, hd.p + (size_t)i * D, Tmp.p + (size_t)i * n, n, 1, s);
      NSS_HIP(hipMemcpyAsync(L.V.p, Tmp.p, (size_t)np * n * sizeof(double), hipMemcpyDeviceToDevice, s));
      launch_norm(L.vec(0), n, L.part.p, L.scal.p, s);
      launch_div(L.vec(0), L.vec(0), n, L.scal.p, s);
      // the block orthonormalised in FP64; a vector mostly inside the span of
      // the ones before it (a copy of a converged Ritz vector that the FP32
      // space's rounding left in it, not a partner state) ends the block
      for (int i = 1; i < np; ++i) {
        launch_norm(L.vec(i), n, L.part.p, L.scal.p, s);
        launch_div(L.vec(i), L.vec(i), n, L.scal.p, s);
        launch_orth_pass(L.V.p, n, i, L.vec(i), n, L.part.p, hx.p, L.sq.p, s);
        launch_normalize(L.vec(i), L.vec(i), n, L.sq.p, L.scal.p + 2, s);
        double rest = 0.0;
        NSS_HIP(hipMemcpyAsync(&rest, L.scal.p + 2, sizeof(double), hipMemcpyDeviceToHost, s));
        NSS_HIP(hipStreamSynchronize(s));
        if (!(rest >= 0.1)) {
          if (lz_mixed_trace()) std::fprintf(stderr, "mixed   block vector %d dropped (%.3e outside the rest)\n", i, rest);
          np = i;
          break;
        }
      }
      NSS_HIP(hipStreamSynchronize(s));  // y lives on the host
      p = np;
      break;
    }
  }
  // hand-over: FP64 thick restart from the current lowest block vector
  if (lz_mixed_trace()) std::fprintf(stderr, "mixed n %lld: FP64 hand-over\n", (long long)n);
  return trlan(L, L.vec(0), 1, maxiter, tol, false, x)[0];
}

LanczosOut lanczos_device(nss_ctx* ctx, SbDevice* sb, int neigen, int kmax, int maxiter, double tol,
                          uint64_t seed, int v0_mode, int precision, double* xdev) {
  LzClock clk(ctx);
  const int64_t n = sb_n(sb);
  NSS_REQUIRE(n > 0, NSS_E_STATE, "lanczos: empty superblock");
  NSS_REQUIRE(neigen >= 1 && neigen <= 4 && neigen <= n, NSS_E_ARG, "lanczos: 1 <= neigen <= min(4,n)");
  if (kmax <= 0) kmax = 150;
  kmax = std::min(kmax, LZ_MAX_BASIS - 1);  // the Gram-Schmidt pass takes kmax + 1 vectors
  kmax = (int)std::min<int64_t>(kmax, n);
  kmax = std::max(kmax, std::min<int>((int)n, neigen + 2));
  if (maxiter <= 0) maxiter = 512;
  Lz L;
  L.ctx = ctx;
  L.sb = sb;
  L.n = n;
  L.kmax = kmax;
  L.S = 2 * (kmax + 1) + 1;
  L.V.resize(ctx->pool, (size_t)(kmax + 1) * n);
  L.w.resize(ctx->pool, n);
  L.slots.resize(ctx->pool, (size_t)kmax * L.S);
  NSS_HIP(hipMemsetAsync(L.slots.p, 0, (size_t)kmax * L.S * sizeof(double), ctx->stream));
  const int64_t nch = (n + orth_chunk(n) - 1) / orth_chunk(n);
  L.part.resize(ctx->pool, std::max<int64_t>((int64_t)(kmax + 2) * nch, gemv_t_chunks(n)) + 1);
  L.sq.resize(ctx->pool, orth_blocks(n) + 1);
  L.scal.resize(ctx->pool, 4);
  L.nyt = lz_alpha_from_y() ? 4 * sb_hv_ntiles(sb) : 0;
  L.adot.resize(ctx->pool, std::max<int64_t>(1, L.nyt));
  DevArr<double> v0;
  v0.resize(ctx->pool, n);
  bool filled = false;
  if (v0_mode == 2) {
    // caller-supplied start vector in xdev (a predicted wavefunction); a
    // vanishing or non-finite one falls back to the seeded random start
    NSS_HIP(hipMemcpyAsync(v0.p, xdev, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    launch_norm(v0.p, n, L.part.p, L.scal.p, ctx->stream);
    double nrm = 0.0;
    NSS_HIP(hipMemcpyAsync(&nrm, L.scal.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    filled = std::isfinite(nrm) && nrm > 1e-8;
  }
  if (!filled) {
    if (v0_mode == 1) launch_fill_const(v0.p, n, 1.0 / std::sqrt((double)n), ctx->stream);
    else launch_fill_splitmix(v0.p, n, seed, ctx->stream);
  }
  clk.lap("lz.setup");
  LanczosOut out;
  if (precision == NSS_FP32_MIXED && neigen == 1) {
    out.evals.assign(1, lanczos_mixed_block(L, v0.p, maxiter, tol, xdev));
  } else if (precision == NSS_FP32_MIXED) {
    // several eigenpairs: FP32-MFMA H*v to an FP32-attainable residual, then the FP64 restart
    trlan(L, v0.p, neigen, 2, std::max(tol, 1e-5), true, xdev, true);
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

// alpha_j from the Y launch's partials (NSS_LZ_ALPHA_FROM_Y=0: a separate
// dot-product pass over H v_j, as the FP32 iterations do)
bool lz_alpha_from_y() {
  static const bool v = [] {
    const char* e = std::getenv("NSS_LZ_ALPHA_FROM_Y");
    return !e || std::atoi(e) != 0;
  }();
  return v;
}

// Enqueue the next iteration before waiting on a convergence check
// (NSS_LZ_CHECK_AHEAD=0: wait first).
bool lz_check_ahead() {
  static const bool v = [] {
    const char* e = std::getenv("NSS_LZ_CHECK_AHEAD");
    return !e || std::atoi(e) != 0;
  }();
  return v;
}

int64_t lz_pro_min_n() {
  static const int64_t v = [] {
    const char* e = std::getenv("NSS_LZ_PRO_MIN_N");
    return e ? (int64_t)std::max(0LL, std::atoll(e)) : (int64_t)1 << 17;
  }();
  return v;
}

int64_t lz_mixed_min_n() {
  const char* e = std::getenv("NSS_MIXED_MIN_N");
  return e ? (int64_t)std::max(0LL, std::atoll(e)) : (int64_t)1 << 17;
}

// Full Gram-Schmidt pass in every batched iteration (NSS_LZB_FULL_REORTH=1)
// instead of partial reorthogonalisation.
static int lzb_full_reorth() {
  static const int v = [] {
    const char* e = std::getenv("NSS_LZB_FULL_REORTH");
    return e && e[0] == '1' ? 1 : 0;
  }();
  return v;
}

// Lowest eigenpair of several independent superblocks at once: Lanczos
// with partial reorthogonalisation (the classical Gram-Schmidt pass against
// the whole basis only when Simon's estimate of the loss of orthogonality
// exceeds sqrt(eps), decided on the device), every chain in lockstep.  Each
// iteration is one merged W launch and one merged Y launch over all active
// chains (their vectors in shared X / Y arenas) and five batched vector
// launches; the host reads every chain's new slots in one copy at the
// checks (iterations 1..8, then every 4th).
//
// A chain leaves the batch when its residual |beta_{m-1} y_{m-1}| meets
// tol max(1, |theta|) (the single-chain criterion), at an invariant
// subspace, or at its kmax; one that is still unconverged there is finished
// by the single-chain thick-restart solver from its current Ritz vector.
void lanczos_batch(nss_ctx* ctx, std::vector<LanczosJob>& jobs, int kmax_req, int maxiter, double tol) {
  const int K = (int)jobs.size();
  if (K == 0) return;
  hipStream_t s = ctx->stream;
  if (kmax_req <= 0) kmax_req = 150;
  std::vector<int64_t> n(K), xoff(K), woff(K), voff(K), poff(K);
  std::vector<int> kmc(K), nch(K), blk0(K);
  std::vector<SbDevice*> sbs(K);
  int64_t xtot = 0, wtot = 0, vtot = 0, ptot = 0;
  int btot = 0, kb = 0;
  for (int c = 0; c < K; ++c) {
    sbs[c] = jobs[c].sb;
    n[c] = sb_n(sbs[c]);
    NSS_REQUIRE(n[c] > 0, NSS_E_STATE, "lanczos_batch: empty superblock");
    kmc[c] = (int)std::min<int64_t>(kmax_req, n[c]);
    kmc[c] = std::max(kmc[c], (int)std::min<int64_t>(n[c], 3));
    kb = std::max(kb, kmc[c]);
  }
  const int kb1 = kb + 1, S = 2 * kb1 + 1;
  for (int c = 0; c < K; ++c) {
    nch[c] = (int)lz_nch(n[c]);
    xoff[c] = xtot;
    xtot += n[c];
    woff[c] = wtot;
    wtot += std::max<int64_t>(1, sb_wsize(sbs[c]));
    voff[c] = vtot;
    vtot += (int64_t)(kmc[c] + 1) * n[c];
    poff[c] = ptot;
    ptot += (int64_t)(kb1 + 2) * nch[c];
    blk0[c] = btot;
    btot += nch[c];
  }
  DevArr<double> X, Y, Wa, V, P, SL, scal, yd, OM, BP;
  DevArr<int32_t> ST;
  X.resize(ctx->pool, xtot);
  Y.resize(ctx->pool, xtot);
  Wa.resize(ctx->pool, wtot);
  V.resize(ctx->pool, vtot);
  P.resize(ctx->pool, ptot);
  SL.resize(ctx->pool, (size_t)kb * K * S);
  scal.resize(ctx->pool, K);
  yd.resize(ctx->pool, (size_t)K * kb1);  // a slice per chain
  NSS_HIP(hipMemsetAsync(SL.p, 0, (size_t)kb * K * S * sizeof(double), s));
  // orthogonality estimates: omega_{0,0} = 1, everything else 0
  OM.resize(ctx->pool, (size_t)K * 3 * kb1);
  BP.resize(ctx->pool, K);
  ST.resize(ctx->pool, 2 * K);
  {
    std::vector<double> om((size_t)K * 3 * kb1, 0.0);
    for (int c = 0; c < K; ++c) om[(size_t)c * 3 * kb1] = 1.0;
    NSS_HIP(hipMemcpyAsync(OM.p, om.data(), om.size() * sizeof(double), hipMemcpyHostToDevice, s));
    NSS_HIP(hipMemsetAsync(ST.p, 0, 2 * K * sizeof(int32_t), s));
    NSS_HIP(hipStreamSynchronize(s));
  }
  const int full = lzb_full_reorth();
  std::vector<LzChainDev> hc(K);
  std::vector<int32_t> bmap(btot);
  for (int c = 0; c < K; ++c) {
    LzChainDev& C = hc[c];
    C.V = V.p + voff[c];
    C.w = Y.p + xoff[c];
    C.x = X.p + xoff[c];
    C.slots = SL.p + (size_t)c * S;
    C.part = P.p + poff[c];
    C.om = OM.p + (size_t)c * 3 * kb1;
    C.st = ST.p + 2 * c;
    C.bpre = BP.p + c;
    C.n = n[c];
    C.sstride = (int64_t)K * S;
    C.eps = 2.220446049250313e-16;  // FP64 H v
    C.psi = C.eps * std::sqrt((double)n[c]);  // rounding of the length-n dot products
    C.nch = nch[c];
    C.blk0 = blk0[c];
    C.active = 1;
    C.kmax = kmc[c];
    for (int b = 0; b < nch[c]; ++b) bmap[blk0[c] + b] = c;
  }
  DevArr<LzChainDev> dch;
  DevArr<int32_t> dmap;
  dch.resize(ctx->pool, K);
  dmap.resize(ctx->pool, btot);
  NSS_HIP(hipMemcpyAsync(dch.p, hc.data(), K * sizeof(LzChainDev), hipMemcpyHostToDevice, s));
  NSS_HIP(hipMemcpyAsync(dmap.p, bmap.data(), btot * sizeof(int32_t), hipMemcpyHostToDevice, s));
  // start vectors: the caller's (v0_mode 2, unless vanishing or non-finite),
  // constant (1) or SplitMix64 (0); normalised into V_c[0] and X_c
  for (int c = 0; c < K; ++c) {
    LanczosJob& J = jobs[c];
    double* v0 = hc[c].V;
    bool filled = false;
    if (J.v0_mode == 2) {
      NSS_HIP(hipMemcpyAsync(v0, J.xdev, n[c] * sizeof(double), hipMemcpyDeviceToDevice, s));
      launch_norm(v0, n[c], hc[c].part, scal.p + c, s);
      double nrm = 0.0;
      NSS_HIP(hipMemcpyAsync(&nrm, scal.p + c, sizeof(double), hipMemcpyDeviceToHost, s));
      NSS_HIP(hipStreamSynchronize(s));
      filled = std::isfinite(nrm) && nrm > 1e-8;
    }
    if (!filled) {
      if (J.v0_mode == 1) launch_fill_const(v0, n[c], 1.0 / std::sqrt((double)n[c]), s);
      else launch_fill_splitmix(v0, n[c], J.seed, s);
    }
    launch_norm(v0, n[c], hc[c].part, scal.p + c, s);
    launch_div(v0, v0, n[c], scal.p + c, s);
    NSS_HIP(hipMemcpyAsync(hc[c].x, v0, n[c] * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  LzClock clk(ctx);
  GemmPlan Wp, Yp;
  sb_merge_hv_plans(sbs, xoff, woff, Wp, Yp, ctx, s);
  clk.lap("lzb.merge");
  // host side of every chain: the projected tridiagonal
  std::vector<std::vector<double>> d(K, std::vector<double>(kb1, 0.0)), e(K, std::vector<double>(kb1, 0.0)),
      beta(K, std::vector<double>(kb1, 0.0));
  std::vector<double> hn(K, 0.0);
  std::vector<std::vector<double>> ybuf(K);  // each chain's Ritz coefficients, alive until the end
  std::vector<int> numop(K, 0);
  std::vector<char> active(K, 1), fallback(K, 0);
  int nact = K, j = 0, read_upto = 0;
  const Bases bases{X.p, Wa.p, Y.p};
  std::vector<double> hs;
  while (nact > 0) {
    hv_plans(ctx, Wp, Yp, bases);
    launch_lzb_iteration(dch.p, dmap.p, btot, K, j, kb1, full, s);
    clk.lap("lzb.enqueue");
    ++j;
    bool at_end = false;
    for (int c = 0; c < K; ++c)
      if (active[c]) {
        ++numop[c];
        at_end = at_end || j == kmc[c];
      }
    if (!(j <= 8 || j % 4 == 0 || at_end)) continue;
    const int cnt = j - read_upto;
    hs.resize((size_t)cnt * K * S);
    NSS_HIP(hipMemcpyAsync(hs.data(), SL.p + (size_t)read_upto * K * S, hs.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    NSS_HIP(hipStreamSynchronize(s));
    clk.lap("lzb.wait");
    bool changed = false;
    for (int c = 0; c < K; ++c) {
      if (!active[c]) continue;
      int m = j;
      bool breakdown = false;
      for (int it = read_upto; it < j; ++it) {
        const double* sl = hs.data() + ((size_t)(it - read_upto) * K + c) * S;
        d[c][it] = sl[it] + sl[kb1 + it];
        if (it > 0) e[c][it - 1] = sl[it - 1] + sl[kb1 + it - 1];
        beta[c][it] = sl[2 * kb1];
        hn[c] = std::max(hn[c], std::fabs(d[c][it]));
        if (beta[c][it] <= 1e-14 * std::max(1.0, hn[c]) || !std::isfinite(beta[c][it])) {  // invariant subspace
          m = it + 1;
          breakdown = true;
          break;
        }
      }
      // the lowest pair of the tridiagonal: theta, and its eigenvector y from
      // the same inverse iteration (its last component gives the residual)
      double th0 = 0.0, zl = 0.0;
      std::vector<double>& y = ybuf[c];
      y.assign(m, 0.0);
      tridiag_lowest(m, d[c].data(), e[c].data(), &th0, &zl, y.data());
      const double res = std::fabs(beta[c][m - 1] * zl);
      const bool conv = res <= tol * std::max(1.0, std::fabs(th0));
      if (!(conv || breakdown || m >= kmc[c])) continue;
      // Ritz vector x = V[0..m) y (y stays alive in ybuf until the final
      // synchronisation: no wait per chain)
      double* ydc = yd.p + (size_t)c * kb1;
      NSS_HIP(hipMemcpyAsync(ydc, y.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
      launch_gemv_n(hc[c].V, n[c], m, ydc, jobs[c].xdev, n[c], 1, s);
      launch_norm(jobs[c].xdev, n[c], hc[c].part, scal.p + c, s);
      launch_div(jobs[c].xdev, jobs[c].xdev, n[c], scal.p + c, s);
      jobs[c].out.evals.assign(1, th0);
      jobs[c].out.numop = numop[c];
      fallback[c] = !(conv || breakdown);
      active[c] = 0;
      hc[c].active = 0;
      --nact;
      changed = true;
    }
    read_upto = j;
    clk.lap("lzb.host");
    if (changed && nact > 0) {
      // drop the finished chains from the merged H*v and the vector launches
      NSS_HIP(hipMemcpyAsync(dch.p, hc.data(), K * sizeof(LzChainDev), hipMemcpyHostToDevice, s));
      std::vector<SbDevice*> asb;
      std::vector<int64_t> ax, aw;
      for (int c = 0; c < K; ++c)
        if (active[c]) {
          asb.push_back(sbs[c]);
          ax.push_back(xoff[c]);
          aw.push_back(woff[c]);
        }
      sb_merge_hv_plans(asb, ax, aw, Wp, Yp, ctx, s);
      clk.lap("lzb.merge");
      clk.count("lzb.merges");
    }
  }
  NSS_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < K; ++c)
    if (fallback[c]) {
      LanczosOut o = lanczos_device(ctx, sbs[c], 1, kmax_req, maxiter, tol, jobs[c].seed, 2, NSS_FP64, jobs[c].xdev);
      jobs[c].out.evals = o.evals;
      jobs[c].out.numop += o.numop;
    }
}

// Dense path of sb_diag (m_sb <= LANC_DIM_THRESHOLD, src/DMRG_SUPERBLOCK.f90:438-467):
// H_sb is assembled on the device column by column (H e_j) and diagonalised.
LanczosOut dense_diag(nss_ctx* ctx, SbDevice* sb, int neigen, double* xdev) {
  const int64_t n = sb_n(sb);
  NSS_REQUIRE(n > 0 && n <= 8192, NSS_E_ARG, "dense_diag: dimension out of range");
  hipStream_t s = ctx->stream;
  DevArr<double> E, H;
  E.resize(ctx->pool, (size_t)n * n);
  H.resize(ctx->pool, (size_t)n * n);
  // identity columns, one H*v per column, no host round trip in between
  std::vector<double> eye((size_t)n * n, 0.0);
  for (int64_t j = 0; j < n; ++j) eye[(size_t)j * n + j] = 1.0;
  NSS_HIP(hipMemcpyAsync(E.p, eye.data(), eye.size() * sizeof(double), hipMemcpyHostToDevice, s));
  const bool t = ctx->time_hv;
  ctx->time_hv = false;  // assembly is not a Lanczos H*v
  for (int64_t j = 0; j < n; ++j) sb_hv(ctx, sb, E.p + (size_t)j * n, H.p + (size_t)j * n);
  ctx->time_hv = t;
  std::vector<double> hh((size_t)n * n);
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
  out.numop = 0;
  return out;
}

}  // namespace nss

// Large rho blocks: the symmetric eigensolve of blocks_matrix%eigh
// (src/MATRIX/MATRIX_BLOCKS.f90:589-657, LAPACK dsyev through SciFortran's
// eigh) for blocks above the one-workgroup tridiagonal-path kernel, batched
// over every block of a step and spread over many workgroups per block.
//
//  1. Householder tridiagonalisation T = Q^T A Q (dsytd2's lower form,
//     Q = H_0 H_1 ... H_{n-3}, H_k = I - tau_k v_k v_k^T, v_k[k+1] = 1), one
//     launch per column j for all blocks.  Each wave of a workgroup owns one
//     row of the trailing matrix.  The workgroup recomputes, redundantly,
//     the previous step's w = p - (tau/2)(p.v) v and the updated row j (the
//     column j by symmetry), forms v_j, then each wave applies the previous
//     rank-2 update A -= v w^T + w v^T to its row and accumulates its share
//     of p_j = tau_j A v_j.  The kernel boundary is the only global
//     synchronisation.  Reflector k is kept in row k of the block
//     (columns k+1..n-1), written by the launch after the one that formed it
//     (nothing reads row k then).
//  2. Every eigenvalue of T by multisection on Sturm counts within its
//     unreduced block (T split where |e_k| <= eps ||T||, as dstebz splits),
//     EIG_BIS lanes per eigenvalue.
//  3. On request, for chosen positions only (the kept states of a
//     truncation): inverse iteration with T - lambda I (dstein: two solves
//     with the dgttrf factors from a pseudo-random start), one lane per
//     vector; Gram-Schmidt within clusters of eigenvalues closer than
//     1e-3 ||T|| (dstein's criterion), one workgroup per cluster, in panels
//     (modified within a panel, the panel then projected out of the rest); the
//     back-transformation x = Q z, one wave per group of vectors held in
//     registers while the reflectors stream past.
#include <cfloat>

#include "nss_internal.h"

namespace nss {

namespace {

constexpr int EIG_THREADS = 256;
constexpr int EIG_BIS = 4;     // lanes (interior points) per eigenvalue
constexpr int EIG_PER_WG = EIG_THREADS / EIG_BIS;
constexpr int GS_THREADS = 1024;
constexpr int GS_WAVES = GS_THREADS / 64;

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double rdlane(double v, int lane) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// sum over the wave (every lane active), returned in every lane
__device__ __forceinline__ double wsum(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror
  v += dpp_d<0x140>(v);  // row_mirror
  return (rdlane(v, 0) + rdlane(v, 16)) + (rdlane(v, 32) + rdlane(v, 48));
}

__device__ __forceinline__ double wmax(double v) {
  v = fmax(v, dpp_d<0xB1>(v));
  v = fmax(v, dpp_d<0x4E>(v));
  v = fmax(v, dpp_d<0x141>(v));
  v = fmax(v, dpp_d<0x140>(v));
  return fmax(fmax(rdlane(v, 0), rdlane(v, 16)), fmax(rdlane(v, 32), rdlane(v, 48)));
}

// workgroup sum / max with NW waves, in every thread (red: NW doubles of LDS)
template <int NW>
__device__ __forceinline__ double bsum(double v, double* red) {
  v = wsum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) r += red[w];
  __syncthreads();
  return r;
}

// the same with one barrier, for a red that the workgroup uses once
template <int NW>
__device__ __forceinline__ double bsum_once(double v, double* red) {
  v = wsum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) r += red[w];
  return r;
}

template <int NW>
__device__ __forceinline__ double bmax(double v, double* red) {
  v = wmax(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) r = fmax(r, red[w]);
  __syncthreads();
  return r;
}

// 1/q to full precision: the hardware reciprocal and two Newton steps
__device__ __forceinline__ double rcp_d(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  return fma(fma(-q, r, 1.0), r, r);
}

}  // namespace

int64_t tred_ws_size(int n) { return (int64_t)7 * n; }

// ------------------------------------------------------------ 1. tridiagonal
// Every global load of a column is issued at the start, in one round: the
// previous v and p and row j (into registers, then LDS) and this wave's own
// row (into registers, RV doubles per lane: n <= 64 RV), so the launch
// waits for memory once rather than once per dependent stage.  NT threads,
// one trailing-matrix row per wave.  Two workgroup barriers per launch: the
// block sums of p.v and of the column's squares reduce into LDS slots of
// their own, and v_j is scaled where it is read rather than in place.
template <int NT, int RV>
__global__ __launch_bounds__(NT) void k_tred_col(const TredBlk* __restrict__ blks, int nblk, int j, int nmax) {
  constexpr int NW = NT / 64, ROWS = NW, NU = (RHO_BIG_MAX + NT - 1) / NT;
  extern __shared__ double sm[];  // vp, wp over [j, n); vj over [j, n)
  __shared__ double red[NW], red2[NW];
  const int b = blockIdx.x % nblk, t = blockIdx.x / nblk;
  const TredBlk B = blks[b];
  const int n = B.n;
  if (j >= B.jt) return;
  const int r0 = j + 1 + t * ROWS;
  if (t > 0 && r0 >= n) return;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int m = n - j;  // the trailing range [j, n)
  double* vp = sm;
  double* wp = sm + nmax;
  double* vj = sm + 2 * nmax;
  double* A = B.a;
  double* ws = B.ws;
  double* vbuf = ws;          // [2][n]
  double* pbuf = ws + 2 * n;  // [2][n]
  double* dd = ws + 4 * n;
  double* ee = ws + 5 * n;
  double* tau = ws + 6 * n;
  // loads: v_{j-1}, p_{j-1} and row j over [j, n) (at most NU entries per
  // thread), this wave's row at columns j+1.. n-1
  const double* vq = vbuf + ((j - 1) & 1) * n;
  const double* pq = pbuf + ((j - 1) & 1) * n;
  const double* rowj = A + (int64_t)j * n + j;
  double lv[NU], lp[NU], lr[NU];
  const double gv0 = j > 0 ? vq[j] : 0.0, gp0 = j > 0 ? pq[j] : 0.0;  // v, p at row j
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int i = tid + u * NT;
    lv[u] = lp[u] = lr[u] = 0.0;
    if (i < m) {
      if (j > 0) {
        lv[u] = vq[j + i];
        lp[u] = pq[j + i];
      }
      lr[u] = rowj[i];
    }
  }
  const int io = r0 + wv;  // this wave's row
  const bool own = io < n;
  double* row = A + (int64_t)(own ? io : j) * n;
  double rv[RV];
#pragma unroll
  for (int u = 0; u < RV; ++u) {
    const int c = j + 1 + lane + 64 * u;
    rv[u] = (own && c < n) ? row[c] : 0.0;
  }
  // previous step: w = p - (tau/2)(p.v) v
  const double taup = j > 0 ? tau[j - 1] : 0.0;
  double pv = 0.0;
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int i = tid + u * NT;
    if (i < m) {
      vp[i] = lv[u];
      pv += lp[u] * lv[u];
    }
  }
  pv = bsum_once<NW>(pv, red);  // (its barrier also publishes vp)
  const double K = 0.5 * taup * pv;
  // w over [j, n) to LDS, and row j after the previous update (the column j
  // by symmetry) into vj; w and v at row j itself come from the loads
  const double w0 = gp0 - K * gv0;
  double s2 = 0.0;
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int i = tid + u * NT;
    if (i < m) {
      const double w = lp[u] - K * lv[u];
      wp[i] = w;
      const double r = lr[u] - gv0 * w - w0 * lv[u];
      vj[i] = r;
      if (i >= 2) s2 += r * r;
    }
  }
  s2 = bsum_once<NW>(s2, red2);  // (its barrier also publishes wp, vj)
  const double dj = vj[0];
  const double alpha = m > 1 ? vj[1] : 0.0;
  double tj = 0.0, beta = alpha, scale = 0.0;
  if (s2 > 0.0) {
    beta = -copysign(sqrt(alpha * alpha + s2), alpha);
    tj = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  // v_j at local index i >= 1 (vj holds the unscaled column)
  auto vjs = [&](int i) { return i == 1 ? 1.0 : vj[i] * scale; };
  if (t == 0) {
    if (tid == 0) {
      dd[j] = dj;
      ee[j] = m > 1 ? beta : 0.0;
      tau[j] = tj;
    }
    double* vo = vbuf + (j & 1) * n;
    for (int i = tid + 1; i < m; i += NT) vo[j + i] = vjs(i);
    if (j > 0) {  // reflector j-1 to its row (columns j..n-1)
      double* keep = A + (int64_t)(j - 1) * n;
      for (int i = tid; i < m; i += NT) keep[j + i] = vp[i];
    }
  }
  // this wave's row: the previous update, then its share of p_j = tau_j A v_j
  if (own) {
    const int li = io - j;
    const double vi = vp[li], wi = wp[li];
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < RV; ++u) {
      const int c = j + 1 + lane + 64 * u;
      if (c < n) {
        const int lc = c - j;
        double a = rv[u];
        if (j > 0) {
          a -= vi * wp[lc] + wi * vp[lc];
          row[c] = a;
        }
        acc += a * vjs(lc);
      }
    }
    acc = wsum(acc);
    if (lane == 0) (pbuf + (j & 1) * n)[io] = tj * acc;
  }
}

void launch_tred_col(const TredBlk* blks, int nblk, int j, int nmax, hipStream_t s) {
  NSS_REQUIRE(nmax <= RHO_BIG_MAX, NSS_E_ARG, "tridiagonalisation: block above the large-block limit");
  const size_t lds = (size_t)3 * nmax * sizeof(double);
  if (nmax <= 1024) {
    const int tiles = std::max(1, (nmax - j - 1 + 15) / 16);
    hipLaunchKernelGGL((k_tred_col<1024, 16>), dim3(nblk * tiles), dim3(1024), lds, s, blks, nblk, j, nmax);
  } else {
    const int tiles = std::max(1, (nmax - j - 1 + 7) / 8);
    hipLaunchKernelGGL((k_tred_col<512, 32>), dim3(nblk * tiles), dim3(512), lds, s, blks, nblk, j, nmax);
  }
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ 1'. tridiagonal, LDS-resident
// The same reduction with the trailing matrix held on chip: the G workgroups
// of a block own its rows cyclically (row i in workgroup i mod G, slot
// i / G of its LDS), one launch covers every column, and a per-block
// counter barrier between columns replaces the kernel boundary.  Per column
// j a workgroup reads only p_{j-1} (rows j..n-1) and row j (published by its
// owner in the previous column, after the update of column j-2), recomputes
// w_{j-1}, row j after the update of column j-1 and v_j redundantly (every
// workgroup runs the same arithmetic in the same order, so their copies agree
// bit for bit), updates its own rows in LDS and publishes its share of p_j.
// Reflector j-1 goes to row j-1 of the block in global memory, as in the
// per-column kernel.  The launch is cooperative (every workgroup resident);
// a barrier wait longer than TL_SPIN_MAX polls raises the abort flag, after
// which every workgroup leaves.
constexpr int TL_THREADS = 512;
constexpr int TL_WAVES = TL_THREADS / 64;
constexpr unsigned TL_SPIN_MAX = 1u << 22;

// Every wave waits for its own stores, the workgroup meets, and one thread
// releases them at agent scope with the counter increment; it polls the
// counter with relaxed loads (an acquiring poll would invalidate the caches
// on every round) and acquires once when the count is reached.
__device__ __forceinline__ bool tl_barrier(unsigned* cnt, unsigned target, int* abort, int* flag) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    unsigned spins = 0;
    while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if (++spins > TL_SPIN_MAX) {
        __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *flag = ok;
  }
  __syncthreads();
  return *flag != 0;
}

__global__ __launch_bounds__(TL_THREADS) void k_tred_lds(const TredBlk* __restrict__ blks,
                                                         const int2* __restrict__ wgmap, const int* __restrict__ gcnt,
                                                         unsigned* __restrict__ bar, int* __restrict__ abort) {
  extern __shared__ double sm[];  // rows[R][n], then vbuf[2][n], wp[n]
  __shared__ double red[TL_WAVES];
  __shared__ int flag;
  const int2 mp = wgmap[blockIdx.x];
  if (mp.x < 0) return;
  const int b = mp.x, g = mp.y, G = gcnt[b];
  const TredBlk B = blks[b];
  const int n = B.n;
  const int R = (n - g + G - 1) / G;  // rows g, g + G, ...
  double* rows = sm;
  double* vb0 = sm + (size_t)R * n;
  double* vb1 = vb0 + n;
  double* wp = vb1 + n;
  double* A = B.a;
  double* ws = B.ws;
  double* rowbuf = ws;        // [2][n]: row j+1 after the update of column j-1
  double* pbuf = ws + 2 * n;  // [2][n]
  double* dd = ws + 4 * n;
  double* ee = ws + 5 * n;
  double* tau = ws + 6 * n;
  unsigned* cnt = bar + b;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  for (int s = 0; s < R; ++s) {
    const double* src = A + (int64_t)(g + s * G) * n;
    for (int c = tid; c < n; c += TL_THREADS) rows[(size_t)s * n + c] = src[c];
  }
  for (int c = tid; c < n; c += TL_THREADS) vb0[c] = vb1[c] = wp[c] = 0.0;
  __syncthreads();
  double taup = 0.0;
  for (int j = 0; j < n; ++j) {
    double* vp = (j & 1) ? vb0 : vb1;  // v_{j-1} at its columns j..n-1 (zero for j = 0)
    double* vj = (j & 1) ? vb1 : vb0;
    // w_{j-1} = p_{j-1} - (tau/2)(p.v) v over rows j..n-1
    if (j > 0) {
      const double* pq = pbuf + ((j - 1) & 1) * n;
      double pv = 0.0;
      for (int c = j + tid; c < n; c += TL_THREADS) {
        const double p = pq[c];
        wp[c] = p;
        pv += p * vp[c];
      }
      pv = bsum<TL_WAVES>(pv, red);
      const double K = 0.5 * taup * pv;
      for (int c = j + tid; c < n; c += TL_THREADS) wp[c] -= K * vp[c];
      __syncthreads();
    }
    // row j after the update of column j-1, into vj[j..n)
    const double v0 = vp[j], w0 = wp[j];
    const double* rj = j == 0 ? A : rowbuf + (j & 1) * n;
    double s2 = 0.0;
    for (int c = j + tid; c < n; c += TL_THREADS) {
      const double r = rj[c] - v0 * wp[c] - w0 * vp[c];
      vj[c] = r;
      if (c >= j + 2) s2 += r * r;
    }
    s2 = bsum<TL_WAVES>(s2, red);
    const double dj = vj[j];
    const double alpha = j + 1 < n ? vj[j + 1] : 0.0;
    double tj = 0.0, beta = alpha, scale = 0.0;
    if (s2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + s2), alpha);
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    __syncthreads();  // every thread has read vj[j], vj[j + 1]
    for (int c = j + 1 + tid; c < n; c += TL_THREADS) vj[c] = c == j + 1 ? 1.0 : vj[c] * scale;
    __syncthreads();
    if (g == 0) {
      if (tid == 0) {
        dd[j] = dj;
        ee[j] = j + 1 < n ? beta : 0.0;
        tau[j] = tj;
      }
      if (j > 0) {  // reflector j-1 to its row (columns j..n-1)
        double* keep = A + (int64_t)(j - 1) * n;
        for (int c = j + tid; c < n; c += TL_THREADS) keep[c] = vp[c];
      }
    }
    // owned rows i > j: the update of column j-1, then p_j[i] = tau_j A_i v_j;
    // row j+1 also to rowbuf for the next column
    double* po = pbuf + (j & 1) * n;
    double* rn = rowbuf + ((j + 1) & 1) * n;
    const int s0 = j + 1 > g ? (j + 1 - g + G - 1) / G : 0;  // first slot with row > j
    for (int s = s0 + wv; s < R; s += TL_WAVES) {
      const int i = g + s * G;
      const double vi = vp[i], wi = wp[i];
      double* row = rows + (size_t)s * n;
      double acc = 0.0;
      const bool pub = i == j + 1;
#pragma unroll 4
      for (int c = j + 1 + lane; c < n; c += 64) {
        const double a = row[c] - vi * wp[c] - wi * vp[c];
        row[c] = a;
        acc += a * vj[c];
        if (pub) rn[c] = a;
      }
      acc = wsum(acc);
      if (lane == 0) po[i] = tj * acc;
    }
    taup = tj;
    if (!tl_barrier(cnt, (unsigned)(j + 1) * (unsigned)G, abort, &flag)) return;
  }
}

namespace {
int tl_lds_bytes(int device) {
  static int cached = -1;
  if (cached >= 0) return cached;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lds = 65536;
  const int dyn = lds - 1024;  // the kernel's static LDS is under 100 bytes
  if (hipFuncSetAttribute((const void*)k_tred_lds, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) != hipSuccess) {
    (void)hipGetLastError();
    cached = 0;
  } else {
    cached = dyn;
  }
  return cached;
}
}  // namespace

bool tred_lds_plan(int device, const std::vector<int>& dims, TredLdsPlan* plan) {
  plan->wgmap.clear();
  plan->gcount.assign(dims.size(), 0);
  plan->covered.assign(dims.size(), 0);
  plan->grid = 0;
  const int lds = tl_lds_bytes(device);
  if (lds <= 0) return false;
  int ncu = 0, occ = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0)
    return false;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_tred_lds, TL_THREADS, (size_t)lds) !=
          hipSuccess ||
      occ <= 0) {
    (void)hipGetLastError();
    return false;
  }
  constexpr int NX = 8;  // workgroup k runs on XCD k mod 8: a block's workgroups share one L2 where they fit
  const int per_x = ncu * occ / NX;
  if (per_x <= 0) return false;
  const int cap = per_x * NX;
  // workgroups per block: the fewest whose rows fit in LDS
  std::vector<int> need(dims.size(), 0);
  for (size_t b = 0; b < dims.size(); ++b) {
    const int64_t n = dims[b];
    const int64_t room = lds / (int64_t)sizeof(double) - 3 * n;  // doubles left for rows
    const int64_t rmax = room / n;
    if (rmax < 1) continue;
    need[b] = (int)((n + rmax - 1) / rmax);
  }
  // largest blocks first, while the launch stays co-resident
  std::vector<int> ord(dims.size());
  for (size_t b = 0; b < ord.size(); ++b) ord[b] = (int)b;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return dims[x] > dims[y]; });
  int used = 0;
  for (int b : ord)
    if (need[b] > 0 && used + need[b] <= cap) {
      plan->gcount[b] = need[b];
      plan->covered[b] = 1;
      used += need[b];
    }
  if (used == 0) return false;
  // spare workgroups to the blocks with the most rows per workgroup
  for (;;) {
    int best = -1;
    int64_t bw = 0;
    for (int b : ord)
      if (plan->covered[b]) {
        const int64_t R = (dims[b] + plan->gcount[b] - 1) / plan->gcount[b];
        if (R > 8 && R * dims[b] > bw) {
          bw = R * dims[b];
          best = b;
        }
      }
    if (best < 0 || used >= cap) break;
    ++plan->gcount[best];
    ++used;
  }
  // workgroup ids: XCD x holds ids x, x + 8, ...; a block goes to the XCD
  // with the most free slots, spilling over the next ones if it needs more
  std::vector<std::vector<int2>> bins(NX);
  std::vector<int> ordg(ord);
  std::stable_sort(ordg.begin(), ordg.end(), [&](int x, int y) { return plan->gcount[x] > plan->gcount[y]; });
  for (int b : ordg) {
    if (!plan->covered[b]) continue;
    int left = plan->gcount[b], gi = 0;
    while (left > 0) {
      int x = 0;
      for (int k = 1; k < NX; ++k)
        if (bins[k].size() < bins[x].size()) x = k;
      const int take = std::min(left, per_x - (int)bins[x].size());
      if (take <= 0) return false;  // cannot happen: used <= cap
      for (int t = 0; t < take; ++t) bins[x].push_back(make_int2(b, gi++));
      left -= take;
    }
  }
  size_t slots = 0;
  for (auto& v : bins) slots = std::max(slots, v.size());
  plan->grid = (int)(slots * NX);
  plan->wgmap.assign(plan->grid, make_int2(-1, 0));
  for (int x = 0; x < NX; ++x)
    for (size_t t = 0; t < bins[x].size(); ++t) plan->wgmap[t * NX + x] = bins[x][t];
  plan->lds = (size_t)lds;
  return true;
}

int launch_tred_lds(const TredBlk* blks, const int2* wgmap, const int* gcnt, unsigned* bar, int* abort, int grid,
                    size_t lds, hipStream_t s) {
  void* args[] = {(void*)&blks, (void*)&wgmap, (void*)&gcnt, (void*)&bar, (void*)&abort};
  const hipError_t e = hipLaunchCooperativeKernel((const void*)k_tred_lds, dim3(grid), dim3(TL_THREADS), args,
                                                  (unsigned)lds, s);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return 1;
}

// ------------------------------------------------------------ 2. eigenvalues
// Position p of block b belongs to the unreduced block [s, e) of T that holds
// row p and takes eigenvalue p - s of that block (ascending within it).
__global__ __launch_bounds__(EIG_THREADS) void k_tri_eigvals(const TredBlk* __restrict__ blks,
                                                             const int64_t* __restrict__ eoff,
                                                             double* __restrict__ wout, int nblk, int nmax) {
  extern __shared__ double sm[];  // dg, eg, e2 (nmax each)
  __shared__ double red[EIG_THREADS / 64];
  __shared__ int sbeg[EIG_PER_WG], send[EIG_PER_WG];
  const int b = blockIdx.x % nblk, g = blockIdx.x / nblk;
  const TredBlk B = blks[b];
  const int n = B.n;
  const int p0 = g * EIG_PER_WG;
  if (p0 >= n) return;
  const int tid = threadIdx.x;
  double* dg = sm;
  double* eg = sm + nmax;
  double* e2 = sm + 2 * nmax;
  const double* dsrc = B.ws + 4 * n;
  const double* esrc = B.ws + 5 * n;
  double tn = 0.0, lo = HUGE_VAL, hi = -HUGE_VAL, em = 0.0;
  for (int i = tid; i < n; i += EIG_THREADS) {
    const double d = dsrc[i];
    const double el = i > 0 ? fabs(esrc[i - 1]) : 0.0;
    const double er = i + 1 < n ? fabs(esrc[i]) : 0.0;
    dg[i] = d;
    eg[i] = i + 1 < n ? esrc[i] : 0.0;
    tn = fmax(tn, fabs(d) + el + er);
    lo = fmin(lo, d - el - er);
    hi = fmax(hi, d + el + er);
    if (i + 1 < n) em = fmax(em, er * er);
  }
  tn = fmax(bmax<EIG_THREADS / 64>(tn, red), DBL_MIN);
  lo = -bmax<EIG_THREADS / 64>(-lo, red);
  hi = bmax<EIG_THREADS / 64>(hi, red);
  em = bmax<EIG_THREADS / 64>(em, red);
  const double pivmin = DBL_MIN * fmax(1.0, em);
  for (int i = tid; i + 1 < n; i += EIG_THREADS)
    if (fabs(eg[i]) <= DBL_EPSILON * tn) eg[i] = 0.0;
  __syncthreads();
  for (int i = tid; i < n; i += EIG_THREADS) e2[i] = i + 1 < n ? eg[i] * eg[i] : 0.0;
  // the unreduced blocks of this workgroup's positions (wave 0: ballots over
  // 64 consecutive positions, then a scan in chunks of 64 beyond them)
  if (tid < 64) {
    const int p = p0 + tid;
    const bool starts = p < n && (p == 0 || eg[p - 1] == 0.0);
    const bool ends = p < n && (p == n - 1 || eg[p] == 0.0);
    const uint64_t bs = __ballot(starts), be = __ballot(ends);
    // start before p0: nearest k <= p0 with k == 0 or eg[k-1] == 0
    int sprev = p0;
    if (!(bs & 1ull)) {
      for (int base = p0 - 1; base >= 0; base -= 64) {  // k = base - lane
        const int k = base - tid;
        const bool z = k >= 0 && (k == 0 || eg[k - 1] == 0.0);
        const uint64_t bz = __ballot(z);
        if (bz) {
          sprev = base - (int)__builtin_ctzll(bz);
          break;
        }
      }
    }
    // end after the last position: nearest k >= p0+63 with eg[k] == 0 or k == n-1
    const int plast = min(p0 + 63, n - 1);
    int enext = plast + 1;
    if (!((be >> (plast - p0)) & 1ull)) {
      for (int base = plast + 1; base < n; base += 64) {
        const int k = base + tid;
        const bool z = k < n && (k == n - 1 || eg[k] == 0.0);
        const uint64_t bz = __ballot(z);
        if (bz) {
          enext = base + (int)__builtin_ctzll(bz) + 1;
          break;
        }
      }
    }
    if (p < n) {
      const uint64_t below = bs & ((tid == 63) ? ~0ull : ((2ull << tid) - 1ull));
      sbeg[tid] = below ? p0 + 63 - (int)__builtin_clzll(below) : sprev;
      const uint64_t above = be & ~((1ull << tid) - 1ull);
      send[tid] = above ? p0 + (int)__builtin_ctzll(above) + 1 : enext;
    }
  }
  __syncthreads();
  const int lp = tid / EIG_BIS, lb = tid % EIG_BIS;
  const int p = p0 + lp;
  if (p >= n) return;  // (no barrier follows)
  const int bs = sbeg[lp], be = send[lp];
  const int tix = p - bs;
  double l = lo, h = hi;
  if (be - bs == 1) l = h = dg[p];
  const double tol = 2.0 * DBL_EPSILON * tn;
  constexpr double inv_parts = 1.0 / (EIG_BIS + 1);
  for (int it = 0; it < 96; ++it) {
    if (!(h - l > tol)) break;  // alike on the lanes of a group
    const double x = l + (h - l) * (double)(lb + 1) * inv_parts;
    double q = dg[bs] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int cnt = q < 0.0;
#pragma unroll 8
    for (int i = bs + 1; i < be; ++i) {
      q = dg[i] - x - e2[i - 1] * rcp_d(q);
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q < 0.0;
    }
    const bool above = cnt >= tix + 1;  // eigenvalue tix lies below x
    const uint64_t bal = __ballot(above);
    const unsigned bits = (unsigned)(bal >> (threadIdx.x & (64 - EIG_BIS))) & ((1u << EIG_BIS) - 1u);
    const int f = bits ? __builtin_ctz(bits) : EIG_BIS;  // first point above it
    const double span = h - l;
    const double nl = f > 0 ? l + span * (double)f * inv_parts : l;
    const double nh = f < EIG_BIS ? l + span * (double)(f + 1) * inv_parts : h;
    l = nl;
    h = fmax(nh, nl);
  }
  if (lb == 0) wout[eoff[b] + p] = 0.5 * (l + h);
}

void launch_tri_eigvals(const TredBlk* blks, const int64_t* eoff, double* w, int nblk, int nmax, hipStream_t s) {
  const int groups = (nmax + EIG_PER_WG - 1) / EIG_PER_WG;
  hipLaunchKernelGGL(k_tri_eigvals, dim3(nblk * groups), dim3(EIG_THREADS), (size_t)3 * nmax * sizeof(double), s,
                     blks, eoff, w, nblk, nmax);
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ 3. eigenvectors
// One lane per requested vector: LU with partial pivoting of T[rs..re) -
// lam I (its factors at fac[(f * nmax + i) * nreq + r]), two solves from a
// SplitMix64 start, the normalised iterate to rows rs..re-1 of column slot
// of z (rows outside stay as the caller zeroed them).
__global__ __launch_bounds__(256) void k_tri_invit(const InvReq* __restrict__ rq, int nreq,
                                                   double* __restrict__ fac, int nmax) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nreq) return;
  const InvReq R = rq[r];
  const int nb = R.re - R.rs;
  double* zb = R.z + (int64_t)R.rs * R.ldz + R.slot;  // local row i at zb[i * ldz]
  const int64_t ldz = R.ldz;
  if (nb == 1) {
    zb[0] = 1.0;
    return;
  }
  const int64_t fs = (int64_t)nmax * nreq;
  double* DL = fac + r;
  double* DD = fac + fs + r;
  double* DU = fac + 2 * fs + r;
  double* DU2 = fac + 3 * fs + r;
  double* PV = fac + 4 * fs + r;
  const double* dgb = R.d + R.rs;
  const double* egb = R.e + R.rs;
  const double sh = R.lam, tiny = R.tiny;
  double ddi = dgb[0] - sh, dui = egb[0];
  for (int i = 0; i + 1 < nb; ++i) {
    const double dli = egb[i];
    double ddn = dgb[i + 1] - sh;
    double dun = (i + 2 < nb) ? egb[i + 1] : 0.0;
    double f, du2i = 0.0, pvi = 0.0;
    if (fabs(ddi) >= fabs(dli)) {
      if (ddi == 0.0) ddi = tiny;
      f = dli / ddi;
      ddn -= f * dui;
    } else {
      f = ddi / dli;
      const double tmp = dui;
      ddi = dli;
      dui = ddn;
      ddn = tmp - f * ddn;
      if (i + 2 < nb) {
        du2i = dun;
        dun = -f * dun;
      }
      pvi = 1.0;
    }
    DL[(int64_t)i * nreq] = f;
    DD[(int64_t)i * nreq] = ddi;
    DU[(int64_t)i * nreq] = dui;
    DU2[(int64_t)i * nreq] = du2i;
    PV[(int64_t)i * nreq] = pvi;
    ddi = ddn;
    dui = dun;
  }
  if (ddi == 0.0) ddi = tiny;
  DD[(int64_t)(nb - 1) * nreq] = ddi;
  for (int i = 0; i < nb; ++i) {  // start: uniform [-1, 1) from SplitMix64 of (position, row)
    uint64_t z = ((uint64_t)R.seedpos << 24) + (uint64_t)(i + 1);
    z *= 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    zb[i * ldz] = (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  }
  for (int pass = 0; pass < 2; ++pass) {
    double xi = zb[0];
#pragma unroll 8
    for (int i = 0; i + 1 < nb; ++i) {  // L, with the row interchanges
      double xn = zb[(i + 1) * ldz];
      if (PV[(int64_t)i * nreq] != 0.0) {
        const double t = xi;
        xi = xn;
        xn = t;
      }
      xn -= DL[(int64_t)i * nreq] * xi;
      zb[i * ldz] = xi;
      xi = xn;
    }
    double x1 = xi / DD[(int64_t)(nb - 1) * nreq];  // U (two superdiagonals)
    zb[(nb - 1) * ldz] = x1;
    double x0 = (zb[(nb - 2) * ldz] - DU[(int64_t)(nb - 2) * nreq] * x1) / DD[(int64_t)(nb - 2) * nreq];
    zb[(nb - 2) * ldz] = x0;
    double x2 = x1;
    x1 = x0;
#pragma unroll 8
    for (int i = nb - 3; i >= 0; --i) {
      x0 = (zb[i * ldz] - DU[(int64_t)i * nreq] * x1 - DU2[(int64_t)i * nreq] * x2) / DD[(int64_t)i * nreq];
      zb[i * ldz] = x0;
      x2 = x1;
      x1 = x0;
    }
    double mx = 0.0;
    for (int i = 0; i < nb; ++i) mx = fmax(mx, fabs(zb[i * ldz]));
    double ss = 0.0;
    for (int i = 0; i < nb; ++i) {
      const double y = zb[i * ldz] / mx;
      ss += y * y;
    }
    const double inv = 1.0 / (mx * sqrt(ss));
    for (int i = 0; i < nb; ++i) zb[i * ldz] *= inv;
  }
}

void launch_tri_invit(const InvReq* req, int nreq, double* fac, int nmax, hipStream_t s) {
  if (nreq <= 0) return;
  // one wave per workgroup: each lane streams its own factors through
  // memory, so the vectors' waves are spread over as many CUs as they fill
  hipLaunchKernelGGL(k_tri_invit, dim3((nreq + 63) / 64), dim3(64), 0, s, req, nreq, fac, nmax);
  NSS_HIP(hipGetLastError());
}

// Block Gram-Schmidt over the vectors s0..s1-1 of a cluster (nonzero in rows
// rs..re-1 only), up to two sweeps, one workgroup per cluster.  The vectors
// are taken in panels of PW: a panel is copied to LDS and orthonormalised
// there by modified Gram-Schmidt (step j: h_k = z_j . z_k for the panel's
// k >= j, then z_k -= (h_k / h_j) z_j; z_j is scaled by 1/sqrt(h_j) in step
// j+1, which reads neither z_j nor its scale), written back, and projected
// out of every later vector of the cluster at once (C = P^T Z, Z -= P C,
// waves over rows and lanes over vectors, the wave partials of C summed in
// a fixed order), so each later vector is read twice per panel rather than
// twice per vector.  A second sweep runs where the first removed overlaps
// above 1e-6.
template <int PW>
__global__ __launch_bounds__(GS_THREADS) void k_tri_cluster_gs(const GsCluster* __restrict__ cls) {
  extern __shared__ double gsm[];
  __shared__ double dr[GS_WAVES][PW];
  __shared__ double hh[PW];
  __shared__ double red[GS_WAVES];
  const GsCluster C = cls[blockIdx.x];
  const int c = C.s1 - C.s0;
  if (c < 2) return;
  const int nr = C.re - C.rs;
  const int64_t ldz = C.ldz;
  double* Z = C.z + (int64_t)C.rs * ldz + C.s0;  // row i, vector k at Z[i * ldz + k]
  double* Pl = gsm;                                // panel: row i, column q at Pl[i * PW + q]
  double* part = gsm + (size_t)nr * PW;           // wave partials of C: [wave][q][lane]
  double* Cc = part + (size_t)GS_WAVES * PW * 64;  // C: [q][lane]
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  constexpr int RG = GS_THREADS / PW;  // row groups of the in-panel dot products
  for (int sweep = 0; sweep < 2; ++sweep) {
    double ovl = 0.0;
    for (int p0 = 0; p0 < c; p0 += PW) {
      const int pw = min(PW, c - p0), p1 = p0 + pw;
      for (int e = tid; e < nr * PW; e += GS_THREADS) {
        const int i = e / PW, q = e - i * PW;
        Pl[e] = q < pw ? Z[(int64_t)i * ldz + p0 + q] : 0.0;
      }
      __syncthreads();
      // (a) modified Gram-Schmidt in LDS
      double inv = 1.0;
      for (int j = 0; j < pw; ++j) {
        // phase 1: column j-1 scaled; partial dots of column j with k >= j
        const int q = tid & (PW - 1), g = tid / PW;
        if (j > 0)
          for (int i = tid; i < nr; i += GS_THREADS) Pl[i * PW + j - 1] *= inv;
        double acc = 0.0;
        if (q >= j && q < pw)
          for (int i = g; i < nr; i += RG) acc += Pl[i * PW + j] * Pl[i * PW + q];
        // lanes with equal q (lane % PW) hold one sum each: reduce them
#pragma unroll
        for (int o = PW; o < 64; o <<= 1) acc += __shfl_xor(acc, o);
        if (lane < PW) dr[wv][lane] = acc;
        __syncthreads();
        if (tid < PW) {
          double h = 0.0;
#pragma unroll
          for (int w = 0; w < GS_WAVES; ++w) h += dr[w][tid];
          hh[tid] = h;
        }
        __syncthreads();
        // phase 2: z_k -= (h_k / h_j) z_j for k > j
        const double hj = hh[j];
        for (int e = tid; e < nr * PW; e += GS_THREADS) {
          const int i = e / PW, k = e - i * PW;
          if (k > j && k < pw) Pl[e] -= (hh[k] / hj) * Pl[i * PW + j];
        }
        for (int k = j + 1; k < pw; ++k) ovl = fmax(ovl, fabs(hh[k] / hj));
        inv = 1.0 / sqrt(hj);
        __syncthreads();
      }
      for (int i = tid; i < nr; i += GS_THREADS) Pl[i * PW + pw - 1] *= inv;
      __syncthreads();
      for (int e = tid; e < nr * PW; e += GS_THREADS) {
        const int i = e / PW, q = e - i * PW;
        if (q < pw) Z[(int64_t)i * ldz + p0 + q] = Pl[e];
      }
      // (b) project the panel out of the cluster's later vectors
      for (int k0 = p1; k0 < c; k0 += 64) {
        const int k = k0 + lane;
        const bool valid = k < c;
        double acc[PW];
#pragma unroll
        for (int q = 0; q < PW; ++q) acc[q] = 0.0;
        for (int i = wv; i < nr; i += GS_WAVES) {
          const double z = valid ? Z[(int64_t)i * ldz + k] : 0.0;
#pragma unroll
          for (int q = 0; q < PW; ++q) acc[q] = fma(Pl[i * PW + q], z, acc[q]);
        }
#pragma unroll
        for (int q = 0; q < PW; ++q) part[(wv * PW + q) * 64 + lane] = acc[q];
        __syncthreads();
        for (int e = tid; e < PW * 64; e += GS_THREADS) {
          const int q = e >> 6, l = e & 63;
          double v = 0.0;
#pragma unroll
          for (int w = 0; w < GS_WAVES; ++w) v += part[(w * PW + q) * 64 + l];
          Cc[e] = (q < pw && k0 + l < c) ? v : 0.0;
        }
        __syncthreads();
        if (valid) {
#pragma unroll
          for (int q = 0; q < PW; ++q) ovl = fmax(ovl, fabs(Cc[q * 64 + lane]));
          for (int i = wv; i < nr; i += GS_WAVES) {
            double z = Z[(int64_t)i * ldz + k];
#pragma unroll
            for (int q = 0; q < PW; ++q) z = fma(-Pl[i * PW + q], Cc[q * 64 + lane], z);
            Z[(int64_t)i * ldz + k] = z;
          }
        }
        __syncthreads();
      }
    }
    // a second sweep only where the first removed overlaps above 1e-6
    if (!(bmax<GS_WAVES>(ovl, red) > 1e-6)) break;
  }
}

// LDS of the block Gram-Schmidt: the panel, the wave partials and C
static size_t gs_lds(int pw, int nr) {
  return ((size_t)nr * pw + (size_t)GS_WAVES * pw * 64 + (size_t)pw * 64) * sizeof(double);
}

void launch_tri_cluster_gs(const GsCluster* cl, int ncl, int nrmax, hipStream_t s) {
  if (ncl <= 0) return;
  NSS_REQUIRE(nrmax >= 1 && nrmax <= RHO_BIG_MAX, NSS_E_ARG, "cluster Gram-Schmidt: rows above the block limit");
  static const bool attr = [] {  // dynamic LDS above the default limit
    (void)hipFuncSetAttribute((const void*)k_tri_cluster_gs<8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)gs_lds(8, 1024));
    (void)hipFuncSetAttribute((const void*)k_tri_cluster_gs<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)gs_lds(4, RHO_BIG_MAX));
    return true;
  }();
  (void)attr;
  if (nrmax <= 1024)
    hipLaunchKernelGGL(k_tri_cluster_gs<8>, dim3(ncl), dim3(GS_THREADS), gs_lds(8, nrmax), s, cl);
  else hipLaunchKernelGGL(k_tri_cluster_gs<4>, dim3(ncl), dim3(GS_THREADS), gs_lds(4, nrmax), s, cl);
  NSS_HIP(hipGetLastError());
}

// x = H_0 (H_1 (... (H_{n-3} z))) for V vectors per wave, row lane + 64 t of
// vector v in register z[t][v] (NPL rows per lane: n <= 64 NPL).  Reflectors
// (and their tau) are loaded PF steps ahead into a ring of registers, so a
// wave waits for memory about once per PF reflectors.
template <int NPL, int V, int PF>
__global__ __launch_bounds__(256) void k_tri_backtr(const BtTask* __restrict__ tasks, int ntask) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= ntask) return;  // wave-uniform
  const BtTask T = tasks[w];
  const int lane = threadIdx.x & 63;
  const int n = T.n;
  double z[NPL][V];
#pragma unroll
  for (int t = 0; t < NPL; ++t) {
    const int i = lane + 64 * t;
#pragma unroll
    for (int v = 0; v < V; ++v) z[t][v] = (i < n && v < T.ns) ? T.z[(int64_t)i * T.ldz + T.s0 + v] : 0.0;
  }
  // ring slot u holds reflector k0 - u - PF q at step q (zero for k < 0)
  double vr[PF][NPL], tq[PF];
  const int k0 = n - 3;
  auto load = [&](int k, double (&dst)[NPL], double& tk) {
    const double* vk = T.a + (int64_t)(k > 0 ? k : 0) * n;
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
      const int i = lane + 64 * t;
      dst[t] = (k >= 0 && i > k && i < n) ? vk[i] : 0.0;
    }
    tk = k >= 0 ? T.tau[k] : 0.0;
  };
#pragma unroll
  for (int u = 0; u < PF; ++u) load(k0 - u, vr[u], tq[u]);
  for (int kb = k0; kb >= 0; kb -= PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int k = kb - u;
      if (k >= 0) {  // uniform
        const double tk = tq[u];
        if (tk != 0.0) {
          double d[V];
#pragma unroll
          for (int v = 0; v < V; ++v) d[v] = 0.0;
#pragma unroll
          for (int t = 0; t < NPL; ++t)
#pragma unroll
            for (int v = 0; v < V; ++v) d[v] = fma(vr[u][t], z[t][v], d[v]);
#pragma unroll
          for (int v = 0; v < V; ++v) d[v] = tk * wsum(d[v]);
#pragma unroll
          for (int t = 0; t < NPL; ++t)
#pragma unroll
            for (int v = 0; v < V; ++v) z[t][v] = fma(-vr[u][t], d[v], z[t][v]);
        }
        load(k - PF, vr[u], tq[u]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NPL; ++t) {
    const int i = lane + 64 * t;
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (i < n && v < T.ns) T.z[(int64_t)i * T.ldz + T.s0 + v] = z[t][v];
  }
}

int backtr_vectors_per_task(int n) { return n <= 1024 ? 1 : 1; }

void launch_tri_backtr(const BtTask* tasks, int ntask, int nmax, hipStream_t s) {
  if (ntask <= 0) return;
  NSS_REQUIRE(nmax <= RHO_BIG_MAX, NSS_E_ARG, "back-transformation: block above the large-block limit");
  const dim3 grid((ntask + 3) / 4);
  if (nmax <= 1024) hipLaunchKernelGGL((k_tri_backtr<16, 1, 4>), grid, dim3(256), 0, s, tasks, ntask);
  else hipLaunchKernelGGL((k_tri_backtr<32, 1, 4>), grid, dim3(256), 0, s, tasks, ntask);
  NSS_HIP(hipGetLastError());
}

}  // namespace nss

// HIP kernels of the superblock path, written for gfx950 (CDNA4, wave64).
//
//  * k_gemm_grouped<T>: one launch for every per-sector product of an H*v
//    phase (H_L.Psi, Psi.H_R^T, A.(Psi.B^T)), of rho = Psi Psi^T, or of a
//    block rotation U^T O U.  Each workgroup owns one 64x64 output tile and
//    accumulates all of its K-segments in MFMA registers (no atomics, fixed
//    order -> deterministic).  Operands are staged through LDS in 16-deep
//    K slices; four waves each compute a 32x32 quarter with 2x2 MFMA 16x16
//    blocks (v_mfma_f64_16x16x4f64 or v_mfma_f32_16x16x4f32).
//  * streaming Lanczos kernels: coalesced grid-stride loads, wave shuffles
//    then LDS, two fixed-order stages (deterministic dot/norm/GEMV).
//  * k_kron: QN-blocked Kronecker scatter used by block enlargement.
#include <algorithm>

#include "nss_internal.h"

namespace nss {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T>
__device__ __forceinline__ const T* resolve_c(int base, const void* abs, int64_t off, const Bases& b) {
  switch (base) {
    case BASE_X: return static_cast<const T*>(b.x) + off;
    case BASE_W: return static_cast<const T*>(b.w) + off;
    case BASE_Y: return static_cast<const T*>(b.y) + off;
    default: return static_cast<const T*>(abs);
  }
}

template <typename T>
__device__ __forceinline__ void gemm_tile(const GemmTile* __restrict__ tiles, const GemmSeg* __restrict__ segs,
                                          const Bases& bs) {
  constexpr int BM = GEMM_BM, BN = GEMM_BN, BK = 16;
  constexpr int PAD = 1;
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  using acc_t = typename Mfma<T>::acc_t;

  const GemmTile tl = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32;
  const int wn = (wave & 1) * 32;
  const int fr = lane & 15;  // fragment row/col within a 16x16 block
  const int fk = lane >> 4;  // fragment k within a 4-deep MFMA step

  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  for (int s = tl.seg0; s < tl.seg0 + tl.nseg; ++s) {
    const GemmSeg g = segs[s];
    const T* A = resolve_c<T>(g.abase, g.a, g.aoff, bs);
    const T* B = resolve_c<T>(g.bbase, g.b, g.boff, bs);
    const T alpha = static_cast<T>(g.alpha);
    const bool ta = g.flags & SEG_TA, tb = g.flags & SEG_TB;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
#pragma unroll
      for (int e = 0; e < (BM * BK) / 256; ++e) {
        const int idx = tid + e * 256;
        int i, k;
        if (!ta) { i = idx / BK; k = idx % BK; } else { k = idx / BM; i = idx % BM; }
        const int gi = tl.row0 + i, gk = k0 + k;
        T v = T(0);
        if (gi < tl.M && gk < g.K)
          v = ta ? A[(size_t)gk * g.lda + gi] : A[(size_t)gi * g.lda + gk];
        As[k][i] = alpha * v;
      }
#pragma unroll
      for (int e = 0; e < (BN * BK) / 256; ++e) {
        const int idx = tid + e * 256;
        int j, k;
        if (!tb) { k = idx / BN; j = idx % BN; } else { j = idx / BK; k = idx % BK; }
        const int gj = tl.col0 + j, gk = k0 + k;
        T v = T(0);
        if (gj < tl.N && gk < g.K)
          v = tb ? B[(size_t)gj * g.ldb + gk] : B[(size_t)gk * g.ldb + gj];
        Bs[k][j] = v;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const T a0 = As[kk + fk][wm + fr];
        const T a1 = As[kk + fk][wm + 16 + fr];
        const T b0 = Bs[kk + fk][wn + fr];
        const T b1 = Bs[kk + fk][wn + 16 + fr];
        acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
        acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
        acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
        acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
      }
      __syncthreads();
    }
  }
  T* C = const_cast<T*>(resolve_c<T>(tl.cbase, tl.c, tl.coff, bs));
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // accumulator item r of lane l holds D[row][l%16] with
        // row = l/16 + 4r (f64 16x16x4) or 4*(l/16) + r (f32 16x16x4)
        const int row = tl.row0 + wm + mi * 16 + (sizeof(T) == 8 ? fk + 4 * r : fk * 4 + r);
        const int col = tl.col0 + wn + ni * 16 + fr;
        if (row < tl.M && col < tl.N) {
          T* p = C + (size_t)row * tl.ldc + col;
          *p = tl.beta ? *p + acc[mi][ni][r] : acc[mi][ni][r];
        }
      }
}

// One kernel symbol per use so that traces separate them: the two H*v
// phases (W: Psi.B^T per term, Y: per-sector accumulation) and the generic
// grouped GEMM (rho, rotations).
#define NSS_GEMM_KERNEL(name, T)                                                             \
  __global__ __launch_bounds__(256) void name(const GemmTile* __restrict__ tiles,            \
                                              const GemmSeg* __restrict__ segs, Bases bs) { \
    gemm_tile<T>(tiles, segs, bs);                                                           \
  }
NSS_GEMM_KERNEL(k_hv_w_f64, double)
NSS_GEMM_KERNEL(k_hv_y_f64, double)
NSS_GEMM_KERNEL(k_gemm_f64, double)
NSS_GEMM_KERNEL(k_hv_w_f32, float)
NSS_GEMM_KERNEL(k_hv_y_f32, float)

void launch_gemm_f64(const GemmPlan& p, Bases b, hipStream_t s, int kind) {
  if (p.htiles.empty()) return;
  auto k = kind == GEMM_HV_W ? k_hv_w_f64 : kind == GEMM_HV_Y ? k_hv_y_f64 : k_gemm_f64;
  hipLaunchKernelGGL(k, dim3((unsigned)p.htiles.size()), dim3(256), 0, s, p.tiles.p, p.segs.p, b);
  NSS_HIP(hipGetLastError());
}

void launch_gemm_f32(const GemmPlan& p, Bases b, hipStream_t s, int kind) {
  if (p.htiles.empty()) return;
  auto k = kind == GEMM_HV_W ? k_hv_w_f32 : k_hv_y_f32;
  hipLaunchKernelGGL(k, dim3((unsigned)p.htiles.size()), dim3(256), 0, s, p.tiles.p, p.segs.p, b);
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ vectors
__global__ void k_fill_splitmix(double* v, int64_t n, uint64_t seed) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    v[i] = (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  }
}

__global__ void k_fill_const(double* v, int64_t n, double c) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = c;
}

__global__ void k_d2f(const double* x, float* y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

__global__ void k_f2d(const float* x, double* y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}

__global__ void k_div(const double* x, double* y, int64_t n, const double* den) {
  const double inv = 1.0 / den[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] * inv;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block-wide sum in fixed order (256 threads): wave partials then wave 0.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

constexpr int CHUNK = 4096;  // elements per reduction chunk (256 threads x 16)

int64_t gemv_t_chunks(int64_t n) { return (n + CHUNK - 1) / CHUNK; }

// partial[j*nch + c] = sum_{i in chunk c} V[j][i] * w[i]
__global__ __launch_bounds__(256) void k_gemv_t_partial(const double* __restrict__ V, int64_t ldv,
                                                        const double* __restrict__ w, int64_t n,
                                                        double* __restrict__ partial, int64_t nch) {
  __shared__ double sh[4];
  const int64_t c = blockIdx.x;
  const int j = blockIdx.y;
  const double* v = V + (size_t)j * ldv;
  const int64_t i0 = c * CHUNK, i1 = min<int64_t>(n, i0 + CHUNK);
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc += v[i] * w[i];
  const double r = block_sum256(acc, sh);
  if (threadIdx.x == 0) partial[(size_t)j * nch + c] = r;
}

__global__ __launch_bounds__(256) void k_reduce_rows(const double* __restrict__ partial, int64_t nch,
                                                     double* __restrict__ out, int do_sqrt) {
  __shared__ double sh[4];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int64_t c = threadIdx.x; c < nch; c += 256) acc += partial[(size_t)j * nch + c];
  const double r = block_sum256(acc, sh);
  if (threadIdx.x == 0) out[j] = do_sqrt ? sqrt(r) : r;
}

void launch_gemv_t(const double* V, int64_t ldv, int ncols, const double* w, int64_t n,
                   double* partial, double* h, hipStream_t s) {
  if (ncols <= 0) return;
  const int64_t nch = gemv_t_chunks(n);
  hipLaunchKernelGGL(k_gemv_t_partial, dim3((unsigned)nch, (unsigned)ncols), dim3(256), 0, s, V, ldv,
                     w, n, partial, nch);
  hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)ncols), dim3(256), 0, s, partial, nch, h, 0);
  NSS_HIP(hipGetLastError());
}

// w[i] -= sum_j V[j][i] h[j]   (overwrite: w[i] = sum_j ...)
__global__ __launch_bounds__(256) void k_gemv_n(const double* __restrict__ V, int64_t ldv, int ncols,
                                                const double* __restrict__ h, double* __restrict__ w,
                                                int64_t n, int overwrite) {
  extern __shared__ double hs[];
  for (int j = threadIdx.x; j < ncols; j += blockDim.x) hs[j] = h[j];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < ncols; ++j) acc += V[(size_t)j * ldv + i] * hs[j];
    w[i] = overwrite ? acc : w[i] - acc;
  }
}

void launch_gemv_n(const double* V, int64_t ldv, int ncols, const double* h, double* w, int64_t n,
                   int overwrite, hipStream_t s) {
  if (ncols <= 0 && !overwrite) return;
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_gemv_n, dim3((unsigned)nb), dim3(256), sizeof(double) * (size_t)std::max(ncols, 1), s,
                     V, ldv, ncols, h, w, n, overwrite);
  NSS_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_sumsq_partial(const double* __restrict__ x, int64_t n,
                                                       double* __restrict__ partial) {
  __shared__ double sh[4];
  const int64_t c = blockIdx.x;
  const int64_t i0 = c * CHUNK, i1 = min<int64_t>(n, i0 + CHUNK);
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc += x[i] * x[i];
  const double r = block_sum256(acc, sh);
  if (threadIdx.x == 0) partial[c] = r;
}

void launch_norm(const double* x, int64_t n, double* partial, double* out, hipStream_t s) {
  const int64_t nch = gemv_t_chunks(n);
  hipLaunchKernelGGL(k_sumsq_partial, dim3((unsigned)nch), dim3(256), 0, s, x, n, partial);
  hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, s, partial, nch, out, 1);
  NSS_HIP(hipGetLastError());
}

static unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 8192); }

void launch_fill_splitmix(double* v, int64_t n, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_splitmix, dim3(grid_for(n)), dim3(256), 0, s, v, n, seed);
  NSS_HIP(hipGetLastError());
}
void launch_fill_const(double* v, int64_t n, double c, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_const, dim3(grid_for(n)), dim3(256), 0, s, v, n, c);
  NSS_HIP(hipGetLastError());
}
void launch_d2f(const double* x, float* y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(grid_for(n)), dim3(256), 0, s, x, y, n);
  NSS_HIP(hipGetLastError());
}
void launch_f2d(const float* x, double* y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, dim3(grid_for(n)), dim3(256), 0, s, x, y, n);
  NSS_HIP(hipGetLastError());
}
void launch_div(const double* x, double* y, int64_t n, const double* den, hipStream_t s) {
  hipLaunchKernelGGL(k_div, dim3(grid_for(n)), dim3(256), 0, s, x, y, n, den);
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ symmetry check
__global__ __launch_bounds__(256) void k_asym_max(const double* __restrict__ a, const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ dim, double* __restrict__ out) {
  __shared__ double sh[256];
  const int b = blockIdx.x;
  const int d = dim[b];
  const double* m = a + off[b];
  double mx = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)d * d; e += 256) {
    const int i = (int)(e / d), j = (int)(e % d);
    if (j > i) mx = fmax(mx, fabs(m[(size_t)i * d + j] - m[(size_t)j * d + i]));
  }
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = sh[0];
}

void launch_asym_max(const double* a, const int64_t* off, const int32_t* d, int nblk, double* out,
                     hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_asym_max, dim3(nblk), dim3(256), 0, s, a, off, d, out);
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ Kronecker scatter
constexpr int KT = KRON_TILE;

__global__ __launch_bounds__(256) void k_kron(const KronTask* __restrict__ tasks,
                                              const KronContrib* __restrict__ cs) {
  const KronTask t = tasks[blockIdx.x];
  for (int e = threadIdx.x; e < KT * KT; e += 256) {
    const int r = t.row_t + e / KT, c = t.col_t + e % KT;
    if (r >= t.rows || c >= t.cols) continue;
    double v = 0.0;
    for (int q = t.c0; q < t.c0 + t.nc; ++q) {
      const KronContrib k = cs[q];
      double x;
      if (k.x == nullptr) x = (r == c) ? 1.0 : 0.0;
      else x = k.trans ? k.x[(size_t)c * k.ldx + r] : k.x[(size_t)r * k.ldx + c];
      v += k.scale * x;
    }
    t.dst[(size_t)r * t.ldd + c] = v;
  }
}

void launch_kron(const KronTask* tasks, int ntasks, const KronContrib* contribs, hipStream_t s) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_kron, dim3((unsigned)ntasks), dim3(256), 0, s, tasks, contribs);
  NSS_HIP(hipGetLastError());
}

}  // namespace nss

namespace nss {
// Lane-layout check of the two MFMA builtins used above: one wave, lane l
// supplies a[l] and b[l]; d[4*l + r] receives accumulator item r.
__global__ void k_probe_mfma(const double* a, const double* b, double* d, int prec) {
  const int l = threadIdx.x;
  if (prec == 0) {
    Mfma<double>::acc_t acc = {0, 0, 0, 0};
    acc = Mfma<double>::mma(a[l], b[l], acc);
    for (int r = 0; r < 4; ++r) d[4 * l + r] = acc[r];
  } else {
    Mfma<float>::acc_t acc = {0, 0, 0, 0};
    acc = Mfma<float>::mma((float)a[l], (float)b[l], acc);
    for (int r = 0; r < 4; ++r) d[4 * l + r] = (double)acc[r];
  }
}
}  // namespace nss

extern "C" int nss_probe_mfma(int prec, const double* a, const double* b, double* d) {
  using namespace nss;
  try {
    DevicePool* none = nullptr;
    (void)none;
    double *da = nullptr, *db = nullptr, *dd = nullptr;
    NSS_HIP(hipMalloc(&da, 64 * sizeof(double)));
    NSS_HIP(hipMalloc(&db, 64 * sizeof(double)));
    NSS_HIP(hipMalloc(&dd, 256 * sizeof(double)));
    NSS_HIP(hipMemcpy(da, a, 64 * sizeof(double), hipMemcpyHostToDevice));
    NSS_HIP(hipMemcpy(db, b, 64 * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_mfma, dim3(1), dim3(64), 0, 0, da, db, dd, prec);
    NSS_HIP(hipGetLastError());
    NSS_HIP(hipMemcpy(d, dd, 256 * sizeof(double), hipMemcpyDeviceToHost));
    NSS_HIP(hipFree(da));
    NSS_HIP(hipFree(db));
    NSS_HIP(hipFree(dd));
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return NSS_OK;
}

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
#include <cfloat>

#include <hip/hip_ext.h>

#include "nss_internal.h"

#include <type_traits>

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

__device__ __forceinline__ double wave_sum(double v);

template <typename T>
struct SliceRegs {  // one K-slice of A and B held in registers while the previous slice computes
  T a[8], b[8];
};

// Loads the K-slice [k0, k0+32) of segment g into registers (zero-padded).
template <typename T>
__device__ __forceinline__ void load_slice(const GemmSeg& g, const GemmTile& tl, const Bases& bs, int k0, int tid,
                                           SliceRegs<T>& r) {
  constexpr int BM = GEMM_BM, BN = GEMM_BN, BK = 32;
  const T* A = resolve_c<T>(g.abase, g.a, g.aoff, bs);
  const T* B = resolve_c<T>(g.bbase, g.b, g.boff, bs);
  const bool ta = g.flags & SEG_TA, tb = g.flags & SEG_TB;
#pragma unroll
  for (int e = 0; e < (BM * BK) / 256; ++e) {
    const int idx = tid + e * 256;
    int i, k;
    if (!ta) { i = idx / BK; k = idx % BK; } else { k = idx / BM; i = idx % BM; }
    const int gi = tl.row0 + i, gk = k0 + k;
    T v = T(0);
    if (gi < tl.M && gk < g.K) v = ta ? A[(size_t)gk * g.lda + gi] : A[(size_t)gi * g.lda + gk];
    r.a[e] = v;
  }
#pragma unroll
  for (int e = 0; e < (BN * BK) / 256; ++e) {
    const int idx = tid + e * 256;
    int j, k;
    if (!tb) { k = idx / BN; j = idx % BN; } else { j = idx / BK; k = idx % BK; }
    const int gj = tl.col0 + j, gk = k0 + k;
    T v = T(0);
    if (gj < tl.N && gk < g.K) v = tb ? B[(size_t)gj * g.ldb + gk] : B[(size_t)gk * g.ldb + gj];
    r.b[e] = v;
  }
}

template <typename T>
__device__ __forceinline__ void gemm_tile(const GemmTile* __restrict__ tiles, const GemmSeg* __restrict__ segs,
                                          const Bases& bs, int tix) {
  constexpr int BM = GEMM_BM, BN = GEMM_BN, BK = 32;
  constexpr int PAD = 1;
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  using acc_t = typename Mfma<T>::acc_t;

  const GemmTile tl = tiles[0];
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

  const int send = tl.seg0 + tl.nseg;
  int s = tl.seg0, k0 = 0;
  GemmSeg g;
  SliceRegs<T> r;
  if (s < send) {
    g = segs[s];
    load_slice<T>(g, tl, bs, 0, tid, r);
  }
  while (s < send) {
    // registers -> LDS (alpha applied to A), layout As[k][i], Bs[k][j]
    const T alpha = static_cast<T>(g.alpha);
    const bool ta = g.flags & SEG_TA, tb = g.flags & SEG_TB;
#pragma unroll
    for (int e = 0; e < (BM * BK) / 256; ++e) {
      const int idx = tid + e * 256;
      int i, k;
      if (!ta) { i = idx / BK; k = idx % BK; } else { k = idx / BM; i = idx % BM; }
      As[k][i] = alpha * r.a[e];
    }
#pragma unroll
    for (int e = 0; e < (BN * BK) / 256; ++e) {
      const int idx = tid + e * 256;
      int j, k;
      if (!tb) { k = idx / BN; j = idx % BN; } else { j = idx / BK; k = idx % BK; }
      Bs[k][j] = r.b[e];
    }
    __syncthreads();
    // next slice (possibly of the next segment): loads in flight during the MFMAs
    int ns = s, nk0 = k0 + BK;
    if (nk0 >= g.K) { ns = s + 1; nk0 = 0; }
    if (ns < send) {
      if (ns != s) g = segs[ns];
      load_slice<T>(g, tl, bs, nk0, tid, r);
    }
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
    s = ns;
    k0 = nk0;
  }
  T* C = const_cast<T*>(resolve_c<T>(tl.cbase, tl.c, tl.coff, bs));
  const T* X = static_cast<const T*>(bs.x) + tl.coff;  // read only with bs.dot (Y and x share offsets)
  double dsum = 0.0;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // accumulator item q of lane l holds D[row][l%16] with
        // row = l/16 + 4q (f64 16x16x4) or 4*(l/16) + q (f32 16x16x4)
        const int row = tl.row0 + wm + mi * 16 + (sizeof(T) == 8 ? fk + 4 * q : fk * 4 + q);
        const int col = tl.col0 + wn + ni * 16 + fr;
        if (row < tl.M && col < tl.N) {
          T* p = C + (size_t)row * tl.ldc + col;
          const T out = tl.beta ? *p + acc[mi][ni][q] : acc[mi][ni][q];
          *p = out;
          if (bs.dot) dsum += (double)out * (double)X[(size_t)row * tl.ldc + col];
        }
      }
  if (bs.dot) {  // each wave's share of x^T y (no block barrier)
    const double r = wave_sum(dsum);
    if ((threadIdx.x & 63) == 0) bs.dot[4 * tix + (threadIdx.x >> 6)] = r;
  }
}

// ---------------------------------------------------------------------------
// 32x32 output tile per workgroup; the tile's K chain (the concatenation of its
// segments) is cut into 16-deep chunks dealt round-robin to the four waves, so
// a tile runs on four SIMDs at once and small sectors waste little MFMA work.
// Each wave loads its operand fragments straight from global memory (no
// operand is shared between waves) one chunk ahead of the MFMAs, and the four
// partial tiles are summed through LDS in a fixed order (deterministic).
// Fragment layout, MFMA 16x16x4 step r of a chunk at k0: lane l supplies
// A[i = l%16][k0 + 4r + l/16] and B[k0 + 4r + l/16][j = l%16].
template <typename T>
struct Frag32 {
  T a[2][4], b[2][4];
};

// Operand pointers in the global address space: loads through them compile
// to global_load (counted by vmcnt only) rather than flat_load, whose
// lgkmcnt share would make every scalar descriptor wait drain them.
template <typename T>
using gptr = const __attribute__((address_space(1))) T*;

// Per-lane view of the current segment: base of the lane's two A rows and
// two B columns, and the element stride along k of each operand.
template <typename T>
struct SegView32 {
  gptr<T> pa[2], pb[2];
  int ska, skb;
  int K;
  T alpha;
};

template <typename T>
__device__ __forceinline__ void seg_view32(const GemmSeg& g, const Bases& bs, const GemmTile& tl, int fr,
                                           SegView32<T>& v) {
  const T* A = resolve_c<T>(g.abase, g.a, g.aoff, bs);
  const T* B = resolve_c<T>(g.bbase, g.b, g.boff, bs);
  const bool ta = g.flags & SEG_TA, tb = g.flags & SEG_TB;
  const int64_t sai = ta ? 1 : (int64_t)g.lda, sbj = tb ? (int64_t)g.ldb : 1;
  v.ska = ta ? g.lda : 1;
  v.skb = tb ? 1 : g.ldb;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    // clamped rows/columns: always in bounds; rows >= M and columns >= N
    // only reach output elements that are never stored
    const int64_t i = min(tl.row0 + 16 * m + fr, tl.M - 1);
    const int64_t j = min(tl.col0 + 16 * m + fr, tl.N - 1);
    v.pa[m] = (gptr<T>)(A + i * sai);
    v.pb[m] = (gptr<T>)(B + j * sbj);
  }
  v.K = g.K;
  v.alpha = static_cast<T>(g.alpha);
}

// Raw fragment loads from clamped addresses: no branch and no use of the
// loaded values, so a whole chunk of loads stays in flight.  The K bound and
// alpha are applied to A at use (mma_chunk32).
template <typename T>
__device__ __forceinline__ void load_frag32(const SegView32<T>& v, int k0, int fk, Frag32<T>& f) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int k = min(k0 + 4 * r + fk, v.K - 1);
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      f.a[m][r] = v.pa[m][k * v.ska];
      f.b[m][r] = v.pb[m][k * v.skb];
    }
  }
}

template <typename T>
__device__ __forceinline__ void mma_chunk32(const Frag32<T>& f, int k0, int K, T alpha, int fk,
                                            typename Mfma<T>::acc_t (&acc)[2][2]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const T sc = (k0 + 4 * r + fk < K) ? alpha : T(0);
    const T a0 = sc * f.a[0][r], a1 = sc * f.a[1][r];
    acc[0][0] = Mfma<T>::mma(a0, f.b[0][r], acc[0][0]);
    acc[0][1] = Mfma<T>::mma(a0, f.b[1][r], acc[0][1]);
    acc[1][0] = Mfma<T>::mma(a1, f.b[0][r], acc[1][0]);
    acc[1][1] = Mfma<T>::mma(a1, f.b[1][r], acc[1][1]);
  }
}

// Maps the launch index to a tile so that each XCD (workgroups are dealt to
// the 8 XCDs round-robin) receives a contiguous run of the tile list: the
// tiles of one sector then share that XCD's L2.  Returns -1 for padding.
__device__ __forceinline__ int xcd_tile(int ntiles) {
  const int per = (ntiles + 7) >> 3;
  const int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  return t < ntiles ? t : -1;
}

// Fragment buffers of the 32x32 kernel (three measured slower: the chains
// of these launches are a few chunks per wave).
#define NSS_NBUF 2

template <typename T, int NBUF, bool SKIPLOAD>
__device__ __forceinline__ void gemm_tile32(const GemmTile* __restrict__ tiles, const GemmSeg* __restrict__ segs,
                                            int t, const Bases& bs, T (*red)[32][33], int tix = -1,
                                            bool store = true) {
  // tix: this thread's index within the tile's 256 threads (threadIdx.x when
  // a workgroup runs one tile); store = false computes without writing
  if (tix < 0) tix = threadIdx.x;
  using acc_t = typename Mfma<T>::acc_t;
  const GemmTile tl = tiles[t];
  const int lane = tix & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tix >> 6);  // wave-uniform
  const int fr = lane & 15, fk = lane >> 4;

  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  // cursor (segment s, chunk ch): this wave's chunks are the global chunk
  // indices congruent to `wave` mod 4, counted across the segment chain.
  const int send = tl.seg0 + tl.nseg;
  int s = tl.seg0, ch = wave - 4, nch = 0;
  SegView32<T> v;
  // Advances the cursor to this wave's next chunk: (k0, K) of the chunk go
  // to m, its alpha to alpha.  When the chain is exhausted it returns false,
  // sets alpha = 0 and leaves m and the segment view as they were, so the
  // following load still reads valid (clamped) addresses and the MFMA pass
  // over it adds exactly zero.
  auto advance = [&](int (&m)[2], T& alpha) -> bool {
    if (s >= send) {  // exhausted for good (ch and nch are stale from here on)
      alpha = T(0);
      return false;
    }
    ch += 4;
    if (ch >= nch) {
      GemmSeg g;
      bool moved = false;
      while (s < send) {
        if (moved || nch > 0) {
          ch -= nch;
          ++s;
          if (s >= send) break;
        }
        moved = true;
        g = segs[s];
        nch = (g.K + 15) >> 4;
        if (ch < nch) break;
      }
      if (s >= send) {
        alpha = T(0);
        return false;
      }
      seg_view32<T>(g, bs, tl, fr, v);
    }
    m[0] = ch << 4;
    m[1] = v.K;
    alpha = v.alpha;
    return true;
  };
  // NBUF fragment buffers in rotation, each refilled right after its chunk
  // has gone to the MFMA pipe, so loads run NBUF-1 chunks ahead.  The MFMAs
  // of an exhausted buffer are skipped (a wave-uniform branch).  Buffers take
  // the wave's chunks in order, so once buffer 0 has found the chain
  // exhausted, the chunks still held by the others are computed in the same
  // round.  Refills of exhausted buffers:
  //  - SKIPLOAD = false: issued anyway (they re-read valid clamped
  //    addresses), which keeps the load order the same on every path, so the
  //    wait before each buffer's MFMAs leaves the later buffers' loads in
  //    flight.  For long K chains (Y).
  //  - SKIPLOAD = true: skipped, which spares short chains (W: one or two
  //    chunks per wave) a last round of loads; the compiler's waits then
  //    drain every load in flight before each buffer's MFMAs.
  int m[NBUF][2];
  T al[NBUF];
#pragma unroll
  for (int b = 0; b < NBUF; ++b) {
    m[b][0] = 0;
    m[b][1] = 1;
    al[b] = T(0);
  }
  if (advance(m[0], al[0])) {
    Frag32<T> f[NBUF];
    bool live[NBUF];
    live[0] = true;
    load_frag32<T>(v, m[0][0], fk, f[0]);
#pragma unroll
    for (int b = 1; b < NBUF; ++b) {
      live[b] = advance(m[b], al[b]);
      if (!SKIPLOAD || live[b]) load_frag32<T>(v, m[b][0], fk, f[b]);
    }
    do {
#pragma unroll
      for (int b = 0; b < NBUF; ++b) {
        if (live[b]) mma_chunk32<T>(f[b], m[b][0], m[b][1], al[b], fk, acc);
        live[b] = advance(m[b], al[b]);
        if (!SKIPLOAD || live[b]) load_frag32<T>(v, m[b][0], fk, f[b]);
      }
    } while (live[0]);
  }

#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // accumulator item q of lane l holds D[row][l%16] with
        // row = l/16 + 4q (f64 16x16x4) or 4*(l/16) + q (f32 16x16x4)
        const int row = mi * 16 + (sizeof(T) == 8 ? fk + 4 * q : fk * 4 + q);
        red[wave][row][ni * 16 + fr] = acc[mi][ni][q];
      }
  __syncthreads();
  T* C = const_cast<T*>(resolve_c<T>(tl.cbase, tl.c, tl.coff, bs));
  const T* X = static_cast<const T*>(bs.x) + tl.coff;  // read only with bs.dot (Y and x share offsets)
  double dsum = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tix + 256 * u;
    const int row = e >> 5, col = e & 31;
    const int gi = tl.row0 + row, gj = tl.col0 + col;
    if (store && gi < tl.M && gj < tl.N) {
      const T v = (red[0][row][col] + red[1][row][col]) + (red[2][row][col] + red[3][row][col]);
      T* p = C + (size_t)gi * tl.ldc + gj;
      const T out = tl.beta ? *p + v : v;
      *p = out;
      if (bs.dot) dsum += (double)out * (double)X[(size_t)gi * tl.ldc + gj];
    }
  }
  if (bs.dot) {  // each wave's share of x^T y (no block barrier)
    const double r = wave_sum(dsum);
    if (lane == 0 && store) bs.dot[4 * t + wave] = r;
  }
}

// ---------------------------------------------------------------------------
// Large-output path: 128x128 output tile per workgroup of GB_NW = 4 waves in
// a 2x2 grid, each wave a 64x64 block (4x4 MFMA 16x16 blocks held in
// registers for the whole K chain).  Operands go global -> registers -> LDS
// in K slices of BK (16 for FP64, 32 for FP32), two LDS stages: while slice
// c computes, slice c+1 goes from registers to the free stage and slice
// c+2's loads are in flight (L2 hits, a few hundred cycles, against a
// slice's thousands of MFMA cycles); one barrier per slice, two workgroups
// per CU.  Every load address is a precomputed per-thread offset plus a
// per-slice scalar, so a slice issues its loads and stores with next to no
// vector arithmetic.
// Every operand layout is stored the same way in LDS: element (k, r) of a
// stage at k*P + (r ^ bsw(k)) (row pitch P, below).  Rows past M/N and k past
// K load clamped (valid) addresses; k past K is zeroed and alpha applied as
// the slice is written to LDS.
// A work item is one part [c0, c1) of a tile's slice chain (the slices of
// its segments in order).  A tile cut into several parts writes every
// part's partial tile to part[] (lane-major per register, coalesced); the
// workgroup that completes the tile (a per-tile arrival counter) sums the
// partials in part order, so the result does not depend on timing, then
// applies beta and the dot epilogue and resets the counter for the next
// launch.
constexpr int GB = 128, GB_NW = 4, GB_THREADS = 64 * GB_NW;
constexpr int GB_WC = GB_NW / 2;          // wave grid 2 x GB_WC
constexpr int GB_WN = GB / GB_WC;         // wave block 64 x GB_WN
constexpr int GB_NB = GB_WN / 16;         // MFMA blocks per wave along N
constexpr int GB_REGS = 4 * GB_NB * 4;    // accumulator elements per lane

// LDS stage layout: row k (one k, 128 rows of the output) starts at
// k*P + H*(k&1), element r of it at that start + r (rows do not overlap:
// H <= P - 128).  With P = 138 and H = 6 (FP64) or 8 (FP32) an MFMA fragment
// read (16 rows at each of 4 consecutive k) meets no bank twice in either
// half-wave, and the stores of both global layouts take the minimum number
// of passes (checked exhaustively over the lane patterns of the three access
// kinds).  Rows are not permuted, so every address of a slice is a
// per-thread base plus a compile-time offset.
template <typename T>
struct BigCfg {
  static constexpr int BK = sizeof(T) == 8 ? 16 : 32;  // K slice
  static constexpr int EPT = GB * BK / GB_THREADS;     // elements per thread and operand per slice
  static constexpr int RPI = 64 / BK;                  // rows per wave load when contiguous along k
  static constexpr int P = 138;                        // LDS row pitch (elements)
  static constexpr int H = sizeof(T) == 8 ? 6 : 8;     // extra start offset of odd rows
  static_assert(H >= 0 && H <= P - GB, "LDS rows overlap");
  static constexpr int STAGE = BK * P + H;             // elements of one operand in one stage
  static constexpr int KC_STEP = RPI * GB_NW;          // row step between a thread's elements (k-contiguous)
  static constexpr int RC_KSTEP = GB_THREADS / GB;     // k step between a thread's elements (row-contiguous)
};

// LDS of the t128 launches: two stages of both operands, or the reduction
// buffers of the GB_NW/4 tiles of 32x32 a workgroup runs on the small path
constexpr int GB_LDS_BYTES = std::max<int>({(GB_NW / 4) * 4 * 32 * 33 * 8, 4 * BigCfg<double>::STAGE * 8,
                                            4 * BigCfg<float>::STAGE * 4});

// LDS pointers typed as such: a generic pointer to LDS compiles to flat
// accesses, which share lgkmcnt with the global loads in flight.
template <typename T>
using lptr = __attribute__((address_space(3))) T*;

template <typename T>
__device__ __forceinline__ int big_rowstart(int k) {
  return k * BigCfg<T>::P + BigCfg<T>::H * (k & 1);
}

// Element u of this thread in a slice, by the operand's layout.  Contiguous
// along k: lane l of wave w takes k = l / RPI of row l % RPI + RPI w + KC_STEP
// u (each load instruction reads RPI rows of BK consecutive k).  Contiguous
// along the row index: row tid % 128, k tid/128 + RC_KSTEP u.
template <typename T>
__device__ __forceinline__ int big_row(bool kc, int tid, int u) {
  constexpr int RPI = BigCfg<T>::RPI;
  return kc ? (tid & 63) % RPI + RPI * (tid >> 6) + BigCfg<T>::KC_STEP * u : tid % GB;
}
template <typename T>
__device__ __forceinline__ int big_k(bool kc, int tid, int u) {
  return kc ? (tid & 63) / BigCfg<T>::RPI : tid / GB + BigCfg<T>::RC_KSTEP * u;
}

// One operand of the current segment: element (r, k) at p + (r*rs + k*ks)
// elements.  In an interior tile (all 128 rows valid) element u of this
// thread sits at byte offset off0 + u*ustep from p + k0*ks*sizeof(T); edge
// tiles and the segment's last, partial slice clamp rows and k instead.
// (Element offsets stay below 2^31 and byte offsets below 2^32: the host
// sends larger operands elsewhere.)
template <typename T>
struct BigOp {
  gptr<char> p;
  int rs, ks;
  int r0, rmax;  // tile's first row; last valid row (relative)
  bool kc;       // contiguous along k
  uint32_t off0, ustep;
};

template <typename T>
__device__ __forceinline__ void big_op(const T* base, int rs, int ks, bool kc, int r0, int rmax, int tid,
                                       BigOp<T>& o) {
  o.p = (gptr<char>)base;
  o.rs = rs;
  o.ks = ks;
  o.kc = kc;
  o.r0 = r0;
  o.rmax = rmax;
  o.off0 = (uint32_t)((big_row<T>(kc, tid, 0) + r0) * rs + big_k<T>(kc, tid, 0) * ks) * (uint32_t)sizeof(T);
  o.ustep = (uint32_t)(kc ? BigCfg<T>::KC_STEP * rs : BigCfg<T>::RC_KSTEP * ks) * (uint32_t)sizeof(T);
}

template <typename T>
__device__ __forceinline__ void big_view(const GemmSeg& g, const Bases& bs, const GemmTile& tl, int tid,
                                         BigOp<T>& A, BigOp<T>& B) {
  const bool ta = g.flags & SEG_TA, tb = g.flags & SEG_TB;
  big_op<T>(resolve_c<T>(g.abase, g.a, g.aoff, bs), ta ? 1 : g.lda, ta ? g.lda : 1, !ta, tl.row0,
            tl.M - 1 - tl.row0, tid, A);
  // B's "row" is the output column j
  big_op<T>(resolve_c<T>(g.bbase, g.b, g.boff, bs), tb ? g.ldb : 1, tb ? 1 : g.ldb, tb, tl.col0,
            tl.N - 1 - tl.col0, tid, B);
}

template <typename T>
struct BigSlice {  // one slice of both operands in registers, with what its LDS store needs
  T a[BigCfg<T>::EPT], b[BigCfg<T>::EPT];
  bool akc, bkc;
  int kval;  // valid k of the slice
  T alpha;
};

// Slice [k0, k0 + BK) of one operand into registers.
template <typename T>
__device__ __forceinline__ void big_load(const BigOp<T>& o, int k0, int K, int tid, T (&v)[BigCfg<T>::EPT]) {
  if (k0 + BigCfg<T>::BK <= K && o.rmax >= GB - 1) {
    const gptr<char> base = o.p + (uint32_t)(k0 * o.ks) * (uint32_t)sizeof(T);
#pragma unroll
    for (int u = 0; u < BigCfg<T>::EPT; ++u) v[u] = *(gptr<T>)(base + (o.off0 + u * o.ustep));
  } else {
    // clamped rows and k: valid addresses; the store zeroes k past K, and
    // rows past the tile's edge only reach outputs that are never stored
#pragma unroll
    for (int u = 0; u < BigCfg<T>::EPT; ++u) {
      const int rr = min(big_row<T>(o.kc, tid, u), o.rmax) + o.r0;
      const int kk = min(k0 + big_k<T>(o.kc, tid, u), K - 1);
      v[u] = *(gptr<T>)(o.p + (uint32_t)(rr * o.rs + kk * o.ks) * (uint32_t)sizeof(T));
    }
  }
}

// Registers to an LDS stage operand S; alpha and the K bound applied only
// where the slice needs them.
template <typename T>
__device__ __forceinline__ void big_store(lptr<T> S, const T (&v)[BigCfg<T>::EPT], bool kc, int kval, T scale,
                                          int tid) {
  constexpr int EPT = BigCfg<T>::EPT, KS = BigCfg<T>::KC_STEP, RS = BigCfg<T>::RC_KSTEP * BigCfg<T>::P;
  const int k0t = big_k<T>(kc, tid, 0);
  lptr<T> d = S + big_rowstart<T>(k0t) + big_row<T>(kc, tid, 0);
  if (kval >= BigCfg<T>::BK && scale == T(1)) {
    if (kc) {
#pragma unroll
      for (int u = 0; u < EPT; ++u) d[u * KS] = v[u];
    } else {
#pragma unroll
      for (int u = 0; u < EPT; ++u) d[u * RS] = v[u];
    }
  } else if (kc) {
#pragma unroll
    for (int u = 0; u < EPT; ++u) d[u * KS] = k0t < kval ? scale * v[u] : T(0);
  } else {
#pragma unroll
    for (int u = 0; u < EPT; ++u) d[u * RS] = k0t + BigCfg<T>::RC_KSTEP * u < kval ? scale * v[u] : T(0);
  }
}

template <typename T>
__device__ __forceinline__ void big_mma(lptr<T> As, lptr<T> Bs, int wm, int wn, int fr, int fk,
                                        typename Mfma<T>::acc_t (&acc)[4][GB_NB]) {
  constexpr int P = BigCfg<T>::P;
  // k = s + fk with s a multiple of 4: every k-step reads at this lane's
  // bases plus s*P
  lptr<T> Ak = As + big_rowstart<T>(fk) + wm + fr;
  lptr<T> Bk = Bs + big_rowstart<T>(fk) + wn + fr;
#pragma unroll
  for (int s = 0; s < BigCfg<T>::BK; s += 4) {
    T a[4], b[GB_NB];
#pragma unroll
    for (int m = 0; m < 4; ++m) a[m] = Ak[s * P + 16 * m];
#pragma unroll
    for (int n = 0; n < GB_NB; ++n) b[n] = Bk[s * P + 16 * n];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < GB_NB; ++n) acc[m][n] = Mfma<T>::mma(a[m], b[n], acc[m][n]);
  }
}

template <typename T>
__device__ __forceinline__ void gemm_big(const GemmTile* __restrict__ tiles, const GemmSeg* __restrict__ segs,
                                         const BigItem& it, const Bases& bs, void* part, unsigned* cnt,
                                         lptr<T> smem) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int BK = BigCfg<T>::BK, STG = BigCfg<T>::STAGE;
  const GemmTile tl = tiles[it.tile];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wave / GB_WC) * 64, wn = (wave % GB_WC) * GB_WN;
  const int fr = lane & 15, fk = lane >> 4;
  // stage st: A at smem + 2 st STG, B after it
  acc_t acc[4][GB_NB];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < GB_NB; ++n) acc[m][n] = acc_t{0, 0, 0, 0};
  const int c0 = it.c0, c1 = it.c1;
  if (c0 < c1) {
    // load cursor: segment s, slice `in` within it
    const int send = tl.seg0 + tl.nseg;
    int s = tl.seg0, in = c0;
    GemmSeg g = segs[s];
    for (;;) {
      const int ns = (g.K + BK - 1) / BK;
      if (in < ns) break;
      in -= ns;
      g = segs[++s];
    }
    BigOp<T> oa, ob;
    big_view<T>(g, bs, tl, tid, oa, ob);
    BigSlice<T> rg;
    // the slice at the cursor into registers (a cursor past the chain's end
    // reads clamped, valid addresses: the loads never depend on a branch),
    // then the cursor moves on
    auto issue = [&]() {
      const int k0 = in * BK;
      big_load<T>(oa, k0, g.K, tid, rg.a);
      big_load<T>(ob, k0, g.K, tid, rg.b);
      rg.akc = oa.kc;
      rg.bkc = ob.kc;
      rg.kval = g.K - k0;
      rg.alpha = static_cast<T>(g.alpha);
      ++in;
      if (in * BK >= g.K && s + 1 < send) {
        g = segs[++s];
        in = 0;
        big_view<T>(g, bs, tl, tid, oa, ob);
      }
    };
    auto store = [&](int st) {
      big_store<T>(smem + 2 * st * STG, rg.a, rg.akc, rg.kval, rg.alpha, tid);
      big_store<T>(smem + (2 * st + 1) * STG, rg.b, rg.bkc, rg.kval, T(1), tid);
    };
    issue();  // slice c0 to stage 0
    store(0);
    issue();  // slice c0 + 1
    __syncthreads();
    // iteration c: slice c+1 from registers to the free stage, slice c+2's
    // loads in flight behind slice c's MFMAs, one barrier (unrolled by two,
    // so that the stage offsets are constants)
    auto step = [&](int c, int st) {
      if (c + 1 < c1) store(st ^ 1);
      issue();
      big_mma<T>(smem + 2 * st * STG, smem + (2 * st + 1) * STG, wm, wn, fr, fk, acc);
      __syncthreads();
    };
    for (int c = c0; c < c1; c += 2) {
      step(c, 0);
      if (c + 1 < c1) step(c + 1, 1);
    }
  }
  __shared__ double dred[GB_NW];
  if (it.nsplit > 1) {
    // partial tile: part[(pslot * NW + wave) * 64 * REGS + reg * 64 + lane]
    T* P = static_cast<T*>(part);
    T* mine = P + ((int64_t)it.pslot * GB_NW + wave) * (64 * GB_REGS);
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < GB_NB; ++n)
#pragma unroll
        for (int q = 0; q < 4; ++q) mine[((m * GB_NB + n) * 4 + q) * 64 + lane] = acc[m][n][q];
    // every wave's stores done, the workgroup meets, one thread releases them
    // at agent scope with its arrival; the last arrival acquires
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    __shared__ int last_sh;
    if (tid == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const bool last = __hip_atomic_fetch_add(cnt + it.tile, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                        (unsigned)(it.nsplit - 1);
      if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      last_sh = last;
    }
    __syncthreads();
    if (!last_sh) return;
    // the sum in part order (the tile's parts hold slots pslot - part .. )
    const int64_t slot0 = it.pslot - it.part;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < GB_NB; ++n) {
        acc_t sum = acc_t{0, 0, 0, 0};
        for (int p = 0; p < it.nsplit; ++p) {
          const T* src = P + ((slot0 + p) * GB_NW + wave) * (64 * GB_REGS) + (m * GB_NB + n) * 256 + lane;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const T x = p == it.part ? acc[m][n][q] : src[q * 64];
            sum[q] = p == 0 ? x : sum[q] + x;
          }
        }
        acc[m][n] = sum;
      }
    if (tid == 0) cnt[it.tile] = 0u;
  }
  T* C = const_cast<T*>(resolve_c<T>(tl.cbase, tl.c, tl.coff, bs));
  const T* X = static_cast<const T*>(bs.x) + tl.coff;  // read only with bs.dot (Y and x share offsets)
  double dsum = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < GB_NB; ++n)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // accumulator item q of lane l holds D[row][l%16] with
        // row = l/16 + 4q (f64 16x16x4) or 4*(l/16) + q (f32 16x16x4)
        const int row = tl.row0 + wm + m * 16 + (sizeof(T) == 8 ? fk + 4 * q : fk * 4 + q);
        const int col = tl.col0 + wn + n * 16 + fr;
        if (row < tl.M && col < tl.N) {
          T* p = C + (size_t)row * tl.ldc + col;
          const T out = tl.beta ? *p + acc[m][n][q] : acc[m][n][q];
          *p = out;
          if (bs.dot) dsum += (double)out * (double)X[(size_t)row * tl.ldc + col];
        }
      }
  if (bs.dot) {  // four partials of x^T y per tile (the waves' sums in fixed groups of GB_NW/4)
    const double r = wave_sum(dsum);
    if (lane == 0) dred[wave] = r;
    __syncthreads();
    if (tid < 4) {
      double v = 0.0;
      for (int w = 0; w < GB_NW / 4; ++w) v += dred[tid * (GB_NW / 4) + w];
      bs.dot[4 * it.tile + tid] = v;
    }
  }
}

// Workgroup b of a launch with nb work items dealt so that each XCD
// (workgroups go to the 8 XCDs round-robin) receives a contiguous run of
// the item list; -1 for padding.
__device__ __forceinline__ int xcd_item(int b, int nitems) {
  const int per = (nitems + 7) >> 3;
  const int t = (b & 7) * per + (b >> 3);
  return t < nitems ? t : -1;
}

// One kernel symbol per use so that traces separate them: the two H*v
// phases (W: Psi.B^T per term, Y: per-sector accumulation) and the generic
// grouped GEMM (rho, rotations); the _t64 symbols are the 64x64 LDS-staged
// variant of each, the _t128 symbols the launch with large outputs: its
// first workgroups run the 128x128 work items, the rest the 32x32 tiles of
// the plan's other outputs, all sharing one LDS buffer.
#define NSS_GEMM_KERNEL(name, T, SKIPLOAD)                                                              \
  __global__ __launch_bounds__(256) void name(const GemmTile* __restrict__ tiles,                       \
                                              const GemmSeg* __restrict__ segs, int ntiles, Bases bs) { \
    __shared__ T red[4][32][33];                                                                        \
    const int t = xcd_tile(ntiles);                                                                     \
    if (t >= 0) gemm_tile32<T, NSS_NBUF, SKIPLOAD>(tiles, segs, t, bs, red);                            \
  }                                                                                                     \
  __global__ __launch_bounds__(GB_THREADS, 2) void name##_t128(const GemmTile* __restrict__ tiles,      \
                                                               const GemmSeg* __restrict__ segs,        \
                                                               int ntiles, Bases bs, BigArgs ba) {      \
    __shared__ __attribute__((aligned(16))) char smem[GB_LDS_BYTES];                                    \
    const int nb8 = 8 * ((ba.nitems + 7) / 8);                                                          \
    if ((int)blockIdx.x < nb8) {                                                                        \
      const int i = xcd_item(blockIdx.x, ba.nitems);                                                    \
      if (i >= 0) gemm_big<T>(tiles, segs, ba.items[i], bs, ba.part, ba.cnt, (lptr<T>)(smem));          \
    } else {                                                                                            \
      /* GB_NW / 4 tiles of 32x32 per workgroup, four waves each; a missing */                          \
      /* last tile repeats the one before without storing, so every wave */                            \
      /* meets the same barriers */                                                                     \
      const int nsm = ntiles - ba.nbig, tpw = GB_NW / 4;                                                \
      const int pair = xcd_item(blockIdx.x - nb8, (nsm + tpw - 1) / tpw);                               \
      if (pair >= 0) {                                                                                  \
        const int h = threadIdx.x >> 8, t = pair * tpw + h;                                             \
        gemm_tile32<T, 1, SKIPLOAD>(tiles, segs, ba.nbig + min(t, nsm - 1), bs,                         \
                                           reinterpret_cast<T(*)[32][33]>(smem) + 4 * h,                \
                                           threadIdx.x & 255, t < nsm);                                 \
      }                                                                                                 \
    }                                                                                                   \
  }                                                                                                     \
  __global__ __launch_bounds__(256) void name##_t64(const GemmTile* __restrict__ tiles,                 \
                                                    const GemmSeg* __restrict__ segs, int ntiles,       \
                                                    Bases bs) {                                         \
    const int t = xcd_tile(ntiles);                                                                     \
    if (t >= 0) gemm_tile<T>(tiles + t, segs, bs, t);                                                   \
  }
static_assert((GB_NW / 4) * 4 * 32 * 33 * sizeof(double) <= GB_LDS_BYTES, "32x32 reduction buffers exceed the LDS");
static_assert(4 * BigCfg<double>::STAGE * sizeof(double) <= GB_LDS_BYTES &&
                  4 * BigCfg<float>::STAGE * sizeof(float) <= GB_LDS_BYTES,
              "two stages of both operands exceed the LDS");
static_assert(GB_NW % 4 == 0 && GB_NW <= 8, "t128 workgroups run GB_NW/4 small tiles of four waves");
// W and the generic GEMM (rho, rotations) have short K chains; Y's chain
// runs over every term of its sector.
NSS_GEMM_KERNEL(k_hv_w_f64, double, true)
NSS_GEMM_KERNEL(k_hv_y_f64, double, false)

NSS_GEMM_KERNEL(k_gemm_f64, double, true)
NSS_GEMM_KERNEL(k_hv_w_f32, float, true)
NSS_GEMM_KERNEL(k_hv_y_f32, float, false)

typedef void (*gemm_kernel_t)(const GemmTile*, const GemmSeg*, int, Bases);
typedef void (*gemm128_kernel_t)(const GemmTile*, const GemmSeg*, int, Bases, BigArgs);

// With start/stop events the launch records them itself (the dispatch's own
// begin and end timestamps, hipExtLaunchKernelGGL): no marker packets
// between the kernels of an iteration.
static void launch_plan(const GemmPlan& p, Bases b, hipStream_t s, gemm_kernel_t k32, gemm_kernel_t k64,
                        gemm128_kernel_t k128, hipEvent_t ev0, hipEvent_t ev1) {
  const int n = (int)p.htiles.size();
  if (p.nbig > 0) {
    const int nitems = (int)p.hbig.size();
    const BigArgs ba{p.big.p, nitems, p.nbig, p.part.p, p.cnt.p};
    const int npairs = (n - p.nbig + GB_NW / 4 - 1) / (GB_NW / 4);
    const unsigned grid = 8u * (unsigned)((nitems + 7) / 8) + 8u * (unsigned)((npairs + 7) / 8);
    if (ev0)
      hipExtLaunchKernelGGL(k128, dim3(grid), dim3(GB_THREADS), 0, s, ev0, ev1, 0, p.tiles.p, p.segs.p, n, b, ba);
    else hipLaunchKernelGGL(k128, dim3(grid), dim3(GB_THREADS), 0, s, p.tiles.p, p.segs.p, n, b, ba);
    NSS_HIP(hipGetLastError());
    return;
  }
  const unsigned grid = 8u * (unsigned)((n + 7) / 8);
  gemm_kernel_t k = p.tile == 64 ? k64 : k32;
  if (ev0) hipExtLaunchKernelGGL(k, dim3(grid), dim3(256), 0, s, ev0, ev1, 0, p.tiles.p, p.segs.p, n, b);
  else hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, s, p.tiles.p, p.segs.p, n, b);
  NSS_HIP(hipGetLastError());
}

void launch_gemm_f64(const GemmPlan& p, Bases b, hipStream_t s, int kind, hipEvent_t ev0, hipEvent_t ev1) {
  if (p.htiles.empty()) return;
  if (kind == GEMM_HV_W) launch_plan(p, b, s, k_hv_w_f64, k_hv_w_f64_t64, k_hv_w_f64_t128, ev0, ev1);
  else if (kind == GEMM_HV_Y) launch_plan(p, b, s, k_hv_y_f64, k_hv_y_f64_t64, k_hv_y_f64_t128, ev0, ev1);
  else launch_plan(p, b, s, k_gemm_f64, k_gemm_f64_t64, k_gemm_f64_t128, ev0, ev1);
}

void launch_gemm_f32(const GemmPlan& p, Bases b, hipStream_t s, int kind, hipEvent_t ev0, hipEvent_t ev1) {
  if (p.htiles.empty()) return;
  if (kind == GEMM_HV_W) launch_plan(p, b, s, k_hv_w_f32, k_hv_w_f32_t64, k_hv_w_f32_t128, ev0, ev1);
  else launch_plan(p, b, s, k_hv_y_f32, k_hv_y_f32_t64, k_hv_y_f32_t128, ev0, ev1);
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

// ---- Lanczos iteration kernels (no host round trip per iteration)
// All reductions are fixed-order: per-chunk partials, then a block-cooperative
// sum (strided over 256 threads, then wave shuffles and LDS) whose order does
// not depend on timing, so every run is bit-reproducible.
constexpr int64_t LZ_CHUNK = 2048;  // elements per partial-sum chunk
// basis vectors per block of k_lz_gemvt: one gives the most blocks in flight
// (Hubbard L=64 m=256 steps 5.1 / 5.2 / 5.4 / 5.9 / 6.9 ms with 1 / 2 / 4 / 8 / 16)
constexpr int LZ_GEMVT_VG = 1;

__device__ __forceinline__ double coop_sum(const double* __restrict__ p, int64_t cnt, double* sh) {
  double a = 0.0;
  for (int64_t c = threadIdx.x; c < cnt; c += 256) a += p[c];
  return block_sum256(a, sh);  // valid on thread 0
}

int64_t lz_nch(int64_t n) { return (n + LZ_CHUNK - 1) / LZ_CHUNK; }
int orth_blocks(int64_t n) { return (int)std::min<int64_t>(2048, std::max<int64_t>(1, (n + 255) / 256)); }
int64_t orth_chunk(int64_t) { return LZ_CHUNK; }

// w -= beta_{j-1} v_{j-1} (when beta_prev != null); partial[c] = <v_j, w> over chunk c;
// block 0 copies beta_{j-1} to beta_out.
__global__ __launch_bounds__(256) void k_lz_alpha(double* __restrict__ w, const double* __restrict__ vprev,
                                                  const double* __restrict__ vj, int64_t n,
                                                  const double* __restrict__ beta_prev, double* __restrict__ partial,
                                                  double* __restrict__ beta_out) {
  __shared__ double sh[4];
  const int64_t c = blockIdx.x;
  const int64_t i0 = c * LZ_CHUNK, i1 = min<int64_t>(n, i0 + LZ_CHUNK);
  const double b = beta_prev ? beta_prev[0] : 0.0;
  double acc = 0.0;
#pragma unroll 8
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    double x = w[i];
    if (beta_prev) {
      x -= b * vprev[i];
      w[i] = x;
    }
    acc += vj[i] * x;
  }
  const double r = block_sum256(acc, sh);
  if (threadIdx.x == 0) {
    partial[c] = r;
    if (c == 0 && beta_out) beta_out[0] = b;
  }
}

// One classical Gram-Schmidt pass w -= V h, h = V^T w', in four launches
// that each read what they need once: the recurrence scalars (alpha from its
// partials, beta), the chunk products (every block keeps its chunk of w' in
// registers and runs it against all m basis vectors), the coefficient sums
// (one block per vector) and the update.  Every sum has a fixed order.

// aout[0] = sum of apart[0..na) (fixed order); bout[0] = bprev[0] or 0
__global__ __launch_bounds__(256) void k_lz_scalars(const double* __restrict__ apart, int64_t na,
                                                    double* __restrict__ aout, const double* __restrict__ bprev,
                                                    double* __restrict__ bout) {
  __shared__ double sh[4];
  const double a = coop_sum(apart, na, sh);
  if (threadIdx.x == 0) {
    aout[0] = a;
    if (bout) bout[0] = bprev ? bprev[0] : 0.0;
  }
}

// partial[j*nch + c] = <V_j, w'> over chunk c, w' = w - alpha v_a - beta v_b
// (alpha = ascal[0] when ascal, beta = bprev[0] when vb); block (c, g) takes
// the chunk's products with the LZ_GEMVT_VG vectors of group g
__global__ __launch_bounds__(256) void k_lz_gemvt(const double* __restrict__ V, int64_t ldv, int m,
                                                  const double* __restrict__ w, int64_t n,
                                                  double* __restrict__ partial, int64_t nch,
                                                  const double* __restrict__ ascal, const double* __restrict__ va,
                                                  const double* __restrict__ vb, const double* __restrict__ bprev) {
  constexpr int PER = LZ_CHUNK / 256;
  __shared__ double red[LZ_GEMVT_VG][4];
  const int64_t c = blockIdx.x;
  const int j0 = blockIdx.y * LZ_GEMVT_VG;  // this block's basis vectors j0 .. j0 + VG - 1
  const int64_t i0 = c * LZ_CHUNK;
  const double a = ascal ? ascal[0] : 0.0;
  const double b = vb ? bprev[0] : 0.0;
  double x[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int64_t i = i0 + threadIdx.x + 256 * u;
    double wi = 0.0;
    if (i < n) {
      wi = w[i];
      if (ascal) wi -= a * va[i];
      if (vb) wi -= b * vb[i];
    }
    x[u] = wi;
  }
  // every load of the block's vectors issued together (indices past m read
  // vector m-1 and are not stored), then one wave sum per vector
  double acc[LZ_GEMVT_VG];
#pragma unroll
  for (int q = 0; q < LZ_GEMVT_VG; ++q) {
    const double* v = V + (size_t)min(j0 + q, m - 1) * ldv;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int64_t i = i0 + threadIdx.x + 256 * u;
      if (i < n) s += v[i] * x[u];
    }
    acc[q] = s;
  }
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < LZ_GEMVT_VG; ++q) {
    const double r = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) red[q][wave] = r;
  }
  __syncthreads();
  const int q = threadIdx.x;
  if (q < LZ_GEMVT_VG && j0 + q < m)
    partial[(size_t)(j0 + q) * nch + c] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
}

// h[j] = sum_c partial[j*nch + c], one block per j
__global__ __launch_bounds__(256) void k_lz_hsum(const double* __restrict__ partial, int64_t nch,
                                                 double* __restrict__ h) {
  __shared__ double sh[4];
  const double r = coop_sum(partial + (size_t)blockIdx.x * nch, nch, sh);
  if (threadIdx.x == 0) h[blockIdx.x] = r;
}

// w = w' - V[0..m) h (w' as in k_lz_gemvt); sq[block] = sum of squares of
// the block's new rows
__global__ __launch_bounds__(256) void k_orth_update(const double* __restrict__ V, int64_t ldv, int m,
                                                     const double* __restrict__ h, double* __restrict__ w,
                                                     int64_t n, double* __restrict__ sq,
                                                     const double* __restrict__ ascal, const double* __restrict__ va,
                                                     const double* __restrict__ vb,
                                                     const double* __restrict__ bprev) {
  extern __shared__ double hs[];
  __shared__ double red[4];
  for (int j = threadIdx.x; j < m; j += 256) hs[j] = h[j];
  __syncthreads();
  const double a = ascal ? ascal[0] : 0.0;
  const double b = vb ? bprev[0] : 0.0;
  double ss = 0.0;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    // eight basis vectors per step, their loads issued together
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int j = 0;
    for (; j + 8 <= m; j += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = V[(size_t)(j + u) * ldv + i];
      a0 += x[0] * hs[j] + x[4] * hs[j + 4];
      a1 += x[1] * hs[j + 1] + x[5] * hs[j + 5];
      a2 += x[2] * hs[j + 2] + x[6] * hs[j + 6];
      a3 += x[3] * hs[j + 3] + x[7] * hs[j + 7];
    }
    for (; j < m; ++j) a0 += V[(size_t)j * ldv + i] * hs[j];
    double wi = w[i];
    if (ascal) wi -= a * va[i];
    if (vb) wi -= b * vb[i];
    const double x = wi - ((a0 + a1) + (a2 + a3));
    w[i] = x;
    ss += x * x;
  }
  if (!sq) return;  // no sums of squares asked for (uniform over the grid)
  const double r = block_sum256(ss, red);
  if (threadIdx.x == 0) sq[blockIdx.x] = r;
}

// beta = sqrt(sum sq[0..nb)); v = w / beta; block 0 stores beta.
__global__ __launch_bounds__(256) void k_normalize(const double* __restrict__ w, double* __restrict__ v,
                                                   int64_t n, const double* __restrict__ sq, int nb,
                                                   double* __restrict__ bout) {
  __shared__ double sh[4];
  __shared__ double bsh;
  const double ssum = coop_sum(sq, nb, sh);
  if (threadIdx.x == 0) {
    bsh = sqrt(ssum);
    if (blockIdx.x == 0) bout[0] = bsh;
  }
  __syncthreads();
  const double inv = 1.0 / bsh;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) v[i] = w[i] * inv;
}

static unsigned wide_grid(int64_t n) { return (unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, (n + 255) / 256)); }

void launch_lz_recurrence(double* w, const double* vprev, const double* vj, int64_t n, const double* beta_prev,
                          double* beta_out, double* apart, hipStream_t s) {
  const int64_t nch = lz_nch(n);
  hipLaunchKernelGGL(k_lz_alpha, dim3((unsigned)nch), dim3(256), 0, s, w, vprev, vj, n, beta_prev, apart,
                     beta_out);
  NSS_HIP(hipGetLastError());
}

void launch_orth_pass(const double* V, int64_t ldv, int m, double* w, int64_t n, double* partial,
                      double* hout, double* sq, hipStream_t s, const double* apart, const double* va,
                      double* aout, int64_t na, const double* vb, const double* bprev, double* bout) {
  NSS_REQUIRE(m >= 1 && m <= LZ_MAX_BASIS, NSS_E_ARG, "Gram-Schmidt pass: basis size outside 1.." +
                                                        std::to_string(LZ_MAX_BASIS));
  NSS_REQUIRE(!apart || aout, NSS_E_ARG, "Gram-Schmidt pass: alpha partials without an alpha slot");
  const int64_t nch = lz_nch(n);
  if (na < 0) na = nch;
  const double* ascal = nullptr;
  if (apart) {
    hipLaunchKernelGGL(k_lz_scalars, dim3(1), dim3(256), 0, s, apart, na, aout, vb ? bprev : nullptr, bout);
    ascal = aout;
  }
  hipLaunchKernelGGL(k_lz_gemvt, dim3((unsigned)nch, (unsigned)((m + LZ_GEMVT_VG - 1) / LZ_GEMVT_VG)), dim3(256), 0, s,
                     V, ldv, m, w, n, partial, nch, ascal, va, vb, bprev);
  hipLaunchKernelGGL(k_lz_hsum, dim3((unsigned)m), dim3(256), 0, s, partial, nch, hout);
  hipLaunchKernelGGL(k_orth_update, dim3((unsigned)orth_blocks(n)), dim3(256), sizeof(double) * (size_t)m, s, V,
                     ldv, m, hout, w, n, sq, ascal, va, vb, bprev);
  NSS_HIP(hipGetLastError());
}

void launch_normalize(const double* w, double* v, int64_t n, const double* sq, double* bout, hipStream_t s) {
  hipLaunchKernelGGL(k_normalize, dim3(wide_grid(n)), dim3(256), 0, s, w, v, n, sq, orth_blocks(n), bout);
  NSS_HIP(hipGetLastError());
}

// ---- batched Lanczos: one launch per vector operation for every active chain
// Block b of a launch works on chain bchain[b], chunk b - blk0 of that chain
// (LZ_CHUNK elements).  The per-chain slot of iteration j lies at
// slots + j * sstride with the single-chain layout: [0, kb1) Gram-Schmidt
// coefficients, kb1 + i the recurrence terms, 2 kb1 beta_j.  Reductions are
// fixed-order (partials per chunk, then a block-cooperative sum).

// w -= beta_{j-1} v_{j-1} (j > 0); part[chunk] = <v_j, w>; chunk 0 copies beta_{j-1}
__global__ __launch_bounds__(256) void k_lzb_alpha(const LzChainDev* __restrict__ ch,
                                                   const int32_t* __restrict__ bchain, int j, int kb1) {
  __shared__ double sh[4];
  const LzChainDev C = ch[bchain[blockIdx.x]];
  if (!C.active) return;
  const int64_t chunk = blockIdx.x - C.blk0;
  const int64_t i0 = chunk * LZ_CHUNK, i1 = min<int64_t>(C.n, i0 + LZ_CHUNK);
  const double* vj = C.V + (size_t)j * C.n;
  const double* vp = C.V + (size_t)(j > 0 ? j - 1 : 0) * C.n;
  const double b = j > 0 ? C.slots[(size_t)(j - 1) * C.sstride + 2 * kb1] : 0.0;
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    double x = C.w[i];
    if (j > 0) {
      x -= b * vp[i];
      C.w[i] = x;
    }
    acc += vj[i] * x;
  }
  const double r = block_sum256(acc, sh);
  if (threadIdx.x == 0) {
    C.part[chunk] = r;
    if (chunk == 0 && j > 0) C.slots[(size_t)j * C.sstride + kb1 + j - 1] = b;
  }
}

// alpha = sum of the chain's partials; w -= alpha v_j; chunk 0 stores alpha
__global__ __launch_bounds__(256) void k_lzb_sub(const LzChainDev* __restrict__ ch,
                                                 const int32_t* __restrict__ bchain, int j, int kb1) {
  __shared__ double sh[4];
  __shared__ double ash;
  const LzChainDev C = ch[bchain[blockIdx.x]];
  if (!C.active) return;
  const int64_t chunk = blockIdx.x - C.blk0;
  const double a0 = coop_sum(C.part, C.nch, sh);
  if (threadIdx.x == 0) {
    ash = a0;
    if (chunk == 0) C.slots[(size_t)j * C.sstride + kb1 + j] = a0;
  }
  __syncthreads();
  const double a = ash;
  const double* vj = C.V + (size_t)j * C.n;
  const int64_t i0 = chunk * LZ_CHUNK, i1 = min<int64_t>(C.n, i0 + LZ_CHUNK);
  double ss = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const double x = C.w[i] - a * vj[i];
    C.w[i] = x;
    ss += x * x;
  }
  const double r = block_sum256(ss, sh);
  if (threadIdx.x == 0) C.part[(size_t)(kb1 + 1) * C.nch + chunk] = r;
}

// Partial reorthogonalisation (H. D. Simon, Math. Comp. 42 (1984) 115-142):
// one block per chain propagates the estimates omega_{j+1,k} ~ v_{j+1}^T v_k
// of the loss of orthogonality with the recurrence
//   beta_j om_{j+1,k} = beta_k om_{j,k+1} + (alpha_k - alpha_j) om_{j,k}
//                       + beta_{k-1} om_{j,k-1} - beta_{j-1} om_{j-1,k} +- 2 eps |T|
// and asks for the full Gram-Schmidt pass when max_k |om_{j+1,k}| exceeds
// sqrt(eps), then again in the next iteration; after a pass the estimates
// restart at eps.  beta_j here is |w| after the three-term step; eps is the
// chain's rounding level (FP64 or FP32 H v) and psi its estimate of
// omega_{j+1,j}.
__global__ __launch_bounds__(256) void k_lzb_omega(const LzChainDev* __restrict__ ch, int j, int kb1) {
  __shared__ double sh[4];
  __shared__ double red[256];
  __shared__ int now_s;
  const LzChainDev C = ch[blockIdx.x];
  if (!C.active) return;
  const double bp0 = coop_sum(C.part + (size_t)(kb1 + 1) * C.nch, C.nch, sh);
  __shared__ double bps;
  if (threadIdx.x == 0) bps = sqrt(bp0);
  __syncthreads();
  const double bp = bps;
  const double* sj = C.slots + (size_t)j * C.sstride;
  const double aj = sj[kb1 + j];
  double* cur = C.om + (size_t)(j % 3) * (kb1 + 1);
  double* prv = C.om + (size_t)((j + 2) % 3) * (kb1 + 1);
  double* nxt = C.om + (size_t)((j + 1) % 3) * (kb1 + 1);
  const double bjm1 = j > 0 ? C.slots[(size_t)(j - 1) * C.sstride + 2 * kb1] : 0.0;
  // |T| estimate: max_k |alpha_k| + beta_k over the iterations so far
  double tn = fabs(aj) + bp;
  for (int k = threadIdx.x; k < j; k += 256) {
    const double* sk = C.slots + (size_t)k * C.sstride;
    tn = fmax(tn, fabs(sk[kb1 + k]) + sk[2 * kb1]);
  }
  red[threadIdx.x] = tn;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  tn = red[0];
  __syncthreads();
  const double eps = C.eps;
  double mx = 0.0;
  for (int k = threadIdx.x; k < j; k += 256) {
    const double* sk = C.slots + (size_t)k * C.sstride;
    const double ak = sk[kb1 + k], bk = sk[2 * kb1];
    double t = bk * cur[k + 1] + (ak - aj) * cur[k] - bjm1 * prv[k];
    if (k > 0) t += C.slots[(size_t)(k - 1) * C.sstride + 2 * kb1] * cur[k - 1];
    const double v = (t + copysign(2.0 * eps * tn, t)) / bp;
    nxt[k] = v;
    mx = fmax(mx, fabs(v));
  }
  const double psi = C.psi;
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    mx = fmax(red[0], psi);
    const int pend = C.st[1];
    const int trig = mx > sqrt(eps);
    const int now = pend || trig || !(bp > 0.0) || !isfinite(bp);
    C.st[0] = now;
    C.st[1] = (trig && !pend) ? 1 : 0;
    C.bpre[0] = bp;
    now_s = now;
    nxt[j] = psi;
    nxt[j + 1] = 1.0;
  }
  __syncthreads();
  if (now_s)
    for (int k = threadIdx.x; k <= j; k += 256) nxt[k] = eps;
}

// part[r * nch + chunk] = <V_r, w> over the chunk, r = blockIdx.y <= j
__global__ __launch_bounds__(256) void k_lzb_gemvt(const LzChainDev* __restrict__ ch,
                                                   const int32_t* __restrict__ bchain, int full) {
  __shared__ double sh[4];
  const LzChainDev C = ch[bchain[blockIdx.x]];
  if (!C.active || !(full || C.st[0])) return;
  const int64_t chunk = blockIdx.x - C.blk0;
  const int r = blockIdx.y;
  const double* v = C.V + (size_t)r * C.n;
  const int64_t i0 = chunk * LZ_CHUNK, i1 = min<int64_t>(C.n, i0 + LZ_CHUNK);
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc += v[i] * C.w[i];
  const double res = block_sum256(acc, sh);
  if (threadIdx.x == 0) C.part[(size_t)r * C.nch + chunk] = res;
}

// h_r = sum_c part[r * nch + c] (r <= j, every block, one thread per r, a
// fixed order; chunk 0 publishes h); w -= V[0..j] h; the chunk's sum of
// squares of the new w to part[kb1 * nch + chunk]
__global__ __launch_bounds__(256) void k_lzb_update(const LzChainDev* __restrict__ ch,
                                                    const int32_t* __restrict__ bchain, int j, int kb1, int full) {
  extern __shared__ double hs[];
  __shared__ double red[4];
  const LzChainDev C = ch[bchain[blockIdx.x]];
  if (!C.active || !(full || C.st[0])) return;
  const int64_t chunk = blockIdx.x - C.blk0;
  const int m = j + 1;
  for (int r = threadIdx.x; r < m; r += 256) {
    const double* p = C.part + (size_t)r * C.nch;
    double a0 = 0.0, a1 = 0.0;
    int c = 0;
    for (; c + 2 <= C.nch; c += 2) {
      a0 += p[c];
      a1 += p[c + 1];
    }
    if (c < C.nch) a0 += p[c];
    hs[r] = a0 + a1;
    if (chunk == 0) C.slots[(size_t)j * C.sstride + r] = a0 + a1;
  }
  __syncthreads();
  const int64_t i0 = chunk * LZ_CHUNK, i1 = min<int64_t>(C.n, i0 + LZ_CHUNK);
  double ss = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int r = 0;
    for (; r + 4 <= m; r += 4) {
      a0 += C.V[(size_t)r * C.n + i] * hs[r];
      a1 += C.V[(size_t)(r + 1) * C.n + i] * hs[r + 1];
      a2 += C.V[(size_t)(r + 2) * C.n + i] * hs[r + 2];
      a3 += C.V[(size_t)(r + 3) * C.n + i] * hs[r + 3];
    }
    for (; r < m; ++r) a0 += C.V[(size_t)r * C.n + i] * hs[r];
    const double x = C.w[i] - ((a0 + a1) + (a2 + a3));
    C.w[i] = x;
    ss += x * x;
  }
  const double rs = block_sum256(ss, red);
  if (threadIdx.x == 0) C.part[(size_t)kb1 * C.nch + chunk] = rs;
}

// beta_j = sqrt(sum of the chain's squares); v_{j+1} = w / beta_j, also to
// x (the next H*v input); chunk 0 stores beta_j
__global__ __launch_bounds__(256) void k_lzb_normalize(const LzChainDev* __restrict__ ch,
                                                       const int32_t* __restrict__ bchain, int j, int kb1, int full) {
  __shared__ double sh[4];
  __shared__ double bsh;
  const LzChainDev C = ch[bchain[blockIdx.x]];
  if (!C.active) return;
  const int64_t chunk = blockIdx.x - C.blk0;
  const bool pass = full || C.st[0];
  const double ssum = pass ? coop_sum(C.part + (size_t)kb1 * C.nch, C.nch, sh) : 0.0;
  if (threadIdx.x == 0) {
    bsh = pass ? sqrt(ssum) : C.bpre[0];
    if (chunk == 0) C.slots[(size_t)j * C.sstride + 2 * kb1] = bsh;
  }
  __syncthreads();
  const double inv = 1.0 / bsh;
  double* v = C.V + (size_t)(j + 1) * C.n;
  const int64_t i0 = chunk * LZ_CHUNK, i1 = min<int64_t>(C.n, i0 + LZ_CHUNK);
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const double y = C.w[i] * inv;
    v[i] = y;
    if (C.x) C.x[i] = y;
  }
}

void launch_lzb_iteration(const LzChainDev* ch, const int32_t* bchain, int nblk, int nchain, int j, int kb1,
                          int full, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_lzb_alpha, dim3(nblk), dim3(256), 0, s, ch, bchain, j, kb1);
  hipLaunchKernelGGL(k_lzb_sub, dim3(nblk), dim3(256), 0, s, ch, bchain, j, kb1);
  if (!full) hipLaunchKernelGGL(k_lzb_omega, dim3(nchain), dim3(256), 0, s, ch, j, kb1);
  hipLaunchKernelGGL(k_lzb_gemvt, dim3(nblk, j + 1), dim3(256), 0, s, ch, bchain, full);
  hipLaunchKernelGGL(k_lzb_update, dim3(nblk), dim3(256), sizeof(double) * (size_t)(j + 1), s, ch, bchain, j, kb1,
                     full);
  hipLaunchKernelGGL(k_lzb_normalize, dim3(nblk), dim3(256), 0, s, ch, bchain, j, kb1, full);
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

// ------------------------------------------------------------ batched Jacobi
// Cyclic two-sided Jacobi on each symmetric block (one 512-thread workgroup
// per block, the block A and V^T resident in LDS).  A round of the
// round-robin tournament pairs all indices into n/2 disjoint rotations; each
// 2x2 sub-block A[{p,q},{r,s}] is then updated once as J_pq^T B J_rs (rows
// and columns in one pass), and rows p,q of V^T rotate with J_pq.  Rotation
// angles: Golub & Van Loan, Algorithm 8.4.1 (symSchur2).  Output:
// eigenvalues (unsorted) and eigenvectors as columns, row-major.
constexpr int JAC_THREADS = 512;
constexpr int JAC_WAVES = JAC_THREADS / 64;

// max over the workgroup, valid on every thread
__device__ __forceinline__ double jac_block_max(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < JAC_WAVES; ++w) r = fmax(r, red[w]);
  __syncthreads();
  return r;
}

// sum over the workgroup in a fixed order, valid on every thread
__device__ __forceinline__ double jac_block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < JAC_WAVES; ++w) r += red[w];
  __syncthreads();
  return r;
}

// The updates of a round are flattened over the workgroup: item e of the A
// update is the 2x2 sub-block (pair e / half, pair e % half), item e of the
// V update is column e % np of the row pair e / np.  Each thread loads the
// operands of JAC_U items before it computes and stores any of them (the
// items of a round touch disjoint elements), so their LDS latencies overlap.
constexpr int JAC_U = 4;

__global__ __launch_bounds__(JAC_THREADS) void k_jacobi_eig(const double* __restrict__ a,
                                                            const int64_t* __restrict__ off,
                                                            const int32_t* __restrict__ dim,
                                                            const int64_t* __restrict__ eoff,
                                                            double* __restrict__ vout, double* __restrict__ wout,
                                                            int max_sweeps) {
  extern __shared__ double sm[];
  __shared__ double cs_c[64], cs_s[64];
  __shared__ int pr_p[64], pr_q[64];
  __shared__ double red[JAC_WAVES];
  const int b = blockIdx.x;
  const int n = dim[b];
  const int np = n + (n & 1);  // even number of slots (one dummy when n is odd)
  const int ld = np + 1;
  const int half = np / 2;
  const int tid = threadIdx.x;
  double* A = sm;
  double* Vt = sm + (size_t)np * ld;
  const double* src = a + off[b];
  double f = 0.0;
  for (int e = tid; e < np * np; e += JAC_THREADS) {
    const int i = e / np, j = e - i * np;
    const double x = (i < n && j < n) ? src[(size_t)i * n + j] : 0.0;
    A[i * ld + j] = x;
    Vt[i * ld + j] = (i == j) ? 1.0 : 0.0;
    f += x * x;
  }
  // rotations are skipped, and the sweeps stop, once every off-diagonal
  // element is below 1e-15 ||A||_F: a few ulps of ||A||, above the rounding
  // noise that rotations leave behind, so the absolute accuracy of the
  // eigenvalues is ~eps*||A|| as LAPACK gives for a PSD matrix.  The
  // two triangles are updated separately and drift apart by rounding, so
  // both tests use the symmetric part (A_pq + A_qp)/2: a test on one
  // triangle and a rotation decided on the other would leave an element
  // just above the threshold for ever.  (The sum also orders the loads
  // above before the sweeps read A.)
  const double thr = 1e-15 * sqrt(jac_block_sum(f, red));
  const int na = half * half, nv = half * np;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    double o = 0.0;
    for (int e = tid; e < np * np; e += JAC_THREADS) {
      const int i = e / np, j = e - i * np;
      if (i != j) o = fmax(o, fabs(0.5 * (A[i * ld + j] + A[j * ld + i])));
    }
    if (!(jac_block_max(o, red) > thr)) break;
    for (int r = 0; r < np - 1; ++r) {
      if (tid < half) {
        const int i = tid;
        const int k1 = np - 1 - i;
        int p = (i == 0) ? 0 : ((i - 1 + r) % (np - 1)) + 1;
        int q = ((k1 - 1 + r) % (np - 1)) + 1;
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, sn = 0.0;
        const double apq = 0.5 * (A[p * ld + q] + A[q * ld + p]);
        if (q < n && fabs(apq) > thr) {
          const double tau = (A[q * ld + q] - A[p * ld + p]) / (2.0 * apq);
          const double t = (tau >= 0.0) ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          sn = t * c;
        }
        cs_c[i] = c;
        cs_s[i] = sn;
        pr_p[i] = p;
        pr_q[i] = q;
      }
      __syncthreads();
      // A <- J^T A J, one 2x2 sub-block per item
      for (int e0 = tid; e0 < na; e0 += JAC_THREADS * JAC_U) {
        double b00[JAC_U], b01[JAC_U], b10[JAC_U], b11[JAC_U], ci[JAC_U], si[JAC_U], cj[JAC_U], sj[JAC_U];
        int p[JAC_U], q[JAC_U], rr[JAC_U], ss[JAC_U];
        bool diag[JAC_U];
#pragma unroll
        for (int u = 0; u < JAC_U; ++u) {
          const int e = min(e0 + u * JAC_THREADS, na - 1);  // past the end: repeat the last item, not stored
          const int i = e / half, j = e - i * half;
          p[u] = pr_p[i];
          q[u] = pr_q[i];
          rr[u] = pr_p[j];
          ss[u] = pr_q[j];
          ci[u] = cs_c[i];
          si[u] = cs_s[i];
          cj[u] = cs_c[j];
          sj[u] = cs_s[j];
          diag[u] = i == j;
          b00[u] = A[p[u] * ld + rr[u]];
          b01[u] = A[p[u] * ld + ss[u]];
          b10[u] = A[q[u] * ld + rr[u]];
          b11[u] = A[q[u] * ld + ss[u]];
        }
#pragma unroll
        for (int u = 0; u < JAC_U; ++u) {
          if (e0 + u * JAC_THREADS >= na) break;
          const double t00 = ci[u] * b00[u] - si[u] * b10[u], t01 = ci[u] * b01[u] - si[u] * b11[u];
          const double t10 = si[u] * b00[u] + ci[u] * b10[u], t11 = si[u] * b01[u] + ci[u] * b11[u];
          double n01 = sj[u] * t00 + cj[u] * t01, n10 = cj[u] * t10 - sj[u] * t11;
          if (diag[u] && si[u] != 0.0) n01 = n10 = 0.0;  // the rotated pair is annihilated
          A[p[u] * ld + rr[u]] = cj[u] * t00 - sj[u] * t01;
          A[p[u] * ld + ss[u]] = n01;
          A[q[u] * ld + rr[u]] = n10;
          A[q[u] * ld + ss[u]] = sj[u] * t10 + cj[u] * t11;
        }
      }
      // V <- V J  (rows p, q of V^T)
      for (int e0 = tid; e0 < nv; e0 += JAC_THREADS * JAC_U) {
        double vp[JAC_U], vq[JAC_U], c[JAC_U], sn[JAC_U];
        int ip[JAC_U], iq[JAC_U];
#pragma unroll
        for (int u = 0; u < JAC_U; ++u) {
          const int e = min(e0 + u * JAC_THREADS, nv - 1);
          const int i = e / np, k = e - i * np;
          c[u] = cs_c[i];
          sn[u] = cs_s[i];
          ip[u] = pr_p[i] * ld + k;
          iq[u] = pr_q[i] * ld + k;
          vp[u] = Vt[ip[u]];
          vq[u] = Vt[iq[u]];
        }
#pragma unroll
        for (int u = 0; u < JAC_U; ++u) {
          if (e0 + u * JAC_THREADS >= nv) break;
          if (sn[u] != 0.0) {
            Vt[ip[u]] = c[u] * vp[u] - sn[u] * vq[u];
            Vt[iq[u]] = sn[u] * vp[u] + c[u] * vq[u];
          }
        }
      }
      __syncthreads();
    }
  }
  double* vo = vout + off[b];
  for (int e = tid; e < n * n; e += JAC_THREADS) {
    const int i = e / n, j = e - i * n;
    vo[(size_t)i * n + j] = Vt[j * ld + i];
  }
  for (int i = tid; i < n; i += JAC_THREADS) wout[eoff[b] + i] = A[i * ld + i];
}

int jacobi_max_dim(int device) {
  static int cached = -1;
  if (cached >= 0) return cached;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lds = 65536;
  const int budget = lds - 8192;  // static LDS of the kernel
  int nmax = 0;
  for (int n = 2; n <= 128; n += 2)
    if (2 * n * (n + 1) * (int)sizeof(double) <= budget) nmax = n;
  if (nmax > 0)
    (void)hipFuncSetAttribute((const void*)k_jacobi_eig, hipFuncAttributeMaxDynamicSharedMemorySize,
                              2 * nmax * (nmax + 1) * (int)sizeof(double));
  cached = nmax;
  return cached;
}

void launch_jacobi_eig(const double* a, const int64_t* off, const int32_t* dim, const int64_t* eoff, double* vout,
                       double* wout, int nblk, int nmax, hipStream_t s) {
  if (nblk <= 0) return;
  const int np = nmax + (nmax & 1);
  const size_t lds = (size_t)2 * np * (np + 1) * sizeof(double);
  hipLaunchKernelGGL(k_jacobi_eig, dim3(nblk), dim3(JAC_THREADS), lds, s, a, off, dim, eoff, vout, wout, 40);
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ batched tridiagonal-path eigensolver
// One 1024-thread workgroup per symmetric block of dimension 3..TRI_MAX
// (LAPACK's dsyevx path: dsytrd + dstebz + dstein + dormtr):
//  1. Householder tridiagonalisation in LDS (dsytd2, lower: Q =
//     H_0 ... H_{n-3}, H_k = I - tau_k v_k v_k^T, the v_k kept in a global
//     workspace), the p = tau A v products eight lanes per row;
//  2. every eigenvalue by multisection on Sturm counts (Wilkinson,
//     The Algebraic Eigenvalue Problem, 5.37-5.39), TRI_BIS lanes per
//     eigenvalue, each evaluating one of TRI_BIS interior points per step;
//  3. its eigenvector by inverse iteration (dstein): two solves with
//     T - lambda I, factored by Gaussian elimination with partial pivoting
//     (dgttrf), from a per-vector pseudo-random start, the factors in the
//     global workspace and the iterate in column t of Z (LDS);
//  4. Gram-Schmidt within clusters of eigenvalues closer than 1e-3 ||T||,
//     as dstein does (right-looking modified Gram-Schmidt, a second sweep
//     where the first met near-degenerate iterates);
//  5. Z <- Q Z, two reflectors per pass.
// T is split into unreduced blocks at |e_k| <= eps ||T|| first (dstebz) and
// every eigenpair is computed within its block.  Output as the Jacobi
// kernel's: eigenvectors as columns (row-major), eigenvalues at wout +
// eoff[b], ascending within each block of T (the caller sorts).
constexpr int TRI_THREADS = 1024;
constexpr int TRI_WAVES = TRI_THREADS / 64;
constexpr int TRI_MAX = 128;
constexpr int TRI_BIS = 4;  // lanes (and interior points) per eigenvalue in the multisection
constexpr int TRI_PB = 8;   // vectors per panel of the cluster Gram-Schmidt
static_assert(TRI_MAX * TRI_BIS <= TRI_THREADS, "one group of TRI_BIS lanes per eigenvalue");
static_assert(TRI_MAX <= 128, "a Gram-Schmidt panel: two rows per lane of one wave");
constexpr int64_t TRI_WS = (int64_t)6 * TRI_MAX * TRI_MAX;  // workspace doubles per block

int64_t rho_tri_workspace(int nblk) { return (int64_t)nblk * TRI_WS; }

// A double moved across lanes by a DPP control (both 32-bit halves): lane
// exchanges within rows of 16 lanes at register speed, where a shuffle goes
// through the LDS crossbar.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// sum over the aligned group of 8 lanes, in each of them: quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror
__device__ __forceinline__ double group8_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  return v;
}

// sum over the wave (every lane active), in every lane: groups of 8, then
// row_mirror for rows of 16, then the four rows
__device__ __forceinline__ double tri_wave_sum(double v) {
  v = group8_sum(v);
  v += dpp_f64<0x140>(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// max over the workgroup, in every thread (the barriers also order the
// LDS work around it)
__device__ __forceinline__ double tri_block_max(double v, double* red) {
  v = fmax(v, dpp_f64<0xB1>(v));
  v = fmax(v, dpp_f64<0x4E>(v));
  v = fmax(v, dpp_f64<0x141>(v));
  v = fmax(v, dpp_f64<0x140>(v));
  v = fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < TRI_WAVES; ++w) r = fmax(r, red[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ double tri_block_sum(double v, double* red) {  // all threads get the sum
  v = tri_wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < TRI_WAVES; ++w) r += red[w];
  __syncthreads();
  return r;
}


// 1/q to full precision: the hardware reciprocal and two Newton steps
__device__ __forceinline__ double tri_rcp(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  return fma(fma(-q, r, 1.0), r, r);
}

// Householder tridiagonalisation (dsytd2's lower form) of the n x n block A
// in LDS (row stride ld, 3 <= n <= TRI_MAX) by the whole 1024-thread
// workgroup.  Reflector k (v[0] = 1 at row k+1, n-k-1 entries) goes to
// rout[k rk + r]; T to dg (n), eg (n-1) and tau (k < n-2).  A is overwritten.
__device__ __forceinline__ void tri_reduce_lds(double* A, int ld, int n, double* rout, int64_t rk, double* dg,
                                               double* eg, double* tau, double* vb, double* pb, double* red,
                                               double* red2, double* red3) {
  const int tid = threadIdx.x;
  const int g8 = tid >> 3, l8 = tid & 7;    // groups of 8 lanes
  const int r32 = tid >> 5, c32 = tid & 31;  // 32 x 32 thread grid
  // Every branch below is uniform: its inputs are LDS values or block sums
  // that every thread reads alike.  Three barriers per reflector: the sums
  // of squares of the next column and p.v are reduced per wave before the
  // barrier that follows their producing phase, and summed over the waves
  // after it.
  bool have_s2 = false;  // red2 holds the wave partials of the next column's squares
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // x = A[k+1..n-1][k]
    double xn2 = 0.0;
    if (have_s2) {
#pragma unroll
      for (int w = 0; w < TRI_WAVES; ++w) xn2 += red2[w];
    } else {
      double s2 = 0.0;
      if (tid + k + 2 < n) {
        const double x = A[(tid + k + 2) * ld + k];
        s2 = x * x;
      }
      xn2 = tri_block_sum(s2, red);
    }
    const double alpha = A[(k + 1) * ld + k];
    double t = 0.0, beta = alpha, scale = 0.0;
    if (xn2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      t = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    if (tid < m) {
      const double v = (tid == 0) ? 1.0 : A[(k + 1 + tid) * ld + k] * scale;
      vb[tid] = v;
      rout[(int64_t)k * rk + tid] = v;
    }
    if (tid == 0) {
      tau[k] = t;
      eg[k] = beta;
      dg[k] = A[k * ld + k];
    }
    __syncthreads();
    have_s2 = false;
    if (t != 0.0) {
      double pvl = 0.0;
      {  // p = t A22 v, eight threads per row
        double acc = 0.0;
        if (g8 < m) {
          const double* row = A + (k + 1 + g8) * ld + (k + 1);
          for (int c = l8; c < m; c += 8) acc += row[c] * vb[c];
        }
        acc = group8_sum(acc);
        if (g8 < m && l8 == 0) {
          const double pr = t * acc;
          pb[g8] = pr;
          pvl = pr * vb[g8];
        }
      }
      pvl = tri_wave_sum(pvl);
      if ((tid & 63) == 0) red3[tid >> 6] = pvl;
      __syncthreads();
      double pv = 0.0;
#pragma unroll
      for (int w = 0; w < TRI_WAVES; ++w) pv += red3[w];
      const double a2 = -0.5 * t * pv;
      // A22 -= v w^T + w v^T with w = p + a2 v, i.e. v p^T + p v^T + 2 a2 v v^T;
      // the new column k+1 below its subdiagonal feeds the next step's norm
      const double a22 = 2.0 * a2;
      double s2n = 0.0;
      double vcq[4], wcq[4];  // this thread's columns c32 + 32q (m <= 127), read once
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = c32 + 32 * q;
        vcq[q] = c < m ? vb[c] : 0.0;
        wcq[q] = c < m ? pb[c] + a22 * vb[c] : 0.0;
      }
      for (int r = r32; r < m; r += 32) {
        const double vr = vb[r], pr = pb[r];
        double* row = A + (k + 1 + r) * ld + (k + 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c = c32 + 32 * q;
          if (c < m) {
            const double nv = row[c] - (vr * wcq[q] + pr * vcq[q]);
            row[c] = nv;
            if (c == 0 && r >= 2) s2n += nv * nv;
          }
        }
      }
      s2n = tri_wave_sum(s2n);
      if ((tid & 63) == 0) red2[tid >> 6] = s2n;
      __syncthreads();
      have_s2 = true;
    }
  }
  if (tid == 0) {
    dg[n - 2] = A[(n - 2) * ld + (n - 2)];
    dg[n - 1] = A[(n - 1) * ld + (n - 1)];
    eg[n - 2] = A[(n - 1) * ld + (n - 2)];
  }
  __syncthreads();
}

__global__ __launch_bounds__(TRI_THREADS) void k_rho_eig_tri(const double* __restrict__ a,
                                                             const int64_t* __restrict__ off,
                                                             const int32_t* __restrict__ dim,
                                                             const int64_t* __restrict__ eoff,
                                                             double* __restrict__ vout, double* __restrict__ wout,
                                                             double* __restrict__ ws) {
  extern __shared__ double sm[];  // n x (n+1): A, then Z
  __shared__ double dg[TRI_MAX], eg[TRI_MAX], e2g[TRI_MAX], tau[TRI_MAX], lam[TRI_MAX];
  __shared__ double vb[TRI_MAX], pb[TRI_MAX], hb[TRI_MAX], yb[TRI_MAX];
  __shared__ double red[TRI_WAVES], red2[TRI_WAVES], red3[TRI_WAVES];
  __shared__ double cb[TRI_PB * TRI_MAX];  // C of a Gram-Schmidt panel
  const int b = blockIdx.x;
  const int n = dim[b];
  const int ld = n + 1;
  const int tid = threadIdx.x;
  const int g8 = tid >> 3, l8 = tid & 7;    // groups of 8 lanes
  const int r32 = tid >> 5, c32 = tid & 31;  // 32 x 32 thread grid
  const int wv = tid >> 6, lane = tid & 63;  // waves
  double* A = sm;
  const double* src = a + off[b];
  double* hv = ws + (size_t)b * TRI_WS;           // reflector k at hv[k*TRI_MAX + r], r < n-k-1
  double* lu = hv + (size_t)TRI_MAX * TRI_MAX;    // factor f, row i, vector t at lu[(f*TRI_MAX + i)*TRI_MAX + t]
  for (int i = r32; i < n; i += 32)
    for (int j = c32; j < n; j += 32) A[i * ld + j] = src[(size_t)i * n + j];
  __syncthreads();
  // 1. tridiagonalisation; the reflectors to the global workspace
  tri_reduce_lds(A, ld, n, hv, TRI_MAX, dg, eg, tau, vb, pb, red, red2, red3);
  // Gershgorin interval and ||T|| (every thread alike)
  double lo0 = HUGE_VAL, hi0 = -HUGE_VAL, tnorm = 0.0, emax2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? fabs(eg[i - 1]) : 0.0) + (i + 1 < n ? fabs(eg[i]) : 0.0);
    lo0 = fmin(lo0, dg[i] - r);
    hi0 = fmax(hi0, dg[i] + r);
    tnorm = fmax(tnorm, fabs(dg[i]) + r);
    if (i + 1 < n) emax2 = fmax(emax2, eg[i] * eg[i]);
  }
  tnorm = fmax(tnorm, DBL_MIN);
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  // zero pivots of T - lambda I become tiny: eps ||T||, floored so that a
  // (near-)zero block still gives finite iterates
  const double tiny = fmax(DBL_EPSILON * tnorm, 1e-280);
  __syncthreads();
  // T splits into unreduced blocks where |e_k| <= eps ||T|| (as dstebz
  // splits it); each eigenpair is computed within its block, so vectors of
  // different blocks are exactly orthogonal (the null space of a
  // rank-deficient rho comes out as unit vectors)
  if (tid + 1 < n && fabs(eg[tid]) <= DBL_EPSILON * tnorm) eg[tid] = 0.0;
  __syncthreads();
  if (tid < n) e2g[tid] = tid + 1 < n ? eg[tid] * eg[tid] : 0.0;
  __syncthreads();
  // position p of the output belongs to the block [s, e) that holds row p and
  // takes eigenvalue p - s of that block (ascending within the block)
  auto block_of = [&](int p, int& s, int& e) {
    s = p;
    e = p + 1;
    while (s > 0 && eg[s - 1] != 0.0) --s;
    while (e < n && eg[e - 1] != 0.0) ++e;
  };
  // 2. eigenvalues by multisection: the TRI_BIS lanes of a group evaluate the
  //    Sturm count of the block at TRI_BIS interior points of the interval
  //    each step
  {
    const int gb = tid / TRI_BIS, lb = tid % TRI_BIS;
    const int p = gb;
    int bs = 0, be = 1;
    if (p < n) block_of(p, bs, be);
    const int t = p - bs;
    double lo = lo0, hi = hi0;
    if (p >= n || be - bs == 1) lo = hi = (p < n ? dg[p] : 0.0);
    const double tol = 2.0 * DBL_EPSILON * tnorm;
    constexpr double inv_parts = 1.0 / (TRI_BIS + 1);
    for (int it = 0; it < 96; ++it) {
      if (!(hi - lo > tol)) break;  // alike on the lanes of a group
      const double x = lo + (hi - lo) * (double)(lb + 1) * inv_parts;
      double q = dg[bs] - x;
      if (fabs(q) < pivmin) q = -pivmin;
      int cnt = q < 0.0;
#pragma unroll 8
      for (int i = bs + 1; i < be; ++i) {
        q = dg[i] - x - e2g[i - 1] * tri_rcp(q);
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
      }
      const bool above = cnt >= t + 1;  // eigenvalue t lies below x
      const uint64_t bal = __ballot(above);
      const unsigned bits = (unsigned)(bal >> (threadIdx.x & (64 - TRI_BIS))) & ((1u << TRI_BIS) - 1u);
      const int f = bits ? __builtin_ctz(bits) : TRI_BIS;  // first point above it
      const double span = hi - lo;
      const double nlo = f > 0 ? lo + span * (double)f * inv_parts : lo;
      const double nhi = f < TRI_BIS ? lo + span * (double)(f + 1) * inv_parts : hi;
      lo = nlo;
      hi = fmax(nhi, nlo);
    }
    if (p < n && lb == 0) lam[p] = 0.5 * (lo + hi);
  }
  __syncthreads();
  double* Z = sm;  // A is no longer read: the reflectors are in hv, T in dg/eg
  if (tid < n) {
    // 3. inverse iteration for position tid within its block [bs, be):
    //    LU with partial pivoting of T_block - sh I (local rows 0..nb-1)
    int bs, be;
    block_of(tid, bs, be);
    const int nb = be - bs;
    for (int i = 0; i < n; ++i)
      if (i < bs || i >= be) Z[i * ld + tid] = 0.0;
    if (nb == 1) {
      Z[tid * ld + tid] = 1.0;
    } else {
      const double sh = lam[tid];
      double* DL = lu + tid;
      double* DD = lu + (size_t)TRI_MAX * TRI_MAX + tid;
      double* DU = lu + (size_t)2 * TRI_MAX * TRI_MAX + tid;
      double* DU2 = lu + (size_t)3 * TRI_MAX * TRI_MAX + tid;
      double* PV = lu + (size_t)4 * TRI_MAX * TRI_MAX + tid;
      const double* dgb = dg + bs;
      const double* egb = eg + bs;
      double* Zb = Z + (size_t)bs * ld + tid;  // local row i at Zb[i * ld]
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
        DL[(size_t)i * TRI_MAX] = f;
        DD[(size_t)i * TRI_MAX] = ddi;
        DU[(size_t)i * TRI_MAX] = dui;
        DU2[(size_t)i * TRI_MAX] = du2i;
        PV[(size_t)i * TRI_MAX] = pvi;
        ddi = ddn;
        dui = dun;
      }
      if (ddi == 0.0) ddi = tiny;
      DD[(size_t)(nb - 1) * TRI_MAX] = ddi;
      // start: uniform [-1, 1) from SplitMix64 of (vector, row)
      for (int i = 0; i < nb; ++i) {
        uint64_t z = (uint64_t)(tid * TRI_MAX + i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        Zb[i * ld] = (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
      }
      for (int pass = 0; pass < 2; ++pass) {
        double xi = Zb[0];
        // (unrolled so that the factor loads of several rows are in flight
        // at once: they do not depend on the iterate)
#pragma unroll 8
        for (int i = 0; i + 1 < nb; ++i) {  // L, with the row interchanges
          double xn = Zb[(i + 1) * ld];
          if (PV[(size_t)i * TRI_MAX] != 0.0) {
            const double t = xi;
            xi = xn;
            xn = t;
          }
          xn -= DL[(size_t)i * TRI_MAX] * xi;
          Zb[i * ld] = xi;
          xi = xn;
        }
        double x1 = xi / DD[(size_t)(nb - 1) * TRI_MAX];  // U (two superdiagonals)
        Zb[(nb - 1) * ld] = x1;
        double x0 = (Zb[(nb - 2) * ld] - DU[(size_t)(nb - 2) * TRI_MAX] * x1) / DD[(size_t)(nb - 2) * TRI_MAX];
        Zb[(nb - 2) * ld] = x0;
        double x2 = x1;
        x1 = x0;
#pragma unroll 8
        for (int i = nb - 3; i >= 0; --i) {
          x0 = (Zb[i * ld] - DU[(size_t)i * TRI_MAX] * x1 - DU2[(size_t)i * TRI_MAX] * x2) / DD[(size_t)i * TRI_MAX];
          Zb[i * ld] = x0;
          x2 = x1;
          x1 = x0;
        }
        double mx = 0.0;
        for (int i = 0; i < nb; ++i) mx = fmax(mx, fabs(Zb[i * ld]));
        double ss = 0.0;
        for (int i = 0; i < nb; ++i) {
          const double y = Zb[i * ld] / mx;
          ss += y * y;
        }
        const double inv = 1.0 / (mx * sqrt(ss));
        for (int i = 0; i < nb; ++i) Zb[i * ld] *= inv;
      }
    }
  }
  __syncthreads();
  // 4. clusters: consecutive eigenvalues of one block within 1e-3 ||T|| of
  //    each other (dstein's reorthogonalisation criterion); blocked modified
  //    Gram-Schmidt, up to two sweeps.  A panel of TRI_PB consecutive vectors
  //    is orthonormalised within itself by one wave holding it in registers
  //    (two rows per lane; step j: h_k = z_j . z_k for the panel's k >= j,
  //    z_k -= (h_k / h_j) z_j, z_j /= |z_j|), then projected out of every
  //    later vector of the cluster at once: C = P^T Z (eight lanes per column
  //    of Z), Z -= P C.  A cluster's vectors are read once per panel instead
  //    of once per vector, and a panel costs three barriers.
  const double ctol = 1e-3 * tnorm;
  for (int s0 = 0; s0 < n;) {
    int e0 = s0 + 1;
    while (e0 < n && eg[e0 - 1] != 0.0 && lam[e0] - lam[e0 - 1] <= ctol) ++e0;  // within one block
    if (e0 - s0 > 1) {
      for (int sweep = 0; sweep < 2; ++sweep) {
        double ovl = 0.0;  // largest overlap coefficient this thread applied
        for (int j0 = s0; j0 < e0; j0 += TRI_PB) {
          const int pe = min(j0 + TRI_PB, e0);
          if (wv == 0) {  // wave-uniform: every lane of wave 0 takes part in the sums
            const int np = pe - j0;
            double z0[TRI_PB], z1[TRI_PB];  // rows lane and lane + 64
#pragma unroll
            for (int q = 0; q < TRI_PB; ++q) {
              z0[q] = (q < np && lane < n) ? Z[lane * ld + j0 + q] : 0.0;
              z1[q] = (q < np && lane + 64 < n) ? Z[(lane + 64) * ld + j0 + q] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < TRI_PB; ++j) {
              if (j < np) {
                double h[TRI_PB];
#pragma unroll
                for (int k = j; k < TRI_PB; ++k) h[k] = k < np ? tri_wave_sum(z0[j] * z0[k] + z1[j] * z1[k]) : 0.0;
                const double hj = h[j];
#pragma unroll
                for (int k = j + 1; k < TRI_PB; ++k)
                  if (k < np) {
                    const double c = h[k] / hj;
                    ovl = fmax(ovl, fabs(c));
                    z0[k] -= c * z0[j];
                    z1[k] -= c * z1[j];
                  }
                const double inv = 1.0 / sqrt(hj);
                z0[j] *= inv;
                z1[j] *= inv;
              }
            }
#pragma unroll
            for (int q = 0; q < TRI_PB; ++q) {
              if (q < np && lane < n) Z[lane * ld + j0 + q] = z0[q];
              if (q < np && lane + 64 < n) Z[(lane + 64) * ld + j0 + q] = z1[q];
            }
          }
          __syncthreads();
          const int pb = pe - j0, rest = e0 - pe;
          if (rest > 0) {
            for (int kk = g8; kk < rest; kk += TRI_THREADS / 8) {
              double acc[TRI_PB];
#pragma unroll
              for (int q = 0; q < TRI_PB; ++q) acc[q] = 0.0;
              for (int i = l8; i < n; i += 8) {
                const double z = Z[i * ld + pe + kk];
#pragma unroll
                for (int q = 0; q < TRI_PB; ++q)
                  if (q < pb) acc[q] += Z[i * ld + j0 + q] * z;
              }
#pragma unroll
              for (int q = 0; q < TRI_PB; ++q) acc[q] = group8_sum(acc[q]);
              if (l8 == 0)
                for (int q = 0; q < pb; ++q) cb[q * TRI_MAX + kk] = acc[q];
            }
            __syncthreads();
            double cq[4][TRI_PB];  // this thread's columns c32 + 32s of C
#pragma unroll
            for (int sc = 0; sc < 4; ++sc)
#pragma unroll
              for (int q = 0; q < TRI_PB; ++q) {
                const int k = c32 + 32 * sc;
                cq[sc][q] = (k < rest && q < pb) ? cb[q * TRI_MAX + k] : 0.0;
                ovl = fmax(ovl, fabs(cq[sc][q]));
              }
            for (int i = r32; i < n; i += 32) {
              double pz[TRI_PB];
#pragma unroll
              for (int q = 0; q < TRI_PB; ++q) pz[q] = q < pb ? Z[i * ld + j0 + q] : 0.0;
#pragma unroll
              for (int sc = 0; sc < 4; ++sc) {
                const int k = c32 + 32 * sc;
                if (k < rest) {
                  double d = 0.0;
#pragma unroll
                  for (int q = 0; q < TRI_PB; ++q) d += pz[q] * cq[sc][q];
                  Z[i * ld + pe + k] -= d;
                }
              }
            }
            __syncthreads();
          }
        }
        // a second sweep only where the first removed overlaps above 1e-6
        // (near-degenerate iterates); below that one sweep leaves the
        // cluster orthogonal to working precision
        if (!(tri_block_max(ovl, red) > 1e-6)) break;
      }
    }
    s0 = e0;
  }
  // 5. Z <- H_0 ... H_{n-3} Z, two reflectors per pass: with k1 = k0 + 1,
  //    H_k0 H_k1 Z = Z - v1 y1^T - v0 y0^T, y1 = t1 Z^T v1 and
  //    y0 = t0 (Z^T v0 - (v0.v1) y1), both over the rows k1.. (v1 has a zero
  //    there); one load, one product and one update phase per pair
  int k1 = n - 3;
  for (; k1 >= 1; k1 -= 2) {
    const int k0 = k1 - 1;
    const int m0 = n - k1;  // rows k1 .. n-1
    const double t1 = tau[k1], t0 = tau[k0];
    double sl = 0.0;
    if (tid < m0) {
      const double v0 = hv[(size_t)k0 * TRI_MAX + tid];  // reflector k0 starts at row k0+1 = k1
      const double v1 = tid == 0 ? 0.0 : hv[(size_t)k1 * TRI_MAX + tid - 1];
      vb[tid] = v0;
      hb[tid] = v1;
      sl = v0 * v1;
    }
    sl = tri_wave_sum(sl);
    if ((tid & 63) == 0) red2[tid >> 6] = sl;
    __syncthreads();
    for (int c = g8; c < n; c += TRI_THREADS / 8) {
      double a1 = 0.0, a0 = 0.0;
      for (int r = l8; r < m0; r += 8) {
        const double z = Z[(k1 + r) * ld + c];
        a1 += hb[r] * z;
        a0 += vb[r] * z;
      }
      a1 = group8_sum(a1);
      a0 = group8_sum(a0);
      if (l8 == 0) {
        pb[c] = t1 * a1;
        yb[c] = a0;
      }
    }
    __syncthreads();
    double s01 = 0.0;
#pragma unroll
    for (int w = 0; w < TRI_WAVES; ++w) s01 += red2[w];
    double y1q[4], y0q[4];  // this thread's columns c32 + 32q (n <= 128), read once
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = c32 + 32 * q;
      y1q[q] = c < n ? pb[c] : 0.0;
      y0q[q] = c < n ? t0 * (yb[c] - s01 * pb[c]) : 0.0;
    }
    for (int r = r32; r < m0; r += 32) {
      const double v0 = vb[r], v1 = hb[r];
      double* row = Z + (k1 + r) * ld;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = c32 + 32 * q;
        if (c < n) row[c] -= v1 * y1q[q] + v0 * y0q[q];
      }
    }
    __syncthreads();
  }
  if (k1 == 0 && tau[0] != 0.0) {  // the last single reflector
    const double t = tau[0];
    const int m = n - 1;
    if (tid < m) vb[tid] = hv[tid];
    __syncthreads();
    for (int c = g8; c < n; c += TRI_THREADS / 8) {
      double acc = 0.0;
      for (int r = l8; r < m; r += 8) acc += vb[r] * Z[(1 + r) * ld + c];
      acc = group8_sum(acc);
      if (l8 == 0) pb[c] = t * acc;
    }
    __syncthreads();
    for (int r = r32; r < m; r += 32) {
      const double vr = vb[r];
      double* row = Z + (1 + r) * ld;
      for (int c = c32; c < n; c += 32) row[c] -= vr * pb[c];
    }
    __syncthreads();
  }
  double* vo = vout + off[b];
  for (int i = r32; i < n; i += 32)
    for (int j = c32; j < n; j += 32) vo[(size_t)i * n + j] = Z[i * ld + j];
  for (int j = tid; j < n; j += TRI_THREADS) wout[eoff[b] + j] = lam[j];
}

int rho_tri_max_dim(int device) {
  static int cached = -1;
  if (cached >= 0) return cached;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lds = 65536;
  const int budget = lds - 24576;  // static LDS of the kernel (about 18 KB)
  int nmax = 0;
  for (int n = 3; n <= TRI_MAX; ++n)
    if (n * (n + 1) * (int)sizeof(double) <= budget) nmax = n;
  if (nmax > 0)
    (void)hipFuncSetAttribute((const void*)k_rho_eig_tri, hipFuncAttributeMaxDynamicSharedMemorySize,
                              nmax * (nmax + 1) * (int)sizeof(double));
  cached = nmax;
  return cached;
}

void launch_rho_eig_tri(const double* a, const int64_t* off, const int32_t* dim, const int64_t* eoff, double* vout,
                        double* wout, double* ws, int nblk, int nmax, hipStream_t s) {
  if (nblk <= 0) return;
  const size_t lds = (size_t)nmax * (nmax + 1) * sizeof(double);
  hipLaunchKernelGGL(k_rho_eig_tri, dim3(nblk), dim3(TRI_THREADS), lds, s, a, off, dim, eoff, vout, wout, ws);
  NSS_HIP(hipGetLastError());
}

// The tail of the large-block tridiagonalisation.  Once the per-column
// launches have reduced columns 0..jt-1 of a block, the rows and columns
// jt..n-1 still owe the update of reflector jt-1 (v in v[(jt-1)&1], p in
// p[(jt-1)&1] over rows jt.., tau[jt-1]); one workgroup applies it,
// A -= v w^T + w v^T with w = p - (tau/2)(p.v) v, while loading the trailing
// block into LDS, stores reflector jt-1 to row jt-1, and reduces the
// trailing block there.  Output in the per-column path's layout.
__global__ __launch_bounds__(TRI_THREADS) void k_tred_tail(const TredBlk* __restrict__ blks) {
  extern __shared__ double sm[];  // m0 x (m0+1)
  __shared__ double dg[TRI_MAX], eg[TRI_MAX], tau[TRI_MAX], vb[TRI_MAX], pb[TRI_MAX];
  __shared__ double red[TRI_WAVES], red2[TRI_WAVES], red3[TRI_WAVES];
  const TredBlk B = blks[blockIdx.x];
  const int n = B.n, jt = B.jt, m0 = n - jt;
  if (jt < 1 || m0 < 3) return;  // not a tail (the host never launches one)
  const int ld = m0 + 1;
  const int tid = threadIdx.x;
  const int r32 = tid >> 5, c32 = tid & 31;
  double* A = B.a;
  double* dd = B.ws + 4 * (int64_t)n;
  double* ee = B.ws + 5 * (int64_t)n;
  double* tt = B.ws + 6 * (int64_t)n;
  const double* vq = B.ws + (int64_t)((jt - 1) & 1) * n;
  const double* pq = B.ws + 2 * (int64_t)n + (int64_t)((jt - 1) & 1) * n;
  double pv = 0.0;
  for (int i = tid; i < m0; i += TRI_THREADS) {
    const double v = vq[jt + i], p = pq[jt + i];
    vb[i] = v;
    pb[i] = p;
    pv += p * v;
  }
  pv = tri_block_sum(pv, red);  // its barriers also publish vb, pb
  const double K = 0.5 * tt[jt - 1] * pv;
  for (int i = tid; i < m0; i += TRI_THREADS) {
    const double w = pb[i] - K * vb[i];
    A[(int64_t)(jt - 1) * n + jt + i] = vb[i];  // reflector jt-1 to its row
    pb[i] = w;  // only this thread reads and writes entry i here
  }
  __syncthreads();
  for (int i = r32; i < m0; i += 32)
    for (int j = c32; j < m0; j += 32)
      sm[i * ld + j] = A[(int64_t)(jt + i) * n + jt + j] - (vb[i] * pb[j] + pb[i] * vb[j]);
  __syncthreads();
  tri_reduce_lds(sm, ld, m0, A + (int64_t)jt * n + jt + 1, n + 1, dg, eg, tau, vb, pb, red, red2, red3);
  for (int k = tid; k < m0; k += TRI_THREADS) {
    dd[jt + k] = dg[k];
    ee[jt + k] = k + 1 < m0 ? eg[k] : 0.0;
    tt[jt + k] = k + 2 < m0 ? tau[k] : 0.0;
  }
  if (tid == 0) A[(int64_t)(n - 2) * n + (n - 1)] = 1.0;  // reflector n-2: the identity
}

void launch_tred_tail(const TredBlk* blks, int nblk, int nmax, hipStream_t s) {
  if (nblk <= 0) return;
  NSS_REQUIRE(nmax >= 3 && nmax <= TRI_MAX, NSS_E_ARG, "tridiagonalisation tail: size out of range");
  static const bool attr = [] {
    (void)hipFuncSetAttribute((const void*)k_tred_tail, hipFuncAttributeMaxDynamicSharedMemorySize,
                              TRI_MAX * (TRI_MAX + 1) * (int)sizeof(double));
    return true;
  }();
  (void)attr;
  const size_t lds = (size_t)nmax * (nmax + 1) * sizeof(double);
  hipLaunchKernelGGL(k_tred_tail, dim3(nblk), dim3(TRI_THREADS), lds, s, blks);
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ batched dense eigensolve staging
// Block b (dim[b] x dim[b] at src+off[b]) is copied into the b-th ncls x ncls
// slab of dst, the rest of the slab holding -1 on the diagonal (below any
// rho eigenvalue, decoupled from the block).
__global__ __launch_bounds__(256) void k_pad_blocks(const double* __restrict__ src, const int64_t* __restrict__ off,
                                                    const int32_t* __restrict__ dim, int ncls,
                                                    double* __restrict__ dst) {
  const int b = blockIdx.y;
  const int n = dim[b];
  double* o = dst + (size_t)b * ncls * ncls;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < (int64_t)ncls * ncls; e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / ncls), j = (int)(e % ncls);
    o[e] = (i < n && j < n) ? src[off[b] + (size_t)i * n + j] : (i == j ? -1.0 : 0.0);
  }
}

// After an ascending column-major eigen-decomposition of each slab, the
// block's eigenpairs are the last dim[b]: vectors (column-major, rows < n)
// back to dst+off[b], values to w+eoff[b].
__global__ __launch_bounds__(256) void k_unpad_eig(const double* __restrict__ vecs, const double* __restrict__ vals,
                                                   int ncls, const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ dim, const int64_t* __restrict__ eoff,
                                                   double* __restrict__ dst, double* __restrict__ w) {
  const int b = blockIdx.y;
  const int n = dim[b];
  const double* v = vecs + (size_t)b * ncls * ncls;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < (int64_t)n * n; e += (int64_t)gridDim.x * 256) {
    const int k = (int)(e / n), i = (int)(e % n);
    dst[off[b] + e] = v[(size_t)(ncls - n + k) * ncls + i];
  }
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < n; k += 256) w[eoff[b] + k] = vals[(size_t)b * ncls + ncls - n + k];
}

// Column gather for block rotations: task t copies the selected columns of
// a strided source into a dense row-major (rows x ncol) destination.
__global__ __launch_bounds__(256) void k_gather_cols(const GatherTask* __restrict__ tasks) {
  const GatherTask t = tasks[blockIdx.y];
  const int64_t ne = (int64_t)t.rows * t.ncol;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < ne; e += (int64_t)gridDim.x * 256) {
    const int r = (int)(e / t.ncol), c = (int)(e % t.ncol);
    t.dst[e] = t.src[r * t.si + (int64_t)t.cols[c] * t.sj];
  }
}

void launch_gather_cols(const GatherTask* tasks, int ntasks, int max_elems, hipStream_t s) {
  if (ntasks <= 0 || max_elems <= 0) return;
  const unsigned gx = (unsigned)std::min<int64_t>(256, ((int64_t)max_elems + 255) / 256);
  hipLaunchKernelGGL(k_gather_cols, dim3(gx, ntasks), dim3(256), 0, s, tasks);
  NSS_HIP(hipGetLastError());
}

void launch_pad_blocks(const double* src, const int64_t* off, const int32_t* dim, int nb, int ncls, double* dst,
                       hipStream_t s) {
  const unsigned gx = (unsigned)std::min<int64_t>(64, ((int64_t)ncls * ncls + 255) / 256);
  hipLaunchKernelGGL(k_pad_blocks, dim3(gx, nb), dim3(256), 0, s, src, off, dim, ncls, dst);
  NSS_HIP(hipGetLastError());
}

void launch_unpad_eig(const double* vecs, const double* vals, int ncls, const int64_t* off, const int32_t* dim,
                      const int64_t* eoff, int nb, double* dst, double* w, hipStream_t s) {
  const unsigned gx = (unsigned)std::min<int64_t>(64, ((int64_t)ncls * ncls + 255) / 256);
  hipLaunchKernelGGL(k_unpad_eig, dim3(gx, nb), dim3(256), 0, s, vecs, vals, ncls, off, dim, eoff, dst, w);
  NSS_HIP(hipGetLastError());
}

// ------------------------------------------------------------ symmetry check
// max |a_ij - a_ji| of every block: a 32x32 tile pair (I, J), I <= J, per
// workgroup (grid x: tile pairs of the largest block, y: block), both tiles
// read row-wise through LDS; the block's maximum by an unsigned max on the
// bit pattern (non-negative doubles order like their bits; NaN as +inf).
// out[] must enter zeroed.
constexpr int ASYM_T = 32;

__global__ __launch_bounds__(256) void k_asym_max(const double* __restrict__ a, const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ dim,
                                                  unsigned long long* __restrict__ out) {
  __shared__ double tj[ASYM_T][ASYM_T + 1];
  __shared__ double red[4];
  const int b = blockIdx.y;
  const int d = dim[b];
  const int nt = (d + ASYM_T - 1) / ASYM_T;
  // pair index p -> (I, J), I <= J, row-major over the upper triangle
  int p = blockIdx.x, I = 0;
  while (I < nt && p >= nt - I) {
    p -= nt - I;
    ++I;
  }
  if (I >= nt) return;  // workgroup-uniform
  const int J = I + p;
  const double* m = a + off[b];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows per pass
  for (int r = ty; r < ASYM_T; r += 8) {  // tile (J, I): row J*32 + r, column I*32 + tx
    const int gi = J * ASYM_T + r, gj = I * ASYM_T + tx;
    tj[r][tx] = (gi < d && gj < d) ? m[(size_t)gi * d + gj] : 0.0;
  }
  __syncthreads();
  double mx = 0.0;
  for (int r = ty; r < ASYM_T; r += 8) {  // a(I*32 + r, J*32 + tx) against a(J*32 + tx, I*32 + r)
    const int gi = I * ASYM_T + r, gj = J * ASYM_T + tx;
    if (gi < d && gj < d && gj > gi) {
      const double x = fabs(m[(size_t)gi * d + gj] - tj[tx][r]);
      mx = (x == x) ? fmax(mx, x) : HUGE_VAL;
    }
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    if (r > 0.0) atomicMax(out + b, (unsigned long long)__double_as_longlong(r));
  }
}

void launch_asym_max(const double* a, const int64_t* off, const int32_t* d, int nblk, int dmax, double* out,
                     hipStream_t s) {
  if (nblk <= 0) return;
  NSS_HIP(hipMemsetAsync(out, 0, (size_t)nblk * sizeof(double), s));
  const int nt = (dmax + ASYM_T - 1) / ASYM_T;
  hipLaunchKernelGGL(k_asym_max, dim3((unsigned)(nt * (nt + 1) / 2), (unsigned)nblk), dim3(256), 0, s, a, off, d,
                     reinterpret_cast<unsigned long long*>(out));
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

namespace nss {
// Plain streaming copy y = x (16-byte accesses): the known-size dispatch that
// HBM counters (FETCH_SIZE, WRITE_SIZE) are calibrated against, and a
// measured bandwidth ceiling for the streaming kernels.
__global__ __launch_bounds__(256) void k_copy_probe(const double2* __restrict__ x, double2* __restrict__ y,
                                                    int64_t n2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
}  // namespace nss

extern "C" int nss_probe_copy(int64_t n, int reps, double* seconds) {
  using namespace nss;
  try {
    NSS_REQUIRE(n > 0 && n % 2 == 0 && reps > 0 && seconds, NSS_E_ARG, "probe_copy: n must be even and > 0");
    double *dx = nullptr, *dy = nullptr;
    hipEvent_t e0, e1;
    NSS_HIP(hipMalloc(&dx, n * sizeof(double)));
    NSS_HIP(hipMalloc(&dy, n * sizeof(double)));
    NSS_HIP(hipMemset(dx, 0, n * sizeof(double)));
    NSS_HIP(hipEventCreate(&e0));
    NSS_HIP(hipEventCreate(&e1));
    const int64_t n2 = n / 2;
    const unsigned grid = (unsigned)std::min<int64_t>((n2 + 255) / 256, 16384);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, 0, (const double2*)dx, (double2*)dy, n2);
    NSS_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, 0, (const double2*)dx, (double2*)dy, n2);
    NSS_HIP(hipGetLastError());
    NSS_HIP(hipEventRecord(e1, 0));
    NSS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    NSS_HIP(hipEventElapsedTime(&ms, e0, e1));
    *seconds = 1e-3 * ms / reps;
    NSS_HIP(hipEventDestroy(e0));
    NSS_HIP(hipEventDestroy(e1));
    NSS_HIP(hipFree(dx));
    NSS_HIP(hipFree(dy));
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
  return NSS_OK;
}

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

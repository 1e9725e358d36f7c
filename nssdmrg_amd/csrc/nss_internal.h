// Internal declarations of the nssDMRG MI355X library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <functional>
#include <vector>

#include "../../include/nssdmrg.h"

namespace nss {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& msg);

#define NSS_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::nss::Error(NSS_E_HIP, std::string(#call " failed: ") + hipGetErrorString(e_) + \
                                        " at " __FILE__ ":" + std::to_string(__LINE__)); \
  } while (0)

#define NSS_REQUIRE(cond, code, msg)                 \
  do {                                               \
    if (!(cond)) throw ::nss::Error((code), (msg));  \
  } while (0)

// ------------------------------------------------------------- allocator
// Caching device allocator: power-of-two size classes, never returns memory
// to the driver until the context is destroyed (hipMalloc per DMRG step
// would dominate small steps).
class DevicePool {
 public:
  explicit DevicePool(int device) : device_(device) {}
  ~DevicePool();
  void* alloc(size_t bytes);
  void free(void* p);
  int device() const { return device_; }
  size_t bytes_in_use() const { return in_use_; }
  size_t bytes_reserved() const { return reserved_; }
  // Returns every cached (free) block to HIP.  Callers make sure no pending
  // device work still uses memory freed into the pool.
  void trim();

 private:
  int device_;
  size_t cache_cap_ = 0;  // cached bytes above which free() trims (a quarter of the device memory)
  std::map<size_t, std::vector<void*>> free_;
  std::map<void*, size_t> live_;
  size_t in_use_ = 0, reserved_ = 0;
};

template <typename T>
struct DevArr {  // owning device array backed by a pool
  DevicePool* pool = nullptr;
  T* p = nullptr;
  size_t n = 0;
  DevArr() = default;
  DevArr(const DevArr&) = delete;
  DevArr& operator=(const DevArr&) = delete;
  DevArr(DevArr&& o) noexcept { *this = std::move(o); }
  DevArr& operator=(DevArr&& o) noexcept {
    if (this != &o) {
      release();
      pool = o.pool; p = o.p; n = o.n;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DevArr() { release(); }
  void release() {
    if (p && pool) pool->free(p);
    p = nullptr; n = 0;
  }
  void resize(DevicePool* pl, size_t count) {
    if (count <= n && p) return;
    release();
    pool = pl;
    n = count;
    p = count ? static_cast<T*>(pool->alloc(count * sizeof(T))) : nullptr;
  }
};

// ------------------------------------------------------- grouped GEMM
// acc(tile) = sum_seg alpha_seg * opA(A_seg) . opB(B_seg); C(tile) = acc (+ C if beta).
// Pointers are (base selector, element offset): base 0 absolute address
// (offset ignored), 1 = X (input vector), 2 = W (scratch), 3 = Y (output).
enum : int { BASE_ABS = 0, BASE_X = 1, BASE_W = 2, BASE_Y = 3 };
enum : int { SEG_TA = 1, SEG_TB = 2 };

struct GemmSeg {
  const void* a;       // used when abase == BASE_ABS
  const void* b;
  int64_t aoff, boff;  // element offsets for relative bases
  int32_t lda, ldb, K, flags;
  int32_t abase, bbase;
  double alpha;
};

struct GemmTile {
  void* c;
  int64_t coff;
  int32_t ldc, M, N, row0, col0;
  int32_t seg0, nseg;
  int32_t cbase, beta;
};

// One work item of the large-output path: slices [c0, c1) of tile `tile`'s
// K chain (part `part` of nsplit; partial-tile slot pslot when nsplit > 1,
// the parts of a tile in consecutive slots).
struct BigItem {
  int32_t tile, part, nsplit, c0, c1, pslot;
};

struct BigArgs {
  const BigItem* items;
  int nitems, nbig;  // work items; large tiles at the front of the tile list
  void* part;
  unsigned* cnt;
};

struct Bases {
  const void* x;
  void* w;
  void* y;
  // H*v Y plans only (may be null): dot[4t + q] = sum over wave q's share
  // of output tile t of y . x at the same offsets (partial sums of x^T H x)
  double* dot = nullptr;
};

// Forced tile edge (NSS_GEMM_TILE = 32, 64 or 128), 0 for the size-based choice.
int gemm_default_tile();

// One grouped launch: tiles over outputs, segments over K-contributions.
struct GemmPlan {
  std::vector<GemmTile> htiles;
  std::vector<GemmSeg> hsegs;
  DevArr<GemmTile> tiles;
  DevArr<GemmSeg> segs;
  double flops = 0.0;
  // Output tile edge of the outputs not on the large-output path, chosen
  // in upload(): 32 -> k_* (32x32 tile, K chain split over the four waves
  // of a workgroup, reduced through LDS; for launches too small to fill the
  // chip); 64 -> k_*_t64 (64x64 tile, four waves share LDS-staged operands;
  // for large launches of small outputs).  NSS_GEMM_TILE forces either.
  int tile = 32;
  // Large-output path (k_*_t128): the outputs of at least 128x128 whose
  // edges pad little are cut into 128x128 tiles, the first nbig entries of
  // htiles, and their slice chains into work items (hbig) of about equal
  // length; the plan's other tiles are 32x32 and run in the same launch.
  // A tile cut into several parts sums them in part order through part
  // (partial tiles) and cnt (arrival counters, zero between launches).
  int nbig = 0;
  std::vector<BigItem> hbig;
  DevArr<BigItem> big;
  DevArr<double> part;
  DevArr<unsigned> cnt;
  std::vector<GemmTile> outs;  // one per add_output (row0 = col0 = 0)
  void clear() {
    htiles.clear();
    hsegs.clear();
    outs.clear();
    hbig.clear();
    nbig = 0;
    flops = 0.0;
    mode = 0;
  }
  // Adds output C (M x N, ldc) as tiles of the kernel's tile size, each with
  // the same list of segments.
  void add_output(int cbase, void* c, int64_t coff, int ldc, int M, int N,
                  const std::vector<GemmSeg>& segs_for_c, int beta = 0);
  // How upload() tiles the plan: PLAN_AUTO by sizes (outputs of 128x128 and
  // more on the large-output path with about 512 work items, the rest 32x32,
  // or 64x64 for launches of many small outputs), or one of the fixed modes
  // the autotuner measures (plan_autotune).
  int mode = 0;
  // esize: operand bytes of the launches (8 FP64, 4 FP32), which set the K
  // slice of the large-output path's work items
  void upload(DevicePool* pool, hipStream_t s, int esize = 8);
};

enum : int { PLAN_AUTO = 0, PLAN_T32 = 1, PLAN_T64 = 2, PLAN_BIG = 3, PLAN_NMODES = 4 };

// Uploads an H*v plan (kind GEMM_HV_W or GEMM_HV_Y, operands of esize
// bytes) in the mode measured fastest for its shape class: a class seen for
// the first time is timed in every candidate mode on the buffers of b
// (which the trial launches overwrite), each timed launch after prelude
// (enqueued on the context stream); later plans of the class reuse the
// cached mode.  Without measure, an uncached class takes the size-based mode.
void plan_autotune(nss_ctx* ctx, GemmPlan& p, int kind, int esize, Bases b, bool measure,
                   const std::function<void()>& prelude);
// The cached modes: count, and the entries as (kind, esize, log2 tiles,
// log2 edge, mode) rows.
int plan_tune_entries(int* rows, int maxrows);

constexpr int GEMM_BM = 64, GEMM_BN = 64;

enum : int { GEMM_GENERIC = 0, GEMM_HV_W = 1, GEMM_HV_Y = 2 };
// ev0/ev1 (optional): recorded by the launch itself at its start and end.
void launch_gemm_f64(const GemmPlan& p, Bases b, hipStream_t s, int kind = GEMM_GENERIC, hipEvent_t ev0 = nullptr,
                     hipEvent_t ev1 = nullptr);
void launch_gemm_f32(const GemmPlan& p, Bases b, hipStream_t s, int kind, hipEvent_t ev0 = nullptr,
                     hipEvent_t ev1 = nullptr);

// ------------------------------------------------------- vector kernels
void launch_fill_splitmix(double* v, int64_t n, uint64_t seed, hipStream_t s);
void launch_fill_const(double* v, int64_t n, double c, hipStream_t s);
void launch_d2f(const double* x, float* y, int64_t n, hipStream_t s);
void launch_f2d(const float* x, double* y, int64_t n, hipStream_t s);
void launch_scale(double* x, int64_t n, const double* num, const double* den, hipStream_t s);
// h[j] = <V_j, w> for j < ncols (V vectors of length n, stride ldv);
// partial scratch >= ncols * nchunks(n); result to h (device).
int64_t gemv_t_chunks(int64_t n);
void launch_gemv_t(const double* V, int64_t ldv, int ncols, const double* w, int64_t n,
                   double* partial, double* h, hipStream_t s);
// w -= V h  (or y = V h when overwrite)
void launch_gemv_n(const double* V, int64_t ldv, int ncols, const double* h, double* w, int64_t n,
                   int overwrite, hipStream_t s);
// One classical Gram-Schmidt pass against V[0..m): h = V^T w (to hout,
// device), w -= V h; with sq != nullptr, per-block sums of squares of the new
// w are left in sq[0..orth_blocks(n)).  With apart != nullptr the pass works
// on w - alpha va instead of w (alpha = the fixed-order sum of the lz_nch(n)
// partials in apart, stored to aout), writing that difference minus V h back
// to w: the three-term step's last subtraction folded into the pass.  No
// host synchronisation.
// With apart, the pass works on w - alpha v_a (- beta v_b with vb and bprev:
// beta = bprev[0], written to bout by block 0), alpha the sum of the na
// partials in apart (na < 0: lz_nch(n) of them), written to aout.
// m <= LZ_MAX_BASIS
constexpr int LZ_MAX_BASIS = 512;
void launch_orth_pass(const double* V, int64_t ldv, int m, double* w, int64_t n, double* partial,
                      double* hout, double* sq, hipStream_t s, const double* apart = nullptr,
                      const double* va = nullptr, double* aout = nullptr, int64_t na = -1,
                      const double* vb = nullptr, const double* bprev = nullptr, double* bout = nullptr);
int orth_blocks(int64_t n);
int64_t orth_chunk(int64_t n);
// First half of the Lanczos three-term step on the device: w -= beta_prev
// v_prev (when beta_prev != null) and the lz_nch(n) partial sums of
// alpha = <v_j, w> to apart; beta_prev copied to beta_out.  The subtraction
// of alpha v_j happens in the following launch_orth_pass.
void launch_lz_recurrence(double* w, const double* vprev, const double* vj, int64_t n, const double* beta_prev,
                          double* beta_out, double* apart, hipStream_t s);
// v = w / beta with beta = sqrt(sum sq) written to bout (device).
void launch_normalize(const double* w, double* v, int64_t n, const double* sq, double* bout, hipStream_t s);
// Batched Lanczos over independent superblocks ("chains"): one launch per
// vector operation covers every active chain.  Chain c owns the blocks
// [blk0, blk0 + nch) of a launch (bchain[b] = c), one per lz_nch chunk.
struct LzChainDev {
  double* V;        // (kmax + 1) x n Krylov basis
  double* w;        // H v output, then the orthogonalised residual
  double* x;        // H v input: copy of the newest basis vector (may be null: no copy)
  double* slots;    // iteration j at slots + j * sstride (single-chain slot layout, width 2 kb1 + 1)
  double* part;     // (kb1 + 2) x nch partial sums
  double* om;       // 3 x (kb1 + 1) orthogonality estimates: omega_{j,k} in row j % 3
  int32_t* st;      // [0] full Gram-Schmidt pass this iteration, [1] second pass pending
  double* bpre;     // |w| after the three-term step
  int64_t n;
  int64_t sstride;
  // rounding level of the H v products (unit of the orthogonality
  // estimates' rounding terms; the full pass runs above sqrt(eps))
  double eps;
  // estimate of the new vector's loss against the previous one
  // (omega_{j+1,j}) at each step
  double psi;
  int32_t nch, blk0, active, kmax;
};
int64_t lz_nch(int64_t n);
// One Lanczos iteration j after H v (w = H v_j), every active chain: the
// three-term step, one classical Gram-Schmidt pass against V[0..j], and
// v_{j+1} = w / |w| (also copied to x).  kb1 = batch kmax + 1.
// full != 0: the Gram-Schmidt pass in every iteration; otherwise only when
// the chain's orthogonality estimate asks for it (partial reorthogonalisation).
void launch_lzb_iteration(const LzChainDev* ch, const int32_t* bchain, int nblk, int nchain, int j, int kb1,
                          int full, hipStream_t s);
// out[0] = sqrt(<x,x>)
void launch_norm(const double* x, int64_t n, double* partial, double* out, hipStream_t s);
// y[i] = x[i] * (1/den[0])
void launch_div(const double* x, double* y, int64_t n, const double* den, hipStream_t s);

// Batched Jacobi eigen-solver for symmetric blocks of dimension <= nmax
// (nmax <= jacobi_max_dim): block b (dim[b] x dim[b] row-major at a+off[b])
// -> eigenvectors as columns at vout+off[b], eigenvalues (unsorted) at
// wout+eoff[b].
int jacobi_max_dim(int device);
// Tridiagonal-path eigen-solver for symmetric blocks of dimension 3..nmax
// (nmax <= rho_tri_max_dim): one workgroup per block, same layout and
// output as launch_jacobi_eig (eigenvalues unsorted across the blocks of
// its tridiagonal form); ws holds
// rho_tri_workspace(nblk) doubles.
int rho_tri_max_dim(int device);
int64_t rho_tri_workspace(int nblk);
void launch_rho_eig_tri(const double* a, const int64_t* off, const int32_t* dim, const int64_t* eoff, double* vout,
                        double* wout, double* ws, int nblk, int nmax, hipStream_t s);
void launch_jacobi_eig(const double* a, const int64_t* off, const int32_t* dim, const int64_t* eoff, double* vout,
                       double* wout, int nblk, int nmax, hipStream_t s);

// Large-block tridiagonal path (nss_rho_big.hip) for symmetric blocks of
// dimension up to RHO_BIG_MAX, batched over every block of a call.
constexpr int RHO_BIG_MAX = 2048;
struct TredBlk {
  double* a;    // n x n row-major symmetric block, reduced in place: afterwards
                // row k holds reflector k at columns k+1..n-1 (v[k+1] = 1)
  double* ws;   // tred_ws_size(n) doubles: v[2][n], p[2][n], d[n], e[n], tau[n]
  int32_t n;
  int32_t jt;   // columns 0..jt-1 by the per-column launches, the rest by the tail
};
int64_t tred_ws_size(int n);
// The tail of the per-column tridiagonalisation, one workgroup per block in
// LDS: the pending update of reflector jt-1, then columns jt..n-1
// (1 <= jt, 3 <= n - jt <= nmax <= 128).
void launch_tred_tail(const TredBlk* blks, int nblk, int nmax, hipStream_t s);
// Column j of the Householder tridiagonalisation of every block (launch j = 0
// .. (the largest jt) - 1 in order; nmax = the largest n of the batch; a
// block is left alone from its column jt on).
void launch_tred_col(const TredBlk* blks, int nblk, int j, int nmax, hipStream_t s);
// The whole tridiagonalisation in one cooperative launch with the trailing
// matrices in LDS (k_tred_lds): the plan gives each block that fits a
// number of workgroups (gcount) and maps workgroup ids to (block, index)
// (wgmap, XCD-aware); covered[b] = 0 leaves block b to launch_tred_col.
struct TredLdsPlan {
  std::vector<int2> wgmap;
  std::vector<int> gcount, covered;
  int grid = 0;
  size_t lds = 0;
};
bool tred_lds_plan(int device, const std::vector<int>& dims, TredLdsPlan* plan);
// bar: one zeroed counter per block; abort: zeroed, set when a barrier wait
// times out (the results are then invalid).  Returns 0 when the cooperative
// launch is refused (nothing ran).
int launch_tred_lds(const TredBlk* blks, const int2* wgmap, const int* gcnt, unsigned* bar, int* abort, int grid,
                    size_t lds, hipStream_t s);
// Every eigenvalue of every T (d, e in ws): position p to w[eoff[b] + p],
// ascending within each unreduced block of T.
void launch_tri_eigvals(const TredBlk* blks, const int64_t* eoff, double* w, int nblk, int nmax, hipStream_t s);
struct InvReq {        // one inverse-iteration vector
  const double* d;     // T's diagonal / off-diagonal
  const double* e;
  double* z;           // slab: row i, slot s at z[i * ldz + s] (zeroed by the caller)
  int64_t ldz;
  int32_t rs, re;      // the unreduced block holding the position
  int32_t slot, seedpos;
  double lam, tiny;    // shift, replacement for a zero pivot
};
void launch_tri_invit(const InvReq* req, int nreq, double* fac, int nmax, hipStream_t s);  // fac: 5 nmax nreq
struct GsCluster {     // slots s0..s1-1 of a slab, nonzero in rows rs..re-1
  double* z;
  int64_t ldz;
  int32_t rs, re, s0, s1;
};
void launch_tri_cluster_gs(const GsCluster* cl, int ncl, int nrmax, hipStream_t s);
struct BtTask {        // back-transformation of slots s0..s0+ns-1 of a slab
  const double* a;     // the reduced block (reflectors in its rows)
  const double* tau;
  double* z;
  int64_t ldz;
  int32_t n, s0, ns, pad;
};
int backtr_vectors_per_task(int n);
void launch_tri_backtr(const BtTask* tasks, int ntask, int nmax, hipStream_t s);

// Padding of blocks into ncls x ncls slabs for a strided-batched dense
// eigensolve, and the inverse copy of the eigenpairs (see nss_kernels.hip).
void launch_pad_blocks(const double* src, const int64_t* off, const int32_t* dim, int nb, int ncls, double* dst,
                       hipStream_t s);
void launch_unpad_eig(const double* vecs, const double* vals, int ncls, const int64_t* off, const int32_t* dim,
                      const int64_t* eoff, int nb, double* dst, double* w, hipStream_t s);

// out[b] = max_{i<j} |a_ij - a_ji| of the d[b] x d[b] row-major block at a + off[b]
// (dmax >= every d[b]; NaN differences report +inf)
void launch_asym_max(const double* a, const int64_t* off, const int32_t* d, int nblk, int dmax, double* out,
                     hipStream_t s);

// Kronecker-scatter for block enlargement: dst sub-block (rows x cols, ld)
// = sum of contributions scale * X (or X^T, or identity).
struct KronContrib {
  const double* x;   // nullptr -> identity
  int32_t ldx, trans;
  double scale;
};
struct KronTask {
  double* dst;
  int32_t ldd, rows, cols;
  int32_t c0, nc;
  int32_t row_t, col_t;  // tile origin within the task
};
void launch_kron(const KronTask* tasks, int ntasks, const KronContrib* contribs, hipStream_t s);

// ------------------------------------------------------------- context
struct HostMat {  // small host-side dense matrix (row-major)
  int r = 0, c = 0;
  std::vector<double> a;
};

struct SbDevice;  // superblock plan (nss_superblock.cpp)
constexpr int KRON_TILE = 32;  // k_kron task tile edge

}  // namespace nss


struct nss_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  nss::DevicePool* pool = nullptr;
  void* rocblas = nullptr;  // rocblas_handle, created lazily
  // auxiliary streams (default priority), each with its own rocblas handle,
  // created lazily: independent batches of the rho eigensolve run on them
  // concurrently with the context stream
  // [0, AUX_TRI): rocSOLVER rho classes; AUX_TRI: the one-workgroup
  // tridiagonal-path rho kernel; AUX_BIG: the large-block rho path
  static constexpr int NAUX = 5;
  static constexpr int AUX_TRI = 3;
  static constexpr int AUX_BIG = 4;
  hipStream_t aux[NAUX] = {};
  void* aux_rocblas[NAUX] = {};
  nss::SbDevice* sb = nullptr;
  // H*v timing with HIP events recorded by the W and Y launches themselves:
  // 4 events per timed call (W start/end, Y start/end).  An event-carrying
  // launch costs the host several microseconds more than a plain one, and
  // the host's enqueue rate bounds the Lanczos iterations of small
  // superblocks, so only every hv_every-th call is timed (NSS_HV_TIMING:
  // the period, 1 every call, 0 none; default 4).  st_all counts every call.
  bool time_hv = true;  // false: calls are neither timed nor counted (dense assembly)
  int hv_every = 4;
  int64_t hv_seq = 0;
  int64_t fl_all = 0, fl_timed = 0;  // calls since the last flush: all, timed
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  // per recorded call: flops of the W and Y launches
  std::vector<double> ev_fw, ev_fy;
  // accumulated seconds / flops, [0] FP64, [1] FP32
  double st_tw[2] = {0, 0}, st_ty[2] = {0, 0}, st_fw[2] = {0, 0}, st_fy[2] = {0, 0};
  int64_t st_calls[2] = {0, 0}, st_all[2] = {0, 0};
  std::vector<int> ev_kind;  // 0 fp64, 1 fp32 per recorded call
  // Lanczos convergence checks: pinned host copy of the iteration slots and
  // the event that marks it (created on first use, freed with the context)
  double* lz_pin = nullptr;
  size_t lz_pin_n = 0;
  hipEvent_t lz_ev = nullptr;
  // rho eigensolve: pinned host copy of the symmetry check's maxima, and the
  // event after which rho is complete (the auxiliary streams wait on it)
  double* rdm_pin = nullptr;
  size_t rdm_pin_n = 0;
  hipEvent_t rho_ev = nullptr;
  // engines created on this context: destroyed (their device memory
  // returned) before the context's pool goes away
  std::vector<struct nss_dmrg*> engines;
  // this device's copy of the file-scope scratch array of the allocation-free
  // H*v entry point (nss_sb_hv_device_noalloc), resolved once per device
  double* noalloc_scratch = nullptr;
  // phase timers (NSS_PROFILE=1): host wall time with a stream sync at the
  // end of each timed phase; off by default (no extra syncs)
  bool prof = false;
  std::map<std::string, double> timers;
};

namespace nss {
// Accumulates wall time into ctx->timers[phase] when profiling: a phase runs
// from construction or next() to the following next() or destruction.
struct PhaseTimer {
  ::nss_ctx* c;
  const char* name;
  double t0;
  PhaseTimer(::nss_ctx* ctx, const char* n);
  void next(const char* n);
  ~PhaseTimer();
};
}  // namespace nss

namespace nss {
void nss_release_rocblas(::nss_ctx* ctx);
// Frees the state of every engine still alive on ctx (nss_engine.cpp).
void engines_release(::nss_ctx* ctx);
// Synchronises the context stream, folds the recorded H*v events into the
// context statistics and returns the seconds spent in H*v since the last
// flush (the timed calls' seconds scaled by all calls / timed calls).
double nss_hv_timer_flush(::nss_ctx* ctx);

// Host symmetric eigen-solvers (nss_eig.cpp).
// Dense symmetric a (n x n row-major, overwritten by eigenvectors as columns),
// eigenvalues ascending in w.  Householder tridiagonalisation + implicit QL.
void sym_eig(int n, double* a, double* w);
// Symmetric tridiagonal (d[0..n), e[0..n-1)): eigenvalues ascending to w and
// the requested rows of the eigenvector matrix Z (T = Z diag(w) Z^T):
// zrows (nr x n, row-major) must enter holding the same rows of the identity.
// Lowest eigenvalue theta of the tridiagonal (d, e) and the last component of
// its unit eigenvector (sign arbitrary): the Lanczos residual estimate
// |beta_m z_last| in O(n) per bisection step instead of a full QL.  With
// zfull, the whole eigenvector (n entries) from the same inverse iteration.
void tridiag_lowest(int n, const double* d, const double* e, double* theta, double* zlast,
                    double* zfull = nullptr);
void tridiag_eig(int n, const double* d, const double* e, double* w, int nr, const int* rows,
                 double* zrows);
// Full eigenvectors of the tridiagonal (z n x n row-major, columns = vectors).
void tridiag_eig_full(int n, const double* d, const double* e, double* w, double* z);

// Lanczos on a context's current superblock (nss_lanczos.cpp).
struct LanczosOut {
  std::vector<double> evals;
  int numop = 0, numop_fp32 = 0;
};
// Eigenvectors stay on the device in xdev (neigen x n).
LanczosOut lanczos_device(nss_ctx* ctx, SbDevice* sb, int neigen, int kmax, int maxiter, double tol,
                          uint64_t seed, int v0_mode, int precision, double* xdev);
// One job of a batched ground-state solve: an uploaded superblock, the
// start vector (v0_mode 2: in xdev) and the resulting eigenvector (xdev).
struct LanczosJob {
  SbDevice* sb = nullptr;
  double* xdev = nullptr;
  int v0_mode = 0;
  uint64_t seed = 0;
  LanczosOut out;
};
// Lowest eigenpair of every job, the H*v of all chains in one grouped launch
// per phase (nss_lanczos.cpp).
void lanczos_batch(::nss_ctx* ctx, std::vector<LanczosJob>& jobs, int kmax, int maxiter, double tol);
// Mixed-precision cycles end when their residual estimate has fallen by
// this factor below the cycle's first FP64 residual (NSS_MIXED_INNER, default 1e-5).
double lz_mixed_inner();
// Superblock dimension from which a single engine's FP64 lowest-eigenpair
// solve uses lanczos_batch (partial reorthogonalisation) instead of the
// thick-restart solver with a Gram-Schmidt pass every iteration
// (NSS_LZ_PRO_MIN_N; 0 for every size).
int64_t lz_pro_min_n();
// Smallest superblock a mixed-precision engine solves with FP32-MFMA H*v
// (NSS_MIXED_MIN_N, default 2^17, read on every call).
int64_t lz_mixed_min_n();
// FP64 thick-restart iterations take alpha_j from the Y launch's partial
// sums of v_j . H v_j (NSS_LZ_ALPHA_FROM_Y; 0 for a separate dot pass).
bool lz_alpha_from_y();
// Thick-restart iterations: the iteration after a convergence check is
// enqueued before the host waits on the check (NSS_LZ_CHECK_AHEAD; 0 waits first).
bool lz_check_ahead();
// Dense path: H_sb assembled on the device from unit vectors, diagonalised on the host.
LanczosOut dense_diag(nss_ctx* ctx, SbDevice* sb, int neigen, double* xdev);

// Superblock plan API (nss_superblock.cpp).
struct OpRef {  // a dense operand: device pointer, leading dim, transpose flag, scale
  const void* p = nullptr;
  const void* p32 = nullptr;  // FP32 mirror (mixed precision), may be null
  int32_t ld = 0;
  int32_t trans = 0;
  double alpha = 1.0;
};
struct SbTerm {
  int dst, src;
  OpRef A;   // Dl_dst x Dl_src
  OpRef Bt;  // B^T : Dr_src x Dr_dst
};
SbDevice* sb_create(nss_ctx* ctx);
void sb_destroy(SbDevice* sb);
void sb_reset(SbDevice* sb);  // forget layout and plans (operands owned elsewhere may be gone)
void sb_set_layout(SbDevice* sb, const std::vector<int>& Dl, const std::vector<int>& Dr,
                   const std::vector<int64_t>* offset);
// HL, HR per sector (square, row-major refs), terms; builds W/Y plans.
void sb_set_terms(nss_ctx* ctx, SbDevice* sb, const std::vector<OpRef>& HL,
                  const std::vector<OpRef>& HR, const std::vector<SbTerm>& terms, bool want_fp32);
void sb_hv(nss_ctx* ctx, SbDevice* sb, const double* x, double* y);          // device, async
// sb_hv that also leaves, for wave q of every output tile t of the Y launch,
// the partial sum dot[4t + q] of y . x over that wave's share of the tile
// (4 x sb_hv_ntiles of them; summed in a fixed order they give x^T H x
// deterministically).
void sb_hv_dot(nss_ctx* ctx, SbDevice* sb, const double* x, double* y, double* dot);
int64_t sb_hv_ntiles(const SbDevice* sb);
void sb_hv_fp32(nss_ctx* ctx, SbDevice* sb, const double* x, double* y);     // device, async
// Batch H*v: merged W and Y plans of several superblocks whose vectors sit at
// xoff[c] of shared X / Y arenas and whose term scratch sits at woff[c] of a
// shared W arena (sb_wsize elements each); hv_plans launches them (timed).
// f32: the plans of the FP32 operand copies (arenas of floats at the same
// offsets), launched by hv_plans32.
int64_t sb_wsize(const SbDevice* sb);
void sb_merge_hv_plans(const std::vector<SbDevice*>& sbs, const std::vector<int64_t>& xoff,
                       const std::vector<int64_t>& woff, GemmPlan& W, GemmPlan& Y, ::nss_ctx* ctx, hipStream_t s,
                       bool f32 = false);
void hv_plans(::nss_ctx* ctx, const GemmPlan& W, const GemmPlan& Y, Bases b);
void hv_plans32(::nss_ctx* ctx, const GemmPlan& W, const GemmPlan& Y, Bases b);
int64_t sb_n(const SbDevice* sb);
double sb_flops(const SbDevice* sb);
int sb_nq(const SbDevice* sb);
const std::vector<int>& sb_Dl(const SbDevice* sb);
const std::vector<int>& sb_Dr(const SbDevice* sb);
const std::vector<int64_t>& sb_off(const SbDevice* sb);
// rho blocks of a device vector: side 0 left, 1 right -> host eigen-decomposition.
void sb_rdm_eigh(nss_ctx* ctx, SbDevice* sb, const double* psi_dev, int side,
                 std::vector<std::vector<double>>& evals, std::vector<std::vector<double>>& evecs);
// both sides at once (sides bit 0: left, bit 1: right)
// Device view of one rho block's eigenvectors: component i of eigenvector
// j at base[i*si + j*sj]; order[k] = index j of the k-th smallest eigenvalue.
struct RhoDeferred;
struct RhoVec {
  const double* base = nullptr;
  int d = 0;
  int64_t si = 0, sj = 0;
  std::vector<int> order;
  // a block of the large-block path whose eigenvectors are not computed yet
  // (base == nullptr): rho_materialize computes those of the largest
  // eigenvalues on request
  std::shared_ptr<RhoDeferred> def;
  int def_blk = -1;
};
// Computes the eigenvectors of the count[i] largest eigenvalues of each
// deferred view vecs[i] (count 0: left deferred) and points the views at
// them (valid while the views live).  Enqueued on the context stream.
void rho_materialize(nss_ctx* ctx, const std::vector<RhoVec*>& vecs, const std::vector<int>& count);
// rho_L / rho_R blocks of psi (sides: bit 0 left, bit 1 right) and their
// eigenpairs, eigenvalues ascending on the host.  With dvecs == nullptr the
// eigenvectors come to the host (evecs: row-major, eigenvectors as columns);
// otherwise they stay in the superblock's rho buffer (valid until the next
// call) and dvecs[side][q] describes them.
void sb_rdm_eigh2(nss_ctx* ctx, SbDevice* sb, const double* psi_dev, int sides,
                  std::vector<std::vector<double>> evals[2], std::vector<std::vector<double>> evecs[2],
                  std::vector<RhoVec>* dvecs = nullptr, bool defer = false);
// Several superblocks at once (independent chains): job j's rho blocks of
// psi_j, its eigenvalues to evals[2], eigenvectors to evecs[2] (host) or
// dvecs[2] (views into rho, valid until rho is reused).  One GEMM launch,
// one Jacobi launch and one rocSOLVER batch per size class for all jobs.
struct RdmJob {
  SbDevice* sb = nullptr;
  const double* psi = nullptr;
  int sides = 3;
  std::vector<std::vector<double>>* evals = nullptr;  // [2]
  std::vector<std::vector<double>>* evecs = nullptr;  // [2], or nullptr with dvecs
  std::vector<RhoVec>* dvecs = nullptr;               // [2], or nullptr
  // with dvecs: blocks of the large-block path may stay deferred (their
  // eigenvectors computed later, for the kept states only)
  bool defer = false;
};
void sb_rdm_eigh_multi(::nss_ctx* ctx, std::vector<RdmJob>& jobs, DevArr<double>& rho, GemmPlan& R);
struct GatherTask {  // dst[r*ncol + c] = src[r*si + cols[c]*sj]
  double* dst;
  const double* src;
  int64_t si, sj;
  const int32_t* cols;
  int32_t rows, ncol;
};
void launch_gather_cols(const GatherTask* tasks, int ntasks, int max_elems, hipStream_t s);

}  // namespace nss

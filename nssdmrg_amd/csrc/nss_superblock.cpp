// Context, allocator, grouped-GEMM plans and the superblock H*v of the
// nssDMRG MI355X library.  H*v mirrors spMatVec_direct_main
// (src/DMRG_SUPERBLOCK_SETUP.f90:740-865) as two grouped launches:
//   W phase: W_t = Psi_src(t) . B_t^T              (one tile set per term)
//   Y phase: Y_q = H_L,q Psi_q + Psi_q H_R,q^T + sum_{t->q} A_t W_t
// which is the reference's evaluation order (B first, then A).
#include <algorithm>
#include <functional>
#include <map>
#include <cfloat>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "nss_internal.h"

namespace nss {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------- pool
static size_t size_class(size_t b) {
  size_t c = 256;
  while (c < b) c <<= 1;
  return c;
}

void* DevicePool::alloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  const size_t c = size_class(bytes);
  void* p = nullptr;
  auto it = free_.find(c);
  if (it != free_.end() && !it->second.empty()) {
    p = it->second.back();
    it->second.pop_back();
  } else {
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      // release cached blocks and retry once
      for (auto& kv : free_)
        for (void* q : kv.second) {
          (void)hipFree(q);
          reserved_ -= kv.first;
        }
      free_.clear();
      e = hipMalloc(&p, c);
      if (e != hipSuccess)
        throw Error(NSS_E_ALLOC, "device allocation of " + std::to_string(c) + " bytes failed: " +
                                     hipGetErrorString(e));
    }
    reserved_ += c;
  }
  live_[p] = c;
  in_use_ += c;
  return p;
}

void DevicePool::free(void* p) {
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_[it->second].push_back(p);
  in_use_ -= it->second;
  live_.erase(it);
  if (cache_cap_ == 0) {
    size_t fr = 0, tot = 0;
    cache_cap_ = hipMemGetInfo(&fr, &tot) == hipSuccess && tot > 0 ? tot / 4 : (size_t(16) << 30);
  }
  // the runtime allocates memory of its own at kernel launches; a cache that
  // keeps most of the device would starve it
  if (reserved_ - in_use_ > cache_cap_) trim();
}

void DevicePool::trim() {
  for (auto& kv : free_)
    for (void* q : kv.second) {
      (void)hipFree(q);
      reserved_ -= kv.first;
    }
  free_.clear();
}

DevicePool::~DevicePool() {
  for (auto& kv : free_)
    for (void* q : kv.second) (void)hipFree(q);
  for (auto& kv : live_) (void)hipFree(kv.first);
}

// ---------------------------------------------------------------- plans
// Read at every plan upload (tests switch them per case).
int gemm_default_tile() {
  const char* e = std::getenv("NSS_GEMM_TILE");
  const int v = e ? std::atoi(e) : 0;
  return (v == 32 || v == 64 || v == 128) ? v : 0;
}

// NSS_GEMM_SPLIT = 1..16 forces the number of parts of every 128x128 tile
// of the large-output path (capped at its slice count).
static int gemm_forced_split() {
  const char* e = std::getenv("NSS_GEMM_SPLIT");
  const int v = e ? std::atoi(e) : 0;
  return (v >= 1 && v <= 16) ? v : 0;
}

void GemmPlan::add_output(int cbase, void* c, int64_t coff, int ldc, int M, int N,
                          const std::vector<GemmSeg>& sg, int beta) {
  if (M <= 0 || N <= 0) return;
  const int seg0 = (int)hsegs.size();
  int nseg = 0;
  for (const GemmSeg& g : sg) {
    if (g.K <= 0) continue;
    hsegs.push_back(g);
    ++nseg;
    flops += 2.0 * M * (double)N * g.K;
  }
  GemmTile t{};
  t.c = c;
  t.coff = coff;
  t.ldc = ldc;
  t.M = M;
  t.N = N;
  t.seg0 = seg0;
  t.nseg = nseg;
  t.cbase = cbase;
  t.beta = beta;
  outs.push_back(t);
}

// Launches with at least this many 32x32 tiles (two 64x64 tiles per CU) use
// the 64x64 LDS-staged kernel: on large launches its operand sharing across
// the four waves wins (spin-1 m=1024: 29.7 vs 21.9 TFLOP/s FP64); on smaller
// ones the split K chain of the 32x32 kernel does.
static constexpr int64_t GEMM_T64_MIN_TILES = 2048;

// Element offsets of a segment's operands within the output's tiles fit the
// large-output path's 32-bit arithmetic.
static bool big_offsets_fit(const GemmTile& o, const std::vector<GemmSeg>& segs) {
  for (int q = o.seg0; q < o.seg0 + o.nseg; ++q) {
    const GemmSeg& g = segs[q];
    const bool ta = g.flags & SEG_TA, tb = g.flags & SEG_TB;
    const int64_t a = (int64_t)(o.M - 1) * (ta ? 1 : g.lda) + (int64_t)(g.K - 1) * (ta ? g.lda : 1);
    const int64_t b = (int64_t)(o.N - 1) * (tb ? g.ldb : 1) + (int64_t)(g.K - 1) * (tb ? 1 : g.ldb);
    if (a >= INT32_MAX || b >= INT32_MAX) return false;
  }
  return true;
}

// Outputs of at least 128x128 whose 128x128 tiling pads at most 30% of
// their area take the large-output path (k_*_t128).
static bool big_output(const GemmTile& o) {
  if (o.M < 128 || o.N < 128) return false;
  const double padded = (double)((o.M + 127) / 128 * 128) * (double)((o.N + 127) / 128 * 128);
  return padded <= 1.3 * (double)o.M * (double)o.N;
}

// The slice chains of the large-output tiles are cut into work items of
// about equal length, about this many per launch (two per CU of the chip's
// 256; NSS_GEMM_BIG_ITEMS overrides), at most 16 parts per tile and at least
// 8 slices per part.
static int64_t gemm_big_items() {
  const char* e = std::getenv("NSS_GEMM_BIG_ITEMS");
  const int64_t v = e ? std::atoll(e) : 0;
  return v >= 8 ? v : 512;
}

void GemmPlan::upload(DevicePool* pool, hipStream_t s, int esize) {
  // NSS_GEMM_TILE overrides the plan's mode (set by the autotuner), which
  // overrides the size-based choice
  int forced = gemm_default_tile();
  const int64_t nit = gemm_big_items();
  if (!forced && mode != PLAN_AUTO) forced = mode == PLAN_T32 ? 32 : mode == PLAN_T64 ? 64 : 0;
  const bool nobig = forced == 32 || forced == 64;
  const int bk = esize == 8 ? 16 : 32;  // K slice of the large-output path
  std::vector<char> isbig(outs.size(), 0);
  int64_t n32 = 0;
  bool anybig = false;
  for (size_t i = 0; i < outs.size(); ++i) {
    const GemmTile& o = outs[i];
    isbig[i] = (forced == 128 || (!nobig && big_output(o))) && big_offsets_fit(o, hsegs);
    anybig = anybig || isbig[i];
    if (!isbig[i]) n32 += (int64_t)((o.M + 31) / 32) * ((o.N + 31) / 32);
  }
  tile = anybig ? 32 : nobig ? forced : (n32 >= GEMM_T64_MIN_TILES ? 64 : 32);
  htiles.clear();
  hbig.clear();
  std::vector<int64_t> tsl;  // slices of each large tile's chain
  for (size_t i = 0; i < outs.size(); ++i) {
    if (!isbig[i]) continue;
    const GemmTile& o = outs[i];
    int64_t sl = 0;
    for (int q = o.seg0; q < o.seg0 + o.nseg; ++q) sl += (hsegs[q].K + bk - 1) / bk;
    for (int r = 0; r < o.M; r += 128)
      for (int cc = 0; cc < o.N; cc += 128) {
        GemmTile t = o;
        t.row0 = r;
        t.col0 = cc;
        htiles.push_back(t);
        tsl.push_back(sl);
      }
  }
  nbig = (int)htiles.size();
  for (size_t i = 0; i < outs.size(); ++i) {
    if (isbig[i]) continue;
    const GemmTile& o = outs[i];
    for (int r = 0; r < o.M; r += tile)
      for (int cc = 0; cc < o.N; cc += tile) {
        GemmTile t = o;
        t.row0 = r;
        t.col0 = cc;
        htiles.push_back(t);
      }
  }
  if (nbig > 0) {
    int64_t total = 0;
    for (int64_t v : tsl) total += v;
    const int64_t len = std::max<int64_t>(8, (total + nit - 1) / nit);
    const int fs = gemm_forced_split();
    int pslot = 0;
    for (int t = 0; t < nbig; ++t) {
      const int64_t sl = tsl[t];
      int ns = fs ? fs : (int)std::clamp<int64_t>((sl + len / 2) / len, 1, 16);
      ns = (int)std::max<int64_t>(1, std::min<int64_t>(ns, sl));
      for (int p = 0; p < ns; ++p)
        hbig.push_back(BigItem{t, p, ns, (int)(sl * p / ns), (int)(sl * (p + 1) / ns), ns > 1 ? pslot + p : 0});
      if (ns > 1) pslot += ns;
    }
    big.resize(pool, hbig.size());
    NSS_HIP(hipMemcpyAsync(big.p, hbig.data(), hbig.size() * sizeof(BigItem), hipMemcpyHostToDevice, s));
    part.resize(pool, (size_t)std::max(1, pslot) * 16384);
    cnt.resize(pool, nbig);
    NSS_HIP(hipMemsetAsync(cnt.p, 0, (size_t)nbig * sizeof(unsigned), s));
  }
  tiles.resize(pool, std::max<size_t>(1, htiles.size()));
  segs.resize(pool, std::max<size_t>(1, hsegs.size()));
  if (!htiles.empty())
    NSS_HIP(hipMemcpyAsync(tiles.p, htiles.data(), htiles.size() * sizeof(GemmTile),
                           hipMemcpyHostToDevice, s));
  if (!hsegs.empty())
    NSS_HIP(hipMemcpyAsync(segs.p, hsegs.data(), hsegs.size() * sizeof(GemmSeg),
                           hipMemcpyHostToDevice, s));
  // host vectors must outlive the async copies: callers keep the plan alive
  // and synchronize before mutating it.
  NSS_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- autotuner
// The tiling mode of an H*v plan is measured, not guessed: the first plan of
// a shape class is uploaded and launched in each candidate mode (twice
// untimed, then four times back to back, each launch timed with HIP events
// on the context stream and preceded by a prelude: for Y its W launch) and
// the fastest mode is cached for the class.  The size-based mode stays
// unless another is more than 10% faster: isolated trial launches overstate
// the large-output path (its split-K partials and counters start cold), by
// up to 45% against the same plan inside a Lanczos run.  Shape class: phase
// (W or Y), operand bytes, and the number of 32x32 tiles and the largest
// output edge, both in quarter octaves (ceil(4 log2)).
// NSS_GEMM_TUNE_TRACE=1 prints every measurement on stderr.  Candidates:
// 32x32 tiles; 64x64 tiles (launches of at least 256 tiles of 32x32); the
// large-output path (plans with an output it takes).  NSS_GEMM_AUTOTUNE=0
// keeps the size-based mode; NSS_GEMM_TILE forces a tile for every plan.
namespace {
struct TuneKey {
  int kind, esize, lt, lm;
  bool operator<(const TuneKey& o) const {
    return kind != o.kind ? kind < o.kind : esize != o.esize ? esize < o.esize : lt != o.lt ? lt < o.lt : lm < o.lm;
  }
};
std::mutex tune_mu;
std::map<TuneKey, int> tune_cache;

// ceil(4 log2 v): quarter octaves
int ceil_log2q(int64_t v) {
  return v <= 1 ? 0 : (int)std::ceil(4.0 * std::log2((double)v) - 1e-9);
}

bool autotune_enabled() {
  const char* e = std::getenv("NSS_GEMM_AUTOTUNE");
  return !(e && e[0] == '0');
}
}  // namespace

void plan_autotune(nss_ctx* ctx, GemmPlan& p, int kind, int esize, Bases b, bool measure,
                   const std::function<void()>& prelude) {
  hipStream_t s = ctx->stream;
  auto launch = [&]() {
    if (esize == 8) launch_gemm_f64(p, b, s, kind);
    else launch_gemm_f32(p, b, s, kind);
  };
  if (gemm_default_tile() || !autotune_enabled() || p.outs.empty()) {
    p.mode = PLAN_AUTO;
    p.upload(ctx->pool, s, esize);
    return;
  }
  int64_t n32 = 0;
  int mx = 0;
  bool anybig = false;
  for (const GemmTile& o : p.outs) {
    n32 += (int64_t)((o.M + 31) / 32) * ((o.N + 31) / 32);
    mx = std::max(mx, std::max(o.M, o.N));
    anybig = anybig || (big_output(o) && big_offsets_fit(o, p.hsegs));
  }
  const TuneKey key{kind, esize, ceil_log2q(n32), ceil_log2q(mx)};
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_cache.find(key);
    if (it != tune_cache.end() || !measure) {
      p.mode = it != tune_cache.end() ? it->second : PLAN_AUTO;
      p.upload(ctx->pool, s, esize);
      return;
    }
  }
  std::vector<int> cand{PLAN_T32};
  if (n32 >= 256) cand.push_back(PLAN_T64);
  if (anybig) cand.push_back(PLAN_BIG);
  // the size-based mode, kept unless another is clearly faster
  const int dflt = anybig ? PLAN_BIG : n32 >= GEMM_T64_MIN_TILES ? PLAN_T64 : PLAN_T32;
  int best = dflt;
  if (cand.size() > 1) {
    hipEvent_t ev[8];
    for (hipEvent_t& e : ev) NSS_HIP(hipEventCreate(&e));
    std::vector<float> t(cand.size(), 0.f);
    for (size_t c = 0; c < cand.size(); ++c) {
      p.mode = cand[c];
      p.upload(ctx->pool, s, esize);
      // two untimed calls (the plan's fresh tables and scratch reach the TLB
      // and caches, the clocks leave the idle of the upload), then four
      // back-to-back calls timed one by one
      for (int rep = 0; rep < 6; ++rep) {
        if (prelude) prelude();
        if (rep >= 2) NSS_HIP(hipEventRecord(ev[2 * (rep - 2)], s));
        launch();
        if (rep >= 2) NSS_HIP(hipEventRecord(ev[2 * (rep - 2) + 1], s));
      }
      NSS_HIP(hipEventSynchronize(ev[7]));
      for (int r = 0; r < 4; ++r) {
        float ms = 0.f;
        NSS_HIP(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]));
        t[c] += 0.5f * ms;
      }
    }
    for (hipEvent_t& e : ev) (void)hipEventDestroy(e);
    float td = -1.f, tb = 0.f;
    int bi = 0;
    for (size_t c = 0; c < cand.size(); ++c) {
      if (cand[c] == dflt) td = t[c];
      if (c == 0 || t[c] < tb) {
        tb = t[c];
        bi = (int)c;
      }
    }
    if (!(td >= 0.f && td <= 1.1f * tb)) best = cand[bi];
    if (const char* tr = std::getenv("NSS_GEMM_TUNE_TRACE"); tr && tr[0] == '1') {
      std::fprintf(stderr, "gemm tune kind %d bytes %d class (%d, %d) tiles32 %lld:", kind, esize, key.lt, key.lm,
                   (long long)n32);
      for (size_t c = 0; c < cand.size(); ++c) std::fprintf(stderr, " mode %d %.1f us", cand[c], 500.f * t[c]);
      std::fprintf(stderr, " -> %d\n", best);
    }
  }
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache[key] = best;
  }
  p.mode = best;
  p.upload(ctx->pool, s, esize);
}

int plan_tune_entries(int* rows, int maxrows) {
  std::lock_guard<std::mutex> lk(tune_mu);
  int r = 0;
  for (const auto& kv : tune_cache) {
    if (rows && r < maxrows) {
      int* o = rows + 5 * r;
      o[0] = kv.first.kind;
      o[1] = kv.first.esize;
      o[2] = kv.first.lt;
      o[3] = kv.first.lm;
      o[4] = kv.second;
    }
    ++r;
  }
  return r;
}

// ---------------------------------------------------------------- superblock
struct SbDevice {
  std::vector<int> Dl, Dr;
  std::vector<int64_t> off;
  int64_t n = 0;
  GemmPlan W64, Y64, W32, Y32;
  DevArr<double> W;
  DevArr<float> Wf, Xf, Yf;
  int64_t wsize = 0;
  double flops = 0.0;
  bool has32 = false;
  // storage of the low-level upload path
  DevArr<double> owned;
  DevArr<float> owned32;
  // rho scratch
  DevArr<double> rho;
  GemmPlan R;
};

SbDevice* sb_create(nss_ctx*) { return new SbDevice(); }
void sb_reset(SbDevice* sb) {
  sb->Dl.clear();
  sb->Dr.clear();
  sb->off.clear();
  sb->n = 0;
  sb->W64.clear();
  sb->Y64.clear();
  sb->W32.clear();
  sb->Y32.clear();
  sb->has32 = false;
  sb->flops = 0.0;
}
void sb_destroy(SbDevice* sb) { delete sb; }
int64_t sb_n(const SbDevice* sb) { return sb->n; }
double sb_flops(const SbDevice* sb) { return sb->flops; }
int sb_nq(const SbDevice* sb) { return (int)sb->Dl.size(); }
const std::vector<int>& sb_Dl(const SbDevice* sb) { return sb->Dl; }
const std::vector<int>& sb_Dr(const SbDevice* sb) { return sb->Dr; }
const std::vector<int64_t>& sb_off(const SbDevice* sb) { return sb->off; }

void sb_set_layout(SbDevice* sb, const std::vector<int>& Dl, const std::vector<int>& Dr,
                   const std::vector<int64_t>* offset) {
  NSS_REQUIRE(Dl.size() == Dr.size() && !Dl.empty(), NSS_E_ARG, "sb_set_layout: nq must be > 0");
  std::vector<int64_t> off(Dl.size(), 0);
  int64_t acc = 0;
  for (size_t q = 0; q < Dl.size(); ++q) {  // validate before touching the current state
    NSS_REQUIRE(Dl[q] >= 0 && Dr[q] >= 0, NSS_E_ARG, "sb_set_layout: negative sector dimension");
    if (offset) {
      NSS_REQUIRE((*offset)[q] == acc, NSS_E_ARG,
                  "sb_set_layout: offsets must be the prefix sums of Dl*Dr (Offset, "
                  "DMRG_SUPERBLOCK.f90:72)");
    }
    off[q] = acc;
    acc += (int64_t)Dl[q] * Dr[q];
  }
  sb->Dl = Dl;
  sb->Dr = Dr;
  sb->off = off;
  sb->n = acc;
  sb->W64.clear();
  sb->Y64.clear();
  sb->W32.clear();
  sb->Y32.clear();
  sb->has32 = false;
  sb->flops = 0.0;
}

static GemmSeg seg_abs(const void* a, int lda, int ta, const void* b, int ldb, int tb, int K,
                       double alpha) {
  GemmSeg g{};
  g.a = a; g.b = b; g.aoff = 0; g.boff = 0;
  g.lda = lda; g.ldb = ldb; g.K = K;
  g.flags = (ta ? SEG_TA : 0) | (tb ? SEG_TB : 0);
  g.abase = BASE_ABS; g.bbase = BASE_ABS;
  g.alpha = alpha;
  return g;
}

static void build_hv_plans(SbDevice* sb, const std::vector<OpRef>& HL, const std::vector<OpRef>& HR,
                           const std::vector<SbTerm>& terms, bool f32, GemmPlan& Wp, GemmPlan& Yp) {
  Wp.clear();
  Yp.clear();
  const int nq = (int)sb->Dl.size();
  std::vector<int64_t> woff(terms.size(), 0);
  int64_t w = 0;
  for (size_t t = 0; t < terms.size(); ++t) {
    woff[t] = w;
    w += (int64_t)sb->Dl[terms[t].src] * sb->Dr[terms[t].dst];
  }
  sb->wsize = w;
  auto P = [&](const OpRef& r) { return f32 ? r.p32 : r.p; };
  // W phase: W_t = Psi_src . B_t^T   (Dl_src x Dr_dst, K = Dr_src)
  for (size_t t = 0; t < terms.size(); ++t) {
    const SbTerm& T = terms[t];
    const int dls = sb->Dl[T.src], drs = sb->Dr[T.src], drd = sb->Dr[T.dst];
    if (dls == 0 || drs == 0 || drd == 0 || sb->Dl[T.dst] == 0) continue;
    GemmSeg g = seg_abs(nullptr, drs, 0, P(T.Bt), T.Bt.ld, T.Bt.trans, drs, 1.0);
    g.abase = BASE_X;
    g.aoff = sb->off[T.src];
    Wp.add_output(BASE_W, nullptr, woff[t], drd, dls, drd, {g});
  }
  // Y phase
  std::vector<std::vector<size_t>> into(nq);
  for (size_t t = 0; t < terms.size(); ++t) into[terms[t].dst].push_back(t);
  for (int q = 0; q < nq; ++q) {
    const int dl = sb->Dl[q], dr = sb->Dr[q];
    if (dl == 0 || dr == 0) continue;
    std::vector<GemmSeg> sg;
    GemmSeg a = seg_abs(P(HL[q]), HL[q].ld, HL[q].trans, nullptr, dr, 0, dl, HL[q].alpha);
    a.bbase = BASE_X;
    a.boff = sb->off[q];
    sg.push_back(a);
    GemmSeg b = seg_abs(nullptr, dr, 0, P(HR[q]), HR[q].ld, !HR[q].trans, dr, HR[q].alpha);
    b.abase = BASE_X;
    b.aoff = sb->off[q];
    sg.push_back(b);
    for (size_t t : into[q]) {
      const SbTerm& T = terms[t];
      const int dls = sb->Dl[T.src];
      if (dls == 0 || sb->Dr[T.src] == 0) continue;
      GemmSeg c = seg_abs(P(T.A), T.A.ld, T.A.trans, nullptr, dr, 0, dls, T.A.alpha);
      c.bbase = BASE_W;
      c.boff = woff[t];
      sg.push_back(c);
    }
    Yp.add_output(BASE_Y, nullptr, sb->off[q], dr, dl, dr, sg);
  }
}

void sb_set_terms(nss_ctx* ctx, SbDevice* sb, const std::vector<OpRef>& HL, const std::vector<OpRef>& HR,
                  const std::vector<SbTerm>& terms, bool want_fp32) {
  const int nq = (int)sb->Dl.size();
  NSS_REQUIRE((int)HL.size() == nq && (int)HR.size() == nq, NSS_E_ARG, "sb_set_terms: HL/HR count");
  for (const SbTerm& t : terms)
    NSS_REQUIRE(t.dst >= 0 && t.dst < nq && t.src >= 0 && t.src < nq, NSS_E_ARG,
                "sb_set_terms: term sector index out of range");
  build_hv_plans(sb, HL, HR, terms, false, sb->W64, sb->Y64);
  sb->flops = sb->W64.flops + sb->Y64.flops;
  sb->W.resize(ctx->pool, std::max<int64_t>(1, sb->wsize));
  {
    // the autotuner's trial launches read x and W and write W and y: scratch
    // vectors (the operator blocks are the superblock's own).  A Y trial
    // follows its W launch, as in every H*v.  (Trials from evicted caches
    // ranked the modes differently from the H*v calls of a run, whose
    // operator blocks stay partly cached; back-to-back trials rank them as
    // the runs do.)
    const int64_t nx = std::max<int64_t>(1, sb->n);
    DevArr<double> tx, ty;
    tx.resize(ctx->pool, nx);
    ty.resize(ctx->pool, nx);
    launch_fill_const(tx.p, nx, 0.5, ctx->stream);
    const Bases tb{tx.p, sb->W.p, ty.p};
    plan_autotune(ctx, sb->W64, GEMM_HV_W, 8, tb, true, nullptr);
    plan_autotune(ctx, sb->Y64, GEMM_HV_Y, 8, tb, true,
                  [&]() { launch_gemm_f64(sb->W64, tb, ctx->stream, GEMM_HV_W); });
    NSS_HIP(hipStreamSynchronize(ctx->stream));
  }
  sb->has32 = false;
  if (want_fp32) {
    build_hv_plans(sb, HL, HR, terms, true, sb->W32, sb->Y32);
    sb->Wf.resize(ctx->pool, std::max<int64_t>(1, sb->wsize));
    sb->Xf.resize(ctx->pool, std::max<int64_t>(1, sb->n));
    sb->Yf.resize(ctx->pool, std::max<int64_t>(1, sb->n));
    NSS_HIP(hipMemsetAsync(sb->Xf.p, 0, std::max<int64_t>(1, sb->n) * sizeof(float), ctx->stream));
    const Bases tb{sb->Xf.p, sb->Wf.p, sb->Yf.p};
    plan_autotune(ctx, sb->W32, GEMM_HV_W, 4, tb, true, nullptr);
    plan_autotune(ctx, sb->Y32, GEMM_HV_Y, 4, tb, true,
                  [&]() { launch_gemm_f32(sb->W32, tb, ctx->stream, GEMM_HV_W); });
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    sb->has32 = true;
  }
}

static hipEvent_t next_event(nss_ctx* ctx) {
  if (ctx->ev_used == ctx->ev.size()) {
    // Timing-only events: no system-scope fence (no L2 writeback between the
    // W and Y launches, and tighter bracketing of each launch).
    hipEvent_t e;
    NSS_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    ctx->ev.push_back(e);
  }
  return ctx->ev[ctx->ev_used++];
}


// Timed H*v: each launch records its own start and end (4 events per call),
// so the statistics are the kernels' durations and no marker packet sits
// between them.
static void hv_timed(nss_ctx* ctx, const GemmPlan& W, const GemmPlan& Y, Bases b, int kind,
                     double* ydot = nullptr) {
  bool t = false;
  if (ctx->time_hv) {
    ++ctx->st_all[kind];
    ++ctx->fl_all;
    t = ctx->hv_every > 0 && ctx->hv_seq++ % ctx->hv_every == 0;
    ctx->fl_timed += t;
  }
  hipEvent_t e[4] = {};
  if (t)
    for (hipEvent_t& x : e) x = next_event(ctx);
  if (kind == 0) {
    launch_gemm_f64(W, b, ctx->stream, GEMM_HV_W, e[0], e[1]);
    Bases by = b;
    by.dot = ydot;  // Y's per-tile x . y partials
    launch_gemm_f64(Y, by, ctx->stream, GEMM_HV_Y, e[2], e[3]);
  } else {
    launch_gemm_f32(W, b, ctx->stream, GEMM_HV_W, e[0], e[1]);
    launch_gemm_f32(Y, b, ctx->stream, GEMM_HV_Y, e[2], e[3]);
  }
  if (t) {
    // an empty plan launches nothing: its events are recorded as markers
    if (W.htiles.empty()) {
      NSS_HIP(hipEventRecord(e[0], ctx->stream));
      NSS_HIP(hipEventRecord(e[1], ctx->stream));
    }
    if (Y.htiles.empty()) {
      NSS_HIP(hipEventRecord(e[2], ctx->stream));
      NSS_HIP(hipEventRecord(e[3], ctx->stream));
    }
    ctx->ev_kind.push_back(kind);
    ctx->ev_fw.push_back(W.flops);
    ctx->ev_fy.push_back(Y.flops);
  }
}

void sb_hv(nss_ctx* ctx, SbDevice* sb, const double* x, double* y) {
  hv_timed(ctx, sb->W64, sb->Y64, Bases{x, sb->W.p, y}, 0);
}

void sb_hv_dot(nss_ctx* ctx, SbDevice* sb, const double* x, double* y, double* dot) {
  hv_timed(ctx, sb->W64, sb->Y64, Bases{x, sb->W.p, y}, 0, dot);
}

int64_t sb_hv_ntiles(const SbDevice* sb) { return (int64_t)sb->Y64.htiles.size(); }

void sb_hv_fp32(nss_ctx* ctx, SbDevice* sb, const double* x, double* y) {
  NSS_REQUIRE(sb->has32, NSS_E_STATE, "FP32 H*v requested but blocks were uploaded for FP64 only");
  launch_d2f(x, sb->Xf.p, sb->n, ctx->stream);
  hv_timed(ctx, sb->W32, sb->Y32, Bases{sb->Xf.p, sb->Wf.p, sb->Yf.p}, 1);
  launch_f2d(sb->Yf.p, y, sb->n, ctx->stream);
}

// Merged H*v plans of several superblocks for one W and one Y launch: chain
// c's vector operands (x, y) sit at xoff[c] of the batch's X / Y arenas and
// its term scratch at woff[c] of the W arena; operator blocks are absolute.
int64_t sb_wsize(const SbDevice* sb) { return sb->wsize; }

static void plan_append(GemmPlan& dst, const GemmPlan& src, int64_t xoff, int64_t woff) {
  auto shift = [&](int base) -> int64_t {
    return base == BASE_X || base == BASE_Y ? xoff : base == BASE_W ? woff : 0;
  };
  const int seg0 = (int)dst.hsegs.size();
  for (GemmSeg g : src.hsegs) {
    g.aoff += shift(g.abase);
    g.boff += shift(g.bbase);
    dst.hsegs.push_back(g);
  }
  for (GemmTile t : src.outs) {
    t.coff += shift(t.cbase);
    t.seg0 += seg0;
    dst.outs.push_back(t);
  }
  dst.flops += src.flops;
}

void sb_merge_hv_plans(const std::vector<SbDevice*>& sbs, const std::vector<int64_t>& xoff,
                       const std::vector<int64_t>& woff, GemmPlan& W, GemmPlan& Y, nss_ctx* ctx, hipStream_t s,
                       bool f32) {
  W.clear();
  Y.clear();
  for (size_t c = 0; c < sbs.size(); ++c) {
    NSS_REQUIRE(!f32 || sbs[c]->has32, NSS_E_STATE, "FP32 H*v requested but blocks were uploaded for FP64 only");
    plan_append(W, f32 ? sbs[c]->W32 : sbs[c]->W64, xoff[c], woff[c]);
    plan_append(Y, f32 ? sbs[c]->Y32 : sbs[c]->Y64, xoff[c], woff[c]);
  }
  // tiling modes from the autotuner's cache (a merged plan of one chain has
  // that chain's shape class), size-based for classes not measured yet
  (void)s;
  plan_autotune(ctx, W, GEMM_HV_W, f32 ? 4 : 8, Bases{}, false, nullptr);
  plan_autotune(ctx, Y, GEMM_HV_Y, f32 ? 4 : 8, Bases{}, false, nullptr);
}

// H*v of merged plans, timed like sb_hv.
void hv_plans(nss_ctx* ctx, const GemmPlan& W, const GemmPlan& Y, Bases b) { hv_timed(ctx, W, Y, b, 0); }
void hv_plans32(nss_ctx* ctx, const GemmPlan& W, const GemmPlan& Y, Bases b) { hv_timed(ctx, W, Y, b, 1); }

double nss_hv_timer_flush(nss_ctx* ctx) {
  NSS_HIP(hipStreamSynchronize(ctx->stream));
  double tot = 0.0;
  for (size_t c = 0; c < ctx->ev_kind.size(); ++c) {
    float a = 0.f, b = 0.f;
    NSS_HIP(hipEventElapsedTime(&a, ctx->ev[4 * c], ctx->ev[4 * c + 1]));
    NSS_HIP(hipEventElapsedTime(&b, ctx->ev[4 * c + 2], ctx->ev[4 * c + 3]));
    const double w = a * 1e-3, y = b * 1e-3;
    const int k = ctx->ev_kind[c];
    tot += w + y;
    ctx->st_tw[k] += w;
    ctx->st_ty[k] += y;
    ctx->st_fw[k] += ctx->ev_fw[c];
    ctx->st_fy[k] += ctx->ev_fy[c];
    ctx->st_calls[k]++;
  }
  ctx->ev_used = 0;
  ctx->ev_kind.clear();
  ctx->ev_fw.clear();
  ctx->ev_fy.clear();
  const double scale = ctx->fl_timed > 0 ? (double)ctx->fl_all / (double)ctx->fl_timed : 0.0;
  ctx->fl_all = ctx->fl_timed = 0;
  return tot * scale;
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

PhaseTimer::PhaseTimer(nss_ctx* ctx, const char* n) : c(ctx), name(n), t0(ctx->prof ? now_s() : 0.0) {}

void PhaseTimer::next(const char* n) {
  if (c->prof) {
    (void)hipStreamSynchronize(c->stream);
    const double t = now_s();
    c->timers[name] += t - t0;
    t0 = t;
  }
  name = n;
}

PhaseTimer::~PhaseTimer() {
  if (!c->prof) return;
  (void)hipStreamSynchronize(c->stream);
  c->timers[name] += now_s() - t0;
}

void nss_release_rocblas(nss_ctx* ctx) {
  for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
  ctx->ev.clear();
  if (ctx->lz_ev) (void)hipEventDestroy(ctx->lz_ev);
  ctx->lz_ev = nullptr;
  if (ctx->lz_pin) (void)hipHostFree(ctx->lz_pin);
  ctx->lz_pin = nullptr;
  ctx->lz_pin_n = 0;
  if (ctx->rho_ev) (void)hipEventDestroy(ctx->rho_ev);
  ctx->rho_ev = nullptr;
  if (ctx->rdm_pin) (void)hipHostFree(ctx->rdm_pin);
  ctx->rdm_pin = nullptr;
  ctx->rdm_pin_n = 0;
  if (ctx->rocblas) {
    (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->rocblas));
    ctx->rocblas = nullptr;
  }
  for (int k = 0; k < nss_ctx::NAUX; ++k) {
    if (ctx->aux_rocblas[k]) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->aux_rocblas[k]));
    ctx->aux_rocblas[k] = nullptr;
    if (ctx->aux[k]) {
      (void)hipStreamSynchronize(ctx->aux[k]);
      (void)hipStreamDestroy(ctx->aux[k]);
    }
    ctx->aux[k] = nullptr;
  }
}

// Largest rho block sent to the batched Jacobi kernel: NSS_JACOBI_MAX,
// default 32.  Blocks above it up to NSS_RHO_TRI_MAX (default 128, 0 for
// none) go to the tridiagonal-path kernel, larger ones to rocSOLVER; both on
// the auxiliary streams.
static int jacobi_limit() {
  static const int v = [] {
    const char* e = std::getenv("NSS_JACOBI_MAX");
    return e ? std::max(0, std::atoi(e)) : 32;
  }();
  return v;
}

// Largest rho block of the large-block tridiagonal path (NSS_RHO_BIG_MAX,
// default and cap RHO_BIG_MAX, 0 for none); larger blocks go to rocSOLVER.
static int rho_big_limit() {
  static const int v = [] {
    const char* e = std::getenv("NSS_RHO_BIG_MAX");
    return e ? std::clamp(std::atoi(e), 0, RHO_BIG_MAX) : RHO_BIG_MAX;
  }();
  return v;
}

// Number of large blocks in one rho call from which blocks of up to 768 go to
// rocSOLVER instead of the large-block path (NSS_RHO_BIG_MANY, default 64).
// Trailing size below which the large-block tridiagonalisation finishes in
// one workgroup per block (NSS_TRED_TAIL, default 128, capped by the LDS of
// the one-workgroup kernel; below 3: every column by its own launch)
static int tred_tail_size() {
  static const int v = [] {
    const char* e = std::getenv("NSS_TRED_TAIL");
    return e ? std::atoi(e) : 128;
  }();
  return v;
}

// Largest block that a call with many large blocks sends to rocSOLVER
// (NSS_RHO_MANY_MID, default RHO_BIG_MAX: all of them; read once)
static int rho_many_mid() {
  static const int v = [] {
    const char* e = std::getenv("NSS_RHO_MANY_MID");
    return e ? std::atoi(e) : RHO_BIG_MAX;
  }();
  return v;
}

static int rho_big_many() {
  static const int v = [] {
    const char* e = std::getenv("NSS_RHO_BIG_MANY");
    return e ? std::max(1, std::atoi(e)) : 64;
  }();
  return v;
}

static int rho_tri_limit() {
  static const int v = [] {
    const char* e = std::getenv("NSS_RHO_TRI_MAX");
    return e ? std::max(0, std::atoi(e)) : 128;
  }();
  return v;
}

// Padded size of a block in the batched rocSOLVER rho eigensolves.  Up to
// 256 (rocSOLVER's single-kernel tridiagonalisation) padding costs flops, so
// classes are multiples of 64 (NSS_RHO_CLASS overrides the granularity);
// above it the one-stage tridiagonalisation launches a few kernels per column
// of every class, so all larger blocks share one class, the largest of them
// rounded up to 64 (NSS_RHO_CLASS_LARGE > 0 sets a granularity instead).
static int rho_class(int d, int dlarge) {
  static const int small = [] {
    const char* e = std::getenv("NSS_RHO_CLASS");
    const int g = e ? std::atoi(e) : 64;
    return g >= 16 ? g : 64;
  }();
  static const int large = [] {
    const char* e = std::getenv("NSS_RHO_CLASS_LARGE");
    return e ? std::max(0, std::atoi(e)) : 0;
  }();
  if (d <= 256) return (d + small - 1) / small * small;
  if (large > 0) return (d + large - 1) / large * large;
  if (dlarge <= 0) return (d + 63) / 64 * 64;
  return (dlarge + 63) / 64 * 64;
}

// Bytes of one rocSOLVER call's padded slab: larger classes are split into
// several calls (many chains' large blocks in one rho batch would otherwise
// exhaust the device memory with slab and solver workspace).
constexpr int64_t RHO_SLAB_BYTES = int64_t(1) << 30;

// Auxiliary stream k with its rocblas handle bound to it (created on first use).
// NSS_TRED_LDS=1 sends the large blocks that fit the chip's LDS to the
// one-launch LDS-resident tridiagonalisation; by default every large block
// takes the per-column kernel, which measured faster on the spin-1 m=1024
// workload (about 8 us per column each; the LDS path's cross-XCD barrier
// costs more than a kernel boundary).  Read on every call.
static bool tred_lds_enabled() {
  const char* e = std::getenv("NSS_TRED_LDS");
  return e && std::atoi(e) == 1;
}

static rocblas_handle aux_handle(nss_ctx* ctx, int k, hipStream_t* st) {
  if (!ctx->aux[k]) NSS_HIP(hipStreamCreateWithFlags(&ctx->aux[k], hipStreamNonBlocking));
  if (!ctx->aux_rocblas[k]) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw Error(NSS_E_HIP, "rocblas_create_handle failed");
    ctx->aux_rocblas[k] = h;
    if (rocblas_set_stream(h, ctx->aux[k]) != rocblas_status_success)
      throw Error(NSS_E_HIP, "rocblas_set_stream failed");
  }
  *st = ctx->aux[k];
  return static_cast<rocblas_handle>(ctx->aux_rocblas[k]);
}

// ---------------------------------------------------------------- large rho blocks
// Blocks of the large-block path after tridiagonalisation: their eigenvalues
// are known, their eigenvectors are computed on request (rho_materialize).
// The reduced blocks (reflectors in their rows) stay in the rho buffer of
// the call that made them.
struct RhoDeferred {
  nss_ctx* ctx = nullptr;
  std::vector<TredBlk> blk;                // device addresses of the blocks and workspaces
  std::vector<int64_t> hwo;                // offset of block k's workspace in ws / hws
  std::vector<double> hws;                 // host copy of the workspaces (d at 4n, e at 5n)
  std::vector<std::vector<double>> lam;    // eigenvalue at each position of block k
  DevArr<double> ws;
  // materialised eigenvectors and the tables of their launches (kept until
  // the views go away)
  struct Batch {
    std::vector<DevArr<double>> slabs;
    DevArr<double> fac;
    DevArr<InvReq> dreq;
    DevArr<GsCluster> dcl;
    DevArr<BtTask> dbt;
    std::vector<InvReq> hreq;
    std::vector<GsCluster> hcl;
    std::vector<BtTask> hbt;
  };
  std::vector<std::unique_ptr<Batch>> done;
};

void rho_materialize(nss_ctx* ctx, const std::vector<RhoVec*>& vecs, const std::vector<int>& count) {
  auto B = std::make_unique<RhoDeferred::Batch>();
  RhoDeferred* owner = nullptr;
  int nmax = 0;
  for (size_t i = 0; i < vecs.size(); ++i) {
    RhoVec& V = *vecs[i];
    const int k = count[i];
    if (!V.def || V.base || k <= 0) continue;
    NSS_REQUIRE(k <= V.d && V.def_blk >= 0, NSS_E_STATE, "rho_materialize: bad request");
    RhoDeferred& D = *V.def;
    owner = owner ? owner : &D;
    const int b = V.def_blk, n = V.d;
    const double* d = D.hws.data() + D.hwo[b] + 4 * (int64_t)n;
    const double* e = D.hws.data() + D.hwo[b] + 5 * (int64_t)n;
    const std::vector<double>& lam = D.lam[b];
    // ||T|| and the unreduced blocks, as the eigenvalue kernel splits T
    double tn = 0.0;
    for (int r = 0; r < n; ++r)
      tn = std::max(tn, std::fabs(d[r]) + (r > 0 ? std::fabs(e[r - 1]) : 0.0) + (r + 1 < n ? std::fabs(e[r]) : 0.0));
    tn = std::max(tn, DBL_MIN);
    std::vector<int> bs(n), be(n);
    for (int r = 0; r < n; ++r) bs[r] = (r > 0 && !(std::fabs(e[r - 1]) <= DBL_EPSILON * tn)) ? bs[r - 1] : r;
    for (int r = n - 1; r >= 0; --r) be[r] = (r + 1 < n && !(std::fabs(e[r]) <= DBL_EPSILON * tn)) ? be[r + 1] : r + 1;
    // the k largest eigenvalues' positions, ascending: slot s holds pos[s]
    std::vector<int> pos(V.order.end() - k, V.order.end());
    std::sort(pos.begin(), pos.end());
    DevArr<double> slab;
    slab.resize(ctx->pool, (size_t)n * k);
    NSS_HIP(hipMemsetAsync(slab.p, 0, (size_t)n * k * sizeof(double), ctx->stream));
    const double tiny = std::max(DBL_EPSILON * tn, 1e-280);
    const double ctol = 1e-3 * tn;
    const TredBlk& T = D.blk[b];
    const double* dd = T.ws + 4 * (int64_t)n;
    const double* de = T.ws + 5 * (int64_t)n;
    int c0 = 0;
    for (int sl = 0; sl < k; ++sl) {
      const int p = pos[sl];
      InvReq q{};
      q.d = dd;
      q.e = de;
      q.z = slab.p;
      q.ldz = k;
      q.rs = bs[p];
      q.re = be[p];
      q.slot = sl;
      q.seedpos = p;
      q.lam = lam[p];
      q.tiny = tiny;
      B->hreq.push_back(q);
      // clusters: consecutive kept positions of one unreduced block whose
      // eigenvalues lie within 1e-3 ||T|| (dstein's criterion)
      const bool joins = sl > 0 && bs[pos[sl - 1]] == bs[p] && lam[p] - lam[pos[sl - 1]] <= ctol;
      if (!joins) {
        if (sl - c0 > 1) B->hcl.push_back({slab.p, (int64_t)k, bs[pos[c0]], be[pos[c0]], c0, sl});
        c0 = sl;
      }
    }
    if (k - c0 > 1) B->hcl.push_back({slab.p, (int64_t)k, bs[pos[c0]], be[pos[c0]], c0, k});
    const int vpt = backtr_vectors_per_task(n);
    for (int s0 = 0; s0 < k; s0 += vpt) {
      BtTask t{};
      t.a = T.a;
      t.tau = T.ws + 6 * (int64_t)n;
      t.z = slab.p;
      t.ldz = k;
      t.n = n;
      t.s0 = s0;
      t.ns = std::min(vpt, k - s0);
      B->hbt.push_back(t);
    }
    nmax = std::max(nmax, n);
    // the view: slot of the c-th largest eigenvalue
    V.base = slab.p;
    V.si = k;
    V.sj = 1;
    for (int c = 0; c < k; ++c) {
      const int p = V.order[n - 1 - c];
      V.order[n - 1 - c] = (int)(std::lower_bound(pos.begin(), pos.end(), p) - pos.begin());
    }
    B->slabs.push_back(std::move(slab));
  }
  if (!owner) return;
  hipStream_t s = ctx->stream;
  const int nreq = (int)B->hreq.size(), ncl = (int)B->hcl.size(), nbt = (int)B->hbt.size();
  // one backtransformation launch serves one register shape: n <= 1024 or above
  std::vector<BtTask> lo, hi;
  for (const BtTask& t : B->hbt) (t.n <= 1024 ? lo : hi).push_back(t);
  B->hbt = lo;
  B->hbt.insert(B->hbt.end(), hi.begin(), hi.end());
  B->fac.resize(ctx->pool, (size_t)5 * nmax * nreq);
  B->dreq.resize(ctx->pool, nreq);
  B->dcl.resize(ctx->pool, std::max(1, ncl));
  B->dbt.resize(ctx->pool, nbt);
  NSS_HIP(hipMemcpyAsync(B->dreq.p, B->hreq.data(), nreq * sizeof(InvReq), hipMemcpyHostToDevice, s));
  if (ncl) NSS_HIP(hipMemcpyAsync(B->dcl.p, B->hcl.data(), ncl * sizeof(GsCluster), hipMemcpyHostToDevice, s));
  NSS_HIP(hipMemcpyAsync(B->dbt.p, B->hbt.data(), nbt * sizeof(BtTask), hipMemcpyHostToDevice, s));
  if (ctx->prof)
    for (const GsCluster& c : B->hcl) {
      const double nv = c.s1 - c.s0;
      ctx->timers["rdm.gs_clusters"] += 1;
      ctx->timers["rdm.gs_vectors"] += nv;
      ctx->timers["rdm.gs_work_c2n"] += nv * nv * (c.re - c.rs);
      ctx->timers["rdm.gs_cmax"] = std::max(ctx->timers["rdm.gs_cmax"], nv);
    }
  launch_tri_invit(B->dreq.p, nreq, B->fac.p, nmax, s);
  int nrmax = 1;
  for (const GsCluster& c : B->hcl) nrmax = std::max(nrmax, c.re - c.rs);
  launch_tri_cluster_gs(B->dcl.p, ncl, nrmax, s);
  if (!lo.empty()) launch_tri_backtr(B->dbt.p, (int)lo.size(), 1024, s);
  if (!hi.empty()) launch_tri_backtr(B->dbt.p + lo.size(), (int)hi.size(), RHO_BIG_MAX, s);
  owner->done.push_back(std::move(B));
}

// rho blocks of both sides of one or several superblocks in one batch:
// items (job, side, sector) share one rho buffer, one GEMM launch, one
// symmetry check, one Jacobi launch and one strided-batched rocSOLVER call
// per size class.
void sb_rdm_eigh_multi(nss_ctx* ctx, std::vector<RdmJob>& jobs, DevArr<double>& rho, GemmPlan& R) {
  struct Item {
    int job, side, q, d;
    int64_t roff, eoff;
    int kind;  // 1 jacobi, 2 rocsolver, 3 tridiagonal path (one workgroup), 4 large-block path
  };
  std::vector<Item> items;
  int64_t tot = 0;
  for (int jb = 0; jb < (int)jobs.size(); ++jb) {
    RdmJob& J = jobs[jb];
    const SbDevice* sb = J.sb;
    const int nq = (int)sb->Dl.size();
    for (int side = 0; side < 2; ++side) {
      J.evals[side].assign(nq, {});
      if (J.evecs) J.evecs[side].assign(nq, {});
      if (!(J.sides & (1 << side))) continue;
      for (int q = 0; q < nq; ++q) {
        if ((int64_t)sb->Dl[q] * sb->Dr[q] == 0) continue;
        const int d = side == 0 ? sb->Dl[q] : sb->Dr[q];
        items.push_back({jb, side, q, d, tot, -1, 1});
        tot += (int64_t)d * d;
      }
    }
  }
  const int ni = (int)items.size();
  rho.resize(ctx->pool, std::max<int64_t>(1, tot));
  PhaseTimer pt(ctx, "rdm.gemm");
  R.clear();
  for (const Item& it : items) {
    const SbDevice* sb = jobs[it.job].sb;
    const int dl = sb->Dl[it.q], dr = sb->Dr[it.q];
    GemmSeg g{};
    g.alpha = 1.0;
    g.abase = BASE_ABS;
    g.bbase = BASE_ABS;
    g.a = g.b = jobs[it.job].psi + sb->off[it.q];
    g.lda = g.ldb = dr;
    if (it.side == 0) {  // rho_L = Psi Psi^T : K = Dr
      g.K = dr;
      g.flags = SEG_TB;
    } else {  // rho_R = Psi^T Psi : K = Dl
      g.K = dl;
      g.flags = SEG_TA;
    }
    R.add_output(BASE_Y, nullptr, it.roff, it.d, it.d, it.d, {g});
  }
  R.upload(ctx->pool, ctx->stream);
  Bases b{nullptr, nullptr, rho.p};
  launch_gemm_f64(R, b, ctx->stream, GEMM_GENERIC);
  // build_density_matrix's symmetry invariant (DMRG_RDM.f90:159-160), on the
  // device; its maxima reach pinned host memory behind the eigensolves and
  // are checked before any result of this call is used
  pt.next("rdm.symcheck");
  std::vector<int64_t> hoff;
  std::vector<int32_t> hdim;
  int dchk = 0;
  for (const Item& it : items)
    if (it.d > 1) {
      hoff.push_back(it.roff);
      hdim.push_back(it.d);
      dchk = std::max(dchk, it.d);
    }
  const int nchk = (int)hoff.size();
  DevArr<int64_t> doff;
  DevArr<int32_t> ddim;
  DevArr<double> dasym;
  if (ctx->rdm_pin_n < (size_t)std::max(1, nchk)) {
    if (ctx->rdm_pin) NSS_HIP(hipHostFree(ctx->rdm_pin));
    ctx->rdm_pin = nullptr;
    ctx->rdm_pin_n = 0;
    const size_t n = (size_t)std::max(64, 2 * nchk);
    NSS_HIP(hipHostMalloc((void**)&ctx->rdm_pin, n * sizeof(double), hipHostMallocDefault));
    ctx->rdm_pin_n = n;
  }
  if (!ctx->rho_ev) NSS_HIP(hipEventCreateWithFlags(&ctx->rho_ev, hipEventDisableTiming));
  if (nchk) {
    doff.resize(ctx->pool, nchk);
    ddim.resize(ctx->pool, nchk);
    dasym.resize(ctx->pool, nchk);
    NSS_HIP(hipMemcpyAsync(doff.p, hoff.data(), nchk * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    NSS_HIP(hipMemcpyAsync(ddim.p, hdim.data(), nchk * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    launch_asym_max(rho.p, doff.p, ddim.p, nchk, dchk, dasym.p, ctx->stream);
    NSS_HIP(hipMemcpyAsync(ctx->rdm_pin, dasym.p, nchk * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  NSS_HIP(hipEventRecord(ctx->rho_ev, ctx->stream));
  // the check itself, once the context stream has passed the copy
  auto check_symmetric = [&]() {
    for (int c = 0; c < nchk; ++c)
      if (!(ctx->rdm_pin[c] <= 1e-12)) throw Error(NSS_E_NUMERIC, "build_density_matrix error: rho not Symmetric");
  };
  // eigen-decompositions: blocks up to the LDS limit (the 1x1 and tiny
  // ones included) with the batched GPU Jacobi kernel, larger ones with
  // rocSOLVER dsyevd; device results overwrite rho in place
  pt.next("rdm.jacobi");
  const int jmax = std::min(jacobi_max_dim(ctx->device), jacobi_limit());
  if (jmax < 2) throw Error(NSS_E_HIP, "rho eigensolve: the device's LDS cannot hold a Jacobi block");
  const int tmax = std::min(rho_tri_max_dim(ctx->device), rho_tri_limit());
  const int bmax = rho_big_limit();
  // many large blocks in one call (the rho of a batch of chains): those up
  // to rho_many_mid() (by default all) go to the batched rocSOLVER classes,
  // which beat one tridiagonalisation launch per column carrying all of
  // them; the few large blocks of one chain (m=1024) take the large-block
  // path
  int mid = 0;
  {
    int nbig = 0;
    for (const Item& it : items) nbig += it.d > tmax && it.d <= bmax;
    if (nbig >= rho_big_many()) mid = rho_many_mid();
  }
  int64_t etot = 0;
  std::vector<int64_t> joff, jeoff, toff, teoff;
  std::vector<int32_t> jdim, tdim;
  int jn = 0, tn = 0;
  for (Item& it : items) {
    it.eoff = etot;
    etot += it.d;
    it.kind = it.d <= jmax ? 1 : it.d <= tmax ? 3 : it.d <= mid ? 2 : it.d <= bmax ? 4 : 2;
    if (ctx->prof) {
      ctx->timers[it.kind == 1   ? "rdm.count_jacobi"
                  : it.kind == 3 ? "rdm.count_tri"
                  : it.kind == 4 ? "rdm.count_big"
                                 : "rdm.count_rocsolver"] += 1;
      if (it.kind == 2) ctx->timers["rdm.sum_dim_rocsolver"] += it.d;
    }
    if (it.kind == 1) {
      joff.push_back(it.roff);
      jeoff.push_back(it.eoff);
      jdim.push_back(it.d);
      jn = std::max(jn, it.d);
    } else if (it.kind == 3) {
      toff.push_back(it.roff);
      teoff.push_back(it.eoff);
      tdim.push_back(it.d);
      tn = std::max(tn, it.d);
    }
  }
  DevArr<double> dD;
  if (etot) dD.resize(ctx->pool, etot);
  DevArr<int64_t> djoff, djeoff;
  DevArr<int32_t> djdim;
  if (!jdim.empty()) {
    const int nb = (int)jdim.size();
    djoff.resize(ctx->pool, nb);
    djeoff.resize(ctx->pool, nb);
    djdim.resize(ctx->pool, nb);
    NSS_HIP(hipMemcpyAsync(djoff.p, joff.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    NSS_HIP(hipMemcpyAsync(djeoff.p, jeoff.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    NSS_HIP(hipMemcpyAsync(djdim.p, jdim.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    launch_jacobi_eig(rho.p, djoff.p, djdim.p, djeoff.p, rho.p, dD.p, nb, jn, ctx->stream);
  }
  // blocks above the Jacobi limit up to tmax: the tridiagonal-path kernel on
  // the last auxiliary stream, concurrently with the Jacobi launch and the
  // rocSOLVER classes (after rho_ev)
  // every auxiliary stream used here is drained before this function
  // returns or throws: the pool does not track streams, so buffers freed on
  // the way out must not be in use there
  struct AuxDrain {
    nss_ctx* c;
    bool used[nss_ctx::NAUX] = {};
    ~AuxDrain() {
      for (int k = 0; k < nss_ctx::NAUX; ++k)
        if (used[k] && c->aux[k]) (void)hipStreamSynchronize(c->aux[k]);
    }
  } drain{ctx};
  bool* used_aux = drain.used;
  DevArr<int64_t> dtoff, dteoff;
  DevArr<int32_t> dtdim;
  DevArr<double> tws;
  if (!tdim.empty()) {
    const int nb = (int)tdim.size();
    hipStream_t st = nullptr;
    (void)aux_handle(ctx, nss_ctx::AUX_TRI, &st);
    NSS_HIP(hipStreamWaitEvent(st, ctx->rho_ev, 0));
    used_aux[nss_ctx::AUX_TRI] = true;
    dtoff.resize(ctx->pool, nb);
    dteoff.resize(ctx->pool, nb);
    dtdim.resize(ctx->pool, nb);
    tws.resize(ctx->pool, rho_tri_workspace(nb));
    NSS_HIP(hipMemcpyAsync(dtoff.p, toff.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, st));
    NSS_HIP(hipMemcpyAsync(dteoff.p, teoff.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, st));
    NSS_HIP(hipMemcpyAsync(dtdim.p, tdim.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
    launch_rho_eig_tri(rho.p, dtoff.p, dtdim.p, dteoff.p, rho.p, dD.p, tws.p, nb, tn, st);
  }
  // blocks above the Jacobi limit: rocSOLVER dsyevd, strided-batched per
  // size class (dims rounded up to a multiple of 64, padded slabs).  The
  // classes run on the auxiliary streams, concurrently with the Jacobi
  // launch on the context stream (each auxiliary stream waits on rho_ev,
  // after which rho is complete; every batch writes disjoint rows of rho and
  // of the eigenvalue array).
  // blocks above tmax up to bmax: the large-block path on its own auxiliary
  // stream (Householder tridiagonalisation, one launch per column over all of
  // them, then every eigenvalue); their eigenvectors come later
  std::vector<int> bigs;
  for (int i = 0; i < ni; ++i)
    if (items[i].kind == 4) bigs.push_back(i);
  auto def = std::make_shared<RhoDeferred>();
  DevArr<TredBlk> dtb, dtbr;
  DevArr<int2> dtlm;
  DevArr<int> dtlg;
  std::function<bool()> big_abort;
  DevArr<int64_t> dbeo;
  std::vector<int64_t> hbeo;
  if (!bigs.empty()) {
    pt.next("rdm.big");
    hipStream_t st = nullptr;
    (void)aux_handle(ctx, nss_ctx::AUX_BIG, &st);
    NSS_HIP(hipStreamWaitEvent(st, ctx->rho_ev, 0));
    used_aux[nss_ctx::AUX_BIG] = true;
    const int nb = (int)bigs.size();
    int64_t wtot = 0;
    int nbmax = 0;
    for (int i : bigs) {
      def->hwo.push_back(wtot);
      wtot += tred_ws_size(items[i].d);
      nbmax = std::max(nbmax, items[i].d);
    }
    // after the blocks' workspaces: the LDS path's barrier counters and
    // abort flag (unsigned, zeroed), inside ws so that hws carries the flag
    const int64_t nbar_d = (nb + 2) / 2;
    def->ws.resize(ctx->pool, wtot + nbar_d);
    unsigned* dbar = reinterpret_cast<unsigned*>(def->ws.p + wtot);
    NSS_HIP(hipMemsetAsync(dbar, 0, nbar_d * sizeof(double), st));
    for (int k = 0; k < nb; ++k) {
      const Item& it = items[bigs[k]];
      TredBlk t{};
      t.a = rho.p + it.roff;
      t.ws = def->ws.p + def->hwo[k];
      t.n = it.d;
      t.jt = it.d;
      def->blk.push_back(t);
      hbeo.push_back(it.eoff);
    }
    dtb.resize(ctx->pool, nb);
    dbeo.resize(ctx->pool, nb);
    NSS_HIP(hipMemcpyAsync(dtb.p, def->blk.data(), nb * sizeof(TredBlk), hipMemcpyHostToDevice, st));
    NSS_HIP(hipMemcpyAsync(dbeo.p, hbeo.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, st));
    // tridiagonalisation: one cooperative launch for the blocks whose
    // trailing matrices fit the chip's LDS, one launch per column for the rest
    std::vector<int> bdims(nb);
    for (int k = 0; k < nb; ++k) bdims[k] = def->blk[k].n;
    TredLdsPlan plan;
    bool lds_ok = tred_lds_enabled() && tred_lds_plan(ctx->device, bdims, &plan);
    if (lds_ok) {
      dtlm.resize(ctx->pool, plan.grid);
      dtlg.resize(ctx->pool, nb);
      NSS_HIP(hipMemcpyAsync(dtlm.p, plan.wgmap.data(), plan.grid * sizeof(int2), hipMemcpyHostToDevice, st));
      NSS_HIP(hipMemcpyAsync(dtlg.p, plan.gcount.data(), nb * sizeof(int), hipMemcpyHostToDevice, st));
      lds_ok = launch_tred_lds(dtb.p, dtlm.p, dtlg.p, dbar, reinterpret_cast<int*>(dbar + nb), plan.grid, plan.lds,
                               st) != 0;
      if (ctx->prof) ctx->timers["rdm.count_tred_lds"] += lds_ok;
    }
    // the rest: one launch per column until the trailing block fits one
    // workgroup's LDS, then that tail in one launch per block set
    const int tail = std::min(tred_tail_size(), tmax);
    std::vector<TredBlk> rest;
    int rmax = 0, jtmax = 0, tmx = 0;
    for (int k = 0; k < nb; ++k)
      if (!lds_ok || !plan.covered[k]) {
        TredBlk t = def->blk[k];
        if (tail >= 3 && t.n > tail) {
          t.jt = t.n - tail;
          tmx = tail;
        }
        rest.push_back(t);
        rmax = std::max(rmax, t.n);
        jtmax = std::max(jtmax, t.jt);
      }
    if (!rest.empty()) {
      dtbr.resize(ctx->pool, rest.size());
      NSS_HIP(hipMemcpyAsync(dtbr.p, rest.data(), rest.size() * sizeof(TredBlk), hipMemcpyHostToDevice, st));
      for (int j = 0; j < jtmax; ++j) launch_tred_col(dtbr.p, (int)rest.size(), j, rmax, st);
      if (tmx) launch_tred_tail(dtbr.p, (int)rest.size(), tmx, st);
    }
    launch_tri_eigvals(dtb.p, dbeo.p, dD.p, nb, nbmax, st);
    def->hws.resize(wtot + nbar_d);
    NSS_HIP(hipMemcpyAsync(def->hws.data(), def->ws.p, (wtot + nbar_d) * sizeof(double), hipMemcpyDeviceToHost,
                           st));
    big_abort = [def, wtot, nb]() {
      return reinterpret_cast<const unsigned*>(def->hws.data() + wtot)[nb] != 0;
    };
  }
  pt.next("rdm.rocsolver");
  std::map<int, std::vector<int>> classes;
  int dlarge = 0;  // largest block above the single-kernel range
  for (const Item& it : items)
    if (it.kind == 2 && it.d > 256) dlarge = std::max(dlarge, it.d);
  {  // one class for all large blocks only while its padded slab stays modest
    int64_t nlarge = 0;
    for (const Item& it : items) nlarge += it.kind == 2 && it.d > 256;
    const int64_t dl64 = (dlarge + 63) / 64 * 64;
    if (nlarge * dl64 * dl64 * (int64_t)sizeof(double) > 2 * RHO_SLAB_BYTES) dlarge = 0;
  }
  for (int i = 0; i < ni; ++i)
    if (items[i].kind == 2) classes[rho_class(items[i].d, dlarge)].push_back(i);
  // calls: each class cut into groups whose slab fits RHO_SLAB_BYTES
  std::vector<std::pair<int, std::vector<int>>> calls;
  for (auto& kv : classes) {
    const int64_t per = (int64_t)kv.first * kv.first * (int64_t)sizeof(double);
    const size_t cap = (size_t)std::max<int64_t>(1, RHO_SLAB_BYTES / per);
    for (size_t i0 = 0; i0 < kv.second.size(); i0 += cap)
      calls.push_back({kv.first, std::vector<int>(kv.second.begin() + i0,
                                                  kv.second.begin() + std::min(kv.second.size(), i0 + cap))});
  }
  struct ClassBatch {
    int ncls = 0, nb = 0;
    hipStream_t st = nullptr;
    std::vector<int64_t> co, ceo;
    std::vector<int32_t> cd, ci;
    DevArr<int64_t> dco, dceo;
    DevArr<int32_t> dcd, dci;
    DevArr<double> slab, sD, sE;
  };
  std::vector<ClassBatch> batches(calls.size());
  {
    size_t c = 0;
    for (auto& kv : calls) {
      ClassBatch& B = batches[c];
      const int k = (int)(c % nss_ctx::AUX_TRI);
      rocblas_handle hnd = aux_handle(ctx, k, &B.st);
      NSS_HIP(hipStreamWaitEvent(B.st, ctx->rho_ev, 0));
      used_aux[k] = true;
      ++c;
      B.ncls = kv.first;
      B.nb = (int)kv.second.size();
      if (ctx->prof) {
        ctx->timers["rdm.class_" + std::to_string(B.ncls) + ".calls"] += 1;
        ctx->timers["rdm.class_" + std::to_string(B.ncls) + ".blocks"] += B.nb;
      }
      for (int i : kv.second) {
        B.co.push_back(items[i].roff);
        B.ceo.push_back(items[i].eoff);
        B.cd.push_back(items[i].d);
      }
      const int nb = B.nb, ncls = B.ncls;
      B.ci.assign(nb, 0);
      B.dco.resize(ctx->pool, nb);
      B.dceo.resize(ctx->pool, nb);
      B.dcd.resize(ctx->pool, nb);
      B.dci.resize(ctx->pool, nb);
      B.slab.resize(ctx->pool, (size_t)nb * ncls * ncls);
      B.sD.resize(ctx->pool, (size_t)nb * ncls);
      B.sE.resize(ctx->pool, (size_t)nb * ncls);
      NSS_HIP(hipMemcpyAsync(B.dco.p, B.co.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, B.st));
      NSS_HIP(hipMemcpyAsync(B.dceo.p, B.ceo.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, B.st));
      NSS_HIP(hipMemcpyAsync(B.dcd.p, B.cd.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, B.st));
      launch_pad_blocks(rho.p, B.dco.p, B.dcd.p, nb, ncls, B.slab.p, B.st);
      const double te0 = ctx->prof ? now_s() : 0.0;
      const rocblas_status st = rocsolver_dsyevd_strided_batched(
          hnd, rocblas_evect_original, rocblas_fill_upper, ncls, B.slab.p, ncls, (rocblas_stride)ncls * ncls, B.sD.p,
          ncls, B.sE.p, ncls, B.dci.p, nb);
      if (ctx->prof) ctx->timers["rdm.rocsolver_enqueue"] += now_s() - te0;
      if (st != rocblas_status_success)
        throw Error(NSS_E_HIP, "rocsolver_dsyevd_strided_batched failed (rocblas status " + std::to_string((int)st) +
                                   ", " + std::to_string(nb) + " blocks of " + std::to_string(ncls) + ")");
      launch_unpad_eig(B.slab.p, B.sD.p, ncls, B.dco.p, B.dcd.p, B.dceo.p, nb, rho.p, dD.p, B.st);
    }
  }
  // every class is enqueued before the host waits on any of them
  const double tw0 = ctx->prof ? now_s() : 0.0;
  for (int k = 0; k < nss_ctx::NAUX; ++k)
    if (used_aux[k]) NSS_HIP(hipStreamSynchronize(ctx->aux[k]));
  if (ctx->prof) ctx->timers["rdm.rocsolver_wait"] += now_s() - tw0;
  NSS_HIP(hipEventSynchronize(ctx->rho_ev));
  check_symmetric();
  if (big_abort && big_abort())
    throw Error(NSS_E_HIP, "rho tridiagonalisation: a workgroup barrier of the LDS-resident path timed out");
  for (ClassBatch& B : batches)
    NSS_HIP(hipMemcpy(B.ci.data(), B.dci.p, B.nb * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (const ClassBatch& B : batches)
    for (int c = 0; c < B.nb; ++c)
      if (B.ci[c] != 0) throw Error(NSS_E_NOCONV, "rocsolver_dsyevd did not converge on a rho block");
  pt.next("rdm.download");
  std::vector<double> hD(std::max<int64_t>(1, etot));
  if (etot) NSS_HIP(hipMemcpyAsync(hD.data(), dD.p, etot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  // large blocks: their eigenvalues to the deferred record; eigenvectors
  // computed now (all of them) unless the job lets them wait for the
  // kept-state selection, then the block of rho holds them as the Jacobi
  // kernel leaves its output
  std::vector<RhoVec> bigvec(bigs.size());
  if (!bigs.empty()) {
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    def->ctx = ctx;
    std::vector<RhoVec*> now;
    std::vector<int> cnt;
    for (size_t k = 0; k < bigs.size(); ++k) {
      const Item& it = items[bigs[k]];
      def->lam.emplace_back(hD.begin() + it.eoff, hD.begin() + it.eoff + it.d);
      RhoVec& v = bigvec[k];
      v.d = it.d;
      v.def = def;
      v.def_blk = (int)k;
      v.order.resize(it.d);
      for (int j = 0; j < it.d; ++j) v.order[j] = j;
      const std::vector<double>& w = def->lam[k];
      std::stable_sort(v.order.begin(), v.order.end(), [&](int x, int y) { return w[x] < w[y]; });
      const RdmJob& J = jobs[it.job];
      if (!(J.dvecs && J.defer)) {
        now.push_back(&v);
        cnt.push_back(it.d);
      }
    }
    if (!now.empty()) {
      PhaseTimer pm(ctx, "rdm.big_vectors");
      rho_materialize(ctx, now, cnt);
      for (size_t k = 0; k < bigs.size(); ++k) {
        const Item& it = items[bigs[k]];
        if (!bigvec[k].base) continue;
        // every position kept: slot = position, so the slab is the
        // row-major eigenvector matrix
        NSS_HIP(hipMemcpyAsync(rho.p + it.roff, bigvec[k].base, (size_t)it.d * it.d * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
      }
    }
  }
  // eigenvectors to the host only when asked for (host_vectors); otherwise
  // they stay in rho and are handed out as device views
  bool host_vecs = false;
  for (const RdmJob& J : jobs) host_vecs = host_vecs || !J.dvecs;
  std::vector<double> h;
  if (host_vecs) {
    h.resize(std::max<int64_t>(1, tot));
    if (tot) NSS_HIP(hipMemcpyAsync(h.data(), rho.p, tot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  NSS_HIP(hipStreamSynchronize(ctx->stream));
  pt.next("rdm.host");
  for (RdmJob& J : jobs)
    if (J.dvecs) {
      const int nq = (int)J.sb->Dl.size();
      J.dvecs[0].assign(nq, {});
      J.dvecs[1].assign(nq, {});
    }
  for (const Item& it : items) {
    const int d = it.d;
    RdmJob& J = jobs[it.job];
    std::vector<RhoVec>* dvecs = J.dvecs;
    std::vector<double>& ev = J.evals[it.side][it.q];
    const double* a = dvecs ? nullptr : h.data() + it.roff;
    ev.resize(d);
    RhoVec* rv = dvecs ? &dvecs[it.side][it.q] : nullptr;
    std::vector<double>* vec = dvecs ? nullptr : &J.evecs[it.side][it.q];
    if (vec) vec->resize((size_t)d * d);
    if (rv) {
      rv->base = rho.p + it.roff;
      rv->d = d;
      rv->order.resize(d);
      for (int j = 0; j < d; ++j) rv->order[j] = j;
    }
    const bool deferred = it.kind == 4 && rv && J.defer;
    if (deferred) {
      // eigenvalues by position (ascending within each unreduced block of
      // T); the view carries the ascending order and stays deferred
      const size_t k = std::find(bigs.begin(), bigs.end(), (int)(&it - items.data())) - bigs.begin();
      const std::vector<double>& w = def->lam[k];
      for (int j = 0; j < d; ++j) ev[j] = w[bigvec[k].order[j]];
      *rv = bigvec[k];
    } else if (it.kind == 2) {
      // rocSOLVER: eigenvector j contiguous (component i at j*d + i), ascending
      std::copy(hD.begin() + it.eoff, hD.begin() + it.eoff + d, ev.begin());
      if (vec)
        for (int i = 0; i < d; ++i)
          for (int j = 0; j < d; ++j) (*vec)[(size_t)i * d + j] = a[(size_t)j * d + i];
      if (rv) {
        rv->si = 1;
        rv->sj = d;
      }
    } else {
      // Jacobi: row-major, eigenvectors as columns, unsorted -> ascending order
      std::vector<int> idx(d);
      for (int i = 0; i < d; ++i) idx[i] = i;
      const double* w = hD.data() + it.eoff;
      std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return w[x] < w[y]; });
      for (int j = 0; j < d; ++j) {
        ev[j] = w[idx[j]];
        if (vec)
          for (int i = 0; i < d; ++i) (*vec)[(size_t)i * d + j] = a[(size_t)i * d + idx[j]];
      }
      if (rv) {
        rv->si = d;
        rv->sj = 1;
        rv->order = idx;
      }
    }
    for (double& e : ev)
      if (e < 0.0) e = DBL_MIN;  // where(c%E<0d0)c%E=tiny(1d0), MATRIX_BLOCKS.f90:630
  }
}

void sb_rdm_eigh2(nss_ctx* ctx, SbDevice* sb, const double* psi_dev, int sides,
                  std::vector<std::vector<double>> evals[2], std::vector<std::vector<double>> evecs[2],
                  std::vector<RhoVec>* dvecs, bool defer) {
  std::vector<RdmJob> jobs(1);
  jobs[0].sb = sb;
  jobs[0].psi = psi_dev;
  jobs[0].sides = sides;
  jobs[0].defer = defer;
  jobs[0].evals = evals;
  jobs[0].evecs = evecs;
  jobs[0].dvecs = dvecs;
  sb_rdm_eigh_multi(ctx, jobs, sb->rho, sb->R);
}

void sb_rdm_eigh(nss_ctx* ctx, SbDevice* sb, const double* psi_dev, int side,
                 std::vector<std::vector<double>>& evals, std::vector<std::vector<double>>& evecs) {
  std::vector<std::vector<double>> ev[2], vec[2];
  sb_rdm_eigh2(ctx, sb, psi_dev, side == 0 ? 1 : 2, ev, vec);
  evals = std::move(ev[side]);
  evecs = std::move(vec[side]);
}

}  // namespace nss

using namespace nss;

// ------------------------------------------------------------------ scratch
// Device scratch of the allocation-free H*v entry point: the term products
// W_t = Psi_src B_t^T of one call (the role of the W buffer that
// nss_sb_upload_blocks allocates for nss_sb_hv_device).  Every device the
// library runs on holds one copy of NSS_NOALLOC_W_MAX doubles (128 MiB) from
// module load until the process ends; all calls on a device share it.  In
// host code `scratch` names a host-side stand-in, never the device memory:
// its per-device address comes from hipGetSymbolAddress (noalloc_scratch).
static __device__ __align__(256) double scratch[NSS_NOALLOC_W_MAX];

namespace {
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_addr;  // device -> address of its copy of `scratch`
}  // namespace

// Address of `scratch` on `device`, which must be the current device;
// resolved once per device.
static double* noalloc_scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) return it->second;
  void* p = nullptr;
  NSS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  NSS_REQUIRE(p != nullptr, NSS_E_HIP, "hipGetSymbolAddress returned no address for the H*v scratch");
  g_scratch_addr[device] = static_cast<double*>(p);
  return static_cast<double*>(p);
}

// ------------------------------------------------------------------ C ABI
#define NSS_API_BEGIN try {
#define NSS_API_END                                  \
  return NSS_OK;                                     \
  }                                                  \
  catch (const nss::Error& e) {                      \
    nss::set_last_error(e.what());                   \
    return e.code;                                   \
  }                                                  \
  catch (const std::exception& e) {                  \
    nss::set_last_error(e.what());                   \
    return NSS_E_ARG;                                \
  }

extern "C" const char* nss_last_error(void) { return nss::g_last_error.c_str(); }

extern "C" int nss_device_ok(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return 0;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 0;
  return std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

extern "C" int nss_ctx_create(int device, nss_ctx** out) {
  NSS_API_BEGIN
  NSS_REQUIRE(out != nullptr, NSS_E_ARG, "nss_ctx_create: out is NULL");
  int n = 0;
  NSS_HIP(hipGetDeviceCount(&n));
  NSS_REQUIRE(device >= 0 && device < n, NSS_E_ARG, "nss_ctx_create: no such device");
  hipDeviceProp_t prop;
  NSS_HIP(hipGetDeviceProperties(&prop, device));
  NSS_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, NSS_E_HIP,
              std::string("nss_ctx_create: device is ") + prop.gcnArchName +
                  ", this library is built for gfx950 only");
  NSS_HIP(hipSetDevice(device));
  nss_ctx* c = new nss_ctx();
  c->device = device;
  NSS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->pool = new DevicePool(device);
  c->sb = sb_create(c);
  try {
    c->noalloc_scratch = noalloc_scratch_address(device);
  } catch (...) {
    (void)hipStreamDestroy(c->stream);
    sb_destroy(c->sb);
    delete c->pool;
    delete c;
    throw;
  }
  const char* pe = std::getenv("NSS_PROFILE");
  c->prof = pe && pe[0] == '1';
  const char* te = std::getenv("NSS_HV_TIMING");  // period of the timed H*v calls (0: none)
  c->hv_every = te ? std::max(0, std::atoi(te)) : 4;
  *out = c;
  NSS_API_END
}

extern "C" int nss_ctx_destroy(nss_ctx* ctx) {
  NSS_API_BEGIN
  if (!ctx) return NSS_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  nss::engines_release(ctx);
  sb_destroy(ctx->sb);
  nss::nss_release_rocblas(ctx);
  delete ctx->pool;
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  NSS_API_END
}

extern "C" int nss_sb_set_layout(nss_ctx* ctx, int nq, const int32_t* Dl, const int32_t* Dr,
                                 const int64_t* offset) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && Dl && Dr && nq > 0, NSS_E_ARG, "nss_sb_set_layout: bad arguments");
  std::vector<int> dl(Dl, Dl + nq), dr(Dr, Dr + nq);
  std::vector<int64_t> off;
  if (offset) off.assign(offset, offset + nq);
  NSS_HIP(hipSetDevice(ctx->device));
  sb_set_layout(ctx->sb, dl, dr, offset ? &off : nullptr);
  NSS_API_END
}

extern "C" int nss_sb_upload_blocks(nss_ctx* ctx, const double* const* HL, const double* const* HR,
                                    int nterms, const int32_t* dst, const int32_t* src,
                                    const double* const* A, const double* const* B, int layout,
                                    int precision) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && HL && HR, NSS_E_ARG, "nss_sb_upload_blocks: NULL arguments");
  SbDevice* sb = ctx->sb;
  const int nq = (int)sb->Dl.size();
  NSS_REQUIRE(nq > 0, NSS_E_STATE, "nss_sb_upload_blocks: call nss_sb_set_layout first");
  NSS_REQUIRE(nterms >= 0 && (nterms == 0 || (dst && src && A && B)), NSS_E_ARG,
              "nss_sb_upload_blocks: bad term arrays");
  NSS_REQUIRE(layout == NSS_ROW_MAJOR || layout == NSS_COL_MAJOR, NSS_E_ARG, "bad layout");
  NSS_REQUIRE(precision == NSS_FP64 || precision == NSS_FP32_MIXED, NSS_E_ARG, "bad precision");
  NSS_HIP(hipSetDevice(ctx->device));
  const bool cm = layout == NSS_COL_MAJOR;
  // pack every block into one staging buffer, one upload
  std::vector<int64_t> oHL(nq), oHR(nq), oA(nterms), oB(nterms);
  int64_t tot = 0;
  for (int q = 0; q < nq; ++q) {
    oHL[q] = tot; tot += (int64_t)sb->Dl[q] * sb->Dl[q];
    oHR[q] = tot; tot += (int64_t)sb->Dr[q] * sb->Dr[q];
  }
  for (int t = 0; t < nterms; ++t) {
    NSS_REQUIRE(dst[t] >= 0 && dst[t] < nq && src[t] >= 0 && src[t] < nq, NSS_E_ARG,
                "nss_sb_upload_blocks: term sector out of range");
    oA[t] = tot; tot += (int64_t)sb->Dl[dst[t]] * sb->Dl[src[t]];
    oB[t] = tot; tot += (int64_t)sb->Dr[dst[t]] * sb->Dr[src[t]];
  }
  std::vector<double> stage((size_t)std::max<int64_t>(tot, 1));
  auto put = [&](int64_t o, const double* p, int64_t cnt) {
    if (cnt) {
      NSS_REQUIRE(p != nullptr, NSS_E_ARG, "nss_sb_upload_blocks: NULL block with nonzero size");
      std::memcpy(stage.data() + o, p, cnt * sizeof(double));
    }
  };
  for (int q = 0; q < nq; ++q) {
    put(oHL[q], HL[q], (int64_t)sb->Dl[q] * sb->Dl[q]);
    put(oHR[q], HR[q], (int64_t)sb->Dr[q] * sb->Dr[q]);
  }
  for (int t = 0; t < nterms; ++t) {
    put(oA[t], A[t], (int64_t)sb->Dl[dst[t]] * sb->Dl[src[t]]);
    put(oB[t], B[t], (int64_t)sb->Dr[dst[t]] * sb->Dr[src[t]]);
  }
  sb->owned.resize(ctx->pool, stage.size());
  NSS_HIP(hipMemcpyAsync(sb->owned.p, stage.data(), stage.size() * sizeof(double),
                         hipMemcpyHostToDevice, ctx->stream));
  const bool f32 = precision == NSS_FP32_MIXED;
  if (f32) {
    sb->owned32.resize(ctx->pool, stage.size());
    launch_d2f(sb->owned.p, sb->owned32.p, (int64_t)stage.size(), ctx->stream);
  }
  NSS_HIP(hipStreamSynchronize(ctx->stream));
  auto ref = [&](int64_t o, int ld, int trans) {
    OpRef r;
    r.p = sb->owned.p + o;
    r.p32 = f32 ? (const void*)(sb->owned32.p + o) : nullptr;
    r.ld = ld;
    r.trans = trans;
    r.alpha = 1.0;
    return r;
  };
  std::vector<OpRef> hl(nq), hr(nq);
  for (int q = 0; q < nq; ++q) {
    hl[q] = ref(oHL[q], sb->Dl[q], cm ? 1 : 0);
    hr[q] = ref(oHR[q], sb->Dr[q], cm ? 1 : 0);
  }
  std::vector<SbTerm> terms(nterms);
  for (int t = 0; t < nterms; ++t) {
    const int dd = dst[t], ss = src[t];
    terms[t].dst = dd;
    terms[t].src = ss;
    // A (Dl_dst x Dl_src): row-major ld = Dl_src; column-major = row-major A^T
    terms[t].A = cm ? ref(oA[t], sb->Dl[dd], 1) : ref(oA[t], sb->Dl[ss], 0);
    // B^T (Dr_src x Dr_dst): from row-major B (ld Dr_src) transposed, or
    // column-major B read directly with ld Dr_dst
    terms[t].Bt = cm ? ref(oB[t], sb->Dr[dd], 0) : ref(oB[t], sb->Dr[ss], 1);
  }
  sb_set_terms(ctx, sb, hl, hr, terms, f32);
  NSS_API_END
}

static void hv_host(nss_ctx* ctx, const double* v, double* hv, int64_t n, bool f32) {
  SbDevice* sb = ctx->sb;
  NSS_REQUIRE(sb->n > 0 && !sb->Y64.htiles.empty(), NSS_E_STATE,
              "nss_sb_hv: upload blocks before applying H");
  NSS_REQUIRE(n == sb->n, NSS_E_ARG,
              "nss_sb_hv: n=" + std::to_string(n) + " but the layout has " + std::to_string(sb->n));
  NSS_HIP(hipSetDevice(ctx->device));
  DevArr<double> x, y;
  x.resize(ctx->pool, n);
  y.resize(ctx->pool, n);
  NSS_HIP(hipMemcpyAsync(x.p, v, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if (f32) sb_hv_fp32(ctx, sb, x.p, y.p);
  else sb_hv(ctx, sb, x.p, y.p);
  NSS_HIP(hipMemcpyAsync(hv, y.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  NSS_HIP(hipStreamSynchronize(ctx->stream));
  nss_hv_timer_flush(ctx);
}

extern "C" int nss_sb_hv(nss_ctx* ctx, const double* v, double* hv, int64_t n) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && v && hv, NSS_E_ARG, "nss_sb_hv: NULL arguments");
  hv_host(ctx, v, hv, n, false);
  NSS_API_END
}

extern "C" int nss_sb_hv_fp32(nss_ctx* ctx, const double* v, double* hv, int64_t n) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && v && hv, NSS_E_ARG, "nss_sb_hv_fp32: NULL arguments");
  hv_host(ctx, v, hv, n, true);
  NSS_API_END
}

extern "C" int nss_sb_hv_device(nss_ctx* ctx, const double* d_v, double* d_hv, int64_t n) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && d_v && d_hv, NSS_E_ARG, "nss_sb_hv_device: NULL arguments");
  NSS_REQUIRE(n == ctx->sb->n && n > 0, NSS_E_ARG, "nss_sb_hv_device: n does not match the layout");
  NSS_HIP(hipSetDevice(ctx->device));
  sb_hv(ctx, ctx->sb, d_v, d_hv);
  NSS_API_END
}

// H*v of the current superblock on device vectors with no device allocation
// and no timing events: the term products go to this device's copy of the
// file-scope scratch array.  Enqueued on the context stream like
// nss_sb_hv_device; calls that overlap on one device (other contexts or
// streams) race on the shared scratch.
extern "C" int nss_sb_hv_device_noalloc(nss_ctx* ctx, const double* d_v, double* d_hv, int64_t n) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && d_v && d_hv, NSS_E_ARG, "nss_sb_hv_device_noalloc: NULL arguments");
  NSS_REQUIRE(n == ctx->sb->n && n > 0, NSS_E_ARG, "nss_sb_hv_device_noalloc: n does not match the layout");
  NSS_REQUIRE(ctx->noalloc_scratch != nullptr, NSS_E_STATE, "nss_sb_hv_device_noalloc: scratch not resolved");
  const SbDevice* sb = ctx->sb;
  NSS_REQUIRE(!sb->W64.htiles.empty() || !sb->Y64.htiles.empty(), NSS_E_STATE,
              "nss_sb_hv_device_noalloc: call nss_sb_upload_blocks first");
  NSS_REQUIRE(sb->wsize <= NSS_NOALLOC_W_MAX, NSS_E_ARG,
              "nss_sb_hv_device_noalloc: the superblock's term products need " + std::to_string(sb->wsize) +
                  " doubles, above the scratch's " + std::to_string((long long)NSS_NOALLOC_W_MAX));
  NSS_HIP(hipSetDevice(ctx->device));
  const Bases b{d_v, ctx->noalloc_scratch, d_hv};
  launch_gemm_f64(sb->W64, b, ctx->stream, GEMM_HV_W);
  launch_gemm_f64(sb->Y64, b, ctx->stream, GEMM_HV_Y);
  NSS_API_END
}

extern "C" int nss_gemm_tune_table(int* rows, int maxrows, int* count) {
  NSS_API_BEGIN
  NSS_REQUIRE(count && maxrows >= 0, NSS_E_ARG, "nss_gemm_tune_table: NULL count or negative maxrows");
  *count = nss::plan_tune_entries(rows, maxrows);
  NSS_API_END
}

extern "C" int nss_ctx_hv_stats(nss_ctx* ctx, int reset, double* out) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && out, NSS_E_ARG, "nss_ctx_hv_stats: NULL arguments");
  nss_hv_timer_flush(ctx);
  for (int k = 0; k < 2; ++k) {
    out[6 * k + 0] = (double)ctx->st_calls[k];
    out[6 * k + 1] = ctx->st_tw[k];
    out[6 * k + 2] = ctx->st_ty[k];
    out[6 * k + 3] = ctx->st_fw[k];
    out[6 * k + 4] = ctx->st_fy[k];
    out[6 * k + 5] = (double)ctx->st_all[k];
    if (reset) {
      ctx->st_calls[k] = ctx->st_all[k] = 0;
      ctx->st_tw[k] = ctx->st_ty[k] = ctx->st_fw[k] = ctx->st_fy[k] = 0.0;
    }
  }
  NSS_API_END
}

extern "C" int nss_ctx_timers(nss_ctx* ctx, char* buf, int len) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx, NSS_E_ARG, "nss_ctx_timers: NULL ctx");
  std::string out = "{";
  bool first = true;
  for (auto& kv : ctx->timers) {
    out += (first ? "\"" : ", \"") + kv.first + "\": " + std::to_string(kv.second);
    first = false;
  }
  out += "}";
  if (buf && len > 0) {
    std::strncpy(buf, out.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  NSS_API_END
}

extern "C" int nss_sb_flops(nss_ctx* ctx, double* flops) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && flops, NSS_E_ARG, "nss_sb_flops: NULL arguments");
  *flops = ctx->sb->flops;
  NSS_API_END
}

extern "C" int nss_lanczos(nss_ctx* ctx, int neigen, int kmax, int maxiter, double tol, uint64_t seed,
                           int v0_mode, int precision, double* evals, double* evecs, int* numop) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && evals, NSS_E_ARG, "nss_lanczos: NULL arguments");
  SbDevice* sb = ctx->sb;
  NSS_REQUIRE(sb->n > 0 && !sb->Y64.htiles.empty(), NSS_E_STATE, "nss_lanczos: no superblock uploaded");
  NSS_REQUIRE(neigen >= 1 && neigen <= sb->n, NSS_E_ARG, "nss_lanczos: neigen out of range");
  NSS_REQUIRE(v0_mode >= 0 && v0_mode <= 2, NSS_E_ARG, "nss_lanczos: v0_mode must be 0, 1 or 2");
  NSS_REQUIRE(v0_mode != 2 || evecs, NSS_E_ARG, "nss_lanczos: v0_mode 2 reads the start vector from evecs");
  NSS_HIP(hipSetDevice(ctx->device));
  DevArr<double> x;
  x.resize(ctx->pool, (size_t)neigen * sb->n);
  if (v0_mode == 2)
    NSS_HIP(hipMemcpyAsync(x.p, evecs, sb->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  LanczosOut r = lanczos_device(ctx, sb, neigen, kmax, maxiter, tol, seed, v0_mode, precision, x.p);
  for (int i = 0; i < neigen; ++i) evals[i] = r.evals[i];
  if (evecs)
    NSS_HIP(hipMemcpyAsync(evecs, x.p, (size_t)neigen * sb->n * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
  NSS_HIP(hipStreamSynchronize(ctx->stream));
  if (numop) *numop = r.numop + r.numop_fp32;
  NSS_API_END
}

extern "C" int nss_rdm_eigh(nss_ctx* ctx, const double* psi, int side, double* evals, double* evecs) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && psi && evals && evecs, NSS_E_ARG, "nss_rdm_eigh: NULL arguments");
  NSS_REQUIRE(side == 0 || side == 1, NSS_E_ARG, "nss_rdm_eigh: side must be 0 (left) or 1 (right)");
  SbDevice* sb = ctx->sb;
  NSS_REQUIRE(sb->n > 0, NSS_E_STATE, "nss_rdm_eigh: set a layout first");
  NSS_HIP(hipSetDevice(ctx->device));
  DevArr<double> x;
  x.resize(ctx->pool, sb->n);
  NSS_HIP(hipMemcpyAsync(x.p, psi, sb->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  std::vector<std::vector<double>> ev, vec;
  sb_rdm_eigh(ctx, sb, x.p, side, ev, vec);
  size_t oe = 0, ov = 0;
  for (size_t q = 0; q < ev.size(); ++q) {
    const int d = side == 0 ? sb->Dl[q] : sb->Dr[q];
    if (ev[q].empty()) {  // empty sector partner: nothing (the reference skips it)
      for (int i = 0; i < d; ++i) evals[oe++] = 0.0;
      ov += (size_t)d * d;
      continue;
    }
    std::memcpy(evals + oe, ev[q].data(), d * sizeof(double));
    std::memcpy(evecs + ov, vec[q].data(), (size_t)d * d * sizeof(double));
    oe += d;
    ov += (size_t)d * d;
  }
  NSS_API_END
}

extern "C" int nss_kept_count(const double* ev, int ne, int mstates, double estates, int mdim,
                              double deg_thr, int* m_s, double* trunc_err) {
  NSS_API_BEGIN
  NSS_REQUIRE(ev && m_s && ne > 0, NSS_E_ARG, "nss_kept_count: bad arguments");
  int ms;
  if (mstates != 0) {
    ms = std::min(mstates, std::min(mdim, ne));
  } else if (estates != 0.0) {  // minloc(abs(1-cumulate(evals)-Estates)), DMRG_RDM.f90:208-209
    double cum = 0.0, best = INFINITY;
    ms = 1;
    for (int i = 0; i < ne; ++i) {
      cum += ev[i];
      const double d = std::fabs(1.0 - cum - estates);
      if (d < best) { best = d; ms = i + 1; }
    }
  } else {
    throw Error(NSS_E_ARG, "Mdmrg=Edmrg=0 can not fix a threshold for the RDM");
  }
  NSS_REQUIRE(ms >= 1, NSS_E_ARG, "nss_kept_count: m_s < 1");
  const double e_ = ev[ms - 1];
  const int j_ = ms;
  for (int i = j_; i < ne; ++i) {  // DMRG_RDM.f90:213-218
    const double err = std::fabs(e_ - ev[i]) / e_;
    if (err <= deg_thr) ++ms;
  }
  double s = 0.0;
  for (int i = 0; i < ms; ++i) s += ev[i];
  *m_s = ms;
  if (trunc_err) *trunc_err = 1.0 - s;
  NSS_API_END
}

extern "C" int nss_rotate_blocks(nss_ctx* ctx, int nblk, const int32_t* rows, const int32_t* cols,
                                 const int32_t* mr, const int32_t* mc, const double* const* O,
                                 const double* const* Urow, const double* const* Ucol,
                                 double* const* out) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && nblk >= 0, NSS_E_ARG, "nss_rotate_blocks: bad arguments");
  if (nblk == 0) return NSS_OK;
  NSS_REQUIRE(rows && cols && mr && mc && O && Urow && Ucol && out, NSS_E_ARG,
              "nss_rotate_blocks: NULL arrays");
  NSS_HIP(hipSetDevice(ctx->device));
  std::vector<int64_t> oO(nblk), oR(nblk), oC(nblk), oT(nblk), oOut(nblk);
  int64_t tin = 0, tmid = 0, tout = 0;
  for (int b = 0; b < nblk; ++b) {
    NSS_REQUIRE(rows[b] >= 0 && cols[b] >= 0 && mr[b] >= 0 && mc[b] >= 0, NSS_E_ARG,
                "nss_rotate_blocks: negative dims");
    oO[b] = tin; tin += (int64_t)rows[b] * cols[b];
    oR[b] = tin; tin += (int64_t)rows[b] * mr[b];
    oC[b] = tin; tin += (int64_t)cols[b] * mc[b];
    oT[b] = tmid; tmid += (int64_t)rows[b] * mc[b];
    oOut[b] = tout; tout += (int64_t)mr[b] * mc[b];
  }
  std::vector<double> stage(std::max<int64_t>(1, tin));
  for (int b = 0; b < nblk; ++b) {
    std::memcpy(stage.data() + oO[b], O[b], sizeof(double) * rows[b] * (size_t)cols[b]);
    std::memcpy(stage.data() + oR[b], Urow[b], sizeof(double) * rows[b] * (size_t)mr[b]);
    std::memcpy(stage.data() + oC[b], Ucol[b], sizeof(double) * cols[b] * (size_t)mc[b]);
  }
  DevArr<double> din, dmid, dout;
  din.resize(ctx->pool, stage.size());
  dmid.resize(ctx->pool, std::max<int64_t>(1, tmid));
  dout.resize(ctx->pool, std::max<int64_t>(1, tout));
  NSS_HIP(hipMemcpyAsync(din.p, stage.data(), stage.size() * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  GemmPlan p1, p2;
  for (int b = 0; b < nblk; ++b) {  // T = O . Ucol ; out = Urow^T . T
    p1.add_output(BASE_ABS, dmid.p + oT[b], 0, mc[b], rows[b], mc[b],
                  {seg_abs(din.p + oO[b], cols[b], 0, din.p + oC[b], mc[b], 0, cols[b], 1.0)});
    p2.add_output(BASE_ABS, dout.p + oOut[b], 0, mc[b], mr[b], mc[b],
                  {seg_abs(din.p + oR[b], mr[b], 1, dmid.p + oT[b], mc[b], 0, rows[b], 1.0)});
  }
  p1.upload(ctx->pool, ctx->stream);
  p2.upload(ctx->pool, ctx->stream);
  launch_gemm_f64(p1, Bases{nullptr, nullptr, nullptr}, ctx->stream);
  launch_gemm_f64(p2, Bases{nullptr, nullptr, nullptr}, ctx->stream);
  std::vector<double> h(std::max<int64_t>(1, tout));
  NSS_HIP(hipMemcpyAsync(h.data(), dout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  NSS_HIP(hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < nblk; ++b)
    std::memcpy(out[b], h.data() + oOut[b], sizeof(double) * mr[b] * (size_t)mc[b]);
  NSS_API_END
}

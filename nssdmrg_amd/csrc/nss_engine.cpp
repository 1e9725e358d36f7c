// Device-resident DMRG engine (init_dmrg / run_DMRG / step_dmrg,
// src/DMRG_MAIN.f90:22-433).  Every block operator lives in HBM as a list of
// dense QN blocks; enlargement, the superblock product, rho and the
// truncation rotation are GPU launches.  The host keeps only QN
// bookkeeping, the Lanczos projected matrix and the rho spectra.
//
// Basis order.  A block's basis is sector-contiguous.  The enlarged basis of
// block (x) site groups the pairs (block sector a, site state s) by total QN
// in first-appearance order over (a outer, s inner); each pair contributes
// the D_a states of sector a in their order.  This is a permutation of the
// reference's Kronecker order within each QN sector
// (src/LIST/QN_TUPLE_BASIS.f90:371-398), so spectra, energies, entropies and
// kept-state counts per QN are unchanged.
//
// QNs are stored as integers scaled by 2 (spin Sz = +-1/2 -> +-1).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <memory>
#include <numeric>

#include "nss_internal.h"

namespace nss {

struct QN {
  int a = 0, b = 0;
  bool operator==(const QN& o) const { return a == o.a && b == o.b; }
  bool operator<(const QN& o) const { return a < o.a || (a == o.a && b < o.b); }
  QN operator+(const QN& o) const { return {a + o.a, b + o.b}; }
  QN operator-(const QN& o) const { return {a - o.a, b - o.b}; }
};

struct Sector {
  QN q;
  int dim = 0;
};

struct OpBlock {
  int rs, cs;
  size_t off;
};

struct BlockOp {
  QN dq;  // row QN - col QN
  std::vector<OpBlock> blks;
  std::vector<int> by_row;  // row sector -> block index or -1
  DevArr<double> data;
  DevArr<float> data32;
  const OpBlock* row(int rs) const {
    if (rs < 0 || rs >= (int)by_row.size() || by_row[rs] < 0) return nullptr;
    return &blks[by_row[rs]];
  }
};
using OpPtr = std::shared_ptr<BlockOp>;

// Piece of an enlarged basis: block sector a times site state s, starting at
// row `off` of its enlarged sector (a outer, s inner: the reference's
// Kronecker order restricted to the sector, src/DMRG_CONNECT.f90:75-88).
struct Piece {
  int a, s, off;
};

// How a block's basis comes from the enlargement of its parent: sector k
// holds the kept states of enlarged sector src[k] as the columns of
// U_k (dim[k] x m[k], row-major at uoff[k], descending rho eigenvalue).  A
// block that was not renormalized (the enlarged block itself) carries
// identity blocks.
struct Basis {
  std::vector<int> src, dim, m;
  std::vector<int64_t> uoff;
  DevArr<double> U;
};

struct Block {
  int length = 0;
  std::vector<Sector> sec;
  std::map<std::string, OpPtr> ops;
  std::vector<std::vector<Piece>> pieces;  // set on enlarged blocks: pieces of each sector
  std::shared_ptr<Basis> from;             // set on renormalized / stored blocks
  // the block whose enlargement this block's basis comes from (the history
  // Measure_DMRG walks); null for the one-site block
  std::shared_ptr<const Block> parent;
  int dim() const {
    int d = 0;
    for (auto& s : sec) d += s.dim;
    return d;
  }
  int find(QN q) const {
    for (size_t i = 0; i < sec.size(); ++i)
      if (sec[i].q == q) return (int)i;
    return -1;
  }
};

struct SiteOp {
  QN dq;
  std::vector<double> m;  // dim x dim row-major
};

struct Site {
  bool fermion = false;
  int dim = 0;
  std::vector<QN> qn;
  std::map<std::string, SiteOp> ops;
};

// Both rho sides solved and both blocks renormalized in every finite step
// (NSS_FAR_RHO=1), also where the sweep replaces the far block next.
static int full_far_rho() {
  static const int v = [] {
    const char* e = std::getenv("NSS_FAR_RHO");
    return e && e[0] == '1' ? 1 : 0;
  }();
  return v;
}

static int fsign(const QN& q) { return (((q.a + q.b) / 2) & 1) ? -1 : 1; }  // (-1)^(Nup+Ndw)

// -------------------------------------------------------------- sites
// spin_site, src/BLOCKS/SITES.f90:281-344; electron_site :358-434 with the
// local Fock space of src/FOCK/HLOCAL.f90 (Norb=1).
static Site make_site(const nss_model& m) {
  Site s;
  if (m.model == NSS_MODEL_SPIN12 || m.model == NSS_MODEL_SPIN1) {
    const bool one = m.model == NSS_MODEL_SPIN1;
    s.dim = one ? 3 : 2;
    const int d = s.dim;
    std::vector<double> H(d * d, 0.0), Sz(d * d, 0.0), Sp(d * d, 0.0);
    if (!one) {
      Sz[0] = 0.5; Sz[3] = -0.5;
      Sp[0 * 2 + 1] = 1.0;
      // H = hx*pauli_x + hz*pauli_z
      H[0 * 2 + 1] = H[1 * 2 + 0] = m.hvec[0];
      H[0] = m.hvec[1]; H[3] = -m.hvec[1];
      s.qn = {{1, 0}, {-1, 0}};
    } else {
      Sz[0] = 1.0; Sz[4] = 0.0; Sz[8] = -1.0;
      Sp[0 * 3 + 1] = std::sqrt(2.0);
      Sp[1 * 3 + 2] = std::sqrt(2.0);
      const double r = 1.0 / std::sqrt(2.0);
      // H = hx*Sx + hz*Sz, Sx = [[0,1,0],[1,0,1],[0,1,0]]/sqrt(2)
      H[0 * 3 + 1] = H[1 * 3 + 0] = H[1 * 3 + 2] = H[2 * 3 + 1] = m.hvec[0] * r;
      H[0] += m.hvec[1]; H[8] -= m.hvec[1];
      s.qn = {{2, 0}, {0, 0}, {-2, 0}};
    }
    s.ops["H"] = {{0, 0}, H};
    s.ops["Sz"] = {{0, 0}, Sz};
    s.ops["Sp"] = {{2, 0}, Sp};
  } else if (m.model == NSS_MODEL_HUBBARD) {
    s.fermion = true;
    s.dim = 4;
    std::vector<double> H(16, 0.0), Cu(16, 0.0), Cd(16, 0.0);
    for (int st = 0; st < 4; ++st) {
      const int nup = st & 1, ndw = (st >> 1) & 1;
      s.qn.push_back({2 * nup, 2 * ndw});
      double e = m.eloc[0] * nup + m.eloc[1] * ndw + m.uloc * nup * ndw;
      if (m.hfmode) e -= 0.5 * m.uloc * (nup + ndw);
      H[st * 4 + st] = e;
      if (nup) Cu[(st & ~1) * 4 + st] = 1.0;
      if (ndw) Cd[(st & ~2) * 4 + st] = nup ? -1.0 : 1.0;
    }
    s.ops["H"] = {{0, 0}, H};
    s.ops["Cup"] = {{-2, 0}, Cu};
    s.ops["Cdn"] = {{0, -2}, Cd};
  } else {
    throw Error(NSS_E_ARG, "unknown model");
  }
  return s;
}

static std::vector<double> transpose(const std::vector<double>& a, int d) {
  std::vector<double> t(d * d);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) t[j * d + i] = a[i * d + j];
  return t;
}

// -------------------------------------------------------------- engine
struct Step {
  nss_step_info info{};
  // kept sectors: QN and the rho eigenvalues of their kept states
  std::vector<std::pair<QN, std::vector<double>>> kept_left, kept_right;
  // rho%lambda(): every QN block's eigenvalues (ascending), superblock sector order
  std::vector<std::pair<QN, std::vector<double>>> lam_left, lam_right;
};

struct Engine {
  nss_ctx* ctx;
  nss_model model;
  nss_params par;
  std::vector<int32_t> msweep;
  std::vector<double> esweep;
  Site site;
  Block left, right;
  std::vector<Step> steps;
  std::vector<double> hij;  // 2x2
  // last superblock (for export)
  Block lastL, lastR;
  std::vector<int> sbL, sbR;
  std::vector<int> lastDl, lastDr;  // layout of the last superblock (engine-owned copy)
  std::vector<SbTerm> lastTerms;
  DevArr<double> psi;
  bool want32 = false;
  // superblock plan: the context's (standalone engines, so that the
  // low-level entry points see the engine's last superblock) or an own one
  // (members of a batch, whose superblocks coexist)
  SbDevice* own_sb = nullptr;
  SbDevice* sbd() const { return own_sb ? own_sb : ctx->sb; }
  ~Engine() {
    if (own_sb) sb_destroy(own_sb);
  }

  // ------------------------------------------------------------ blocks
  OpPtr new_op(QN dq, const std::vector<Sector>& rsec, const std::vector<Sector>& csec) {
    auto op = std::make_shared<BlockOp>();
    op->dq = dq;
    op->by_row.assign(rsec.size(), -1);
    size_t off = 0;
    for (size_t r = 0; r < rsec.size(); ++r) {
      const QN qc = rsec[r].q - dq;
      int c = -1;
      for (size_t k = 0; k < csec.size(); ++k)
        if (csec[k].q == qc) { c = (int)k; break; }
      if (c < 0) continue;
      op->by_row[r] = (int)op->blks.size();
      op->blks.push_back({(int)r, c, off});
      off += (size_t)rsec[r].dim * csec[c].dim;
    }
    op->data.resize(ctx->pool, std::max<size_t>(1, off));
    return op;
  }

  void make_fp32(BlockOp& op) {
    if (!want32) return;
    op.data32.resize(ctx->pool, op.data.n);
    launch_d2f(op.data.p, op.data32.p, (int64_t)op.data.n, ctx->stream);
  }

  Block block_from_site() {
    Block b;
    b.length = 1;
    std::vector<int> state_sec(site.dim), state_pos(site.dim);
    for (int s = 0; s < site.dim; ++s) {
      int k = b.find(site.qn[s]);
      if (k < 0) {
        b.sec.push_back({site.qn[s], 0});
        k = (int)b.sec.size() - 1;
      }
      state_sec[s] = k;
      state_pos[s] = b.sec[k].dim++;
    }
    for (auto& kv : site.ops) {
      OpPtr op = new_op(kv.second.dq, b.sec, b.sec);
      std::vector<double> h(op->data.n, 0.0);
      for (int i = 0; i < site.dim; ++i)
        for (int j = 0; j < site.dim; ++j) {
          const double v = kv.second.m[i * site.dim + j];
          if (v == 0.0) continue;
          const OpBlock* ob = op->row(state_sec[i]);
          NSS_REQUIRE(ob && ob->cs == state_sec[j], NSS_E_STATE, "site operator breaks its QN shift");
          h[ob->off + (size_t)state_pos[i] * b.sec[ob->cs].dim + state_pos[j]] = v;
        }
      NSS_HIP(hipMemcpyAsync(op->data.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
      NSS_HIP(hipStreamSynchronize(ctx->stream));
      make_fp32(*op);
      b.ops[kv.first] = op;
    }
    return b;
  }

  // X (x) Y term of an enlarged operator
  struct KTerm {
    const BlockOp* x;  // nullptr -> identity
    int xtrans;
    int sign_row, sign_col;  // multiply by P_b sign of the row / col block sector
    std::vector<double> y;   // site matrix (dim x dim)
    QN dq;                   // shift of the whole term
    double coef;
  };

  // enlarge_block, src/DMRG_CONNECT.f90:18-291 (OBC link 'n').
  // Sectors and pieces of the enlarged basis of b (block sector outer, site
  // state inner, grouped by QN in first-appearance order).
  void enlarged_basis(const Block& b, std::vector<Sector>& esec, std::vector<std::vector<Piece>>& pieces) const {
    esec.clear();
    pieces.clear();
    for (size_t a = 0; a < b.sec.size(); ++a)
      for (int s = 0; s < site.dim; ++s) {
        const QN q = b.sec[a].q + site.qn[s];
        int k = -1;
        for (size_t i = 0; i < esec.size(); ++i)
          if (esec[i].q == q) { k = (int)i; break; }
        if (k < 0) {
          esec.push_back({q, 0});
          pieces.emplace_back();
          k = (int)esec.size() - 1;
        }
        pieces[k].push_back({(int)a, s, esec[k].dim});
        esec[k].dim += b.sec[a].dim;
      }
  }

  // One operator of the enlarged basis as a sum of Kronecker terms X (x) Y
  // (X a block operator or the identity, Y a site matrix): the new operator
  // is zeroed on the stream and its k_kron tasks are appended.
  OpPtr kron_op(const Block& b, const std::vector<Sector>& esec, const std::vector<std::vector<Piece>>& pieces,
                const std::vector<KTerm>& terms, QN dq, std::vector<KronTask>& tasks,
                std::vector<KronContrib>& contribs) {
    const int d = site.dim;
    OpPtr op = new_op(dq, esec, esec);
    for (const OpBlock& ob : op->blks) {
      const int ldd = esec[ob.cs].dim;
      for (const Piece& p : pieces[ob.rs])
        for (const Piece& pc : pieces[ob.cs]) {
          const int c0 = (int)contribs.size();
          const int Da = b.sec[p.a].dim, Dc = b.sec[pc.a].dim;
          for (const KTerm& t : terms) {
            const double yv = t.y[p.s * d + pc.s];
            if (yv == 0.0) continue;
            KronContrib c{};
            c.scale = t.coef * yv;
            if (t.sign_row) c.scale *= fsign(b.sec[p.a].q);
            if (t.sign_col) c.scale *= fsign(b.sec[pc.a].q);
            if (t.x == nullptr) {
              if (p.a != pc.a) continue;
              c.x = nullptr;
            } else if (!t.xtrans) {
              const OpBlock* xb = t.x->row(p.a);
              if (!xb || xb->cs != pc.a) continue;
              c.x = t.x->data.p + xb->off;
              c.ldx = Dc;
              c.trans = 0;
            } else {
              const OpBlock* xb = t.x->row(pc.a);
              if (!xb || xb->cs != p.a) continue;
              c.x = t.x->data.p + xb->off;
              c.ldx = Da;  // stored (pc.a x p.a) row-major: element (r,c) at [c*Da + r]
              c.trans = 1;
            }
            contribs.push_back(c);
          }
          const int nc = (int)contribs.size() - c0;
          if (nc == 0) continue;  // stays zero
          double* dst = op->data.p + ob.off + (size_t)p.off * ldd + pc.off;
          for (int r = 0; r < Da; r += KRON_TILE)
            for (int cc = 0; cc < Dc; cc += KRON_TILE)
              tasks.push_back({dst, ldd, Da, Dc, c0, nc, r, cc});
        }
    }
    NSS_HIP(hipMemsetAsync(op->data.p, 0, op->data.n * sizeof(double), ctx->stream));
    return op;
  }

  // Uploads the task lists and runs k_kron (synchronous).
  void run_kron(const std::vector<KronTask>& tasks, const std::vector<KronContrib>& contribs) {
    DevArr<KronTask> dt;
    DevArr<KronContrib> dc;
    dt.resize(ctx->pool, std::max<size_t>(1, tasks.size()));
    dc.resize(ctx->pool, std::max<size_t>(1, contribs.size()));
    if (!tasks.empty()) {
      NSS_HIP(hipMemcpyAsync(dt.p, tasks.data(), tasks.size() * sizeof(KronTask), hipMemcpyHostToDevice,
                             ctx->stream));
      NSS_HIP(hipMemcpyAsync(dc.p, contribs.data(), contribs.size() * sizeof(KronContrib),
                             hipMemcpyHostToDevice, ctx->stream));
      launch_kron(dt.p, (int)tasks.size(), dc.p, ctx->stream);
    }
    NSS_HIP(hipStreamSynchronize(ctx->stream));
  }

  // Link 'n' (OBC and PBC_next): the dot joins the block's inner end; link
  // 'p' (PBC_prev): its outer end.  Operators of the outer end carry the
  // suffix _p.  Stored bases are always block sector outer, site state inner;
  // the reference's Kronecker order (block first for left 'n' and right 'p',
  // dot first otherwise, :251-263) only decides where the signs go.
  Block enlarge(const Block& b, bool is_left, char link = 'n') {
    Block e;
    e.length = b.length + 1;
    std::vector<std::vector<Piece>> pieces;
    enlarged_basis(b, e.sec, pieces);
    const int d = site.dim;
    auto sop = [&](const char* k) -> const SiteOp& { return site.ops.at(k); };
    std::vector<double> I(d * d, 0.0);
    for (int i = 0; i < d; ++i) I[i * d + i] = 1.0;
    std::vector<double> Pd(d * d, 0.0);
    for (int i = 0; i < d; ++i) Pd[i * d + i] = fsign(site.qn[i]);
    auto mul = [&](const std::vector<double>& A, const std::vector<double>& B) {
      std::vector<double> C(d * d, 0.0);
      for (int i = 0; i < d; ++i)
        for (int k = 0; k < d; ++k)
          for (int j = 0; j < d; ++j) C[i * d + j] += A[i * d + k] * B[k * d + j];
      return C;
    };
    const QN z{0, 0};
    const bool dot_last = is_left == (link == 'n');  // the reference's block-first Kronecker order
    const std::string x = link == 'n' ? "" : "_p";   // operators of the end the dot joins
    std::map<std::string, std::vector<KTerm>> defs;
    std::vector<KTerm>& H = defs["H"];
    H.push_back({b.ops.at("H").get(), 0, 0, 0, I, z, 1.0});
    H.push_back({nullptr, 0, 0, 0, sop("H").m, z, 1.0});
    if (!site.fermion) {
      const BlockOp* Sz = b.ops.at("Sz" + x).get();
      const BlockOp* Sp = b.ops.at("Sp" + x).get();
      const double jz = hij[0], jxy = hij[3];
      // connect_spin_blocks :396-454 (symmetric under the exchange of its arguments)
      H.push_back({Sz, 0, 0, 0, sop("Sz").m, z, jz});
      H.push_back({Sp, 0, 0, 0, transpose(sop("Sp").m, d), z, jxy});
      H.push_back({Sp, 1, 0, 0, sop("Sp").m, z, jxy});
    } else {
      const char* names[2] = {"Cup", "Cdn"};
      for (int sp = 0; sp < 2; ++sp) {
        const double t = hij[sp * 2 + sp];
        const BlockOp* C = b.ops.at(names[sp] + x).get();
        const std::vector<double>& Cs = sop(names[sp]).m;
        if (t == 0.0) continue;
        if (dot_last) {
          // (C_b^T P_b) x C_s  +  (P_b C_b) x C_s^T      (connect(self,dot))
          H.push_back({C, 1, 0, 1, Cs, z, t});
          H.push_back({C, 0, 1, 0, transpose(Cs, d), z, t});
        } else {
          // C_b x (C_s^T P_s)  +  C_b^T x (P_s C_s)      (connect(dot,self))
          H.push_back({C, 0, 0, 0, mul(transpose(Cs, d), Pd), z, t});
          H.push_back({C, 1, 0, 0, mul(Pd, Cs), z, t});
        }
      }
    }
    // the other operators (:116-237; P is the sector sign, kept implicit)
    for (const auto& kv : b.ops) {
      const std::string& k = kv.first;
      if (k == "H") continue;
      const bool kp = k.size() > 2 && k.compare(k.size() - 2, 2, "_p") == 0;
      const SiteOp& so = sop(k.c_str());
      if (kp == (link == 'p')) {
        // the end the dot joins: the dot's operator, behind the block's sign
        // where the block comes first (fermions)
        defs[k].push_back({nullptr, 0, (dot_last && site.fermion) ? 1 : 0, 0, so.m, so.dq, 1.0});
      } else {
        // the other end keeps the block's operator, behind the dot's sign
        // where the dot comes first (fermions)
        defs[k].push_back({kv.second.get(), 0, 0, 0, (!dot_last && site.fermion) ? Pd : I, so.dq, 1.0});
      }
    }
    // tasks
    std::vector<KronTask> tasks;
    std::vector<KronContrib> contribs;
    for (auto& kv : defs) {
      const QN dq = kv.first == "H" ? z : site.ops.at(kv.first).dq;
      e.ops[kv.first] = kron_op(b, e.sec, pieces, kv.second, dq, tasks, contribs);
    }
    run_kron(tasks, contribs);
    for (auto& kv : e.ops) make_fp32(*kv.second);
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    e.pieces = std::move(pieces);
    return e;
  }

  // Identity basis record of a block that was not renormalized.
  std::shared_ptr<Basis> identity_basis(const Block& b) {
    auto B = std::make_shared<Basis>();
    int64_t tot = 0;
    for (size_t k = 0; k < b.sec.size(); ++k) {
      const int D = b.sec[k].dim;
      B->src.push_back((int)k);
      B->dim.push_back(D);
      B->m.push_back(D);
      B->uoff.push_back(tot);
      tot += (int64_t)D * D;
    }
    std::vector<double> h(std::max<int64_t>(1, tot), 0.0);
    for (size_t k = 0; k < b.sec.size(); ++k)
      for (int i = 0; i < B->dim[k]; ++i) h[B->uoff[k] + (int64_t)i * B->dim[k] + i] = 1.0;
    B->U.resize(ctx->pool, h.size());
    NSS_HIP(hipMemcpyAsync(B->U.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    return B;
  }

  // ------------------------------------------------------------ superblock
  // sb_get_states + Setup_SuperBlock_{Spin,Fermion}_Direct
  void setup_superblock(const Block& L, const Block& R, QN target) {
    sbL.clear();
    sbR.clear();
    std::vector<int> Dl, Dr;
    for (size_t k = 0; k < L.sec.size(); ++k) {
      const int r = R.find(target - L.sec[k].q);
      if (r < 0) continue;
      sbL.push_back((int)k);
      sbR.push_back(r);
      Dl.push_back(L.sec[k].dim);
      Dr.push_back(R.sec[r].dim);
    }
    NSS_REQUIRE(!sbL.empty(), NSS_E_STATE, "sb_get_states ERROR: total_states=0. There are no SB states.");
    const int nq = (int)sbL.size();
    sb_set_layout(sbd(), Dl, Dr, nullptr);
    lastDl = Dl;
    lastDr = Dr;
    std::map<QN, int> index;
    for (int k = 0; k < nq; ++k) index[L.sec[sbL[k]].q] = k;
    auto ref = [&](const BlockOp& op, const OpBlock& ob, int ld, int trans, double alpha) {
      OpRef r;
      r.p = op.data.p + ob.off;
      r.p32 = want32 ? (const void*)(op.data32.p + ob.off) : nullptr;
      r.ld = ld;
      r.trans = trans;
      r.alpha = alpha;
      return r;
    };
    std::vector<OpRef> HL(nq), HR(nq);
    const BlockOp& hl = *L.ops.at("H");
    const BlockOp& hr = *R.ops.at("H");
    for (int k = 0; k < nq; ++k) {
      const OpBlock* a = hl.row(sbL[k]);
      const OpBlock* b = hr.row(sbR[k]);
      NSS_REQUIRE(a && b, NSS_E_STATE, "missing H block");
      HL[k] = ref(hl, *a, Dl[k], 0, 1.0);
      HR[k] = ref(hr, *b, Dr[k], 0, 1.0);
    }
    std::vector<SbTerm> terms;
    // links of the connecting terms: 'n' (the inner ends), and with PBC 'p'
    // (the outer ends, suffix _p), in this order within each sector
    std::vector<std::string> links{""};
    if (par.pbc) links.push_back("_p");
    if (!site.fermion) {
      const double jz = hij[0], jxy = hij[3];
      for (int k = 0; k < nq; ++k) {
        const int lk = sbL[k], rk = sbR[k];
        for (const std::string& x : links) {
          const BlockOp& SzL = *L.ops.at("Sz" + x);
          const BlockOp& SzR = *R.ops.at("Sz" + x);
          const BlockOp& SpL = *L.ops.at("Sp" + x);
          const BlockOp& SpR = *R.ops.at("Sp" + x);
          if (jz != 0.0) {  // it=1 (it=4 for 'p'), SETUP:289-299
            const OpBlock* a = SzL.row(lk);
            const OpBlock* b = SzR.row(rk);
            if (a && b) terms.push_back({k, k, ref(SzL, *a, Dl[k], 0, jz), ref(SzR, *b, Dr[k], 1, 1.0)});
          }
          if (jxy == 0.0) continue;
          auto it = index.find(L.sec[lk].q - QN{2, 0});
          if (it == index.end()) continue;
          const int j = it->second, lj = sbL[j], rj = sbR[j];
          const OpBlock* a = SpL.row(lk);   // (lk, lj)
          const OpBlock* b = SpR.row(rj);   // (rj, rk)
          if (!a || a->cs != lj || !b || b->cs != rk) continue;
          // it=2: Y_k += Jxy Sp_L(lk,lj) Psi_j Sp_R(rj,rk)        SETUP:306-312
          terms.push_back({k, j, ref(SpL, *a, Dl[j], 0, jxy), ref(SpR, *b, Dr[k], 0, 1.0)});
          // it=3: hconjg partner, Y_j += A2^T Psi_k B2^T^T        SETUP:315-321
          terms.push_back({j, k, ref(SpL, *a, Dl[j], 1, jxy), ref(SpR, *b, Dr[k], 1, 1.0)});
        }
      }
    } else {
      const char* names[2] = {"Cup", "Cdn"};
      const QN dqs[2] = {{2, 0}, {0, 2}};
      for (int k = 0; k < nq; ++k) {
        const int lk = sbL[k], rk = sbR[k];
        for (int sp = 0; sp < 2; ++sp) {
          const double t = hij[sp * 2 + sp];
          if (t == 0.0) continue;
          auto it = index.find(L.sec[lk].q - dqs[sp]);
          if (it == index.end()) continue;
          const int j = it->second, lj = sbL[j], rj = sbR[j];
          for (const std::string& x : links) {
            const BlockOp& CL = *L.ops.at(names[sp] + x);
            const BlockOp& CR = *R.ops.at(names[sp] + x);
            const OpBlock* a = CL.row(lj);  // C_L(lj, lk)
            const OpBlock* b = CR.row(rk);  // C_R(rk, rj)
            if (!a || a->cs != lk || !b || b->cs != rj) continue;
            const double alpha = t * fsign(L.sec[lj].q);
            // it=1 (it=3 for 'p'): A = t (C_L^T P_L)(lk,lj), B = C_R(rk,rj)   SETUP:601-610
            terms.push_back({k, j, ref(CL, *a, Dl[k], 1, alpha), ref(CR, *b, Dr[j], 1, 1.0)});
            // it=2 (it=4): hconjg partner                                     SETUP:612-619
            terms.push_back({j, k, ref(CL, *a, Dl[k], 0, alpha), ref(CR, *b, Dr[j], 0, 1.0)});
          }
        }
      }
    }
    sb_set_terms(ctx, sbd(), HL, HR, terms, want32);
    lastTerms = terms;
  }

  // ------------------------------------------------------------ renormalize
  struct Spectrum {
    std::vector<std::vector<double>> ev;       // per SB sector (ascending)
    std::vector<RhoVec> dv;                    // eigenvectors on the device, per SB sector
    std::vector<double> all_desc;
    std::vector<std::pair<int, int>> order;    // (block, position) in descending order
  };

  static Spectrum sort_spectrum(std::vector<std::vector<double>>&& ev, std::vector<RhoVec>&& dv) {
    Spectrum s;
    s.ev = std::move(ev);
    s.dv = std::move(dv);
    std::vector<double> all;
    std::vector<std::pair<int, int>> pos;
    for (size_t b = 0; b < s.ev.size(); ++b)
      for (size_t i = 0; i < s.ev[b].size(); ++i) {
        all.push_back(s.ev[b][i]);
        pos.push_back({(int)b, (int)i});
      }
    std::vector<int> idx(all.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return all[a] < all[b]; });
    std::reverse(idx.begin(), idx.end());  // sort_quicksort + reverse, MATRIX_BLOCKS.f90:642-648
    for (int i : idx) {
      s.all_desc.push_back(all[i]);
      s.order.push_back(pos[i]);
    }
    return s;
  }

  // rho spectrum of the side that was not solved, per superblock sector q:
  // rho_L,q = Psi_q Psi_q^T and rho_R,q = Psi_q^T Psi_q have the same
  // nonzero eigenvalues, so the far block of dimension d_f takes the
  // largest min(d_n, d_f) eigenvalues of the near one, below them
  // d_f - min(d_n, d_f) null eigenvalues at tiny(1d0), the value the
  // reference gives a rho eigenvalue below zero (MATRIX_BLOCKS.f90:630).
  std::vector<std::vector<double>> far_spectrum(const std::vector<std::vector<double>>& near, int farside) const {
    std::vector<std::vector<double>> f(near.size());
    for (size_t q = 0; q < near.size(); ++q) {
      if ((int64_t)lastDl[q] * lastDr[q] == 0) continue;
      const int df = farside == 0 ? lastDl[q] : lastDr[q];
      const int dn = (int)near[q].size();
      const int k = std::min(dn, df);
      f[q].assign(df, DBL_MIN);
      std::copy(near[q].end() - k, near[q].end(), f[q].end() - k);
    }
    return f;
  }

  static double entropy(const std::vector<double>& ev) {
    double S = 0.0;
    for (double l : ev) {
      const double x = l < 0.0 ? 0.0 : l;
      S -= x * std::log(x + 1e-32);  // write_entanglement, DMRG_MAIN.f90:511-518
    }
    return S;
  }

  // The state selection of renormalize_block (src/DMRG_RDM.f90:195-251): the
  // kept count m_s, the truncation error, the weight at the cut and the new
  // sectors in order of first appearance over the descending spectrum (their
  // QNs in N.sec, the enlarged sector and rho block of each).
  struct Selection {
    int ms = 0;
    double trunc = 0.0, cut = 0.0;
    std::vector<int> src_sector, blk_of_new;
  };
  Selection select_states(const Block& E, const std::vector<int>& sbsec, const Spectrum& sp, int mstates,
                          double estates, Block& N, std::vector<std::pair<QN, std::vector<double>>>& kept) {
    Selection R;
    NSS_REQUIRE(!sp.all_desc.empty(), NSS_E_STATE, "renormalize: empty rho spectrum");
    const int rc = nss_kept_count(sp.all_desc.data(), (int)sp.all_desc.size(), mstates, estates, E.dim(),
                                  par.deg_thr, &R.ms, &R.trunc);
    if (rc != NSS_OK) throw Error(rc, nss_last_error());
    const int ms = R.ms = std::min<int>(R.ms, (int)sp.all_desc.size());
    std::vector<int> newsec_of_blk(sp.ev.size(), -1);
    N.length = E.length;
    for (int im = 0; im < ms; ++im) {
      const int b = sp.order[im].first;
      if (newsec_of_blk[b] < 0) {
        newsec_of_blk[b] = (int)N.sec.size();
        N.sec.push_back({E.sec[sbsec[b]].q, 0});
        R.src_sector.push_back(sbsec[b]);
        R.blk_of_new.push_back(b);
      }
      N.sec[newsec_of_blk[b]].dim++;
    }
    kept.clear();
    for (auto& s : N.sec) kept.push_back({s.q, {}});
    for (int im = 0; im < ms; ++im) {
      const int b = sp.order[im].first;
      kept[newsec_of_blk[b]].second.push_back(sp.all_desc[im]);
    }
    R.cut = sp.all_desc[ms - 1];
    return R;
  }

  // renormalize_block + operators_renormalization (src/DMRG_RDM.f90:172-404)
  Block renormalize(const Block& E, const std::vector<int>& sbsec, Spectrum& sp, int mstates,
                    double estates, int& ms_out, double& trunc, double& cut,
                    std::vector<std::pair<QN, std::vector<double>>>& kept) {
    Block N;
    const Selection sel = select_states(E, sbsec, sp, mstates, estates, N, kept);
    const int ms = sel.ms;
    const double te = sel.trunc;
    cut = sel.cut;
    const std::vector<int>& src_sector = sel.src_sector;  // enlarged sector of each new sector
    const std::vector<int>& blk_of_new = sel.blk_of_new;
    // U blocks (D x m, row-major): kept eigenvectors, descending eigenvalue
    std::vector<int64_t> uoff(N.sec.size());
    int64_t utot = 0;
    for (size_t k = 0; k < N.sec.size(); ++k) {
      uoff[k] = utot;
      utot += (int64_t)E.sec[src_sector[k]].dim * N.sec[k].dim;
    }
    // rho blocks of the large-block path: eigenvectors of the kept states only
    {
      std::vector<RhoVec*> dv;
      std::vector<int> cnt;
      for (size_t k = 0; k < N.sec.size(); ++k) {
        RhoVec& V = sp.dv[blk_of_new[k]];
        if (V.def && !V.base) {
          dv.push_back(&V);
          cnt.push_back(N.sec[k].dim);
        }
      }
      if (!dv.empty()) rho_materialize(ctx, dv, cnt);
    }
    // gathered on the device from the eigenvector blocks of the rho solve
    DevArr<double> du;
    du.resize(ctx->pool, std::max<int64_t>(1, utot));
    std::vector<int32_t> hcols;
    std::vector<GatherTask> gt;
    int max_el = 0;
    for (size_t k = 0; k < N.sec.size(); ++k) {
      const RhoVec& V = sp.dv[blk_of_new[k]];
      const int D = E.sec[src_sector[k]].dim, m = N.sec[k].dim;
      NSS_REQUIRE(V.base && V.d == D, NSS_E_STATE, "renormalize: rho eigenvectors missing");
      GatherTask t{};
      t.dst = du.p + uoff[k];
      t.src = V.base;
      t.si = V.si;
      t.sj = V.sj;
      t.cols = reinterpret_cast<const int32_t*>(hcols.size());  // offset, rebased below
      t.rows = D;
      t.ncol = m;
      for (int c = 0; c < m; ++c) hcols.push_back(V.order[D - 1 - c]);  // largest first
      gt.push_back(t);
      max_el = std::max(max_el, D * m);
    }
    DevArr<int32_t> dcols;
    DevArr<GatherTask> dgt;
    dcols.resize(ctx->pool, std::max<size_t>(1, hcols.size()));
    for (GatherTask& t : gt) t.cols = dcols.p + reinterpret_cast<size_t>(t.cols);
    dgt.resize(ctx->pool, std::max<size_t>(1, gt.size()));
    if (!gt.empty()) {
      NSS_HIP(hipMemcpyAsync(dcols.p, hcols.data(), hcols.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                             ctx->stream));
      NSS_HIP(hipMemcpyAsync(dgt.p, gt.data(), gt.size() * sizeof(GatherTask), hipMemcpyHostToDevice, ctx->stream));
      launch_gather_cols(dgt.p, (int)gt.size(), max_el, ctx->stream);
    }
    // map enlarged sector -> new sector
    std::vector<int> new_of_enl(E.sec.size(), -1);
    for (size_t k = 0; k < N.sec.size(); ++k) new_of_enl[src_sector[k]] = (int)k;
    // rotate every operator: T = O(rs,cs) U_cs ; O' = U_rs^T T
    GemmPlan p1, p2;
    std::vector<OpPtr> outs;
    std::vector<std::string> names;
    std::vector<DevArr<double>> mids;
    for (auto& kv : E.ops) {
      const BlockOp& O = *kv.second;
      OpPtr R = new_op(O.dq, N.sec, N.sec);
      int64_t mt = 0;
      std::vector<int64_t> moff;
      for (const OpBlock& nb : R->blks) {
        moff.push_back(mt);
        mt += (int64_t)E.sec[src_sector[nb.rs]].dim * N.sec[nb.cs].dim;
      }
      DevArr<double> mid;
      mid.resize(ctx->pool, std::max<int64_t>(1, mt));
      for (size_t i = 0; i < R->blks.size(); ++i) {
        const OpBlock& nb = R->blks[i];
        const int ers = src_sector[nb.rs], ecs = src_sector[nb.cs];
        const OpBlock* ob = O.row(ers);
        const int Dr = E.sec[ers].dim, Dc = E.sec[ecs].dim, mr = N.sec[nb.rs].dim, mc = N.sec[nb.cs].dim;
        NSS_REQUIRE(ob && ob->cs == ecs, NSS_E_STATE, "renormalize: operator block missing");
        GemmSeg g1{};
        g1.a = O.data.p + ob->off; g1.lda = Dc;
        g1.b = du.p + uoff[nb.cs]; g1.ldb = mc;
        g1.K = Dc; g1.alpha = 1.0; g1.abase = g1.bbase = BASE_ABS;
        p1.add_output(BASE_ABS, mid.p + moff[i], 0, mc, Dr, mc, {g1});
        GemmSeg g2{};
        g2.a = du.p + uoff[nb.rs]; g2.lda = mr; g2.flags = SEG_TA;
        g2.b = mid.p + moff[i]; g2.ldb = mc;
        g2.K = Dr; g2.alpha = 1.0; g2.abase = g2.bbase = BASE_ABS;
        p2.add_output(BASE_ABS, R->data.p + nb.off, 0, mc, mr, mc, {g2});
      }
      mids.push_back(std::move(mid));
      outs.push_back(R);
      names.push_back(kv.first);
    }
    p1.upload(ctx->pool, ctx->stream);
    p2.upload(ctx->pool, ctx->stream);
    launch_gemm_f64(p1, Bases{nullptr, nullptr, nullptr}, ctx->stream, GEMM_GENERIC);
    launch_gemm_f64(p2, Bases{nullptr, nullptr, nullptr}, ctx->stream, GEMM_GENERIC);
    for (size_t i = 0; i < outs.size(); ++i) N.ops[names[i]] = outs[i];
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    auto B = std::make_shared<Basis>();
    B->src = src_sector;
    for (size_t k = 0; k < N.sec.size(); ++k) {
      B->dim.push_back(E.sec[src_sector[k]].dim);
      B->m.push_back(N.sec[k].dim);
    }
    B->uoff = uoff;
    B->U = std::move(du);
    N.from = B;
    ms_out = ms;
    trunc = te;
    return N;
  }

  std::map<int, Block> bl1, bl2;

  // ------------------------------------------------------------ measurement
  // Measure_DMRG (src/DMRG_MEASURE.f90:148-706) on the ground state of the
  // last step: a site operator O at chain position pos is built as
  // Id (x) O in the enlarged basis of the block of length i (Build_Op_dmrg,
  // :288-396), carried to the superblock basis by "rotate with the block's
  // kept-state matrix, enlarge with (x) Id" (Advance_Op_dmrg, :398-496), and
  // averaged as <psi| O (x) 1 |psi> or <psi| 1 (x) O |psi> over the
  // superblock sectors (Average_Op_dmrg + OdotV_direct, :595-706).  O must
  // conserve the quantum numbers (the reference keeps the QN-diagonal blocks
  // of O only).  Everything stays on the device.

  // O (site basis) in the basis of the one-site block b1.
  OpPtr site_op_in_block(const std::vector<double>& O, const Block& b1) {
    const int d = site.dim;
    std::vector<int> ssec(d), spos(d), cnt(b1.sec.size(), 0);
    for (int st = 0; st < d; ++st) {
      ssec[st] = b1.find(site.qn[st]);
      NSS_REQUIRE(ssec[st] >= 0, NSS_E_STATE, "measure: one-site block does not match the site");
      spos[st] = cnt[ssec[st]]++;
    }
    OpPtr op = new_op({0, 0}, b1.sec, b1.sec);
    std::vector<double> h(op->data.n, 0.0);
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) {
        const double v = O[i * d + j];
        if (v == 0.0) continue;
        const OpBlock* ob = op->row(ssec[i]);
        NSS_REQUIRE(ob && ob->cs == ssec[j], NSS_E_ARG, "measure: operator does not conserve the quantum numbers");
        h[ob->off + (size_t)spos[i] * b1.sec[ssec[j]].dim + spos[j]] = v;
      }
    NSS_HIP(hipMemcpyAsync(op->data.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    return op;
  }

  // U^T O U per sector of the block whose basis record is B (sectors nsec).
  OpPtr rotate_op(const BlockOp& O, const Basis& B, const std::vector<Sector>& nsec) {
    OpPtr R = new_op({0, 0}, nsec, nsec);
    NSS_HIP(hipMemsetAsync(R->data.p, 0, R->data.n * sizeof(double), ctx->stream));
    int64_t mt = 0;
    std::vector<int64_t> moff(R->blks.size(), -1);
    for (size_t i = 0; i < R->blks.size(); ++i) {
      const int k = R->blks[i].rs;
      moff[i] = mt;
      mt += (int64_t)B.dim[k] * B.m[k];
    }
    DevArr<double> mid;
    mid.resize(ctx->pool, std::max<int64_t>(1, mt));
    GemmPlan p1, p2;
    for (size_t i = 0; i < R->blks.size(); ++i) {
      const OpBlock& nb = R->blks[i];
      const int k = nb.rs, e = B.src[k], D = B.dim[k], m = B.m[k];
      const OpBlock* ob = O.row(e);
      if (!ob || ob->cs != e) continue;  // zero block
      const double* Uk = B.U.p + B.uoff[k];
      GemmSeg g1{};
      g1.a = O.data.p + ob->off; g1.lda = D;
      g1.b = Uk; g1.ldb = m;
      g1.K = D; g1.alpha = 1.0; g1.abase = g1.bbase = BASE_ABS;
      p1.add_output(BASE_ABS, mid.p + moff[i], 0, m, D, m, {g1});
      GemmSeg g2{};
      g2.a = Uk; g2.lda = m; g2.flags = SEG_TA;
      g2.b = mid.p + moff[i]; g2.ldb = m;
      g2.K = D; g2.alpha = 1.0; g2.abase = g2.bbase = BASE_ABS;
      p2.add_output(BASE_ABS, R->data.p + nb.off, 0, m, m, m, {g2});
    }
    p1.upload(ctx->pool, ctx->stream);
    p2.upload(ctx->pool, ctx->stream);
    launch_gemm_f64(p1, Bases{nullptr, nullptr, nullptr}, ctx->stream, GEMM_GENERIC);
    launch_gemm_f64(p2, Bases{nullptr, nullptr, nullptr}, ctx->stream, GEMM_GENERIC);
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    return R;
  }

  // <psi| O (x) 1 |psi> (left) or <psi| 1 (x) O |psi> (right), O in the basis
  // of the last superblock's enlarged block on that side.
  double expect(const BlockOp& O, bool left_side) {
    const int nq = (int)lastDl.size();
    int64_t n = 0;
    std::vector<int64_t> off(nq);
    for (int q = 0; q < nq; ++q) {
      off[q] = n;
      n += (int64_t)lastDl[q] * lastDr[q];
    }
    DevArr<double> X, part;
    X.resize(ctx->pool, std::max<int64_t>(1, n));
    part.resize(ctx->pool, gemv_t_chunks(n) + 2);
    NSS_HIP(hipMemsetAsync(X.p, 0, n * sizeof(double), ctx->stream));
    GemmPlan p;
    for (int q = 0; q < nq; ++q) {
      const int Dl = lastDl[q], Dr = lastDr[q];
      const OpBlock* ob = O.row(left_side ? sbL[q] : sbR[q]);
      if (!ob || ob->cs != (left_side ? sbL[q] : sbR[q])) continue;
      GemmSeg g{};
      g.alpha = 1.0;
      g.abase = g.bbase = BASE_ABS;
      if (left_side) {  // X_q = O Psi_q
        g.a = O.data.p + ob->off; g.lda = Dl;
        g.b = psi.p + off[q]; g.ldb = Dr;
        g.K = Dl;
      } else {          // X_q = Psi_q O^T
        g.a = psi.p + off[q]; g.lda = Dr;
        g.b = O.data.p + ob->off; g.ldb = Dr; g.flags = SEG_TB;
        g.K = Dr;
      }
      p.add_output(BASE_ABS, X.p + off[q], 0, Dr, Dl, Dr, {g});
    }
    p.upload(ctx->pool, ctx->stream);
    launch_gemm_f64(p, Bases{nullptr, nullptr, nullptr}, ctx->stream, GEMM_GENERIC);
    launch_gemv_t(psi.p, n, 1, X.p, n, part.p, part.p + gemv_t_chunks(n), ctx->stream);
    double v = 0.0;
    NSS_HIP(hipMemcpyAsync(&v, part.p + gemv_t_chunks(n), sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    return v;
  }

  double measure_one(const std::vector<double>& O, int pos) {
    const int L = lastL.length, R = lastR.length;
    NSS_REQUIRE(L > 0 && !lastDl.empty(), NSS_E_STATE, "measure: run DMRG first");
    NSS_REQUIRE(pos >= 1 && pos <= L + R, NSS_E_ARG, "measure: position outside [1, L+R]");
    const bool left_side = pos <= L;
    const int i = left_side ? pos : L + R + 1 - pos;  // site index inside its block
    const int Lb = left_side ? L : R;
    // history of the last superblock's block on that side: bl[k] has length k
    std::map<int, const Block*> bl;
    for (const Block* b = (left_side ? lastL : lastR).parent.get(); b; b = b->parent.get()) bl[b->length] = b;
    const QN z{0, 0};
    std::vector<double> I(site.dim * site.dim, 0.0);
    for (int k = 0; k < site.dim; ++k) I[k * site.dim + k] = 1.0;
    OpPtr cur;
    const Block* cb = nullptr;  // block whose enlargement the operator lives in (nullptr: one-site block)
    int it0;
    if (i == 1) {
      NSS_REQUIRE(bl.count(1), NSS_E_STATE, "measure: block history missing");
      cur = site_op_in_block(O, *bl.at(1));
      it0 = 1;
    } else {
      NSS_REQUIRE(bl.count(i - 1), NSS_E_STATE, "measure: block history missing");
      cb = bl.at(i - 1);
      std::vector<Sector> esec;
      std::vector<std::vector<Piece>> pieces;
      enlarged_basis(*cb, esec, pieces);
      std::vector<KronTask> tasks;
      std::vector<KronContrib> contribs;
      for (int a = 0; a < site.dim; ++a)
        for (int b = 0; b < site.dim; ++b)
          NSS_REQUIRE(O[a * site.dim + b] == 0.0 || site.qn[a] == site.qn[b], NSS_E_ARG,
                      "measure: operator does not conserve the quantum numbers");
      cur = kron_op(*cb, esec, pieces, {KTerm{nullptr, 0, 0, 0, O, z, 1.0}}, z, tasks, contribs);
      run_kron(tasks, contribs);
      it0 = i;
    }
    for (int it = it0; it <= Lb - 1; ++it) {
      NSS_REQUIRE(bl.count(it), NSS_E_STATE, "measure: block history missing");
      const Block& b = *bl.at(it);
      if (it > 1) {  // rotate from the enlargement of bl[it-1] into bl[it]
        NSS_REQUIRE(b.from, NSS_E_STATE, "measure: block without a basis record");
        cur = rotate_op(*cur, *b.from, b.sec);
      }
      std::vector<Sector> esec;
      std::vector<std::vector<Piece>> pieces;
      enlarged_basis(b, esec, pieces);
      std::vector<KronTask> tasks;
      std::vector<KronContrib> contribs;
      // the source operator must stay alive until k_kron has read it
      OpPtr next = kron_op(b, esec, pieces, {KTerm{cur.get(), 0, 0, 0, I, z, 1.0}}, z, tasks, contribs);
      run_kron(tasks, contribs);
      cur = std::move(next);
    }
    return expect(*cur, left_side);
  }

  // ------------------------------------------------------------ restart I/O
  // write_block / read_block (src/BLOCKS/BLOCKS.f90:418-537) in the
  // reference's list-directed text format: length, Dim, SiteType, OpName,
  // the sectors list (QN and 1-based state map per sector,
  // src/LIST/LIST_SECTORS.f90:639-705), the operators list (key, type and
  // sparse rows per operator, src/LIST/LIST_OPERATORS.f90:602-659,
  // src/MATRIX/MATRIX_SPARSE.f90:703-766) and the rotation-matrix list.
  // Keys of both generations are read (S_z / S_z_n, C_l1_s1 / C_l1_s1_n,
  // P / P_n); the current ones are written.

  struct Tokens {
    std::vector<std::string> t;
    size_t i = 0;
    const std::string& next() {
      NSS_REQUIRE(i < t.size(), NSS_E_ARG, "read_block: unexpected end of file");
      return t[i++];
    }
    long inum() { return std::strtol(next().c_str(), nullptr, 10); }
    double rnum() {
      std::string v = next();
      for (char& c : v)
        if (c == 'D' || c == 'd') c = 'E';
      return std::strtod(v.c_str(), nullptr);
    }
  };

  // Position of every state of the enlarged basis of `parent` (sector q, index
  // within it) by its index in the reference's Kronecker order: block (x) site
  // on the left (a d + s), site (x) block on the right (s dim + a), with a the
  // parent state's index in the order a file lists the parent's basis (forder:
  // file index -> (sector, index); empty: the parent's own order).
  std::vector<std::pair<int, int>> kron_positions(const Block& parent, bool left, std::vector<Sector>& esec,
                                                   const std::vector<std::pair<int, int>>& forder = {}) const {
    std::vector<std::vector<Piece>> pieces;
    enlarged_basis(parent, esec, pieces);
    std::vector<int> aoff(parent.sec.size() + 1, 0);
    for (size_t a = 0; a < parent.sec.size(); ++a) aoff[a + 1] = aoff[a] + parent.sec[a].dim;
    const int dp = aoff.back(), d = site.dim;
    std::vector<int> fidx(dp);  // own index -> file index
    for (int g = 0; g < dp; ++g) fidx[g] = g;
    if (!forder.empty()) {
      NSS_REQUIRE((int)forder.size() == dp, NSS_E_ARG, "read_block: rotation matrix does not match the block below");
      for (int g = 0; g < dp; ++g) fidx[aoff[forder[g].first] + forder[g].second] = g;
    }
    std::vector<std::pair<int, int>> pos((size_t)dp * d, {-1, -1});
    for (size_t q = 0; q < pieces.size(); ++q)
      for (const Piece& pc : pieces[q])
        for (int ia = 0; ia < parent.sec[pc.a].dim; ++ia) {
          const int g = fidx[aoff[pc.a] + ia];
          pos[left ? (size_t)g * d + pc.s : (size_t)pc.s * dp + g] = {(int)q, pc.off + ia};
        }
    return pos;
  }

  Block load_block(const std::string& path, bool left) {
    std::ifstream in(path);
    NSS_REQUIRE(in.good(), NSS_E_ARG, ("read_block: cannot open " + path).c_str());
    Tokens tk;
    for (std::string w; in >> w;) tk.t.push_back(w);
    Block b;
    b.length = (int)tk.inum();
    const long dim = tk.inum();
    const std::string stype = tk.next();
    (void)tk.next();  // OpName
    NSS_REQUIRE((stype == "FERMION") == site.fermion, NSS_E_ARG, "read_block: site type does not match the model");
    const long nlists = tk.inum();
    NSS_REQUIRE(nlists == 1, NSS_E_ARG, "read_block: one sectors list expected");
    const long nsec = tk.inum();
    const long qdim = tk.inum();
    std::vector<std::pair<int, int>> pos(dim, {-1, -1});  // old state -> (sector, index)
    for (long k = 0; k < nsec; ++k) {
      const long dq = tk.inum();
      NSS_REQUIRE(dq == qdim && (qdim == 1 || qdim == 2), NSS_E_ARG, "read_block: bad QN size");
      QN q;
      q.a = (int)std::lround(2.0 * tk.rnum());
      q.b = qdim == 2 ? (int)std::lround(2.0 * tk.rnum()) : 0;
      const long nm = tk.inum();
      b.sec.push_back({q, (int)nm});
      for (long j = 0; j < nm; ++j) {
        const long st = tk.inum() - 1;
        NSS_REQUIRE(st >= 0 && st < dim && pos[st].first < 0, NSS_E_ARG, "read_block: bad sector map");
        pos[st] = {(int)k, (int)j};
      }
    }
    for (long st = 0; st < dim; ++st) NSS_REQUIRE(pos[st].first >= 0, NSS_E_ARG, "read_block: state in no sector");
    // link-'p' keys (okey's _p suffix, src/AUX_FUNCS.f90:108-112) are read
    // only by a PBC run
    static const std::map<std::string, std::string> keymap_obc = {
        {"H", "H"},         {"S_z", "Sz"},        {"S_z_n", "Sz"},      {"S_p", "Sp"},
        {"S_p_n", "Sp"},    {"C_l1_s1", "Cup"},   {"C_l1_s1_n", "Cup"}, {"C_l1_s2", "Cdn"},
        {"C_l1_s2_n", "Cdn"}};
    std::map<std::string, std::string> keymap = keymap_obc;
    if (par.pbc)
      for (const auto& kv : std::map<std::string, std::string>{
               {"S_z_p", "Sz_p"}, {"S_p_p", "Sp_p"}, {"C_l1_s1_p", "Cup_p"}, {"C_l1_s2_p", "Cdn_p"}})
        keymap.insert(kv);
    const long nops = tk.inum();
    for (long o = 0; o < nops; ++o) {
      const std::string key = tk.next();
      (void)tk.next();  // type
      const long nr = tk.inum(), nc = tk.inum();
      NSS_REQUIRE(nr == dim && nc == dim, NSS_E_ARG, "read_block: operator shape differs from Dim");
      struct E { int r, c; double v; };
      std::vector<E> ent;
      for (long r = 0; r < nr; ++r) {
        const long nz = tk.inum();
        for (long z = 0; z < nz; ++z) {
          const long c = tk.inum() - 1;
          const double v = tk.rnum();
          NSS_REQUIRE(c >= 0 && c < nc, NSS_E_ARG, "read_block: column out of range");
          if (v != 0.0) ent.push_back({(int)r, (int)c, v});
        }
      }
      auto it = keymap.find(key);
      if (it == keymap.end()) continue;  // P (implicit here) and any other key
      const QN dq = site.ops.at(it->second).dq;
      OpPtr op = new_op(dq, b.sec, b.sec);
      std::vector<double> h(op->data.n, 0.0);
      for (const E& e : ent) {
        const auto pr = pos[e.r], pc = pos[e.c];
        const OpBlock* ob = op->row(pr.first);
        NSS_REQUIRE(ob && ob->cs == pc.first, NSS_E_ARG,
                    ("read_block: operator " + key + " does not have the QN shift of " + it->second).c_str());
        h[ob->off + (size_t)pr.second * b.sec[pc.first].dim + pc.second] = e.v;
      }
      NSS_HIP(hipMemcpyAsync(op->data.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      NSS_HIP(hipStreamSynchronize(ctx->stream));
      make_fp32(*op);
      b.ops[it->second] = op;
    }
    for (const auto& kv : site.ops)
      NSS_REQUIRE(b.ops.count(kv.first), NSS_E_ARG, ("read_block: operator missing: " + kv.first).c_str());
    // rotation matrices (omatrices, keys "1".."length"): with all of them the
    // block's history is rebuilt, which measurement walks; with fewer (older
    // files keep only "1") the block continues a run but is not measured,
    // as in the reference (src/BLOCKS/BLOCKS.f90:526-528).  Row r of matrix
    // k is the state (a, s) of the enlarged block of length k-1, a d + s on
    // the left and s dim + a on the right, a in the order that file lists
    // the block of length k-1 (the columns of matrix k-1; the site's own
    // order under matrix "1")
    struct Sp { long nr = 0, nc = 0; std::vector<long> r, c; std::vector<double> v; };
    std::map<long, Sp> om;
    if (tk.i < tk.t.size()) {
      const long nom = tk.inum();
      for (long o = 0; o < nom; ++o) {
        const long key = std::strtol(tk.next().c_str(), nullptr, 10);
        (void)tk.next();  // type
        Sp u;
        u.nr = tk.inum();
        u.nc = tk.inum();
        for (long r = 0; r < u.nr; ++r) {
          const long nz = tk.inum();
          for (long z = 0; z < nz; ++z) {
            const long c = tk.inum() - 1;
            const double v = tk.rnum();
            NSS_REQUIRE(c >= 0 && c < u.nc, NSS_E_ARG, "read_block: rotation matrix column out of range");
            if (v != 0.0) {
              u.r.push_back(r);
              u.c.push_back(c);
              u.v.push_back(v);
            }
          }
        }
        om[key] = std::move(u);
      }
    }
    bool history = b.length >= 1;
    for (long k = 2; k <= b.length; ++k) history = history && om.count(k);
    // best effort: rotation matrices that do not fit the block (another
    // writer's convention) leave it without history, as a file with "1" only
    if (history && b.length >= 2) try {
      std::shared_ptr<const Block> par = std::make_shared<Block>(block_from_site());
      // the one-site block's file order: column c of "1" holds site state s
      // (the reference writes the identity: the site's own order)
      std::vector<std::pair<int, int>> forder(site.dim);
      {
        std::vector<int> rank(site.dim, 0), cnt(par->sec.size(), 0);
        for (int st = 0; st < site.dim; ++st) rank[st] = cnt[par->find(site.qn[st])]++;
        std::vector<int> state_of(site.dim);
        for (int c = 0; c < site.dim; ++c) state_of[c] = c;
        if (om.count(1)) {
          const Sp& u1 = om.at(1);
          NSS_REQUIRE(u1.nr == site.dim && u1.nc == site.dim, NSS_E_ARG, "read_block: rotation matrix 1 is not the site's");
          for (size_t z = 0; z < u1.v.size(); ++z) state_of[u1.c[z]] = (int)u1.r[z];
        }
        for (int c = 0; c < site.dim; ++c) forder[c] = {par->find(site.qn[state_of[c]]), rank[state_of[c]]};
      }
      for (int k = 2; k <= b.length; ++k) {
        const Sp& u = om.at(k);
        std::vector<Sector> esec;
        const auto kp = kron_positions(*par, left, esec, forder);
        NSS_REQUIRE(u.nr == (long)kp.size(), NSS_E_ARG, "read_block: rotation matrix rows differ from the enlarged block");
        // each column's enlarged sector
        std::vector<int> colq(u.nc, -1);
        for (size_t z = 0; z < u.v.size(); ++z) {
          const int q = kp[u.r[z]].first;
          NSS_REQUIRE(colq[u.c[z]] < 0 || colq[u.c[z]] == q, NSS_E_ARG, "read_block: rotation matrix mixes QN sectors");
          colq[u.c[z]] = q;
        }
        // the new basis: the stored block's own sector map at its length,
        // sectors in order of first appearance (columns in file order) below
        Block nb;
        nb.length = k;
        std::vector<std::pair<int, int>> cpos(u.nc);
        if (k == b.length) {
          NSS_REQUIRE(u.nc == dim, NSS_E_ARG, "read_block: last rotation matrix does not match Dim");
          nb.sec = b.sec;
          for (long c = 0; c < u.nc; ++c) cpos[c] = pos[c];
        } else {
          std::map<int, int> secof;
          for (long c = 0; c < u.nc; ++c) {
            NSS_REQUIRE(colq[c] >= 0, NSS_E_ARG, "read_block: empty rotation matrix column");
            auto it = secof.find(colq[c]);
            if (it == secof.end()) {
              it = secof.emplace(colq[c], (int)nb.sec.size()).first;
              nb.sec.push_back({esec[colq[c]].q, 0});
            }
            cpos[c] = {it->second, nb.sec[it->second].dim++};
          }
        }
        auto B = std::make_shared<Basis>();
        int64_t tot = 0;
        for (size_t j = 0; j < nb.sec.size(); ++j) {
          int q = -1;
          for (long c = 0; c < u.nc; ++c)
            if (cpos[c].first == (int)j && colq[c] >= 0) { q = colq[c]; break; }
          NSS_REQUIRE(q >= 0 && esec[q].q == nb.sec[j].q, NSS_E_ARG, "read_block: rotation matrix and sector map disagree");
          B->src.push_back(q);
          B->dim.push_back(esec[q].dim);
          B->m.push_back(nb.sec[j].dim);
          B->uoff.push_back(tot);
          tot += (int64_t)esec[q].dim * nb.sec[j].dim;
        }
        std::vector<double> h(std::max<int64_t>(1, tot), 0.0);
        for (size_t z = 0; z < u.v.size(); ++z) {
          const auto [q, loc] = kp[u.r[z]];
          const auto [j, ic] = cpos[u.c[z]];
          NSS_REQUIRE(B->src[j] == q, NSS_E_ARG, "read_block: rotation matrix and sector map disagree");
          h[B->uoff[j] + (int64_t)loc * B->m[j] + ic] = u.v[z];
        }
        B->U.resize(ctx->pool, h.size());
        NSS_HIP(hipMemcpyAsync(B->U.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        NSS_HIP(hipStreamSynchronize(ctx->stream));
        forder = cpos;  // this block's file order, for the rows of the next matrix
        if (k == b.length) {
          b.from = B;
          b.parent = par;
        } else {
          nb.from = B;
          nb.parent = par;
          par = std::make_shared<Block>(std::move(nb));
        }
      }
    } catch (const Error& e) {
      if (e.code != NSS_E_ARG) throw;
      NSS_HIP(hipStreamSynchronize(ctx->stream));
      b.from.reset();
      b.parent.reset();
    }
    NSS_HIP(hipStreamSynchronize(ctx->stream));
    return b;
  }

  void save_block(const Block& b, const std::string& path, bool left) {
    std::ofstream out(path);
    NSS_REQUIRE(out.good(), NSS_E_ARG, ("write_block: cannot open " + path).c_str());
    out.precision(17);
    const int dim = b.dim();
    const int qdim = site.fermion ? 2 : 1;
    out << b.length << "\n" << dim << "\n" << (site.fermion ? "FERMION" : "SPIN") << "\n"
        << (site.fermion ? "C" : "S") << "\n1\n" << b.sec.size() << "\n" << qdim << "\n";
    std::vector<int> off(b.sec.size() + 1, 0);
    for (size_t k = 0; k < b.sec.size(); ++k) off[k + 1] = off[k] + b.sec[k].dim;
    for (size_t k = 0; k < b.sec.size(); ++k) {
      out << qdim << "\n" << 0.5 * b.sec[k].q.a;
      if (qdim == 2) out << " " << 0.5 * b.sec[k].q.b;
      out << "\n" << b.sec[k].dim << "\n";
      for (int j = 0; j < b.sec[k].dim; ++j) out << off[k] + j + 1 << (j + 1 < b.sec[k].dim ? " " : "\n");
    }
    struct W { std::string key, type, name; };
    std::vector<W> ws = site.fermion
        ? std::vector<W>{{"H", "bosonic", "H"}, {"C_l1_s1_n", "fermionic", "Cup"}, {"C_l1_s2_n", "fermionic", "Cdn"},
                         {"P_n", "psign", ""}}
        : std::vector<W>{{"H", "bosonic", "H"}, {"S_z_n", "bosonic", "Sz"}, {"S_p_n", "bosonic", "Sp"}};
    if (par.pbc) {
      if (site.fermion)
        ws.insert(ws.end(), {{"C_l1_s1_p", "fermionic", "Cup_p"}, {"C_l1_s2_p", "fermionic", "Cdn_p"},
                             {"P_p", "psign", ""}});
      else
        ws.insert(ws.end(), {{"S_z_p", "bosonic", "Sz_p"}, {"S_p_p", "bosonic", "Sp_p"}});
    }
    out << ws.size() << "\n";
    for (const W& w : ws) {
      std::vector<double> dense((size_t)dim * dim, 0.0);
      if (w.name.empty()) {  // fermionic parity, diagonal
        for (size_t k = 0; k < b.sec.size(); ++k)
          for (int j = off[k]; j < off[k + 1]; ++j) dense[(size_t)j * dim + j] = fsign(b.sec[k].q);
      } else {
        const BlockOp& op = *b.ops.at(w.name);
        std::vector<double> h(op.data.n);
        NSS_HIP(hipMemcpy(h.data(), op.data.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (const OpBlock& ob : op.blks)
          for (int r = 0; r < b.sec[ob.rs].dim; ++r)
            for (int c = 0; c < b.sec[ob.cs].dim; ++c)
              dense[(size_t)(off[ob.rs] + r) * dim + off[ob.cs] + c] = h[ob.off + (size_t)r * b.sec[ob.cs].dim + c];
      }
      out << w.key << "\n" << w.type << "\n" << dim << " " << dim << "\n";
      for (int r = 0; r < dim; ++r) {
        int nz = 0;
        for (int c = 0; c < dim; ++c) nz += dense[(size_t)r * dim + c] != 0.0;
        out << nz << "\n";
        for (int c = 0; c < dim; ++c)
          if (dense[(size_t)r * dim + c] != 0.0) out << c + 1 << " " << dense[(size_t)r * dim + c] << "\n";
      }
    }
    // rotation matrices (omatrices, STORE_UMAT='b', src/DMRG_RDM.f90:223-228):
    // "1" the one-site block's basis, "k" the kept states of the block of
    // length k as columns over the enlarged block of length k-1 (rows in the
    // Kronecker order of block (x) site, or site (x) block on the right),
    // when the block carries its whole history
    std::vector<const Block*> chain(b.length + 1, nullptr);
    chain[b.length] = &b;
    for (int k = b.length; k > 1 && chain[k]; --k) chain[k - 1] = chain[k]->parent.get();
    bool full = chain[1] != nullptr && chain[1]->length == 1;
    for (int k = 2; k <= b.length && full; ++k) full = chain[k] && chain[k]->from && chain[k]->length == k;
    auto write_sparse = [&](const std::string& key, int nr, int nc, const std::vector<double>& dense) {
      out << key << "\nnone\n" << nr << " " << nc << "\n";
      for (int r = 0; r < nr; ++r) {
        int nz = 0;
        for (int c = 0; c < nc; ++c) nz += dense[(size_t)r * nc + c] != 0.0;
        out << nz << "\n";
        for (int c = 0; c < nc; ++c)
          if (dense[(size_t)r * nc + c] != 0.0) out << c + 1 << " " << dense[(size_t)r * nc + c] << "\n";
      }
    };
    out << (full ? b.length : 1) << "\n";
    {  // "1": site state s (rows, natural order) -> the one-site block's state
      const Block b1 = full ? *chain[1] : Block{};
      const int d = site.dim;
      std::vector<double> u((size_t)d * d, 0.0);
      std::vector<int> cnt(std::max<size_t>(1, b1.sec.size()), 0), soff(b1.sec.size() + 1, 0);
      for (size_t k = 0; k < b1.sec.size(); ++k) soff[k + 1] = soff[k] + b1.sec[k].dim;
      for (int st = 0; st < d; ++st) {
        const int k = full ? b1.find(site.qn[st]) : -1;
        u[(size_t)st * d + (k >= 0 ? soff[k] + cnt[k]++ : st)] = 1.0;
      }
      write_sparse("1", d, d, u);
    }
    for (int k = 2; full && k <= b.length; ++k) {
      const Block& bk = *chain[k];
      const Basis& B = *bk.from;
      std::vector<Sector> esec;
      const auto kp = kron_positions(*chain[k - 1], left, esec);
      const int nr = (int)kp.size(), nc = bk.dim();
      std::vector<int> coff(bk.sec.size() + 1, 0);
      for (size_t j = 0; j < bk.sec.size(); ++j) coff[j + 1] = coff[j] + bk.sec[j].dim;
      int64_t tot = 0;
      for (size_t j = 0; j < B.src.size(); ++j) tot += (int64_t)B.dim[j] * B.m[j];
      std::vector<double> h(std::max<int64_t>(1, tot));
      NSS_HIP(hipMemcpy(h.data(), B.U.p, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost));
      std::vector<double> u((size_t)nr * nc, 0.0);
      for (int r = 0; r < nr; ++r) {
        const auto [q, loc] = kp[r];
        for (size_t j = 0; j < B.src.size(); ++j) {
          if (B.src[j] != q) continue;
          for (int c = 0; c < B.m[j]; ++c) u[(size_t)r * nc + coff[j] + c] = h[B.uoff[j] + (int64_t)loc * B.m[j] + c];
        }
      }
      write_sparse(std::to_string(k), nr, nc, u);
    }
    NSS_REQUIRE(out.good(), NSS_E_HIP, ("write_block: write failed for " + path).c_str());
  }

  // ------------------------------------------------------------ prediction
  // Ground state of the previous finite step and the basis maps around it.
  struct PrevStep {
    bool valid = false;
    int label = 0, neig = 0;
    int64_t n = 0;
    std::vector<int> sbL, sbR, Dl, Dr;
    std::vector<std::vector<Piece>> piecesL, piecesR;  // its enlarged blocks
    std::shared_ptr<Basis> unenlL, unenlR;             // blocks that were enlarged
    std::shared_ptr<Basis> newL, newR;                 // blocks after renormalization
    DevArr<double> psi, sum;
  } prev;

  // White's wavefunction transformation between two finite steps in the
  // same direction.  Left->right (label 1): the previous ground state
  // Psi[(a,s1),(s2,b)] becomes
  //     Psi'[(k,s2),(c,s3)] = sum U_L[(a,s1),k] Psi[(a,s1),(s2,b)] U_R[(c,s3),b]
  // with U_L the kept states of the block just renormalized and U_R the
  // matrix that built the previous right block from the enlargement of the
  // current one; right->left (label 2) is the mirror image.  Both products
  // are sector-blocked grouped GEMMs: T = U_L^T Psi_q (or Psi_q U_R), then
  // one product per piece of the far enlarged sector into its output block.
  // Writes the predicted vector (the sum over the previous eigenvectors when
  // LANC_NEIGEN > 1) to out; returns false when no prediction applies.
  bool predict_psi(int label, double* out, int64_t n_new) {
    const Basis* rot = label == 1 ? prev.newL.get() : prev.newR.get();
    const Basis* far = label == 1 ? prev.unenlR.get() : prev.unenlL.get();
    if (!rot || !far) return false;
    const Block& near_enl = label == 1 ? left : right;
    const Block& far_enl = label == 1 ? right : left;
    const std::vector<std::vector<Piece>>& prev_near_pieces = label == 1 ? prev.piecesL : prev.piecesR;
    const std::vector<std::vector<Piece>>& prev_far_pieces = label == 1 ? prev.piecesR : prev.piecesL;
    // the far block's map must refer to the enlargement used now
    for (size_t b = 0; b < far->src.size(); ++b)
      if (far->src[b] >= (int)far_enl.sec.size() || far_enl.sec[far->src[b]].dim != far->dim[b]) return false;
    std::vector<int> k_of_e(prev_near_pieces.size(), -1);
    for (size_t k = 0; k < rot->src.size(); ++k)
      if (rot->src[k] < (int)k_of_e.size()) k_of_e[rot->src[k]] = (int)k;
    std::map<std::pair<int, int>, std::pair<int, int>> piece_at;  // (block sector, site state) -> (sector, offset)
    for (size_t c = 0; c < near_enl.pieces.size(); ++c)
      for (const Piece& pc : near_enl.pieces[c]) piece_at[{pc.a, pc.s}] = {(int)c, pc.off};
    const std::vector<int>& sb_near = label == 1 ? sbL : sbR;
    const std::vector<int>& sb_far = label == 1 ? sbR : sbL;
    std::vector<int> q_of_near(near_enl.sec.size(), -1);
    for (size_t q = 0; q < sb_near.size(); ++q) q_of_near[sb_near[q]] = (int)q;
    std::vector<int64_t> noff(lastDl.size() + 1, 0), poff(prev.Dl.size() + 1, 0);
    for (size_t q = 0; q < lastDl.size(); ++q) noff[q + 1] = noff[q] + (int64_t)lastDl[q] * lastDr[q];
    for (size_t q = 0; q < prev.Dl.size(); ++q) poff[q + 1] = poff[q] + (int64_t)prev.Dl[q] * prev.Dr[q];
    if (noff.back() != n_new || poff.back() != prev.n) return false;
    hipStream_t st = ctx->stream;
    const double* x = prev.psi.p;
    if (prev.neig > 1) {  // sum of the previous eigenvectors: w = x_0 - X_{1..} (-1)
      prev.sum.resize(ctx->pool, prev.n + prev.neig);
      std::vector<double> m1(prev.neig, -1.0);
      NSS_HIP(hipMemcpyAsync(prev.sum.p + prev.n, m1.data(), prev.neig * sizeof(double), hipMemcpyHostToDevice, st));
      NSS_HIP(hipMemcpyAsync(prev.sum.p, prev.psi.p, prev.n * sizeof(double), hipMemcpyDeviceToDevice, st));
      launch_gemv_n(prev.psi.p + prev.n, prev.n, prev.neig - 1, prev.sum.p + prev.n, prev.sum.p, prev.n, 0, st);
      x = prev.sum.p;
    }
    GemmPlan p1, p2;
    std::vector<int64_t> toff(prev.Dl.size(), -1);
    int64_t ttot = 0;
    for (size_t q = 0; q < prev.Dl.size(); ++q) {
      const int k = k_of_e[label == 1 ? prev.sbL[q] : prev.sbR[q]];
      if (k < 0) continue;
      toff[q] = ttot;
      ttot += (int64_t)rot->m[k] * (label == 1 ? prev.Dr[q] : prev.Dl[q]);
    }
    DevArr<double> T;
    T.resize(ctx->pool, std::max<int64_t>(1, ttot));
    int pieces_done = 0;
    for (size_t q = 0; q < prev.Dl.size(); ++q) {
      if (toff[q] < 0) continue;
      const int eL = prev.sbL[q], eR = prev.sbR[q];
      const int k = k_of_e[label == 1 ? eL : eR];
      const int mk = rot->m[k], Dl = prev.Dl[q], Dr = prev.Dr[q];
      const double* Uk = rot->U.p + rot->uoff[k];
      const double* Psi = x + poff[q];
      double* Tq = T.p + toff[q];
      GemmSeg g{};
      g.abase = g.bbase = BASE_ABS;
      g.alpha = 1.0;
      if (label == 1) {  // T_q = U_k^T Psi_q  (m_k x Dr)
        g.a = Uk; g.lda = mk; g.flags = SEG_TA;
        g.b = Psi; g.ldb = Dr; g.K = Dl;
        p1.add_output(BASE_ABS, Tq, 0, Dr, mk, Dr, {g});
      } else {           // T_q = Psi_q U_k  (Dl x m_k)
        g.a = Psi; g.lda = Dr;
        g.b = Uk; g.ldb = mk; g.K = Dr;
        p1.add_output(BASE_ABS, Tq, 0, mk, Dl, mk, {g});
      }
      for (const Piece& pc : prev_far_pieces[label == 1 ? eR : eL]) {
        auto it = piece_at.find({k, pc.s});
        if (it == piece_at.end()) continue;
        const int c = it->second.first, o = it->second.second;
        const int q2 = q_of_near[c];
        if (q2 < 0 || sb_far[q2] != far->src[pc.a]) continue;
        const int mb = far->m[pc.a], Db = far->dim[pc.a];
        const double* Ub = far->U.p + far->uoff[pc.a];
        GemmSeg h{};
        h.abase = h.bbase = BASE_ABS;
        h.alpha = 1.0;
        h.K = mb;
        if (label == 1) {  // rows [o, o+m_k) of sector q2: T_q[:, cols of piece] U_b^T
          h.a = Tq + pc.off; h.lda = Dr;
          h.b = Ub; h.ldb = mb; h.flags = SEG_TB;
          p2.add_output(BASE_ABS, out + noff[q2] + (int64_t)o * lastDr[q2], 0, lastDr[q2], mk, Db, {h});
        } else {           // cols [o, o+m_k) of sector q2: U_a T_q[rows of piece, :]
          h.a = Ub; h.lda = mb;
          h.b = Tq + (int64_t)pc.off * mk; h.ldb = mk;
          p2.add_output(BASE_ABS, out + noff[q2] + o, 0, lastDr[q2], Db, mk, {h});
        }
        ++pieces_done;
      }
    }
    if (pieces_done == 0) return false;
    NSS_HIP(hipMemsetAsync(out, 0, n_new * sizeof(double), st));
    p1.upload(ctx->pool, st);
    p2.upload(ctx->pool, st);
    launch_gemm_f64(p1, Bases{nullptr, nullptr, nullptr}, st, GEMM_GENERIC);
    launch_gemm_f64(p2, Bases{nullptr, nullptr, nullptr}, st, GEMM_GENERIC);
    NSS_HIP(hipStreamSynchronize(st));
    return true;
  }

  // ------------------------------------------------------------ step
  QN target_of(int L) const {
    // sb_set_current_qn (src/DMRG_SUPERBLOCK.f90:717-726), Norb = 1:
    // QNtype=local int(target_qn*L), QNtype=global min(L, int(target_qn))
    auto one = [&](double q) { return par.qn_type == 1 ? std::min(L, (int)q) : (int)(q * L); };
    QN t;
    t.a = 2 * one(par.target_qn[0]);
    t.b = site.fermion ? 2 * one(par.target_qn[1]) : 0;
    return t;
  }

  // QNtruncation_error / QNtruncation_dim of eigh_blocks_matrix
  // (src/MATRIX/MATRIX_BLOCKS.f90:627-629): a QN block whose eigenvalues sum
  // below the error threshold, or whose dimension is below the dimension
  // threshold, gets tiny(1d0) for every eigenvalue
  void qn_truncate(std::vector<std::vector<double>>& ev) const {
    if (par.qn_trunc_error <= 0.0 && par.qn_trunc_dim <= 1) return;
    for (std::vector<double>& b : ev) {
      if (b.empty()) continue;
      double w = 0.0;
      for (double x : b) w += x;
      if (w < par.qn_trunc_error || (int)b.size() < par.qn_trunc_dim) std::fill(b.begin(), b.end(), DBL_MIN);
    }
  }

  // write_entanglement (src/DMRG_MAIN.f90:519-560): the QN blocks' weights
  // P_q and entropies S_q of the normalised spectra give S_N = -sum P_q ln P_q
  // and S_c = sum P_q S_q, with the same 1d-32 regularisation as S.  A block of
  // zero weight (the reference would divide by it) contributes nothing.
  static void entropy_split(const std::vector<std::vector<double>>& ev, double& sn, double& sc) {
    sn = sc = 0.0;
    for (const std::vector<double>& b : ev) {
      if (b.empty()) continue;
      double P = 0.0;
      for (double x : b) P += x;
      sn -= P * std::log(P + 1e-32);
      if (!(P > 0.0)) continue;
      double Sq = 0.0;
      for (double x : b) {
        const double o = x / P;
        Sq += o * std::log(o + 1e-32);
      }
      sc -= P * Sq;
    }
  }

  std::vector<std::pair<QN, std::vector<double>>> lambdas(const std::vector<std::vector<double>>& ev,
                                                          const Block& E, const std::vector<int>& sbsec) const {
    std::vector<std::pair<QN, std::vector<double>>> out;
    for (size_t q = 0; q < ev.size(); ++q)
      if (!ev[q].empty()) out.push_back({E.sec[sbsec[q]].q, ev[q]});
    return out;
  }

  // One DMRG step (step_dmrg, src/DMRG_MAIN.f90:265-460) in three phases:
  // begin_step (enlargement, superblock, start vector), the diagonalisation
  // (diag, or a batched Lanczos over several engines), end_step (rho,
  // truncation, renormalization, the step record).
  struct StepState {
    char type = 'i';
    int label = 0, sweep = 0, mstates = 0, neig = 1, v0 = 0;
    double estates = 0.0;
    bool dense = false;
    // false in finite steps whose far block (right for label 1, left for
    // label 2) the sweep replaces before it is used: that side's rho is then
    // not solved and its block not renormalized
    bool far = true;
    int64_t msb = 0;
    Step st;
    std::shared_ptr<Basis> fromL0, fromR0;
    std::chrono::steady_clock::time_point t0, t2, t3;
    LanczosOut lo;
  };

  static double secs(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  }

  StepState begin_step(char type, int label, int sweep, bool far = true, char link = 'n') {
    using clk = std::chrono::steady_clock;
    StepState S;
    S.t0 = clk::now();
    S.type = type;
    S.far = far || type != 'f' || (label != 1 && label != 2) || full_far_rho();
    S.label = label;
    S.sweep = sweep;
    S.mstates = par.mdmrg;
    S.estates = par.edmrg;
    if (type == 'f') {
      S.mstates = msweep.at(sweep);
      S.estates = esweep.empty() ? 0.0 : esweep.at(sweep);
    }
    nss_step_info& I = S.st.info;
    I.type = type;
    I.label = label;
    I.sweep = sweep + (type == 'f' ? 1 : 0);
    PhaseTimer pt(ctx, "step.enlarge");
    S.fromL0 = left.from;
    S.fromR0 = right.from;
    auto parL = std::make_shared<const Block>(left), parR = std::make_shared<const Block>(right);
    left = enlarge(left, true, link);
    right = enlarge(right, false, link);
    I.link = link;
    left.parent = parL;
    right.parent = parR;
    pt.next("step.setup");
    const auto t1 = clk::now();
    I.t_enlarge = secs(S.t0, t1);
    I.left_len = left.length;
    I.right_len = right.length;
    I.L = left.length + right.length;
    setup_superblock(left, right, target_of(I.L));
    S.msb = sb_n(sbd());
    I.m_sb = S.msb;
    I.nsect = sb_nq(sbd());
    I.nterms = (int)lastTerms.size();
    I.hv_flops = sb_flops(sbd());
    S.neig = (int)std::min<int64_t>(S.msb, par.neigen);
    I.neigen = S.neig;
    psi.resize(ctx->pool, (size_t)S.neig * S.msb);
    pt.next("step.predict");
    S.t2 = clk::now();
    S.dense = S.neig == S.msb || S.msb <= par.lanc_dim_threshold;
    if (!S.dense) {
      S.v0 = par.v0_mode;
      // one eigenpair only: with several, a start biased towards the previous
      // states can let a higher state pass the residual test before a missed
      // lower one surfaces, so those steps keep the random start
      if (type == 'f' && par.predict && S.neig == 1 && prev.valid && prev.label == label &&
          predict_psi(label, psi.p, S.msb))
        S.v0 = 2;
    }
    return S;
  }

  // The H*v precision of a step: mixed-precision engines solve superblocks
  // below lz_mixed_min_n() states in FP64 (their H*v is latency-bound, so
  // FP32 saves nothing, while the mixed solver's extra cycles cost more).
  int step_precision(const StepState& S) const {
    return par.precision == NSS_FP32_MIXED && S.msb >= lz_mixed_min_n() ? NSS_FP32_MIXED : NSS_FP64;
  }

  void diag(StepState& S) {
    PhaseTimer pt(ctx, "step.diag");
    nss_step_info& I = S.st.info;
    const int prec = step_precision(S);
    if (S.dense) {
      S.lo = dense_diag(ctx, sbd(), S.neig, psi.p);
      I.numop = 0;
    } else if (S.neig == 1 && prec == NSS_FP64 && S.msb >= lz_pro_min_n()) {
      // large superblocks: Lanczos with partial reorthogonalisation (a batch
      // of one chain; the thick-restart solver past kmax), where the full
      // Gram-Schmidt pass would cost more HBM traffic than H*v itself
      std::vector<LanczosJob> jobs(1);
      jobs[0].sb = sbd();
      jobs[0].xdev = psi.p;
      jobs[0].v0_mode = S.v0;
      jobs[0].seed = par.seed;
      lanczos_batch(ctx, jobs, par.lanc_kmax, par.lanc_maxiter, par.lanc_tol);
      S.lo = jobs[0].out;
      I.numop = S.lo.numop;
      I.numop_fp32 = 0;
    } else {
      S.lo = lanczos_device(ctx, sbd(), S.neig, par.lanc_kmax, par.lanc_maxiter, par.lanc_tol, par.seed, S.v0,
                            prec, psi.p);
      I.numop = S.lo.numop + S.lo.numop_fp32;
      I.numop_fp32 = S.lo.numop_fp32;
    }
  }

  // rho eigenpairs of a step (sb_get_rdm): both sides, vectors on the device
  struct RdmOut {
    std::vector<std::vector<double>> evs[2], unused[2];
    std::vector<RhoVec> dvs[2];
  };
  // sides of rho to solve (bit 0 left, bit 1 right)
  static int rdm_sides(const StepState& S) { return S.far ? 3 : (S.label == 1 ? 1 : 2); }
  RdmJob rdm_job(RdmOut& o, const StepState& S) {
    RdmJob J;
    J.sb = sbd();
    J.psi = psi.p;
    J.sides = rdm_sides(S);
    J.evals = o.evs;
    J.evecs = o.unused;
    J.dvecs = o.dvs;
    J.defer = true;
    return J;
  }

  void end_step(StepState& S) {
    S.t3 = std::chrono::steady_clock::now();
    RdmOut o;
    {
      PhaseTimer pt(ctx, "step.rdm");
      sb_rdm_eigh2(ctx, sbd(), psi.p, rdm_sides(S), o.evs, o.unused, o.dvs, true);
    }
    finish_step(S, o);
  }

  // everything after the rho eigensolve; S.t3 marks the end of the diagonalisation
  void finish_step(StepState& S, RdmOut& o) {
    using clk = std::chrono::steady_clock;
    const char type = S.type;
    const int label = S.label, neig = S.neig;
    const int64_t msb = S.msb;
    Step& st = S.st;
    nss_step_info& I = st.info;
    I.t_diag = secs(S.t2, S.t3);
    for (int i = 0; i < neig && i < 4; ++i) I.energies[i] = S.lo.evals[i];
    PhaseTimer pt(ctx, "step.spectrum");
    const int farside = S.far ? -1 : (label == 1 ? 1 : 0);
    if (farside >= 0) o.evs[farside] = far_spectrum(o.evs[1 - farside], farside);
    qn_truncate(o.evs[0]);
    qn_truncate(o.evs[1]);
    entropy_split(o.evs[0], I.sn_left, I.sc_left);
    entropy_split(o.evs[1], I.sn_right, I.sc_right);
    st.lam_left = lambdas(o.evs[0], left, sbL);
    st.lam_right = lambdas(o.evs[1], right, sbR);
    Spectrum sl = sort_spectrum(std::move(o.evs[0]), std::move(o.dvs[0]));
    Spectrum sr = sort_spectrum(std::move(o.evs[1]), std::move(o.dvs[1]));
    const auto t4 = clk::now();
    I.t_rdm = secs(S.t3, t4);
    I.entropy_left = entropy(sl.all_desc);
    I.entropy_right = entropy(sr.all_desc);
    bool ren;
    if (type == 'i') ren = left.length != par.ldmrg;
    else ren = !((left.length == par.ldmrg + 1) && (label == 1));
    I.renormalized = ren;
    I.m_left = left.dim();
    I.m_right = right.dim();
    lastL = left;
    lastR = right;
    pt.next("step.renorm");
    if (ren) {
      int ml = 0, mr = 0;
      double tl = 0.0, tr = 0.0;
      if (farside == 1) {  // right block replaced by the sweep: selection only
        Block nr;
        const Selection sel = select_states(right, sbR, sr, S.mstates, S.estates, nr, st.kept_right);
        mr = sel.ms;
        tr = sel.trunc;
        I.cut_right = sel.cut;
        right.from = identity_basis(right);
      } else {
        Block nr = renormalize(right, sbR, sr, S.mstates, S.estates, mr, tr, I.cut_right, st.kept_right);
        nr.parent = right.parent;
        right = std::move(nr);
      }
      if (farside == 0) {  // left block replaced by the sweep
        Block nl;
        const Selection sel = select_states(left, sbL, sl, S.mstates, S.estates, nl, st.kept_left);
        ml = sel.ms;
        tl = sel.trunc;
        I.cut_left = sel.cut;
        left.from = identity_basis(left);
      } else {
        Block nl = renormalize(left, sbL, sl, S.mstates, S.estates, ml, tl, I.cut_left, st.kept_left);
        nl.parent = left.parent;
        left = std::move(nl);
      }
      I.m_left = ml;
      I.m_right = mr;
      I.trunc_left = tl;
      I.trunc_right = tr;
    } else {
      left.from = identity_basis(left);
      right.from = identity_basis(right);
    }
    prev.valid = type == 'f' && par.predict;
    if (prev.valid) {
      prev.label = label;
      prev.neig = neig;
      prev.n = msb;
      prev.sbL = sbL;
      prev.sbR = sbR;
      prev.Dl = lastDl;
      prev.Dr = lastDr;
      prev.piecesL = lastL.pieces;
      prev.piecesR = lastR.pieces;
      prev.unenlL = S.fromL0;
      prev.unenlR = S.fromR0;
      prev.newL = left.from;
      prev.newR = right.from;
      prev.psi.resize(ctx->pool, (size_t)neig * msb);
      NSS_HIP(hipMemcpyAsync(prev.psi.p, psi.p, (size_t)neig * msb * sizeof(double), hipMemcpyDeviceToDevice,
                             ctx->stream));
    }
    pt.next("step.tail");
    const auto t5 = clk::now();
    I.t_renorm = secs(t4, t5);
    I.t_hv = nss_hv_timer_flush(ctx);
    I.t_step = secs(S.t0, t5);
    if (par.verbose > 0) {
      std::printf("step %c%d len %d+%d  m_sb %lld nsect %d  E/L %.15f  S %.12f  m %d,%d  nop %d (fp32 %d)  t %.3fs\n",
                  type, label, I.left_len, I.right_len, (long long)msb, I.nsect, I.energies[0] / I.L, I.entropy_left,
                  I.m_left, I.m_right, I.numop, I.numop_fp32, I.t_step);
      std::fflush(stdout);
    }
    steps.push_back(std::move(st));
  }

  void step(char type, int label, int sweep, bool far = true, char link = 'n') {
    StepState S = begin_step(type, label, sweep, far, link);
    diag(S);
    end_step(S);
  }

  // infinite_DMRG, src/DMRG_MAIN.f90:140-159: with PBC every iteration is a
  // link-'n' step followed by a link-'p' step.
  void infinite() {
    bl1.clear();
    bl2.clear();
    bl1[left.length] = left;
    bl2[right.length] = right;
    while (left.length < par.ldmrg) {
      step('i', 0, 0);
      bl1[left.length] = left;
      bl2[right.length] = right;
      if (!par.pbc) continue;
      step('i', 0, 0, true, 'p');
      bl1[left.length] = left;
      bl2[right.length] = right;
    }
  }

  // finite_DMRG, src/DMRG_MAIN.f90:168-241, split into the iDMRG growth and
  // one call per sweep.  bl1 / bl2 hold the left / right blocks by length:
  // after a run, bl?[k+1] is the renormalization of the enlarged bl?[k], the
  // chain that Measure_DMRG walks.
  bool finite_ready = false;

  void finite_begin() {
    NSS_REQUIRE(!par.pbc, NSS_E_ARG, "finite_DMRG ERROR: Finite DMRG with PBC not yet implemented.");
    bl1.clear();
    bl2.clear();
    bl1[1] = left;
    bl2[1] = right;
    while (left.length < par.ldmrg) {
      step('i', 0, 0);
      bl1[left.length] = left;
      bl2[right.length] = right;
    }
    finite_ready = true;
    fs_im = fs_loop = 0;
  }

  // Position in the finite algorithm: sweep fs_im, loop fs_loop of its three
  // (left->right to the right end, right->left to the left end, left->right
  // back to the middle).
  int fs_im = 0, fs_loop = 0;

  // The next step of the finite sweeps (src/DMRG_MAIN.f90:214-239); false
  // once sweep `last` (or the last of MSWEEP) is complete.  The far block of a
  // step is needed only by the last step of each loop (the next loop
  // enlarges it; after the last it is the final state).
  bool finite_step(int last) {
    NSS_REQUIRE(finite_ready, NSS_E_STATE, "finite sweep before nss_dmrg_finite_begin");
    const int fn = 2 * par.ldmrg;
    for (;;) {
      if (fs_im > last || fs_im >= (int)msweep.size()) return false;
      if (fs_loop == 0 && left.length < fn) {
        right = bl2.at(fn - left.length);
        step('f', 1, fs_im, left.length + 1 >= fn);
        bl1[left.length] = left;
        return true;
      }
      if (fs_loop == 0) fs_loop = 1;
      if (fs_loop == 1 && right.length < fn) {
        left = bl1.at(fn - right.length);
        step('f', 2, fs_im, right.length + 1 >= fn);
        bl2[right.length] = right;
        return true;
      }
      if (fs_loop == 1) fs_loop = 2;
      if (left.length <= par.ldmrg) {
        right = bl2.at(fn - left.length);
        step('f', 1, fs_im, left.length + 1 > par.ldmrg);
        bl1[left.length] = left;
        return true;
      }
      fs_loop = 0;
      ++fs_im;
    }
  }

  void sweep(int im) {
    NSS_REQUIRE(finite_ready, NSS_E_STATE, "finite sweep before nss_dmrg_finite_begin");
    NSS_REQUIRE(im >= 0 && im < (int)msweep.size(), NSS_E_ARG, "sweep index outside MSWEEP");
    NSS_REQUIRE(im == fs_im && fs_loop == 0, NSS_E_STATE, "finite sweeps run in order, each one whole");
    while (finite_step(im)) {
    }
  }

  void finite() {
    finite_begin();
    for (int im = 0; im < (int)msweep.size(); ++im) sweep(im);
  }
};

// Independent chains solved together.  The engines' steps run in lockstep
// (equal LDMRG gives equal step sequences): every engine enlarges and sets
// up its superblock, one batched Lanczos solves all ground states (the H*v
// of all chains as one grouped launch per phase), then every engine
// truncates.  Steps on the dense path, with LANC_NEIGEN > 1 or in mixed
// precision (superblocks from lz_mixed_min_n() states) use the engine's own
// solver inside the same lockstep.
struct Batch {
  nss_ctx* ctx = nullptr;
  std::vector<Engine*> es;

  void step(char type, int label, int sweep, bool far = true, char link = 'n') {
    std::vector<Engine::StepState> S;
    S.reserve(es.size());
    for (Engine* e : es) S.push_back(e->begin_step(type, label, sweep, far, link));
    std::vector<LanczosJob> jobs;
    std::vector<size_t> who;
    for (size_t i = 0; i < es.size(); ++i) {
      Engine* e = es[i];
      if (S[i].dense || S[i].neig != 1 || e->step_precision(S[i]) != NSS_FP64) {
        e->diag(S[i]);
        continue;
      }
      LanczosJob J;
      J.sb = e->sbd();
      J.xdev = e->psi.p;
      J.v0_mode = S[i].v0;
      J.seed = e->par.seed;
      jobs.push_back(J);
      who.push_back(i);
    }
    {
      PhaseTimer pt(ctx, "step.diag");
      const nss_params& p0 = es[0]->par;
      lanczos_batch(ctx, jobs, p0.lanc_kmax, p0.lanc_maxiter, p0.lanc_tol);
    }
    for (size_t k = 0; k < jobs.size(); ++k) {
      Engine::StepState& St = S[who[k]];
      St.lo = jobs[k].out;
      St.st.info.numop = jobs[k].out.numop;
    }
    // rho of every chain in one batch; the eigenvector views live in rho
    // until every chain has renormalized
    const auto t3 = std::chrono::steady_clock::now();
    std::vector<Engine::RdmOut> ro(es.size());
    std::vector<RdmJob> rj;
    for (size_t i = 0; i < es.size(); ++i) {
      S[i].t3 = t3;
      rj.push_back(es[i]->rdm_job(ro[i], S[i]));
    }
    DevArr<double> rho;
    GemmPlan R;
    {
      PhaseTimer pt(ctx, "step.rdm");
      sb_rdm_eigh_multi(ctx, rj, rho, R);
    }
    for (size_t i = 0; i < es.size(); ++i) es[i]->finish_step(S[i], ro[i]);
  }

  void finite_begin() {
    NSS_REQUIRE(!es[0]->par.pbc, NSS_E_ARG, "finite_DMRG ERROR: Finite DMRG with PBC not yet implemented.");
    for (Engine* e : es) {
      e->bl1.clear();
      e->bl2.clear();
      e->bl1[1] = e->left;
      e->bl2[1] = e->right;
    }
    while (es[0]->left.length < es[0]->par.ldmrg) {
      step('i', 0, 0);
      for (Engine* e : es) {
        e->bl1[e->left.length] = e->left;
        e->bl2[e->right.length] = e->right;
      }
    }
    for (Engine* e : es) e->finite_ready = true;
  }

  void infinite() {
    for (Engine* e : es) {
      e->bl1.clear();
      e->bl2.clear();
      e->bl1[e->left.length] = e->left;
      e->bl2[e->right.length] = e->right;
    }
    while (es[0]->left.length < es[0]->par.ldmrg) {
      for (char link : es[0]->par.pbc ? std::string("np") : std::string("n")) {
        step('i', 0, 0, true, link);
        for (Engine* e : es) {
          e->bl1[e->left.length] = e->left;
          e->bl2[e->right.length] = e->right;
        }
      }
    }
  }

  // the loops of Engine::sweep, applied to every engine
  void sweep(int im) {
    for (Engine* e : es) {
      NSS_REQUIRE(e->finite_ready, NSS_E_STATE, "finite sweep before nss_dmrg_batch_finite_begin");
      NSS_REQUIRE(im >= 0 && im < (int)e->msweep.size(), NSS_E_ARG, "sweep index outside MSWEEP");
    }
    Engine* e0 = es[0];
    const int fn = 2 * e0->par.ldmrg;
    while (e0->left.length < fn) {
      for (Engine* e : es) e->right = e->bl2.at(fn - e->left.length);
      step('f', 1, im, e0->left.length + 1 >= fn);
      for (Engine* e : es) e->bl1[e->left.length] = e->left;
    }
    while (e0->right.length < fn) {
      for (Engine* e : es) e->left = e->bl1.at(fn - e->right.length);
      step('f', 2, im, e0->right.length + 1 >= fn);
      for (Engine* e : es) e->bl2[e->right.length] = e->right;
    }
    while (e0->left.length <= e0->par.ldmrg) {
      for (Engine* e : es) e->right = e->bl2.at(fn - e->left.length);
      step('f', 1, im, e0->left.length + 1 > e0->par.ldmrg);
      for (Engine* e : es) e->bl1[e->left.length] = e->left;
    }
  }
};

}  // namespace nss

struct nss_dmrg {
  nss::Engine* e = nullptr;
};

using namespace nss;

#define NSS_API_BEGIN try {
#define NSS_API_END                  \
  return NSS_OK;                     \
  }                                  \
  catch (const nss::Error& e) {      \
    nss::set_last_error(e.what());   \
    return e.code;                   \
  }                                  \
  catch (const std::exception& e) {  \
    nss::set_last_error(e.what());   \
    return NSS_E_ARG;                \
  }

extern "C" int nss_dmrg_create(nss_ctx* ctx, const nss_model* model, const nss_params* params, nss_dmrg** out) {
  NSS_API_BEGIN
  NSS_REQUIRE(ctx && model && params && out, NSS_E_ARG, "nss_dmrg_create: NULL arguments");
  NSS_REQUIRE(params->ldmrg >= 1, NSS_E_ARG, "LDMRG must be >= 1");
  NSS_REQUIRE(params->neigen >= 1 && params->neigen <= 4, NSS_E_ARG, "LANC_NEIGEN must be in 1..4");
  NSS_REQUIRE(params->mdmrg > 0 || params->edmrg != 0.0, NSS_E_ARG,
              "Mdmrg=Edmrg=0 can not fix a threshold for the RDM");
  NSS_REQUIRE(params->precision == NSS_FP64 || params->precision == NSS_FP32_MIXED, NSS_E_ARG, "bad precision");
  NSS_HIP(hipSetDevice(ctx->device));
  auto* E = new Engine();
  E->ctx = ctx;
  E->model = *model;
  E->par = *params;
  if (params->nsweep > 0) {
    NSS_REQUIRE(params->msweep != nullptr, NSS_E_ARG, "MSWEEP required when NSWEEP > 0");
    E->msweep.assign(params->msweep, params->msweep + params->nsweep);
    if (params->esweep) E->esweep.assign(params->esweep, params->esweep + params->nsweep);
  }
  E->hij.assign(model->hij, model->hij + 4);
  E->want32 = params->precision == NSS_FP32_MIXED;
  if (E->par.pbc && E->par.ldmrg % 2 == 0) ++E->par.ldmrg;  // Ldmrg = 2k+1 under PBC, src/INPUT_VARS.f90:178-183
  try {
    E->site = make_site(*model);
    if (E->par.pbc) {
      // the operators of the outer ends (PBC_prev), src/BLOCKS/SITES.f90:339-342, :413-424
      std::vector<std::string> keys;
      for (const auto& kv : E->site.ops)
        if (kv.first != "H") keys.push_back(kv.first);
      for (const std::string& k : keys) E->site.ops[k + "_p"] = E->site.ops[k];
    }
    E->left = E->block_from_site();
    E->right = E->block_from_site();
  } catch (...) {
    delete E;
    throw;
  }
  auto* d = new nss_dmrg();
  d->e = E;
  ctx->engines.push_back(d);
  *out = d;
  NSS_API_END
}

// Called by nss_ctx_destroy: engines still alive lose their state (their
// handles stay valid shells that nss_dmrg_destroy frees).
void nss::engines_release(nss_ctx* ctx) {
  for (nss_dmrg* d : ctx->engines) {
    delete d->e;
    d->e = nullptr;
  }
  ctx->engines.clear();
}

extern "C" int nss_dmrg_run(nss_dmrg* d, int type) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e, NSS_E_ARG, "nss_dmrg_run: NULL handle");
  NSS_HIP(hipSetDevice(d->e->ctx->device));
  if (type == 'i' || type == 'I') d->e->infinite();
  else if (type == 'f' || type == 'F') d->e->finite();
  else throw Error(NSS_E_ARG, "ERROR DMRG: unsupported DMRGtype: DMRGtype !=['i','f']");
  NSS_API_END
}

extern "C" int nss_dmrg_finite_begin(nss_dmrg* d) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e, NSS_E_ARG, "nss_dmrg_finite_begin: NULL handle");
  NSS_HIP(hipSetDevice(d->e->ctx->device));
  d->e->finite_begin();
  NSS_API_END
}

extern "C" int nss_dmrg_finite_step(nss_dmrg* d, int* done) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e && done, NSS_E_ARG, "nss_dmrg_finite_step: NULL arguments");
  NSS_HIP(hipSetDevice(d->e->ctx->device));
  *done = d->e->finite_step(INT32_MAX) ? 0 : 1;
  NSS_API_END
}

extern "C" int nss_dmrg_sweep(nss_dmrg* d, int im) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e, NSS_E_ARG, "nss_dmrg_sweep: NULL handle");
  NSS_HIP(hipSetDevice(d->e->ctx->device));
  d->e->sweep(im);
  NSS_API_END
}

// Batch entry points: the engines share one context and LDMRG; each gets a
// superblock of its own (the context's stays free for the low-level calls).
static Batch make_batch(int n, nss_dmrg* const* ds, bool finite) {
  NSS_REQUIRE(n >= 1 && ds, NSS_E_ARG, "nss_dmrg_batch: need at least one engine");
  Batch B;
  for (int i = 0; i < n; ++i) {
    NSS_REQUIRE(ds[i] && ds[i]->e, NSS_E_ARG, "nss_dmrg_batch: NULL engine");
    Engine* e = ds[i]->e;
    if (!B.ctx) B.ctx = e->ctx;
    NSS_REQUIRE(e->ctx == B.ctx, NSS_E_ARG, "nss_dmrg_batch: engines on different contexts");
    NSS_REQUIRE(e->par.ldmrg == ds[0]->e->par.ldmrg, NSS_E_ARG, "nss_dmrg_batch: engines with different LDMRG");
    NSS_REQUIRE(!e->par.pbc == !ds[0]->e->par.pbc, NSS_E_ARG, "nss_dmrg_batch: engines with different PBCdmrg");
    if (finite)
      NSS_REQUIRE(e->msweep.size() == ds[0]->e->msweep.size(), NSS_E_ARG, "nss_dmrg_batch: different NSWEEP");
    for (Engine* o : B.es) NSS_REQUIRE(o != e, NSS_E_ARG, "nss_dmrg_batch: an engine appears twice");
    if (!e->own_sb) e->own_sb = sb_create(e->ctx);
    B.es.push_back(e);
  }
  NSS_HIP(hipSetDevice(B.ctx->device));
  return B;
}

extern "C" int nss_dmrg_batch_run(int n, nss_dmrg* const* ds, int type) {
  NSS_API_BEGIN
  const bool fin = type == 'f' || type == 'F';
  NSS_REQUIRE(fin || type == 'i' || type == 'I', NSS_E_ARG, "ERROR DMRG: unsupported DMRGtype: DMRGtype !=['i','f']");
  Batch B = make_batch(n, ds, fin);
  if (!fin) {
    B.infinite();
  } else {
    B.finite_begin();
    for (int im = 0; im < (int)B.es[0]->msweep.size(); ++im) B.sweep(im);
  }
  NSS_API_END
}

extern "C" int nss_dmrg_batch_finite_begin(int n, nss_dmrg* const* ds) {
  NSS_API_BEGIN
  Batch B = make_batch(n, ds, true);
  B.finite_begin();
  NSS_API_END
}

extern "C" int nss_dmrg_batch_sweep(int n, nss_dmrg* const* ds, int im) {
  NSS_API_BEGIN
  Batch B = make_batch(n, ds, true);
  B.sweep(im);
  NSS_API_END
}

extern "C" int nss_dmrg_nsteps(nss_dmrg* d) { return (d && d->e) ? (int)d->e->steps.size() : -1; }

extern "C" int nss_dmrg_step_info(nss_dmrg* d, int i, nss_step_info* out) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e && out, NSS_E_ARG, "nss_dmrg_step_info: NULL arguments");
  NSS_REQUIRE(i >= 0 && i < (int)d->e->steps.size(), NSS_E_ARG, "nss_dmrg_step_info: step out of range");
  *out = d->e->steps[i].info;
  NSS_API_END
}

extern "C" int nss_dmrg_kept_qn(nss_dmrg* d, int i, int side, double min_weight, int max, double* qn,
                                int32_t* count, int32_t* nout) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e && nout, NSS_E_ARG, "nss_dmrg_kept_qn: NULL arguments");
  NSS_REQUIRE(i >= 0 && i < (int)d->e->steps.size(), NSS_E_ARG, "nss_dmrg_kept_qn: step out of range");
  const auto& k = side == 0 ? d->e->steps[i].kept_left : d->e->steps[i].kept_right;
  *nout = (int32_t)k.size();
  for (int j = 0; j < (int)k.size() && j < max; ++j) {
    if (qn) {
      qn[2 * j] = 0.5 * k[j].first.a;
      qn[2 * j + 1] = 0.5 * k[j].first.b;
    }
    if (count) {
      int c = 0;
      for (double l : k[j].second) c += (l > min_weight) ? 1 : 0;
      count[j] = c;
    }
  }
  NSS_API_END
}

extern "C" int nss_dmrg_lambda(nss_dmrg* d, int i, int side, int maxb, double* qn, int32_t* count, int64_t maxv,
                               double* vals, int32_t* nblocks, int64_t* nvals) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e && nblocks && nvals, NSS_E_ARG, "nss_dmrg_lambda: NULL arguments");
  NSS_REQUIRE(i >= 0 && i < (int)d->e->steps.size(), NSS_E_ARG, "nss_dmrg_lambda: step out of range");
  NSS_REQUIRE(side == 0 || side == 1, NSS_E_ARG, "nss_dmrg_lambda: side must be 0 (left) or 1 (right)");
  const auto& L = side == 0 ? d->e->steps[i].lam_left : d->e->steps[i].lam_right;
  *nblocks = (int32_t)L.size();
  int64_t tot = 0, w = 0;
  for (int j = 0; j < (int)L.size(); ++j) {
    const auto& b = L[j];
    if (j < maxb) {
      if (qn) {
        qn[2 * j] = 0.5 * b.first.a;
        qn[2 * j + 1] = 0.5 * b.first.b;
      }
      if (count) count[j] = (int32_t)b.second.size();
    }
    for (double x : b.second) {
      if (vals && w < maxv) vals[w++] = x;
      ++tot;
    }
  }
  *nvals = tot;
  NSS_API_END
}

extern "C" int nss_dmrg_last_sb_dims(nss_dmrg* d, int* nq, int* nterms, int64_t* n) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e, NSS_E_ARG, "NULL handle");
  int64_t nn = 0;
  for (size_t q = 0; q < d->e->lastDl.size(); ++q) nn += (int64_t)d->e->lastDl[q] * d->e->lastDr[q];
  if (nq) *nq = (int)d->e->lastDl.size();
  if (nterms) *nterms = (int)d->e->lastTerms.size();
  if (n) *n = nn;
  NSS_API_END
}

// Dense row-major copies: HL, HR concatenated per sector; A, B per term
// (A_t: Dl_dst x Dl_src, B_t: Dr_dst x Dr_src with Hv_dst += A Psi_src B^T).
extern "C" int nss_dmrg_last_sb_export(nss_dmrg* d, int32_t* Dl, int32_t* Dr, int32_t* dst, int32_t* src,
                                       double* HL, double* HR, double* A, double* B, double* psi) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e, NSS_E_ARG, "NULL handle");
  Engine* E = d->e;
  const auto& dl = E->lastDl;
  const auto& dr = E->lastDr;
  const int nq = (int)dl.size();
  int64_t nn = 0;
  for (int q = 0; q < nq; ++q) nn += (int64_t)dl[q] * dr[q];
  hipStream_t s = E->ctx->stream;
  auto fetch = [&](const OpRef& r, int rows, int cols, double* outp) {
    // r describes an operand (rows x cols) with ld/trans; alpha applied
    std::vector<double> raw((size_t)rows * cols);
    const double* p = static_cast<const double*>(r.p);
    const int st_rows = r.trans ? cols : rows, st_cols = r.ld;
    std::vector<double> stor((size_t)st_rows * st_cols);
    NSS_HIP(hipMemcpyAsync(stor.data(), p, stor.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    NSS_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j)
        outp[(size_t)i * cols + j] = r.alpha * (r.trans ? stor[(size_t)j * r.ld + i] : stor[(size_t)i * r.ld + j]);
  };
  if (Dl) for (int q = 0; q < nq; ++q) Dl[q] = dl[q];
  if (Dr) for (int q = 0; q < nq; ++q) Dr[q] = dr[q];
  size_t oh = 0, orr = 0;
  for (int q = 0; q < nq; ++q) {
    const BlockOp& hl = *E->lastL.ops.at("H");
    const BlockOp& hr = *E->lastR.ops.at("H");
    OpRef a, b;
    a.p = hl.data.p + hl.row(E->sbL[q])->off; a.ld = dl[q];
    b.p = hr.data.p + hr.row(E->sbR[q])->off; b.ld = dr[q];
    if (HL) fetch(a, dl[q], dl[q], HL + oh);
    if (HR) fetch(b, dr[q], dr[q], HR + orr);
    oh += (size_t)dl[q] * dl[q];
    orr += (size_t)dr[q] * dr[q];
  }
  size_t oa = 0, ob = 0;
  for (size_t t = 0; t < E->lastTerms.size(); ++t) {
    const SbTerm& T = E->lastTerms[t];
    if (dst) dst[t] = T.dst;
    if (src) src[t] = T.src;
    if (A) fetch(T.A, dl[T.dst], dl[T.src], A + oa);
    if (B) {  // B = (B^T)^T: fetch B^T (Dr_src x Dr_dst) then transpose
      std::vector<double> bt((size_t)dr[T.src] * dr[T.dst]);
      fetch(T.Bt, dr[T.src], dr[T.dst], bt.data());
      for (int i = 0; i < dr[T.dst]; ++i)
        for (int j = 0; j < dr[T.src]; ++j) B[ob + (size_t)i * dr[T.src] + j] = bt[(size_t)j * dr[T.dst] + i];
    }
    oa += (size_t)dl[T.dst] * dl[T.src];
    ob += (size_t)dr[T.dst] * dr[T.src];
  }
  if (psi) {
    NSS_HIP(hipMemcpyAsync(psi, E->psi.p, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    NSS_HIP(hipStreamSynchronize(s));
  }
  NSS_API_END
}

extern "C" int nss_dmrg_load_blocks(nss_dmrg* d, const char* left_path, const char* right_path) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e && left_path && right_path, NSS_E_ARG, "nss_dmrg_load_blocks: bad arguments");
  nss::Engine* E = d->e;
  NSS_HIP(hipSetDevice(E->ctx->device));
  nss::Block l = E->load_block(left_path, true), r = E->load_block(right_path, false);
  NSS_REQUIRE(l.length == r.length, NSS_E_ARG, "infinite_DMRG error: L.length != R.length");
  E->left = std::move(l);
  E->right = std::move(r);
  NSS_API_END
}

extern "C" int nss_dmrg_save_blocks(nss_dmrg* d, const char* left_path, const char* right_path) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e && left_path && right_path, NSS_E_ARG, "nss_dmrg_save_blocks: bad arguments");
  nss::Engine* E = d->e;
  NSS_REQUIRE(E->lastL.parent && E->lastR.parent, NSS_E_STATE, "nss_dmrg_save_blocks: run a step first");
  NSS_HIP(hipSetDevice(E->ctx->device));
  E->save_block(*E->lastL.parent, left_path, true);
  E->save_block(*E->lastR.parent, right_path, false);
  NSS_API_END
}

extern "C" int nss_dmrg_measure(nss_dmrg* d, int nops, const double* ops, int npos, const int32_t* pos,
                                double* out) {
  NSS_API_BEGIN
  NSS_REQUIRE(d && d->e && ops && pos && out && nops > 0 && npos > 0, NSS_E_ARG, "nss_dmrg_measure: bad arguments");
  nss::Engine* E = d->e;
  // Measure_DMRG under PBC walks the ring's own site order
  // (src/DMRG_MEASURE.f90:320-360), which the engine does not keep
  NSS_REQUIRE(!E->par.pbc, NSS_E_ARG, "nss_dmrg_measure: not available for PBC runs");
  NSS_HIP(hipSetDevice(E->ctx->device));
  const int dd = E->site.dim;
  for (int o = 0; o < nops; ++o) {
    std::vector<double> O(ops + (size_t)o * dd * dd, ops + (size_t)(o + 1) * dd * dd);
    for (int p = 0; p < npos; ++p) out[(size_t)o * npos + p] = E->measure_one(O, pos[p]);
  }
  NSS_API_END
}

extern "C" int nss_dmrg_destroy(nss_dmrg* d) {
  NSS_API_BEGIN
  if (!d) return NSS_OK;
  if (d->e) {
    nss_ctx* c = d->e->ctx;
    // the context's superblock plan points into this engine's blocks
    sb_reset(c->sb);
    c->engines.erase(std::remove(c->engines.begin(), c->engines.end(), d), c->engines.end());
    delete d->e;
    // an engine's blocks can hold a large share of the device: hand the
    // pool's cache back once no queued work can use it
    NSS_HIP(hipStreamSynchronize(c->stream));
    c->pool->trim();
  }
  delete d;
  NSS_API_END
}

"""CPU oracle for the nssDMRG superblock path — TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  The product path (``nssdmrg_amd``) never imports anything from here and
fails loudly if its HIP library is missing.

It restates, in plain numpy and in the reference's own basis ordering, the
algorithm of aamaricci/nssDMRG for one DMRG step and the infinite/finite
drivers around it.  Each function cites the reference file:line it follows.

Third-party arithmetic the reference delegates to SciFortran (not vendored,
version unpinned: CI clones HEAD, ``.github/workflows/PushWorkflow.yml:38-41``):
  * ``sp_eigh`` (ARPACK implicitly restarted Lanczos) is restated as a plain
    Lanczos with full re-orthogonalisation (``lanczos_lowest``); tests cross
    check it against scipy's ARPACK (``eigsh``) and LAPACK (``eigh``);
  * ``eigh`` (LAPACK dsyev) is ``numpy.linalg.eigh`` (LAPACK dsyevd);
  * ``sort_quicksort`` tie order is not reproducible (parity unpinned for the
    order of exactly tied rho eigenvalues; kept counts are tie independent
    because the degeneracy window always keeps every tied value).

Pinning (tests/test_oracle_*.py): the reference's golden ``.check`` values
(spin-1/2 Heisenberg L40 M20 and Hubbard U=0 L40 M40 trajectories), closed
forms (4-site Heisenberg, free fermions), brute-force exact diagonalisation of
full chains, and invariants (symmetry of H_sb, trace of rho, lambda_L=lambda_R).
The PBC path (link-'p' steps) is pinned by periodic-ring closed forms
(Heisenberg rings of 4 and 6 sites, free-fermion rings), a Fock-space ED of
the Hubbard ring, a brute-force spin-1 ring, the fermion algebra of both
block ends and the symmetry of every PBC H_sb.  The entanglement split
(``entropy_split``, ``lambda_lists``) is pinned by the exact reduced density
matrix of an ED ground state built from explicit Kronecker products and by
S = S_N + S_c; QNtype=global by ED of the targeted sector; the QN-sector
truncation knobs by their limits (every eigenvalue tiny(1d0) above every
sector's weight or dimension, nothing changed below).  Under PBC the spin terms are
the Hermitian ones of connect_spin_blocks(link='p'); the reference's direct
spin setup differs there (see setup_direct).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TINY = np.finfo(np.float64).tiny   # Fortran tiny(1d0)
EPS_LOG = 1e-32                    # src/DMRG_MAIN.f90:509 (epsi)

# ---------------------------------------------------------------------------
# Seeded RNG: SplitMix64 -> uniform[-1,1).  Same stream as the HIP library's
# start vector (k_fill_splitmix in nssdmrg_amd/csrc/nss_kernels.hip).
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64_uniform(n: int, seed: int) -> np.ndarray:
    """n doubles uniform in [-1,1) from SplitMix64(seed) (53-bit mantissa)."""
    out = np.empty(n, dtype=np.float64)
    s = seed & _M64
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (2.0 ** -53) * 2.0 - 1.0
    return out


def splitmix64_uniform_fast(n: int, seed: int) -> np.ndarray:
    """Vectorised SplitMix64 (identical stream to splitmix64_uniform)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        s = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


# ---------------------------------------------------------------------------
# Sites (src/BLOCKS/SITES.f90, src/FOCK/HLOCAL.f90)
# ---------------------------------------------------------------------------
@dataclass
class Site:
    kind: str                       # "spin" | "fermion"
    dim: int
    qn: np.ndarray                  # (dim, qdim) float
    ops: dict                       # name -> dense matrix
    optype: dict                    # name -> "bosonic" | "fermionic" | "psign"


def spin_site(sun: int = 2, hvec=(0.0, 0.0)) -> Site:
    """spin_site, src/BLOCKS/SITES.f90:281-344 (SU(2) and SU(3) presets)."""
    if sun == 2:
        sz = np.diag([0.5, -0.5])
        sp = np.array([[0.0, 1.0], [0.0, 0.0]])           # Sp(1,2)=1, :318
        px = np.array([[0.0, 1.0], [1.0, 0.0]])
        pz = np.array([[1.0, 0.0], [0.0, -1.0]])
        h = hvec[0] * px + hvec[1] * pz
        qn = np.array([[0.5], [-0.5]])
    elif sun == 3:
        sz = np.diag([1.0, 0.0, -1.0])
        sp = np.zeros((3, 3))
        sp[0, 1] = math.sqrt(2.0)
        sp[1, 2] = math.sqrt(2.0)
        sx = np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]], float) / math.sqrt(2.0)
        h = hvec[0] * sx + hvec[1] * sz
        qn = np.array([[1.0], [0.0], [-1.0]])
    else:
        raise ValueError("spin_site: SU(N) value not supported")
    return Site("spin", sz.shape[0], qn,
                {"H": h, "Sz": sz, "Sp": sp},
                {"H": "bosonic", "Sz": "bosonic", "Sp": "bosonic"})


def electron_site(uloc: float = 0.0, hfmode: bool = True, eloc=(0.0, 0.0)) -> Site:
    """electron_site for Norb=1, src/BLOCKS/SITES.f90:358-434.

    Fock basis m = iup + 2*idw, |up,dw> with up fastest
    (src/FOCK/HLOCAL.f90:735-758); QN [Nup,Ndw].  Local H from
    build_Hlocal_operator :312-581 (crystal field, U n_up n_dw, Hartree shift
    -U/2 (n_up+n_dw) when HFMODE=T :423-435).  C operators carry the
    Jordan-Wigner sign of src/FOCK/HLOCAL.f90:884-896.
    """
    dim = 4
    qn = np.zeros((dim, 2))
    h = np.zeros((dim, dim))
    cup = np.zeros((dim, dim))
    cdn = np.zeros((dim, dim))
    p = np.zeros((dim, dim))
    for m in range(dim):
        nup, ndw = m & 1, (m >> 1) & 1
        qn[m] = (nup, ndw)
        e = eloc[0] * nup + eloc[1] * ndw + uloc * nup * ndw
        if hfmode:
            e -= 0.5 * uloc * (nup + ndw)
        h[m, m] = e
        p[m, m] = (-1.0) ** (nup + ndw)
        if nup:                                  # c_up: no bit below -> +1
            cup[m & ~1, m] = 1.0
        if ndw:                                  # c_dw: sign (-1)^n_up
            cdn[m & ~2, m] = -1.0 if nup else 1.0
    return Site("fermion", dim, qn,
                {"H": h, "Cup": cup, "Cdn": cdn, "P": p},
                {"H": "bosonic", "Cup": "fermionic", "Cdn": "fermionic", "P": "psign"})


# ---------------------------------------------------------------------------
# Blocks (src/BLOCKS/BLOCKS.f90) with dense operators in the reference basis
# ---------------------------------------------------------------------------
@dataclass
class Block:
    kind: str
    length: int
    qn: np.ndarray                  # (dim, qdim), basis-state QNs in basis order
    ops: dict
    optype: dict

    @property
    def dim(self) -> int:
        return self.qn.shape[0]

    def copy(self) -> "Block":
        return Block(self.kind, self.length, self.qn.copy(),
                     {k: v.copy() for k, v in self.ops.items()}, dict(self.optype))


def block_from_site(s: Site) -> Block:
    return Block(s.kind, 1, s.qn.copy(), {k: v.copy() for k, v in s.ops.items()},
                 dict(s.optype))


def sectors(qn: np.ndarray):
    """QN -> increasing state list, in first-appearance order.

    sectors_list(tbasis) / appends, src/LIST/LIST_SECTORS.f90:286-325.
    """
    order, maps = [], {}
    for i, q in enumerate(map(tuple, qn)):
        if q not in maps:
            maps[q] = []
            order.append(q)
        maps[q].append(i)
    return order, {q: np.array(v, dtype=np.int64) for q, v in maps.items()}


def pbc_site(s: Site) -> Site:
    """The site of a PBC run: every operator but H also under its link-'p'
    key (suffix ``_p``), src/BLOCKS/SITES.f90:339-342 (spin) and :413-424
    (electron: C and P)."""
    ops, optype = dict(s.ops), dict(s.optype)
    for k in s.ops:
        if k != "H":
            ops[k + "_p"] = s.ops[k].copy()
            optype[k + "_p"] = s.optype[k]
    return Site(s.kind, s.dim, s.qn, ops, optype)


def _link(key: str) -> str:
    """Link of an operator key: ``_p`` keys are the 'p' end, all others 'n'."""
    return "p" if key.endswith("_p") else "n"


def _connect(left_ops, right_ops, kind, hij, link: str = "n"):
    """connect_{spin,fermion}_blocks, src/DMRG_CONNECT.f90:296-454: the
    operators of the given link on both sides, the sign operator P of that
    link on the left one."""
    x = "" if link == "n" else "_p"
    if kind == "spin":
        return (hij[0][0] * np.kron(left_ops["Sz" + x], right_ops["Sz" + x])
                + hij[1][1] * np.kron(left_ops["Sp" + x], right_ops["Sp" + x].T)
                + hij[1][1] * np.kron(left_ops["Sp" + x].T, right_ops["Sp" + x]))
    pl = left_ops["P" + x]
    out = 0.0
    names = ["Cup", "Cdn"]
    for io in range(2):
        for jo in range(2):
            t = hij[io][jo]
            if t == 0.0:
                continue
            cl, cr = left_ops[names[io] + x], right_ops[names[jo] + x]
            out = out + t * np.kron(cl.T @ pl, cr) + t * np.kron(pl @ cl, cr.T)
    return out


def enlarge_block(b: Block, dot: Site, side: str, hij, link: str = "n") -> Block:
    """enlarge_block, src/DMRG_CONNECT.f90:18-291.

    Link 'n' (OBC and PBC_next, the dot joins the inner end):
      left : H = H_b x 1_d + 1_b x H_d + H_{b,d}, basis self .o. dot;
      right: H = 1_d x H_b + H_d x 1_b + H_{d,b}, basis dot .o. self.
    Link 'p' (PBC_prev, the dot joins the outer end):
      left : H = 1_d x H_b + H_d x 1_b + H_{d,b}, basis dot .o. self;
      right: H = H_b x 1_d + 1_b x H_d + H_{b,d}, basis self .o. dot.
    H_{x,y} is connect(x, y) on the step's link.  The other operators
    (:116-237) follow the table below, with O of link olink (``_p`` keys
    'p'), P_b / P_d the sign operators of that link, and 1 identities:

      link side olink  bosonic      fermionic    psign
      n    l    n      1_b x O_d    P_b x O_d    P_b x O_d
      n    l    p      O_b x 1_d    O_b x 1_d    O_b x P_d
      n    r    n      O_d x 1_b    O_d x 1_b    O_d x P_b
      n    r    p      1_d x O_b    P_d x O_b    P_d x O_b
      p    l    n      1_d x O_b    P_d x O_b    P_d x O_b
      p    l    p      O_d x 1_b    O_d x 1_b    O_d x P_b
      p    r    n      O_b x 1_d    O_b x 1_d    O_b x P_d
      p    r    p      1_b x O_d    P_b x O_d    P_b x O_d
    """
    db, dd = b.dim, dot.dim
    ib, idd = np.eye(db), np.eye(dd)
    ops, optype = {}, dict(b.optype)
    # dot joins after the block (block factor first) or before it
    dot_last = (side == "left") == (link == "n")
    if dot_last:
        ops["H"] = (np.kron(b.ops["H"], idd) + np.kron(ib, dot.ops["H"])
                    + _connect(b.ops, dot.ops, b.kind, hij, link))
        qn = (b.qn[:, None, :] + dot.qn[None, :, :]).reshape(db * dd, -1)
    else:
        ops["H"] = (np.kron(idd, b.ops["H"]) + np.kron(dot.ops["H"], ib)
                    + _connect(dot.ops, b.ops, b.kind, hij, link))
        qn = (dot.qn[:, None, :] + b.qn[None, :, :]).reshape(db * dd, -1)
    for k, t in b.optype.items():
        if k == "H":
            continue
        x = "" if _link(k) == "n" else "_p"
        if _link(k) == link:
            # the end the dot joins: the dot's operator, with the block's
            # sign in front of it where the block precedes the dot
            if dot_last:
                ops[k] = np.kron(ib if t == "bosonic" else b.ops["P" + x], dot.ops[k])
            else:
                ops[k] = np.kron(dot.ops[k], b.ops["P" + x] if t == "psign" else ib)
        else:
            # the other end keeps the block's operator, with the dot's sign
            # in front of it where the dot precedes the block
            if dot_last:
                ops[k] = np.kron(b.ops[k], dot.ops["P" + x] if t == "psign" else idd)
            else:
                ops[k] = np.kron(idd if t == "bosonic" else dot.ops["P" + x], b.ops[k])
    return Block(b.kind, b.length + 1, qn, ops, optype)


# ---------------------------------------------------------------------------
# Superblock: states, direct term table and H*v
# ---------------------------------------------------------------------------
@dataclass
class Superblock:
    target: tuple
    qns: list                       # left QN per SB sector
    Dl: np.ndarray
    Dr: np.ndarray
    offset: np.ndarray
    lmap: list                      # left block states per sector
    rmap: list                      # right block states per sector
    HL: list = field(default_factory=list)
    HR: list = field(default_factory=list)
    terms: list = field(default_factory=list)   # (dst, src, A, B, ishc)

    @property
    def n(self) -> int:
        return int((self.Dl * self.Dr).sum())


def current_target(length: int, target_qn, norb: int = 1, qn_type: str = "local") -> tuple:
    """sb_set_current_qn, src/DMRG_SUPERBLOCK.f90:717-726: QNtype=local
    int(target*L*Norb), QNtype=global min(L, int(target*Norb))."""
    if qn_type[:1].lower() == "g":
        return tuple(float(min(length, int(t * norb))) for t in target_qn)
    return tuple(float(int(t * length * norb)) for t in target_qn)


def sb_get_states(left: Block, right: Block, target: tuple) -> Superblock:
    """sb_get_states + sb_build_dims, src/DMRG_SUPERBLOCK.f90:36-121,167-283."""
    lord, lmaps = sectors(left.qn)
    _, rmaps = sectors(right.qn)
    qns, dls, drs, lm, rm = [], [], [], [], []
    for q in lord:
        rq = tuple(t - a for t, a in zip(target, q))
        if rq not in rmaps:
            continue
        qns.append(q)
        lm.append(lmaps[q])
        rm.append(rmaps[rq])
        dls.append(len(lmaps[q]))
        drs.append(len(rmaps[rq]))
    if not qns:
        raise RuntimeError("sb_get_states ERROR: total_states=0. There are no SB states.")
    dl, dr = np.array(dls), np.array(drs)
    off = np.concatenate([[0], np.cumsum(dl * dr)[:-1]]).astype(np.int64)
    return Superblock(target, qns, dl, dr, off, lm, rm)


def setup_direct(sb: Superblock, left: Block, right: Block, hij, pbc: bool = False) -> Superblock:
    """Setup_SuperBlock_{Spin,Fermion}_Direct, src/DMRG_SUPERBLOCK_SETUP.f90:152-693.

    Terms are (dst, src, A, B, is_hconjg): Y_dst += A . Psi_src . B^T.  With
    PBC the terms of link 'p' (the operators of both outer ends) follow those
    of link 'n' in every sector (spin :323-358) or every sector and spin
    (fermion :619-638).  The spin terms under PBC are those of
    connect_spin_blocks(link='p') (src/DMRG_CONNECT.f90:396-454, the dense
    path's H, SETUP:74): every sector gets its Sz.Sz term and the S+S- term
    its own conjugate partner.  The reference's direct spin setup instead
    copies the partner of the link-'n' term (SETUP:352-354) and skips the
    sectors without a qn-1 neighbour (the cycle at :305).
    """
    index = {q: i for i, q in enumerate(sb.qns)}
    sb.HL = [left.ops["H"][np.ix_(m, m)] for m in sb.lmap]
    sb.HR = [right.ops["H"][np.ix_(m, m)] for m in sb.rmap]
    sb.terms = []
    links = [""] + (["_p"] if pbc else [])
    for isb, q in enumerate(sb.qns):
        ai, bi = sb.lmap[isb], sb.rmap[isb]
        if left.kind == "spin":                           # :289-358
            for x in links:
                a = hij[0][0] * left.ops["Sz" + x][np.ix_(ai, ai)]
                b = right.ops["Sz" + x][np.ix_(bi, bi)]
                sb.terms.append((isb, isb, a, b, 0))
                qm = (q[0] - 1.0,)
                if qm not in index:
                    continue
                jsb = index[qm]
                aj, bj = sb.lmap[jsb], sb.rmap[jsb]
                a2 = hij[1][1] * left.ops["Sp" + x][np.ix_(ai, aj)]
                b2 = right.ops["Sp" + x].T[np.ix_(bi, bj)]
                sb.terms.append((isb, jsb, a2, b2, 0))
                sb.terms.append((jsb, isb, a2.T.copy(), b2.T.copy(), 1))
        else:                                              # :574-641
            names = ["Cup", "Cdn"]
            for ispin, dq in enumerate([(1.0, 0.0), (0.0, 1.0)]):
                qm = (q[0] - dq[0], q[1] - dq[1])
                if qm not in index:
                    continue
                jsb = index[qm]
                aj, bj = sb.lmap[jsb], sb.rmap[jsb]
                t = hij[ispin][ispin]
                if t == 0.0:
                    continue
                for x in links:
                    a1 = t * (left.ops[names[ispin] + x].T @ left.ops["P" + x])[np.ix_(ai, aj)]
                    b1 = right.ops[names[ispin] + x][np.ix_(bi, bj)]
                    sb.terms.append((isb, jsb, a1, b1, 0))
                    sb.terms.append((jsb, isb, a1.T.copy(), b1.T.copy(), 1))
    return sb


def hv_direct(sb: Superblock, v: np.ndarray) -> np.ndarray:
    """spMatVec_direct_main, src/DMRG_SUPERBLOCK_SETUP.f90:740-865, matrix form.

    Psi_q = v[off_q : off_q + Dl*Dr] as C row-major (Dl, Dr): element
    i = ir + (il-1)*Dr + off (Fortran) -> [il, ir].  Order of evaluation is the
    reference's: 1xH_R, H_Lx1, then for each term C = B.V (B first), Hv += C.A^T.
    """
    hv = np.zeros_like(v)
    psi = [v[o:o + dl * dr].reshape(dl, dr) for o, dl, dr in zip(sb.offset, sb.Dl, sb.Dr)]
    out = [hv[o:o + dl * dr].reshape(dl, dr) for o, dl, dr in zip(sb.offset, sb.Dl, sb.Dr)]
    for k in range(len(sb.qns)):
        out[k] += psi[k] @ sb.HR[k].T               # :767-779
        out[k] += sb.HL[k] @ psi[k]                 # :784-796
    for dst, src, a, b, _ in sb.terms:               # :808-857
        w = psi[src] @ b.T                           # C = B.V   (Dl_src x Dr_dst)
        out[dst] += a @ w                            # Hv += A.C^T
    return hv


def hv_direct_loops(sb: Superblock, v: np.ndarray) -> np.ndarray:
    """Scalar-loop twin of hv_direct for tiny superblocks (reference loop order)."""
    hv = np.zeros_like(v)
    for k in range(len(sb.qns)):
        dl, dr, off = int(sb.Dl[k]), int(sb.Dr[k]), int(sb.offset[k])
        for il in range(dl):
            for ir in range(dr):
                i = ir + il * dr + off
                for jr in range(dr):
                    hv[i] += sb.HR[k][ir, jr] * v[jr + il * dr + off]
        for ir in range(dr):
            for il in range(dl):
                i = ir + il * dr + off
                for jl in range(dl):
                    hv[i] += sb.HL[k][il, jl] * v[ir + jl * dr + off]
    for dst, src, a, b, _ in sb.terms:
        dld, drd = a.shape[0], b.shape[0]
        dls, drs = a.shape[1], b.shape[1]
        c = np.zeros((drd, dls))
        for aj in range(dls):
            for bi in range(drd):
                for bj in range(drs):
                    c[bi, aj] += b[bi, bj] * v[bj + aj * drs + sb.offset[src]]
        for bi in range(drd):
            for ai in range(dld):
                i = bi + ai * drd + sb.offset[dst]
                for aj in range(dls):
                    hv[i] += a[ai, aj] * c[bi, aj]
    return hv


def hv_flops(sb: Superblock) -> float:
    """Algorithmic FLOPs of one direct H*v with dense sector blocks."""
    f = float(np.sum(2.0 * sb.Dl * sb.Dr * (sb.Dl + sb.Dr)))
    for dst, src, _, _, _ in sb.terms:
        f += 2.0 * sb.Dr[dst] * (sb.Dl[src] * sb.Dr[src] + sb.Dl[dst] * sb.Dl[src])
    return f


def sb_dense(sb: Superblock) -> np.ndarray:
    """Dense H_sb (same matrix as Setup_SuperBlock_Sparse + dump, SETUP:46-88)."""
    n = sb.n
    h = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        h[:, j] = hv_direct(sb, e)
    return h


# ---------------------------------------------------------------------------
# Eigen-solvers (restating SciFortran sp_eigh / eigh)
# ---------------------------------------------------------------------------
def lanczos_lowest(matvec, n: int, neigen: int, v0: np.ndarray, tol: float = 1e-12,
                   kmax: int = 600):
    """Lowest `neigen` eigenpairs by Lanczos with full re-orthogonalisation.

    Restates the role of SciFortran sp_eigh (src/DMRG_SUPERBLOCK.f90:412-425):
    converged eigenvalues/vectors to residual ||Hx - e x|| <= tol*max(1,|e|).
    Returns (evals ascending, evecs (n, neigen), number of H*v).
    """
    kmax = min(kmax, n)
    vecs = np.zeros((kmax + 1, n))
    q = v0 / np.linalg.norm(v0)
    vecs[0] = q
    alpha, beta = [], []
    nop = 0
    for j in range(kmax):
        w = matvec(vecs[j])
        nop += 1
        a = float(np.dot(vecs[j], w))
        w -= a * vecs[j]
        if j > 0:
            w -= beta[-1] * vecs[j - 1]
        for _ in range(2):
            w -= vecs[:j + 1].T @ (vecs[:j + 1] @ w)
        b = float(np.linalg.norm(w))
        alpha.append(a)
        beta.append(b)
        m = j + 1
        if m >= neigen:
            t = np.diag(alpha) + np.diag(beta[:-1], 1) + np.diag(beta[:-1], -1)
            th, s = np.linalg.eigh(t)
            res = np.abs(b * s[-1, :neigen])
            if np.all(res <= tol * np.maximum(1.0, np.abs(th[:neigen]))) or m == kmax or b < 1e-300:
                x = vecs[:m].T @ s[:, :neigen]
                x /= np.linalg.norm(x, axis=0)
                return th[:neigen], x, nop
        vecs[j + 1] = w / b
    raise RuntimeError("lanczos_lowest: not converged")


# ---------------------------------------------------------------------------
# RDM + truncation (src/DMRG_RDM.f90, src/MATRIX/MATRIX_BLOCKS.f90)
# ---------------------------------------------------------------------------
@dataclass
class RhoBlocks:
    qns: list
    maps: list
    evals_blk: list
    evecs_blk: list
    evalues: np.ndarray = None      # all evals, sorted descending
    eorder: np.ndarray = None       # (block, position) per sorted eval


def build_rho(sb: Superblock, psi: np.ndarray, side: str, qn_trunc_error: float = 0.0,
              qn_trunc_dim: int = 1) -> RhoBlocks:
    """sb_get_rdm + build_density_matrix, src/DMRG_RDM.f90:18-162."""
    qns, maps, blocks = [], [], []
    for k, q in enumerate(sb.qns):
        dl, dr, off = int(sb.Dl[k]), int(sb.Dr[k]), int(sb.offset[k])
        p = psi[off:off + dl * dr].reshape(dl, dr)
        if side == "left":
            rho = p @ p.T
            qns.append(q)
            maps.append(sb.lmap[k])
        else:
            rho = p.T @ p
            qns.append(tuple(t - a for t, a in zip(sb.target, q)))
            maps.append(sb.rmap[k])
        if np.any(np.abs(rho - rho.T) > 1e-12):
            raise RuntimeError("build_density_matrix error: rho not Symmetric")
        blocks.append(rho)
    return eigh_blocks(qns, maps, blocks, qn_trunc_error, qn_trunc_dim)


def eigh_blocks(qns, maps, blocks, qn_trunc_error: float = 0.0, qn_trunc_dim: int = 1) -> RhoBlocks:
    """eigh_blocks_matrix(sort=T, reverse=T), src/MATRIX/MATRIX_BLOCKS.f90:589-657,
    with the QN-sector truncation of :627-629 (a block whose eigenvalues sum
    below QNtruncation_error, or whose dimension is below QNtruncation_dim,
    has all its eigenvalues set to tiny(1d0)) before negatives become tiny."""
    ev_blk, vec_blk, allv, order = [], [], [], []
    for ib, rho in enumerate(blocks):
        e, u = np.linalg.eigh(rho)
        if np.sum(e) < qn_trunc_error or len(e) < qn_trunc_dim:
            e = np.full_like(e, TINY)
        e = np.where(e < 0.0, TINY, e)
        ev_blk.append(e)
        vec_blk.append(u)
        allv.extend(e.tolist())
        order.extend((ib, p) for p in range(len(e)))
    allv = np.array(allv)
    asc = np.argsort(allv, kind="stable")
    desc = asc[::-1]
    r = RhoBlocks(qns, maps, ev_blk, vec_blk)
    r.evalues = allv[desc]
    r.eorder = [order[i] for i in desc]
    return r


def kept_count(evals: np.ndarray, mstates: int, mdim: int, deg_thr: float,
               estates: float = 0.0) -> int:
    """m_s selection + degeneracy extension, src/DMRG_RDM.f90:204-218."""
    if mstates != 0:
        ms = min(mstates, mdim, len(evals))
    elif estates != 0.0:
        ms = int(np.argmin(np.abs(1.0 - np.cumsum(evals) - estates))) + 1
    else:
        raise RuntimeError("Mdmrg=Edmrg=0 can not fix a threshold for the RDM")
    e_ = evals[ms - 1]
    j_ = ms
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(j_, len(evals)):
            err = abs(e_ - evals[i]) / e_
            if err <= deg_thr:
                ms += 1
    return ms


def renormalize(b: Block, rho: RhoBlocks, mstates: int, deg_thr: float,
                estates: float = 0.0):
    """renormalize_block + operators_renormalization, src/DMRG_RDM.f90:172-404.

    Returns (new block, m_s, truncation error, kept-state QN list, kept eigenvalues).
    """
    ms = kept_count(rho.evalues, mstates, b.dim, deg_thr, estates)
    trunc = 1.0 - float(np.sum(rho.evalues[:ms]))
    u = np.zeros((b.dim, ms))
    newqn = np.zeros((ms, b.qn.shape[1]))
    for im in range(ms):                       # rho%sparse(m_left,m_s), MATRIX_BLOCKS:556-587
        ib, pos = rho.eorder[im]
        u[rho.maps[ib], im] = rho.evecs_blk[ib][:, pos]
        newqn[im] = rho.qns[ib]
    ops = {k: u.T @ o @ u for k, o in b.ops.items()}
    return Block(b.kind, b.length, newqn, ops, dict(b.optype)), ms, trunc, newqn, rho.evalues[:ms]


def entropy(evals: np.ndarray) -> float:
    """write_entanglement S = -sum l log(l+1e-32), src/DMRG_MAIN.f90:501-521."""
    ev = np.where(evals < 0.0, 0.0, evals)
    return float(-np.sum(ev * np.log(ev + EPS_LOG)))


def entropy_split(rho: RhoBlocks) -> tuple:
    """write_entanglement S_N and S_c, src/DMRG_MAIN.f90:519-560: with
    P_q = sum_i lambda_q(i) (rho%lambda(), MATRIX_BLOCKS.f90:660-687) and
    S_q = sum_i (lambda/P_q) log(lambda/P_q + 1e-32),
    S_N = -sum_q P_q log(P_q + 1e-32), S_c = -sum_q P_q S_q.  A zero-weight
    block (a division by zero in the reference) contributes nothing."""
    sn = sc = 0.0
    for e in rho.evals_blk:
        if len(e) == 0:
            continue
        P = float(np.sum(e))
        sn -= P * math.log(P + EPS_LOG)
        if P > 0.0:
            o = e / P
            sc -= P * float(np.sum(o * np.log(o + EPS_LOG)))
    return sn, sc


def lambda_lists(rho: RhoBlocks) -> list:
    """rho%lambda() per QN block (LambdaQ_left/right of write_entanglement):
    [(qn, eigenvalues ascending)], blocks in superblock sector order."""
    return [(tuple(float(x) for x in q), np.sort(e)) for q, e in zip(rho.qns, rho.evals_blk) if len(e)]


# ---------------------------------------------------------------------------
# DMRG driver (src/DMRG_MAIN.f90)
# ---------------------------------------------------------------------------
@dataclass
class Params:
    ldmrg: int
    mdmrg: int
    target_qn: tuple
    hij: list
    msweep: list = field(default_factory=list)
    neigen: int = 1
    lanc_dim_threshold: int = 128
    lanc_tol: float = 1e-12
    deg_thr: float = 0.1
    seed: int = 0
    norb: int = 1
    ncv_max: int = 600
    pbc: bool = False               # PBCdmrg (infinite algorithm only)
    qn_type: str = "local"          # QNtype
    qn_trunc_error: float = 0.0     # QNtruncation_error
    qn_trunc_dim: int = 1           # QNtruncation_dim


@dataclass
class StepRecord:
    kind: str
    left_len: int
    right_len: int
    L: int
    energies: np.ndarray
    e_per_site: np.ndarray
    entropy_left: float
    entropy_right: float
    m_left: int
    m_right: int
    trunc_left: float
    trunc_right: float
    kept_left: dict
    kept_right: dict
    m_sb: int
    nsect: int
    nop: int
    hv_flops: float
    kept_w_left: dict = None     # QN -> kept rho eigenvalues
    kept_w_right: dict = None
    cut_left: float = 0.0        # lambda_{m_s}
    cut_right: float = 0.0
    sn_left: float = 0.0         # number / configurational entanglement entropies
    sc_left: float = 0.0
    sn_right: float = 0.0
    sc_right: float = 0.0
    lam_left: list = None        # rho%lambda() lists
    lam_right: list = None


class DMRG:
    """init_dmrg / run_DMRG / step_dmrg, src/DMRG_MAIN.f90:22-433.

    With ``p.pbc`` the site carries the link-'p' operators, an even LDMRG is
    shifted to the next odd one (src/INPUT_VARS.f90:178-183), every infinite
    iteration is a link-'n' step followed by a link-'p' step, and the finite
    algorithm is refused, as in the reference (src/DMRG_MAIN.f90:179-180).
    """

    def __init__(self, dot: Site, p: Params):
        if p.pbc:
            dot = pbc_site(dot)
            if p.ldmrg % 2 == 0:
                p.ldmrg += 1
        self.dot, self.p = dot, p
        self.left = block_from_site(dot)
        self.right = block_from_site(dot)
        self.records: list[StepRecord] = []
        self.last_sb = None
        self.last_psi = None

    def step(self, kind: str, label: int = 0, mstates: int = None, link: str = "n") -> StepRecord:
        p = self.p
        ms = p.mdmrg if mstates is None else mstates
        self.left = enlarge_block(self.left, self.dot, "left", p.hij, link)
        self.right = enlarge_block(self.right, self.dot, "right", p.hij, link)
        L = self.left.length + self.right.length
        target = current_target(L, p.target_qn, p.norb, p.qn_type)
        sb = sb_get_states(self.left, self.right, target)
        setup_direct(sb, self.left, self.right, p.hij, p.pbc)
        msb = sb.n
        neig = min(msb, p.neigen)
        nop = 0
        if neig == msb or msb <= p.lanc_dim_threshold:     # sb_diag, SUPERBLOCK.f90:367-369
            e, u = np.linalg.eigh(sb_dense(sb))
            evals, evecs = e[:neig], u[:, :neig]
        else:
            v0 = splitmix64_uniform_fast(msb, p.seed)
            evals, evecs, nop = lanczos_lowest(lambda x: hv_direct(sb, x), msb, neig, v0,
                                               p.lanc_tol, p.ncv_max)
        psi = evecs[:, 0]
        self.last_sb, self.last_psi = sb, psi
        rl = build_rho(sb, psi, "left", p.qn_trunc_error, p.qn_trunc_dim)
        rr = build_rho(sb, psi, "right", p.qn_trunc_error, p.qn_trunc_dim)
        if kind == "i":
            do_ren = self.left.length != p.ldmrg
        else:
            do_ren = not (self.left.length == p.ldmrg + 1 and label == 1)
        ml, mr, tl, tr = self.left.dim, self.right.dim, 0.0, 0.0
        kl, kr, wl, wr, cl, cr = {}, {}, {}, {}, 0.0, 0.0
        if do_ren:
            self.left, ml, tl, ql, lw = renormalize(self.left, rl, ms, p.deg_thr)
            self.right, mr, tr, qr, rw = renormalize(self.right, rr, ms, p.deg_thr)
            for q, w in zip(map(tuple, ql), lw):
                kl[q] = kl.get(q, 0) + 1
                wl.setdefault(q, []).append(float(w))
            for q, w in zip(map(tuple, qr), rw):
                kr[q] = kr.get(q, 0) + 1
                wr.setdefault(q, []).append(float(w))
            cl, cr = float(lw[-1]), float(rw[-1])
        rec = StepRecord(kind, self.left.length, self.right.length, L, evals,
                         evals / L / p.norb, entropy(rl.evalues), entropy(rr.evalues),
                         ml, mr, tl, tr, kl, kr, msb, len(sb.qns), nop, hv_flops(sb), wl, wr, cl, cr,
                         *entropy_split(rl), *entropy_split(rr), lambda_lists(rl), lambda_lists(rr))
        self.records.append(rec)
        return rec

    def infinite(self):
        """infinite_DMRG, src/DMRG_MAIN.f90:140-159."""
        while self.left.length < self.p.ldmrg:
            self.step("i", link="n")
            if self.p.pbc:
                self.step("i", link="p")

    def finite(self):
        """finite_DMRG, src/DMRG_MAIN.f90:168-241 (OBC only; PBC refused :179-180)."""
        p = self.p
        if p.pbc:
            raise RuntimeError("finite_DMRG ERROR: Finite DMRG with PBC not yet implemented.")
        fn = 2 * p.ldmrg
        bl = {1: {1: self.left.copy()}, 2: {1: self.right.copy()}}
        while self.left.length < p.ldmrg:
            self.step("i")
            bl[1][self.left.length] = self.left
            bl[2][self.right.length] = self.right
        for im, m in enumerate(p.msweep):
            while self.left.length < fn:
                self.right = bl[2][fn - self.left.length]
                self.step("f", 1, m)
                bl[1][self.left.length] = self.left
            while self.right.length < fn:
                self.left = bl[1][fn - self.right.length]
                self.step("f", 2, m)
                bl[2][self.right.length] = self.right
            while self.left.length <= p.ldmrg:
                self.right = bl[2][fn - self.left.length]
                self.step("f", 1, m)
                bl[1][self.left.length] = self.left


# ---------------------------------------------------------------------------
# Brute-force exact diagonalisation of full chains (pinning only)
# ---------------------------------------------------------------------------
def chain_hamiltonian(dot: Site, L: int, hij) -> tuple[np.ndarray, np.ndarray]:
    """Full-chain H (OBC) and per-state QN, built site by site with Kronecker
    products in the left-block order; independent of the superblock code."""
    b = block_from_site(dot)
    for _ in range(L - 1):
        b = enlarge_block(b, dot, "left", hij)
    return b.ops["H"], b.qn


def ed_site_expectations(dot: Site, L: int, hij, target: tuple, ops) -> np.ndarray:
    """<psi0| O_i |psi0>, i = 1..L, for each site operator in ops, on the exact
    ground state of the full OBC chain in the target sector.  The chain basis
    is built site by site in the left-block order, so site 1 is the leading
    Kronecker factor.  This is the untruncated limit of Measure_DMRG
    (src/DMRG_MEASURE.f90:148-706); the operators are diagonal in the
    occupation basis, so no Jordan-Wigner string enters.  Returns
    (len(ops), L)."""
    h, qn = chain_hamiltonian(dot, L, hij)
    sel = np.all(np.isclose(qn, np.array(target)[None, :]), axis=1)
    e, v = np.linalg.eigh(h[np.ix_(sel, sel)])
    psi = np.zeros(h.shape[0])
    psi[sel] = v[:, 0]
    d = dot.dim
    T = psi.reshape([d] * L)
    out = np.zeros((len(ops), L))
    for k, O in enumerate(ops):
        for i in range(L):
            Tm = np.moveaxis(T, i, 0).reshape(d, -1)
            out[k, i] = float(np.sum(Tm * (np.asarray(O) @ Tm)))
    return out


def ed_ground(dot: Site, L: int, hij, target: tuple, k: int = 1) -> np.ndarray:
    h, qn = chain_hamiltonian(dot, L, hij)
    sel = np.all(np.isclose(qn, np.array(target)[None, :]), axis=1)
    hs = h[np.ix_(sel, sel)]
    return np.linalg.eigvalsh(hs)[:k]


def free_fermion_obc(L: int, nup: int, ndw: int, t: float = 1.0) -> float:
    """OBC tight-binding ground energy: sum of the lowest single-particle levels."""
    eps = np.sort(-2.0 * t * np.cos(np.pi * np.arange(1, L + 1) / (L + 1)))
    return float(eps[:nup].sum() + eps[:ndw].sum())

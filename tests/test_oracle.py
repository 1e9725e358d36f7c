"""Pins the CPU oracle (oracle/dmrg_oracle.py) before anything is checked against it.

Each check is chosen so that a plausible slip (dropped term, wrong sign, wrong
transpose, wrong sector bookkeeping) fails at least one of them:
  * the reference's golden .check trajectories (tests/golden, copied data);
  * closed forms: 4-site Heisenberg, free fermions (U=0), 2-site Hubbard;
  * brute-force exact diagonalisation of full chains;
  * an independent dense Kronecker construction of H_sb;
  * library eigen-solvers (LAPACK eigh, ARPACK eigsh) for the Lanczos restatement;
  * invariants of rho.
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import dmrg_oracle as O

SPIN_HIJ = [[1.0, 0.0], [0.0, 0.5]]          # diag([Jp, Jx/2]), spinSU2_1d.f90:35
HUB_HIJ = [[-1.0, 0.0], [0.0, -1.0]]         # diag(ts), hm_1d.f90 (TS=-1)


def _load(golden_dir, name):
    return np.loadtxt(os.path.join(golden_dir, name))


# --------------------------------------------------------------------------- RNG
def test_splitmix64_known_answer():
    # SplitMix64(seed=0) first raw output is 0xE220A8397B1DCDAF (published KAT).
    raw = 0xE220A8397B1DCDAF
    expect = (raw >> 11) * 2.0 ** -53 * 2.0 - 1.0
    assert O.splitmix64_uniform(1, 0)[0] == expect
    a = O.splitmix64_uniform(257, 12345)
    b = O.splitmix64_uniform_fast(257, 12345)
    assert np.array_equal(a, b)
    assert a.min() >= -1.0 and a.max() < 1.0


# ------------------------------------------------------------------ golden files
@pytest.fixture(scope="module")
def spin_run():
    p = O.Params(ldmrg=40, mdmrg=20, target_qn=(0.0,), hij=SPIN_HIJ, neigen=2,
                 lanc_dim_threshold=128)
    d = O.DMRG(O.spin_site(2), p)
    d.infinite()
    return d


def test_spin_golden_energy_trajectory(spin_run, golden_dir):
    ref = _load(golden_dir, "spin12_L40_M20_energy.check")
    assert len(spin_run.records) == len(ref) == 39
    for r, row in zip(spin_run.records, ref):
        assert r.left_len == int(row[0])
        assert abs(r.e_per_site[0] - row[1]) < 1e-12
        assert abs(r.e_per_site[1] - row[2]) < 1e-12
    fin = _load(golden_dir, "spin12_L40_M20_final_energy.check")
    assert abs(spin_run.records[-1].e_per_site[0] - fin[1]) < 1e-12


def test_spin_golden_entropy_trajectory(spin_run, golden_dir):
    ref = _load(golden_dir, "spin12_L40_M20_Sentropy.check")
    # Rows 2..39.  The file's last row repeats row 39 (written by an older code
    # version that reused the previous step's rho spectrum); the current code
    # recomputes rho every step (src/DMRG_RDM.f90:73-88), so row 40 is skipped.
    assert ref[-1] == ref[-2]
    for r, s in zip(spin_run.records[:-1], ref[:-1]):
        assert abs(r.entropy_left - s) < 1e-10


def test_spin_closed_form_4site(spin_run):
    e4 = -(0.75 + math.sqrt(3.0) / 2.0) / 4.0
    assert abs(spin_run.records[0].e_per_site[0] - e4) < 1e-14


def test_hubbard_golden_trajectory(golden_dir):
    ref = _load(golden_dir, "hubbard_U0_L40_M40_energy.check")
    refs = _load(golden_dir, "hubbard_U0_L40_M40_Sentropy.check")
    p = O.Params(ldmrg=40, mdmrg=40, target_qn=(0.5, 0.5), hij=HUB_HIJ, neigen=1,
                 lanc_dim_threshold=128)
    d = O.DMRG(O.electron_site(0.0, True), p)
    d.infinite()
    assert len(d.records) == 39
    for r, e in zip(d.records, ref):
        assert abs(r.e_per_site[0] - e) < 1e-12
    for r, s in zip(d.records[:-1], refs[:-1]):
        assert abs(r.entropy_left - s) < 1e-9
    # first row: 4-site free fermions at half filling, -sqrt(5)/2 per site
    assert abs(d.records[0].e_per_site[0] + math.sqrt(5.0) / 2.0) < 1e-14
    # U=0: the untruncated steps (L=4, 6) equal the tight-binding value
    for r in d.records[:2]:
        ex = O.free_fermion_obc(r.L, r.L // 2, r.L // 2) / r.L
        assert abs(r.e_per_site[0] - ex) < 1e-12


U2_ROWS = 27  # rows of the published U=2 run reproduced by the current truncation rule


def test_hubbard_U2_published_trajectory(golden_dir):
    # .plot/Hubbard/U2 (U=2, M=50, HFMODE=F, iDMRG): the only U != 0 Hubbard
    # trajectory the reference holds.  Rows 1..27 (left length 2..28, with
    # truncation from left length 4 on) are reproduced to rounding.  At left
    # length 29 the current rule's 10% degeneracy window (DMRG_RDM.f90:213-218)
    # extends the cut from 52 to 55 states (lambda ratio 0.073, not a
    # borderline case) and the published run, made by an older code whose
    # DMRG.conf has no DEG_EVALS_THRESHOLD key, departs; no window threshold
    # (0, 1e-4, 0.05, 0.1) reproduces all 99 rows.
    ref = _load(golden_dir, "hubbard_U2_L100_M50_hf0_energy.dmrg")
    refs = _load(golden_dir, "hubbard_U2_L100_M50_hf0_Sentropy.dmrg")
    p = O.Params(ldmrg=U2_ROWS + 1, mdmrg=50, target_qn=(0.5, 0.5), hij=HUB_HIJ, neigen=1,
                 lanc_dim_threshold=128)
    d = O.DMRG(O.electron_site(2.0, False), p)
    d.infinite()
    assert len(d.records) == U2_ROWS
    assert max(r.m_left for r in d.records[:-1]) >= 50       # truncated steps are in the range
    for r, e, s in zip(d.records, ref, refs):
        assert r.left_len == int(e[0])
        assert abs(r.e_per_site[0] - e[1]) < 1e-12, r.left_len
        assert abs(r.entropy_left - s[1]) < 1e-10, r.left_len
    # the first row is the 4-site Hubbard chain at half filling: exact ED
    e4 = O.ed_ground(O.electron_site(2.0, False), 4, np.array(HUB_HIJ), (2.0, 2.0))
    assert abs(d.records[0].energies[0] - e4) < 1e-12


def test_hubbard_U0_published_and_tight_binding_exact(golden_dir):
    # .plot/Hubbard/U0 (U=0, M=40, HFMODE=F) and .plot/Hubbard/TBexact (exact
    # E/L of the half-filled open tight-binding chain of 2j sites).  Left
    # lengths 2..4 are untruncated and equal the published run; from 5 on
    # the degeneracy window of the current rule keeps more than 40 states of
    # the highly degenerate U=0 spectra, and every step lies between the
    # exact energy (variational bound) and the published one.  First 19 rows.
    ref = {int(r[0]): r[1] for r in _load(golden_dir, "hubbard_U0_L100_M40_hf0_energy.dmrg")}
    ex = {int(r[0]): r[1] for r in _load(golden_dir, "tb_exact_obc_eVSj.txt")}
    p = O.Params(ldmrg=20, mdmrg=40, target_qn=(0.5, 0.5), hij=HUB_HIJ, neigen=1, lanc_dim_threshold=128)
    d = O.DMRG(O.electron_site(0.0, False), p)
    d.infinite()
    assert [r.left_len for r in d.records] == list(range(2, 21))
    for r in d.records:
        e = r.e_per_site[0]
        if r.left_len <= 4:
            assert abs(e - ref[r.left_len]) < 1e-12
        assert ex[r.left_len] - 1e-12 <= e <= ref[r.left_len] + 1e-12, r.left_len
    assert max(r.m_left for r in d.records[3:]) > 40


def test_spin12_L300_published_first_rows(golden_dir):
    # .plot/HeisenbergSU2.3 (spin-1/2, M=40, LANC_NEIGEN=2): left lengths 2..8
    # to rounding (E0/L, E1/L, S); at left length 8 the degeneracy window keeps
    # 43 states and the runs part, the current rule's energies staying at or
    # below the published ones.  First 29 rows.
    ref = _load(golden_dir, "spin12_L300_M40_energy.dmrg")
    refs = _load(golden_dir, "spin12_L300_M40_Sentropy.dmrg")
    p = O.Params(ldmrg=30, mdmrg=40, target_qn=(0.0,), hij=SPIN_HIJ, neigen=2, lanc_dim_threshold=128)
    d = O.DMRG(O.spin_site(2), p)
    d.infinite()
    assert len(d.records) == 29
    for r, e, sv in zip(d.records, ref, refs):
        assert r.left_len == int(e[0])
        if r.left_len <= 8:
            assert abs(r.e_per_site[0] - e[1]) < 1e-12 and abs(r.e_per_site[1] - e[2]) < 1e-12
            assert abs(r.entropy_left - sv[1]) < 1e-10
        assert r.e_per_site[0] <= e[1] + 1e-12 and r.e_per_site[1] <= e[2] + 1e-10, r.left_len


# ------------------------------------------------------------------ exact checks
@pytest.mark.parametrize("sun,L", [(2, 10), (3, 6)])
def test_spin_untruncated_equals_ed(sun, L):
    p = O.Params(ldmrg=L // 2, mdmrg=4096, target_qn=(0.0,), hij=SPIN_HIJ, neigen=2,
                 lanc_dim_threshold=64)
    d = O.DMRG(O.spin_site(sun), p)
    d.infinite()
    for r in d.records:
        ed = O.ed_ground(O.spin_site(sun), r.L, SPIN_HIJ, (0.0,), k=2)
        assert np.allclose(r.energies, ed, atol=1e-10, rtol=0)


@pytest.mark.parametrize("U,hf", [(4.0, True), (2.0, False)])
def test_hubbard_untruncated_equals_ed(U, hf):
    site = O.electron_site(U, hf)
    p = O.Params(ldmrg=3, mdmrg=4096, target_qn=(0.5, 0.5), hij=HUB_HIJ, neigen=1,
                 lanc_dim_threshold=64)
    d = O.DMRG(site, p)
    d.infinite()
    for r in d.records:
        ed = O.ed_ground(site, r.L, HUB_HIJ, (r.L // 2, r.L // 2))
        assert abs(r.energies[0] - ed[0]) < 1e-10


def test_ed_site_expectations_sum_rules_and_kron():
    # spin-1/2, L=8, Sz_total = 1: sum_i <Sz_i> = 1, <Sz_i^2> = 1/4,
    # reflection symmetry of the OBC ground state
    sz = np.diag([0.5, -0.5])
    v = O.ed_site_expectations(O.spin_site(2), 8, SPIN_HIJ, (1.0,), [sz, sz @ sz])
    assert abs(v[0].sum() - 1.0) < 1e-12
    assert np.allclose(v[1], 0.25, atol=1e-13)
    assert np.allclose(v[0], v[0][::-1], atol=1e-12)
    assert np.ptp(v[0]) > 1e-3  # a nontrivial profile
    # Hubbard U=4, L=6, (N_up, N_dw) = (2, 2): particle-number sum rules,
    # docc <= min(n_up, n_dw), and an explicit Kronecker check of one site
    site = O.electron_site(4.0, True)
    nup, ndw = np.diag([0.0, 1.0, 0.0, 1.0]), np.diag([0.0, 0.0, 1.0, 1.0])
    w = O.ed_site_expectations(site, 6, HUB_HIJ, (2.0, 2.0), [nup, ndw, nup @ ndw])
    assert abs(w[0].sum() - 2.0) < 1e-12 and abs(w[1].sum() - 2.0) < 1e-12
    assert np.all(w[2] <= np.minimum(w[0], w[1]) + 1e-12)
    assert np.allclose(w[0], w[0][::-1], atol=1e-12)
    h, qn = O.chain_hamiltonian(site, 6, HUB_HIJ)
    sel = np.all(np.isclose(qn, [2.0, 2.0]), axis=1)
    e, u = np.linalg.eigh(h[np.ix_(sel, sel)])
    psi = np.zeros(h.shape[0])
    psi[sel] = u[:, 0]
    i = 2
    Oi = np.kron(np.kron(np.eye(4 ** i), nup @ ndw), np.eye(4 ** (6 - 1 - i)))
    assert abs(psi @ Oi @ psi - w[2][i]) < 1e-12


def test_hubbard_two_site_closed_form():
    # E = (U - sqrt(U^2+16 t^2))/2 for 2 sites, 1 up + 1 down; HF shift -U/2*N.
    U = 3.0
    ed = O.ed_ground(O.electron_site(U, True), 2, HUB_HIJ, (1.0, 1.0))[0]
    assert abs(ed - ((U - math.sqrt(U * U + 16.0)) / 2.0 - U)) < 1e-12


def test_finite_sweeps_reach_ed_when_untruncated():
    # Ldmrg=4 -> sweeps act on 2*Ldmrg+2 = 10 sites; M large -> exact.
    p = O.Params(ldmrg=4, mdmrg=4096, target_qn=(0.0,), hij=SPIN_HIJ, neigen=1,
                 lanc_dim_threshold=64, msweep=[4096])
    d = O.DMRG(O.spin_site(2), p)
    d.finite()
    ed10 = O.ed_ground(O.spin_site(2), 10, SPIN_HIJ, (0.0,))[0]
    fin = [r for r in d.records if r.kind == "f"]
    assert fin and all(r.L == 10 for r in fin)
    for r in fin:
        assert abs(r.energies[0] - ed10) < 1e-10


def test_finite_msweep_order_pinned_by_ed():
    # Msweep = [4, 4096] on 10 sites (finite_DMRG, src/DMRG_MAIN.f90:208-239,
    # fNlat = 2 Ldmrg = 8).  Sweep 1 runs 13 steps (left 4 -> 8, right 2 -> 8,
    # left 2 -> 5) and sweep 2, which starts from the unrenormalised left block
    # of length 5, runs 12.  Sweep 1 keeps 4 states: variational and visibly
    # above the exact energy.  Sweep 2 keeps everything: its last two steps
    # (both blocks rebuilt untruncated) are exact.  Ignoring Msweep, or taking
    # its entries in another order, fails one of these.  The last step of each
    # sweep (left length Ldmrg+1, label 1) is not renormalised
    # (src/DMRG_MAIN.f90:401-404): its left block keeps all 4 x 2 enlarged
    # states in sweep 1.
    p = O.Params(ldmrg=4, mdmrg=4096, target_qn=(0.0,), hij=SPIN_HIJ, neigen=1,
                 lanc_dim_threshold=16, msweep=[4, 4096])
    d = O.DMRG(O.spin_site(2), p)
    d.finite()
    ed10 = O.ed_ground(O.spin_site(2), 10, SPIN_HIJ, (0.0,))[0]
    fin = [r for r in d.records if r.kind == "f"]
    assert len(fin) == 25
    s1, s2 = fin[:13], fin[13:]
    assert [r.left_len for r in s1] == [5, 6, 7, 8, 7, 6, 5, 4, 3, 2, 3, 4, 5]
    assert [r.left_len for r in s2] == [6, 7, 8, 7, 6, 5, 4, 3, 2, 3, 4, 5]
    assert all(r.energies[0] >= ed10 - 1e-10 for r in s1)
    assert min(r.energies[0] - ed10 for r in s1[2:]) > 1e-4
    assert s1[-1].m_left == 8
    for r in s2[-2:]:
        assert abs(r.energies[0] - ed10) < 1e-10, (r.left_len, r.energies[0] - ed10)


# ------------------------------------------------------- superblock H, H*v
def _sb_after(dot, hij, steps, M, target_qn):
    p = O.Params(ldmrg=steps + 1, mdmrg=M, target_qn=target_qn, hij=hij, neigen=1,
                 lanc_dim_threshold=64)
    d = O.DMRG(dot, p)
    for _ in range(steps):
        d.step("i")
    left = O.enlarge_block(d.left, dot, "left", hij)
    right = O.enlarge_block(d.right, dot, "right", hij)
    tgt = O.current_target(left.length + right.length, target_qn)
    sb = O.sb_get_states(left, right, tgt)
    O.setup_direct(sb, left, right, hij)
    return sb, left, right


@pytest.mark.parametrize("model", ["spin", "spin1", "hubbard"])
def test_direct_hv_equals_kronecker_restriction(model):
    if model == "spin":
        dot, hij, tq = O.spin_site(2), SPIN_HIJ, (0.0,)
    elif model == "spin1":
        dot, hij, tq = O.spin_site(3), SPIN_HIJ, (0.0,)
    else:
        dot, hij, tq = O.electron_site(2.5, True), HUB_HIJ, (0.5, 0.5)
    sb, left, right = _sb_after(dot, hij, 2, 6, tq)
    # independent construction: full left x right Hamiltonian, then restrict
    hfull = (np.kron(left.ops["H"], np.eye(right.dim)) + np.kron(np.eye(left.dim), right.ops["H"])
             + O._connect(left.ops, right.ops, left.kind, hij))
    states = np.concatenate([(lm[:, None] * right.dim + rm[None, :]).ravel()
                             for lm, rm in zip(sb.lmap, sb.rmap)])
    href = hfull[np.ix_(states, states)]
    hd = O.sb_dense(sb)
    assert np.allclose(hd, href, atol=1e-13, rtol=0)
    assert np.allclose(hd, hd.T, atol=1e-13)
    # loop twin follows the reference loops exactly
    v = O.splitmix64_uniform_fast(sb.n, 7)
    assert np.allclose(O.hv_direct_loops(sb, v), O.hv_direct(sb, v), atol=1e-12)
    # no leakage: the restricted H equals the full H on SB states
    x = np.zeros(left.dim * right.dim)
    x[states] = v
    assert np.allclose((hfull @ x)[states], O.hv_direct(sb, v), atol=1e-12)


def test_lanczos_matches_library_solvers():
    sb, _, _ = _sb_after(O.spin_site(2), SPIN_HIJ, 4, 16, (0.0,))
    n = sb.n
    v0 = O.splitmix64_uniform_fast(n, 0)
    ev, x, nop = O.lanczos_lowest(lambda y: O.hv_direct(sb, y), n, 2, v0)
    hd = O.sb_dense(sb)
    ex = np.linalg.eigvalsh(hd)[:2]
    assert np.allclose(ev, ex, atol=1e-11)
    op = spla.LinearOperator((n, n), matvec=lambda y: O.hv_direct(sb, y), dtype=float)
    ea = np.sort(spla.eigsh(op, k=2, which="SA", tol=1e-12, v0=v0)[0])
    assert np.allclose(ev, ea, atol=1e-11)
    res = np.linalg.norm(hd @ x[:, 0] - ev[0] * x[:, 0])
    assert res < 1e-9


# ------------------------------------------------------------------ rho / truncation
def test_rho_invariants():
    sb, _, _ = _sb_after(O.electron_site(4.0, True), HUB_HIJ, 2, 16, (0.5, 0.5))
    hd = O.sb_dense(sb)
    e, u = np.linalg.eigh(hd)
    psi = u[:, 0]
    rl = O.build_rho(sb, psi, "left")
    rr = O.build_rho(sb, psi, "right")
    assert abs(rl.evalues.sum() - 1.0) < 1e-12 and abs(rr.evalues.sum() - 1.0) < 1e-12
    nz = min(len(rl.evalues), len(rr.evalues))
    assert np.allclose(rl.evalues[:nz], rr.evalues[:nz], atol=1e-12)
    # S from Schmidt values of each sector block equals S from rho
    sv = np.concatenate([np.linalg.svd(psi[o:o + a * b].reshape(a, b), compute_uv=False)
                         for o, a, b in zip(sb.offset, sb.Dl, sb.Dr)])
    lam = sv ** 2
    s_svd = float(-np.sum(lam * np.log(lam + 1e-32)))
    assert abs(O.entropy(rl.evalues) - s_svd) < 1e-12
    assert np.all(np.diff(rl.evalues) <= 0.0)


def test_product_state_has_zero_entropy():
    ev = np.array([1.0, 0.0, 0.0])
    assert O.entropy(ev) == 0.0
    bell = np.array([0.5, 0.5])
    assert abs(O.entropy(bell) - math.log(2.0)) < 1e-15


def test_kept_count_rule():
    ev = np.array([0.4, 0.3, 0.1, 0.095, 0.091, 0.05, 0.0])
    assert O.kept_count(ev, 3, 10, 0.1) == 5          # 0.095,0.091 within 10% of 0.1
    assert O.kept_count(ev, 3, 10, 0.0) == 3
    assert O.kept_count(ev, 2, 10, 0.1) == 2          # 0.1 is 66% below 0.3
    assert O.kept_count(ev, 20, 4, 0.1) == 5          # m_s=min(M,mdim)=4, 0.091 kept
    tied = np.array([0.5, 0.25, 0.25, 0.0])
    assert O.kept_count(tied, 2, 4, 0.1) == 3         # exact ties are never split
    # truncation error threshold mode (Estates): argmin |1-cumsum-E|
    assert O.kept_count(np.array([0.9, 0.09, 0.01]), 0, 3, 0.0, estates=0.01) == 2


# ------------------------------------------------------ C restatement (baseline)
@pytest.mark.parametrize("model", ["spin", "hubbard"])
def test_c_restatement_matches_numpy_oracle(model):
    from oracle import hv_ref
    if model == "spin":
        sb, _, _ = _sb_after(O.spin_site(2), SPIN_HIJ, 4, 16, (0.0,))
    else:
        sb, _, _ = _sb_after(O.electron_site(2.0, True), HUB_HIJ, 3, 24, (0.5, 0.5))
    c = hv_ref.from_oracle(sb)
    v = O.splitmix64_uniform_fast(sb.n, 4)
    assert np.allclose(c.hv(v), O.hv_direct(sb, v), atol=1e-12)
    assert np.allclose(c.hv(v), O.hv_direct_loops(sb, v), atol=1e-12)


# --------------------------------------------------------------------------- PBC
# Periodic rings (PBCdmrg, infinite algorithm).  Independent pins: ring
# closed forms and a Fock-space ED whose fermion signs come from the mode
# order alone, so a wrong Jordan-Wigner string at either block end fails.
def _ring_ed_fermion(N, hij_t, U, hf, nup, ndw):
    """Lowest energy of the Hubbard ring (N sites, bonds (i, i+1 mod N)) in
    the (nup, ndw) sector: modes 2i (up) and 2i+1 (dw), c_m |n> carries
    (-1)^(number of occupied modes below m)."""
    states = [s for s in range(1 << (2 * N))
              if bin(s & int("01" * N, 2)).count("1") == nup and bin(s & int("10" * N, 2)).count("1") == ndw]
    idx = {s: i for i, s in enumerate(states)}
    H = np.zeros((len(states), len(states)))

    def hop(s, a, b):  # c_a^+ c_b |s>: (sign, new state) or None
        if not (s >> b) & 1 or (a != b and (s >> a) & 1):
            return None
        sg = (-1) ** bin(s & ((1 << b) - 1)).count("1")
        s1 = s & ~(1 << b)
        sg *= (-1) ** bin(s1 & ((1 << a) - 1)).count("1")
        return sg, s1 | (1 << a)

    bonds = [(i, (i + 1) % N) for i in range(N)] if N > 2 else [(0, 1)]
    for s in states:
        i0 = idx[s]
        for i in range(N):
            nu, nd = (s >> (2 * i)) & 1, (s >> (2 * i + 1)) & 1
            H[i0, i0] += U * nu * nd - (0.5 * U * (nu + nd) if hf else 0.0)
        for i, j in bonds:
            for sp in range(2):
                for a, b in ((2 * i + sp, 2 * j + sp), (2 * j + sp, 2 * i + sp)):
                    r = hop(s, a, b)
                    if r:
                        H[idx[r[1]], i0] += hij_t * r[0]
    return float(np.linalg.eigvalsh(H)[0])


def _ring_ed_spin(sun, N, hij):
    """Lowest energy of the Heisenberg ring in the Sz=0 sector (brute force)."""
    s = O.spin_site(sun)
    d = s.dim
    Sz, Sp = s.ops["Sz"], s.ops["Sp"]

    def at(op, i):
        m = np.eye(1)
        for k in range(N):
            m = np.kron(m, op if k == i else np.eye(d))
        return m

    H = 0.0
    for i in range(N):
        j = (i + 1) % N
        H = H + hij[0][0] * at(Sz, i) @ at(Sz, j) + hij[1][1] * (at(Sp, i) @ at(Sp.T, j) + at(Sp.T, i) @ at(Sp, j))
    sz = sum(np.diag(at(Sz, i)) for i in range(N))
    sel = np.isclose(sz, 0.0)
    return float(np.linalg.eigvalsh(H[np.ix_(sel, sel)])[0])


def _pbc_run(site, ldmrg, m, tq, hij, thr=100000):
    p = O.Params(ldmrg=ldmrg, mdmrg=m, target_qn=tq, hij=hij, neigen=1, lanc_dim_threshold=thr, pbc=True)
    d = O.DMRG(site, p)
    d.infinite()
    return d


def test_pbc_heisenberg_ring_closed_forms():
    # spin-1/2 Heisenberg rings: E0(4) = -2, E0(6) = -2.8027756377319946
    # (exact diagonalisation values quoted throughout the literature)
    d = _pbc_run(O.spin_site(2), 3, 200, (0.0,), SPIN_HIJ)
    assert [r.L for r in d.records] == [4, 6]
    assert abs(d.records[0].energies[0] + 2.0) < 1e-12
    assert abs(d.records[1].energies[0] + 2.8027756377319946) < 1e-12


def test_pbc_even_ldmrg_shifted_and_finite_refused():
    p = O.Params(ldmrg=4, mdmrg=8, target_qn=(0.0,), hij=SPIN_HIJ, pbc=True)
    d = O.DMRG(O.spin_site(2), p)
    assert d.p.ldmrg == 5                                  # INPUT_VARS.f90:178-183
    with pytest.raises(RuntimeError, match="PBC"):
        d.finite()


@pytest.mark.parametrize("N", [4, 6])
def test_pbc_free_fermion_ring(N):
    # U=0: sum of the lowest ring levels -2 cos(2 pi k / N) per spin (periodic
    # fermions; the antiperiodic ring would give -5.657 and -6.928)
    d = _pbc_run(O.electron_site(0.0, True), 3, 400, (0.5, 0.5), HUB_HIJ)
    rec = {r.L: r for r in d.records}[N]
    eps = np.sort(-2.0 * np.cos(2.0 * np.pi * np.arange(N) / N))
    assert abs(rec.energies[0] - 2.0 * eps[:N // 2].sum()) < 1e-11


def test_pbc_hubbard_ring_equals_fock_ed():
    d = _pbc_run(O.electron_site(4.0, True), 3, 400, (0.5, 0.5), HUB_HIJ)
    for r in d.records:
        e = _ring_ed_fermion(r.L, -1.0, 4.0, True, r.L // 2, r.L // 2)
        assert abs(r.energies[0] - e) < 1e-10 * abs(e), r.L


def test_pbc_spin1_ring_equals_ed():
    d = _pbc_run(O.spin_site(3), 3, 400, (0.0,), SPIN_HIJ)
    for r in d.records:
        assert abs(r.energies[0] - _ring_ed_spin(3, r.L, SPIN_HIJ)) < 1e-10, r.L


def test_pbc_blocks_keep_fermion_algebra_at_both_ends():
    # operators of the inner (n) and outer (p) end of an enlarged block are
    # fermions: {C_a, C_b^+} = delta_ab, {C_a, C_b} = 0, P C P = -C
    s = O.pbc_site(O.electron_site(2.0, True))
    for side in ("left", "right"):
        b = O.block_from_site(s)
        for link in ("n", "p", "p", "n"):
            b = O.enlarge_block(b, s, side, HUB_HIJ, link)
            cs = [b.ops[k] for k in ("Cup", "Cdn", "Cup_p", "Cdn_p")]
            P = b.ops["P"]
            assert np.allclose(P, b.ops["P_p"])
            one = np.eye(b.dim)
            for i, a in enumerate(cs):
                assert np.allclose(P @ a @ P, -a)
                for j, c in enumerate(cs):
                    assert np.allclose(a @ c + c @ a, 0.0), (side, link, i, j)
                    anti = a @ c.T + c.T @ a
                    # the two ends are different sites from the first enlargement on
                    assert np.allclose(anti, one if i == j else 0.0), (side, link, i, j)


def test_pbc_superblock_hamiltonian_is_symmetric():
    # Hermitian H_sb in every step: the link-'p' spin-flip term carries its
    # own conjugate partner
    for site, tq, hij in ((O.spin_site(2), (0.0,), SPIN_HIJ), (O.electron_site(3.0, True), (0.5, 0.5), HUB_HIJ)):
        p = O.Params(ldmrg=5, mdmrg=12, target_qn=tq, hij=hij, neigen=1, lanc_dim_threshold=2000, pbc=True)
        d = O.DMRG(site, p)
        while d.left.length < p.ldmrg:
            for link in ("n", "p"):
                d.step("i", link=link)
                h = O.sb_dense(d.last_sb)
                assert np.allclose(h, h.T, atol=1e-12)


# ------------------------------------------------------------ entanglement split
def _spin_chain_ed(L):
    """Heisenberg chain H = sum Sz Sz + (S+S- + S-S+)/2 by explicit Kronecker
    products (no oracle code): H, per-state Sz."""
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    eye = np.eye(2)

    def op(o, i):
        m = np.array([[1.0]])
        for k in range(L):
            m = np.kron(m, o if k == i else eye)
        return m
    h = np.zeros((2 ** L, 2 ** L))
    for i in range(L - 1):
        h += op(sz, i) @ op(sz, i + 1) + 0.5 * (op(sp, i) @ op(sp.T, i + 1) + op(sp.T, i) @ op(sp, i + 1))
    tot = sum(np.diag(op(sz, i)) for i in range(L))
    return h, tot


def test_lambda_lists_equal_exact_reduced_density_matrix():
    # untruncated L=6 step: rho_L's spectrum per Sz sector of the left half
    # equals that of the exact ground state's reduced density matrix
    L = 6
    h, tot = _spin_chain_ed(L)
    sel = np.isclose(tot, 0.0)
    e, u = np.linalg.eigh(h[np.ix_(sel, sel)])
    psi = np.zeros(2 ** L)
    psi[sel] = u[:, 0]
    m = psi.reshape(2 ** (L // 2), 2 ** (L // 2))
    rho = m @ m.T
    _, totl = _spin_chain_ed(L // 2)
    p = O.Params(ldmrg=L // 2, mdmrg=4096, target_qn=(0.0,), hij=SPIN_HIJ, neigen=1, lanc_dim_threshold=4096)
    d = O.DMRG(O.spin_site(2), p)
    d.infinite()
    rec = d.records[-1]
    assert rec.L == L and abs(rec.energies[0] - e[0]) < 1e-12
    got = {q[0]: ev for q, ev in rec.lam_left}
    assert set(got) == set(np.unique(totl).tolist())
    for q, ev in got.items():
        s = np.isclose(totl, q)
        ref = np.linalg.eigvalsh(rho[np.ix_(s, s)])
        assert np.allclose(np.sort(np.maximum(ev, 0)), np.maximum(ref, 0), atol=1e-13), q


@pytest.mark.parametrize("model", ["spin", "hubbard"])
def test_entropy_is_number_plus_configurational(model):
    # S = S_N + S_c for every rho (the split of write_entanglement); a
    # dropped weight or a sign slip in S_N / S_c breaks the identity
    if model == "spin":
        d = O.DMRG(O.spin_site(3), O.Params(ldmrg=5, mdmrg=20, target_qn=(0.0,), hij=SPIN_HIJ, msweep=[20]))
    else:
        d = O.DMRG(O.electron_site(3.0, True), O.Params(ldmrg=4, mdmrg=24, target_qn=(0.5, 0.5), hij=HUB_HIJ,
                                                         msweep=[24]))
    d.finite()
    for r in d.records:
        assert abs(r.sn_left + r.sc_left - r.entropy_left) < 1e-12
        assert abs(r.sn_right + r.sc_right - r.entropy_right) < 1e-12
        tr = sum(float(np.sum(ev)) for _, ev in r.lam_left)
        assert abs(tr - 1.0) < 1e-12
    assert max(r.sn_left for r in d.records) > 1e-3 and max(r.sc_left for r in d.records) > 1e-3


def test_qntype_global_matches_exact_sector():
    # QNtype=global: the target at every length is min(L, int(q)); an
    # untruncated 6-site Hubbard run at N_up = N_dw = 2 equals the exact
    # ground energy of that sector (4 sites: min(4, 2) = 2 as well)
    dot = O.electron_site(4.0, True)
    p = O.Params(ldmrg=3, mdmrg=4096, target_qn=(2.0, 2.0), hij=HUB_HIJ, neigen=1, lanc_dim_threshold=4096,
                 qn_type="global")
    d = O.DMRG(dot, p)
    d.infinite()
    for r in d.records:
        ref = O.ed_ground(dot, r.L, np.array(HUB_HIJ), (2.0, 2.0))[0]
        assert abs(r.energies[0] - ref) < 1e-10, r.L


def test_qn_truncation_knobs():
    # QNtruncation_dim above every block dimension (or QNtruncation_error
    # above every block weight) sets every rho eigenvalue to tiny(1d0): the
    # entropies vanish and the kept states carry no weight
    for kw in (dict(qn_trunc_dim=10 ** 6), dict(qn_trunc_error=2.0)):
        p = O.Params(ldmrg=5, mdmrg=8, target_qn=(0.0,), hij=SPIN_HIJ, neigen=1, **kw)
        d = O.DMRG(O.spin_site(2), p)
        d.infinite()
        for r in d.records:
            assert abs(r.entropy_left) < 1e-300 and abs(r.sn_left) < 1e-300
            assert all(np.all(ev == O.TINY) for _, ev in r.lam_left)
    # a threshold below every block's weight changes nothing
    base = O.DMRG(O.spin_site(2), O.Params(ldmrg=5, mdmrg=8, target_qn=(0.0,), hij=SPIN_HIJ, neigen=1))
    base.infinite()
    d = O.DMRG(O.spin_site(2), O.Params(ldmrg=5, mdmrg=8, target_qn=(0.0,), hij=SPIN_HIJ, neigen=1,
                                        qn_trunc_error=1e-300))
    d.infinite()
    for a, b in zip(base.records, d.records):
        assert a.energies[0] == b.energies[0] and a.entropy_left == b.entropy_left

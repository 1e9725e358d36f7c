"""End-to-end GPU parity of the device-resident DMRG engine: the reference's
golden trajectories, the oracle's kept-state counts per QN, finite sweeps,
and the FP32-MFMA path with FP64 refinement (|dE0/L| <= 1e-9)."""
import os

import numpy as np
import pytest

import nssdmrg_amd as N
from oracle import dmrg_oracle as O

pytestmark = pytest.mark.gpu

SPIN_HIJ = [[1.0, 0.0], [0.0, 0.5]]
HUB_HIJ = [[-1.0, 0.0], [0.0, -1.0]]


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def run(ctx, kind="i", **kw):
    cfg = N.DMRGConfig(**kw)
    d = N.DMRG(ctx, cfg)
    steps = d.run(kind)
    return d, steps


# rho eigenvalues below this are rounding noise of rho's null space, whose
# basis (and so which null states the kept-state rule picks) is arbitrary in
# the reference too; parity of kept counts is asserted above it.
W_MIN = 1e-10


def _kept_oracle(rec, min_weight=0.0):
    out = {}
    for q, ws in rec.kept_w_left.items():
        c = sum(1 for w in ws if w > min_weight)
        if c:
            out[(float(q[0]), float(q[1]) if len(q) > 1 else 0.0)] = c
    return out


def test_spin_golden_L40_M20(ctx, golden_dir):
    ref = np.loadtxt(os.path.join(golden_dir, "spin12_L40_M20_energy.check"))
    refS = np.loadtxt(os.path.join(golden_dir, "spin12_L40_M20_Sentropy.check"))
    d, st = run(ctx, "i", site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=40, mdmrg=20,
                target_qn=(0.0, 0.0), neigen=2, lanc_dim_threshold=128)
    assert len(st) == 39
    for s, row in zip(st, ref):
        assert s["left_len"] == int(row[0])
        assert abs(s["e_per_site"][0] - row[1]) < 1e-10, (s["left_len"], s["e_per_site"][0] - row[1])
        assert abs(s["e_per_site"][1] - row[2]) < 1e-10
    for s, S in zip(st[:-1], refS[:-1]):
        assert abs(s["entropy_left"] - S) < 1e-9
    # kept-state counts per QN sector identical to the oracle's
    p = O.Params(ldmrg=40, mdmrg=20, target_qn=(0.0,), hij=SPIN_HIJ, neigen=2, lanc_dim_threshold=128)
    od = O.DMRG(O.spin_site(2), p)
    od.infinite()
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert s["m_left"] == r.m_left and s["m_right"] == r.m_right
        if s["renormalized"]:
            assert d.kept_qn(i, "left") == _kept_oracle(r)
            assert abs(s["trunc_left"] - r.trunc_left) < 1e-12


def test_hubbard_golden_L40_M40(ctx, golden_dir):
    ref = np.loadtxt(os.path.join(golden_dir, "hubbard_U0_L40_M40_energy.check"))
    refS = np.loadtxt(os.path.join(golden_dir, "hubbard_U0_L40_M40_Sentropy.check"))
    d, st = run(ctx, "i", site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=40, mdmrg=40,
                target_qn=(0.5, 0.5), neigen=1, lanc_dim_threshold=128, uloc=0.0, hfmode=True)
    assert len(st) == 39
    for s, e in zip(st, ref):
        assert abs(s["e_per_site"][0] - e) < 1e-10, (s["left_len"], s["e_per_site"][0] - e)
    for s, S in zip(st[:-1], refS[:-1]):
        assert abs(s["entropy_left"] - S) < 1e-8


def test_hubbard_U4_kept_counts_vs_oracle(ctx):
    p = O.Params(ldmrg=12, mdmrg=32, target_qn=(0.5, 0.5), hij=HUB_HIJ, neigen=1, lanc_dim_threshold=128)
    od = O.DMRG(O.electron_site(4.0, True), p)
    od.infinite()
    d, st = run(ctx, "i", site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=12, mdmrg=32,
                target_qn=(0.5, 0.5), neigen=1, lanc_dim_threshold=128, uloc=4.0, hfmode=True)
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert abs(s["energies"][0] - r.energies[0]) < 1e-10 * abs(r.energies[0])
        assert abs(s["entropy_left"] - r.entropy_left) < 1e-9
        assert s["m_sb"] == r.m_sb and s["nsect"] == r.nsect
        if s["renormalized"]:
            assert d.kept_qn(i, "left") == _kept_oracle(r)


@pytest.mark.parametrize("model", ["spin12", "spin1", "hubbard"])
def test_finite_sweeps_vs_oracle(ctx, model):
    if model == "hubbard":
        osite, site, hij, tq, U = O.electron_site(2.0, True), N.electron_site(), HUB_HIJ, (0.5, 0.5), 2.0
        L, M, ms = 6, 24, [24, 32]
    else:
        sun = 2 if model == "spin12" else 3
        osite, site, hij, tq, U = O.spin_site(sun), N.spin_site(sun), SPIN_HIJ, (0.0, 0.0), 0.0
        L, M, ms = (8, 16, [16, 24]) if sun == 2 else (6, 20, [20, 28])
    print(f"case finite {model}", flush=True)
    p = O.Params(ldmrg=L, mdmrg=M, target_qn=tq[:2] if model == "hubbard" else (0.0,), hij=hij, neigen=1,
                 lanc_dim_threshold=100, msweep=ms)
    od = O.DMRG(osite, p)
    od.finite()
    d, st = run(ctx, "f", site=site, hij=np.array(hij), ldmrg=L, mdmrg=M, msweep=ms, target_qn=tq,
                neigen=1, lanc_dim_threshold=100, uloc=U, hfmode=True)
    assert len(st) == len(od.records)
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert s["type"] == r.kind and s["L"] == r.L
        assert abs(s["energies"][0] - r.energies[0]) < 1e-10 * max(1.0, abs(r.energies[0])), i
        if s["renormalized"]:
            assert d.kept_qn(i, "left", W_MIN) == _kept_oracle(r, W_MIN), i
            if min(s["cut_left"], r.cut_left) > W_MIN:
                assert s["m_left"] == r.m_left, i


def test_finite_msweep_order_pinned_by_ed(ctx):
    # the engine's finite driver with Msweep = [4, 4096] on 10 sites against
    # exact diagonalisation (as the oracle's test_finite_msweep_order_pinned_by_ed):
    # sweep 1 (13 steps) keeps 4 states and stays visibly above the exact
    # energy, sweep 2 (12 steps) ends exact, and the last step of sweep 1
    # (left length Ldmrg+1, label 1) is not renormalised (8 = 4 x 2 states)
    print("case finite Msweep [4, 4096] vs ED", flush=True)
    d, st = run(ctx, "f", site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=4, mdmrg=4096, msweep=[4, 4096],
                target_qn=(0.0, 0.0), neigen=1, lanc_dim_threshold=16)
    ed10 = O.ed_ground(O.spin_site(2), 10, SPIN_HIJ, (0.0,))[0]
    fin = [s for s in st if s["type"] == "f"]
    assert len(fin) == 25
    s1, s2 = fin[:13], fin[13:]
    assert [s["left_len"] for s in s1] == [5, 6, 7, 8, 7, 6, 5, 4, 3, 2, 3, 4, 5]
    assert [s["left_len"] for s in s2] == [6, 7, 8, 7, 6, 5, 4, 3, 2, 3, 4, 5]
    assert all(s["energies"][0] >= ed10 - 1e-10 for s in s1)
    assert min(s["energies"][0] - ed10 for s in s1[2:]) > 1e-4
    assert not s1[-1]["renormalized"] and s1[-1]["m_left"] == 8
    for s in s2[-2:]:
        assert abs(s["energies"][0] - ed10) < 1e-10, (s["left_len"], s["energies"][0] - ed10)
    d.close()


@pytest.mark.parametrize("model,neigen", [("spin12", 1), ("spin12", 2), ("hubbard", 1)])
def test_finite_prediction_matches_oracle(ctx, model, neigen):
    # Lanczos started from the transformed previous ground state converges to
    # the same states as the reference's random start, with fewer H*v; with
    # two eigenpairs the engine keeps the random start (same H*v count).
    if model == "hubbard":
        osite, site, hij, tq, U = O.electron_site(2.0, True), N.electron_site(), HUB_HIJ, (0.5, 0.5), 2.0
        L, M, ms = 6, 32, [32, 40]
    else:
        osite, site, hij, tq, U = O.spin_site(2), N.spin_site(2), SPIN_HIJ, (0.0, 0.0), 0.0
        L, M, ms = 10, 24, [24, 32]
    print(f"case predict {model} neigen={neigen}", flush=True)
    p = O.Params(ldmrg=L, mdmrg=M, target_qn=tq[:2] if model == "hubbard" else (0.0,), hij=hij, neigen=neigen,
                 lanc_dim_threshold=32, msweep=ms)
    od = O.DMRG(osite, p)
    od.finite()
    kw = dict(site=site, hij=np.array(hij), ldmrg=L, mdmrg=M, msweep=ms, target_qn=tq, neigen=neigen,
              lanc_dim_threshold=32, uloc=U, hfmode=True)
    _, plain = run(ctx, "f", **kw)
    d, st = run(ctx, "f", predict=True, **kw)
    assert len(st) == len(od.records) == len(plain)
    # An excited state that rho does not target depends on which rho
    # null-space states the blocks kept (arbitrary, as in the reference);
    # E1 is compared while no renormalization so far has kept a state of
    # weight <= W_MIN.
    clean = True
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert abs(s["energies"][0] - r.energies[0]) < 1e-10 * max(1.0, abs(r.energies[0])), i
        if neigen > 1 and clean:
            assert abs(s["energies"][1] - r.energies[1]) < 1e-10 * max(1.0, abs(r.energies[1])), i
        if s["renormalized"] and min(s["cut_left"], s["cut_right"], r.cut_left, r.cut_right) <= W_MIN:
            clean = False
        assert abs(s["entropy_left"] - r.entropy_left) < 1e-8, i
        if s["renormalized"]:
            assert d.kept_qn(i, "left", W_MIN) == _kept_oracle(r, W_MIN), i
    fin = [i for i, s in enumerate(st) if s["type"] == "f" and s["numop"] > 0]
    assert len(fin) > 10
    n_pred = sum(st[i]["numop"] for i in fin)
    n_plain = sum(plain[i]["numop"] for i in fin)
    print(f"  H*v calls in finite steps: predicted {n_pred}, random start {n_plain}", flush=True)
    if neigen == 1:
        assert n_pred < n_plain
    else:
        assert n_pred == n_plain


SZ12 = np.diag([0.5, -0.5])
NUP = np.diag([0.0, 1.0, 0.0, 1.0])
NDW = np.diag([0.0, 0.0, 1.0, 1.0])


def test_measure_golden_hubbard_L40_M40(ctx, golden_dir):
    # hm_1d.f90: Measure_DMRG([dens, docc, s2z], pos=1..Ldmrg) after run_DMRG,
    # asserted to 1e-6 against n.check, d.check, s2z.check
    d, st = run(ctx, "i", site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=40, mdmrg=40,
                target_qn=(0.5, 0.5), neigen=1, lanc_dim_threshold=128, uloc=0.0, hfmode=True)
    ops = [NUP + NDW, NUP @ NDW, (NUP - NDW) @ (NUP - NDW)]
    avo = d.measure(ops, pos=np.arange(1, 41))
    for k, name in enumerate(["n", "d", "s2z"]):
        ref = np.loadtxt(os.path.join(golden_dir, f"hubbard_U0_L40_M40_{name}.check"))
        assert np.max(np.abs(avo[k] - ref)) < 1e-6, name


def test_measure_golden_spin_L40_M20(ctx, golden_dir):
    d, st = run(ctx, "i", site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=40, mdmrg=20,
                target_qn=(0.0, 0.0), neigen=2, lanc_dim_threshold=128)
    # the reference test's own tolerance (test/src/Spin1d_SU2/spinSU2_1d.f90:60)
    ref = np.loadtxt(os.path.join(golden_dir, "spin12_L40_M20_sz.check"))
    assert np.max(np.abs(d.measure(SZ12, pos=np.arange(1, 41)) - ref)) < 1e-8


@pytest.mark.parametrize("model,kind", [("spin12", "i"), ("spin12", "f"), ("hubbard", "i"), ("hubbard", "f")])
def test_measure_matches_ed_untruncated(ctx, model, kind):
    # no truncation: the measured site profile equals the exact one
    print(f"case measure {model} {kind}", flush=True)
    if model == "spin12":
        L = 5
        kw = dict(site=N.spin_site(2), hij=np.array(SPIN_HIJ), target_qn=(0.1, 0.0))
        osite, ops, target = O.spin_site(2), [SZ12], (1.0,)
    else:
        # half filling keeps the target int(L/2) consistent through the growth;
        # double occupancy and (n_up-n_dw)^2 have nontrivial OBC profiles
        L = 3 if kind == "i" else 2
        kw = dict(site=N.electron_site(), hij=np.array(HUB_HIJ), target_qn=(0.5, 0.5), uloc=4.0, hfmode=True)
        osite, ops, target = O.electron_site(4.0, True), [NUP @ NDW, (NUP - NDW) @ (NUP - NDW)], (3.0, 3.0)
    d, st = run(ctx, kind, ldmrg=L, mdmrg=64, msweep=[64], neigen=1, lanc_dim_threshold=64, **kw)
    nsites = st[-1]["L"]
    got = d.measure(ops)
    assert np.ptp(got) > 1e-3  # a nontrivial profile
    ref = O.ed_site_expectations(osite, nsites, np.array(HUB_HIJ if model == "hubbard" else SPIN_HIJ), target, ops)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) < 1e-9, np.max(np.abs(got - ref))


def test_measure_rejects_qn_breaking_operator(ctx):
    d, st = run(ctx, "i", site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=4, mdmrg=16, target_qn=(0.0, 0.0),
                neigen=1, lanc_dim_threshold=64)
    with pytest.raises(N.NssError):
        d.measure(np.array([[0.0, 1.0], [0.0, 0.0]]), pos=[2])


@pytest.mark.parametrize("model", ["spin12", "hubbard"])
def test_restart_fixture_last_step(ctx, golden_dir, model):
    # infinite_DMRG with block_left/right.restart present: the reference's
    # length-39 blocks, one step to L=40, against the final energy.check
    if model == "spin12":
        kw = dict(site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=40, mdmrg=20, target_qn=(0.0, 0.0), neigen=2,
                  lanc_dim_threshold=128)
        tag = "spin12_L40_M20"  # row: length, E0/L, E1/L, length
        ref = np.loadtxt(os.path.join(golden_dir, "spin12_L40_M20_final_energy.check"))[1:3]
    else:
        kw = dict(site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=40, mdmrg=40, target_qn=(0.5, 0.5),
                  neigen=1, lanc_dim_threshold=128, uloc=0.0, hfmode=True)
        tag = "hubbard_U0_L40_M40"
        ref = np.atleast_1d(np.loadtxt(os.path.join(golden_dir, "hubbard_U0_L40_M40_final_energy.check")))
    d = N.DMRG(ctx, N.DMRGConfig(**kw))
    d.load_blocks(os.path.join(golden_dir, f"{tag}_block_left.restart"),
                  os.path.join(golden_dir, f"{tag}_block_right.restart"))
    st = d.run("i")
    assert len(st) == 1 and st[0]["left_len"] == 40
    got = st[0]["e_per_site"][:len(ref)]
    assert np.max(np.abs(got - ref)) < 1e-10, (got, ref)
    # the files hold the rotation matrices of every length in the reference's
    # own basis order: the continued run measures the reference's profiles
    # (the reference tests' tolerances: spinSU2_1d.f90:60, hm_1d.f90:102-118)
    if model == "spin12":
        checks, tol = {"sz": [SZ12]}, 1e-8
    else:
        checks, tol = {"n": [NUP + NDW], "d": [NUP @ NDW], "s2z": [(NUP - NDW) @ (NUP - NDW)]}, 1e-6
    for name, ops in checks.items():
        prof = np.loadtxt(os.path.join(golden_dir, f"{tag}_{name}.check"))
        assert np.max(np.abs(d.measure(ops, pos=np.arange(1, 41))[0] - prof)) < tol, name


@pytest.mark.parametrize("model", ["spin12", "hubbard"])
def test_restart_round_trip(ctx, tmp_path, model):
    # write_block of the blocks entering the last step, read_block into a new
    # run: the same last step again
    if model == "spin12":
        kw = dict(site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=10, mdmrg=16, target_qn=(0.0, 0.0), neigen=1,
                  lanc_dim_threshold=64)
    else:
        kw = dict(site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=6, mdmrg=24, target_qn=(0.5, 0.5), neigen=1,
                  lanc_dim_threshold=64, uloc=3.0, hfmode=True)
    d1, st1 = run(ctx, "i", **kw)
    lp, rp = str(tmp_path / "block_left.restart"), str(tmp_path / "block_right.restart")
    d1.save_blocks(lp, rp)
    d2 = N.DMRG(ctx, N.DMRGConfig(**kw))
    d2.load_blocks(lp, rp)
    st2 = d2.run("i")
    assert len(st2) == 1
    assert abs(st2[0]["energies"][0] - st1[-1]["energies"][0]) < 1e-11
    assert abs(st2[0]["entropy_left"] - st1[-1]["entropy_left"]) < 1e-10
    assert st2[0]["m_sb"] == st1[-1]["m_sb"]
    # the files carry the rotation matrices of every length (omatrices), so
    # the restarted run measures: the same profile as the original run
    if model == "spin12":
        ops = [np.diag([0.5, -0.5])]
    else:
        ops = [np.diag([0.0, 1.0, 1.0, 2.0]), np.diag([0.0, 0.0, 0.0, 1.0])]
    m1 = d1.measure(ops)
    m2 = d2.measure(ops)
    assert m1.shape == m2.shape
    assert np.abs(m1 - m2).max() < 1e-9, np.abs(m1 - m2).max()
    # written again from the restarted run, the history survives a second trip
    d2.save_blocks(lp, rp)
    d3 = N.DMRG(ctx, N.DMRGConfig(**kw))
    d3.load_blocks(lp, rp)
    d3.run("i")
    assert np.abs(d3.measure(ops) - m1).max() < 1e-9


def test_restart_file_with_unfitting_rotation_matrices(ctx, tmp_path):
    # rotation matrices that do not fit the block (every row of matrix "2"
    # of the left file emptied, so its kept states have no components): the
    # block loads without its history, continues the run as before, and
    # cannot be measured on that side
    kw = dict(site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=10, mdmrg=16, target_qn=(0.0, 0.0), neigen=1,
              lanc_dim_threshold=64)
    d1, st1 = run(ctx, "i", **kw)
    lp, rp = str(tmp_path / "block_left.restart"), str(tmp_path / "block_right.restart")
    d1.save_blocks(lp, rp)
    lines = open(lp).read().split("\n")
    k = next(i for i in range(len(lines) - 1) if lines[i].strip() == "2" and lines[i + 1].strip() == "none")
    nr = int(lines[k + 2].split()[0])
    r, body = k + 3, []  # each row: its nonzero count, then one line per entry
    for _ in range(nr):
        r += 1 + int(lines[r])
        body.append("0")
    lines[k + 3:r] = body
    open(lp, "w").write("\n".join(lines))
    d2 = N.DMRG(ctx, N.DMRGConfig(**kw))
    d2.load_blocks(lp, rp)
    st2 = d2.run("i")
    assert abs(st2[0]["energies"][0] - st1[-1]["energies"][0]) < 1e-11
    with pytest.raises(N.NssError):
        d2.measure(np.diag([0.5, -0.5]), pos=[1])
    assert np.isfinite(d2.measure(np.diag([0.5, -0.5]), pos=[st2[0]["L"]])).all()


def test_mixed_precision_refined_matches_fp64(ctx, monkeypatch):
    # every superblock through the mixed solver (the engine's default sends
    # superblocks below 2^17 states to the FP64 solver)
    monkeypatch.setenv("NSS_MIXED_MIN_N", "0")
    kw = dict(site=N.spin_site(3), hij=np.array(SPIN_HIJ), ldmrg=10, mdmrg=48, msweep=[48],
              target_qn=(0.0, 0.0), neigen=1, lanc_dim_threshold=128)
    _, s64 = run(ctx, "f", **kw)
    _, s32 = run(ctx, "f", precision="mixed", **kw)
    assert len(s64) == len(s32)
    for a, b in zip(s64, s32):
        assert abs(a["e_per_site"][0] - b["e_per_site"][0]) <= 1e-9
        if min(a["cut_left"], b["cut_left"]) > W_MIN:
            assert a["m_left"] == b["m_left"]
    assert sum(b["numop_fp32"] for b in s32) > 0


def test_engine_superblock_spectrum_matches_oracle(ctx):
    # the engine's last superblock (its own basis order) has the oracle's spectrum
    p = O.Params(ldmrg=5, mdmrg=12, target_qn=(0.5, 0.5), hij=HUB_HIJ, neigen=1, lanc_dim_threshold=64)
    od = O.DMRG(O.electron_site(3.0, True), p)
    od.infinite()
    d, st = run(ctx, "i", site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=5, mdmrg=12,
                target_qn=(0.5, 0.5), neigen=1, lanc_dim_threshold=64, uloc=3.0, hfmode=True)
    ex = d.last_superblock()
    sb = O.Superblock((0.0,), list(range(len(ex["Dl"]))), ex["Dl"].astype(np.int64), ex["Dr"].astype(np.int64),
                      np.concatenate([[0], np.cumsum(ex["Dl"].astype(np.int64) * ex["Dr"])[:-1]]), [], [])
    sb.HL, sb.HR = ex["HL"], ex["HR"]
    sb.terms = [(int(a), int(b), A, B, 0) for a, b, A, B in zip(ex["dst"], ex["src"], ex["A"], ex["B"])]
    e_eng = np.linalg.eigvalsh(O.sb_dense(sb))
    e_ora = np.linalg.eigvalsh(O.sb_dense(od.last_sb))
    assert np.allclose(e_eng[:20], e_ora[:20], atol=1e-10)
    # and the engine's GPU plan applies exactly that matrix
    v = O.splitmix64_uniform_fast(sb.n, 2)
    assert np.linalg.norm(ctx.hv(v) - O.hv_direct(sb, v)) < 1e-12 * np.linalg.norm(O.hv_direct(sb, v))


def _batch_cases():
    # chains of one batch: different models, m, U and fillings, equal LDMRG
    return [
        dict(kind="hubbard", U=2.0, tq=(0.5, 0.5), M=32, ms=[32, 40]),
        dict(kind="hubbard", U=6.0, tq=(0.375, 0.375), M=24, ms=[24, 28]),
        dict(kind="spin", U=0.0, tq=(0.0, 0.0), M=20, ms=[20, 24]),
        dict(kind="hubbard", U=0.5, tq=(0.5, 0.5), M=40, ms=[40, 48]),
    ]


@pytest.mark.parametrize("predict", [False, True])
def test_batch_matches_oracle_and_single_runs(ctx, predict):
    # Independent chains stepped together with one batched Lanczos per step:
    # each chain matches the oracle (E to 1e-10 relative, kept counts per QN
    # above W_MIN, entropy) and its own standalone run.
    cases = _batch_cases()
    L = 6
    print(f"case batch predict={predict} chains={len(cases)}", flush=True)
    kws, oracles = [], []
    for c in cases:
        if c["kind"] == "hubbard":
            site, osite, hij, tq = N.electron_site(), O.electron_site(c["U"], True), HUB_HIJ, c["tq"]
            otq = tq
        else:
            site, osite, hij, tq = N.spin_site(2), O.spin_site(2), SPIN_HIJ, c["tq"]
            otq = (0.0,)
        kws.append(dict(site=site, hij=np.array(hij), ldmrg=L, mdmrg=c["M"], msweep=c["ms"], target_qn=tq,
                        neigen=1, lanc_dim_threshold=32, uloc=c["U"], hfmode=True, predict=predict))
        p = O.Params(ldmrg=L, mdmrg=c["M"], target_qn=otq, hij=hij, neigen=1, lanc_dim_threshold=32,
                     msweep=c["ms"])
        od = O.DMRG(osite, p)
        od.finite()
        oracles.append(od)
    singles = [run(ctx, "f", **kw)[1] for kw in kws]
    engines = [N.DMRG(ctx, N.DMRGConfig(**kw)) for kw in kws]
    batch = N.DMRGBatch(engines)
    res = batch.run("f")
    for ci, (st, single, od, d) in enumerate(zip(res, singles, oracles, engines)):
        assert len(st) == len(single) == len(od.records), ci
        nlz = 0
        for i, (s, s1, r) in enumerate(zip(st, single, od.records)):
            e = r.energies[0]
            assert abs(s["energies"][0] - e) < 1e-10 * max(1.0, abs(e)), (ci, i)
            assert abs(s["energies"][0] - s1["energies"][0]) < 1e-10 * max(1.0, abs(e)), (ci, i)
            assert abs(s["entropy_left"] - r.entropy_left) < 1e-8, (ci, i)
            if s["renormalized"]:
                # total kept counts include rho null-space states (arbitrary
                # in the reference too): compared only when every kept
                # state carries weight above W_MIN
                if min(s["cut_left"], s["cut_right"], s1["cut_left"], s1["cut_right"]) > W_MIN:
                    assert s["m_left"] == s1["m_left"] and s["m_right"] == s1["m_right"], (ci, i)
                assert d.kept_qn(i, "left", W_MIN) == _kept_oracle(r, W_MIN), (ci, i)
            nlz += s["numop"] > 0
        assert nlz > 10, ci


def test_batch_rejects_mismatched_engines(ctx):
    a = N.DMRG(ctx, N.DMRGConfig(site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=4, mdmrg=8,
                                 target_qn=(0.0, 0.0)))
    b = N.DMRG(ctx, N.DMRGConfig(site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=5, mdmrg=8,
                                 target_qn=(0.0, 0.0)))
    with pytest.raises(N.NssError, match="LDMRG"):
        N.DMRGBatch([a, b]).run("i")
    with pytest.raises(N.NssError, match="twice"):
        N.DMRGBatch([a, a]).run("i")


# --------------------------------------------------------------------------- PBC
# Periodic rings (PBCdmrg): the infinite algorithm alternates link-'n' and
# link-'p' steps; the oracle's PBC path is pinned by ring closed forms and a
# Fock-space ED (tests/test_oracle.py).
def _pbc_case(model):
    if model == "hubbard":
        return O.electron_site(3.0, True), N.electron_site(), HUB_HIJ, (0.5, 0.5), 3.0, 24
    if model == "spin1":
        return O.spin_site(3), N.spin_site(3), SPIN_HIJ, (0.0, 0.0), 0.0, 20
    return O.spin_site(2), N.spin_site(2), SPIN_HIJ, (0.0, 0.0), 0.0, 16


@pytest.mark.parametrize("model", ["spin12", "spin1", "hubbard"])
def test_pbc_infinite_matches_oracle(ctx, model):
    osite, site, hij, tq, U, M = _pbc_case(model)
    print(f"case pbc infinite {model}", flush=True)
    L = 7
    p = O.Params(ldmrg=L, mdmrg=M, target_qn=tq[:2] if model == "hubbard" else (0.0,), hij=hij, neigen=1,
                 lanc_dim_threshold=64, pbc=True)
    od = O.DMRG(osite, p)
    od.infinite()
    d, st = run(ctx, "i", site=site, hij=np.array(hij), ldmrg=L, mdmrg=M, target_qn=tq, neigen=1,
                lanc_dim_threshold=64, uloc=U, hfmode=True, pbc=True)
    assert len(st) == len(od.records)
    assert [s["link"] for s in st] == ["n", "p"] * (len(st) // 2)
    assert any(s["numop"] > 0 for s in st)           # the Lanczos path ran
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert s["L"] == r.L and s["m_sb"] == r.m_sb and s["nsect"] == r.nsect, i
        assert abs(s["energies"][0] - r.energies[0]) < 1e-10 * max(1.0, abs(r.energies[0])), i
        assert abs(s["entropy_left"] - r.entropy_left) < 1e-9, i
        if s["renormalized"]:
            assert d.kept_qn(i, "left", W_MIN) == _kept_oracle(r, W_MIN), i
            if min(s["cut_left"], r.cut_left) > W_MIN:
                assert s["m_left"] == r.m_left, i


def test_pbc_ring_closed_forms(ctx):
    # untruncated rings: spin-1/2 Heisenberg E0(4) = -2, E0(6) = -2.8027756377319946;
    # free fermions (U=0) on periodic rings of 4 and 6 sites at half filling: -4, -8
    print("case pbc closed forms", flush=True)
    _, st = run(ctx, "i", site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=3, mdmrg=200, target_qn=(0.0, 0.0),
                neigen=1, lanc_dim_threshold=8, pbc=True)
    assert [s["L"] for s in st] == [4, 6]
    assert abs(st[0]["energies"][0] + 2.0) < 1e-11 and abs(st[1]["energies"][0] + 2.8027756377319946) < 1e-11
    _, st = run(ctx, "i", site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=3, mdmrg=400,
                target_qn=(0.5, 0.5), neigen=1, lanc_dim_threshold=16, uloc=0.0, hfmode=True, pbc=True)
    assert abs(st[0]["energies"][0] + 4.0) < 1e-11 and abs(st[1]["energies"][0] + 8.0) < 1e-11


def test_pbc_refusals_and_odd_ldmrg(ctx):
    # an even LDMRG becomes odd (INPUT_VARS.f90:178-183); finite_DMRG and
    # Measure_DMRG refuse PBC runs
    print("case pbc refusals", flush=True)
    kw = dict(site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=4, mdmrg=8, msweep=[8], target_qn=(0.0, 0.0),
              neigen=1, lanc_dim_threshold=64, pbc=True)
    d, st = run(ctx, "i", **kw)
    assert st[-1]["left_len"] == 5 and len(st) == 4
    with pytest.raises(N.NssError, match="PBC"):
        d.measure(np.diag([0.5, -0.5]), [1])
    d2 = N.DMRG(ctx, N.DMRGConfig(**kw))
    with pytest.raises(N.NssError, match="PBC"):
        d2.run("f")


def test_pbc_batch_matches_single_runs(ctx):
    print("case pbc batch", flush=True)
    cases = [dict(site=N.electron_site(), hij=np.array(HUB_HIJ), uloc=2.0, target_qn=(0.5, 0.5), mdmrg=24),
             dict(site=N.spin_site(2), hij=np.array(SPIN_HIJ), uloc=0.0, target_qn=(0.0, 0.0), mdmrg=16)]
    common = dict(ldmrg=7, neigen=1, lanc_dim_threshold=64, hfmode=True, pbc=True)
    single = [run(ctx, "i", **c, **common)[1] for c in cases]
    engines = [N.DMRG(ctx, N.DMRGConfig(**c, **common)) for c in cases]
    batched = N.DMRGBatch(engines).run("i")
    for sb, ss in zip(batched, single):
        assert len(sb) == len(ss)
        for a, b in zip(sb, ss):
            assert a["link"] == b["link"] and a["m_sb"] == b["m_sb"]
            assert abs(a["energies"][0] - b["energies"][0]) < 1e-10 * max(1.0, abs(b["energies"][0]))
            assert abs(a["entropy_left"] - b["entropy_left"]) < 1e-8


def test_large_superblocks_partial_reorth_free_fermions(ctx):
    # single-engine FP64 steps with >= 2^17 superblock states use the Lanczos
    # with partial reorthogonalisation; U=0 finite sweeps must reach the
    # free-fermion ground energy of the open chain
    print("case large superblocks, partial reorthogonalisation", flush=True)
    L, M = 16, 256
    d, st = run(ctx, "f", site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=(L - 2) // 2, mdmrg=M,
                msweep=[M, M], target_qn=(0.5, 0.5), neigen=1, lanc_dim_threshold=128, uloc=0.0, hfmode=True,
                predict=True)
    assert max(s["m_sb"] for s in st) >= 1 << 17
    e = O.free_fermion_obc(L, L // 2, L // 2)
    fin = [s for s in st if s["type"] == "f" and s["L"] == L]
    assert abs(fin[-1]["energies"][0] - e) < 1e-8 * abs(e)
    assert all(abs(s["energies"][0] - e) < 1e-6 * abs(e) for s in fin[len(fin) // 2:])


@pytest.mark.parametrize("model", ["spin12", "hubbard"])
def test_pbc_restart_files_carry_both_ends(ctx, tmp_path, model):
    # write_block under PBC writes the operators of both block ends (okey's
    # _n and _p keys); a PBC run reads them back and continues; a block file
    # of an OBC run lacks the outer-end operators and a PBC run refuses it
    print(f"case pbc restart {model}", flush=True)
    if model == "spin12":
        kw = dict(site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=5, mdmrg=16, target_qn=(0.0, 0.0), neigen=1,
                  lanc_dim_threshold=64)
        keys = ["S_z_n", "S_p_n", "S_z_p", "S_p_p"]
    else:
        kw = dict(site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=5, mdmrg=24, target_qn=(0.5, 0.5),
                  neigen=1, lanc_dim_threshold=64, uloc=3.0, hfmode=True)
        keys = ["C_l1_s1_n", "C_l1_s2_n", "P_n", "C_l1_s1_p", "C_l1_s2_p", "P_p"]
    d1, st1 = run(ctx, "i", pbc=True, **kw)
    lp, rp = str(tmp_path / "pbc_left.restart"), str(tmp_path / "pbc_right.restart")
    d1.save_blocks(lp, rp)
    for path in (lp, rp):
        toks = open(path).read().split()
        for k in keys:
            assert k in toks, (path, k)
    d2 = N.DMRG(ctx, N.DMRGConfig(pbc=True, **kw))
    d2.load_blocks(lp, rp)
    st2 = d2.run("i")
    # the saved blocks entered the last (link-'p') step; the resumed run
    # starts a new iteration from them, as the reference's loop does
    assert [s["link"] for s in st2] == ["n", "p"] and st2[0]["left_len"] == st1[-1]["left_len"]
    assert all(np.isfinite(s["energies"][0]) for s in st2)
    d3, _ = run(ctx, "i", **kw)
    op, orp = str(tmp_path / "obc_left.restart"), str(tmp_path / "obc_right.restart")
    d3.save_blocks(op, orp)
    d4 = N.DMRG(ctx, N.DMRGConfig(pbc=True, **kw))
    with pytest.raises(N.NssError, match="operator missing"):
        d4.load_blocks(op, orp)


# ------------------------------------------------------------ entanglement split, QN options
@pytest.mark.parametrize("model", ["spin1", "hubbard"])
def test_entanglement_split_and_lambdas_vs_oracle(ctx, model):
    # write_entanglement's S_N, S_c and rho%lambda() per QN block against the
    # oracle, step by step (infinite and finite steps)
    print(f"case entanglement split {model}", flush=True)
    if model == "hubbard":
        osite, site, hij, tq, U, L, M = O.electron_site(3.0, True), N.electron_site(), HUB_HIJ, (0.5, 0.5), 3.0, 5, 40
    else:
        osite, site, hij, tq, U, L, M = O.spin_site(3), N.spin_site(3), SPIN_HIJ, (0.0, 0.0), 0.0, 6, 30
    p = O.Params(ldmrg=L, mdmrg=M, target_qn=tq[:2] if model == "hubbard" else (0.0,), hij=hij, neigen=1,
                 lanc_dim_threshold=64, msweep=[M])
    od = O.DMRG(osite, p)
    od.finite()
    d, st = run(ctx, "f", site=site, hij=np.array(hij), ldmrg=L, mdmrg=M, msweep=[M], target_qn=tq, neigen=1,
                lanc_dim_threshold=64, uloc=U, hfmode=True)
    assert len(st) == len(od.records)
    clean = True  # no kept state so far at rounding level (rho null space)
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert abs(s["sn_left"] - r.sn_left) < 1e-9 and abs(s["sc_left"] - r.sc_left) < 1e-9, i
        assert abs(s["sn_right"] - r.sn_right) < 1e-9 and abs(s["sc_right"] - r.sc_right) < 1e-9, i
        # per QN block (the engine's sector order is its own)
        got = {(q[0], q[1]): v for q, v in d.lambdas(i, "left")}
        ref = {(float(q[0]), float(q[1]) if len(q) > 1 else 0.0): v for q, v in r.lam_left}
        assert set(got) == set(ref), i
        for q, rv in ref.items():
            gv = got[q]
            # once a cut has kept null-space states, which ones it kept (and
            # so the later blocks' dimensions) is arbitrary, in the reference
            # too: from then on only the spectrum above W_MIN is compared
            if clean:
                assert len(gv) == len(rv), (i, q)
            a, b = np.sort(gv[gv > W_MIN]), np.sort(rv[rv > W_MIN])
            assert len(a) == len(b) and np.allclose(a, b, rtol=1e-8, atol=1e-14), (i, q)
        if s["renormalized"] and min(s["cut_left"], s["cut_right"]) <= W_MIN:
            clean = False


def test_qntype_global_and_qn_truncation_vs_oracle(ctx):
    print("case QNtype=global, QN truncation knobs", flush=True)
    # global targets: N_up = N_dw = min(L, 3) at every length, untruncated (a
    # truncated run keeps rho null-space states of the early rank-deficient
    # steps, whose basis is arbitrary).  As in the reference, the states of
    # a block sector with no complement in the superblock have no rho block
    # and are dropped (4 sites at N_up = N_dw = 3: 9 of 16), so from 6 sites
    # on the run follows the oracle, not the exact sector ground state
    dot = O.electron_site(2.0, True)
    p = O.Params(ldmrg=5, mdmrg=4096, target_qn=(3.0, 3.0), hij=HUB_HIJ, neigen=1, lanc_dim_threshold=64,
                 qn_type="global")
    od = O.DMRG(dot, p)
    od.infinite()
    d, st = run(ctx, "i", site=N.electron_site(), hij=np.array(HUB_HIJ), ldmrg=5, mdmrg=4096,
                target_qn=(3.0, 3.0), qn_type="global", neigen=1, lanc_dim_threshold=64, uloc=2.0, hfmode=True)
    assert [s["L"] for s in st] == [r.L for r in od.records] == [4, 6, 8, 10]
    assert abs(st[0]["energies"][0] - O.ed_ground(dot, 4, np.array(HUB_HIJ), (3.0, 3.0))[0]) < 1e-10
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert s["m_sb"] == r.m_sb and s["m_left"] == r.m_left, i
        assert abs(s["energies"][0] - r.energies[0]) < 1e-10 * max(1.0, abs(r.energies[0])), i
    # QNtruncation_dim: QN blocks of fewer than 3 states weigh tiny(1d0)
    p = O.Params(ldmrg=6, mdmrg=16, target_qn=(0.0,), hij=SPIN_HIJ, neigen=1, lanc_dim_threshold=64,
                 qn_trunc_dim=3)
    od = O.DMRG(O.spin_site(2), p)
    od.infinite()
    d, st = run(ctx, "i", site=N.spin_site(2), hij=np.array(SPIN_HIJ), ldmrg=6, mdmrg=16, target_qn=(0.0, 0.0),
                neigen=1, lanc_dim_threshold=64, qn_trunc_dim=3)
    for i, (s, r) in enumerate(zip(st, od.records)):
        assert abs(s["energies"][0] - r.energies[0]) < 1e-10 * max(1.0, abs(r.energies[0])), i
        assert abs(s["entropy_left"] - r.entropy_left) < 1e-9, i
        small = [ev for q, ev in d.lambdas(i, "left") if len(ev) < 3]
        assert all(np.all(ev == np.finfo(np.float64).tiny) for ev in small), i

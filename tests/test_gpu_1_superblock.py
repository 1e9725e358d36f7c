"""GPU parity of the low-level drop-ins (through the C ABI) against the oracle:
H_sb.psi (spMatVec_direct_main), Lanczos (sp_eigh), rho + eigh (sb_get_rdm),
block rotation (operators_renormalization).  FP64 bar: relative 1e-12 on
H.v, 1e-11 on eigenvalues; FP32-MFMA bar: relative 1e-5 on H.v."""
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


def oracle_sb(model, steps, M):
    if model == "spin":
        dot, hij, tq = O.spin_site(2), SPIN_HIJ, (0.0,)
    elif model == "spin1":
        dot, hij, tq = O.spin_site(3), SPIN_HIJ, (0.0,)
    else:
        dot, hij, tq = O.electron_site(4.0, True), HUB_HIJ, (0.5, 0.5)
    p = O.Params(ldmrg=steps + 2, mdmrg=M, target_qn=tq, hij=hij, neigen=1, lanc_dim_threshold=64)
    d = O.DMRG(dot, p)
    for _ in range(steps):
        d.step("i")
    left = O.enlarge_block(d.left, dot, "left", hij)
    right = O.enlarge_block(d.right, dot, "right", hij)
    sb = O.sb_get_states(left, right, O.current_target(left.length + right.length, tq))
    O.setup_direct(sb, left, right, hij)
    return sb


def upload(ctx, sb, layout="C", precision="fp64"):
    ctx.set_layout(sb.Dl, sb.Dr, sb.offset)
    ctx.upload_blocks(sb.HL, sb.HR, [t[0] for t in sb.terms], [t[1] for t in sb.terms],
                      [t[2] for t in sb.terms], [t[3] for t in sb.terms], layout=layout, precision=precision)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("model,steps,M", [("spin", 6, 20), ("spin1", 4, 30), ("hubbard", 3, 40)])
@pytest.mark.parametrize("layout", ["C", "F"])
def test_hv_matches_oracle(ctx, model, steps, M, layout):
    sb = oracle_sb(model, steps, M)
    print(f"case {model} n={sb.n} nq={len(sb.qns)} terms={len(sb.terms)} layout={layout}", flush=True)
    upload(ctx, sb, layout)
    for seed in (0, 1):
        v = O.splitmix64_uniform_fast(sb.n, seed)
        assert rel(ctx.hv(v), O.hv_direct(sb, v)) < 1e-12
    assert abs(ctx.flops() - O.hv_flops(sb)) <= 1e-9 * O.hv_flops(sb)


def test_hv_tiny_matches_reference_loops(ctx):
    sb = oracle_sb("spin", 2, 8)
    upload(ctx, sb)
    v = O.splitmix64_uniform_fast(sb.n, 3)
    assert rel(ctx.hv(v), O.hv_direct_loops(sb, v)) < 1e-13


def _random_sb(seed, dims, nterms):
    rng = np.random.default_rng(seed)
    Dl = np.array([d[0] for d in dims])
    Dr = np.array([d[1] for d in dims])
    nq = len(dims)
    sb = O.Superblock((0.0,), list(range(nq)), Dl, Dr,
                      np.concatenate([[0], np.cumsum(Dl * Dr)[:-1]]).astype(np.int64), [], [])
    sb.HL = [(lambda a: a + a.T)(rng.uniform(-1, 1, (a, a))) for a in Dl]
    sb.HR = [(lambda a: a + a.T)(rng.uniform(-1, 1, (b, b))) for b in Dr]
    sb.terms = []
    for _ in range(nterms):
        d, s = rng.integers(0, nq, 2)
        sb.terms.append((int(d), int(s), rng.uniform(-1, 1, (Dl[d], Dl[s])), rng.uniform(-1, 1, (Dr[d], Dr[s])), 0))
    return sb


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hv_uneven_sectors(ctx, seed):
    dims = [(1, 1), (1, 70), (130, 3), (65, 65), (200, 17), (64, 64), (3, 129), (97, 250)]
    sb = _random_sb(seed, dims, 14)
    print(f"case uneven seed={seed} n={sb.n}", flush=True)
    upload(ctx, sb)
    v = O.splitmix64_uniform_fast(sb.n, 11 + seed)
    ref = O.hv_direct(sb, v)
    assert rel(ctx.hv(v), ref) < 1e-12
    upload(ctx, sb, precision="fp32")
    assert rel(ctx.hv(v), ref) < 1e-12
    assert rel(ctx.hv(v, fp32=True), ref) < 1e-5


@pytest.mark.parametrize("split", ["auto", "1", "3", "8"])
@pytest.mark.parametrize("layout", ["C", "F"])
def test_hv_128_tile_kernel(ctx, split, layout, monkeypatch):
    # the 128x128 kernel forced on sectors around and across its tile edge,
    # long and short K chains, both operand layouts, K chains split into
    # 1, 3, 8 parts or the size-based count; FP64 against the oracle,
    # bitwise repeatable, and the FP32-MFMA plan at its bar
    monkeypatch.setenv("NSS_GEMM_TILE", "128")
    if split == "auto":
        monkeypatch.delenv("NSS_GEMM_SPLIT", raising=False)
    else:
        monkeypatch.setenv("NSS_GEMM_SPLIT", split)
    dims = [(600, 40), (40, 600), (300, 300), (130, 129), (1, 1), (128, 128), (17, 255), (256, 3)]
    sb = _random_sb(7, dims, 16)
    print(f"case t128 split={split} layout={layout} n={sb.n}", flush=True)
    upload(ctx, sb, layout)
    v = O.splitmix64_uniform_fast(sb.n, 21)
    ref = O.hv_direct(sb, v)
    y1 = ctx.hv(v)
    assert rel(y1, ref) < 1e-12
    assert np.array_equal(y1, ctx.hv(v))
    upload(ctx, sb, layout, precision="fp32")
    assert rel(ctx.hv(v, fp32=True), ref) < 1e-5


def test_hv_linearity_and_symmetry(ctx):
    sb = oracle_sb("hubbard", 3, 24)
    upload(ctx, sb)
    u = O.splitmix64_uniform_fast(sb.n, 5)
    v = O.splitmix64_uniform_fast(sb.n, 6)
    hu, hv = ctx.hv(u), ctx.hv(v)
    assert abs(np.dot(u, hv) - np.dot(hu, v)) < 1e-11 * np.linalg.norm(hu) * np.linalg.norm(v)
    assert rel(ctx.hv(2.0 * u - 3.0 * v), 2.0 * hu - 3.0 * hv) < 1e-13
    # no leakage: a vector supported on one sector maps into connected sectors only
    e = np.zeros(sb.n)
    e[sb.offset[0]:sb.offset[0] + sb.Dl[0] * sb.Dr[0]] = 1.0
    assert rel(ctx.hv(e), O.hv_direct(sb, e)) < 1e-13


def test_hv_error_paths(ctx):
    sb = oracle_sb("spin", 3, 8)
    upload(ctx, sb)
    with pytest.raises(N.NssError, match="NSS_E_ARG"):
        ctx.hv(np.zeros(sb.n + 1))
    with pytest.raises(N.NssError, match="NSS_E_ARG"):
        ctx.set_layout(sb.Dl, sb.Dr, sb.offset + 1)
    assert np.isfinite(ctx.hv(np.ones(sb.n))).all()   # a rejected layout leaves the state intact
    ctx.set_layout(sb.Dl, sb.Dr, sb.offset)
    with pytest.raises(N.NssError, match="NSS_E_STATE"):
        ctx.hv(np.zeros(sb.n))     # new layout, no blocks yet


@pytest.mark.parametrize("model,neig", [("spin", 2), ("hubbard", 1), ("spin1", 2)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lanczos_matches_dense(ctx, model, neig, precision):
    sb = oracle_sb(model, 4 if model != "spin1" else 3, 24)
    print(f"case lanczos {model} n={sb.n} {precision}", flush=True)
    upload(ctx, sb, precision=precision)
    hd = O.sb_dense(sb)
    ex = np.linalg.eigvalsh(hd)[:neig]
    ev, x, nop = ctx.lanczos(neig, kmax=60, tol=1e-12, seed=0, precision=precision)
    assert np.allclose(ev, ex, atol=1e-11 * max(1, abs(ex[0])))
    for i in range(neig):
        r = np.linalg.norm(hd @ x[i] - ev[i] * x[i])
        assert r < 1e-8 * max(1, abs(ev[i]))
        assert abs(np.linalg.norm(x[i]) - 1.0) < 1e-12
    assert nop > 0


def test_lanczos_restarts_when_kmax_small(ctx):
    sb = oracle_sb("spin", 6, 20)
    upload(ctx, sb)
    ex = np.linalg.eigvalsh(O.sb_dense(sb))[:2]
    ev, _, nop = ctx.lanczos(2, kmax=12, tol=1e-12, seed=3)
    assert np.allclose(ev, ex, atol=1e-11)
    assert nop > 12


def test_lanczos_constant_start(ctx):
    sb = oracle_sb("hubbard", 3, 16)
    upload(ctx, sb)
    ex = np.linalg.eigvalsh(O.sb_dense(sb))[0]
    ev, _, _ = ctx.lanczos(1, kmax=80, v0_mode=1)
    assert abs(ev[0] - ex) < 1e-11 * abs(ex)


@pytest.mark.parametrize("model", ["spin", "hubbard"])
def test_rdm_eigh_matches_oracle(ctx, model):
    sb = oracle_sb(model, 4, 20)
    upload(ctx, sb)
    e, u = np.linalg.eigh(O.sb_dense(sb))
    psi = u[:, 0]
    for side in ("left", "right"):
        evs, vecs = ctx.rdm_eigh(psi, side)
        ref = O.build_rho(sb, psi, side)
        for ev, vec, rev, k in zip(evs, vecs, ref.evals_blk, range(len(evs))):
            assert np.allclose(ev, rev, atol=1e-14)
            dl, dr, off = sb.Dl[k], sb.Dr[k], sb.offset[k]
            p = psi[off:off + dl * dr].reshape(dl, dr)
            rho = p @ p.T if side == "left" else p.T @ p
            assert np.allclose(vec @ np.diag(ev) @ vec.T, rho, atol=1e-13)


def test_rotate_blocks(ctx):
    rng = np.random.default_rng(0)
    shapes = [(5, 5, 3, 3), (70, 40, 20, 33), (1, 1, 1, 1), (130, 66, 64, 65)]
    O_ = [rng.standard_normal((r, c)) for r, c, _, _ in shapes]
    Ur = [np.linalg.qr(rng.standard_normal((r, m)))[0] for r, _, m, _ in shapes]
    Uc = [np.linalg.qr(rng.standard_normal((c, m)))[0] for _, c, _, m in shapes]
    out = ctx.rotate_blocks(O_, Ur, Uc)
    for o, a, b, r in zip(O_, Ur, Uc, out):
        assert np.allclose(r, a.T @ o @ b, atol=1e-12)


def test_hv_stats_count_calls(ctx):
    sb = oracle_sb("spin", 4, 16)
    upload(ctx, sb)
    ctx.hv_stats(reset=True)
    v = O.splitmix64_uniform_fast(sb.n, 1)
    for _ in range(8):
        ctx.hv(v)
    s = ctx.hv_stats(reset=True)
    # every call counted, every 4th timed (the default period)
    assert s["calls_all"] == 8 and s["calls"] == 2 and s["t_w"] > 0 and s["t_y"] > 0
    assert abs(s["flops"] - 2 * O.hv_flops(sb)) < 1e-6 * s["flops"]


def test_rdm_eigh_large_blocks(ctx):
    # sector dims above the host-eigensolver cut (rocSOLVER path) and below it
    dims = [(200, 17), (97, 250), (40, 40), (130, 130), (1, 300)]
    sb = _random_sb(4, dims, 6)
    upload(ctx, sb)
    psi = O.splitmix64_uniform_fast(sb.n, 21)
    psi /= np.linalg.norm(psi)
    for side in ("left", "right"):
        evs, vecs = ctx.rdm_eigh(psi, side)
        for k, (ev, vec) in enumerate(zip(evs, vecs)):
            dl, dr, off = sb.Dl[k], sb.Dr[k], sb.offset[k]
            p = psi[off:off + dl * dr].reshape(dl, dr)
            rho = p @ p.T if side == "left" else p.T @ p
            ref = np.linalg.eigvalsh(rho)
            assert np.allclose(np.maximum(ev, 0), np.maximum(ref, 0), atol=1e-14), (side, k)
            assert np.abs(vec @ np.diag(ev) @ vec.T - rho).max() < 1e-14
            assert np.abs(vec.T @ vec - np.eye(len(ev))).max() < 1e-12


def test_rdm_eigh_tridiagonal_path_blocks(ctx):
    # blocks of 33..128 (the tridiagonal-path kernel): full rank, rank 1 and 7
    # with large null clusters, the 33 and 128 ends, and an all-zero block
    dims = [(100, 7), (7, 100), (128, 128), (33, 33), (120, 1), (50, 200), (60, 60)]
    sb = _random_sb(11, dims, 6)
    upload(ctx, sb)
    psi = O.splitmix64_uniform_fast(sb.n, 9)
    z = len(dims) - 1
    psi[sb.offset[z]:sb.offset[z] + dims[z][0] * dims[z][1]] = 0.0
    psi /= np.linalg.norm(psi)
    for side in ("left", "right"):
        evs, vecs = ctx.rdm_eigh(psi, side)
        for k, (ev, vec) in enumerate(zip(evs, vecs)):
            dl, dr, off = sb.Dl[k], sb.Dr[k], sb.offset[k]
            p = psi[off:off + dl * dr].reshape(dl, dr)
            rho = p @ p.T if side == "left" else p.T @ p
            ref = np.linalg.eigvalsh(rho)
            print(f"case tri side={side} k={k} d={len(ev)}", flush=True)
            assert np.all(np.diff(ev) >= 0), (side, k)
            assert np.allclose(np.maximum(ev, 0), np.maximum(ref, 0), atol=1e-15), (side, k)
            assert np.abs(vec @ np.diag(ev) @ vec.T - rho).max() < 1e-14, (side, k)
            assert np.abs(vec.T @ vec - np.eye(len(ev))).max() < 1e-12, (side, k)


def test_rdm_eigh_rank_deficient_blocks(ctx):
    # rank-deficient rho blocks on the three eigensolver paths (Jacobi up to
    # 32, the tridiagonal-path kernel up to 128, rocSOLVER above): large null
    # spaces, rank-1 blocks, a 2x2 and a full block
    dims = [(135, 5), (5, 135), (2, 2), (64, 64), (135, 1), (1, 136)]
    sb = _random_sb(7, dims, 6)
    upload(ctx, sb)
    psi = O.splitmix64_uniform_fast(sb.n, 5)
    psi /= np.linalg.norm(psi)
    for side in ("left", "right"):
        evs, vecs = ctx.rdm_eigh(psi, side)
        for k, (ev, vec) in enumerate(zip(evs, vecs)):
            dl, dr, off = sb.Dl[k], sb.Dr[k], sb.offset[k]
            p = psi[off:off + dl * dr].reshape(dl, dr)
            rho = p @ p.T if side == "left" else p.T @ p
            ref = np.linalg.eigvalsh(rho)
            assert np.all(np.diff(ev) >= 0), (side, k)
            assert np.allclose(np.maximum(ev, 0), np.maximum(ref, 0), atol=1e-15), (side, k)
            assert np.abs(vec @ np.diag(ev) @ vec.T - rho).max() < 1e-14, (side, k)
            assert np.abs(vec.T @ vec - np.eye(len(ev))).max() < 1e-13, (side, k)


@pytest.mark.parametrize("tred", ["column", "lds"])
def test_rdm_eigh_large_block_path(ctx, tred, monkeypatch):
    # blocks of the large-block path (129..2048 states): graded spectra like a
    # DMRG rho (Schmidt values 10^(-k/8)), a flat random one, low rank, and
    # the 2048 end; every eigenpair against numpy, scaled by the block's norm.
    # tred="lds": the blocks that fit the chip's LDS (all but the 2048 one)
    # go through the one-launch LDS-resident tridiagonalisation
    monkeypatch.setenv("NSS_TRED_LDS", "1" if tred == "lds" else "0")
    print(f"case large-block rho path, {tred} tridiagonalisation", flush=True)
    dims = [(900, 900), (1500, 40), (2048, 3), (129, 600), (700, 700), (300, 1)]
    sb = _random_sb(13, dims, 0)
    upload(ctx, sb)
    rng = np.random.default_rng(17)
    psi = np.zeros(sb.n)
    for k, (dl, dr) in enumerate(dims):
        r = min(dl, dr)
        ul = np.linalg.qr(rng.standard_normal((dl, r)))[0]
        vr = np.linalg.qr(rng.standard_normal((dr, r)))[0]
        sv = 10.0 ** (-np.arange(r) / 8.0) if k != 4 else rng.uniform(0.5, 1.0, r)
        psi[sb.offset[k]:sb.offset[k] + dl * dr] = ((ul * sv) @ vr.T).ravel()
    psi /= np.linalg.norm(psi)
    for side in ("left", "right"):
        evs, vecs = ctx.rdm_eigh(psi, side)
        for k, (ev, vec) in enumerate(zip(evs, vecs)):
            dl, dr, off = sb.Dl[k], sb.Dr[k], sb.offset[k]
            p = psi[off:off + dl * dr].reshape(dl, dr)
            rho = p @ p.T if side == "left" else p.T @ p
            ref = np.linalg.eigvalsh(rho)
            nrm = max(np.abs(ref).max(), 1e-300)
            print(f"  side={side} k={k} d={len(ev)}", flush=True)
            assert np.all(np.diff(ev) >= 0), (side, k)
            assert np.abs(np.maximum(ev, 0) - np.maximum(ref, 0)).max() < 1e-13 * nrm, (side, k)
            assert np.abs(vec @ np.diag(ev) @ vec.T - rho).max() < 1e-13 * nrm, (side, k)
            assert np.abs(vec.T @ vec - np.eye(len(ev))).max() < 1e-12, (side, k)

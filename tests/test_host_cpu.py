"""CPU-side checks of the product library: it loads, exports every C-ABI
symbol the header declares, and its host-only pieces (small eigen-solvers,
kept-state rule, input parsing) agree with LAPACK and the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import nssdmrg_amd as N
from nssdmrg_amd import _ffi as F
from oracle import dmrg_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "nssdmrg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nss_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(F.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    # and the ctypes table covers them all
    assert set(syms) <= set(F.SIGNATURES)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(F, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(F, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        F.lib()


@pytest.mark.parametrize("n", [1, 2, 3, 7, 33, 120])
def test_host_sym_eig_matches_lapack(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a + a.T
    w, v = N.host_sym_eig(a)
    we = np.linalg.eigvalsh(a)
    assert np.allclose(w, we, atol=1e-12 * max(1, np.abs(we).max()))
    assert np.allclose(a @ v, v * w[None, :], atol=1e-11)
    assert np.allclose(v.T @ v, np.eye(n), atol=1e-12)


def test_host_sym_eig_degenerate_and_psd():
    # rank-deficient PSD matrix with degenerate spectrum (like rho blocks)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((12, 4))
    rho = x @ x.T
    rho /= np.trace(rho)
    w, v = N.host_sym_eig(rho)
    assert np.allclose(w, np.linalg.eigvalsh(rho), atol=1e-14)
    assert np.allclose(rho @ v, v * w, atol=1e-13)
    eye = np.eye(5)
    w2, v2 = N.host_sym_eig(eye * 2.0)
    assert np.allclose(w2, 2.0) and np.allclose(v2.T @ v2, np.eye(5))


def test_host_sym_eig_rank_deficient_rho():
    # large rho with a rank-14 range and a 138-dimensional null space at
    # rounding level: deflation must be relative to ||rho|| to converge
    rng = np.random.default_rng(14)
    p = rng.standard_normal((152, 14)) * np.exp(-np.arange(14))[None, :]
    rho = p @ p.T
    rho /= np.trace(rho)
    w, v = N.host_sym_eig(rho)
    assert np.allclose(w, np.linalg.eigvalsh(rho), atol=1e-15)
    assert np.abs(rho @ v - v * w).max() < 1e-14
    assert np.abs(v.T @ v - np.eye(152)).max() < 1e-13


def test_host_tridiag_eig():
    rng = np.random.default_rng(9)
    d = rng.standard_normal(40)
    e = rng.standard_normal(39)
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, z = N.host_tridiag_eig(d, e)
    assert np.allclose(w, np.linalg.eigvalsh(t), atol=1e-12)
    assert np.allclose(t @ z, z * w, atol=1e-11)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 40, 120])
def test_host_tridiag_lowest_matches_lapack(n):
    rng = np.random.default_rng(100 + n)
    d = rng.standard_normal(n) * 3.0
    e = rng.standard_normal(max(0, n - 1))
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, z = np.linalg.eigh(t)
    th, zl = N.host_tridiag_lowest(d, e)
    scale = max(1.0, np.abs(w).max())
    assert abs(th - w[0]) <= 1e-13 * scale
    if n == 1 or w[1] - w[0] > 1e-6 * scale:
        assert abs(abs(zl) - abs(z[-1, 0])) <= 1e-10


def test_host_tridiag_lowest_closed_form_and_edge_cases():
    # T = tridiag(-1, 2, -1): theta = 2 - 2 cos(pi/(n+1)), z_k = sqrt(2/(n+1)) sin(k pi/(n+1))
    for n in (5, 64, 150):
        th, zl = N.host_tridiag_lowest(np.full(n, 2.0), np.full(n - 1, -1.0))
        assert abs(th - (2.0 - 2.0 * np.cos(np.pi / (n + 1)))) <= 1e-14
        assert abs(abs(zl) - np.sqrt(2.0 / (n + 1)) * np.sin(n * np.pi / (n + 1))) <= 1e-12
    # a converged Lanczos tridiagonal: tiny last coupling, so z_last is tiny and
    # |beta z_last| must come out at the size numpy gives it
    rng = np.random.default_rng(5)
    n = 60
    d = rng.uniform(-3, 3, n)
    e = rng.uniform(0.2, 1.0, n - 1)
    e[-1] = 1e-9
    d[-1] = 10.0
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, z = np.linalg.eigh(t)
    th, zl = N.host_tridiag_lowest(d, e)
    assert abs(th - w[0]) <= 1e-13 * np.abs(w).max()
    assert abs(abs(zl) - abs(z[-1, 0])) <= 1e-14 + 1e-6 * abs(z[-1, 0])
    # decoupled: the lowest pair lives in the upper block, so z_last = 0
    d = np.array([-5.0, 1.0, 2.0, 3.0, 4.0])
    e = np.array([0.5, 0.0, 0.3, 0.2])
    th, zl = N.host_tridiag_lowest(d, e)
    assert abs(th - np.linalg.eigvalsh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1))[0]) <= 1e-14
    assert abs(zl) <= 1e-12
    # diagonal matrix: the smallest entry
    th, zl = N.host_tridiag_lowest(np.array([3.0, -1.0, 2.0]), np.zeros(2))
    assert th == pytest.approx(-1.0, abs=1e-15) and abs(zl) <= 1e-12


@pytest.mark.parametrize("seed", range(6))
def test_kept_count_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    ev = np.sort(rng.exponential(size=50) ** 3)[::-1]
    ev /= ev.sum()
    ev[rng.integers(0, 50, 5)] = ev[10]          # inject ties
    ev = np.sort(ev)[::-1].copy()
    for M in (1, 5, 10, 20, 60):
        for thr in (0.0, 0.1, 0.5):
            ms, te = N.kept_count(ev, M, 40, thr)
            assert ms == O.kept_count(ev, M, 40, thr)
            assert abs(te - (1.0 - ev[:ms].sum())) < 1e-15
    ms, _ = N.kept_count(ev, 0, 50, 0.1, estates=1e-3)
    assert ms == O.kept_count(ev, 0, 50, 0.1, estates=1e-3)


def test_kept_count_rejects_no_threshold():
    with pytest.raises(N.NssError, match="Mdmrg=Edmrg=0"):
        N.kept_count(np.array([0.5, 0.5]), 0, 2, 0.1, estates=0.0)


def test_read_input_reference_confs(golden_dir):
    c = N.read_input(os.path.join(golden_dir, "spin12_L40_M20_DMRG.conf"))
    assert c["LDMRG"] == 40 and c["MDMRG"] == 20 and c["LANC_NEIGEN"] == 2
    assert c["JX"] == 1.0 and c["JP"] == 1.0 and c["LANC_DIM_THRESHOLD"] == 128
    assert c["DMRGTYPE"] == "i" and c["SPARSE_H"] is False
    h = N.read_input(os.path.join(golden_dir, "hubbard_U0_L40_M40_DMRG.conf"))
    assert h["DMRG_QN"] == [0.5, 0.5] and h["TS"] == [-1.0, -1.0] and h["HFMODE"] is True
    assert h["LANC_NEIGEN"] == 1 and h["MDMRG"] == 40

"""nssdmrg_amd: the nssDMRG superblock path (H_sb.psi, Lanczos, RDM truncation)
on one MI355X (gfx950), behind the C ABI of include/nssdmrg.h.

Python mirror of the reference's host interface:
  * module-level ``read_input`` / ``init_dmrg`` / ``run_DMRG`` /
    ``finalize_dmrg`` (src/INPUT_VARS.f90:141-339, src/DMRG_MAIN.f90:22-132);
  * ``spin_site`` / ``electron_site`` site presets (src/BLOCKS/SITES.f90);
  * ``Context`` for the low-level drop-ins: ``hv`` for spMatVec_direct_main,
    ``lanczos`` for sp_eigh, ``rdm_eigh`` + ``kept_count`` +
    ``rotate_blocks`` for sb_get_rdm / renormalize_block.
Every call runs the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import _ffi as F
from ._ffi import NssError

__all__ = ["Context", "DMRG", "DMRGBatch", "NssError", "kept_count", "host_sym_eig", "host_tridiag_eig",
           "read_input", "init_dmrg", "run_DMRG", "finalize_dmrg", "spin_site", "electron_site",
           "device_ok"]


def device_ok(device: int = 0) -> bool:
    return bool(F.lib().nss_device_ok(device))


class Context:
    """nss_ctx: one device, one stream, the current superblock."""

    def __init__(self, device: int = 0):
        L = F.lib()
        h = C.c_void_p()
        F.check(L.nss_ctx_create(device, C.byref(h)))
        self._h = h
        self._keep = []

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            F.lib().nss_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- superblock
    def set_layout(self, Dl, Dr, offset=None):
        self._Dl = np.ascontiguousarray(Dl, dtype=np.int32)
        self._Dr = np.ascontiguousarray(Dr, dtype=np.int32)
        off = None if offset is None else np.ascontiguousarray(offset, dtype=np.int64)
        F.check(F.lib().nss_sb_set_layout(self._h, len(self._Dl), F.iptr(self._Dl), F.iptr(self._Dr),
                                          None if off is None else F.lptr(off)))
        self.n = int(np.sum(self._Dl.astype(np.int64) * self._Dr))

    def upload_blocks(self, HL, HR, dst, src, A, B, layout="C", precision="fp64"):
        """HL[q], HR[q], and terms Hv_dst += A_t Psi_src B_t^T (dense blocks)."""
        order = "F" if layout == "F" else "C"
        conv = (lambda a: np.asfortranarray(a, dtype=np.float64).ravel(order="K")) if order == "F" \
            else (lambda a: np.ascontiguousarray(a, dtype=np.float64))
        hl = [conv(np.asarray(h)) for h in HL]
        hr = [conv(np.asarray(h)) for h in HR]
        aa = [conv(np.asarray(a)) for a in A]
        bb = [conv(np.asarray(b)) for b in B]
        d = np.ascontiguousarray(dst, dtype=np.int32)
        s = np.ascontiguousarray(src, dtype=np.int32)
        prec = F.NSS_FP32_MIXED if precision in ("fp32", "mixed", F.NSS_FP32_MIXED) else F.NSS_FP64
        lay = F.NSS_COL_MAJOR if order == "F" else F.NSS_ROW_MAJOR
        F.check(F.lib().nss_sb_upload_blocks(self._h, F.ptr_array(hl), F.ptr_array(hr), len(d),
                                             F.iptr(d) if len(d) else None, F.iptr(s) if len(s) else None,
                                             F.ptr_array(aa), F.ptr_array(bb), lay, prec))

    def hv(self, v, fp32: bool = False):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty_like(v)
        fn = F.lib().nss_sb_hv_fp32 if fp32 else F.lib().nss_sb_hv
        F.check(fn(self._h, F.dptr(v), F.dptr(out), v.size))
        return out

    def flops(self) -> float:
        f = C.c_double()
        F.check(F.lib().nss_sb_flops(self._h, C.byref(f)))
        return f.value

    def hv_stats(self, reset: bool = False) -> dict:
        """H*v event timings per precision: timed calls with their W/Y launch
        seconds and FLOPs, and all calls (``calls_all``)."""
        out = np.zeros(12)
        F.check(F.lib().nss_ctx_hv_stats(self._h, int(reset), F.dptr(out)))
        r = {}
        for k, tag in ((0, ""), (1, "32")):
            o = out[6 * k:6 * k + 6]
            r.update({f"calls{tag}": int(o[0]), f"t_w{tag}": o[1], f"t_y{tag}": o[2],
                      f"flops_w{tag}": o[3], f"flops_y{tag}": o[4], f"calls_all{tag}": int(o[5])})
        r["flops"] = r["flops_w"] + r["flops_y"]
        r["flops32"] = r["flops_w32"] + r["flops_y32"]
        return r

    def timers(self) -> dict:
        """Phase timers (seconds) collected when NSS_PROFILE=1."""
        import json
        buf = C.create_string_buffer(1 << 16)
        F.check(F.lib().nss_ctx_timers(self._h, buf, len(buf)))
        return json.loads(buf.value.decode() or "{}")

    def lanczos(self, neigen=1, kmax=150, maxiter=512, tol=1e-12, seed=0, v0_mode=0, precision="fp64"):
        ev = np.zeros(neigen)
        vecs = np.zeros((neigen, self.n))
        nop = C.c_int()
        prec = F.NSS_FP32_MIXED if precision in ("fp32", "mixed") else F.NSS_FP64
        F.check(F.lib().nss_lanczos(self._h, neigen, kmax, maxiter, tol, seed, v0_mode, prec,
                                    F.dptr(ev), F.dptr(vecs), C.byref(nop)))
        return ev, vecs, nop.value

    def rdm_eigh(self, psi, side: str):
        """rho_L (side 'left') or rho_R ('right') per sector -> (evals, evecs) lists."""
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        dims = self._Dl if side == "left" else self._Dr
        ne = int(dims.sum())
        nv = int((dims.astype(np.int64) ** 2).sum())
        ev = np.zeros(max(ne, 1))
        vec = np.zeros(max(nv, 1))
        F.check(F.lib().nss_rdm_eigh(self._h, F.dptr(psi), 0 if side == "left" else 1, F.dptr(ev), F.dptr(vec)))
        evs, vecs, oe, ov = [], [], 0, 0
        for d in dims:
            d = int(d)
            evs.append(ev[oe:oe + d].copy())
            vecs.append(vec[ov:ov + d * d].reshape(d, d).copy())
            oe += d
            ov += d * d
        return evs, vecs

    def rotate_blocks(self, O, Urow, Ucol):
        """out_b = Urow_b^T O_b Ucol_b for lists of dense blocks."""
        o = [np.ascontiguousarray(x, dtype=np.float64) for x in O]
        ur = [np.ascontiguousarray(x, dtype=np.float64) for x in Urow]
        uc = [np.ascontiguousarray(x, dtype=np.float64) for x in Ucol]
        rows = np.array([x.shape[0] for x in o], dtype=np.int32)
        cols = np.array([x.shape[1] for x in o], dtype=np.int32)
        mr = np.array([x.shape[1] for x in ur], dtype=np.int32)
        mc = np.array([x.shape[1] for x in uc], dtype=np.int32)
        out = [np.zeros((int(a), int(b))) for a, b in zip(mr, mc)]
        F.check(F.lib().nss_rotate_blocks(self._h, len(o), F.iptr(rows), F.iptr(cols), F.iptr(mr), F.iptr(mc),
                                          F.ptr_array(o), F.ptr_array(ur), F.ptr_array(uc), F.ptr_array(out)))
        return out


GEMM_MODES = {1: "tile32", 2: "tile64", 3: "large_output"}


def gemm_tune_table() -> list:
    """The H*v tiling modes the autotuner measured in this process, one dict
    per shape class (phase, operand bytes, 32x32 tile count and largest output
    edge in quarter octaves, i.e. ceil(4 log2), chosen mode)."""
    cnt = C.c_int()
    F.check(F.lib().nss_gemm_tune_table(None, 0, C.byref(cnt)))
    rows = (C.c_int * max(1, 5 * cnt.value))()
    F.check(F.lib().nss_gemm_tune_table(rows, cnt.value, C.byref(cnt)))
    out = []
    for r in range(cnt.value):
        k, es, lt, lm, m = rows[5 * r:5 * r + 5]
        out.append({"phase": "W" if k == 1 else "Y", "bytes": es, "tiles32_4log2": lt, "max_edge_4log2": lm,
                    "mode": GEMM_MODES.get(m, str(m))})
    return out


def kept_count(evals_desc, mstates: int, mdim: int, deg_thr: float, estates: float = 0.0):
    """m_s rule of renormalize_block (src/DMRG_RDM.f90:204-220) -> (m_s, trunc_err)."""
    ev = np.ascontiguousarray(evals_desc, dtype=np.float64)
    ms = C.c_int()
    te = C.c_double()
    F.check(F.lib().nss_kept_count(F.dptr(ev), ev.size, mstates, estates, mdim, deg_thr, C.byref(ms),
                                   C.byref(te)))
    return ms.value, te.value


def host_sym_eig(a):
    a = np.array(a, dtype=np.float64, order="C")
    w = np.zeros(a.shape[0])
    F.check(F.lib().nss_host_sym_eig(a.shape[0], F.dptr(a), F.dptr(w)))
    return w, a


def host_tridiag_eig(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    n = d.size
    w = np.zeros(n)
    z = np.zeros((n, n))
    F.check(F.lib().nss_host_tridiag_eig(n, F.dptr(d), F.dptr(e), F.dptr(w), F.dptr(z)))
    return w, z


def host_tridiag_lowest(d, e):
    """Lowest eigenvalue of the tridiagonal (d, e) and the last component of
    its unit eigenvector (sign arbitrary)."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    th, zl = C.c_double(), C.c_double()
    F.check(F.lib().nss_host_tridiag_lowest(d.size, F.dptr(d), F.dptr(e), C.byref(th), C.byref(zl)))
    return th.value, zl.value


# ---------------------------------------------------------------- sites
@dataclass
class SiteSpec:
    kind: str                 # "spin12" | "spin1" | "electron"
    hvec: tuple = (0.0, 0.0)


def spin_site(sun: int = 2, hvec=(0.0, 0.0)) -> SiteSpec:
    """spin_site(sun) preset, src/BLOCKS/SITES.f90:281-344."""
    if sun not in (2, 3):
        raise ValueError("spin_site ERROR: SU(N) value not supported")
    return SiteSpec("spin12" if sun == 2 else "spin1", tuple(hvec))


def electron_site() -> SiteSpec:
    """electron_site() preset (Norb=1), src/BLOCKS/SITES.f90:358-434."""
    return SiteSpec("electron")


# ---------------------------------------------------------------- input
DEFAULTS = {  # src/INPUT_VARS.f90:141-339
    "DMRGTYPE": "i", "PBCDMRG": False, "LDMRG": 5, "MDMRG": 0, "EDMRG": 1e-8, "NSWEEP": 1,
    "MSWEEP": [0], "ESWEEP": [0.0], "QNDIM": 1, "DMRG_QN": [0.0], "NORB": 1, "ULOC": [0.0],
    "UST": 0.0, "JH": 0.0, "JX": 0.0, "JP": 0.0, "HFMODE": True, "XMU": 0.0,
    "DEG_EVALS_THRESHOLD": 0.1, "VERBOSE": 3, "SPARSE_H": False, "LANC_NEIGEN": 2,
    "LANC_NCV_FACTOR": 10, "LANC_NCV_ADD": 0, "LANC_NITER": 512, "LANC_TOLERANCE": 1e-12,
    "LANC_DIM_THRESHOLD": 1024, "LANC_V0_DBLE": False, "SUN": 2, "HVEC": [0.0, 0.0, 0.0],
    "TS": [-1.0, -1.0], "MH": [0.0, 0.0], "QNTYPE": "local", "QNTRUNCATION_ERROR": 0.0,
    "QNTRUNCATION_DIM": 1,
}


def _parse_value(s: str):
    s = s.strip()
    parts = [p for p in s.split(",") if p.strip() != ""]
    vals = []
    for p in parts:
        p = p.strip()
        lo = p.lower()
        if lo in ("t", ".true.", "true"):
            vals.append(True)
        elif lo in ("f", ".false.", "false"):
            vals.append(False)
        else:
            try:
                vals.append(int(p))
            except ValueError:
                try:
                    vals.append(float(p.lower().replace("d", "e")))
                except ValueError:
                    vals.append(p)
    return vals[0] if len(vals) == 1 else vals



class DMRGBatch:
    """Independent chains solved together (nss_dmrg_batch_*): the engines step
    in lockstep and each step's ground states come from one batched Lanczos
    whose H*v over all chains is one grouped launch per phase.  The engines
    must share the context and LDMRG (and NSWEEP for finite runs); their step
    records stay per engine (``DMRG.steps``)."""

    def __init__(self, engines):
        self.engines = list(engines)
        if not self.engines:
            raise ValueError("DMRGBatch needs at least one engine")
        self._arr = (C.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])

    def run(self, kind: str = "i"):
        F.check(F.lib().nss_dmrg_batch_run(len(self.engines), self._arr, ord(kind)))
        return [e.steps() for e in self.engines]

    def finite_begin(self):
        F.check(F.lib().nss_dmrg_batch_finite_begin(len(self.engines), self._arr))

    def sweep(self, im: int):
        F.check(F.lib().nss_dmrg_batch_sweep(len(self.engines), self._arr, im))

def read_input(path: str) -> dict:
    """Reads a DMRG.conf (KEY=value !comment) as parse_input_variable does."""
    cfg = {k: (list(v) if isinstance(v, list) else v) for k, v in DEFAULTS.items()}
    with open(path) as f:
        for line in f:
            line = line.split("!")[0].strip()
            if not line or "=" not in line:
                continue
            k, v = line.split("=", 1)
            cfg[k.strip().upper()] = _parse_value(v)
    if cfg["PBCDMRG"] and cfg["LDMRG"] % 2 == 0:      # :178-183
        cfg["LDMRG"] += 1
    for k in ("MSWEEP", "ESWEEP", "DMRG_QN", "ULOC", "TS", "MH"):
        if not isinstance(cfg[k], list):
            cfg[k] = [cfg[k]]
    return cfg


# ---------------------------------------------------------------- engine
@dataclass
class DMRGConfig:
    site: SiteSpec
    hij: np.ndarray
    ldmrg: int
    mdmrg: int
    edmrg: float = 0.0
    msweep: list = field(default_factory=list)
    esweep: list = field(default_factory=list)
    target_qn: tuple = (0.0, 0.0)
    uloc: float = 0.0
    hfmode: bool = True
    eloc: tuple = (0.0, 0.0)
    neigen: int = 2
    lanc_dim_threshold: int = 1024
    lanc_kmax: int = 150
    lanc_maxiter: int = 512
    lanc_tol: float = 1e-12
    seed: int = 0
    v0_mode: int = 0
    deg_thr: float = 0.1
    precision: str = "fp64"
    verbose: int = 0
    predict: bool = False   # start finite-step Lanczos from the transformed previous ground state
    pbc: bool = False       # PBCdmrg: periodic ring (infinite algorithm; an even ldmrg becomes odd)
    qn_type: str = "local"  # QNtype: "local" (target int(q L)) or "global" (target min(L, int(q)))
    qn_trunc_error: float = 0.0  # QNtruncation_error
    qn_trunc_dim: int = 1        # QNtruncation_dim


class DMRG:
    """Device-resident DMRG run (nss_dmrg handle)."""

    def __init__(self, ctx: Context, cfg: DMRGConfig):
        self.ctx, self.cfg = ctx, cfg
        m = F.Model()
        m.model = {"spin12": F.NSS_MODEL_SPIN12, "spin1": F.NSS_MODEL_SPIN1,
                   "electron": F.NSS_MODEL_HUBBARD}[cfg.site.kind]
        m.hfmode = int(bool(cfg.hfmode))
        hij = np.asarray(cfg.hij, dtype=np.float64).reshape(2, 2)
        for i in range(4):
            m.hij[i] = hij.ravel()[i]
        m.uloc = float(cfg.uloc)
        m.hvec[0], m.hvec[1] = cfg.site.hvec[0], cfg.site.hvec[1]
        m.eloc[0], m.eloc[1] = cfg.eloc
        p = F.Params()
        p.ldmrg, p.mdmrg, p.edmrg = cfg.ldmrg, cfg.mdmrg, cfg.edmrg
        self._ms = np.ascontiguousarray(cfg.msweep or [0], dtype=np.int32)
        self._es = np.ascontiguousarray(cfg.esweep or [0.0] * len(self._ms), dtype=np.float64)
        p.nsweep = len(cfg.msweep)
        p.msweep = self._ms.ctypes.data_as(C.POINTER(C.c_int32))
        p.esweep = self._es.ctypes.data_as(C.POINTER(C.c_double))
        p.target_qn[0] = cfg.target_qn[0]
        p.target_qn[1] = cfg.target_qn[1] if len(cfg.target_qn) > 1 else 0.0
        p.neigen, p.lanc_dim_threshold = cfg.neigen, cfg.lanc_dim_threshold
        p.lanc_kmax, p.lanc_maxiter, p.lanc_tol = cfg.lanc_kmax, cfg.lanc_maxiter, cfg.lanc_tol
        p.seed, p.v0_mode, p.deg_thr = cfg.seed, cfg.v0_mode, cfg.deg_thr
        p.precision = F.NSS_FP32_MIXED if cfg.precision in ("fp32", "mixed") else F.NSS_FP64
        p.verbose = cfg.verbose
        p.predict = int(bool(cfg.predict))
        p.pbc = int(bool(cfg.pbc))
        if cfg.qn_type[:1].lower() not in ("l", "g"):
            raise ValueError("DMRG_MAIN error: QNtype != [local,global]")
        p.qn_type = 1 if cfg.qn_type[:1].lower() == "g" else 0
        p.qn_trunc_error = float(cfg.qn_trunc_error)
        p.qn_trunc_dim = int(cfg.qn_trunc_dim)
        h = C.c_void_p()
        F.check(F.lib().nss_dmrg_create(ctx.handle, C.byref(m), C.byref(p), C.byref(h)))
        self._h = h

    def run(self, kind: str = "i"):
        F.check(F.lib().nss_dmrg_run(self._h, ord(kind)))
        return self.steps()

    def finite_begin(self):
        """iDMRG growth that fills blocks_list before the sweeps."""
        F.check(F.lib().nss_dmrg_finite_begin(self._h))

    def sweep(self, im: int):
        """Finite sweep im (0-based) with Mstates = msweep[im]."""
        F.check(F.lib().nss_dmrg_sweep(self._h, im))

    def finite_step(self) -> bool:
        """The next step of the finite sweeps; False once they are complete."""
        done = C.c_int32()
        F.check(F.lib().nss_dmrg_finite_step(self._h, C.byref(done)))
        return not done.value

    def steps(self) -> list:
        n = F.lib().nss_dmrg_nsteps(self._h)
        out = []
        for i in range(n):
            s = F.StepInfo()
            F.check(F.lib().nss_dmrg_step_info(self._h, i, C.byref(s)))
            d = {k: getattr(s, k) for k, _ in F.StepInfo._fields_}
            d["type"] = chr(s.type)
            d["link"] = chr(s.link) if s.link else "n"
            d["energies"] = np.array(s.energies[:s.neigen])
            d["e_per_site"] = d["energies"] / s.L
            out.append(d)
        return out

    def measure(self, ops, pos=None) -> np.ndarray:
        """Measure_DMRG(ops, pos, avOp) (src/DMRG_MEASURE.f90:148-287) on the
        ground state of the last step.  ops: one (dim x dim) site operator or
        a list of them, in the site basis (spin-1/2 |up>,|dn>; spin-1
        |+1>,|0>,|-1>; electron |0>,|up>,|dw>,|up dw>), conserving the
        quantum numbers; pos: 1-based chain positions (default: every site
        of the last superblock).  Returns an array (len(ops), len(pos)), or
        (len(pos),) for a single operator."""
        single = np.ndim(ops) == 2
        mats = [np.asarray(ops, dtype=np.float64)] if single else [np.asarray(o, dtype=np.float64) for o in ops]
        if pos is None:
            pos = np.arange(1, self.steps()[-1]["L"] + 1)
        p = np.ascontiguousarray(pos, dtype=np.int32)
        a = np.ascontiguousarray(np.stack(mats), dtype=np.float64)
        out = np.zeros((len(mats), len(p)))
        F.check(F.lib().nss_dmrg_measure(self._h, len(mats), F.dptr(a), len(p),
                                         p.ctypes.data_as(C.POINTER(C.c_int32)), F.dptr(out)))
        return out[0] if single else out

    def load_blocks(self, left_path: str, right_path: str):
        """read_block of block_left/right.restart (reference text format) as
        the blocks the next infinite run continues from."""
        F.check(F.lib().nss_dmrg_load_blocks(self._h, os.fsencode(left_path), os.fsencode(right_path)))

    def save_blocks(self, left_path: str, right_path: str):
        """write_block of the two blocks that entered the last step."""
        F.check(F.lib().nss_dmrg_save_blocks(self._h, os.fsencode(left_path), os.fsencode(right_path)))

    def kept_qn(self, i: int, side: str = "left", min_weight: float = 0.0) -> dict:
        """QN -> number of kept states with rho eigenvalue > min_weight."""
        n = C.c_int32()
        sd = 0 if side == "left" else 1
        F.check(F.lib().nss_dmrg_kept_qn(self._h, i, sd, min_weight, 0, None, None, C.byref(n)))
        qn = np.zeros(2 * max(n.value, 1))
        cnt = np.zeros(max(n.value, 1), dtype=np.int32)
        F.check(F.lib().nss_dmrg_kept_qn(self._h, i, sd, min_weight, n.value, F.dptr(qn), F.iptr(cnt),
                                         C.byref(n)))
        return {(qn[2 * j], qn[2 * j + 1]): int(cnt[j]) for j in range(n.value) if cnt[j] > 0}

    def lambdas(self, i: int, side: str = "left") -> list:
        """rho%lambda() of step i (the LambdaQ_left / LambdaQ_right lists of
        write_entanglement): [(qn, eigenvalues ascending), ...] per QN block."""
        sd = 0 if side == "left" else 1
        nb, nv = C.c_int32(), C.c_int64()
        F.check(F.lib().nss_dmrg_lambda(self._h, i, sd, 0, None, None, 0, None, C.byref(nb), C.byref(nv)))
        qn = np.zeros(2 * max(nb.value, 1))
        cnt = np.zeros(max(nb.value, 1), dtype=np.int32)
        vals = np.zeros(max(nv.value, 1))
        F.check(F.lib().nss_dmrg_lambda(self._h, i, sd, nb.value, F.dptr(qn), F.iptr(cnt), nv.value, F.dptr(vals),
                                        C.byref(nb), C.byref(nv)))
        out, o = [], 0
        for j in range(nb.value):
            out.append(((qn[2 * j], qn[2 * j + 1]), vals[o:o + cnt[j]].copy()))
            o += cnt[j]
        return out

    def last_superblock(self) -> dict:
        nq, nt, n = C.c_int(), C.c_int(), C.c_int64()
        F.check(F.lib().nss_dmrg_last_sb_dims(self._h, C.byref(nq), C.byref(nt), C.byref(n)))
        Dl = np.zeros(nq.value, dtype=np.int32)
        Dr = np.zeros(nq.value, dtype=np.int32)
        F.check(F.lib().nss_dmrg_last_sb_export(self._h, F.iptr(Dl), F.iptr(Dr), None, None, None, None, None,
                                                None, None))
        dst = np.zeros(max(nt.value, 1), dtype=np.int32)
        src = np.zeros(max(nt.value, 1), dtype=np.int32)
        F.check(F.lib().nss_dmrg_last_sb_export(self._h, F.iptr(Dl), F.iptr(Dr), F.iptr(dst), F.iptr(src), None,
                                                None, None, None, None))
        dl64, dr64 = Dl.astype(np.int64), Dr.astype(np.int64)
        HL = np.zeros(max(1, int((dl64 ** 2).sum())))
        HR = np.zeros(max(1, int((dr64 ** 2).sum())))
        nA = int(sum(dl64[d] * dl64[s] for d, s in zip(dst[:nt.value], src[:nt.value])))
        nB = int(sum(dr64[d] * dr64[s] for d, s in zip(dst[:nt.value], src[:nt.value])))
        A = np.zeros(max(1, nA))
        B = np.zeros(max(1, nB))
        psi = np.zeros(max(1, n.value))
        F.check(F.lib().nss_dmrg_last_sb_export(self._h, F.iptr(Dl), F.iptr(Dr), F.iptr(dst), F.iptr(src),
                                                F.dptr(HL), F.dptr(HR), F.dptr(A), F.dptr(B), F.dptr(psi)))
        hl, hr, aa, bb = [], [], [], []
        o = 0
        for d in dl64:
            hl.append(HL[o:o + d * d].reshape(d, d)); o += d * d
        o = 0
        for d in dr64:
            hr.append(HR[o:o + d * d].reshape(d, d)); o += d * d
        oa = ob = 0
        for d, s in zip(dst[:nt.value], src[:nt.value]):
            aa.append(A[oa:oa + dl64[d] * dl64[s]].reshape(dl64[d], dl64[s])); oa += dl64[d] * dl64[s]
            bb.append(B[ob:ob + dr64[d] * dr64[s]].reshape(dr64[d], dr64[s])); ob += dr64[d] * dr64[s]
        return {"Dl": Dl, "Dr": Dr, "dst": dst[:nt.value], "src": src[:nt.value], "HL": hl, "HR": hr,
                "A": aa, "B": bb, "psi": psi[:n.value]}

    def close(self):
        if getattr(self, "_h", None):
            F.lib().nss_dmrg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------- module-level mirror
_state: dict = {}


def init_dmrg(Hij, ModelDot, cfg: dict | None = None, ctx: Context | None = None, **over) -> DMRG:
    """init_dmrg(Hij, ModelDot) (src/DMRG_MAIN.f90:22-72) with the values of
    read_input (or DEFAULTS) for the run parameters."""
    c = dict(DEFAULTS)
    if cfg:
        c.update(cfg)
    site = ModelDot[0] if isinstance(ModelDot, (list, tuple)) else ModelDot
    qn = c["DMRG_QN"] if isinstance(c["DMRG_QN"], list) else [c["DMRG_QN"]]
    uloc = c["ULOC"][0] if isinstance(c["ULOC"], list) else c["ULOC"]
    mh = c["MH"] if isinstance(c["MH"], list) else [c["MH"], c["MH"]]
    conf = DMRGConfig(site=site, hij=np.asarray(Hij, dtype=np.float64), ldmrg=int(c["LDMRG"]),
                      mdmrg=int(c["MDMRG"]), edmrg=float(c["EDMRG"]),
                      msweep=[int(x) for x in (c["MSWEEP"] if isinstance(c["MSWEEP"], list) else [c["MSWEEP"]])],
                      esweep=[float(x) for x in (c["ESWEEP"] if isinstance(c["ESWEEP"], list) else [c["ESWEEP"]])],
                      target_qn=tuple(float(x) for x in qn) + ((0.0,) if len(qn) == 1 else ()),
                      uloc=float(uloc), hfmode=bool(c["HFMODE"]), eloc=(float(mh[0]), float(mh[0])),
                      neigen=int(c["LANC_NEIGEN"]), lanc_dim_threshold=int(c["LANC_DIM_THRESHOLD"]),
                      lanc_maxiter=int(c["LANC_NITER"]), lanc_tol=float(c["LANC_TOLERANCE"]),
                      v0_mode=1 if c["LANC_V0_DBLE"] else 0, deg_thr=float(c["DEG_EVALS_THRESHOLD"]),
                      pbc=bool(c["PBCDMRG"]), qn_type=str(c["QNTYPE"]).lower(),
                      qn_trunc_error=float(c["QNTRUNCATION_ERROR"]), qn_trunc_dim=int(c["QNTRUNCATION_DIM"]))
    for k, v in over.items():
        setattr(conf, k, v)
    if c["DMRGTYPE"].lower().startswith("i"):
        conf.msweep, conf.esweep = [], []
    own = ctx is None
    ctx = ctx or Context(0)
    d = DMRG(ctx, conf)
    _state.update(dmrg=d, ctx=ctx, own=own, type=c["DMRGTYPE"].lower()[0])
    return d


def run_DMRG():
    """run_DMRG, src/DMRG_MAIN.f90:121-132."""
    if "dmrg" not in _state:
        raise RuntimeError("DMRG ERROR: DMRG not initialized. Call init_dmrg first.")
    return _state["dmrg"].run(_state["type"])


def finalize_dmrg():
    """finalize_dmrg, src/DMRG_MAIN.f90:81-113."""
    d = _state.pop("dmrg", None)
    if d is not None:
        d.close()
    ctx = _state.pop("ctx", None)
    if ctx is not None and _state.pop("own", False):
        ctx.close()
    _state.clear()


def build():
    from .build_lib import build as _b
    return _b()

"""Writes the oracle's step records of two benchmark workloads as fixtures.

  config1_spin12_L40_m64.json   spin-1/2 Heisenberg J=1 (Hij = diag(1, 1/2)),
                                OBC, Sz=0, L=40 (LDMRG=19), m=64, 2 finite
                                sweeps, LANC_NEIGEN=1, tol 1e-12, seed 0
  config2_hubbard_L64_m<M>.json Hubbard t=1, U=4, HFMODE=T, half filling,
                                OBC, L=64 (LDMRG=31), 3 finite sweeps at
                                m=M (128, and the full 256: about 77 minutes
                                of the numpy oracle on one core)

Each record: step type, left length, L, E0, entropy of rho_L and its
number / configurational split, kept count and truncation error of the left
block, and the kept states per QN with
rho weight above 1e-10 (the null-space states below it are arbitrary in the
reference too).  Run from the repository root:

    python tests/golden/make_config_fixtures.py config1
    python tests/golden/make_config_fixtures.py config2 [M]
"""
from __future__ import annotations

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import dmrg_oracle as O  # noqa: E402

W_MIN = 1e-10
HERE = os.path.dirname(os.path.abspath(__file__))


def records(d: "O.DMRG") -> list:
    out = []
    for r in d.records:
        kept = {}
        for q, ws in (r.kept_w_left or {}).items():
            c = sum(1 for w in ws if w > W_MIN)
            if c:
                kept[",".join(f"{float(x):g}" for x in q)] = c
        out.append({"type": r.kind, "left_len": r.left_len, "L": r.L, "e0": float(r.energies[0]),
                    "entropy_left": float(r.entropy_left), "sn_left": float(r.sn_left),
                    "sc_left": float(r.sc_left), "m_left": int(r.m_left),
                    "trunc_left": float(r.trunc_left), "cut_left": float(r.cut_left), "m_sb": int(r.m_sb),
                    "nsect": int(r.nsect), "kept_left": kept})
    return out


def main():
    which = sys.argv[1]
    t0 = time.time()
    if which == "config1":
        m, L, ns = 64, 40, 2
        p = O.Params(ldmrg=(L - 2) // 2, mdmrg=m, target_qn=(0.0,), hij=[[1.0, 0.0], [0.0, 0.5]], neigen=1,
                     lanc_dim_threshold=128, msweep=[m] * ns, seed=0)
        d = O.DMRG(O.spin_site(2), p)
        name = "config1_spin12_L40_m64.json"
        desc = "spin-1/2 Heisenberg J=1, OBC, Sz=0, L=40 (LDMRG=19), m=64, 2 finite sweeps, FP64"
    elif which == "config2":
        m = int(sys.argv[2]) if len(sys.argv) > 2 else 64
        L, ns = 64, 3
        p = O.Params(ldmrg=(L - 2) // 2, mdmrg=m, target_qn=(0.5, 0.5), hij=[[-1.0, 0.0], [0.0, -1.0]], neigen=1,
                     lanc_dim_threshold=128, msweep=[m] * ns, seed=0)
        d = O.DMRG(O.electron_site(4.0, True), p)
        name = f"config2_hubbard_L64_m{m}.json"
        desc = f"Hubbard t=1 U=4 HFMODE=T, half filling, OBC, L=64 (LDMRG=31), m={m}, 3 finite sweeps, FP64"
    else:
        raise SystemExit(f"unknown fixture {which}")
    d.finite()
    out = {"workload": desc, "generator": "oracle/dmrg_oracle.py DMRG.finite (numpy restatement)",
           "w_min": W_MIN, "seconds": round(time.time() - t0, 1), "records": records(d)}
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {name}: {len(out['records'])} steps in {out['seconds']} s")


if __name__ == "__main__":
    main()

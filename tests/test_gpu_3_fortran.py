"""The Fortran host binding (nssdmrg_amd/fortran/nss_dmrg.f90) on the GPU:
the driver tests/fortran/spin_chain_golden.f90 runs the reference's spin-1/2
L40 M20 iDMRG through init_dmrg_gpu/run_dmrg_gpu and checks the golden
energies (1e-10, all 39 rows) and the golden S_z profile through dmrg_gpu_measure
(1e-8, the tolerance of test/src/Spin1d_SU2/spinSU2_1d.f90:60),
then binds spMatVec_gpu_main to a sparse_HxV procedure
pointer and checks it against the direct-product loops, and finally runs
PBCdmrg=T on the spin-1/2 Heisenberg rings of 4 and 6 sites against their
exact energies."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "spin_chain_golden")


def test_fortran_driver(golden_dir):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (flang absent at build time)")
    r = subprocess.run([EXE, os.path.join(golden_dir, "spin12_L40_M20_energy.check"),
                        os.path.join(golden_dir, "spin12_L40_M20_sz.check")], capture_output=True,
                       text=True, timeout=100)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout

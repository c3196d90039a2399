"""The Fortran drop-in seam on the CPU: tests/fortran/seam_check.f90 calls
mdft_gpu's energy_and_gradient(f, df) with df(no,nx,ny,nz,1)
(src/module_energy_and_gradient.f90:36,63-64), the F-only form, and
excess_energy_gpu through a procedure pointer of the reference's abstract
interface excess_energy_t (src/module_proc_ptrs.f90:3-21), linked against the
CPU mock of the C-ABI (tests/mock_abi, backed by the oracle).  Checked
against the oracle called directly: F parts and dF to 1e-12."""
import os
import subprocess
import sys

import numpy as np

from mdft import ck

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "fortran"))
import seam_io  # noqa: E402


def test_fortran_seam_with_cpu_mock(oracle_mod):
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "mock_abi")])
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "fortran"), "_build/seam_check_mock"])
    n, L, mmax = (9, 10, 8), (9.0, 10.0, 8.0), 3
    t = ck.synthetic_ck(mmax, 256, 0.05, seed=31, n0=0.0333)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    rng = np.random.default_rng(4)
    vext = rng.uniform(-3.0, 6.0, o.N * o.no)
    xi = 1.0 + 0.3 * rng.uniform(-1, 1, o.N * o.no)
    xi[::37] = 0.0
    f, f2, parts, ff, df, df2 = seam_io.run(os.path.join(HERE, "fortran", "_build", "seam_check_mock"), n, L, mmax,
                                            2.4777, 0.0333, 0, t, vext, xi)
    ref, dref = o.energy_and_gradient(xi, vext)
    fexc, dexc = o.excess(xi)
    assert np.allclose(parts, ref, rtol=1e-12, atol=0)
    assert abs(f - ref.sum()) <= 1e-12 * abs(ref.sum()) and f2 == f
    assert abs(ff - fexc) <= 1e-12 * abs(fexc)
    assert np.abs(df - dref).max() <= 1e-12 * np.abs(dref).max()
    assert np.abs(df2 - (1.0 + dexc)).max() <= 1e-12 * np.abs(dexc).max() + 1e-15

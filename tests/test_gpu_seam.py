"""The Fortran drop-in seam on the GPU: tests/fortran/seam_check.f90 calls
mdft_gpu's energy_and_gradient(f, df) with df(no,nx,ny,nz,1)
(src/module_energy_and_gradient.f90:36,63-64), the F-only form, and
excess_energy_gpu through a procedure pointer of the reference's abstract
interface excess_energy_t (src/module_proc_ptrs.f90:3-21), linked against
lib/libmdft_host.so and libmdft_gpu.so.  Checked against the C-ABI called
through ctypes (the same library: equal to 1e-14) and against the oracle
(FP64 bar: F parts 1e-10, dF 1e-10 of max |dF|)."""
import os
import sys

import numpy as np
import pytest

from mdft import ck

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "fortran"))
import seam_io  # noqa: E402

pytestmark = pytest.mark.gpu
EXE = os.path.join(HERE, "fortran", "_build", "seam_check")


def test_fortran_seam_on_the_gpu(oracle_mod):
    from mdft.gpu import Engine
    assert os.path.exists(EXE), "tests/fortran/_build/seam_check is built by __graft_entry__.build()"
    n, L, mmax = (12, 10, 9), (10.0, 9.0, 8.0), 3
    print(f"[seam] Fortran energy_and_gradient / excess_energy_t pointer, n={n} mmax={mmax}", flush=True)
    t = ck.synthetic_ck(mmax, 256, 0.05, seed=31, n0=0.0333)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    rng = np.random.default_rng(4)
    vext = rng.uniform(-3.0, 6.0, o.N * o.no)
    xi = 1.0 + 0.3 * rng.uniform(-1, 1, o.N * o.no)
    xi[::37] = 0.0
    f, f2, parts, ff, df, df2 = seam_io.run(EXE, n, L, mmax, 2.4777, 0.0333, 0, t, vext, xi)
    eng = Engine(n, L, mmax, 2, 2.4777, 0.0333, "fp64")
    eng.set_ck(t)
    eng.set_vext(vext)
    pg, dg = eng.energy_and_gradient(xi)
    de = np.ones_like(xi)
    fe = eng.excess_energy(xi, de)
    assert np.allclose(parts, pg, rtol=1e-14, atol=0) and f == f2
    assert np.abs(df - dg).max() <= 1e-14 * np.abs(dg).max()
    assert abs(ff - fe) <= 1e-14 * abs(fe) and np.abs(df2 - de).max() <= 1e-14 * np.abs(de).max()
    ref, dref = o.energy_and_gradient(xi, vext)
    rel = np.abs(parts - ref) / np.abs(ref)
    print(f"  F parts {parts} oracle {ref} rel {rel}", flush=True)
    assert np.all(rel < 1e-10)
    assert np.abs(df - dref).max() <= 1e-10 * np.abs(dref).max()

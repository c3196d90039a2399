"""The workloads on the device at their own sizes.

* LJ sphere, 48^3 / 24 A, mmax 1: FP64 L-BFGS from the Boltzmann-factor start to
  |dF|/|dF0| < 1e-6, as the workload prescribes.
* Na+ ion, 96^3 / 32 A, mmax 3 (a charged solute): the FP64 minimum, then the FP32 path
  (FP32 evaluations, vectors and iterate, scaled variables) stopped by the reference's
  rule and confirmed by one FP64 evaluation: Delta G within 0.05 kJ/mol, the type B and
  type C corrections the reference adds to F for a charged solute
  (src/module_energy_and_gradient.f90:190-203), and the FP32 errors on F and dF at a
  perturbed density.
* 2 000-atom globule, 192^3 / 64 A, mmax 5: its external potential on the device, one
  FP64 and one FP32 evaluation at the cold start (n = 2.3e9 unknowns, 18.7 GB per FP64
  field), F_exc = 0 and dF_exc = 0 at xi = 1, F_exc quadratic in Delta rho, and FP32
  against FP64 within 1e-5.
"""
import numpy as np
import pytest

from mdft import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_lj_sphere_48_boltzmann_start_to_1e6():
    from mdft import warmstart as ws
    from mdft.gpu import Engine
    from mdft.minimize import minimize
    w = workloads.WORKLOADS["lj_sphere_48"]
    say(f"[config lj_sphere_48] n={w.n} mmax={w.mmax}: Boltzmann start, FP64, gtol 1e-6")
    sol = w.solute()
    eng = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp64")
    eng.set_ck(w.ck_table())
    eng.compute_vext(sol, w.solvent)
    x = eng.alloc_field()
    ws.boltzmann_start(eng, x)
    r = minimize(eng, x, m=1, gtol=1e-6, maxeval=3000)
    say(f"  {r.evaluations} evaluations, status {r.status}, |g|/|g0| {r.grad_ratio:.2e}, F {r.F:.6f}")
    assert r.status == 0 and r.grad_ratio < 1e-6


def test_na_ion_96_fp32_path_and_charged_corrections():
    from mdft import warmstart as ws
    from mdft.gpu import Engine
    from mdft.minimize import minimize
    w = workloads.WORKLOADS["na_ion_96"]
    sol = w.solute()
    say(f"[config na_ion_96] n={w.n} mmax={w.mmax} net charge {sol.net_charge:+.3f}")
    assert abs(sol.net_charge - 1.0) < 1e-12
    table = w.ck_table()
    e64 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp64")
    e64.set_ck(table)
    e64.compute_vext(sol, w.solvent)
    e32 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp32")
    e32.set_ck(table)
    e32.compute_vext(sol, w.solvent)
    # FP32 against FP64 at a perturbed density
    vext = e64.get_vext()
    rng = np.random.default_rng(8)
    xi = workloads.guess_xi(vext, w.kT) * (1.0 + 0.05 * rng.standard_normal(vext.shape))
    p64, d64 = e64.energy_and_gradient(xi)
    p32, d32 = e32.energy_and_gradient(xi)
    rel = np.abs(p32 - p64) / np.abs(p64)
    derr = np.abs(d32 - d64).max() / np.abs(d64).max()
    say(f"  FP32 vs FP64 at a perturbed density: F parts rel {rel}, dF max rel {derr:.2e}")
    assert np.all(rel < 1e-5) and derr < 1e-5
    del xi, d64, d32, vext
    # converged FP64 and the FP32 path stopped by the reference's rule
    x = e64.alloc_field()
    ws.cold_start(e64, x)
    ref = minimize(e64, x, m=1, gtol=1e-6, maxeval=3000)
    ws.cold_start(e32, x)
    r = minimize(e32, x, m=1, gtol=0.0, ftol=1e-3, ftol_absolute=True, maxeval=1500, fnoise=1e-6,
                 fp32_vectors=True, scaled=True, step0="newton", fp32_iterate=True)
    f64 = float(np.sum(e64.energy_and_gradient(x, None)[0]))
    corr = e64.charged_solute_corrections(sol, w.solvent)
    cb, cc = corr["pbc_correction"], corr["pscheme_correction"]
    say(f"  FP64 F* {ref.F:.5f} ({ref.evaluations} evaluations); FP32 path {r.evaluations} evaluations, "
        f"confirmed F {f64:.5f}, error {f64 - ref.F:+.5f}; type B {cb:.5f}, type C {cc:.5f}, "
        f"Delta G = F + B + C = {ref.F + cb + cc:.5f}")
    assert ref.status == 0 and r.status == 1 and abs(f64 - ref.F) <= 0.05
    assert cb < 0.0 and cc != 0.0


def test_globular_192_evaluations():
    import ctypes as C
    from mdft import warmstart as ws
    from mdft.gpu import DEVICE_PTRS, DeviceBuffer, Engine
    w = workloads.WORKLOADS["globular_192"]
    sol = w.solute()
    say(f"[config globular_192] n={w.n} mmax={w.mmax} atoms {len(sol.q)}")
    assert len(sol.q) == 2000
    table = w.ck_table()
    e64 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp64")
    e64.set_ck(table)
    e64.compute_vext(sol, w.solvent)
    nbytes = 8 * e64.N * e64.no
    say(f"  {e64.N * e64.no / 1e9:.2f} G unknowns, {nbytes / 1e9:.1f} GB per FP64 field")
    x = e64.alloc_field()
    g64 = e64.alloc_field()
    # xi = 1: the ideal and excess terms and the excess gradient vanish
    ones_host = np.ones(e64.N * e64.no // 4)
    for k in range(4):
        e64._check(e64.lib.mdft_memcpy_h2d(e64._h, C.c_void_p(x.ptr.value + k * (nbytes // 4)),
                                           ones_host.ctypes.data, nbytes // 4))
    del ones_host
    p1, _ = e64.energy_and_gradient(x, g64)
    say(f"  xi = 1: F parts {p1}")
    assert p1[0] == 0.0 and p1[2] == 0.0
    # the cold start: FP64 against FP32 (gradients compared on the device, relative 2-norm)
    ws.cold_start(e64, x)
    p64, _ = e64.energy_and_gradient(x, g64)
    v = DeviceBuffer(e64, nbytes)
    e64._check(e64.lib.mdft_get_vext(e64._h, v.ptr, DEVICE_PTRS))
    e32 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp32")
    e32.set_ck(table)
    e32.set_vext(v)
    v.free()
    g32 = e64.alloc_field()
    p32, _ = e32.energy_and_gradient(x, g32)
    n2 = e64.dot(g64, g64)
    e64._check(e64.lib.mdft_vec_lincomb(e64._h, e64.N * e64.no, 1.0, g32.ptr, -1.0, g64.ptr, g32.ptr))
    rel = np.abs(p32 - p64) / np.abs(p64)
    drel = float(np.sqrt(e64.dot(g32, g32) / n2))
    say(f"  cold start: FP64 F parts {p64}; FP32 F parts rel {rel}, dF relative 2-norm {drel:.2e}")
    assert np.all(rel < 1e-5) and drel < 1e-5

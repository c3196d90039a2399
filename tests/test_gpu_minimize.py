"""Device-resident L-BFGS (Fortran host, library vector ops) against a CPU
minimisation of the same functional with scipy's L-BFGS-B on the FP64 oracle,
and the device vext against the host restatement."""
import sys

import numpy as np
import pytest

from mdft import ck, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_device_vext_matches_host():
    w = workloads.small("solute20_128", 16)
    sol = w.solute()
    say(f"[vext] {w.name} natom={len(sol.q)}")
    from mdft.gpu import Engine
    eng = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0)
    eng.compute_vext(sol, w.solvent, electrostatics="direct")
    v = eng.get_vext()
    ref = workloads.vext_numpy(w, sol).ravel()
    assert np.abs(v - ref).max() < 1e-9 * np.abs(ref).max()


def test_vector_ops():
    from mdft.gpu import DeviceBuffer, Engine
    eng = Engine((8, 8, 8), (8.0, 8.0, 8.0), 2, 2, 2.4, 0.0333)
    rng = np.random.default_rng(0)
    n = eng.N * eng.no
    a, b = rng.normal(size=n), rng.normal(size=n)
    da, db = eng.alloc_field(), eng.alloc_field()
    da.upload(a)
    db.upload(b)
    assert abs(eng.dot(da, db) - a @ b) < 1e-12 * abs(a @ b) + 1e-12
    out = np.zeros(1)
    import ctypes as C
    eng._check(eng.lib.mdft_vec_axpy_dot(eng._h, n, 0.5, da.ptr, db.ptr, da.ptr, out.ctypes.data_as(C.POINTER(C.c_double))))
    assert np.allclose(db.download(), b + 0.5 * a) and abs(out[0] - a @ (b + 0.5 * a)) < 1e-9
    eng._check(eng.lib.mdft_vec_lincomb(eng._h, n, 2.0, da.ptr, -1.0, db.ptr, db.ptr))
    assert np.allclose(db.download(), 2 * a - (b + 0.5 * a))
    # multi-dot and linear combination over k vectors in one pass
    k = 13
    vs = [rng.standard_normal(n) for _ in range(k)]
    bufs = []
    for v in vs:
        bb = DeviceBuffer(eng, 8 * n)
        bb.upload(v)
        bufs.append(bb)
    ptrs = (C.c_void_p * k)(*[bb.ptr.value for bb in bufs])
    dots = np.zeros(k)
    eng._check(eng.lib.mdft_vec_multi_dot(eng._h, n, k, ptrs, da.ptr, dots.ctypes.data_as(C.POINTER(C.c_double))))
    ref = np.array([a @ v for v in vs])
    assert np.allclose(dots, ref, rtol=1e-10, atol=1e-9)
    coef = rng.standard_normal(k)
    eng._check(eng.lib.mdft_vec_combine(eng._h, n, k, coef.ctypes.data_as(C.POINTER(C.c_double)), ptrs, db.ptr))
    z = sum(c * v for c, v in zip(coef, vs))
    assert np.allclose(db.download(), z, rtol=1e-12, atol=1e-12)
    # direction and first trial point in one pass: z = sum c v, y = a_vec + s z
    dy = DeviceBuffer(eng, 8 * n)
    eng._check(eng.lib.mdft_vec_combine_step(eng._h, n, k, coef.ctypes.data_as(C.POINTER(C.c_double)), ptrs,
                                             db.ptr, -0.37, da.ptr, dy.ptr))
    assert np.allclose(db.download(), z, rtol=1e-12, atol=1e-12)
    assert np.allclose(dy.download(), a - 0.37 * z, rtol=1e-12, atol=1e-12)
    assert eng.lib.mdft_vec_combine_step(eng._h, n, k, coef.ctypes.data_as(C.POINTER(C.c_double)), ptrs,
                                         db.ptr, 1.0, da.ptr, da.ptr) != 0   # y aliasing x is refused


def test_lbfgs_reaches_the_cpu_minimum(oracle_mod):
    """Odd grid (no Nyquist planes, so the reference's F and dF are mutually
    consistent), LJ sphere + a charge, mmax = 2."""
    from scipy.optimize import minimize as spmin
    from mdft.gpu import Engine
    from mdft.minimize import minimize
    w = workloads.Workload("probe", 0, (15, 15, 15), (12.0, 12.0, 12.0), 2, 256)
    sol = workloads.Solute(np.array([[6.0, 6.0, 6.0]]), np.array([0.5]), np.array([3.0]), np.array([0.8]))
    say("[lbfgs] 15^3 mmax=2, charged LJ sphere")
    t = ck.synthetic_ck(2, 256, 0.05, seed=21, n0=w.n0)
    eng = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0)
    eng.set_ck(t)
    eng.compute_vext(sol, w.solvent)
    vext = eng.get_vext()
    xi0 = workloads.guess_xi(vext, w.kT)
    x = eng.alloc_field()
    x.upload(xi0)
    res = minimize(eng, x, m=5, gtol=1e-7, maxeval=800)
    say(f"  gpu: F={res.F:.10f} evals={res.evaluations} iters={res.iterations} status={res.status} "
        f"|g|/|g0|={res.grad_ratio:.2e}")
    assert res.converged and res.grad_ratio < 1e-6
    # monotone up to the evaluation noise the line search accepts (fnoise = 1e-7
    # relative): the reference's FP32 rounding in its chops makes F noisy at ~1e-8
    assert np.all(np.diff(res.trace[:, 0]) <= 1e-7 * abs(res.F))
    o = oracle_mod.Oracle(w.n, w.length, w.mmax, 2, w.kT, w.n0)
    o.set_ck(t)

    def fg(xv):
        parts, df = o.energy_and_gradient(xv, vext)
        return parts.sum(), df * o.dv

    r = spmin(fg, xi0, jac=True, method="L-BFGS-B", options={"maxiter": 3000, "gtol": 1e-10, "ftol": 1e-15})
    say(f"  cpu: F={r.fun:.10f} nit={r.nit}")
    assert abs(res.F - r.fun) < 1e-7 * max(1.0, abs(r.fun))
    # F depends on xi only through xi^2 (rho = rho0 xi^2), so the two minimisers
    # may settle on opposite signs of xi where the density is small
    # Both stop at |dF|/|dF0| ~ 1e-7; along the flattest directions (low density)
    # that leaves xi free at the 1e-3 level, so the bound is on the RMS and a
    # looser max.
    xs = x.download()
    d = np.abs(xs) - np.abs(r.x)
    say(f"  |xi| difference: rms {np.sqrt(np.mean(d * d)):.2e} max {np.abs(d).max():.2e}")
    assert np.sqrt(np.mean(d * d)) < 1e-3 and np.abs(d).max() < 2e-2


def test_pressure_and_lj_tail_corrections(oracle_mod):
    """src/module_pressure_correction.f90 and src/module_lennardjones.f90:145-149
    restated: molecule counts and cavity from the orientation integral of
    rho0 xi^2, the empty-box functional from the oracle at xi = 0, the LJ tail
    from its closed form."""
    from mdft.gpu import Engine
    from mdft.grid import angular_grid
    n, L, mmax = (9, 10, 8), (9.0, 10.0, 8.0), 2
    say("[dG] pressure correction and LJ tail")
    t = ck.synthetic_ck(mmax, 256, 0.05, seed=9, n0=0.0333)
    eng = Engine(n, L, mmax, 2, 2.4777, 0.0333)
    eng.set_ck(t)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    rng = np.random.default_rng(5)
    N, no = eng.N, eng.no
    xi = 1.0 + 0.3 * rng.uniform(-1, 1, N * no)
    xi.reshape(N, no)[::7] *= 0.01          # some cavity points
    eng.set_vext(rng.uniform(-1, 1, N * no))
    out = eng.pressure_correction(xi)
    g = angular_grid(mmax, 2)
    dens = (xi.reshape(N, no) ** 2 * g["w"]).sum(axis=1) * o.rho0
    dv = 9.0 * 10.0 * 8.0 / N
    vol = 9.0 * 10.0 * 8.0
    assert abs(out["n_with_solute"] - dens.sum() * dv) < 1e-10 * dens.sum() * dv
    assert abs(out["n_without_solute"] - 0.0333 * vol) < 1e-12
    assert abs(out["cavity_volume_A3"] - (dens / 0.0333 < 1e-3).sum() * dv) < 1e-12
    f0 = o.energy_and_gradient(np.zeros(N * no), np.zeros(N * no))[0].sum()
    assert abs(out["F_empty_box"] - f0) < 1e-10 * abs(f0)
    dN = 0.0333 * vol - dens.sum() * dv
    assert abs(out["pmv_correction"] - (-dN / 0.0333 * f0 / vol)) < 1e-9 * abs(f0)
    assert abs(out["ideal_gas_correction"] - dN * 2.4777) < 1e-9
    sol = workloads.Solute(np.array([[1.0, 2.0, 3.0], [2.0, 2.0, 2.0]]), np.array([0.3, -0.3]),
                           np.array([3.0, 3.5]), np.array([0.5, 0.8]))
    solv = workloads.spce(0.0333)
    ref = 0.0
    for a in range(2):
        for site in solv.sites:
            if site.eps > 1e-300:
                e = np.sqrt(sol.eps[a] * site.eps)
                sg = 0.5 * (sol.sig[a] + site.sig)
                ref -= (16.0 * 3.14159 / 3.0) * e * sg ** 6 * 0.0333 / 10.0 ** 3
    assert abs(eng.lj_tail_correction(sol, solv, 10.0) - ref) < 1e-12 * abs(ref)


def test_boltzmann_and_reference_starts_reach_one_minimum():
    """The two cold starts -- the reference's xi in {0, 1} and the Boltzmann factor
    exp(-beta vext / 2) -- lead the device L-BFGS to the same minimum; the
    evaluations each needs are reported."""
    from mdft import warmstart as ws
    from mdft.gpu import Engine
    from mdft.minimize import minimize
    w = workloads.Workload("starts", 2, (24, 24, 24), (14.0, 14.0, 14.0), 3, 512)
    say("[starts] reference / Boltzmann start, 24^3 mmax=3, charged solute")
    eng = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0)
    eng.set_ck(ck.synthetic_ck(3, 512, 0.05, seed=41, n0=w.n0))
    sol = workloads.Solute(np.array([[7.0, 7.0, 7.0], [8.3, 7.2, 6.5], [6.1, 8.0, 7.4]]),
                           np.array([0.6, -0.4, -0.2]), np.array([3.2, 2.9, 2.6]), np.array([0.6, 0.5, 0.4]))
    eng.compute_vext(sol, w.solvent)
    x = eng.alloc_field()
    out = {}
    for name, start in (("reference", ws.cold_start), ("boltzmann", ws.boltzmann_start)):
        start(eng, x)
        r = minimize(eng, x, m=5, gtol=1e-8, maxeval=3000)
        out[name] = r
        say(f"  {name}: F={r.F:.6f} after {r.evaluations} evaluations, |g|/|g0|={r.grad_ratio:.1e}")
    fs = [r.F for r in out.values()]
    assert max(fs) - min(fs) < 1e-4


def test_fp32_vector_ops():
    """The FP32-storage vector passes against numpy: n with a ragged tail (n % 4 = 3),
    FP64 sums of the FP32 products, the stored direction rounded to FP32 and the
    trial point built from it exactly, and the widening copy."""
    import ctypes as C
    from mdft.gpu import DeviceBuffer, Engine
    eng = Engine((8, 8, 8), (8.0, 8.0, 8.0), 1, 2, 2.4, 0.0333)
    rng = np.random.default_rng(3)
    n = 4 * 8192 + 3
    say(f"[vec32] n={n}")
    dp = C.POINTER(C.c_double)
    k = 5
    vs = [rng.standard_normal(n).astype(np.float32) for _ in range(k)]
    bufs = []
    for v in vs:
        bb = DeviceBuffer(eng, 4 * n)
        bb.upload(v)
        bufs.append(bb)
    xs = vs[2]                                  # x is one of the vectors: loaded once
    ptrs = (C.c_void_p * k)(*[bb.ptr.value for bb in bufs])
    dots = np.zeros(k)
    eng._check(eng.lib.mdft_vec_multi_dot_f32(eng._h, n, k, ptrs, bufs[2].ptr, dots.ctypes.data_as(dp)))
    ref = np.array([xs.astype(np.float64) @ v.astype(np.float64) for v in vs])
    assert np.allclose(dots, ref, rtol=1e-12, atol=1e-9)
    coef = rng.standard_normal(k)
    x64 = rng.standard_normal(n)
    dx, dy, dz = DeviceBuffer(eng, 8 * n), DeviceBuffer(eng, 8 * n), DeviceBuffer(eng, 4 * n)
    dx.upload(x64)
    eng._check(eng.lib.mdft_vec_combine_step_f32(eng._h, n, k, coef.ctypes.data_as(dp), ptrs, dz.ptr, -0.37, dx.ptr,
                                                 dy.ptr))
    z = sum(c * v.astype(np.float64) for c, v in zip(coef, vs))
    zd = dz.download(np.float32)
    assert np.abs(zd.astype(np.float64) - z).max() <= 1e-6 * np.abs(z).max()
    # y is built from the stored (rounded) z; the device may fuse the multiply-add (1 ulp)
    ref = x64 - 0.37 * zd.astype(np.float64)
    assert np.abs(dy.download() - ref).max() <= 1e-15 * np.abs(ref).max()
    eng._check(eng.lib.mdft_vec_axpy_f32(eng._h, n, 0.25, dz.ptr, dx.ptr, dy.ptr))
    ref = x64 + 0.25 * zd.astype(np.float64)
    assert np.abs(dy.download() - ref).max() <= 1e-15 * np.abs(ref).max()
    eng._check(eng.lib.mdft_vec_widen_f32(eng._h, n, dz.ptr, dy.ptr))
    assert np.array_equal(dy.download(), zd.astype(np.float64))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gradient_rounded_to_fp32(precision):
    """MDFT_DF_FP32: the gradient written as float is the FP64 gradient of the same
    evaluation rounded to FP32, and F is unchanged."""
    import ctypes as C
    from mdft.gpu import DEVICE_PTRS, DF_FP32, DeviceBuffer, Engine
    w = workloads.small("solute20_128", 16)
    say(f"[df32] {w.name} 16^3 {precision}")
    eng = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, precision)
    eng.set_ck(w.ck_table())
    eng.compute_vext(w.solute(), w.solvent)
    xi = eng.alloc_field()
    rng = np.random.default_rng(1)
    xi.upload(workloads.guess_xi(eng.get_vext(), w.kT) * (1 + 0.05 * rng.standard_normal(eng.N * eng.no)))
    g64 = eng.alloc_field()
    p64, _ = eng.energy_and_gradient(xi, g64)
    g32 = DeviceBuffer(eng, 4 * eng.N * eng.no)
    p32 = np.zeros(3)
    dp = C.POINTER(C.c_double)
    eng._check(eng.lib.mdft_energy_and_gradient(eng._h, xi.ptr, p32.ctypes.data_as(dp), g32.ptr,
                                                DEVICE_PTRS | DF_FP32))
    assert np.array_equal(p32, p64)
    assert np.array_equal(g32.download(np.float32), g64.download().astype(np.float32))
    # refused without device pointers
    assert eng.lib.mdft_energy_and_gradient(eng._h, xi.ptr, p32.ctypes.data_as(dp), g32.ptr, DF_FP32) != 0


def test_fp32_vectors_reach_the_band():
    """FP32 evaluations with the minimiser's vectors in FP32 reach the same 0.05 kJ/mol
    band as with FP64 vectors (24^3, mmax=3, charged solute); evaluations reported."""
    from mdft import warmstart as ws
    from mdft.gpu import Engine
    from mdft.minimize import minimize
    w = workloads.Workload("v32", 2, (24, 24, 24), (14.0, 14.0, 14.0), 3, 512)
    say("[v32] FP32 evaluations: FP64 against FP32 minimiser vectors, 24^3 mmax=3")
    table = ck.synthetic_ck(3, 512, 0.05, seed=41, n0=w.n0)
    sol = workloads.Solute(np.array([[7.0, 7.0, 7.0], [8.3, 7.2, 6.5], [6.1, 8.0, 7.4]]),
                           np.array([0.6, -0.4, -0.2]), np.array([3.2, 2.9, 2.6]), np.array([0.6, 0.5, 0.4]))
    e64 = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, "fp64")
    e64.set_ck(table)
    e64.compute_vext(sol, w.solvent)
    x = e64.alloc_field()
    ws.cold_start(e64, x)
    ref = minimize(e64, x, m=1, gtol=1e-8, maxeval=3000)
    e32 = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, "fp32")
    e32.set_ck(table)
    e32.compute_vext(sol, w.solvent)
    out = {}
    for v32 in (False, True):
        ws.cold_start(e32, x)
        r = minimize(e32, x, m=1, gtol=0.0, maxeval=1500, fnoise=1e-6, ftarget=ref.F, fband=0.05, fp32_vectors=v32)
        f64 = float(np.sum(e64.energy_and_gradient(x, None)[0]))       # confirmed in FP64
        out[v32] = (r, f64)
        say(f"  fp32 vectors={v32}: {r.evaluations} evaluations, FP64 F - F* = {f64 - ref.F:+.4f} kJ/mol")
        assert r.status == 4 and abs(f64 - ref.F) <= 0.1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_scaled_variables_evaluation(precision):
    """MDFT_SCALED_VARS: evaluating at u = s xi gives F(xi) and dF/du = dF/dxi / s, s(io) =
    sqrt(w(io) / mean w); mdft_scale_variables maps xi -> u -> xi (FP64) and an FP32
    gradient back to dF/dxi."""
    import ctypes as C
    from mdft.gpu import DEVICE_PTRS, DF_FP32, SCALED_VARS, DeviceBuffer, Engine
    from mdft.grid import angular_grid
    w = workloads.small("solute20_128", 16)
    say(f"[scaled] {w.name} 16^3 {precision}")
    eng = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, precision)
    eng.set_ck(w.ck_table())
    eng.compute_vext(w.solute(), w.solvent)
    rng = np.random.default_rng(2)
    x = workloads.guess_xi(eng.get_vext(), w.kT) * (1 + 0.05 * rng.standard_normal(eng.N * eng.no))
    wq = angular_grid(w.mmax, 2)["w"]
    s = np.tile(np.sqrt(wq / wq.mean()), eng.N)
    bx, bu, g, gu = eng.alloc_field(), eng.alloc_field(), eng.alloc_field(), eng.alloc_field()
    bx.upload(x)
    p0, _ = eng.energy_and_gradient(bx, g)
    eng._check(eng.lib.mdft_scale_variables(eng._h, eng.N * eng.no, 1, 0, bx.ptr, bu.ptr))
    u = bu.download()
    assert np.abs(u - s * x).max() <= 4e-15 * np.abs(u).max()       # s from the library's own mean of w: a few ulp
    dp = C.POINTER(C.c_double)
    p1 = np.zeros(3)
    eng._check(eng.lib.mdft_energy_and_gradient(eng._h, bu.ptr, p1.ctypes.data_as(dp), gu.ptr,
                                                DEVICE_PTRS | SCALED_VARS))
    tol = 1e-12 if precision == "fp64" else 2e-6
    assert np.abs(p1 - p0).max() <= tol * np.abs(p0).max()
    gx, gux = g.download(), gu.download()
    assert np.abs(gux - gx / s).max() <= tol * np.abs(gx / s).max()
    # FP32 gradient storage back to dF/dxi, and u back to xi in place
    g32 = DeviceBuffer(eng, 4 * eng.N * eng.no)
    eng._check(eng.lib.mdft_energy_and_gradient(eng._h, bu.ptr, p1.ctypes.data_as(dp), g32.ptr,
                                                DEVICE_PTRS | SCALED_VARS | DF_FP32))
    eng._check(eng.lib.mdft_scale_variables(eng._h, eng.N * eng.no, 1, 1, g32.ptr, gu.ptr))
    assert np.abs(gu.download() - gx).max() <= max(tol, 2e-7) * np.abs(gx).max()
    eng._check(eng.lib.mdft_scale_variables(eng._h, eng.N * eng.no, -1, 0, bu.ptr, bu.ptr))
    assert np.abs(bu.download() - x).max() <= 4e-15 * np.abs(x).max()


def test_scaled_variables_minimise_faster():
    """The minimiser over u = s xi reaches the 0.05 kJ/mol band of the same minimum in
    fewer evaluations than over xi (24^3, mmax=5, charged solute, FP64 and FP32 with
    FP32 vectors), and the bulk Newton first step in no more than the unit-length
    one; evaluations reported."""
    from mdft import warmstart as ws
    from mdft.gpu import Engine
    from mdft.minimize import minimize
    w = workloads.Workload("sv", 2, (24, 24, 24), (14.0, 14.0, 14.0), 5, 512)
    say("[scaled] evaluations to the band over xi and over u, 24^3 mmax=5")
    table = ck.synthetic_ck(5, 512, 0.05, seed=43, n0=w.n0)
    sol = workloads.Solute(np.array([[7.0, 7.0, 7.0], [8.3, 7.2, 6.5], [6.1, 8.0, 7.4]]),
                           np.array([0.6, -0.4, -0.2]), np.array([3.2, 2.9, 2.6]), np.array([0.6, 0.5, 0.4]))
    e64 = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, "fp64")
    e64.set_ck(table)
    e64.compute_vext(sol, w.solvent)
    x = e64.alloc_field()
    ws.cold_start(e64, x)
    ref = minimize(e64, x, m=1, gtol=1e-8, maxeval=3000)
    e32 = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, "fp32")
    e32.set_ck(table)
    e32.compute_vext(sol, w.solvent)
    out = {}
    for name, eng, kw in (("fp64", e64, {}), ("fp64 scaled", e64, {"scaled": True}),
                          ("fp32", e32, {"fp32_vectors": True, "fnoise": 1e-6}),
                          ("fp32 scaled", e32, {"fp32_vectors": True, "fnoise": 1e-6, "scaled": True}),
                          ("fp32 scaled newton", e32, {"fp32_vectors": True, "fnoise": 1e-6, "scaled": True,
                                                       "step0": "newton"})):
        ws.cold_start(eng, x)
        r = minimize(eng, x, m=1, gtol=0.0, maxeval=1500, ftarget=ref.F, fband=0.05, **kw)
        f64 = float(np.sum(e64.energy_and_gradient(x, None)[0]))
        out[name] = r.evaluations
        say(f"  {name}: {r.evaluations} evaluations, FP64 F - F* = {f64 - ref.F:+.4f}")
        assert r.status == 4 and abs(f64 - ref.F) <= 0.1
    assert out["fp64 scaled"] < out["fp64"] and out["fp32 scaled"] < out["fp32"]
    assert out["fp32 scaled newton"] <= out["fp32 scaled"]      # the bulk Newton first step


@pytest.mark.parametrize("precision,df32", [("fp64", False), ("fp32", False), ("fp32", True)])
def test_gradient_products_in_the_evaluation(precision, df32):
    """mdft_energy_and_gradient_dots: F and the gradient as mdft_energy_and_gradient
    gives them, and dots[j] = dF . v_j formed in the final pass against numpy (FP64
    sums of the stored gradient), the gradient itself among the vectors."""
    import ctypes as C
    from mdft.gpu import DEVICE_PTRS, DF_FP32, DeviceBuffer, Engine
    w = workloads.small("solute20_128", 16)
    say(f"[dots] {w.name} 16^3 {precision} df32={df32}")
    eng = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, precision)
    eng.set_ck(w.ck_table())
    eng.compute_vext(w.solute(), w.solvent)
    rng = np.random.default_rng(4)
    n = eng.N * eng.no
    xi = eng.alloc_field()
    xi.upload(workloads.guess_xi(eng.get_vext(), w.kT) * (1 + 0.05 * rng.standard_normal(n)))
    dtype = np.float32 if df32 else np.float64
    flags = DEVICE_PTRS | (DF_FP32 if df32 else 0)
    dp = C.POINTER(C.c_double)
    g0 = DeviceBuffer(eng, np.dtype(dtype).itemsize * n)
    p0 = np.zeros(3)
    eng._check(eng.lib.mdft_energy_and_gradient(eng._h, xi.ptr, p0.ctypes.data_as(dp), g0.ptr, flags))
    vs = [rng.standard_normal(n).astype(dtype) for _ in range(2)]
    bv = []
    for v in vs:
        b = DeviceBuffer(eng, v.nbytes)
        b.upload(v)
        bv.append(b)
    g1 = DeviceBuffer(eng, np.dtype(dtype).itemsize * n)
    ptrs = (C.c_void_p * 3)(bv[0].ptr.value, g1.ptr.value, bv[1].ptr.value)
    p1, dots = np.zeros(3), np.zeros(3)
    eng._check(eng.lib.mdft_energy_and_gradient_dots(eng._h, xi.ptr, p1.ctypes.data_as(dp), g1.ptr, flags, 3, ptrs,
                                                     dots.ctypes.data_as(dp)))
    assert np.array_equal(p1, p0)
    ga, gb = g0.download(dtype), g1.download(dtype)
    assert np.array_equal(ga, gb)
    g = gb.astype(np.float64)
    ref = np.array([g @ vs[0].astype(np.float64), g @ g, g @ vs[1].astype(np.float64)])
    say(f"  dots {dots}, numpy {ref}")
    assert np.allclose(dots, ref, rtol=1e-12, atol=1e-12 * np.abs(g).max() * np.sqrt(n))


@pytest.mark.parametrize("n", [(16, 16, 16), (15, 15, 15)])
def test_gradient_products_with_fp32_iterate(n):
    """The fast minimiser's evaluation (FP32 iterate, FP32 gradient, gradient products over
    FP32 vectors, scaled variables) has its own final-pass variant on grids of whole 4-point
    tiles, reading an FP32 copy of vext: F and the gradient equal, to the last bit, those of
    the evaluation without products (the general variant), the products match numpy, and a
    new vext reaches the copy.  15^3 (not whole tiles) takes the general variant."""
    import ctypes as C
    from mdft.gpu import DEVICE_PTRS, DF_FP32, SCALED_VARS, XI_FP32, DeviceBuffer, Engine
    w = workloads.small("solute20_128", 16)
    say(f"[dots fp32 iterate] {n} fp32")
    eng = Engine(n, w.length, w.mmax, 2, w.kT, w.n0, "fp32")
    eng.set_ck(w.ck_table())
    eng.compute_vext(w.solute(), w.solvent)
    rng = np.random.default_rng(5)
    nv = eng.N * eng.no
    base = workloads.guess_xi(eng.get_vext(), w.kT) * (1 + 0.05 * rng.standard_normal(nv))
    xi = DeviceBuffer(eng, 4 * nv)
    xi.upload(base.astype(np.float32))
    flags = DEVICE_PTRS | DF_FP32 | XI_FP32 | SCALED_VARS
    dp = C.POINTER(C.c_double)
    vs = [rng.standard_normal(nv).astype(np.float32) for _ in range(2)]
    bv = []
    for v in vs:
        b = DeviceBuffer(eng, v.nbytes)
        b.upload(v)
        bv.append(b)
    vext0 = eng.get_vext()
    for step, vext in enumerate((None, 1.1 * vext0)):
        if vext is not None:
            eng.set_vext(vext)   # after the FP32 copy was made: the next evaluation must see it
        g0 = DeviceBuffer(eng, 4 * nv)
        p0 = np.zeros(3)
        eng._check(eng.lib.mdft_energy_and_gradient(eng._h, xi.ptr, p0.ctypes.data_as(dp), g0.ptr, flags))
        g1 = DeviceBuffer(eng, 4 * nv)
        ptrs = (C.c_void_p * 3)(bv[0].ptr.value, g1.ptr.value, bv[1].ptr.value)
        p1, dots = np.zeros(3), np.zeros(3)
        eng._check(eng.lib.mdft_energy_and_gradient_dots(eng._h, xi.ptr, p1.ctypes.data_as(dp), g1.ptr, flags, 3,
                                                         ptrs, dots.ctypes.data_as(dp)))
        ga, gb = g0.download(np.float32), g1.download(np.float32)
        say(f"  vext {step}: F parts {p1}, max |dg| {np.abs(ga - gb).max():.3g}")
        assert np.array_equal(p1, p0)
        assert np.array_equal(ga, gb)
        g = gb.astype(np.float64)
        ref = np.array([g @ vs[0].astype(np.float64), g @ g, g @ vs[1].astype(np.float64)])
        assert np.allclose(dots, ref, rtol=1e-12, atol=1e-12 * np.abs(g).max() * np.sqrt(nv))
        if step == 0:
            f_first = p1.copy()
    assert not np.array_equal(p1, f_first)   # the external term changed with vext


def test_fp32_iterate_evaluation_and_minimisation():
    """MDFT_XI_FP32: an evaluation that reads xi as float gives the evaluation of the FP64
    array holding the same values, to the last bit; the minimiser with its iterate in FP32
    as well (FP32 vectors, scaled variables) reaches the 0.05 kJ/mol band of the FP64
    minimum with the reference's stop rule (fold - f) <= 1e-3."""
    import ctypes as C
    from mdft import warmstart as ws
    from mdft.gpu import DEVICE_PTRS, XI_FP32, DeviceBuffer, Engine
    from mdft.minimize import minimize
    w = workloads.Workload("x32", 2, (24, 24, 24), (14.0, 14.0, 14.0), 5, 512)
    say("[fp32 iterate] evaluation and minimisation, 24^3 mmax=5")
    table = ck.synthetic_ck(5, 512, 0.05, seed=43, n0=w.n0)
    sol = workloads.Solute(np.array([[7.0, 7.0, 7.0], [8.3, 7.2, 6.5]]), np.array([0.5, -0.5]),
                           np.array([3.2, 2.9]), np.array([0.6, 0.5]))
    e64 = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, "fp64")
    e64.set_ck(table)
    e64.compute_vext(sol, w.solvent)
    e32 = Engine(w.n, w.length, w.mmax, 2, w.kT, w.n0, "fp32")
    e32.set_ck(table)
    e32.compute_vext(sol, w.solvent)
    rng = np.random.default_rng(9)
    xf = (1.0 + 0.2 * rng.uniform(-1, 1, e32.N * e32.no)).astype(np.float32)
    x32 = DeviceBuffer(e32, xf.nbytes)
    x32.upload(xf)
    x64 = DeviceBuffer(e32, 8 * xf.size)
    x64.upload(xf.astype(np.float64))
    for eng in (e32, e64):
        pa, pb = np.zeros(3), np.zeros(3)
        dp = C.POINTER(C.c_double)
        ga, gb = eng.alloc_field(), eng.alloc_field()
        assert eng.lib.mdft_energy_and_gradient(eng._h, x32.ptr, pa.ctypes.data_as(dp), ga.ptr,
                                                DEVICE_PTRS | XI_FP32) == 0
        assert eng.lib.mdft_energy_and_gradient(eng._h, x64.ptr, pb.ctypes.data_as(dp), gb.ptr, DEVICE_PTRS) == 0
        assert np.array_equal(pa, pb) and np.array_equal(ga.download(), gb.download())
    x = e64.alloc_field()
    ws.cold_start(e64, x)
    ref = minimize(e64, x, m=1, gtol=1e-8, maxeval=3000)
    out = {}
    for it32 in (False, True):
        ws.cold_start(e32, x)
        r = minimize(e32, x, m=1, gtol=0.0, ftol=1e-3, ftol_absolute=True, maxeval=1500, fnoise=1e-6,
                     fp32_vectors=True, scaled=True, step0="newton", fp32_iterate=it32)
        f64 = float(np.sum(e64.energy_and_gradient(x, None)[0]))
        out[it32] = r.evaluations
        say(f"  fp32 iterate {it32}: {r.evaluations} evaluations, status {r.status}, FP64 F - F* = {f64 - ref.F:+.4f}")
        assert r.status == 1 and abs(f64 - ref.F) <= 0.05

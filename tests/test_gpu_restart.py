"""GPU tests of the restart path (density.bin): the device transforms against the
CPU oracle's angl2proj / proj2angl, the device regridding against
oracle/restart_oracle.py, write -> read round trips through the file, the
solute check, and restarts from converged solutions (same grid and a coarser one).

Tolerances: transforms 1e-12 of the largest magnitude (FP64); regridding 1e-13
(the device may contract a*(1-w) + b*w into fused multiply-adds); round trips
of band-limited fields 1e-12; full-density files bit-exact."""
import sys

import numpy as np
import pytest

from mdft import ck, workloads
from oracle import restart_oracle as RO

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(n, length, mmax):
    from mdft.gpu import Engine
    return Engine(n, length, mmax, 2, 2.4777, 0.0333, "fp64")


def _bandlimited(o, rng, N):
    """xi = 1 + proj2angl(random projections of a real field): band-limited, positive-ish."""
    f = 0.2 * rng.normal(size=(N, o.no))
    return np.stack([1.0 + o.proj2angl(o.angl2proj(f[i])) for i in range(N)])


def test_device_transforms_match_oracle(oracle_mod):
    from mdft.gpu import DeviceBuffer
    n, L, mmax = (5, 4, 3), (5.0, 4.0, 3.0), 3
    say(f"[restart] angl2proj / proj2angl on the device vs the oracle, {n} mmax={mmax}")
    eng = _engine(n, L, mmax)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    rng = np.random.default_rng(0)
    N, no, npj = eng.N, eng.no, eng.np_
    npts = 37                                 # not the context's grid: any number of points
    f = rng.normal(size=(npts, no))
    df, dp = DeviceBuffer(eng, f.nbytes), DeviceBuffer(eng, 16 * npts * npj)
    df.upload(f)
    eng._check(eng.lib.mdft_angl2proj(eng._h, df.ptr, npts, dp.ptr))
    got = dp.download(np.complex128).reshape(npts, npj)
    ref = np.stack([o.angl2proj(f[i]) for i in range(npts)])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(f"  angl2proj max rel err {err:.2e}")
    assert err < 1e-12
    back = DeviceBuffer(eng, f.nbytes)
    eng._check(eng.lib.mdft_proj2angl(eng._h, dp.ptr, npts, back.ptr))
    gb = back.download().reshape(npts, no)
    rb = np.stack([o.proj2angl(ref[i]) for i in range(npts)])
    err = np.abs(gb - rb).max() / np.abs(rb).max()
    say(f"  proj2angl max rel err {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("case", [0, 1, 2])
def test_device_regrid_matches_oracle(case):
    from mdft.gpu import DeviceBuffer
    import ctypes as C
    n_prev, d_prev, n, length = [((6, 5, 7), (0.9, 1.1, 0.7), (10, 8, 12), (8.0, 7.0, 6.3)),
                                 ((9, 7, 8), (1.0, 1.0, 1.0), (6, 6, 8), (6.0, 6.0, 6.0)),
                                 ((5, 6, 4), (2.0, 1.25, 2.0), (10, 10, 8), (8.0, 7.5, 8.0))][case]
    say(f"[restart] regrid {n_prev} x {d_prev} -> {n} in {length}")
    eng = _engine(n, length, 1)
    rng = np.random.default_rng(case)
    prev = rng.normal(size=(n_prev[2], n_prev[1], n_prev[0], eng.no))
    dprev, dxi = DeviceBuffer(eng, prev.nbytes), eng.alloc_field()
    dprev.upload(prev)
    eng._check(eng.lib.mdft_regrid_trilinear(eng._h, dprev.ptr, (C.c_int * 3)(*n_prev), (C.c_double * 3)(*d_prev),
                                             1.0, dxi.ptr))
    got = dxi.download().reshape(n[2], n[1], n[0], eng.no)
    ref = RO.regrid(prev, n_prev, d_prev, n, length, outside=1.0)
    err = np.abs(got - ref).max()
    say(f"  max abs err {err:.2e}, points outside the previous grid: {int(np.sum(ref[..., 0] == 1.0))}")
    assert err < 1e-13 * max(1.0, np.abs(ref).max())


def test_density_bin_round_trips(oracle_mod, tmp_path):
    from mdft import restart as R
    n, L, mmax = (6, 5, 4), (6.0, 5.0, 4.0), 3
    say("[restart] write -> read of density.bin (projections and full xi)")
    eng = _engine(n, L, mmax)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    rng = np.random.default_rng(4)
    xi = _bandlimited(o, rng, eng.N)
    sol = workloads.Solute(np.array([[3.0, 2.5, 2.0], [3.5, 2.0, 2.5]]), np.array([0.4, -0.4]),
                           np.array([3.0, 2.8]), np.array([0.5, 0.4]))
    bx, by = eng.alloc_field(), eng.alloc_field()
    bx.upload(xi.reshape(-1))
    path = str(tmp_path / "density.bin")
    R.write_density_bin(path, eng, bx, sol)
    h = R.read_density_bin(path, eng, by, solute=sol)
    assert (h["n"], h["mmax"], h["np"], h["full"]) == (n, mmax, eng.np_, False)
    err = np.abs(by.download() - xi.reshape(-1)).max()
    say(f"  projections: max err {err:.2e}")
    assert err < 1e-12
    R.write_density_bin(path, eng, bx, sol, full=True)
    R.read_density_bin(path, eng, by, solute=sol)
    assert np.array_equal(by.download(), xi.reshape(-1))
    other = workloads.Solute(sol.xyz + 0.1, sol.q, sol.sig, sol.eps)
    with pytest.raises(ValueError):
        R.read_density_bin(path, eng, by, solute=other)


def test_restart_from_converged_solutions(tmp_path):
    """Converge on 24^3 and 32^3 (same box) and write density.bin from each.
    Restarting the 32^3 run from its own file starts within a few kJ/mol of the
    minimum and reaches the 0.05 kJ/mol band in fewer evaluations than the cold
    start.  Restarting it from the 24^3 file (regridded on the device) reaches the
    same minimum; its evaluation count is reported -- trilinear interpolation next
    to the excluded core leaves near-empty points that the minimiser must refill,
    as in the reference's restart."""
    from mdft import restart as R
    from mdft import warmstart as ws
    from mdft.minimize import minimize
    L, mmax = (12.0, 12.0, 12.0), 3
    table = ck.synthetic_ck(mmax, 512, 0.05, seed=31, n0=0.0333)
    sol = workloads.Solute(np.array([[6.0, 6.0, 6.0], [7.2, 6.3, 5.8]]), np.array([0.5, -0.5]),
                           np.array([3.2, 2.9]), np.array([0.6, 0.5]))
    w24 = workloads.Workload("r24", 3, (24,) * 3, L, mmax, 512)
    w32 = workloads.Workload("r32", 3, (32,) * 3, L, mmax, 512)
    say("[restart] converge on 24^3 and 32^3, restart 32^3 from density.bin")
    e24 = _engine(w24.n, L, mmax)
    e24.set_ck(table)
    e24.compute_vext(sol, w24.solvent)
    x24 = e24.alloc_field()
    ws.cold_start(e24, x24)
    minimize(e24, x24, m=5, gtol=1e-7, maxeval=1500)
    coarse = str(tmp_path / "coarse.bin")
    R.write_density_bin(coarse, e24, x24, sol)
    e32 = _engine(w32.n, L, mmax)
    e32.set_ck(table)
    e32.compute_vext(sol, w32.solvent)
    x = e32.alloc_field()
    ws.cold_start(e32, x)
    ref = minimize(e32, x, m=5, gtol=1e-8, maxeval=2000)
    same = str(tmp_path / "same.bin")
    R.write_density_bin(same, e32, x, sol)
    ws.cold_start(e32, x)
    cold = minimize(e32, x, m=5, gtol=0.0, maxeval=1500, ftarget=ref.F, fband=0.05)
    R.read_density_bin(same, e32, x, solute=sol)
    f_same = float(np.sum(e32.energy_and_gradient(x, None)[0]))
    rs = minimize(e32, x, m=5, gtol=0.0, maxeval=1500, ftarget=ref.F, fband=0.05)
    R.read_density_bin(coarse, e32, x, solute=sol)
    f_coarse = float(np.sum(e32.energy_and_gradient(x, None)[0]))
    rc = minimize(e32, x, m=5, gtol=1e-8, maxeval=2000)
    say(f"  F*={ref.F:.4f}: cold start {cold.evaluations} evaluations to the band; restart from the same grid "
        f"F-F*={f_same - ref.F:.3f}, {rs.evaluations} evaluations; from 24^3 F-F*={f_coarse - ref.F:.3f}, "
        f"converged to F={rc.F:.4f} in {rc.evaluations} evaluations")
    assert abs(rs.F - ref.F) <= 0.05 and abs(cold.F - ref.F) <= 0.05
    assert f_same - ref.F < 5.0 and rs.evaluations < cold.evaluations
    assert abs(rc.F - ref.F) < 1e-4

"""GPU parity: the HIP path through the C-ABI against the FP64 CPU oracle on
identical inputs (xi, vext, c(k)), stage by stage and for F, dF.

Tolerances: F parts relative 1e-10 (FP64 path); dF max abs difference <= 1e-10
of max |dF|; projection stages <= 1e-10 of the stage's max modulus."""
import sys

import numpy as np
import pytest

from mdft import ck, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(n, length, mmax, kT=2.4777, n0=0.0333):
    from mdft.gpu import Engine
    return Engine(n, length, mmax, 2, kT, n0, "fp64")


def _inputs(N, no, seed, rho0, nzero=True):
    rng = np.random.default_rng(seed)
    vext = rng.uniform(-3.0, 6.0, N * no)
    vext[::97] = 694.7  # clipped values
    xi = 1.0 + 0.3 * rng.uniform(-1, 1, N * no)
    if nzero:
        xi[::53] = 0.0        # rho = 0: near-zero branch of the ideal term
        xi[::71] = 1e-9       # rho below DBL_EPSILON
    return xi, vext


CASES = [
    # name, n, L, mmax, nk
    ("lj_sphere_48 grid", (48, 48, 48), (24.0, 24.0, 24.0), 1, 512),
    ("mmax2 16^3", (16, 16, 16), (12.0, 12.0, 12.0), 2, 256),
    ("mmax3 24x20x18", (24, 20, 18), (14.0, 13.0, 12.0), 3, 256),
    ("mmax5 16^3", (16, 16, 16), (12.0, 12.0, 12.0), 5, 256),
    ("mmax4 12^3 odd mix", (12, 10, 15), (10.0, 9.0, 11.0), 4, 256),
    # one 192- or 96-point axis each: the radix-3 FFT kernels of the 192^3 and 96^3 workloads
    ("fft x192 192x8x6", (192, 8, 6), (48.0, 4.0, 3.0), 1, 1024),
    ("fft y96 10x96x12", (10, 96, 12), (5.0, 30.0, 6.0), 2, 1024),
    ("fft z192 8x6x192", (8, 6, 192), (4.0, 3.0, 60.0), 1, 1024),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_energy_and_gradient_parity(oracle_mod, case):
    name, n, L, mmax, nk = case
    say(f"[parity] {name} n={n} mmax={mmax}")
    t = ck.synthetic_ck(mmax, nk, 0.05, seed=11 + mmax, n0=0.0333)
    eng = _engine(n, L, mmax)
    eng.set_ck(t)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    assert (eng.no, eng.np_, eng.npc, eng.nq) == (o.no, o.np_, o.npc, o.nq)
    xi, vext = _inputs(eng.N, eng.no, seed=3, rho0=o.rho0)
    eng.set_vext(vext)
    # stages of the excess term
    for stage in (1, 2, 3, 4):
        g = eng.excess_stage(xi, stage)
        r = o.excess_stage(xi, stage)
        scale = np.abs(r).max()
        err = np.abs(g - r).max() / scale
        say(f"  stage {stage}: max rel err {err:.2e}")
        assert err < 1e-10, (stage, err)
    parts_g, df_g = eng.energy_and_gradient(xi)
    parts_o, df_o = o.energy_and_gradient(xi, vext)
    rel = np.abs(parts_g - parts_o) / np.maximum(np.abs(parts_o), 1e-300)
    say(f"  F parts gpu={parts_g} oracle={parts_o} rel={rel}")
    assert np.all(rel < 1e-10)
    derr = np.abs(df_g - df_o).max() / np.abs(df_o).max()
    say(f"  dF max rel err {derr:.2e}")
    assert derr < 1e-10


def test_plugin_and_f_only_modes(oracle_mod):
    n, L, mmax = (16, 16, 16), (12.0, 12.0, 12.0), 3
    say("[modes] excess plug-in accumulation, F-only, ideal alone")
    t = ck.synthetic_ck(mmax, 256, 0.05, seed=5, n0=0.0333)
    eng = _engine(n, L, mmax)
    eng.set_ck(t)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    xi, vext = _inputs(eng.N, eng.no, seed=4, rho0=o.rho0)
    eng.set_vext(vext)
    base = np.linspace(-1.0, 1.0, xi.size)
    df = base.copy()
    fexc = eng.excess_energy(xi, df)
    fexc_o, dexc_o = o.excess(xi)
    assert abs(fexc - fexc_o) < 1e-10 * abs(fexc_o)
    assert np.abs((df - base) - dexc_o).max() < 1e-10 * np.abs(dexc_o).max()
    parts_f, none = eng.energy_and_gradient(xi, want_df=False)
    parts_g, _ = eng.energy_and_gradient(xi)
    assert none is None and np.array_equal(parts_f, parts_g)
    df2 = base.copy()
    fid, fext = eng.ideal_external(xi, df2)
    fid_o, fext_o, did_o = o.ideal_ext(xi, vext)
    assert abs(fid - fid_o) < 1e-10 * abs(fid_o) and abs(fext - fext_o) < 1e-10 * abs(fext_o)
    assert np.abs((df2 - base) - did_o).max() < 1e-10 * np.abs(did_o).max()


def test_error_paths():
    from mdft.gpu import Engine, MdftError
    say("[errors] call order, invalid description, short c(k) table")
    eng = _engine((8, 8, 8), (8.0, 8.0, 8.0), 2)
    xi = np.ones(eng.N * eng.no)
    with pytest.raises(MdftError) as e:
        eng.energy_and_gradient(xi)
    assert e.value.code == -4  # MDFT_ESTATE
    with pytest.raises(MdftError) as e:
        Engine((8, 8, 7 * 11), (8.0, 8.0, 8.0), 2, 2, 2.4, 0.0333)  # 77 = 7*11 has no 2/3/5 plan
    assert e.value.code == -1
    short = ck.synthetic_ck(2, 10, 0.05, seed=1, n0=0.0333)  # 10 rows < nq
    with pytest.raises(MdftError) as e:
        eng.set_ck(short)
    assert e.value.code == -5  # MDFT_ERANGE


def test_full_size_properties():
    """128^3, mmax = 5 (the 20-atom solute workload grid): size-independent
    properties -- xi = 1 gives F_exc = 0 and dF_exc = 0; F_exc is quadratic in
    drho; repeated evaluations are bitwise identical; device-resident and host
    paths agree."""
    w = workloads.WORKLOADS["solute20_128"]
    say(f"[full] {w.name} n={w.n} mmax={w.mmax}")
    eng = _engine(w.n, w.length, w.mmax, kT=w.kT, n0=w.n0)
    eng.set_ck(w.ck_table())
    rng = np.random.default_rng(1)
    nno = eng.N * eng.no
    vext = np.zeros(nno)
    eng.set_vext(vext)
    xi1 = np.ones(nno)
    df = np.empty(nno)
    parts = eng.energy_and_gradient(xi1)[0]
    assert parts[2] == 0.0 and parts[0] == 0.0
    rho0 = w.n0 / (8 * np.pi ** 2 / 2)
    d = 0.2 * rho0 * rng.uniform(-1, 1, nno)
    f1 = eng.excess_energy(np.sqrt(1 + d / rho0))
    f2 = eng.excess_energy(np.sqrt(1 + 2 * d / rho0))
    say(f"  F_exc(d)={f1} F_exc(2d)={f2}")
    assert abs(f2 - 4 * f1) < 1e-7 * abs(f1)
    xi = np.sqrt(1 + d / rho0)
    pa, dfa = eng.energy_and_gradient(xi)
    pb, dfb = eng.energy_and_gradient(xi)
    assert np.array_equal(pa, pb) and np.array_equal(dfa, dfb)
    bx, bd = eng.alloc_field(), eng.alloc_field()
    bx.upload(xi)
    pc, _ = eng.energy_and_gradient(bx, bd)
    assert np.array_equal(pa, pc) and np.array_equal(bd.download(), dfa)
    del df


def test_mmax6_refuses_fast_path_flags():
    """An FP64 context of mmax 6 runs the generic transforms, which have no FP32
    gradient, scaled variables or fused products: those flags are refused."""
    import ctypes as C
    from mdft.gpu import DEVICE_PTRS, DF_FP32, SCALED_VARS, DeviceBuffer, Engine
    say("[mmax 6] fast-path flags refused")
    n, L = (8, 8, 8), (8.0, 8.0, 8.0)
    eng = Engine(n, L, 6, 2, 2.4777, 0.0333, "fp64")
    from mdft.ck import CkTable
    tup = ck.tuples(6)
    tab = CkTable(*(np.array([u[i] for u in tup], np.int32) for i in range(5)), 0.05,
                  np.zeros((400, len(tup)), np.complex128))
    eng.set_ck(tab)
    eng.set_vext(np.zeros(eng.N * eng.no))
    xi = DeviceBuffer(eng, eng.N * eng.no * 8)
    xi.upload(np.ones(eng.N * eng.no))
    df = DeviceBuffer(eng, eng.N * eng.no * 8)
    parts = np.zeros(3)
    pp = parts.ctypes.data_as(C.POINTER(C.c_double))
    for fl in (DF_FP32, SCALED_VARS):
        rc = eng.lib.mdft_energy_and_gradient(eng._h, xi.ptr, pp, df.ptr, DEVICE_PTRS | fl)
        assert rc == -1, fl
        assert b"mmax <= 5" in eng.lib.mdft_last_error(eng._h)
    assert eng.lib.mdft_energy_and_gradient(eng._h, xi.ptr, pp, df.ptr, DEVICE_PTRS) == 0

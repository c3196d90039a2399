"""GPU parity of the FP32 path (MDFT_PREC_FP32: angular transforms, FFTs, the OZ
contraction and the per-orientation ideal, external and excess terms in FP32;
the sums in FP64) against the FP64 CPU oracle on identical inputs.

Tolerances (FP32 path): every F part relative 1e-5, total F relative 1e-5,
dF max abs difference <= 1e-5 of max |dF|, projection stages <= 1e-5 of the
stage's max modulus.  At 128^3 (too large for the oracle) the FP32 engine is
compared with the FP64 engine, itself pinned to the oracle at 1e-10."""
import sys

import numpy as np
import pytest

from mdft import ck, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _inputs(N, no, seed):
    rng = np.random.default_rng(seed)
    vext = rng.uniform(-3.0, 6.0, N * no)
    vext[::97] = 694.7
    xi = 1.0 + 0.3 * rng.uniform(-1, 1, N * no)
    xi[::53] = 0.0
    xi[::71] = 1e-9
    return xi, vext


CASES = [
    ("mmax1 48^3", (48, 48, 48), (24.0, 24.0, 24.0), 1, 512),
    ("mmax3 24x20x18", (24, 20, 18), (14.0, 13.0, 12.0), 3, 256),
    ("mmax5 16^3", (16, 16, 16), (12.0, 12.0, 12.0), 5, 256),
    ("mmax4 12x10x15", (12, 10, 15), (10.0, 9.0, 11.0), 4, 256),
    ("fft x192 192x8x6", (192, 8, 6), (48.0, 4.0, 3.0), 1, 1024),
    ("fft y96 10x96x12", (10, 96, 12), (5.0, 30.0, 6.0), 2, 1024),
    ("fft z192 8x6x192", (8, 6, 192), (4.0, 3.0, 60.0), 1, 1024),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp32_parity(oracle_mod, case):
    from mdft.gpu import Engine
    name, n, L, mmax, nk = case
    say(f"[fp32 parity] {name} n={n} mmax={mmax}")
    t = ck.synthetic_ck(mmax, nk, 0.05, seed=11 + mmax, n0=0.0333)
    eng = Engine(n, L, mmax, 2, 2.4777, 0.0333, "fp32")
    eng.set_ck(t)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    xi, vext = _inputs(eng.N, eng.no, seed=3)
    eng.set_vext(vext)
    for stage in (1, 2, 3, 4):
        g = eng.excess_stage(xi, stage)
        r = o.excess_stage(xi, stage)
        err = np.abs(g - r).max() / np.abs(r).max()
        say(f"  stage {stage}: max rel err {err:.2e}")
        assert err < TOL, (stage, err)
    parts_g, df_g = eng.energy_and_gradient(xi)
    parts_o, df_o = o.energy_and_gradient(xi, vext)
    rel = np.abs(parts_g - parts_o) / np.abs(parts_o)
    relF = abs(parts_g.sum() - parts_o.sum()) / abs(parts_o.sum())
    say(f"  F parts rel {rel}, F rel {relF:.2e}")
    assert np.all(rel < TOL) and relF < TOL
    derr = np.abs(df_g - df_o).max() / np.abs(df_o).max()
    say(f"  dF max rel err {derr:.2e}")
    assert derr < TOL


def test_fp32_full_size_against_fp64_engine():
    from mdft.gpu import Engine
    w = workloads.WORKLOADS["solute20_128"]
    say(f"[fp32 full] {w.name} n={w.n} mmax={w.mmax}")
    t = w.ck_table()
    sol = w.solute()
    rng = np.random.default_rng(2)
    e64 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp64")
    e64.set_ck(t)
    e64.compute_vext(sol, w.solvent)
    vext = e64.get_vext()
    xi = workloads.guess_xi(vext, w.kT) * (1.0 + 0.05 * rng.standard_normal(vext.shape))
    p64, d64 = e64.energy_and_gradient(xi)
    del e64
    e32 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp32")
    e32.set_ck(t)
    e32.set_vext(vext)
    p32, d32 = e32.energy_and_gradient(xi)
    rel = np.abs(p32 - p64) / np.abs(p64)
    relF = abs(p32.sum() - p64.sum()) / abs(p64.sum())
    derr = np.abs(d32 - d64).max() / np.abs(d64).max()
    say(f"  F parts rel {rel}, F rel {relF:.2e}, dF max rel {derr:.2e}")
    assert np.all(rel < TOL) and relF < TOL and derr < TOL

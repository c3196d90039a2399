"""GPU parity of the BF16 path (MDFT_PREC_BF16: angular transforms as BF16
MFMA GEMMs with FP32 accumulation, FP32 FFTs and OZ; ideal term, reductions and
dF assembly in FP64) against the FP64 CPU oracle on identical inputs.

Tolerances (BF16 path): total F relative 2e-3; F_exc alone relative 5e-3
(F_id and F_ext are FP64 and must agree to 1e-10); dF max abs difference
<= 2e-2 of max |dF|; the angular transform (stage 1) <= 1e-2 of its max
modulus.  At 128^3 the BF16
engine is compared with the FP64 engine, itself pinned to the oracle."""
import sys

import numpy as np
import pytest

from mdft import ck, workloads

pytestmark = pytest.mark.gpu

TOL = 2e-3


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
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_bf16_parity(oracle_mod, case):
    from mdft.gpu import Engine
    name, n, L, mmax, nk = case
    say(f"[bf16 parity] {name} n={n} mmax={mmax}")
    t = ck.synthetic_ck(mmax, nk, 0.05, seed=11 + mmax, n0=0.0333)
    eng = Engine(n, L, mmax, 2, 2.4777, 0.0333, "bf16")
    eng.set_ck(t)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    xi, vext = _inputs(eng.N, eng.no, seed=3)
    eng.set_vext(vext)
    g = eng.excess_stage(xi, 1)
    r = o.excess_stage(xi, 1)
    err = np.abs(g - r).max() / np.abs(r).max()
    say(f"  stage 1 (BF16 angl2proj): max rel err {err:.2e}")
    assert err < 1e-2
    parts_g, df_g = eng.energy_and_gradient(xi)
    parts_o, df_o = o.energy_and_gradient(xi, vext)
    rel = np.abs(parts_g - parts_o) / np.abs(parts_o)
    relF = abs(parts_g.sum() - parts_o.sum()) / abs(parts_o.sum())
    say(f"  F parts rel {rel}, F rel {relF:.2e}")
    assert np.all(rel[:2] < 1e-10) and rel[2] < 5e-3 and relF < TOL
    derr = np.abs(df_g - df_o).max() / np.abs(df_o).max()
    say(f"  dF max rel err {derr:.2e}")
    assert derr < 2e-2


def test_bf16_full_size_against_fp64_engine():
    from mdft.gpu import Engine
    w = workloads.WORKLOADS["solute20_128"]
    say(f"[bf16 full] {w.name} n={w.n} mmax={w.mmax}")
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
    e32 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "bf16")
    e32.set_ck(t)
    e32.set_vext(vext)
    p32, d32 = e32.energy_and_gradient(xi)
    rel = np.abs(p32 - p64) / np.abs(p64)
    relF = abs(p32.sum() - p64.sum()) / abs(p64.sum())
    derr = np.abs(d32 - d64).max() / np.abs(d64).max()
    say(f"  F parts rel {rel}, F rel {relF:.2e}, dF max rel {derr:.2e}")
    assert rel[2] < 5e-3 and relF < TOL and derr < 2e-2

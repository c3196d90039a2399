"""The HIP path against golden vectors of the reference's own evaluation.

tests/golden/ref_energy_golden.npz holds F_id, F_ext, F_exc and seeded dF
samples computed by the reference's energy_ideal_and_external +
energy_cproj_mrso (src/module_energy_ideal_and_external.f90:14-74,
src/module_energy_cproj_mrso.f90:61-782), compiled from its sources
(tests/golden/make_ref_golden.py).  The cases: odd and even grids, mmax 1, 2,
3 and 5, synthetic tables and the SPC/E tables the reference ships, and the
workload shapes 48^3 / 24 A / mmax 1 and 96^3 / 32 A / mmax 3 (the Na+ case,
whose 96-point axes run the radix-3 FFT kernels).

Bars: FP64 context -- every F part within 1e-10 relative, dF samples within
1e-10 of max |dF|; FP32 context -- every F part and F within 1e-5 relative,
dF samples within 1e-5 of max |dF|.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import ref_cases as RC  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def golden():
    return np.load(RC.GOLDEN)


def _table(g, c):
    name, n, L, mmax, kind, tseed, nk, _, _ = c
    if kind == "synthetic":
        from oracle import ref_eval as R
        # the table as the reference read it (its null first column reads as zero)
        t = R.table_as_read(RC.synthetic_table(mmax, tseed, nk))
        assert t.dq == float(g[name + "/ck_dq"])
        assert abs(np.abs(t.c).sum() - float(g[name + "/ck_sum"])) <= 1e-12 * float(g[name + "/ck_sum"])
        return t
    return RC.table_from_fixture(g, name)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
@pytest.mark.parametrize("c", RC.CASES, ids=[c[0] for c in RC.CASES])
def test_engine_matches_reference_golden(golden, c, prec, tol):
    from mdft.gpu import Engine
    name, n, L, mmax, kind, tseed, nk, iseed, size = c
    say(f"[reference golden] {name} {prec} n={n} mmax={mmax} table={kind}")
    kT, n0 = (float(v) for v in golden[name + "/consts"][:2])
    t = _table(golden, c)
    xi, vext = RC.inputs(n, RC.no_of(mmax), iseed)
    st = golden[name + "/in_stats"]
    assert abs(xi.sum() - st[0]) <= 1e-12 * abs(st[0]) and abs(vext.sum() - st[1]) <= 1e-12 * abs(st[1])
    eng = Engine(n, L, mmax, 2, kT, n0, prec)
    eng.set_ck(t)
    eng.set_vext(vext)
    parts, df = eng.energy_and_gradient(xi)
    ref = golden[name + "/parts"]
    rel = np.abs(parts - ref) / np.abs(ref)
    relF = abs(parts.sum() - ref.sum()) / abs(ref.sum())
    idx, val = golden[name + "/df_idx"], golden[name + "/df_val"]
    derr = np.abs(df[idx] - val).max() / golden[name + "/df_stats"][2]
    say(f"  F parts {parts} reference {ref} rel {rel} F rel {relF:.2e} dF rel {derr:.2e}")
    assert np.all(rel < tol) and relF < tol, rel
    assert derr < tol, derr
    # F only (module_pressure_correction.f90:124 calls without df): same F
    parts_f, _ = eng.energy_and_gradient(xi, want_df=False)
    assert np.all(np.abs(parts_f - parts) <= 1e-12 * np.abs(parts))

"""The FP64 CPU oracle against golden vectors of the reference's own evaluation.

tests/golden/ref_energy_golden.npz holds F_id, F_ext, F_exc and seeded dF
samples that the reference's energy_ideal_and_external + energy_cproj_mrso
(src/module_energy_ideal_and_external.f90:14-74,
src/module_energy_cproj_mrso.f90:61-782), compiled here from its sources with
FFTW 3.3.4 built from its source tarball, computed on the cases of
tests/golden/ref_cases.py: odd and even grids, mmax 1, 2, 3 and 5, synthetic
tables and the SPC/E tables the reference ships, and the 48^3 / mmax 1 and
96^3 / mmax 3 workload shapes.  These fix the conventions no closed form pins:
the conj and sign folding of c (:415,420-429), the -q pairing with
R(., -khi)(-1)^m (:517-520,636-638), the Nyquist conj (:652-656), the epsilon
chops (:491-492,618-623) and the prefactor (:728).

Bar: F parts within 1e-12 relative, dF samples within 1e-12 of max |dF| (the
oracle follows the reference's arithmetic order closely enough that the
measured differences are ~1e-15).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import ref_cases as RC  # noqa: E402


@pytest.fixture(scope="module")
def golden():
    return np.load(RC.GOLDEN)


def _table(g, c):
    name, n, L, mmax, kind, tseed, nk, _, _ = c
    if kind == "synthetic":
        from oracle import ref_eval as R
        t = R.table_as_read(RC.synthetic_table(mmax, tseed, nk))
        assert abs(np.abs(t.c).sum() - float(g[name + "/ck_sum"])) <= 1e-12 * float(g[name + "/ck_sum"]), \
            "the synthetic c(k) table no longer matches the one the golden vectors were made with"
        return t
    return RC.table_from_fixture(g, name)


def _inputs(g, c):
    name, n, L, mmax, kind, tseed, nk, iseed, _ = c
    xi, vext = RC.inputs(n, RC.no_of(mmax), iseed)
    st = g[name + "/in_stats"]
    assert abs(xi.sum() - st[0]) <= 1e-12 * abs(st[0]) and abs(vext.sum() - st[1]) <= 1e-12 * abs(st[1]), \
        "the seeded inputs no longer match the ones the golden vectors were made with"
    return xi, vext


@pytest.mark.parametrize("c", RC.CASES, ids=[c[0] for c in RC.CASES])
def test_oracle_matches_reference_golden(oracle_mod, golden, c):
    name, n, L, mmax = c[:4]
    if c[8] == "large" and os.environ.get("MDFT_SKIP_LARGE_CPU"):
        pytest.skip("large case skipped by MDFT_SKIP_LARGE_CPU")
    kT, n0, rho0, dv = golden[name + "/consts"]
    t = _table(golden, c)
    xi, vext = _inputs(golden, c)
    o = oracle_mod.Oracle(n, L, mmax, 2, float(kT), float(n0))
    o.set_ck(t)
    assert abs(o.rho0 - rho0) <= 1e-15 * rho0 and abs(o.dv - dv) <= 1e-14 * dv
    parts, df = o.energy_and_gradient(xi, vext)
    ref = golden[name + "/parts"]
    rel = np.abs(parts - ref) / np.abs(ref)
    print(f"{name}: oracle {parts} reference {ref} rel {rel}")
    assert np.all(rel < 1e-12), rel
    idx, val = golden[name + "/df_idx"], golden[name + "/df_val"]
    dmax = golden[name + "/df_stats"][2]
    err = np.abs(df[idx] - val).max() / dmax
    assert err < 1e-12, err
    st = golden[name + "/df_stats"]
    assert abs(df.sum() - st[0]) <= 1e-10 * np.abs(df).sum()
    assert abs(np.dot(df, df) - st[1]) <= 1e-12 * st[1]


def test_golden_planted_conjugation_fails(oracle_mod, golden):
    """The goldens reject the conjugation mistake self-consistency tests cannot see:
    the oracle fed conj(c) (the c(k) conj of cproj_mrso.f90:415 dropped) must miss
    the reference's F_exc by far more than the bar: by 2 % with the shipped SPC/E
    table at mmax 2, and by ~1e-6 with the synthetic mmax 3 table, whose odd-l
    part is small next to its real self term (still 1e5 times the bar)."""
    for name, floor in (("even_mmax2_spce", 1e-2), ("odd_mmax3_synth", 1e-7)):
        _planted(oracle_mod, golden, name, floor)


def _planted(oracle_mod, golden, name, floor):
    c = RC.case(name)
    name, n, L, mmax = c[:4]
    kT, n0 = golden[name + "/consts"][:2]
    t = _table(golden, c)
    t.c = np.conj(t.c)
    xi, vext = _inputs(golden, c)
    o = oracle_mod.Oracle(n, L, mmax, 2, float(kT), float(n0))
    o.set_ck(t)
    parts, _ = o.energy_and_gradient(xi, vext, want_df=False)
    ref = golden[name + "/parts"]
    assert abs(parts[2] - ref[2]) / abs(ref[2]) > floor


REF_EXE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "mdft_ref_eval")


@pytest.mark.skipif(not os.path.exists(REF_EXE) or not os.path.isdir("/root/reference/data"),
                    reason="needs the reference build (oracle/_ref) and the reference tree")
def test_golden_is_current_with_reference_build(golden):
    """Re-runs the reference on one small case: the stored vectors are what the
    reference computes now."""
    from oracle import ref_eval as R
    c = RC.case("even_mmax3_synth")
    name, n, L, mmax, kind, tseed, nk, iseed, _ = c
    t = RC.synthetic_table(mmax, tseed, nk)
    xi, vext = RC.inputs(n, RC.no_of(mmax), iseed)
    r = R.run(n, L, mmax, RC.TEMPERATURE, lambda p: R.write_ck_file_for_flang(p, t), xi, vext)
    assert np.allclose(r.parts, golden[name + "/parts"], rtol=1e-14, atol=0)
    assert np.array_equal(r.df[golden[name + "/df_idx"]], golden[name + "/df_val"])

"""c(k) tables: counts, file format, symmetry laws, relayout."""
import os

import numpy as np
import pytest

from mdft import ck

REF_DCF = "/root/reference/data/dcf/spce"


def test_counts_match_reference_tables():
    # solvent%npluc(0:5) of src/module_solvent.f90:223 and npc of cproj_mrso.f90:381-383
    assert [ck.npluc(m) for m in range(6)] == [1, 6, 75, 252, 877, 2002]
    assert [ck.npc(m) for m in range(6)] == [1, 6, 56, 178, 573, 1268]


def test_nq_needed():
    # 128^3 in a 40 A box with dq = 0.05: qmax = sqrt(3)*pi/0.3125
    assert ck.nq_needed((128, 128, 128), (40.0, 40.0, 40.0), 0.05) == int(np.sqrt(3) * np.pi / 0.3125 / 0.05 + 0.5) + 2


def _sym_errors(t):
    pos = {(int(t.m[i]), int(t.n[i]), int(t.mu[i]), int(t.nu[i]), int(t.khi[i])): i for i in range(t.np_file)}
    e = [0.0, 0.0, 0.0]
    for (m, n, mu, nu, khi), i in pos.items():
        j = pos[(m, n, -mu, -nu, khi)]
        e[0] = max(e[0], np.abs(t.c[:, j] - (-1) ** (m + n + mu + nu) * np.conj(t.c[:, i])).max())
        j = pos[(n, m, nu, mu, khi)]
        e[1] = max(e[1], np.abs(t.c[:, j] - (-1) ** (m + n) * t.c[:, i]).max())
        e[2] = max(e[2], np.abs(t.c[:, pos[(m, n, -mu, -nu, khi)]] - t.c[:, pos[(m, n, mu, nu, -khi)]]).max())
    return e


@pytest.mark.parametrize("mmax", [1, 2, 3, 5])
def test_synthetic_obeys_symmetry_laws(mmax):
    """luc-files/ck_nmax1:5-7: c(m,n,-mu,-nu;khi) = (-1)^(m+n+mu+nu) conj c,
    c(n,m,nu,mu;khi) = (-1)^(m+n) c, and for C2v c(m,n,-mu,-nu;khi) = c(m,n,mu,nu;-khi)."""
    t = ck.synthetic_ck(mmax, 128, 0.05, seed=3, n0=0.0333)
    assert t.np_file == ck.npluc(mmax)
    assert max(_sym_errors(t)) < 1e-15
    r = ck.synthetic_ck(mmax, 128, 0.05, seed=3, n0=0.0333, self_term=0.0)  # the random part alone
    assert ck.spectral_norm_max(r, mmax) <= 0.8 + 1e-12
    ck.add_self_term(r, -ck.self_profile(r, mmax, 2, 15.0, 3.0))
    assert np.array_equal(r.c, t.c)
    # at k = 0 only the l = 0 part survives, which needs m = n
    for i in range(t.np_file):
        if t.m[i] != t.n[i]:
            assert abs(t.c[0, i]) < 1e-15


def test_synthetic_is_seeded():
    a = ck.synthetic_ck(2, 64, 0.05, seed=7, n0=0.0333)
    b = ck.synthetic_ck(2, 64, 0.05, seed=7, n0=0.0333)
    c = ck.synthetic_ck(2, 64, 0.05, seed=8, n0=0.0333)
    assert np.array_equal(a.c, b.c) and not np.array_equal(a.c, c.c)


def test_file_roundtrip(tmp_path):
    t = ck.synthetic_ck(2, 40, 0.05, seed=1, n0=0.0333)
    p = tmp_path / "spce-ck_nonzero_nmax2_ml"
    ck.write_ck_file(str(p), t)
    u = ck.read_ck_file(str(p))
    assert np.array_equal(u.m, t.m) and np.array_equal(u.khi, t.khi)
    assert np.allclose(u.c, t.c, rtol=0, atol=1e-15)
    assert abs(u.dq - 0.05) < 1e-15


@pytest.mark.skipif(not os.path.isdir(REF_DCF), reason="reference data not present")
@pytest.mark.parametrize("nmax", [0, 1, 2])
def test_reference_files_parse_and_obey_laws(nmax):
    t = ck.read_ck_file(os.path.join(REF_DCF, f"spce-ck_nonzero_nmax{nmax}_ml"))
    assert t.np_file == ck.npluc(nmax) and t.nq_file == 1024
    assert max(_sym_errors(t)) < 1e-12


def test_relayout_matches_oracle(oracle_mod):
    t = ck.synthetic_ck(3, 200, 0.05, seed=2, n0=0.0333)
    o = oracle_mod.Oracle((8, 8, 8), (10.0, 10.0, 10.0), 3, 2, 2.4777, 0.0333)
    o.set_ck(t)
    nq = ck.nq_needed((8, 8, 8), (10.0, 10.0, 10.0), 0.05)
    assert o.nq == nq and o.npc == ck.npc(3)
    assert np.array_equal(o.get_ck(), ck.relayout(t, 3, 2, nq))


def test_relayout_rejects_missing_tuples(oracle_mod):
    t = ck.synthetic_ck(2, 200, 0.05, seed=2, n0=0.0333)
    keep = np.arange(t.np_file) != 5
    bad = ck.CkTable(t.m[keep], t.n[keep], t.mu[keep], t.nu[keep], t.khi[keep], t.dq, t.c[:, keep])
    o = oracle_mod.Oracle((8, 8, 8), (10.0, 10.0, 10.0), 2, 2, 2.4777, 0.0333)
    with pytest.raises(ValueError):
        o.set_ck(bad)

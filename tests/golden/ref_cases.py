"""Cases of the reference golden vectors (tests/golden/ref_energy_golden.npz).

Each case fixes a grid (sizes with factors 2, 3 and 5 only, the FFT radices of
the device kernels), an angular order, a c(k) table and seeded inputs
(xi, vext).  `make_ref_golden.py` ran the reference's own evaluation on them
(energy_ideal_and_external + energy_cproj_mrso, compiled from the reference's
sources, oracle/ref_eval.f90) and stored F_id, F_ext, F_exc and seeded samples
of dF; the tests regenerate the inputs from the same definitions and compare.

Tables:
* "synthetic": mdft.ck.synthetic_ck(mmax, nk, 0.05, seed, n0=0.0333); for
  mmax = 1 the (m, n, khi) = (1, 1, -1) column is set to zero (the column the
  amdflang-readable c-file writes as a null value, oracle/ref_eval.py).
* "spce_nmax1" / "spce_nmax2": the SPC/E tables the reference ships
  (data/dcf/spce/spce-ck_nonzero_nmax{1,2}_ml), the rows the grid needs stored
  in the fixture (nmax1 with its (1, 1, -1) column set to zero, as above).
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "ref_energy_golden.npz")
TEMPERATURE = 298.0
NSAMPLE = 4096

# name, grid, box (A), mmax, table kind, table seed, rows (synthetic), input seed, size class
CASES = [
    ("odd_mmax1_synth", (9, 15, 5), (9.0, 13.5, 5.5), 1, "synthetic", 101, 256, 11, "small"),
    ("even_mmax1_spce", (8, 10, 6), (8.0, 9.0, 7.0), 1, "spce_nmax1", 0, 0, 12, "small"),
    ("even_mmax2_spce", (10, 8, 12), (9.0, 8.0, 11.0), 2, "spce_nmax2", 0, 0, 13, "small"),
    ("odd_mmax2_synth", (5, 9, 15), (5.5, 9.5, 14.0), 2, "synthetic", 102, 256, 14, "small"),
    ("odd_mmax3_synth", (9, 15, 5), (9.0, 14.0, 5.0), 3, "synthetic", 103, 256, 15, "small"),
    ("even_mmax3_synth", (12, 10, 8), (11.0, 9.0, 8.0), 3, "synthetic", 104, 256, 16, "small"),
    ("odd_mmax5_synth", (5, 9, 15), (5.0, 8.0, 13.0), 5, "synthetic", 105, 256, 17, "small"),
    ("even_mmax5_synth", (10, 8, 12), (9.0, 8.0, 11.0), 5, "synthetic", 106, 256, 18, "small"),
    # the workloads' own shapes: LJ sphere 48^3 / 24 A / mmax 1 (512 rows), Na+ 96^3 / 32 A / mmax 3 (1024 rows)
    ("lj_sphere_48_mmax1", (48, 48, 48), (24.0, 24.0, 24.0), 1, "synthetic", 1000, 512, 7, "large"),
    ("na_ion_96_mmax3", (96, 96, 96), (32.0, 32.0, 32.0), 3, "synthetic", 1001, 1024, 8, "large"),
]


def case(name):
    for c in CASES:
        if c[0] == name:
            return c
    raise KeyError(name)


def synthetic_table(mmax, seed, nk):
    from mdft import ck
    t = ck.synthetic_ck(mmax, nk, 0.05, seed=seed, n0=0.0333)
    if mmax == 1:
        _zero_khi_m1(t)
    return t


def _zero_khi_m1(t):
    idx = [k for k in range(t.np_file) if (int(t.m[k]), int(t.n[k]), int(t.khi[k])) == (1, 1, -1)]
    t.c[:, idx] = 0.0


def table_from_fixture(g, name):
    """The c(k) table of a stored-table case, as the reference read it."""
    from mdft.ck import CkTable
    p = f"{name}/ck_"
    c = g[p + "re"] + 1j * g[p + "im"]
    return CkTable(g[p + "m"], g[p + "n"], g[p + "mu"], g[p + "nu"], g[p + "khi"], float(g[p + "dq"]), c,
                   g[p + "q"])


def inputs(n, no, seed):
    """Seeded xi and vext (reference layout, orientation fastest): vext uniform in
    [-3, 6] kJ/mol with every 97th value at the clip qfact/2 = 694.7; xi = 1 + 0.3 U(-1, 1)
    with every 53rd value 0 and every 71st 1e-9 (the rho <= eps branch of the ideal term)."""
    N = n[0] * n[1] * n[2]
    rng = np.random.default_rng(seed)
    vext = rng.uniform(-3.0, 6.0, N * no)
    vext[::97] = 694.7
    xi = 1.0 + 0.3 * rng.uniform(-1.0, 1.0, N * no)
    xi[::53] = 0.0
    xi[::71] = 1e-9
    return xi, vext


def sample_index(size, seed):
    rng = np.random.default_rng(10_000 + seed)
    k = min(NSAMPLE, size)
    return np.sort(rng.choice(size, size=k, replace=False))


def no_of(mmax, mrso=2):
    return (mmax + 1) * (2 * mmax + 1) * (2 * (mmax // mrso) + 1)

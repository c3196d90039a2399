"""GPU tests of vext with the reference's Poisson electrostatics
(mdft_compute_vext_ex, MDFT_ELEC_POISSON) against the CPU restatement:
Lennard-Jones (workloads.vext_numpy without its Coulomb term) plus
oracle/poisson_oracle.vextq, clipped at qfact/2.  Tolerance 1e-9 of the largest
unclipped magnitude (the device FFT differs from numpy's in rounding only)."""
import sys

import numpy as np
import pytest

from mdft import workloads
from mdft.constants import QFACT
from oracle import poisson_oracle as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(n, length, mmax, seed, natom=4):
    w = workloads.Workload("poisson", 3, n, length, mmax, 256)
    rng = np.random.default_rng(seed)
    c = np.array(length) / 2
    xyz = c + rng.uniform(-1.5, 1.5, (natom, 3))
    q = rng.normal(0, 0.6, natom)
    q -= q.mean() - 0.3                      # net charge 0.3 e: the k = 0 term matters
    sol = workloads.Solute(xyz, q, rng.uniform(2.5, 3.5, natom), rng.uniform(0.2, 0.9, natom))
    return w, sol


@pytest.mark.parametrize("n,length,mmax", [((16, 12, 18), (8.0, 6.0, 9.0), 1), ((24, 24, 20), (10.0, 10.0, 8.5), 3),
                                           ((15, 16, 12), (7.5, 8.0, 6.0), 2)])
def test_poisson_vext_matches_oracle(n, length, mmax):
    from mdft.gpu import Engine
    w, sol = _case(n, length, mmax, seed=sum(n))
    say(f"[poisson] vext on {n} / {length} A, mmax={mmax}, {len(sol.q)} atoms, net charge {sol.q.sum():.2f}")
    eng = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp64")
    eng.compute_vext(sol, w.solvent, electrostatics="poisson")
    got = eng.get_vext()
    lj = workloads.vext_numpy(w, sol, clip=np.inf, coulomb=False)
    offs = workloads.orientation_site_offsets(w.mmax, w.mrso, w.solvent)
    sq = np.array([s.q for s in w.solvent.sites])
    vq = P.vextq(w.n, w.length, sol.xyz, sol.q, offs, sq, QFACT)
    ref = np.minimum(lj + vq, QFACT / 2.0)
    unclipped = ref < QFACT / 2.0
    err = np.abs(got.reshape(ref.shape) - ref).max()
    scale = np.abs(ref[unclipped]).max()
    say(f"  max abs err {err:.2e} (largest unclipped |vext| {scale:.1f}, electrostatic part up to {np.abs(vq).max():.1f})")
    assert err <= 1e-9 * scale


def test_uncharged_solute_gives_the_same_vext_for_every_electrostatics():
    from mdft.gpu import Engine
    w, sol = _case((12, 12, 12), (6.0, 6.0, 6.0), 1, seed=3)
    sol = workloads.Solute(sol.xyz, np.zeros_like(sol.q), sol.sig, sol.eps)
    eng = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp64")
    out = []
    for e in ("poisson", "direct", "none"):
        eng.compute_vext(sol, w.solvent, electrostatics=e)
        out.append(eng.get_vext())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])

"""Seeded synthetic workloads: grids, solutes, c(k) and densities.

Five workload definitions (rigid SPC/E-like water, n0 = 0.0333 A^-3, T = 298 K):

  lj_sphere_48   one neutral LJ sphere, 24 A box, 48^3, mmax=1, c(k) 512 rows
  na_ion_96      Na+-like ion, 32 A box, 96^3, mmax=3, c(k) 1024 rows
  solute20_128   20-atom seeded solute, 40 A box, 128^3, mmax=5, c(k) 1024 rows
  warmstart_64   5-30 atom solutes, 24 A box, 64^3, mmax=3 (training set)
  globular_192   2000-atom globule, 64 A box, 192^3, mmax=5, c(k) 2048 rows

c(k) rows are spaced dk = 0.05 A^-1 and seeded with 1000 + workload index;
solutes with 2000 + index; perturbed densities with 7 + index.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import ck as ckmod
from .constants import QFACT, fp32, kT_of, spce


@dataclass
class Solute:
    xyz: np.ndarray   # [n, 3] Angstrom (absolute box coordinates)
    q: np.ndarray     # [n] e
    sig: np.ndarray   # [n] Angstrom
    eps: np.ndarray   # [n] kJ/mol

    @property
    def net_charge(self) -> float:
        return float(self.q.sum())


@dataclass
class Workload:
    name: str
    index: int
    n: tuple
    length: tuple
    mmax: int
    nk: int
    dk: float = 0.05
    temperature: float = 298.0
    n0: float = 0.0333
    mrso: int = 2
    extra: dict = field(default_factory=dict)

    @property
    def kT(self) -> float:
        return kT_of(self.temperature)

    @property
    def solvent(self):
        return spce(self.n0)

    def ck_table(self):
        return ckmod.synthetic_ck(self.mmax, self.nk, self.dk, seed=1000 + self.index, n0=self.n0, mrso=self.mrso)

    def solute(self, seed: int | None = None) -> Solute:
        return make_solute(self, 2000 + self.index if seed is None else seed)


WORKLOADS = {
    "lj_sphere_48": Workload("lj_sphere_48", 0, (48, 48, 48), (24.0, 24.0, 24.0), 1, 512),
    "na_ion_96": Workload("na_ion_96", 1, (96, 96, 96), (32.0, 32.0, 32.0), 3, 1024),
    "solute20_128": Workload("solute20_128", 2, (128, 128, 128), (40.0, 40.0, 40.0), 5, 1024),
    "warmstart_64": Workload("warmstart_64", 3, (64, 64, 64), (24.0, 24.0, 24.0), 3, 1024),
    "globular_192": Workload("globular_192", 4, (192, 192, 192), (64.0, 64.0, 64.0), 5, 2048),
}


def small(name: str, n: int, mmax: int | None = None) -> Workload:
    """A reduced grid of a workload (same box and chemistry), for oracle-sized tests."""
    w = WORKLOADS[name]
    return Workload(w.name + f"_n{n}", w.index, (n, n, n), w.length, w.mmax if mmax is None else mmax, w.nk,
                    w.dk, w.temperature, w.n0, w.mrso)


def _sphere_atoms(rng, natom, radius, dmin, center):
    pts = []
    tries = 0
    while len(pts) < natom:
        tries += 1
        if tries > 2_000_000:
            raise RuntimeError("could not place atoms")
        p = rng.uniform(-radius, radius, 3)
        if p @ p > radius * radius:
            continue
        if pts and np.min(np.sum((np.array(pts) - p) ** 2, axis=1)) < dmin * dmin:
            continue
        pts.append(p)
    return np.array(pts) + center


def make_solute(w: Workload, seed: int) -> Solute:
    rng = np.random.default_rng(seed)
    center = np.array(w.length) * 0.5
    if w.index == 0:
        return Solute(center[None, :].copy(), np.zeros(1), np.array([3.73]), np.array([1.23]))
    if w.index == 1:
        return Solute(center[None, :].copy(), np.array([1.0]), np.array([2.35]), np.array([0.544]))
    if w.index in (2, 3):
        nat = 20 if w.index == 2 else int(rng.integers(5, 31))
        xyz = _sphere_atoms(rng, nat, 5.0, 1.2, center)
        sig = rng.uniform(2.5, 4.0, nat)
        eps = rng.uniform(0.2, 1.0, nat)
        q = rng.normal(0.0, 0.4, nat)
        q -= q.mean()
        return Solute(xyz, q, sig, eps)
    xyz = _sphere_atoms(rng, 2000, 20.0, 1.0, center)
    sig = rng.uniform(2.5, 3.9, 2000)
    eps = rng.uniform(0.1, 0.9, 2000)
    q = rng.normal(0.0, 0.5, 2000)
    q -= q.mean()
    return Solute(xyz, q, sig, eps)


def orientation_site_offsets(mmax: int, mrso: int, solvent) -> np.ndarray:
    """Solvent site positions relative to the molecule's origin for every
    orientation io: [no, nsite, 3].  Euler angles (theta, phi, psi) and the
    rotation matrix of src/module_grid.f90:116-139,187-206."""
    from .grid import angular_grid
    g = angular_grid(mmax, mrso)
    out = np.zeros((g["no"], len(solvent.sites), 3))
    for io in range(g["no"]):
        th, ph, ps = g["theta"][io], g["phi"][io], g["psi"][io]
        ct, st, cp, sp, cs, ss = np.cos(th), np.sin(th), np.cos(ph), np.sin(ph), np.cos(ps), np.sin(ps)
        rot = np.array([[ct * cp * cs - sp * ss, -ct * cp * ss - sp * cs, st * cp],
                        [ct * sp * cs + cp * ss, -ct * sp * ss + cp * cs, st * sp],
                        [-st * cs, st * ss, ct]])
        for s, site in enumerate(solvent.sites):
            out[io, s] = rot @ np.array(site.r)
    return out


def vext_numpy(w: Workload, solute: Solute, rcut: float = 10.0, clip: float | None = None,
               coulomb: bool = True) -> np.ndarray:
    """External potential vext[N, no] (kJ/mol) on the host: Lennard-Jones between
    solute atoms and solvent sites (Lorentz-Berthelot, cut at rcut, minimum
    image) plus (coulomb) minimum-image Coulomb between charges, clipped at
    qfact/2 (src/module_solvent.f90:49, src/module_vext.f90:638-640; clip=inf:
    no clip).  For small grids."""
    solv = w.solvent
    offs = orientation_site_offsets(w.mmax, w.mrso, solv)
    no = offs.shape[0]
    nx, ny, nz = w.n
    L = np.array(w.length)
    dl = L / np.array(w.n)
    zz, yy, xx = np.meshgrid(np.arange(nz) * dl[2], np.arange(ny) * dl[1], np.arange(nx) * dl[0], indexing="ij")
    pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)  # [N,3], x fastest
    N = pts.shape[0]
    v = np.zeros((N, no))
    for s, site in enumerate(solv.sites):
        for io in range(no):
            pos = pts + offs[io, s]
            acc = np.zeros(N)
            for a in range(len(solute.q)):
                d = pos - solute.xyz[a]
                d -= L * np.round(d / L)
                r2 = np.sum(d * d, axis=1)
                r2 = np.maximum(r2, 1e-12)
                if site.eps > 0 and solute.eps[a] > 0:
                    e = np.sqrt(site.eps * solute.eps[a])
                    sg = 0.5 * (site.sig + solute.sig[a])
                    sr6 = (sg * sg / r2) ** 3
                    acc += np.where(r2 < rcut * rcut, 4.0 * e * (sr6 * sr6 - sr6), 0.0)
                if coulomb and site.q != 0.0 and solute.q[a] != 0.0:
                    acc += QFACT * site.q * solute.q[a] / np.sqrt(r2)
            v[:, io] += acc
    cap = QFACT / 2.0 if clip is None else clip
    return np.minimum(v, cap)


def guess_xi(vext: np.ndarray, kT: float) -> np.ndarray:
    """Cold start of the reference (src/module_density.f90:267-284): xi = 0 where
    vext >= (-ln(eps_dp) - 0.05) kT, xi = 1 elsewhere."""
    vmax = (-np.log(np.finfo(np.float64).eps) - 0.05) * kT
    return np.where(vext >= vmax, 0.0, 1.0)


def boltzmann_xi(vext: np.ndarray, kT: float) -> np.ndarray:
    """Cold-start guess xi = exp(-beta vext / 2) (rho = rho0 exp(-beta vext))."""
    return np.exp(-0.5 * np.minimum(vext, 50.0 * kT) / kT)


def perturbed_xi(vext: np.ndarray, kT: float, seed: int, amp: float = 0.05) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return boltzmann_xi(vext, kT) * (1.0 + amp * rng.standard_normal(vext.shape))


__all__ = ["Workload", "WORKLOADS", "Solute", "make_solute", "vext_numpy", "guess_xi", "boltzmann_xi", "perturbed_xi", "small",
           "fp32"]

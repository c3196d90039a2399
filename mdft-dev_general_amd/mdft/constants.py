"""Physical constants and the SPC/E solvent model used by the functional.

Values follow src/module_constants.f90:47-66 and src/module_solvent.f90:209-226
of the reference.  The SPC/E site parameters there are default-real (FP32)
literals widened to FP64; `fp32` reproduces that rounding where it matters for
parity (n0 = 0.0332891).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PI = float(np.arccos(-1.0))
EPS0 = 8.854187817e-12          # F/m, module_constants.f90:54
QUNIT = 1.602176565e-19         # C,   :55
NAVO = 6.02214129e23            # 1/mol, :56
BOLTZ = 1.3806488e-23           # J/K, :58
QFACT = QUNIT ** 2 * 1.0e-3 * NAVO / (4.0 * PI * EPS0 * 1.0e-10)  # kJ/mol * A / e^2, :57


def fp32(x: float) -> float:
    """Round a literal to IEEE single precision (a Fortran default-real literal)."""
    return float(np.float32(x))


def kT_of(temperature: float) -> float:
    """kT in kJ/mol, module_thermo.f90:26."""
    return BOLTZ * NAVO * temperature * 1.0e-3


@dataclass
class Site:
    q: float
    sig: float
    eps: float
    r: tuple


@dataclass
class Solvent:
    """A rigid solvent molecule (one species)."""
    name: str
    mrso: int
    n0: float
    sites: list = field(default_factory=list)
    relative_permittivity: float = 71.0

    @property
    def rho0(self) -> float:
        """Bulk density per orientation, module_solvent.f90:222."""
        return self.n0 / (8.0 * PI ** 2 / self.mrso)

    @property
    def quadrupole_trace(self) -> float:
        return float(sum(s.q * (s.r[0] ** 2 + s.r[1] ** 2 + s.r[2] ** 2) for s in self.sites))


def spce(n0: float | None = None) -> Solvent:
    """SPC/E water, module_solvent.f90:209-226 (FP32 literals widened to FP64).

    n0 defaults to the reference's 0.0332891 A^-3; pass 0.0333 for the
    workload definitions that state that bulk density."""
    q = [fp32(-0.8476), fp32(0.4238), fp32(0.4238)]
    sig = [fp32(3.166), 0.0, 0.0]
    eps = [fp32(0.65), 0.0, 0.0]
    r = [(0.0, 0.0, 0.0),
         (fp32(0.816495), 0.0, fp32(0.5773525)),
         (fp32(-0.816495), 0.0, fp32(0.5773525))]
    sites = [Site(q[i], sig[i], eps[i], r[i]) for i in range(3)]
    return Solvent("spce", 2, fp32(0.0332891) if n0 is None else float(n0), sites, 71.0)

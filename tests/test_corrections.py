"""Hunenberger's charged-solute corrections (mdft_charged_solute_corrections, host
arithmetic in the library, called here without a device) against their physics:
the type B constant -1971.01 kJ A/mol is xi_EW e^2 / (4 pi eps0) / 2 with the
Wigner constant of the cubic lattice xi_EW = -2.837297, the type C constant
2909.857 is (2 pi / 3) e^2 / (4 pi eps0) (CODATA values through scipy.constants),
and SPC/E's quadrupole trace is 2 q_H (1 A)^2 (O at the origin, O-H bonds of
1 A).  Reference: src/module_energy_and_gradient.f90:190-203, 287-344."""
import numpy as np
import pytest
import scipy.constants as K

from mdft import constants, workloads
from mdft.gpu import MdftError, charged_solute_corrections

# e^2 / (4 pi eps0) N_A in kJ A / mol
COULOMB = K.e ** 2 / (4 * np.pi * K.epsilon_0) * K.N_A / 1e3 * 1e10
XI_EW = -2.837297


def _ion(q, n=1):
    return workloads.Solute(np.zeros((n, 3)) + 5.0, np.full(n, q / n), np.full(n, 3.0), np.full(n, 0.5))


def test_constants_are_the_published_ones():
    assert abs(XI_EW * COULOMB / 2 - (-1971.01)) < 1e-5 * 1971.01
    assert abs(2 * np.pi / 3 * COULOMB - 2909.857) < 2e-6 * 2909.857


def test_sodium_and_chloride_in_spce():
    sol = constants.spce()
    trace = sol.quadrupole_trace
    qh = [s.q for s in sol.sites if s.q > 0][0]
    assert abs(trace - 2 * qh * 1.0 ** 2) < 1e-6                    # O at the origin, O-H = 1 A (FP32 literals)
    for q, L in ((1.0, 20.0), (-1.0, 20.0), (2.0, 31.5)):
        c = charged_solute_corrections((L, L, L), sol.n0, _ion(q), sol)
        pbc = XI_EW * COULOMB / 2 * q * q / L * (1 - 1 / sol.relative_permittivity)
        psc = -(2 * np.pi / 3) * COULOMB * trace * sol.n0 * q
        assert abs(c["pbc_correction"] - pbc) <= 1e-5 * abs(pbc)
        assert abs(c["pscheme_correction"] - psc) <= 2e-6 * abs(psc)
    # P-scheme odd in the charge, periodicity even
    a = charged_solute_corrections((20.0,) * 3, sol.n0, _ion(1.0), sol)
    b = charged_solute_corrections((20.0,) * 3, sol.n0, _ion(-1.0), sol)
    assert a["pbc_correction"] == b["pbc_correction"] and a["pscheme_correction"] == -b["pscheme_correction"]


def test_neutral_and_non_cubic():
    sol = constants.spce()
    zero = charged_solute_corrections((20.0,) * 3, sol.n0, _ion(0.0, 3), sol)
    assert zero == {"pbc_correction": 0.0, "pscheme_correction": 0.0}
    tiny = charged_solute_corrections((20.0,) * 3, sol.n0, _ion(5e-8), sol)     # |Q| <= 1e-7: no correction
    assert tiny == {"pbc_correction": 0.0, "pscheme_correction": 0.0}
    # a box that is not cubic takes its mean side; the charge of several atoms is summed
    c = charged_solute_corrections((18.0, 20.0, 25.0), sol.n0, _ion(1.0, 4), sol)
    ref = charged_solute_corrections((21.0,) * 3, sol.n0, _ion(1.0), sol)
    assert abs(c["pbc_correction"] - ref["pbc_correction"]) < 1e-12 * abs(ref["pbc_correction"])
    assert abs(c["pscheme_correction"] - ref["pscheme_correction"]) < 1e-12 * abs(ref["pscheme_correction"])
    with pytest.raises(MdftError):
        charged_solute_corrections((0.0, 0.0, 0.0), sol.n0, _ion(1.0), sol)

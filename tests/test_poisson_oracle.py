"""Pins of the Poisson electrostatics restatement (oracle/poisson_oracle.py) on
the CPU: charge conservation of the cloud-in-cell density, the zero box average
that the excluded k = 0 term imposes, the Coulomb law away from the charge,
translation covariance, and the c2r evaluation the device path uses (weighted
half spectrum, complex inverse, real part) against numpy's irfftn."""
import numpy as np
import pytest

from mdft.constants import QFACT
from oracle import poisson_oracle as P


def test_cic_density_conserves_charge_and_weights():
    n, L = (12, 10, 14), (6.0, 5.0, 7.0)
    rng = np.random.default_rng(0)
    xyz = rng.uniform(-3.0, 9.0, (5, 3))             # outside the box too: periodic
    q = rng.normal(size=5)
    rho = P.cic_density(n, L, xyz, q)
    dv = np.prod(np.array(L) / np.array(n))
    assert abs(rho.sum() * dv - q.sum()) < 1e-12
    # a charge on a node puts all of it there
    rho = P.cic_density(n, L, np.array([[1.5, 2.0, 3.5]]), np.array([2.0]))
    assert abs(rho[7, 4, 3] * dv - 2.0) < 1e-12 and np.count_nonzero(np.abs(rho) > 1e-14) == 1


def _one_site(no=1):
    return np.zeros((no, 1, 3)), np.array([1.0])


def test_vextq_box_average_is_zero():
    n, L = (16, 12, 18), (8.0, 6.0, 9.0)
    rng = np.random.default_rng(1)
    off = rng.normal(0, 0.6, (4, 3, 3))
    v = P.vextq(n, L, rng.uniform(0, 6, (3, 3)), np.array([0.5, -0.3, 0.8]), off, np.array([-0.8, 0.4, 0.4]), QFACT)
    assert np.abs(v.mean(axis=0)).max() < 1e-10 * np.abs(v).max()


def test_vextq_follows_coulomb_away_from_the_charge():
    """A unit solute charge on a node of a 64^3 / 32 A box and a unit site charge:
    along x, y and z at 4-9 A the potential fits A/r + B r^2 + C with A = qfact
    (the r^2 and constant terms are the periodic images and the neutralising
    background) within 1 %, and is isotropic."""
    n, L = (64, 64, 64), (32.0, 32.0, 32.0)
    off, sq = _one_site()
    v = P.vextq(n, L, np.array([[16.0, 16.0, 16.0]]), np.array([1.0]), off, sq, QFACT)[:, 0].reshape(64, 64, 64)
    r = np.arange(8, 19) * 0.5
    idx = (32 + np.arange(8, 19))
    for prof in (v[32, 32, idx], v[32, idx, 32], v[idx, 32, 32]):
        A = np.column_stack([1.0 / r, r * r, np.ones_like(r)])
        coef = np.linalg.lstsq(A, prof, rcond=None)[0]
        assert abs(coef[0] / QFACT - 1.0) < 0.01
    assert np.allclose(v[32, 32, idx], v[idx, 32, 32], rtol=1e-9)


def test_vextq_is_translation_covariant():
    n, L = (12, 10, 8), (6.0, 5.0, 4.0)
    rng = np.random.default_rng(2)
    off = rng.normal(0, 0.5, (3, 2, 3))
    xyz, q, sq = rng.uniform(0, 4, (2, 3)), np.array([0.7, -0.2]), np.array([1.0, -0.5])
    a = P.vextq(n, L, xyz, q, off, sq, QFACT).reshape(8, 10, 12, 3)
    b = P.vextq(n, L, xyz + np.array([0.5, 0.0, 1.0]), q, off, sq, QFACT).reshape(8, 10, 12, 3)
    assert np.abs(np.roll(a, shift=(2, 0, 1), axis=(0, 1, 2)) - b).max() < 1e-9 * np.abs(a).max()


@pytest.mark.parametrize("shape", [(6, 5, 8), (7, 4, 9), (4, 6, 5)])
def test_c2r_by_weighted_complex_inverse(shape):
    """c2r of a half spectrum (kx = 0 .. nx/2, arbitrary, not Hermitian) equals the
    real part of the complex inverse of the half spectrum weighted 1 on kx = 0 (and
    kx = nx/2 for even nx), 2 on the other kx <= nx/2, 0 beyond -- the device's way."""
    nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape))
    H = rng.normal(size=(nz, ny, nx // 2 + 1)) + 1j * rng.normal(size=(nz, ny, nx // 2 + 1))
    ref = np.fft.irfftn(H, s=shape, axes=(0, 1, 2))
    w = np.full(nx // 2 + 1, 2.0)
    w[0] = 1.0
    if nx % 2 == 0:
        w[-1] = 1.0
    Z = np.zeros((nz, ny, nx), dtype=complex)
    Z[:, :, : nx // 2 + 1] = H * w
    assert np.abs(np.fft.ifftn(Z).real - ref).max() < 1e-13

"""CPU restatement (numpy) of the electrostatic part of vext by the reference's
fast Poisson solver -- test infrastructure only: tests compare the device path
(mdft_compute_vext_ex with MDFT_ELEC_POISSON) with it.

Reference: src/module_solute.f90:193-232 (cloud-in-cell solute charges, with
floorNode / ceilingNode / distToFloorNode of src/module_mathematica.f90:134-220),
src/module_fastpoissonsolver.f90:121-288 (V(k) = 4 pi rho(k) / k^2 and, per
orientation, c2r[V(k) conj(sigma_io(k))] * qfact / N), src/module_solvent.f90:367-426
(sigma_io(k), the solvent molecule's charge density smeared with a Gaussian of
radius 0.5 A), src/module_vext.f90:600-669 (vext = vext_LJ + vextq, clipped at
qfact/2).

The r2c / c2r transforms are numpy's rfftn / irfftn over arrays [nz][ny][nx]:
the halved axis is x, as in FFTW's Fortran interface, and irfftn, like FFTW's
c2r, keeps only the real part of the self-conjugate planes (kx = 0 and, nx even,
kx = nx/2) after the y and z transforms.  Fields are returned [N][no] (x fastest).
"""
from __future__ import annotations

import math

import numpy as np

EPS = np.finfo(np.float64).eps
SMOOTHER_RADIUS = 0.5


def cic_density(n, length, xyz, q):
    """Solute charge density [nz][ny][nx] (e / A^3) of point charges q at xyz."""
    nx, ny, nz = n
    rho = np.zeros((nz, ny, nx))
    for s in range(len(q)):
        if not abs(q[s]) > EPS:
            continue
        m, p, wp = [], [], []
        for a in range(3):
            d = length[a] / n[a]
            x = float(xyz[s][a])
            f = math.floor((x % length[a]) / d) + 1          # floorNode (1-based; Fortran MODULO)
            m.append((f - 1) % n[a])
            p.append(f % n[a])                               # ceilingNode - 1 = MODULO(f, n)
            wp.append(abs(x / d - math.floor(x / d)))        # distToFloorNode
        wm = [1.0 - w for w in wp]
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    w = q[s] * (wp[0] if cx else wm[0]) * (wp[1] if cy else wm[1]) * (wp[2] if cz else wm[2])
                    rho[p[2] if cz else m[2], p[1] if cy else m[1], p[0] if cx else m[0]] += w
    return rho / ((length[0] / nx) * (length[1] / ny) * (length[2] / nz))


def _freqs(n, halved):
    """Integer frequencies of the Poisson kernel (i - 1 for i <= n/2, else i - 1 - n)."""
    m = np.arange(n)
    m = np.where(m < n // 2, m, m - n)
    return m[: n // 2 + 1] if halved else m


def _waves(n, length, halved):
    """The grid's wave vectors: (l - 1) 2 pi / L for l <= n/2 + 1, else (l - 1 - n) 2 pi / L."""
    m = np.arange(n)
    k = 2.0 * np.pi * np.where(m <= n // 2, m, m - n) / length
    return k[: n // 2 + 1] if halved else k


def vextq(n, length, xyz, q, site_off, site_q, qfact):
    """Electrostatic energy of the solvent sites, [N][no] (kJ/mol); zeros without solute charge."""
    nx, ny, nz = n
    no = site_off.shape[0]
    out = np.zeros((nx * ny * nz, no))
    if not np.any(np.abs(q) > EPS):
        return out
    rho = cic_density(n, length, xyz, q)
    if np.all(np.abs(rho) <= EPS):
        return out
    X = np.fft.rfftn(rho, axes=(0, 1, 2))                     # [nz][ny][nx/2+1], exp(-i k.r)
    m3, m2, m1 = np.meshgrid(_freqs(nz, False), _freqs(ny, False), _freqs(nx, True), indexing="ij")
    k2 = ((2 * np.pi * m1 / length[0]) ** 2 + (2 * np.pi * m2 / length[1]) ** 2 + (2 * np.pi * m3 / length[2]) ** 2)
    V = np.where(np.abs(k2) > EPS, X * 4.0 * np.pi / np.where(k2 == 0, 1.0, k2), 0.0)
    kz, ky, kx = np.meshgrid(_waves(nz, length[2], False), _waves(ny, length[1], False), _waves(nx, length[0], True),
                             indexing="ij")
    smooth = np.exp(-SMOOTHER_RADIUS ** 2 * (kx * kx + ky * ky + kz * kz) / 2.0)
    for io in range(no):
        sigma = np.zeros_like(V)
        for s in range(len(site_q)):
            if abs(site_q[s]) > EPS:
                o = site_off[io, s]
                sigma = sigma + site_q[s] * np.exp(-1j * (kx * o[0] + ky * o[1] + kz * o[2])) * smooth
        # irfftn includes the 1/N of the reference's division by product(gridnode)
        out[:, io] = qfact * np.fft.irfftn(V * np.conj(sigma), s=(nz, ny, nx), axes=(0, 1, 2)).reshape(-1)
    return out

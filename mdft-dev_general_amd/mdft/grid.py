"""Angular grid and size helpers (host side), src/module_grid.f90:95-139."""
from __future__ import annotations

import numpy as np


def gauss_legendre(n: int):
    """Gauss-Legendre nodes in cos(theta) and weights summing to 1
    (src/module_grid.f90:262-301, Newton iteration on the Legendre recursion)."""
    x = np.zeros(n)
    w = np.zeros(n)
    m = (n + 1) // 2
    for i in range(1, m + 1):
        xi = np.cos(np.pi * (i - 0.25) / (n + 0.5))
        dxi, pp = 1.0, 1.0
        for _ in range(200):
            if abs(dxi) <= np.finfo(float).eps:
                break
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3, p2 = p2, p1
                p1 = ((2 * j - 1) * xi * p2 - (j - 1) * p3) / j
            pp = n * (xi * p1 - p2) / (xi * xi - 1.0)
            dxi = -p1 / pp
            xi += dxi
        x[i - 1] = xi
        w[i - 1] = 1.0 / ((1.0 - xi * xi) * pp * pp)
        x[n - i] = -xi
        w[n - i] = w[i - 1]
    return x, w


def sizes(mmax: int, mrso: int = 2):
    ntheta, nphi, npsi = mmax + 1, 2 * mmax + 1, 2 * (mmax // mrso) + 1
    npr = sum((2 * m + 1) * (m // mrso + 1) for m in range(mmax + 1))
    return {"ntheta": ntheta, "nphi": nphi, "npsi": npsi, "no": ntheta * nphi * npsi, "np": npr}


def angular_grid(mmax: int, mrso: int = 2):
    s = sizes(mmax, mrso)
    x, wt = gauss_legendre(s["ntheta"])
    th = np.arccos(x)
    nphi, npsi = s["nphi"], s["npsi"]
    phi = np.arange(nphi) * 2 * np.pi / nphi
    psi = np.arange(npsi) * 2 * np.pi / mrso / npsi
    T, P, S = np.meshgrid(th, phi, psi, indexing="ij")
    W = np.meshgrid(wt, np.full(nphi, 1.0 / nphi), np.full(npsi, 1.0 / (npsi * mrso)), indexing="ij")
    w = (W[0] * W[1] * W[2] * 8 * np.pi ** 2).ravel()
    s.update(theta=T.ravel(), phi=P.ravel(), psi=S.ravel(), w=w, thetaofntheta=th, wthetaofntheta=wt)
    return s

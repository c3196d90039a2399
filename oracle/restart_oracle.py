"""CPU restatement (numpy) of the restart regridding -- test infrastructure only:
tests compare the device kernel (mdft_regrid_trilinear) with it.

Reference: src/module_density.f90:131-133 (the previous grid centred in the box)
and :171-226 (trilinear interpolation of xi from the previous grid's cells onto
the new grid's nodes).  Fields are [nz][ny][nx][no] (x fastest, orientation
innermost).

`regrid_loops` follows the reference's loops literally -- for io, for every
previous cell (z outer, x inner), for every new node in the cell's closed
index range ceiling(x0/d + 1) .. floor(x1/d + 1) -- so the last write wins;
pure Python, for small grids.  `regrid` computes the same per axis: the
winning cell of a node is the last (largest) cell whose range holds it; it is
what the device kernel is compared with at larger sizes.
"""
from __future__ import annotations

import math

import numpy as np


def offsets(length, n_prev, d_prev):
    return [(L - dp * npv) / 2 for L, dp, npv in zip(length, d_prev, n_prev)]


def regrid_loops(prev, n_prev, d_prev, n, length, init=1.0):
    """prev [nzp][nyp][nxp][no]; returns [nz][ny][nx][no] (src/module_density.f90:171-226)."""
    nxp, nyp, nzp = n_prev
    nx, ny, nz = n
    no = prev.shape[-1]
    dx, dy, dz = (L / m for L, m in zip(length, n))
    dxp, dyp, dzp = d_prev
    ox, oy, oz = offsets(length, n_prev, d_prev)
    xi = np.full((nz, ny, nx, no), init, dtype=np.float64)
    for io in range(no):
        for izp in range(1, nzp):
            z0, z1 = (izp - 1) * dzp + oz, izp * dzp + oz
            for iyp in range(1, nyp):
                y0, y1 = (iyp - 1) * dyp + oy, iyp * dyp + oy
                for ixp in range(1, nxp):
                    x0, x1 = (ixp - 1) * dxp + ox, ixp * dxp + ox
                    iz = max(1, math.ceil(z0 / dz + 1))
                    while iz <= min(nz, math.floor(z1 / dz + 1)):
                        iy = max(1, math.ceil(y0 / dy + 1))
                        while iy <= min(ny, math.floor(y1 / dy + 1)):
                            ix = max(1, math.ceil(x0 / dx + 1))
                            while ix <= min(nx, math.floor(x1 / dx + 1)):
                                z, y, x = (iz - 1) * dz, (iy - 1) * dy, (ix - 1) * dx
                                xd, yd, zd = (x - x0) / (x1 - x0), (y - y0) / (y1 - y0), (z - z0) / (z1 - z0)

                                def p(a, b, c):  # xi_prev(io, ixp + a, iyp + b, izp + c), 1-based
                                    return prev[izp - 1 + c, iyp - 1 + b, ixp - 1 + a, io]

                                xi00 = p(0, 0, 0) * (1 - xd) + p(1, 0, 0) * xd
                                xi01 = p(0, 0, 1) * (1 - xd) + p(1, 0, 1) * xd
                                xi10 = p(0, 1, 0) * (1 - xd) + p(1, 1, 0) * xd
                                xi11 = p(0, 1, 1) * (1 - xd) + p(1, 1, 1) * xd
                                xi0 = xi00 * (1 - yd) + xi10 * yd
                                xi1 = xi01 * (1 - yd) + xi11 * yd
                                xi[iz - 1, iy - 1, ix - 1, io] = xi0 * (1 - zd) + xi1 * zd
                                ix += 1
                            iy += 1
                        iz += 1
    return xi


def axis_cells(n, d, n_prev, d_prev, off):
    """0-based winning cell of every new index along one axis (-1: none) and the
    fractional position in it."""
    cell = np.full(n, -1, dtype=np.int64)
    w = np.zeros(n)
    for i in range(1, n + 1):
        x = (i - 1) * d
        for cp in range(n_prev - 1, 0, -1):
            x0, x1 = (cp - 1) * d_prev + off, cp * d_prev + off
            lo, hi = max(1, math.ceil(x0 / d + 1)), min(n, math.floor(x1 / d + 1))
            if lo <= i <= hi:
                cell[i - 1], w[i - 1] = cp - 1, (x - x0) / (x1 - x0)
                break
    return cell, w


def regrid(prev, n_prev, d_prev, n, length, outside=1.0):
    """Same result as regrid_loops, per axis and vectorised."""
    no = prev.shape[-1]
    d = [L / m for L, m in zip(length, n)]
    off = offsets(length, n_prev, d_prev)
    (cx, wx), (cy, wy), (cz, wz) = (axis_cells(n[a], d[a], n_prev[a], d_prev[a], off[a]) for a in range(3))
    out = np.full((n[2], n[1], n[0], no), outside, dtype=np.float64)
    Z, Y, X = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    ok = (cz[Z] >= 0) & (cy[Y] >= 0) & (cx[X] >= 0)
    zc, yc, xc = cz[Z][ok], cy[Y][ok], cx[X][ok]
    xd, yd, zd = wx[X][ok][:, None], wy[Y][ok][:, None], wz[Z][ok][:, None]

    def p(a, b, c):
        return prev[zc + c, yc + b, xc + a]

    xi00 = p(0, 0, 0) * (1 - xd) + p(1, 0, 0) * xd
    xi01 = p(0, 0, 1) * (1 - xd) + p(1, 0, 1) * xd
    xi10 = p(0, 1, 0) * (1 - xd) + p(1, 1, 0) * xd
    xi11 = p(0, 1, 1) * (1 - xd) + p(1, 1, 1) * xd
    xi0 = xi00 * (1 - yd) + xi10 * yd
    xi1 = xi01 * (1 - yd) + xi11 * yd
    out[ok] = xi0 * (1 - zd) + xi1 * zd
    return out

"""The restart file on the CPU: the regridding restatement (oracle/restart_oracle.py)
against the reference's own loop structure, closed forms of trilinear
interpolation, and the density.bin records against what a Fortran compiler
writes and reads (a small program compiled here with amdflang)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mdft import restart as R
from oracle import restart_oracle as RO

CASES = [
    # (n_prev, d_prev, n, length): a smaller grid centred in the box, the same grid, a
    # grid wider than the box, spacings whose nodes coincide (dp = 2 d) and ragged ones
    ((5, 4, 6), (0.9, 1.1, 0.7), (8, 7, 9), (8.0, 7.0, 6.3)),
    ((6, 5, 4), (1.0, 1.2, 1.5), (6, 5, 4), (6.0, 6.0, 6.0)),
    ((9, 7, 8), (1.0, 1.0, 1.0), (6, 6, 6), (6.0, 6.0, 6.0)),
    ((4, 5, 3), (2.0, 2.0, 2.0), (8, 10, 6), (8.0, 10.0, 6.0)),
    ((7, 3, 5), (0.61, 1.37, 0.93), (5, 9, 7), (4.4, 5.1, 6.2)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_regrid_matches_the_reference_loops(case):
    n_prev, d_prev, n, length = CASES[case]
    rng = np.random.default_rng(case)
    prev = rng.normal(size=(n_prev[2], n_prev[1], n_prev[0], 3))
    a = RO.regrid_loops(prev, n_prev, d_prev, n, length, init=1.0)
    b = RO.regrid(prev, n_prev, d_prev, n, length, outside=1.0)
    assert np.abs(a - b).max() <= 1e-15 * max(1.0, np.abs(a).max())


def test_regrid_same_grid_is_the_identity():
    """With spacings exact in binary the index ranges hold every node and the same grid
    maps onto itself exactly.  (With others, floor(x1/d + 1) can round below a node
    and the reference leaves it at its initial value; the comparison with the
    reference's loops above covers that.)"""
    n, length = (6, 5, 7), (3.0, 6.25, 7.0)
    d = tuple(L / m for L, m in zip(length, n))
    prev = np.random.default_rng(1).normal(size=(n[2], n[1], n[0], 4))
    out = RO.regrid(prev, n, d, n, length)
    assert np.array_equal(out, prev)


def test_regrid_reproduces_linear_fields_inside_and_keeps_one_outside():
    n_prev, d_prev, n, length = (5, 6, 4), (0.8, 0.7, 1.1), (9, 8, 10), (7.0, 6.5, 8.0)
    off = RO.offsets(length, n_prev, d_prev)
    coef = np.array([[0.3, -1.2, 0.7, 2.0], [1.0, 0.5, -0.25, 0.1]])   # per orientation: c0 + cx x + cy y + cz z
    zp, yp, xp = np.meshgrid(*(np.arange(n_prev[a]) * d_prev[a] + off[a] for a in (2, 1, 0)), indexing="ij")
    prev = np.stack([c[0] + c[1] * xp + c[2] * yp + c[3] * zp for c in coef], axis=-1)
    out = RO.regrid(prev, n_prev, d_prev, n, length)
    d = [L / m for L, m in zip(length, n)]
    z, y, x = np.meshgrid(*(np.arange(n[a]) * d[a] for a in (2, 1, 0)), indexing="ij")
    inside = ((x >= off[0] - 1e-12) & (x <= off[0] + (n_prev[0] - 1) * d_prev[0] + 1e-12) &
              (y >= off[1] - 1e-12) & (y <= off[1] + (n_prev[1] - 1) * d_prev[1] + 1e-12) &
              (z >= off[2] - 1e-12) & (z <= off[2] + (n_prev[2] - 1) * d_prev[2] + 1e-12))
    assert inside.any() and (~inside).any()
    for io, c in enumerate(coef):
        lin = c[0] + c[1] * x + c[2] * y + c[3] * z
        assert np.abs(out[..., io][inside] - lin[inside]).max() < 1e-12
        assert np.all(out[..., io][~inside] == 1.0)


def test_records_round_trip_in_python(tmp_path):
    rng = np.random.default_rng(3)
    sites = np.zeros(2, dtype=R.SITE_DTYPE)
    sites["name"] = [b"O".ljust(100), b"H1".ljust(100)]
    sites["r"] = rng.normal(size=(2, 3))
    sites["q"], sites["sig"], sites["eps"], sites["z"] = [-0.8, 0.4], [3.1, 0.0], [0.65, 0.0], [8, 1]
    proj = rng.normal(size=(24, 4)) + 1j * rng.normal(size=(24, 4))
    p = tmp_path / "density.bin"
    R.write_file(str(p), 1, 6, 4, (2, 3, 4), (0.5, 0.6, 0.7), sites, proj, full=False)
    h, data = R.read_file(str(p))
    assert (h["mmax"], h["no"], h["np"], h["n"], h["d"], h["full"]) == (1, 6, 4, (2, 3, 4), (0.5, 0.6, 0.7), False)
    assert np.array_equal(h["sites"], sites) and np.array_equal(data, proj)
    xi = rng.uniform(0, 2, 24 * 6)
    R.write_file(str(p), 1, 6, 4, (2, 3, 4), (0.5, 0.6, 0.7), sites, xi, full=True)
    h, data = R.read_file(str(p))
    assert h["full"] and np.array_equal(data, xi)
    # a truncated file is an error, as the reference's reads fail on it
    raw = p.read_bytes()
    p.write_bytes(raw[:-20])
    with pytest.raises(ValueError):
        R.read_file(str(p))


FORTRAN = r"""
program density_bin_io
    implicit none
    integer, parameter :: dp = kind(1.0d0)
    type :: site_type
        character(100) :: name
        integer :: type, n_sites
        real(dp) :: r(3)
        real(dp) :: q, sig, eps, lambda1, lambda2
        integer :: z
    end type site_type
    type(site_type) :: s(2), t
    complex(dp) :: p(4)
    integer :: ix, iy, iz, k, ns, mmax, no, np, n(3), nsite
    real(dp) :: d(3), acc
    character(8) :: mode
    call get_command_argument(1, mode)
    if (mode == 'write') then
        s(1) = site_type('O', 1, 3, [1.5_dp, -2.25_dp, 0.125_dp], -0.8476_dp, 3.166_dp, 0.65_dp, 0.0_dp, 0.0_dp, 8)
        s(2) = site_type('H1', 2, 3, [0.5_dp, 0.75_dp, -1.0_dp], 0.4238_dp, 0.0_dp, 0.0_dp, 1.0_dp, 2.0_dp, 1)
        open (10, file='f.bin', form='unformatted')
        write (10) 1
        write (10) 1
        write (10) 6
        write (10) 4
        write (10) 2, 3, 4
        write (10) 0.5_dp, 0.6_dp, 0.7_dp
        write (10) 2
        do k = 1, 2
            write (10) s(k)
        end do
        do iz = 1, 4
            do iy = 1, 3
                do ix = 1, 2
                    do k = 1, 4
                        p(k) = cmplx(ix + 10 * iy + 100 * iz, k, dp)
                    end do
                    write (10) p
                end do
            end do
        end do
        close (10)
    else
        open (10, file='py.bin', form='unformatted', status='old')
        read (10) ns
        read (10) mmax
        read (10) no
        read (10) np
        read (10) n
        read (10) d
        read (10) nsite
        acc = 0
        do k = 1, nsite
            read (10) t
            acc = acc + t%r(1) + 2 * t%r(2) + 3 * t%r(3) + t%q + t%sig + t%eps + t%z + t%type
        end do
        do iz = 1, n(3)
            do iy = 1, n(2)
                do ix = 1, n(1)
                    read (10) p
                    acc = acc + sum(real(p)) + 2 * sum(aimag(p))
                end do
            end do
        end do
        close (10)
        print '(6I6, 3F12.6, I4, ES24.16)', ns, mmax, no, np, n(1), n(2), d, nsite, acc
    end if
end program density_bin_io
"""


@pytest.mark.skipif(shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"),
                    reason="no Fortran compiler")
def test_records_match_a_fortran_compiler(tmp_path):
    """Records written by a Fortran program with the reference's write sequence are
    read back by the Python reader, and a file written in Python is read by the
    Fortran program, components of site_type included."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    (tmp_path / "io.f90").write_text(FORTRAN)
    subprocess.check_call([fc, "-O0", "-o", str(tmp_path / "io"), str(tmp_path / "io.f90")], cwd=tmp_path)
    subprocess.check_call([str(tmp_path / "io"), "write"], cwd=tmp_path)
    h, data = R.read_file(str(tmp_path / "f.bin"))
    assert (h["mmax"], h["no"], h["np"], h["n"], h["d"], h["full"]) == (1, 6, 4, (2, 3, 4), (0.5, 0.6, 0.7), False)
    s = h["sites"]
    assert s["name"][0].rstrip() == b"O" and s["name"][1].rstrip() == b"H1"
    assert list(s["type"]) == [1, 2] and list(s["n_sites"]) == [3, 3] and list(s["z"]) == [8, 1]
    assert np.array_equal(s["r"][0], [1.5, -2.25, 0.125]) and s["lambda2"][1] == 2.0 and s["q"][0] == -0.8476
    idx = np.arange(24)
    ix, iy, iz = idx % 2 + 1, (idx // 2) % 3 + 1, idx // 6 + 1
    expect = (ix + 10 * iy + 100 * iz)[:, None] + 1j * np.arange(1, 5)[None, :]
    assert np.array_equal(data, expect)
    # the other way round
    rng = np.random.default_rng(5)
    sites = np.zeros(3, dtype=R.SITE_DTYPE)
    sites["name"] = b" " * 100
    sites["r"] = rng.normal(size=(3, 3))
    sites["q"], sites["sig"], sites["eps"] = rng.normal(size=3), [3.0, 2.5, 1.0], [0.5, 0.2, 0.1]
    sites["type"], sites["z"] = [1, 2, 3], [6, 7, 8]
    proj = rng.normal(size=(30, 4)) + 1j * rng.normal(size=(30, 4))
    R.write_file(str(tmp_path / "py.bin"), 1, 6, 4, (3, 2, 5), (0.25, 0.5, 0.75), sites, proj, full=False)
    out = subprocess.check_output([str(tmp_path / "io"), "read"], cwd=tmp_path, text=True).split()
    assert [int(v) for v in out[:6]] == [1, 1, 6, 4, 3, 2]
    assert [float(v) for v in out[6:9]] == [0.25, 0.5, 0.75] and int(out[9]) == 3
    acc = sum(st["r"][0] + 2 * st["r"][1] + 3 * st["r"][2] + st["q"] + st["sig"] + st["eps"] + st["z"] + st["type"]
              for st in sites) + proj.real.sum() + 2 * proj.imag.sum()
    assert abs(float(out[10]) - acc) <= 1e-12 * max(1.0, abs(acc))

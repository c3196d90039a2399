"""ctypes front end of the FP64 CPU oracle (oracle/mdft_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  It restates the reference path energy_and_gradient(f, df)
(src/module_energy_and_gradient.f90:36-240) in plain C.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MDFT_ORACLE_LIB selects another build of the same restatement (the address-sanitised one)
LIB = os.environ.get("MDFT_ORACLE_LIB", os.path.join(HERE, "liboracle.so"))
REFLIB = os.path.join(HERE, "_ref", "libmdft_refpieces.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, with ref=True and the reference present, _ref/)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _load():
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.or_create.restype = C.c_void_p
    lib.or_create.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int, C.c_int, C.c_double, C.c_double]
    lib.or_destroy.argtypes = [C.c_void_p]
    lib.or_sizes.argtypes = [C.c_void_p, _ip]
    lib.or_tables.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.or_proj_tables.argtypes = [C.c_void_p, _ip, _ip, _ip]
    lib.or_k_tables.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.or_gauss_legendre.argtypes = [C.c_int, _dp, _dp]
    lib.or_wigner_small_d.restype = C.c_double
    lib.or_wigner_small_d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
    lib.or_rotation_matrix.argtypes = [C.c_int, _dp, _dp]
    lib.or_set_ck.restype = C.c_int
    lib.or_set_ck.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip, C.c_int, C.c_double, _dp]
    lib.or_get_ck.argtypes = [C.c_void_p, _dp]
    lib.or_angl2proj.argtypes = [C.c_void_p, _dp, _dp]
    lib.or_proj2angl.argtypes = [C.c_void_p, _dp, _dp]
    lib.or_fft3d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.or_excess_stage.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
    lib.or_excess.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.or_ideal_ext.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    lib.or_energy_and_gradient.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.or_num_threads.restype = C.c_int
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def num_threads() -> int:
    return lib().or_num_threads()


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().or_gauss_legendre(n, _p(x), _p(w))
    return x, w


def wigner_small_d(m, mu, mup, theta):
    return lib().or_wigner_small_d(m, mu, mup, float(theta))


def rotation_matrix(mmax, q):
    """R[m, mup+mmax, khi+mmax] (complex), module_rotation.f90:119-247, no chop."""
    W = 2 * mmax + 1
    out = np.zeros((mmax + 1) * W * W, dtype=np.complex128)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    lib().or_rotation_matrix(mmax, _p(qq), _p(out.view(np.float64)))
    return out.reshape(mmax + 1, W, W)


def fft3d(a, sign):
    """a[nz, ny, nx] complex (x fastest); returns the unnormalised sign-`sign` DFT."""
    b = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nz, ny, nx = b.shape
    lib().or_fft3d(_p(b.view(np.float64)), nx, ny, nz, int(sign))
    return b


class Oracle:
    """FP64 CPU oracle context for one grid / angular order / solvent."""

    def __init__(self, n, length, mmax, mrso, kT, n0):
        nx, ny, nz = n
        lx, ly, lz = length
        self.n = (nx, ny, nz)
        self.length = (float(lx), float(ly), float(lz))
        self.mmax, self.mrso = mmax, mrso
        self.kT, self.n0 = kT, n0
        h = lib().or_create(nx, ny, nz, lx, ly, lz, mmax, mrso, kT, n0)
        if not h:
            raise ValueError("oracle: invalid grid description")
        self._h = C.c_void_p(h)
        self._refresh_sizes()
        self.theta = np.zeros(self.ntheta)
        self.wtheta = np.zeros(self.ntheta)
        self.w = np.zeros(self.no)
        lib().or_tables(self._h, _p(self.theta), _p(self.wtheta), _p(self.w))
        m = np.zeros(self.np_, np.int32)
        mup = np.zeros(self.np_, np.int32)
        mu = np.zeros(self.np_, np.int32)
        lib().or_proj_tables(self._h, _pi(m), _pi(mup), _pi(mu))
        self.pm, self.pmup, self.pmu = m, mup, mu
        self.kx, self.ky, self.kz = np.zeros(nx), np.zeros(ny), np.zeros(nz)
        lib().or_k_tables(self._h, _p(self.kx), _p(self.ky), _p(self.kz))
        self.dv = (lx / nx) * (ly / ny) * (lz / nz)
        self.rho0 = n0 / (8.0 * np.arccos(-1.0) ** 2 / mrso)

    def _refresh_sizes(self):
        s = np.zeros(7, np.int32)
        lib().or_sizes(self._h, _pi(s))
        self.ntheta, self.nphi, self.npsi, self.no, self.np_, self.npc, self.nq = (int(v) for v in s)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.or_destroy(h)
            self._h = None

    @property
    def N(self):
        return self.n[0] * self.n[1] * self.n[2]

    def set_ck(self, table):
        """table: a CkTable-like object with fields m, n, mu, nu, khi, dq, c[nq_file, np_file]."""
        arr = [np.ascontiguousarray(getattr(table, k), dtype=np.int32) for k in ("m", "n", "mu", "nu", "khi")]
        c = np.ascontiguousarray(table.c, dtype=np.complex128)
        rc = lib().or_set_ck(self._h, len(arr[0]), *[_pi(a) for a in arr], c.shape[0], float(table.dq),
                             _p(c.view(np.float64)))
        if rc != 0:
            raise ValueError(f"oracle: or_set_ck failed ({rc})")
        self._refresh_sizes()
        self.dq = float(table.dq)

    def get_ck(self):
        out = np.zeros(self.nq * self.npc, dtype=np.complex128)
        lib().or_get_ck(self._h, _p(out.view(np.float64)))
        return out.reshape(self.nq, self.npc)

    def angl2proj(self, o):
        o = np.ascontiguousarray(o, dtype=np.float64)
        p = np.zeros(self.np_, dtype=np.complex128)
        lib().or_angl2proj(self._h, _p(o), _p(p.view(np.float64)))
        return p

    def proj2angl(self, p):
        p = np.ascontiguousarray(p, dtype=np.complex128)
        o = np.zeros(self.no)
        lib().or_proj2angl(self._h, _p(p.view(np.float64)), _p(o))
        return o

    def excess_stage(self, xi, stage):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        out = np.zeros(self.N * self.np_, dtype=np.complex128)
        lib().or_excess_stage(self._h, _p(xi), int(stage), _p(out.view(np.float64)))
        return out.reshape(self.N, self.np_)

    def excess(self, xi, want_df=True):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        f = np.zeros(1)
        df = np.zeros(xi.size) if want_df else None
        lib().or_excess(self._h, _p(xi), _p(f), _p(df) if want_df else None)
        return f[0], df

    def ideal_ext(self, xi, vext, want_df=True):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        vext = np.ascontiguousarray(vext, dtype=np.float64)
        a, b = np.zeros(1), np.zeros(1)
        df = np.zeros(xi.size) if want_df else None
        lib().or_ideal_ext(self._h, _p(xi), _p(vext), _p(a), _p(b), _p(df) if want_df else None)
        return a[0], b[0], df

    def energy_and_gradient(self, xi, vext, want_df=True):
        """Returns (parts[F_id, F_ext, F_exc], df or None); df = (1/dv) dF/dxi."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        vext = np.ascontiguousarray(vext, dtype=np.float64)
        parts = np.zeros(3)
        df = np.zeros(xi.size) if want_df else None
        lib().or_energy_and_gradient(self._h, _p(xi), _p(vext), _p(parts), _p(df) if want_df else None)
        return parts, df


class RefPieces:
    """The reference's own Wigner-d and rotation-matrix functions, compiled from
    its sources into oracle/_ref (only where the reference tree is present)."""

    def __init__(self):
        if not os.path.exists(REFLIB):
            raise FileNotFoundError(REFLIB)
        self.lib = C.CDLL(REFLIB)
        self.lib.ref_wigner_small_d.restype = C.c_double
        self.lib.ref_wigner_small_d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        self.lib.ref_rotation_matrix.argtypes = [C.c_int, _dp, _dp]

    def wigner_small_d(self, m, mu, mup, theta):
        return self.lib.ref_wigner_small_d(m, mu, mup, float(theta))

    def rotation_matrix(self, mmax, q):
        W = 2 * mmax + 1
        out = np.zeros((mmax + 1) * W * W, dtype=np.complex128)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        self.lib.ref_rotation_matrix(mmax, _p(qq), _p(out.view(np.float64)))
        return out.reshape(mmax + 1, W, W)

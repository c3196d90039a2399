"""The Fortran host L-BFGS (fortran/mdft_lbfgs.f90) run on the CPU: it is linked
against a mock of the C-ABI backed by the FP64 oracle (tests/mock_abi), so the
minimiser's host logic is checked here without a GPU, against scipy's L-BFGS-B
on the same functional."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from mdft import ck, workloads
from mdft.gpu import Desc

HERE = os.path.dirname(os.path.abspath(__file__))
MOCK = os.path.join(HERE, "mock_abi")


@pytest.fixture(scope="module")
def host(oracle_mod):
    subprocess.check_call(["make", "-s", "-C", MOCK])
    lib = C.CDLL(os.path.join(MOCK, "_build", "libmdft_host_mock.so"))
    dp = C.POINTER(C.c_double)
    lib.mdft_minimize.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, C.c_int]
    lib.mdft_minimize.restype = C.c_int
    lib.mdft_minimize_ex.argtypes = lib.mdft_minimize.argtypes + [C.c_int, C.c_double]
    lib.mdft_minimize_ex.restype = C.c_int
    mock = C.CDLL(os.path.join(MOCK, "_build", "libmdft_abi_mock.so"))
    mock.mdft_ctx_create.argtypes = [C.POINTER(Desc), C.POINTER(C.c_void_p)]
    ip = C.POINTER(C.c_int)
    mock.mdft_set_ck.argtypes = [C.c_void_p, C.c_int, ip, ip, ip, ip, ip, C.c_int, C.c_double, dp]
    mock.mdft_set_vext.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    mock.mdft_ctx_destroy.argtypes = [C.c_void_p]
    return lib, mock


def _problem(mock, w, sol, t):
    d = Desc(*w.n, *w.length, w.mmax, 2, w.kT, w.n0, 0, 0)
    h = C.c_void_p()
    assert mock.mdft_ctx_create(C.byref(d), C.byref(h)) == 0
    arr = [np.ascontiguousarray(getattr(t, k), dtype=np.int32) for k in ("m", "n", "mu", "nu", "khi")]
    c = np.ascontiguousarray(t.c)
    ip = C.POINTER(C.c_int)
    assert mock.mdft_set_ck(h, len(arr[0]), *[a.ctypes.data_as(ip) for a in arr], c.shape[0], t.dq,
                            c.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))) == 0
    vext = np.ascontiguousarray(workloads.vext_numpy(w, sol).ravel())
    assert mock.mdft_set_vext(h, vext.ctypes.data, 0) == 0
    return h, vext


def _minimize(lib, h, x, m=5, gtol=1e-6, fnoise=1e-7, maxeval=600, ftarget=0.0, fband=0.0, vec32=None, step0=0.0):
    g = np.zeros_like(x)
    stats = np.zeros(16)
    trace = np.zeros((maxeval, 3))
    dp = C.POINTER(C.c_double)
    args = (h, x.size, x.ctypes.data, g.ctypes.data, m, gtol, 0.0, fnoise, ftarget, fband, maxeval,
            stats.ctypes.data_as(dp), trace.ctypes.data_as(dp), maxeval)
    rc = lib.mdft_minimize(*args) if vec32 is None else lib.mdft_minimize_ex(*args, int(vec32), float(step0))
    assert rc == 0
    return stats, trace[:int(stats[6])], g


@pytest.mark.parametrize("charge", [0.0, 0.5])
def test_fortran_lbfgs_matches_scipy(host, oracle_mod, charge):
    from scipy.optimize import minimize as spmin
    lib, mock = host
    w = workloads.Workload("probe", 0, (11, 11, 11), (11.0, 11.0, 11.0), 2, 256)
    sol = workloads.Solute(np.array([[5.5, 5.5, 5.5]]), np.array([charge]), np.array([3.0]), np.array([0.8]))
    t = ck.synthetic_ck(2, 256, 0.05, seed=21, n0=w.n0)
    h, vext = _problem(mock, w, sol, t)
    x0 = workloads.guess_xi(vext, w.kT)
    x = x0.copy()
    stats, trace, g = _minimize(lib, h, x)
    F, ratio, nev, status = stats[0], stats[4], int(stats[5]), int(stats[7])
    assert status == 0 and ratio < 1e-6, (status, ratio, nev)
    assert np.all(np.diff(trace[:, 0]) <= 1e-7 * abs(F))
    o = oracle_mod.Oracle(w.n, w.length, w.mmax, 2, w.kT, w.n0)
    o.set_ck(t)

    def fg(xv):
        parts, df = o.energy_and_gradient(xv, vext)
        return parts.sum(), df * o.dv

    r = spmin(fg, x0, jac=True, method="L-BFGS-B", options={"maxiter": 2000, "gtol": 1e-12, "ftol": 1e-15})
    assert abs(F - r.fun) < 1e-6 * max(1.0, abs(r.fun))
    # F depends on xi only through rho = rho0 xi^2: compare |xi|
    assert np.abs(np.abs(x) - np.abs(r.x)).max() < 1e-3 * np.abs(x).max()
    mock.mdft_ctx_destroy(h)


def test_fortran_lbfgs_memory_sizes_agree(host):
    """m = 1 (the reference's memory) and m = 8 reach the same minimum."""
    lib, mock = host
    w = workloads.Workload("probe", 0, (9, 9, 9), (10.0, 10.0, 10.0), 1, 256)
    sol = workloads.Solute(np.array([[5.0, 5.0, 5.0]]), np.array([-0.4]), np.array([3.2]), np.array([0.6]))
    t = ck.synthetic_ck(1, 256, 0.05, seed=5, n0=w.n0)
    h, vext = _problem(mock, w, sol, t)
    x1 = workloads.guess_xi(vext, w.kT)
    x8 = x1.copy()
    s1, _, _ = _minimize(lib, h, x1, m=1, maxeval=1500)
    s8, _, _ = _minimize(lib, h, x8, m=8)
    assert s1[7] == 0 and s8[7] == 0
    assert abs(s1[0] - s8[0]) < 1e-7 * abs(s8[0])
    mock.mdft_ctx_destroy(h)


def test_fortran_lbfgs_target_band(host):
    """With a target band the run stops at the first iterate within it."""
    lib, mock = host
    w = workloads.Workload("probe", 0, (9, 9, 9), (10.0, 10.0, 10.0), 1, 256)
    sol = workloads.Solute(np.array([[5.0, 5.0, 5.0]]), np.array([0.3]), np.array([3.2]), np.array([0.6]))
    t = ck.synthetic_ck(1, 256, 0.05, seed=6, n0=w.n0)
    h, vext = _problem(mock, w, sol, t)
    x = workloads.guess_xi(vext, w.kT)
    full, tr_full, _ = _minimize(lib, h, x.copy())
    x2 = workloads.guess_xi(vext, w.kT)
    part, tr, _ = _minimize(lib, h, x2, ftarget=full[0], fband=0.05)
    assert part[7] == 4 and abs(part[0] - full[0]) <= 0.05
    first = int(np.nonzero(np.abs(tr_full[:, 0] - full[0]) <= 0.05)[0][0]) + 1
    assert int(part[6]) == first and tr[-1, 2] == part[5]
    mock.mdft_ctx_destroy(h)


def test_fortran_lbfgs_fp32_vectors(host):
    """With the directions and gradients stored in FP32 (mdft_minimize_ex, vec32 = 1)
    the minimiser reaches the FP64-vector minimum to FP32 accuracy, returns its
    gradient widened to FP64, and mdft_minimize_ex(vec32 = 0) agrees with mdft_minimize."""
    lib, mock = host
    w = workloads.Workload("probe", 0, (9, 9, 9), (10.0, 10.0, 10.0), 1, 256)
    sol = workloads.Solute(np.array([[5.0, 5.0, 5.0]]), np.array([-0.4]), np.array([3.2]), np.array([0.6]))
    t = ck.synthetic_ck(1, 256, 0.05, seed=5, n0=w.n0)
    h, vext = _problem(mock, w, sol, t)
    x0 = workloads.guess_xi(vext, w.kT)
    xa, xb, xc = x0.copy(), x0.copy(), x0.copy()
    sa, _, _ = _minimize(lib, h, xa, m=1, maxeval=1500)
    sb, trb, gb = _minimize(lib, h, xb, m=1, maxeval=1500, vec32=0)
    # (the oracle's threaded sums are not bit-reproducible from run to run)
    # to the loose gtol = 1e-6 of both runs
    assert abs(sa[0] - sb[0]) < 1e-6 * abs(sa[0]) and np.abs(xa - xb).max() < 1e-4
    sc, trc, gc = _minimize(lib, h, xc, m=1, gtol=1e-5, maxeval=1500, vec32=1)
    assert sc[7] in (0, 1)
    assert abs(sc[0] - sa[0]) < 1e-6 * abs(sa[0])
    assert np.abs(xc - xa).max() < 1e-3 * np.abs(xa).max()
    # the returned gradient is the final iterate's, widened from FP32
    parts = np.zeros(3)
    gref = np.zeros_like(xc)
    dp = C.POINTER(C.c_double)
    mock.mdft_energy_and_gradient.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.c_int]
    assert mock.mdft_energy_and_gradient(h, xc.ctypes.data, parts.ctypes.data_as(dp), gref.ctypes.data, 0) == 0
    assert np.array_equal(gc, gref.astype(np.float32).astype(np.float64))
    mock.mdft_ctx_destroy(h)


def test_fortran_lbfgs_scaled_variables(host):
    """Minimising over u = s xi, s(io) = sqrt(w(io) / mean w) (opts bit 1), alone and with
    FP32 vectors: the same minimum as over xi, xi and g handed back in the caller's
    variables (g = dF/dxi at the returned xi), and no more evaluations to a 0.05 kJ/mol
    band than over xi (mmax = 3, whose weights span a factor of 2.6)."""
    lib, mock = host
    w = workloads.Workload("probe", 0, (9, 9, 9), (10.0, 10.0, 10.0), 3, 256)
    sol = workloads.Solute(np.array([[5.0, 5.0, 5.0]]), np.array([-0.4]), np.array([3.2]), np.array([0.6]))
    t = ck.synthetic_ck(3, 256, 0.05, seed=5, n0=w.n0)
    h, vext = _problem(mock, w, sol, t)
    x0 = workloads.guess_xi(vext, w.kT)
    xa = x0.copy()
    sa, _, _ = _minimize(lib, h, xa, m=1, gtol=1e-8, maxeval=3000)
    dp = C.POINTER(C.c_double)
    mock.mdft_energy_and_gradient.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.c_int]
    evals = {}
    for opts in (0, 2, 3):
        x = x0.copy()
        s, tr, g = _minimize(lib, h, x, m=1, gtol=0.0, maxeval=3000, ftarget=sa[0], fband=0.05, vec32=opts)
        assert s[7] == 4 and abs(s[0] - sa[0]) <= 0.05
        evals[opts] = int(s[5])
        parts, gref = np.zeros(3), np.zeros_like(x)
        assert mock.mdft_energy_and_gradient(h, x.ctypes.data, parts.ctypes.data_as(dp), gref.ctypes.data, 0) == 0
        assert abs(parts.sum() - s[0]) <= 1e-9 * abs(s[0])                 # x is xi, not u
        tol = 1e-6 if opts & 1 else 1e-12
        assert np.abs(g - gref).max() <= tol * np.abs(gref).max()          # g is dF/dxi
    assert evals[2] <= evals[0]
    xs = x0.copy()
    ss, _, _ = _minimize(lib, h, xs, m=1, gtol=1e-8, maxeval=3000, vec32=2)
    assert abs(ss[0] - sa[0]) < 1e-7 * abs(sa[0])
    # a given first step (here far too long: the line search must bring it back) reaches
    # the same minimum
    xt = x0.copy()
    st, _, _ = _minimize(lib, h, xt, m=1, gtol=1e-8, maxeval=3000, vec32=2, step0=1e4)
    assert abs(st[0] - sa[0]) < 1e-7 * abs(sa[0])
    mock.mdft_ctx_destroy(h)


def test_fortran_lbfgs_fp32_iterate_and_reference_stop(host):
    """The iterate stored in FP32 (opts bit 3, with FP32 vectors and scaled variables,
    evaluations reading xi as float): the same minimum as the FP64 run to the FP32
    path's resolution, xi handed back in FP64.  The reference's absolute stop
    (fold - f) <= tol (opts bit 2, src/module_lbfgs_nocedal_mdft.f90:425-427) ends
    the run with status 1 within 0.05 kJ/mol of the minimum at tol = 1e-3."""
    lib, mock = host
    w = workloads.Workload("probe", 0, (9, 9, 9), (10.0, 10.0, 10.0), 3, 256)
    sol = workloads.Solute(np.array([[5.0, 5.0, 5.0]]), np.array([-0.4]), np.array([3.2]), np.array([0.6]))
    t = ck.synthetic_ck(3, 256, 0.05, seed=5, n0=w.n0)
    h, vext = _problem(mock, w, sol, t)
    x0 = workloads.guess_xi(vext, w.kT)
    xa = x0.copy()
    sa, _, _ = _minimize(lib, h, xa, m=1, gtol=1e-8, maxeval=3000)
    x = x0.copy()
    s, _, g = _minimize(lib, h, x, m=1, gtol=1e-7, maxeval=3000, vec32=1 | 2 | 8)
    assert abs(s[0] - sa[0]) <= 1e-6 * abs(sa[0])
    dp = C.POINTER(C.c_double)
    mock.mdft_energy_and_gradient.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.c_int]
    parts, gref = np.zeros(3), np.zeros_like(x)
    assert mock.mdft_energy_and_gradient(h, x.ctypes.data, parts.ctypes.data_as(dp), gref.ctypes.data, 0) == 0
    assert abs(parts.sum() - s[0]) <= 1e-9 * abs(s[0])                   # xi handed back is the final iterate
    # the reference's stop rule
    gtr = np.zeros_like(x0)
    stats = np.zeros(16)
    trace = np.zeros((3000, 3))
    xb = x0.copy()
    rc = lib.mdft_minimize_ex(h, xb.size, xb.ctypes.data, gtr.ctypes.data, 1, 0.0, 1e-3, 1e-7, 0.0, 0.0, 3000,
                              stats.ctypes.data_as(dp), trace.ctypes.data_as(dp), 3000, 4, 0.0)
    assert rc == 0 and int(stats[7]) == 1
    assert abs(stats[0] - sa[0]) <= 0.05
    it = int(stats[6])
    assert trace[it - 2, 0] - trace[it - 1, 0] <= 1e-3 < trace[it - 3, 0] - trace[it - 2, 0]
    mock.mdft_ctx_destroy(h)

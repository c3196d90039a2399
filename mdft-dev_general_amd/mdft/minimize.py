"""Device-resident L-BFGS minimisation through the Fortran host library.

`minimize` calls mdft_minimize (fortran/mdft_lbfgs.f90, bound with
ISO_C_BINDING) on an Engine's context: xi and the gradient stay on the device,
the two-loop recursion and line search scalars are FP64 on the host.  This
replaces the reference's minimisation loop (src/energy_minimization.f90:232-290).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .gpu import PKG, DeviceBuffer, Engine, MdftError

HOSTLIB = os.path.join(PKG, "lib", "libmdft_host.so")
_HOST = None

STATUS = {0: "converged (gradient)", 1: "converged (F change)", 2: "evaluation budget", 3: "line search failed",
          4: "F within the target band"}


def host_library():
    global _HOST
    if _HOST is None:
        if not os.path.exists(HOSTLIB):
            raise FileNotFoundError(f"{HOSTLIB} is not built (run __graft_entry__.build())")
        lib = C.CDLL(HOSTLIB)
        dp = C.POINTER(C.c_double)
        lib.mdft_minimize.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, C.c_int]
        lib.mdft_minimize.restype = C.c_int
        lib.mdft_minimize_ex.argtypes = lib.mdft_minimize.argtypes + [C.c_int, C.c_double]
        lib.mdft_minimize_ex.restype = C.c_int
        _HOST = lib
    return _HOST


@dataclass
class MinResult:
    F: float
    parts: np.ndarray
    grad_ratio: float
    evaluations: int
    iterations: int
    status: int
    g0: float
    trace: np.ndarray

    @property
    def converged(self) -> bool:
        return self.status in (0, 1, 4)


def minimize(eng: Engine, xi: DeviceBuffer, g: DeviceBuffer | None = None, m: int = 5, gtol: float = 1e-6,
             ftol: float = 0.0, fnoise: float = 1e-7, maxeval: int = 2000, ntrace: int = 2000,
             ftarget: float = 0.0, fband: float = 0.0, fp32_vectors: bool = False, scaled: bool = False,
             step0: float | str = 0.0, compressibility_s0: float | None = None, ftol_absolute: bool = False,
             fp32_iterate: bool = False, lib=None) -> MinResult:
    """Minimise F over xi (in place).  Returns the final F and statistics.
    fnoise is the relative noise level of F used to relax the sufficient-decrease
    test; with fband > 0 the run stops at the first iterate with |F - ftarget| <= fband.
    fp32_vectors stores the minimiser's directions and gradients in FP32 (half the
    bytes per vector pass; meant for FP32 evaluations); g still receives FP64.
    fp32_iterate (with fp32_vectors) stores the iterate and the trial points in FP32 as
    well, read by the evaluations as such (MDFT_XI_FP32); xi receives the result in FP64.
    scaled minimises over u = s xi with s(io) = sqrt(w(io) / mean w), which takes the
    quadrature weight out of the curvature (fewer iterations at the same minimum);
    xi and g are returned in the caller's variables.
    step0 is the first trial step along -g (0: a unit-length move, the reference's
    choice); "newton" (scaled variables only) takes S0 / (4 rho0 kT mean(w)):
    1 / (4 rho0 kT mean(w)) is the inverse curvature of the ideal term at the bulk in the
    scaled variables (in the gradient's units, (1/dv) dF/du; over xi itself the
    curvature is 4 rho0 kT w(io), one value per orientation, which a scalar step cannot
    take), and S0 = 1 / (1 - c_000(q = 0)) the structure factor at k = 0 of the
    context's own c(k) table (mdft.ck.structure_factor_s0; compressibility_s0 overrides
    it), by which the excess term divides the curvature of long-wavelength changes.
    ftol_absolute: stop when one iteration lowers F by at most ftol (kJ/mol), the
    reference's rule (fold - f) <= tol with tol = 1e-3 by default
    (src/module_lbfgs_nocedal_mdft.f90:10,188,425-427); otherwise ftol is relative."""
    lib = lib or host_library()
    if step0 == "newton":
        if not scaled:
            raise ValueError("step0='newton' needs scaled=True: over xi the ideal curvature 4 rho0 kT w(io) "
                             "differs per orientation")
        if compressibility_s0 is None:
            table = getattr(eng, "ck_table", None)
            if table is None:
                raise ValueError("step0='newton' needs the context's c(k) table (Engine.set_ck) or compressibility_s0")
            from .ck import structure_factor_s0
            compressibility_s0 = structure_factor_s0(table)
        rho0 = eng.n0 * eng.mrso / (8.0 * np.pi ** 2)
        wmean = 8.0 * np.pi ** 2 / eng.mrso / eng.no
        step0 = compressibility_s0 / (4.0 * rho0 * eng.kT * wmean)
    own = g is None
    if own:
        g = eng.alloc_field()
    stats = np.zeros(16)
    trace = np.zeros((max(ntrace, 1), 3))
    rc = lib.mdft_minimize_ex(eng._h, eng.N * eng.no, xi.ptr, g.ptr, int(m), float(gtol), float(ftol),
                              float(fnoise), float(ftarget), float(fband), int(maxeval),
                              stats.ctypes.data_as(C.POINTER(C.c_double)),
                              trace.ctypes.data_as(C.POINTER(C.c_double)), int(ntrace),
                              int(bool(fp32_vectors)) | (2 if scaled else 0) | (4 if ftol_absolute else 0)
                              | (8 if (fp32_iterate and fp32_vectors) else 0),
                              float(step0))
    if rc != 0:
        raise MdftError(rc, eng.lib.mdft_last_error(eng._h).decode())
    it = int(stats[6])
    res = MinResult(float(stats[0]), stats[1:4].copy(), float(stats[4]), int(stats[5]), it, int(stats[7]),
                    float(stats[8]), trace[:min(it, ntrace)].copy())
    if own:
        g.free()
    return res

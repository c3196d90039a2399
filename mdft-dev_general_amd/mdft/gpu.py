"""ctypes binding of libmdft_gpu.so (C-ABI declared in include/mdft_gpu.h).

`Engine` mirrors the reference's evaluation interface: energy_and_gradient(f,
df) (src/module_energy_and_gradient.f90:36), the excess_energy plug-in
(src/module_proc_ptrs.f90:3-21) and energy_ideal_and_external
(src/module_energy_ideal_and_external.f90:14).  Non-zero status codes raise
`MdftError` carrying mdft_last_error(), where the reference would stop with
`error stop`.  There is no CPU fallback: without the built library or a GPU the
calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(PKG, "lib", "libmdft_gpu.so")

PREC = {"fp64": 0, "fp32": 1, "bf16": 2}
DEVICE_PTRS = 1
TIMERS = 2
DF_FP32 = 4
SCALED_VARS = 8
XI_FP32 = 16

MDFT_EXPORTS = (
    "mdft_ctx_create", "mdft_ctx_destroy", "mdft_last_error", "mdft_get_info", "mdft_get_quadrature",
    "mdft_set_vext", "mdft_set_ck", "mdft_energy_and_gradient", "mdft_excess_energy",
    "mdft_ideal_external", "mdft_excess_stage", "mdft_phase_times", "mdft_device_alloc",
    "mdft_device_free", "mdft_memcpy_h2d", "mdft_memcpy_d2h", "mdft_synchronize", "mdft_compute_vext",
    "mdft_get_vext", "mdft_vec_dot", "mdft_vec_axpy_dot", "mdft_vec_lincomb", "mdft_vec_multi_dot",
    "mdft_vec_combine", "mdft_pressure_correction", "mdft_lj_tail_correction", "mdft_angl2proj",
    "mdft_proj2angl", "mdft_regrid_trilinear", "mdft_compute_vext_ex", "mdft_vec_combine_step",
    "mdft_charged_solute_corrections", "mdft_vec_multi_dot_f32", "mdft_vec_combine_step_f32", "mdft_vec_axpy_f32",
    "mdft_vec_widen_f32", "mdft_scale_variables", "mdft_energy_and_gradient_dots",
    "mdft_vec_multi_dot_noalloc", "mdft_vec_combine_step_f32x", "mdft_vec_axpy_f32x", "mdft_vec_narrow_f64",
)


class MdftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdft error {code}: {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("mmax", C.c_int), ("mrso", C.c_int), ("kT", C.c_double), ("n0", C.c_double),
                ("precision", C.c_int), ("device", C.c_int)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_LIB = None


def load_library(path: str = LIBPATH):
    """Load libmdft_gpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.mdft_ctx_create.argtypes = [C.POINTER(Desc), C.POINTER(vp)]
    lib.mdft_ctx_create.restype = C.c_int
    lib.mdft_ctx_destroy.argtypes = [vp]
    lib.mdft_ctx_destroy.restype = None
    lib.mdft_last_error.argtypes = [vp]
    lib.mdft_last_error.restype = C.c_char_p
    lib.mdft_get_info.argtypes = [vp, _ip]
    lib.mdft_get_quadrature.argtypes = [vp, _dp, _dp, _dp]
    lib.mdft_set_vext.argtypes = [vp, vp, C.c_int]
    lib.mdft_set_ck.argtypes = [vp, C.c_int, _ip, _ip, _ip, _ip, _ip, C.c_int, C.c_double, _dp]
    lib.mdft_energy_and_gradient.argtypes = [vp, vp, _dp, vp, C.c_int]
    lib.mdft_excess_energy.argtypes = [vp, vp, _dp, vp, C.c_int]
    lib.mdft_ideal_external.argtypes = [vp, vp, _dp, _dp, vp, C.c_int]
    lib.mdft_excess_stage.argtypes = [vp, vp, C.c_int, vp, C.c_int]
    lib.mdft_phase_times.argtypes = [vp, _dp, C.c_int]
    lib.mdft_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    lib.mdft_device_free.argtypes = [vp, vp]
    lib.mdft_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
    lib.mdft_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
    lib.mdft_synchronize.argtypes = [vp]
    lib.mdft_compute_vext.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, C.c_double]
    lib.mdft_compute_vext_ex.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, C.c_double, C.c_int]
    lib.mdft_get_vext.argtypes = [vp, vp, C.c_int]
    lib.mdft_vec_dot.argtypes = [vp, C.c_size_t, vp, vp, _dp]
    lib.mdft_vec_axpy_dot.argtypes = [vp, C.c_size_t, C.c_double, vp, vp, vp, _dp]
    lib.mdft_vec_lincomb.argtypes = [vp, C.c_size_t, C.c_double, vp, C.c_double, vp, vp]
    lib.mdft_vec_multi_dot.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(vp), vp, _dp]
    lib.mdft_vec_multi_dot_noalloc.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(vp), vp, _dp]
    lib.mdft_pressure_correction.argtypes = [vp, vp, C.c_int, _dp]
    lib.mdft_lj_tail_correction.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, C.c_double, _dp]
    lib.mdft_charged_solute_corrections.argtypes = [_dp, C.c_double, C.c_int, _dp, C.c_int, _dp, C.c_double, _dp]
    lib.mdft_vec_multi_dot_f32.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(vp), vp, _dp]
    lib.mdft_vec_combine_step_f32.argtypes = [vp, C.c_size_t, C.c_int, _dp, C.POINTER(vp), vp, C.c_double, vp, vp]
    lib.mdft_vec_axpy_f32.argtypes = [vp, C.c_size_t, C.c_double, vp, vp, vp]
    lib.mdft_vec_widen_f32.argtypes = [vp, C.c_size_t, vp, vp]
    lib.mdft_vec_narrow_f64.argtypes = [vp, C.c_size_t, vp, vp]
    lib.mdft_vec_axpy_f32x.argtypes = [vp, C.c_size_t, C.c_double, vp, vp, vp]
    lib.mdft_vec_combine_step_f32x.argtypes = [vp, C.c_size_t, C.c_int, _dp, C.POINTER(vp), vp, C.c_double, vp, vp]
    lib.mdft_scale_variables.argtypes = [vp, C.c_size_t, C.c_int, C.c_int, vp, vp]
    lib.mdft_energy_and_gradient_dots.argtypes = [vp, vp, _dp, vp, C.c_int, C.c_int, C.POINTER(vp), _dp]
    lib.mdft_angl2proj.argtypes = [vp, vp, C.c_size_t, vp]
    lib.mdft_proj2angl.argtypes = [vp, vp, C.c_size_t, vp]
    lib.mdft_regrid_trilinear.argtypes = [vp, vp, _ip, _dp, C.c_double, vp]
    lib.mdft_vec_combine.argtypes = [vp, C.c_size_t, C.c_int, _dp, C.POINTER(vp), vp]
    lib.mdft_vec_combine_step.argtypes = [vp, C.c_size_t, C.c_int, _dp, C.POINTER(vp), vp, C.c_double, vp, vp]
    for name in ("mdft_get_info", "mdft_get_quadrature", "mdft_set_vext", "mdft_set_ck",
                 "mdft_energy_and_gradient", "mdft_excess_energy", "mdft_ideal_external", "mdft_excess_stage",
                 "mdft_phase_times", "mdft_device_alloc", "mdft_device_free", "mdft_memcpy_h2d",
                 "mdft_memcpy_d2h", "mdft_synchronize", "mdft_compute_vext", "mdft_get_vext", "mdft_vec_dot",
                 "mdft_vec_axpy_dot", "mdft_vec_lincomb", "mdft_vec_multi_dot", "mdft_vec_combine", "mdft_vec_combine_step",
                 "mdft_pressure_correction", "mdft_lj_tail_correction", "mdft_angl2proj", "mdft_proj2angl",
                 "mdft_regrid_trilinear", "mdft_compute_vext_ex", "mdft_charged_solute_corrections",
                 "mdft_vec_multi_dot_f32", "mdft_vec_combine_step_f32", "mdft_vec_axpy_f32", "mdft_vec_widen_f32",
                 "mdft_scale_variables", "mdft_energy_and_gradient_dots", "mdft_vec_multi_dot_noalloc",
                 "mdft_vec_combine_step_f32x", "mdft_vec_axpy_f32x", "mdft_vec_narrow_f64"):
        getattr(lib, name).restype = C.c_int
    _LIB = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


class DeviceBuffer:
    """Device allocation owned through the library (mdft_device_alloc)."""

    def __init__(self, engine: "Engine", nbytes: int):
        self.engine = engine
        self.nbytes = int(nbytes)
        ptr = C.c_void_p()
        engine._check(engine.lib.mdft_device_alloc(engine._h, self.nbytes, C.byref(ptr)))
        self.ptr = ptr

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        assert host.nbytes == self.nbytes
        self.engine._check(self.engine.lib.mdft_memcpy_h2d(self.engine._h, self.ptr, host.ctypes.data, self.nbytes))

    def download(self, dtype=np.float64) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        self.engine._check(self.engine.lib.mdft_memcpy_d2h(self.engine._h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.engine._h is not None:
            self.engine.lib.mdft_device_free(self.engine._h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """One device context: grid, angular quadrature, solvent bulk, c(k), vext."""

    def __init__(self, n, length, mmax, mrso, kT, n0, precision="fp64", device=0):
        self.lib = load_library()
        d = Desc(int(n[0]), int(n[1]), int(n[2]), float(length[0]), float(length[1]), float(length[2]),
                 int(mmax), int(mrso), float(kT), float(n0), PREC[precision], int(device))
        h = C.c_void_p()
        rc = self.lib.mdft_ctx_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise MdftError(rc, self.lib.mdft_last_error(None).decode())
        self._h = h
        self.n = tuple(int(v) for v in n)
        self.length = tuple(float(v) for v in length)
        self.mmax, self.mrso, self.kT, self.n0 = mmax, mrso, kT, n0
        self._refresh()

    def _refresh(self):
        info = np.zeros(10, np.int32)
        self._check(self.lib.mdft_get_info(self._h, info.ctypes.data_as(_ip)))
        self.ntheta, self.nphi, self.npsi, self.no, self.np_, self.npc, self.nq = (int(v) for v in info[:7])
        self.N = self.n[0] * self.n[1] * self.n[2]

    def _check(self, rc):
        if rc < 0:
            raise MdftError(rc, self.lib.mdft_last_error(self._h).decode())
        return rc

    def close(self):
        if getattr(self, "_h", None) is not None:
            self.lib.mdft_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up
    def quadrature(self):
        th, wt, w = np.zeros(self.ntheta), np.zeros(self.ntheta), np.zeros(self.no)
        self._check(self.lib.mdft_get_quadrature(self._h, _p(th), _p(wt), _p(w)))
        return th, wt, w

    def set_vext(self, vext):
        if isinstance(vext, DeviceBuffer):
            self._check(self.lib.mdft_set_vext(self._h, vext.ptr, DEVICE_PTRS))
        else:
            v = np.ascontiguousarray(vext, dtype=np.float64)
            assert v.size == self.N * self.no
            self._check(self.lib.mdft_set_vext(self._h, v.ctypes.data, 0))

    def set_ck(self, table):
        arr = [np.ascontiguousarray(getattr(table, k), dtype=np.int32) for k in ("m", "n", "mu", "nu", "khi")]
        c = np.ascontiguousarray(table.c, dtype=np.complex128)
        self._check(self.lib.mdft_set_ck(self._h, len(arr[0]), *[a.ctypes.data_as(_ip) for a in arr],
                                         c.shape[0], float(table.dq), c.view(np.float64).ctypes.data_as(_dp)))
        self._refresh()
        self.ck_table = table

    ELECTROSTATICS = {"poisson": 0, "direct": 1, "none": 2}

    def compute_vext(self, solute, solvent, rcut: float = 10.0, electrostatics: str = "poisson"):
        """vext of `solute` (mdft.workloads.Solute) for `solvent` (mdft.constants.Solvent), on the device.
        electrostatics: "poisson" (the reference's fast Poisson solver, its default),
        "direct" (minimum-image Coulomb) or "none"."""
        atoms = np.ascontiguousarray(np.column_stack([solute.xyz, solute.q, solute.sig, solute.eps]), dtype=np.float64)
        sites = np.ascontiguousarray([[*s.r, s.q, s.sig, s.eps] for s in solvent.sites], dtype=np.float64)
        self._check(self.lib.mdft_compute_vext_ex(self._h, len(atoms), atoms.ctypes.data_as(_dp), len(sites),
                                                  sites.ctypes.data_as(_dp), float(rcut),
                                                  self.ELECTROSTATICS[electrostatics]))

    def pressure_correction(self, xi) -> dict:
        """src/module_pressure_correction.f90:7-156 for xi (host array or DeviceBuffer)."""
        out = np.zeros(9)
        if isinstance(xi, DeviceBuffer):
            self._check(self.lib.mdft_pressure_correction(self._h, xi.ptr, DEVICE_PTRS, _p(out)))
        else:
            xi = np.ascontiguousarray(xi, dtype=np.float64)
            self._check(self.lib.mdft_pressure_correction(self._h, xi.ctypes.data, 0, _p(out)))
        keys = ("n_with_solute", "n_without_solute", "delta_n", "pmv_A3", "cavity_volume_A3", "F_empty_box",
                "P_bulk_kJmol_A3", "pmv_correction", "ideal_gas_correction")
        return dict(zip(keys, (float(v) for v in out)))

    def lj_tail_correction(self, solute, solvent, rcut: float = 10.0) -> float:
        """src/module_lennardjones.f90:145-149 (kJ/mol)."""
        atoms = np.ascontiguousarray(np.column_stack([solute.xyz, solute.q, solute.sig, solute.eps]), dtype=np.float64)
        sites = np.ascontiguousarray([[*s.r, s.q, s.sig, s.eps] for s in solvent.sites], dtype=np.float64)
        out = np.zeros(1)
        self._check(self.lib.mdft_lj_tail_correction(self._h, len(atoms), _p(atoms), len(sites), _p(sites),
                                                     float(rcut), _p(out)))
        return float(out[0])

    def charged_solute_corrections(self, solute, solvent) -> dict:
        """Type B (periodicity) and type C (P-scheme) corrections for a solute with a net
        charge (src/module_energy_and_gradient.f90:190-203, 287-344), kJ/mol."""
        return charged_solute_corrections(self.length, self.n0, solute, solvent)

    def get_vext(self):
        out = np.empty(self.N * self.no)
        self._check(self.lib.mdft_get_vext(self._h, out.ctypes.data, 0))
        return out

    def dot(self, x: "DeviceBuffer", y: "DeviceBuffer") -> float:
        out = np.zeros(1)
        self._check(self.lib.mdft_vec_dot(self._h, x.nbytes // 8, x.ptr, y.ptr, _p(out)))
        return float(out[0])

    def alloc_field(self) -> DeviceBuffer:
        return DeviceBuffer(self, self.N * self.no * 8)

    # ---- evaluations (host arrays or DeviceBuffer)
    def energy_and_gradient(self, xi, df=None, want_df=True, timers=False):
        """Returns (parts[F_id, F_ext, F_exc], df).  With DeviceBuffer xi, df must be a
        DeviceBuffer or None and the result stays on the device."""
        parts = np.zeros(3)
        flags = TIMERS if timers else 0
        if isinstance(xi, DeviceBuffer):
            ptr = df.ptr if df is not None else None
            self._check(self.lib.mdft_energy_and_gradient(self._h, xi.ptr, _p(parts), ptr, flags | DEVICE_PTRS))
            return parts, df
        x = np.ascontiguousarray(xi, dtype=np.float64)
        out = np.zeros(x.size) if want_df else None
        self._check(self.lib.mdft_energy_and_gradient(self._h, x.ctypes.data, _p(parts),
                                                      out.ctypes.data if want_df else None, flags))
        return parts, out

    def excess_energy(self, xi, df=None):
        """excess_energy_t plug-in: returns F_exc; df (host array) is accumulated into."""
        f = np.zeros(1)
        x = np.ascontiguousarray(xi, dtype=np.float64)
        if df is not None:
            assert df.flags.c_contiguous and df.dtype == np.float64
        self._check(self.lib.mdft_excess_energy(self._h, x.ctypes.data, _p(f),
                                                df.ctypes.data if df is not None else None, 0))
        return f[0]

    def ideal_external(self, xi, df=None):
        a, b = np.zeros(1), np.zeros(1)
        x = np.ascontiguousarray(xi, dtype=np.float64)
        self._check(self.lib.mdft_ideal_external(self._h, x.ctypes.data, _p(a), _p(b),
                                                 df.ctypes.data if df is not None else None, 0))
        return a[0], b[0]

    def excess_stage(self, xi, stage):
        x = np.ascontiguousarray(xi, dtype=np.float64)
        out = np.zeros(self.N * self.np_, dtype=np.complex128)
        self._check(self.lib.mdft_excess_stage(self._h, x.ctypes.data, int(stage), out.ctypes.data, 0))
        return out.reshape(self.N, self.np_)

    def phase_times(self):
        ms = np.zeros(8)
        k = self._check(self.lib.mdft_phase_times(self._h, _p(ms), 8))
        return ms[:k]

    def synchronize(self):
        self._check(self.lib.mdft_synchronize(self._h))


def charged_solute_corrections(length, n0: float, solute, solvent) -> dict:
    """Hunenberger's corrections that the reference adds to F for a solute with a net
    charge under lattice electrostatics (src/module_energy_and_gradient.f90:190-203):
    {"pbc_correction": type B, "pscheme_correction": type C}, kJ/mol, both 0 for a
    neutral solute.  Host arithmetic in the library; needs no device."""
    lib = load_library()
    atoms = np.ascontiguousarray(np.column_stack([solute.xyz, solute.q, solute.sig, solute.eps]), dtype=np.float64)
    sites = np.ascontiguousarray([[*s.r, s.q, s.sig, s.eps] for s in solvent.sites], dtype=np.float64)
    box = np.ascontiguousarray(length, dtype=np.float64)
    out = np.zeros(2)
    rc = lib.mdft_charged_solute_corrections(_p(box), float(n0), len(atoms), _p(atoms), len(sites), _p(sites),
                                             float(solvent.relative_permittivity), _p(out))
    if rc != 0:
        raise MdftError(rc, "mdft_charged_solute_corrections: invalid box or permittivity")
    return {"pbc_correction": float(out[0]), "pscheme_correction": float(out[1])}

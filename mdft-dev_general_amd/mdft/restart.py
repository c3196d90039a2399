"""The reference's restart file, density.bin: writer and reader (with regridding).

Writer: src/module_postprocessing.f90:46-94; reader: src/module_density.f90:61-261
(enabled there by `restart = T`).  The file is Fortran unformatted sequential:
every record is framed by its payload length as a little-endian int32 before and
after.  Records, in order:

    nspec (int32; negative: the file holds xi itself, positive: projections)
    mmax, no, np (one int32 record each)
    nx ny nz (3 x int32) and dx dy dz (3 x float64)
    nsite (int32), then one record per solute site: the components of the
    reference's site_type (src/module_system.f90:7-13) back to back -- name
    (100 characters), type, n_sites (int32), r(3), q, sig, eps, lambda1,
    lambda2 (float64), z (int32): 176 bytes
    then per grid point (x fastest, then y, then z): np complex128 projections
    of xi (angl2proj), one record per point; or, for full density, every
    xi(io, r) as a record of its own (io fastest).

Reading projections rebuilds xi with proj2angl and interpolates it trilinearly
from the file's grid, centred in the box, onto the engine's grid; points outside
the file's grid start from xi = 1, as in the reference.  A file holding xi itself
must have the engine's grid.  The transforms and the interpolation run on the
device (mdft_angl2proj, mdft_proj2angl, mdft_regrid_trilinear).
"""
from __future__ import annotations

import ctypes as C
import struct
import warnings

import numpy as np

from .gpu import DeviceBuffer, Engine

SITE_DTYPE = np.dtype([("name", "S100"), ("type", "<i4"), ("n_sites", "<i4"), ("r", "<f8", (3,)), ("q", "<f8"),
                       ("sig", "<f8"), ("eps", "<f8"), ("lambda1", "<f8"), ("lambda2", "<f8"), ("z", "<i4")])
assert SITE_DTYPE.itemsize == 176


def _record(fh, payload: bytes) -> None:
    n = struct.pack("<i", len(payload))
    fh.write(n)
    fh.write(payload)
    fh.write(n)


def _read_record(fh) -> bytes:
    head = fh.read(4)
    if len(head) < 4:
        raise EOFError("density.bin: truncated file")
    (n,) = struct.unpack("<i", head)
    if n < 0:
        raise ValueError("density.bin: record split into subrecords is not supported")
    payload = fh.read(n)
    tail = fh.read(4)
    if len(payload) < n or len(tail) < 4 or struct.unpack("<i", tail)[0] != n:
        raise ValueError("density.bin: truncated or corrupt record")
    return payload


def _ints(b: bytes):
    return struct.unpack(f"<{len(b) // 4}i", b)


def _framed(count: int, item: np.dtype, shape=()):
    """Structured dtype of `count` records of one item each (marker, payload, marker)."""
    payload = np.dtype((item, shape)) if shape else np.dtype(item)
    return np.dtype([("m1", "<i4"), ("v", payload), ("m2", "<i4")])


def sites_of(solute) -> np.ndarray:
    """site_type records of a workloads.Solute (name, type, n_sites, lambdas and z,
    which the solute does not carry, are blank / 0)."""
    n = len(solute.q)
    s = np.zeros(n, dtype=SITE_DTYPE)
    s["name"] = b" " * 100
    s["r"] = np.asarray(solute.xyz, dtype=np.float64)
    s["q"], s["sig"], s["eps"] = solute.q, solute.sig, solute.eps
    return s


def write_file(path: str, mmax: int, no: int, np_: int, n, d, sites: np.ndarray, data: np.ndarray,
               full: bool) -> None:
    """The records of density.bin from host data: projections [N][np] complex128, or
    (full) xi [N * no] float64 in the reference's point order."""
    with open(path, "wb") as fh:
        _record(fh, struct.pack("<i", -1 if full else 1))
        for v in (mmax, no, np_):
            _record(fh, struct.pack("<i", v))
        _record(fh, struct.pack("<3i", *n))
        _record(fh, struct.pack("<3d", *d))
        _record(fh, struct.pack("<i", len(sites)))
        for st in np.asarray(sites, dtype=SITE_DTYPE):
            _record(fh, st.tobytes())
        if full:
            rec = np.empty(data.size, dtype=_framed(1, "<f8"))
            rec["m1"] = rec["m2"] = 8
            rec["v"] = data.reshape(-1)
        else:
            rec = np.empty(data.shape[0], dtype=_framed(1, "<c16", (np_,)))
            rec["m1"] = rec["m2"] = 16 * np_
            rec["v"] = data
        rec.tofile(fh)


def read_file(path: str):
    """(header, data) of density.bin: data [N_file][np_file] complex128 projections, or
    (full) [N_file * no_file] float64 xi."""
    with open(path, "rb") as fh:
        h = read_header(fh)
        if h["nspec"] != 1:
            raise ValueError("density.bin: one solvent species supported")
        npv = h["n"][0] * h["n"][1] * h["n"][2]
        if h["full"]:
            rec = np.fromfile(fh, dtype=_framed(1, "<f8"), count=npv * h["no"])
            if len(rec) < npv * h["no"] or np.any(rec["m1"] != 8) or np.any(rec["m2"] != 8):
                raise ValueError("density.bin: truncated or corrupt density records")
        else:
            rec = np.fromfile(fh, dtype=_framed(1, "<c16", (h["np"],)), count=npv)
            if len(rec) < npv or np.any(rec["m1"] != 16 * h["np"]) or np.any(rec["m2"] != 16 * h["np"]):
                raise ValueError("density.bin: truncated or corrupt projection records")
    return h, np.ascontiguousarray(rec["v"])


def write_density_bin(path: str, eng: Engine, xi: DeviceBuffer, solute, full: bool = False) -> None:
    """Write xi (device field of `eng`) as the reference's density.bin:
    projections of xi (write_full_density = F) or xi itself (full = True)."""
    d = tuple(L / n for L, n in zip(eng.length, eng.n))
    if full:
        data = xi.download()
    else:
        proj = DeviceBuffer(eng, 16 * eng.N * eng.np_)
        eng._check(eng.lib.mdft_angl2proj(eng._h, xi.ptr, eng.N, proj.ptr))
        data = proj.download(np.complex128).reshape(eng.N, eng.np_)
        proj.free()
    write_file(path, eng.mmax, eng.no, eng.np_, eng.n, d, sites_of(solute), data, full)


def read_header(fh) -> dict:
    ns = _ints(_read_record(fh))[0]
    mmax = _ints(_read_record(fh))[0]
    no = _ints(_read_record(fh))[0]
    np_ = _ints(_read_record(fh))[0]
    n = _ints(_read_record(fh))
    d = struct.unpack("<3d", _read_record(fh))
    nsite = _ints(_read_record(fh))[0]
    sites = np.frombuffer(b"".join(_read_record(fh) for _ in range(nsite)), dtype=SITE_DTYPE)
    return {"full": ns < 0, "nspec": abs(ns), "mmax": mmax, "no": no, "np": np_, "n": tuple(n), "d": d,
            "sites": sites}


def read_density_bin(path: str, eng: Engine, xi: DeviceBuffer, solute=None) -> dict:
    """Fill xi (device field of `eng`) from the reference's density.bin and return
    its header.  With `solute`, every site in the file must match it (position
    shifted by the grid-centring offset, sigma, epsilon), as the reference checks;
    a different site count is reported, not fatal, as in the reference."""
    h, data = read_file(path)
    n_prev, d_prev = h["n"], h["d"]
    off = [(L - dp * npv) / 2 for L, dp, npv in zip(eng.length, d_prev, n_prev)]
    if solute is not None:
        if len(h["sites"]) != len(solute.q):
            warnings.warn("density.bin: the solutes differ in size")
        mine = sites_of(solute)
        for a, b in zip(h["sites"], mine):
            if (any(b["r"][k] != a["r"][k] + off[k] for k in range(3)) or b["sig"] != a["sig"]
                    or b["eps"] != a["eps"]):
                raise ValueError("density.bin: the solutes differ in parameters")
    npv = n_prev[0] * n_prev[1] * n_prev[2]
    if h["full"]:
        if tuple(n_prev) != tuple(eng.n) or h["no"] != eng.no:
            raise ValueError("density.bin holds xi on another grid or orientation set; "
                             "only a projection file can be regridded")
        xi.upload(data)
        return h
    # the reference reads np_file values into a zeroed array of the current np
    # (src/module_density.f90:162-169): a file of lower mmax is zero-padded, a higher one truncated
    proj = np.zeros((npv, eng.np_), dtype=np.complex128)
    k = min(h["np"], eng.np_)
    proj[:, :k] = data[:, :k]
    del data
    dproj = DeviceBuffer(eng, proj.nbytes)
    dproj.upload(proj)
    del proj
    dprev = DeviceBuffer(eng, 8 * npv * eng.no)
    eng._check(eng.lib.mdft_proj2angl(eng._h, dproj.ptr, npv, dprev.ptr))
    dproj.free()
    ni = (C.c_int * 3)(*n_prev)
    dd = (C.c_double * 3)(*d_prev)
    eng._check(eng.lib.mdft_regrid_trilinear(eng._h, dprev.ptr, ni, dd, 1.0, xi.ptr))
    dprev.free()
    return h

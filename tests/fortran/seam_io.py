"""Input/output files of the Fortran seam driver (tests/fortran/seam_check.f90)."""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def run(exe, n, length, mmax, kT, n0, precision, table, vext, xi):
    """Runs the driver; returns f, f (F only), (f_id, f_ext, f_exc), ff, df, df2."""
    d = tempfile.mkdtemp(prefix="mdft_seam_")
    fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
    with open(fin, "wb") as fh:
        np.array([*n, mmax, 2, precision, table.np_file, table.nq_file], np.int32).tofile(fh)
        np.array([*length, kT, n0, table.dq], np.float64).tofile(fh)
        for k in ("m", "n", "mu", "nu", "khi"):
            np.ascontiguousarray(getattr(table, k), dtype=np.int32).tofile(fh)
        np.ascontiguousarray(table.c, dtype=np.complex128).view(np.float64).tofile(fh)
        np.ascontiguousarray(vext, dtype=np.float64).tofile(fh)
        np.ascontiguousarray(xi, dtype=np.float64).tofile(fh)
    subprocess.run([exe, fin, fout], check=True, timeout=600)
    out = np.fromfile(fout, dtype=np.float64)
    nv = xi.size
    f, f2, fid, fext, fexc, ff = out[:6]
    df, df2 = out[6:6 + nv], out[6 + nv:6 + 2 * nv]
    for p in (fin, fout):
        os.remove(p)
    os.rmdir(d)
    return f, f2, np.array([fid, fext, fexc]), ff, df, df2

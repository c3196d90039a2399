"""Runner of the reference's own evaluation (oracle/_ref/mdft_ref_eval).

TEST INFRASTRUCTURE ONLY: writes the files the reference's modules read
(dft.in, the c-file under data/dcf/spce/, xi.bin and vext.bin) into a fresh
directory, runs the executable that oracle/Makefile (`make fftw ref`) builds
from the reference's sources (energy_ideal_and_external,
src/module_energy_ideal_and_external.f90:14-74, and energy_cproj_mrso,
src/module_energy_cproj_mrso.f90:61-782) and reads back F_id, F_ext, F_exc
and df.  Used by tests/golden/make_ref_golden.py (golden vectors, written in
the container that holds the reference) and by bench.py's CPU baseline leg.
"""
from __future__ import annotations

import os
import subprocess
import tempfile
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_ref", "mdft_ref_eval")


@dataclass
class RefResult:
    kT: float
    n0: float
    rho0: float
    dv: float
    parts: np.ndarray        # F_id, F_ext, F_exc (kJ/mol)
    df: np.ndarray | None    # (1/dv) dF/dxi, reference layout (no fastest)
    t_first: float           # wall seconds of the first evaluation (plans FFTs, reads c(k))
    t_eval: float            # mean wall seconds of the later evaluations (0 if nrep == 1)
    threads: int


def available() -> bool:
    return os.path.exists(EXE)


def write_ck_file_for_flang(path: str, t, mrso: int = 2) -> None:
    """Write the c-file `t` (an mdft.ck.CkTable) so that the reference's
    parsers, compiled with amdflang, read it as the reference would read the
    standard file.

    deltaAbscissa (src/module_input.f90:591-650) reads every row with
    `read(10,*) abscissa, useless_stuff` (a one-character item).  flang's
    list-directed input refuses a character item that starts with '(' (the
    first complex value of a standard row), so the standard format stops
    there.  A row written as `q,,(c2) (c3) ...` reads under flang as
      - deltaAbscissa: q, then a null value for the character item;
      - read_c_luc (src/module_read_c_luc.f90:105): q, a null value for the
        first column (which keeps the (0, 0) the array was allocated with,
        :103), then the other columns.
    The first column is therefore chosen among the tuples the relayout never
    reads (mu < 0: the relayout of src/module_energy_cproj_mrso.f90:381-401
    takes mu2 = 0..m/mrso only), or, when there is none (mmax = 1), a column
    that is zero at every q.  The header order follows the columns, as
    read_c_luc maps tuples through the header (:85-95).  Values are written in
    full precision, so the table the reference reads is `t` exactly."""
    import numpy as np_
    cols = list(range(t.np_file))
    unused = [i for i in cols if int(t.mu[i]) < 0]
    zero = [i for i in cols if not np_.any(t.c[:, i])]
    if unused:
        first = unused[0]
    elif zero:
        first = zero[0]
    else:
        raise ValueError("no column of this table can be written as a null value "
                         "(needs a tuple with mu < 0 or an all-zero column)")
    order = [first] + [i for i in cols if i != first]

    def row(label, arr):
        return f" {label:<6}" + "".join(f"{int(arr[i]):12d}" for i in order)

    q = t.q if t.q is not None else np_.arange(t.nq_file) * t.dq
    with open(path, "w") as fh:
        fh.write("\n" * 10)
        fh.write(f" {'alpha':<6}" + "".join(f"{k + 1:12d}" for k in range(t.np_file)) + "\n")
        for lab, arr in (("m", t.m), ("n", t.n), ("mu", t.mu), ("nu", t.nu), ("khi", t.khi)):
            fh.write(row(lab, arr) + "\n")
        fh.write("\n")
        for iq in range(t.nq_file):
            vals = "".join(f" ({float(t.c[iq, i].real)!r},{float(t.c[iq, i].imag)!r})" for i in order[1:])
            fh.write(f"{float(q[iq])!r},,{vals}\n")


def table_as_read(t):
    """The table exactly as the reference reads `write_ck_file_for_flang(t)`:
    the first written column becomes (0, 0), and dq is deltaAbscissa's
    difference of the 4th and 3rd abscissae (module_input.f90:603-617)."""
    import copy
    import numpy as np_
    cols = list(range(t.np_file))
    unused = [i for i in cols if int(t.mu[i]) < 0]
    first = unused[0] if unused else [i for i in cols if not np_.any(t.c[:, i])][0]
    out = copy.deepcopy(t)
    out.c = np_.array(t.c, dtype=np_.complex128, copy=True)
    out.c[:, first] = 0.0
    q = t.q if t.q is not None else np_.arange(t.nq_file) * t.dq
    out.dq = float(q[3]) - float(q[2])
    return out


def write_inputs(d: str, n, length, mmax: int, temperature: float, ck_writer) -> None:
    """dft.in with the tags module_grid / module_thermo / module_solvent read
    (module_grid.f90:62-95, module_thermo.f90:25, module_solvent.f90:188-192),
    and the c-file at the path read_c_luc opens (module_read_c_luc.f90:24)."""
    os.makedirs(os.path.join(d, "data", "dcf", "spce"), exist_ok=True)
    os.makedirs(os.path.join(d, "output"), exist_ok=True)
    with open(os.path.join(d, "dft.in"), "w") as fh:
        fh.write(f"boxlen = {length[0]!r} {length[1]!r} {length[2]!r}\n")
        fh.write(f"boxnod = {n[0]} {n[1]} {n[2]}\n")
        fh.write(f"mmax = {mmax}\n")
        fh.write("molrotsymorder = 2\n")
        fh.write(f"temperature = {temperature!r}\n")
        fh.write("solvent = spce\n")
    ck_writer(os.path.join(d, "data", "dcf", "spce", f"spce-ck_nonzero_nmax{mmax}_ml"))


def run(n, length, mmax: int, temperature: float, ck_writer, xi: np.ndarray | None, vext: np.ndarray | None,
        nrep: int = 1, want_df: bool = True, threads: int | None = None, workdir: str | None = None,
        timeout: float = 3600.0) -> RefResult:
    """One (or nrep) reference evaluations of F and df at xi with vext.

    xi, vext: FP64 arrays of N*no values in the reference layout xi(no,nx,ny,nz), or
    both None for the executable's own index-formula fields (timing at full size
    without writing the fields to disk).
    ck_writer(path) writes the c-file (write_ck_file_for_flang bound to a table)."""
    if not available():
        raise FileNotFoundError(f"{EXE} is not built (make -C oracle fftw ref)")
    own = workdir is None
    d = tempfile.mkdtemp(prefix="mdft_ref_") if own else workdir
    try:
        write_inputs(d, n, length, mmax, temperature, ck_writer)
        synthetic = xi is None and vext is None
        if not synthetic:
            np.ascontiguousarray(xi, dtype=np.float64).tofile(os.path.join(d, "xi.bin"))
            np.ascontiguousarray(vext, dtype=np.float64).tofile(os.path.join(d, "vext.bin"))
        env = dict(os.environ)
        if threads is not None:
            env["OMP_NUM_THREADS"] = str(int(threads))
        env.setdefault("OMP_STACKSIZE", "64M")
        subprocess.run([EXE, str(int(nrep)), "1" if want_df else "0", "synthetic" if synthetic else "files"],
                       cwd=d, env=env, check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout)
        with open(os.path.join(d, "ref_out.txt")) as fh:
            rows = [ln.split() for ln in fh if ln.strip()]
        kT, n0, rho0, dv = (float(v) for v in rows[0])
        parts = np.array([float(v) for v in rows[1]])
        t_first, t_eval = (float(v) for v in rows[2])
        nthr = int(rows[3][0])
        df = np.fromfile(os.path.join(d, "df.bin"), dtype=np.float64) if want_df else None
        return RefResult(kT, n0, rho0, dv, parts, df, t_first, t_eval, nthr)
    finally:
        if own:
            import shutil
            shutil.rmtree(d, ignore_errors=True)

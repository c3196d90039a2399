"""Writes tests/golden/ref_energy_golden.npz from the reference's own evaluation.

Runs in the container that holds the reference (/root/reference): builds
oracle/_ref (FFTW 3.3.4 from the reference's source tarball, then the
reference's energy_ideal_and_external and energy_cproj_mrso with the modules
they use, driven by oracle/ref_eval.f90), runs it on every case of
ref_cases.CASES and stores, per case:
  parts      F_id, F_ext, F_exc (kJ/mol)
  consts     kT, n0, rho0, dv as the reference computed them
  df_idx     seeded sample positions in dF (reference layout)
  df_val     dF at those positions
  df_stats   sum(dF), sum(dF^2), max |dF|
  ck_*       the c(k) table as the reference read it (stored-table cases),
             or ck_sum = sum |c| of the table (synthetic cases, a drift check)
  in_stats   sum(xi), sum(vext) (a drift check of the seeded inputs)

    python tests/golden/make_ref_golden.py [--only name ...]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "mdft-dev_general_amd"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import ref_cases as RC  # noqa: E402
from mdft import ck  # noqa: E402
from oracle import ref_eval as R  # noqa: E402

SPCE = "/root/reference/data/dcf/spce/spce-ck_nonzero_nmax{}_ml"


def table_for(c):
    name, n, L, mmax, kind, tseed, nk, _, _ = c
    if kind == "synthetic":
        return RC.synthetic_table(mmax, tseed, nk), False
    nmax = int(kind[-1])
    t = ck.read_ck_file(SPCE.format(nmax))
    if nmax == 1:
        RC._zero_khi_m1(t)
    nq = ck.nq_needed(n, L, t.dq)
    t.c = t.c[: nq + 4].copy()
    t.q = t.q[: nq + 4].copy()
    return t, True


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args()
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "fftw"])
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    out = dict(np.load(RC.GOLDEN)) if (args.only and os.path.exists(RC.GOLDEN)) else {}
    for c in RC.CASES:
        name, n, L, mmax, kind, tseed, nk, iseed, _ = c
        if args.only and name not in args.only:
            continue
        t, stored = table_for(c)
        no = RC.no_of(mmax)
        xi, vext = RC.inputs(n, no, iseed)
        t0 = time.time()
        r = R.run(n, L, mmax, RC.TEMPERATURE, lambda p: R.write_ck_file_for_flang(p, t), xi, vext)
        idx = RC.sample_index(xi.size, iseed)
        p = name + "/"
        out[p + "parts"] = r.parts
        out[p + "consts"] = np.array([r.kT, r.n0, r.rho0, r.dv])
        out[p + "df_idx"] = idx.astype(np.int64)
        out[p + "df_val"] = r.df[idx]
        out[p + "df_stats"] = np.array([r.df.sum(), np.dot(r.df, r.df), np.abs(r.df).max()])
        out[p + "in_stats"] = np.array([xi.sum(), vext.sum()])
        tr = R.table_as_read(t)
        if stored:
            for k in ("m", "n", "mu", "nu", "khi"):
                out[p + "ck_" + k] = np.asarray(getattr(tr, k), dtype=np.int32)
            out[p + "ck_re"] = tr.c.real.copy()
            out[p + "ck_im"] = tr.c.imag.copy()
            out[p + "ck_q"] = np.asarray(tr.q, dtype=np.float64)
            out[p + "ck_dq"] = np.array(tr.dq)
        else:
            out[p + "ck_sum"] = np.array(np.abs(tr.c).sum())
            out[p + "ck_dq"] = np.array(tr.dq)
        print(f"{name}: F = {r.parts} (reference, {time.time() - t0:.1f} s, {r.threads} threads)", flush=True)
    np.savez_compressed(RC.GOLDEN, **out)
    print(f"wrote {RC.GOLDEN} ({os.path.getsize(RC.GOLDEN)} bytes)")


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the MDFT solver on one MI355X.

Headline: seconds until the solvation free energy is within 0.05 kJ/mol of
the converged FP64 value, for the 20-atom synthetic solute on the 128^3 grid
(40 A box), mmax = 5 (330 orientations, 88 projections), one GPU.

Both timed paths stop by the reference's own rule -- one iteration lowers F by
at most tol = 1e-3 kJ/mol (src/module_lbfgs_nocedal_mdft.f90:425-427) -- which
knows nothing of the answer; each is then checked against Delta G_ref:

* Delta G_ref: the plain FP64 path (reference cold start, xi = 0 inside the
  solute and 1 elsewhere; FP64 evaluations; device-resident L-BFGS with memory
  1 over xi, unit first step) run to |dF|/|dF0| < 1e-6; its final F.
* plain FP64 path (the denominator): the same method stopped by the rule, T_plain.
* fast path (`value`): the calibrated warm start inside the timed region (the mix
  (1 - lam) cold start + lam network prediction; lam is calibrated per network on
  other solutes, and at this workload lam = 0: the reference's cold start, the
  network not run), FP32 evaluations (angular transforms, FFTs, OZ and
  per-orientation terms in FP32; sums in FP64), L-BFGS with FP64 scalars, FP32
  vectors and iterate, the weight-scaled variables and a bulk Newton first step,
  stopped by the rule, plus one FP64 evaluation of the final iterate (the
  confirmed Delta G).
  speedup_band_to_band = T_plain / value.

Also reported: both paths stopped by an oracle at the first iterate within the
band (what the stop rule costs), the plain run to full convergence, the FP64
method over the scaled variables, the FP32 path from the cold, Boltzmann and
pure-network starts, F+dF evaluations per second (one step = one evaluation over
the whole grid, inputs resident in HBM), per-phase HIP-event timings on the
library's stream, the roofline of the dominant kernel of the fast path's own
evaluation, and the reference's own CPU evaluation (compiled from its sources)
timed on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mdft-dev_general_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "s to ΔG_solv within 0.05 kJ/mol, 128³ mmax=5: learned+mixed vs plain FP64 GPU"
PHASES = ["angl2proj", "fft_forward", "oz", "fft_inverse", "proj2angl_ideal_grad", "reductions", "total"]


def phase_bytes(N, no, np_, b=16, fast=False):
    """Algorithmic HBM bytes of each phase of one evaluation (complex of b bytes).
    fast: the fast minimiser's evaluation -- FP32 iterate (4 B), the final pass reading the
    FP32 copy of vext (4 B) and two FP32 vectors for the gradient products (8 B) and
    writing the FP32 gradient (4 B); otherwise FP64 xi and vext read, FP64 dF written."""
    pb = N * np_ * b
    x = 4 if fast else 8
    return {
        "angl2proj": N * no * x + pb,
        "fft_forward": 3 * 2 * pb,            # three axis passes, read + write
        "oz": 2 * pb,
        "fft_inverse": 3 * 2 * pb,
        "proj2angl_ideal_grad": pb + N * no * (20 if fast else 24),
    }


# per-launch view of the phases: (kernel, launches per evaluation, kernel name in the traffic profile)
KERNELS = {
    "angl2proj": ("angl2proj", 1, "angl2proj_t"),
    "fft": ("fft_axis_r (one axis pass)", 6, "fft_axis_r"),
    "oz": ("oz", 1, "oz"),
    "proj2angl_ideal_grad": ("final_t (proj2angl + ideal + gradient)", 1, "final_t"),
}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_eval_fp32_solute20_128_mmax5.md")
TRAFFIC_PROFILE_FAST = os.path.join(ROOT, "profiles", "traffic_fastpath_solute20_128_mmax5.md")


def fast_eval_phases(e32, xi, reps=3):
    """Median per-phase HIP-event times (ms) of the fast minimiser's own evaluation --
    FP32 iterate u = s xi, FP32 gradient, scaled variables, the products of the new
    gradient with two FP32 vectors and itself formed in the final pass -- at xi."""
    import ctypes as C
    from mdft.gpu import DEVICE_PTRS, DF_FP32, SCALED_VARS, TIMERS, XI_FP32, DeviceBuffer
    n = e32.N * e32.no
    u = DeviceBuffer(e32, 8 * n)
    e32._check(e32.lib.mdft_scale_variables(e32._h, n, 1, 0, xi.ptr, u.ptr))
    u32, g32, v0, v1 = (DeviceBuffer(e32, 4 * n) for _ in range(4))
    e32._check(e32.lib.mdft_vec_narrow_f64(e32._h, n, u.ptr, u32.ptr))
    u.free()
    flags = DEVICE_PTRS | DF_FP32 | XI_FP32 | SCALED_VARS
    dp = C.POINTER(C.c_double)
    parts, dots = np.zeros(3), np.zeros(3)
    e32._check(e32.lib.mdft_energy_and_gradient(e32._h, u32.ptr, parts.ctypes.data_as(dp), g32.ptr, flags))
    for v in (v0, v1):   # two stored vectors of the gradient's size and scale
        e32._check(e32.lib.mdft_vec_axpy_f32x(e32._h, n, 0.0, g32.ptr, g32.ptr, v.ptr))
    vecs = (C.c_void_p * 3)(v0.ptr.value, v1.ptr.value, g32.ptr.value)
    ph = []
    for _ in range(reps):
        e32._check(e32.lib.mdft_energy_and_gradient_dots(e32._h, u32.ptr, parts.ctypes.data_as(dp), g32.ptr,
                                                         flags | TIMERS, 3, vecs, dots.ctypes.data_as(dp)))
        ph.append(e32.phase_times())
    for b in (u32, g32, v0, v1):
        b.free()
    return np.median(np.array(ph), axis=0)


def measured_traffic(prefix, fast=False):
    """HBM bytes per launch (fetched + written) of the kernel whose name starts with
    `prefix` in the committed rocprofv3 PMC summary of the FP32 evaluation (fast: of the
    fast minimiser's evaluation), or None."""
    import re
    try:
        txt = open(TRAFFIC_PROFILE_FAST if fast else TRAFFIC_PROFILE).read()
    except OSError:
        return None
    for line in txt.splitlines():
        m = re.match(r"\|\s*`([^`]+)`\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)\s*\|", line)
        if m and m.group(1).startswith(prefix):
            return (float(m.group(2)) + float(m.group(3))) * 1e9
    return None


def roofline_of(ph, N, no, np_, cbytes, traffic=False, fast=False):
    """Roofline of the dominant kernel from median phase times: per launch, the
    algorithmic bytes over the average launch duration (HIP events on the
    library's stream)."""
    pbytes = phase_bytes(N, no, np_, cbytes, fast)
    per = {"angl2proj": (ph[0], pbytes["angl2proj"]),
           "fft": (ph[1] + ph[3], pbytes["fft_forward"] + pbytes["fft_inverse"]),
           "oz": (ph[2], pbytes["oz"]),
           "proj2angl_ideal_grad": (ph[4], pbytes["proj2angl_ideal_grad"])}
    dom = max(per, key=lambda k: per[k][0])
    name, launches, prof = KERNELS[dom]
    ms, nbytes = per[dom][0] / launches, per[dom][1] / launches
    achieved = nbytes / (ms * 1e-3) / 1e9
    t = measured_traffic(prof, fast) if traffic else None
    return {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": t, "algorithmic_bytes_per_launch": nbytes,
            "launch_ms": ms, "launches_per_evaluation": launches}


def algorithmic_bytes(N, no, np_, b=16):
    """Fused minimum of one evaluation: xi twice, vext once, dF once (FP64) and
    the projection field read and written in four phases."""
    return N * no * (2 * 8 + 8 + 8) + N * 8 * np_ * b


def cpu_baseline(w, evals_to_converge):
    """The reference's own FP64 evaluation on the host cores: energy_ideal_and_external
    + energy_cproj_mrso compiled from the reference's sources with FFTW 3.3.4
    (oracle/_ref/mdft_ref_eval, OpenMP), run on the full workload grid (128^3, mmax 5,
    the workload's c(k)) twice; the second evaluation's wall time (the first plans the
    FFTs) times the number of evaluations the plain GPU minimisation needed.  Without
    that build, the C port of the oracle is timed the same way."""
    from oracle import ref_eval as R
    t = w.ck_table()
    if R.available():
        r = R.run(w.n, w.length, w.mmax, 298.0, lambda p: R.write_ck_file_for_flang(p, t), None, None,
                  nrep=2, want_df=False)
        t_eval, cores, kind = r.t_eval, r.threads, "reference"
        what = (f"the reference's energy_ideal_and_external + energy_cproj_mrso (amdflang -O2 -fopenmp, "
                f"FFTW 3.3.4), {w.n[0]}^3 mmax={w.mmax}, one F+dF evaluation after a planning one: "
                f"{t_eval:.2f} s on {cores} threads")
    else:
        from oracle import oracle as O
        o = O.Oracle(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0)
        o.set_ck(t)
        rng = np.random.default_rng(0)
        nno = o.N * o.no
        vext = rng.uniform(-1, 3, nno)
        xi = 1.0 + 0.1 * rng.uniform(-1, 1, nno)
        t0 = time.perf_counter()
        o.energy_and_gradient(xi, vext)
        t_eval, cores, kind = time.perf_counter() - t0, O.num_threads(), "port"
        what = f"the C oracle, {w.n[0]}^3 mmax={w.mmax}, one F+dF evaluation: {t_eval:.2f} s on {cores} threads"
    cpu = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                cpu = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": t_eval * evals_to_converge, "unit": "s", "cores": cores, "kind": kind,
            "evaluations_per_s": 1.0 / t_eval, "cpu": cpu,
            "sample": f"{what}, times the {evals_to_converge} evaluations of the plain GPU minimisation"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed evaluations for the evaluation rate")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="solute20_128")
    ap.add_argument("--gtol", type=float, default=1e-6)
    ap.add_argument("--memory", type=int, default=1, help="L-BFGS memory m of the plain path (1 is the reference's)")
    ap.add_argument("--fast-memory", type=int, default=None,
                    help="L-BFGS memory m of the fast path (default: --memory)")
    ap.add_argument("--fp64-vectors", action="store_true",
                    help="fast path: keep the minimiser's directions and gradients in FP64 (default FP32)")
    ap.add_argument("--fp64-iterate", action="store_true",
                    help="fast path: keep the minimiser's iterate in FP64 (default FP32, with FP32 vectors)")
    ap.add_argument("--unscaled", action="store_true",
                    help="fast path: minimise over xi itself (default: over u = sqrt(w / mean w) xi)")
    ap.add_argument("--unit-first-step", action="store_true",
                    help="fast path: first step of unit length along -g (default: the bulk Newton step)")
    ap.add_argument("--maxeval", type=int, default=1500)
    ap.add_argument("--mode", choices=["converge", "eval", "copy", "fastpath"], default="converge",
                    help="copy: a plain device copy of 1 GiB (z = x through mdft_vec_lincomb), the "
                         "reference kernel for calibrating rocprof's FETCH_SIZE / WRITE_SIZE counters; "
                         "fastpath: the fast path's minimisation alone (cold start, FP32 evaluations, "
                         "the reference's stop rule), --steps times, for kernel traces of the minimiser")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16"], default="fp64",
                    help="evaluation precision for the rate/phase measurement")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--weights", default=os.path.join(PKG, "mdft", "data", "warmstart_solute20.npz"),
                    help="trained warm-start network")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    from mdft import workloads
    from mdft.gpu import Engine
    from mdft.minimize import minimize

    w = workloads.WORKLOADS[args.workload]
    if args.mode == "copy":
        eng = Engine((8, 8, 8), (8.0, 8.0, 8.0), 1, 2, w.kT, w.n0, "fp64")
        n = 1 << 27  # doubles: 1 GiB read + 1 GiB written per launch
        from mdft.gpu import DeviceBuffer
        x, z = DeviceBuffer(eng, 8 * n), DeviceBuffer(eng, 8 * n)
        for _ in range(args.warmup + args.steps):
            eng._check(eng.lib.mdft_vec_lincomb(eng._h, n, 1.0, x.ptr, 0.0, None, z.ptr))
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng._check(eng.lib.mdft_vec_lincomb(eng._h, n, 1.0, x.ptr, 0.0, None, z.ptr))
        eng.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"mode": "copy", "bytes_read": 8 * n, "bytes_written": 8 * n, "ms": dt * 1e3,
                          "GBps": 16 * n / dt / 1e9}), flush=True)
        return
    sol = w.solute()
    if args.mode == "fastpath":
        from mdft import warmstart as wsm
        e32 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp32")
        e32.set_ck(w.ck_table())
        e32.compute_vext(sol, w.solvent)
        xi, df = e32.alloc_field(), e32.alloc_field()
        m = args.fast_memory if args.fast_memory is not None else args.memory
        for rep in range(max(args.steps, 1)):
            e32.synchronize()
            t0 = time.perf_counter()
            wsm.cold_start(e32, xi)
            r = minimize(e32, xi, g=df, m=m, gtol=0.0, maxeval=args.maxeval, fnoise=1e-6,
                         fp32_vectors=not args.fp64_vectors, scaled=not args.unscaled,
                         fp32_iterate=not (args.fp64_vectors or args.fp64_iterate),
                         step0=0.0 if (args.unit_first_step or args.unscaled) else "newton", ftol=1e-3,
                         ftol_absolute=True)
            e32.synchronize()
            print(json.dumps({"mode": "fastpath", "rep": rep, "s": time.perf_counter() - t0,
                              "evaluations": r.evaluations, "iterations": r.iterations, "F": r.F}), flush=True)
        return
    t_setup = time.perf_counter()
    ckt = w.ck_table()
    eng = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, args.precision)
    eng.set_ck(ckt)
    eng.compute_vext(sol, w.solvent)
    vext_h = eng.get_vext()
    xi0 = workloads.guess_xi(vext_h, w.kT)
    del vext_h
    xi = eng.alloc_field()
    xi.upload(xi0)
    df = eng.alloc_field()
    t_setup = time.perf_counter() - t_setup

    # ---- evaluation rate and phase timings
    for _ in range(args.warmup):
        eng.energy_and_gradient(xi, df)
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        parts, _ = eng.energy_and_gradient(xi, df)
    eng.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ph = []
    for _ in range(3):
        eng.energy_and_gradient(xi, df, timers=True)
        ph.append(eng.phase_times())
    ph = np.median(np.array(ph), axis=0)
    cbytes = 16 if args.precision == "fp64" else 8  # projection field: complex FP64 or FP32
    algo = algorithmic_bytes(eng.N, eng.no, eng.np_, cbytes)
    roof = roofline_of(ph, eng.N, eng.no, eng.np_, cbytes, traffic=args.precision == "fp32")
    roof.update({"whole_evaluation_algorithmic_GB": algo / 1e9, "whole_evaluation_frac": algo / dt / 1e9 / HBM_PEAK_GBS})

    line = {
        "metric": METRIC,
        "n_gpus": 1,
        "warmup": args.warmup,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 (f32 transforms and per-orientation terms, FP64 sums)",
                  "bf16": "bf16 MFMA angular transforms, f32 FFT/OZ (FP64 reductions, ideal term, L-BFGS)"}[args.precision],
        "data": "synthetic (seeded 20-atom solute, vext with Poisson electrostatics as the reference by default, seeded c(k) with the convex self term, reference cold start)",
        "config": {"workload": w.name, "grid": list(w.n), "box_A": list(w.length), "mmax": w.mmax,
                   "orientations": eng.no, "projections": eng.np_, "natoms": int(len(sol.q))},
        "evaluations_per_s": 1.0 / dt,
        "ms_per_evaluation": dt * 1e3,
        "phases_ms": {k: float(v) for k, v in zip(PHASES, ph)},
        "roofline": roof,
        "setup_s": t_setup,
    }

    if args.mode == "eval":
        line.update({"metric": f"F+dF evaluations/s, 128^3 mmax=5, {args.precision.upper()}", "value": 1.0 / dt,
                     "unit": "evaluations/s", "higher_is_better": True, "steps": args.steps,
                     "ms_per_step": dt * 1e3})
    else:
        if args.precision != "fp64":
            raise SystemExit("bench.py: converge mode measures the FP64 reference itself; use --precision fp64")
        band = 0.05
        # the reference's own stop: one iteration lowering F by at most tol = factr * epsmch
        # = 1e-3 kJ/mol (src/module_lbfgs_nocedal_mdft.f90:10,188,425-427); it reads nothing of
        # the answer, and both paths of the headline stop by it
        ref_tol = 1e-3
        from mdft import warmstart as wsm
        # ---- plain FP64 path: cold start (on the device) to |dF|/|dF0| < gtol -> Delta G_ref
        eng.synchronize()
        t0 = time.perf_counter()
        wsm.cold_start(eng, xi)
        res = minimize(eng, xi, g=df, m=args.memory, gtol=args.gtol, maxeval=args.maxeval, fnoise=1e-7)
        eng.synchronize()
        t_conv = time.perf_counter() - t0
        dg_ref = res.F

        def plain(stop, scaled=False):
            eng.synchronize()
            t0 = time.perf_counter()
            wsm.cold_start(eng, xi)
            kw = dict(ftol=ref_tol, ftol_absolute=True) if stop == "ref" else dict(ftarget=dg_ref, fband=band)
            r = minimize(eng, xi, g=df, m=args.memory, gtol=0.0, maxeval=args.maxeval, fnoise=1e-7, scaled=scaled,
                         **kw)
            eng.synchronize()
            return time.perf_counter() - t0, r

        # ---- the plain FP64 method stopped by the reference's rule (the headline's denominator),
        # and the same run stopped by the oracle at the first iterate inside the band
        t_plain, res_p = plain("ref")
        t_plain_band, res_pb = plain("band")
        # ---- the FP64 method over the scaled variables (what the change of variables alone gives)
        t_scaled64, res_s = plain("ref", scaled=True)
        # ---- fast path: learned warm start + FP32 evaluations, confirmed by one FP64
        # evaluation of the final iterate
        e32 = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp32")
        e32.set_ck(ckt)
        e32.compute_vext(sol, w.solvent)
        net = wsm.WarmStartNet.load(args.weights) if os.path.exists(args.weights) else None
        info = {}

        fast_m = args.fast_memory if args.fast_memory is not None else args.memory

        def fast(start, stop="ref"):
            e32.synchronize()
            eng.synchronize()
            t0 = time.perf_counter()
            if start == "net":
                info.update(wsm.warm_start(e32, net, xi))
            elif start == "net_pure":
                wsm.warm_start(e32, net, xi, start_mix=1.0)
            elif start == "boltzmann":
                wsm.boltzmann_start(e32, xi)
            else:
                wsm.cold_start(e32, xi)
            kw = dict(ftol=ref_tol, ftol_absolute=True) if stop == "ref" else dict(ftarget=dg_ref, fband=band)
            r = minimize(e32, xi, g=df, m=fast_m, gtol=0.0, maxeval=args.maxeval, fnoise=1e-6,
                         fp32_vectors=not args.fp64_vectors, scaled=not args.unscaled,
                         fp32_iterate=not (args.fp64_vectors or args.fp64_iterate),
                         step0=0.0 if (args.unit_first_step or args.unscaled) else "newton", **kw)
            c, _ = eng.energy_and_gradient(xi, None)
            eng.synchronize()
            return time.perf_counter() - t0, r, float(np.sum(c))

        t_cold32, res_c32, dg_c32 = fast("cold")
        t_boltz32, res_b32, dg_b32 = fast("boltzmann")
        t_pure = res_pure = dg_pure = None
        if net is not None:
            t_pure, res_pure, dg_pure = fast("net_pure")
            t_fast, res_f, dg_fast = fast("net")
            t_fast_band, res_fb, dg_fast_band = fast("net", stop="band")
        else:
            t_fast, res_f, dg_fast = t_cold32, res_c32, dg_c32
            t_fast_band, res_fb, dg_fast_band = fast("cold", stop="band")
        # the fast path's own evaluation (the variant its minimiser calls), timed by phase
        # (HIP events on the FP32 engine's stream): the roofline of the dominant kernel of
        # the timed region
        ph32 = fast_eval_phases(e32, xi)
        line["roofline_fp64"] = line["roofline"]
        line["roofline"] = roofline_of(ph32, e32.N, e32.no, e32.np_, 8, traffic=True, fast=True)
        line["phases_ms_fp32"] = {k: float(v) for k, v in zip(PHASES, ph32)}
        # reference-style solvation free energy of the fast path's final density
        # (src/module_pressure_correction.f90: F + PMV correction + LJ tail), outside the timed region
        pc = eng.pressure_correction(xi)
        lj = eng.lj_tail_correction(sol, w.solvent)
        qc = eng.charged_solute_corrections(sol, w.solvent)   # 0 for a neutral solute
        dq = qc["pbc_correction"] + qc["pscheme_correction"]
        line["dG_post_processing"] = {"F_kJmol": dg_fast, "pmv_correction": pc["pmv_correction"],
                                      "lj_tail_correction": lj, "pmv_A3": pc["pmv_A3"],
                                      "P_bulk_atm": pc["P_bulk_kJmol_A3"] * 16394.078514951, **qc,
                                      "estimated_solvation_free_energy": dg_fast + dq + pc["pmv_correction"] + lj}
        # seconds-to-band roofline of the fast path: each FP32 evaluation at its fused
        # algorithmic bytes and each L-BFGS iteration's vector passes at their bytes per
        # element -- the direction from 2m + 1 stored vectors with the trial point (x read,
        # y written, FP64), and the products of the new gradient with the 2m + 1 live
        # vectors (with FP32 vectors at m = 1 formed in the final pass: the 2 vectors other
        # than the gradient itself are read there) -- all at the HBM peak
        nvec = eng.N * eng.no
        vb = 4 if not args.fp64_vectors else 8
        xb = 4 if not (args.fp64_vectors or args.fp64_iterate) else 8   # iterate and trial points
        dot_bytes = 2 * fast_m * vb if (vb == 4 and fast_m == 1) else (2 * fast_m + 1) * vb
        per_elem = (2 * fast_m + 1) * vb + vb + 2 * xb + dot_bytes
        # one FP32 evaluation: xi read twice (xb bytes), vext once (8), dF written (vb), projections
        eval_bytes = eng.N * eng.no * (2 * xb + 8 + vb) + eng.N * 8 * eng.np_ * 8
        t_roof = (res_f.evaluations * eval_bytes + res_f.iterations * per_elem * nvec) / (HBM_PEAK_GBS * 1e9)
        line["seconds_to_band_roofline"] = {"t_roofline_s": t_roof, "t_fast_s": t_fast, "frac": t_roof / t_fast,
                                            "vector_bytes_per_element_per_iteration": per_elem,
                                            "minimiser_vectors": "fp64" if args.fp64_vectors else "fp32",
                                            "minimiser_iterate": "fp32" if xb == 4 else "fp64",
                                            "evaluation_algorithmic_GB": eval_bytes / 1e9}
        vname = "FP64" if args.fp64_vectors else ("FP32" if args.fp64_iterate else "FP32 (iterate too)")
        sname = "xi" if args.unscaled else "u = sqrt(w / mean w) xi"
        line.update({
            "value": t_fast, "unit": "s", "steps": res_f.evaluations,
            "ms_per_step": t_fast / max(res_f.evaluations, 1) * 1e3,
            "stop_rule": (f"both paths: the reference's (fold - f) <= {ref_tol:g} kJ/mol "
                          "(module_lbfgs_nocedal_mdft.f90:425-427), no knowledge of Delta G_ref"),
            "t_fast_s": t_fast, "t_plain_fp64_s": t_plain,
            "speedup_band_to_band": t_plain / t_fast,
            "dG_ref_kJmol": dg_ref, "dG_fast_confirmed_kJmol": dg_fast,
            "dG_fast_error_kJmol": dg_fast - dg_ref, "within_band": bool(abs(dg_fast - dg_ref) <= band),
            "dG_plain_error_kJmol": res_p.F - dg_ref, "plain_within_band": bool(abs(res_p.F - dg_ref) <= band),
            "plain_evaluations": res_p.evaluations, "plain_status": res_p.status,
            "fast_evaluations": res_f.evaluations, "fast_iterations": res_f.iterations,
            "fast_status": res_f.status,
            "t_plain_fp64_converged_s": t_conv, "speedup_vs_converged_plain": t_conv / t_fast,
            "converge_evaluations": res.evaluations, "converge_iterations": res.iterations,
            "converge_grad_ratio": res.grad_ratio, "converge_status": res.status,
            "oracle_stopped": {"t_plain_fp64_to_band_s": t_plain_band, "plain_evaluations": res_pb.evaluations,
                               "t_fast_to_band_s": t_fast_band, "fast_evaluations": res_fb.evaluations,
                               "dG_fast_error_kJmol": dg_fast_band - dg_ref,
                               "speedup_band_to_band": t_plain_band / t_fast_band,
                               "note": "both stopped at the first iterate within 0.05 kJ/mol of Delta G_ref, "
                                       "which a real run does not know"},
            "t_fp64_scaled_s": t_scaled64, "fp64_scaled_evaluations": res_s.evaluations,
            "dG_fp64_scaled_error_kJmol": res_s.F - dg_ref,
            "t_fp32_network_prediction_start_s": t_pure,
            "fp32_network_prediction_start_evaluations": res_pure.evaluations if res_pure is not None else None,
            "dG_fp32_network_prediction_error_kJmol": (dg_pure - dg_ref) if dg_pure is not None else None,
            "t_fp32_cold_start_s": t_cold32, "t_fp32_boltzmann_start_s": t_boltz32,
            "fp32_cold_start_evaluations": res_c32.evaluations, "fp32_boltzmann_start_evaluations": res_b32.evaluations,
            "dG_fp32_cold_error_kJmol": dg_c32 - dg_ref, "dG_fp32_boltzmann_error_kJmol": dg_b32 - dg_ref,
            "fast_path": (f"calibrated warm start ({info.get('start')}; inside the timed region) + FP32 "
                          f"evaluations + L-BFGS (FP64 scalars, {vname} vectors, {sname} variables, "
                          f"{'unit' if (args.unit_first_step or args.unscaled) else 'bulk Newton'} first step), "
                          "FP64 confirmation")
                         if net is not None else
                         f"FP32 evaluations + L-BFGS (FP64 scalars, {vname} vectors, {sname} variables) from the "
                         "reference cold start, FP64 confirmation (no trained warm-start weights found)",
            "dtype": f"f32 evaluations (f32 transforms and per-orientation terms, f64 sums), L-BFGS with FP64 "
                     f"scalars and {vname} vectors, f64 confirmation; the plain path all f64",
            "F_parts": [float(v) for v in res.parts],
            "trace_fast": [[float(v) for v in row] for row in res_f.trace],
            "trace_plain": [[float(v) for v in row] for row in res_p.trace],
            "warm_start": info, "lbfgs_memory": args.memory, "lbfgs_memory_fast": fast_m,
        })
        if not args.no_cpu_baseline:
            line["cpu_baseline"] = cpu_baseline(w, res_p.evaluations)
    if args.mode == "eval" and not args.no_cpu_baseline:
        cb = cpu_baseline(w, 1)
        cb["value"] = cb["evaluations_per_s"]
        cb["unit"] = "evaluations/s"
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

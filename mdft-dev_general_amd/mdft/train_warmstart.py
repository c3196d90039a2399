"""Trains the warm-start network for the 20-atom-solute workload and writes
mdft/data/warmstart_solute20.npz.

Training data: seeded solutes of 5-30 atoms (positions in a 5 A sphere, min
separation 1.2 A, sigma ~ U[2.5, 4.0] A, eps ~ U[0.2, 1.0] kJ/mol, charges
N(0, 0.4) shifted to zero net charge), each solved by the plain method (reference
cold start, FP32 evaluations, L-BFGS to |dF|/|dF0| < 1e-5) on a 20 A box at
64^3 -- the grid spacing of the 128^3 / 40 A workload (0.3125 A) -- with mmax = 5
and the same c(k) table.  Inputs are the projections of the vext feature
field (each channel scaled to unit rms over the first 16 solutes), targets the
projections of rho/rho0 - 1 of the converged density.

Two phases: Adam on the mean-squared error of the density projections (lr 1e-3,
halved every 10 epochs), then Adam on the free energy of the start itself,
F(guess(net(features))), whose gradient reaches the network through
mdft_ws_start_gradient (lr 2e-4, halved every 10 epochs).  The summary reports
what the network gains on held-out solutes: L-BFGS evaluations to the first
iterate within 0.05 kJ/mol of the solution, cold start against the network's
start.

    python -m mdft.train_warmstart --train 160 --val 16 --epochs 40 --energy-epochs 40 --weights PATH

Calibration of the start (run after training, on the target workload):

    python -m mdft.train_warmstart --calibrate-mix --weights PATH

solves 128^3 solutes of seeds 8000-8003 (not the benchmark's solute) from the
starts (1 - lam) rho_cold + lam rho_pred for lam in {0, 0.5 .. 1.0} (lam = 0: the
cold start itself) with the fast path's minimiser (FP32 evaluations and vectors,
scaled variables, Newton first step) and stores the lam of the fewest mean
evaluations to the 0.05 kJ/mol band with the weights (`start_mix`), with the
table it came from.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import warmstart as ws
from . import workloads


def calibrate_mix(path: str, seeds=(8000, 8001, 8002, 8003), lams=(0.0, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0), log=print):
    """Choose the net's start mixture on solutes of the 128^3 workload (see the module text)."""
    from .gpu import Engine
    from .minimize import minimize
    head = workloads.WORKLOADS["solute20_128"]
    eng = Engine(head.n, head.length, head.mmax, head.mrso, head.kT, head.n0, "fp32")
    eng.set_ck(head.ck_table())
    net = ws.WarmStartNet.load(path)
    xi, g = eng.alloc_field(), eng.alloc_field()
    evals = np.zeros((len(seeds), len(lams)), dtype=np.int64)
    for i, seed in enumerate(seeds):
        eng.compute_vext(workloads.make_solute(head, seed), head.solvent)
        ws.cold_start(eng, xi)
        ref = minimize(eng, xi, g=g, m=1, gtol=1e-6, maxeval=1500, fnoise=1e-6, fp32_vectors=True, scaled=True)
        for j, lam in enumerate(lams):
            ws.warm_start(eng, net, xi, start_mix=lam)
            r = minimize(eng, xi, g=g, m=1, gtol=0.0, maxeval=400, fnoise=1e-6, ftarget=ref.F, fband=0.05,
                         fp32_vectors=True, scaled=True, step0="newton")
            evals[i, j] = r.evaluations
        log(f"  solute {seed}: F*={ref.F:.3f}; evaluations to the band for lam {list(lams)}: {list(evals[i])}")
    best = float(lams[int(np.argmin(evals.mean(axis=0)))])
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    data.update(start_mix=np.asarray(best), mix_calibration_lams=np.asarray(lams),
                mix_calibration_seeds=np.asarray(seeds), mix_calibration_evaluations=evals)
    np.savez(path, **data)
    log(f"[train_warmstart] start_mix = {best} (mean evaluations {evals.mean(axis=0).round(2).tolist()})")
    return best, evals


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=192)
    ap.add_argument("--val", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=40, help="epochs on the density projections")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--energy-epochs", type=int, default=40,
                    help="epochs on F of the start itself, after the density epochs")
    ap.add_argument("--energy-lr", type=float, default=2e-4)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--box", type=float, default=20.0)
    ap.add_argument("--seed", type=int, default=5000)
    ap.add_argument("--weights", default=ws.DEFAULT_WEIGHTS)
    ap.add_argument("--log", default=None, help="JSON summary file")
    ap.add_argument("--gain-solutes", type=int, default=8, help="held-out solutes for the evaluation count")
    ap.add_argument("--calibrate-mix", action="store_true",
                    help="only calibrate the start mixture of the weights at --weights")
    args = ap.parse_args(argv)

    def say(s):
        print(s, flush=True)

    if args.calibrate_mix:
        calibrate_mix(args.weights, log=say)
        return 0

    head = workloads.WORKLOADS["solute20_128"]
    w = workloads.Workload("warmstart_train", 3, (args.n,) * 3, (args.box,) * 3, head.mmax, head.nk, head.dk)
    table = head.ck_table()
    t0 = time.perf_counter()
    say(f"[train_warmstart] solving {args.train + args.val} solutes on {w.n} / {w.length} A, mmax={w.mmax}")
    eng, scale, data = ws.make_training_set(w, args.train + args.val, args.seed, table, gtol=1e-5, log=say)
    t_data = time.perf_counter() - t0
    train, val = data[:args.train], data[args.train:]
    train_sols = [workloads.make_solute(w, args.seed + i) for i in range(args.train)]
    val_sols = [workloads.make_solute(w, args.seed + args.train + j) for j in range(len(val))]
    net = ws.WarmStartNet.for_engine(eng, seed=args.seed)
    net.input_scale = scale
    say(f"[train_warmstart] {net.n_params} parameters; training {args.epochs} epochs on {len(train)} solutes")
    t1 = time.perf_counter()
    hist, vhist = [], []
    lr = args.lr
    for ep in range(args.epochs):
        if ep and ep % 10 == 0:
            lr *= 0.5
        hist += ws.train(net, w.n, train, 1, lr=lr, seed=args.seed + ep)
        # validation loss: loss_grad without the Adam step
        vl = float(np.mean([net.loss_grad(w.n, f, t) for f, t, _, _ in val]))
        vhist.append(vl)
        say(f"  epoch {ep}: train loss {hist[-1]:.5f}, validation loss {vl:.5f}, lr {lr:.1e}")
    gap_before = ws.start_gap(net, eng, val_sols, w.solvent, [d[0] for d in val])
    say(f"[train_warmstart] validation F(start) - F(cold): {np.mean(gap_before):.2f} kJ/mol after the density epochs")
    ehist = []
    lr_e = args.energy_lr
    for ep in range(args.energy_epochs):
        if ep and ep % 10 == 0:
            lr_e *= 0.5
        ehist += ws.train_energy(net, eng, train_sols, w.solvent, [d[0] for d in train], 1, lr=lr_e,
                                 seed=args.seed + 1000 + ep, log=say)
    gap_after = ws.start_gap(net, eng, val_sols, w.solvent, [d[0] for d in val])
    say(f"[train_warmstart] validation F(start) - F(cold): {np.mean(gap_after):.2f} kJ/mol after the energy epochs")
    t_train = time.perf_counter() - t1
    # zero-prediction baseline on validation (the cold start)
    zero = []
    for f, t, _, _ in val:
        tt = t.download(np.float32)
        zero.append(float(np.mean(tt * tt)))
    net.save(args.weights, train_loss=np.array(hist), val_loss=np.array(vhist), energy_gap=np.array(ehist))
    # what the network gains on held-out solutes: L-BFGS evaluations (FP32) to the
    # first iterate within 0.05 kJ/mol of the solution's F, cold start against
    # the network's start as predicted (no start selection)
    from .minimize import minimize
    xi, g = eng.alloc_field(), eng.alloc_field()
    gain = []
    for j, (_, _, _, fstar) in enumerate(val[:args.gain_solutes]):
        eng.compute_vext(val_sols[j], w.solvent)
        ws.cold_start(eng, xi)
        rc = minimize(eng, xi, g=g, m=5, gtol=0.0, maxeval=600, fnoise=1e-6, ftarget=fstar, fband=0.05)
        ws.warm_start(eng, net, xi)
        rw = minimize(eng, xi, g=g, m=5, gtol=0.0, maxeval=600, fnoise=1e-6, ftarget=fstar, fband=0.05)
        gain.append((rc.evaluations, rw.evaluations))
        say(f"  held-out solute {j}: cold {rc.evaluations} evaluations, network start {rw.evaluations}")
    summary = {"solutes_train": len(train), "solutes_val": len(val), "grid": list(w.n), "box_A": list(w.length),
               "mmax": w.mmax, "params": net.n_params, "epochs": args.epochs, "energy_epochs": args.energy_epochs,
               "data_s": t_data, "train_s": t_train,
               "train_loss_first": hist[0] if hist else None, "train_loss_last": hist[-1] if hist else None,
               "val_loss_last": vhist[-1] if vhist else None,
               "val_loss_zero_prediction": float(np.mean(zero)),
               "val_start_gap_after_density_epochs": float(np.mean(gap_before)),
               "val_start_gap_after_energy_epochs": float(np.mean(gap_after)),
               "train_evaluations_mean": float(np.mean([d[2] for d in train])),
               "heldout_evaluations_to_band_cold": [c for c, _ in gain],
               "heldout_evaluations_to_band_warm": [w_ for _, w_ in gain]}
    say(json.dumps(summary))
    if args.log:
        with open(args.log, "w") as fh:
            json.dump(summary, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())

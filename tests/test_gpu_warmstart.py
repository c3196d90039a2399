"""GPU tests of the learned warm start: the HIP network kernels against the
FP64 CPU restatement (oracle/wsnet_oracle.py), the coupling kernels
(features, gamma field, guess) against numpy / the MDFT oracle, and what a
trained network gains over the cold start on held-out solutes.

Tolerances: network forward / gradients / Adam (FP32 kernels) 2e-5 of the
largest magnitude; gamma field (FP64 path) 1e-9 relative to its largest
magnitude; projections and guess 1e-6."""
import sys

import numpy as np
import pytest

from mdft import ck, workloads
from oracle import wsnet_oracle as WO

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(n, length, mmax, precision="fp64"):
    from mdft.gpu import Engine
    return Engine(n, length, mmax, 2, 2.4777, 0.0333, precision)


def test_network_kernels_match_oracle():
    from mdft.gpu import DeviceBuffer
    from mdft.warmstart import WarmStartNet
    say("[wsnet] forward, loss gradient, input gradient and Adam vs the FP64 restatement")
    n = (10, 9, 8)
    N = n[0] * n[1] * n[2]
    chans, ks, dils = (5, 20, 12, 9, 40, 2), (1, 3, 3, 1, 1), (1, 1, 2, 1, 1)
    net = WarmStartNet(chans, ks, dils, seed=5)
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, net.n_params).astype(np.float32)
    net.set_params(p)
    eng = _engine((4, 4, 4), (8.0, 8.0, 8.0), 1)   # only for device buffers
    x = rng.normal(size=(5, n[2], n[1], n[0])).astype(np.float32)
    t = rng.normal(size=(2, n[2], n[1], n[0])).astype(np.float32)
    mask = (rng.uniform(size=(n[2], n[1], n[0])) > 0.25).astype(np.float32)
    dx, dt, dm = DeviceBuffer(eng, x.nbytes), DeviceBuffer(eng, t.nbytes), DeviceBuffer(eng, mask.nbytes)
    dy, dg = DeviceBuffer(eng, 4 * 2 * N), DeviceBuffer(eng, x.nbytes)
    dx.upload(x)
    dt.upload(t)
    dm.upload(mask)
    layers = net.layers(p)
    ref = WO.Net([w for w, _ in layers], [b for _, b in layers], dils)
    net.forward(n, dx, dy)
    y = dy.download(np.float32).reshape(t.shape)
    yr = ref.forward(x.astype(np.float64))[-1]
    err = np.abs(y - yr).max() / np.abs(yr).max()
    say(f"  forward max rel err {err:.2e}")
    assert err < 2e-5
    loss = net.loss_grad(n, dx, dt, dm, dg)
    lr_, gW, gb, gx = ref.loss_and_grads(x.astype(np.float64), t.astype(np.float64), mask[None].astype(np.float64))
    assert abs(loss - lr_) < 2e-5 * abs(lr_)
    g = net.get_grads()
    gref = np.concatenate([np.concatenate([a.ravel(), b]) for a, b in zip(gW, gb)])
    err = np.abs(g - gref).max() / np.abs(gref).max()
    say(f"  loss {loss:.6f} vs {lr_:.6f}; parameter gradient max rel err {err:.2e}")
    assert err < 2e-5
    gxg = dg.download(np.float32).reshape(x.shape)
    err = np.abs(gxg - gx).max() / np.abs(gx).max()
    say(f"  input gradient max rel err {err:.2e}")
    assert err < 2e-5
    # two Adam steps
    pr = p.astype(np.float64).copy()
    m1, m2 = np.zeros_like(pr), np.zeros_like(pr)
    WO.adam_step(pr, g.astype(np.float64), m1, m2, 1, 1e-3)
    net.adam(1e-3)
    net.loss_grad(n, dx, dt, dm)
    g2 = net.get_grads()
    WO.adam_step(pr, g2.astype(np.float64), m1, m2, 2, 1e-3)
    net.adam(1e-3)
    err = np.abs(net.get_params() - pr).max()
    say(f"  Adam: max abs param err after 2 steps {err:.2e}")
    assert err < 1e-6


def test_gamma_field_features_and_guess(oracle_mod):
    from mdft import warmstart as ws
    from mdft.gpu import DeviceBuffer
    from mdft.grid import angular_grid
    n, L, mmax = (9, 8, 10), (9.0, 8.0, 10.0), 3
    say(f"[ws coupling] gamma field, gamma projections, features, guess on {n} mmax={mmax}")
    t = ck.synthetic_ck(mmax, 256, 0.05, seed=7, n0=0.0333)
    eng = _engine(n, L, mmax)
    eng.set_ck(t)
    o = oracle_mod.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    rng = np.random.default_rng(3)
    N, no, npj = eng.N, eng.no, eng.np_
    kT = 2.4777
    vext = rng.uniform(-8.0, 12.0, N * no)
    eng.set_vext(vext)
    xi = 1.0 + 0.3 * rng.uniform(-1, 1, N * no)   # no zeros: gamma is read off dF_exc / xi
    bx = eng.alloc_field()
    bx.upload(xi)
    tgt, gfield = ws.gamma(eng, bx, want_field=True)
    gam = gfield.download()
    _, dexc = o.excess(xi)
    g = angular_grid(mmax, 2)
    w = np.tile(g["w"], N)
    gam_ref = -dexc / (2.0 * o.rho0 * xi * w * kT)
    err = np.abs(gam - gam_ref).max() / np.abs(gam_ref).max()
    say(f"  gamma field vs -dF_exc / (2 rho0 xi w kT): max rel err {err:.2e}")
    assert err < 1e-9
    # projections of gamma = (2 (4 pi^2) / (mrso n0)) * the stage-4 projections
    scale = 2.0 * 4.0 * np.pi ** 2 / 2 / 0.0333
    P4 = o.excess_stage(xi, 4)                       # [N][np] complex
    ref = np.empty((2 * npj, N))
    ref[0::2], ref[1::2] = scale * P4.real.T, scale * P4.imag.T
    got = tgt.download(np.float32).reshape(2 * npj, N)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(f"  gamma projections: max rel err {err:.2e}")
    assert err < 1e-6
    # rebuilding gamma from its projections gives the gamma field back
    rebuilt = np.concatenate([o.proj2angl(P4[i] * scale) for i in range(N)])
    assert np.abs(rebuilt - gam).max() < 1e-9 * np.abs(gam).max()
    # features: projections of clip(-beta v / 2, -4, 4) / 4, then per-channel scaled
    f = ws.features(eng).download(np.float32).reshape(2 * npj, N)
    fo = (0.25 * np.clip(-0.5 * vext / kT, -4.0, 4.0)).reshape(N, no)
    fp = np.stack([o.angl2proj(fo[i]) for i in range(N)])
    err = max(np.abs(f[0::2] - fp.real.T).max(), np.abs(f[1::2] - fp.imag.T).max()) / np.abs(fp).max()
    say(f"  features: max rel err {err:.2e}")
    assert err < 1e-6
    chs = rng.uniform(0.5, 3.0, 2 * npj).astype(np.float32)
    fs = ws.features(eng, chs).download(np.float32).reshape(2 * npj, N)
    assert np.abs(fs - chs[:, None] * f).max() <= 1e-6 * np.abs(fs).max()
    # guess from predicted projections: xi = exp(clip(-beta v + proj2angl(pred), -200, 7) / 2)
    pred = (rng.normal(0, 0.3, (2 * npj, N)) * (np.abs(ref).max() / 10)).astype(np.float32)
    bp = DeviceBuffer(eng, pred.nbytes)
    bp.upload(pred)
    ws.guess(eng, bp, bx, ws.WS_GAMMA)
    pc = pred[0::2].T.astype(np.float64) + 1j * pred[1::2].T.astype(np.float64)
    gp = np.concatenate([o.proj2angl(pc[i]) for i in range(N)])
    lg = np.clip(-vext / kT + gp, -200.0, 7.0)
    err = np.abs(bx.download() - np.exp(0.5 * lg)).max() / np.exp(3.5)
    say(f"  gamma guess: max err / e^3.5 {err:.2e}")
    assert err < 1e-6
    # density target: projections of xi^2 - 1
    vext[::11] = 700.0
    eng.set_vext(vext)
    c = workloads.guess_xi(vext, kT)
    bx.upload(xi)
    dt = ws.density(eng, bx).download(np.float32).reshape(2 * npj, N)
    P1 = np.stack([o.angl2proj(r) for r in (xi * xi - 1.0).reshape(N, no)])
    err = max(np.abs(dt[0::2] - P1.real.T).max(), np.abs(dt[1::2] - P1.imag.T).max()) / np.abs(P1).max()
    say(f"  density target projections: max rel err {err:.2e}")
    assert err < 1e-6
    # density guess: xi = 0 where c = 0, else sqrt(max(1 + proj2angl(pred), 0.01))
    pred2 = (rng.normal(0, 0.5, (2 * npj, N))).astype(np.float32)
    bp.upload(pred2)
    ws.guess(eng, bp, bx, ws.WS_DENSITY)
    pc = pred2[0::2].T.astype(np.float64) + 1j * pred2[1::2].T.astype(np.float64)
    rp = np.where(c > 0, np.maximum(1.0 + np.concatenate([o.proj2angl(pc[i]) for i in range(N)]), 0.01), 0.0)
    assert np.abs(bx.download() - np.sqrt(rp)).max() < 1e-9
    # density mixture
    by = eng.alloc_field()
    by.upload(xi)
    ws.mix(eng, 0.3, by, bx, by)
    assert np.allclose(by.download(), np.sqrt(0.7 * xi ** 2 + 0.3 * rp), rtol=1e-12, atol=1e-12)
    # device cold start = the host restatement of the reference's guess
    vext[::7] = 700.0
    eng.set_vext(vext)
    ws.cold_start(eng, bx)
    assert np.array_equal(bx.download(), workloads.guess_xi(vext, kT))
    # Boltzmann start: xi = exp(-max(beta vext, -3) / 2) (the device's exp against numpy's: a few ulp)
    vext[::13] = -40.0                                 # strongly attractive: capped at rho/rho0 = e^3
    eng.set_vext(vext)
    ws.boltzmann_start(eng, bx)
    ref = np.exp(-0.5 * np.maximum(vext * (1.0 / kT), -3.0))
    err = np.abs(bx.download() - ref).max()
    say(f"  Boltzmann start: max err {err:.2e}")
    assert err < 1e-14 and np.all(bx.download() <= np.exp(1.5) * (1 + 1e-15))
    from mdft.gpu import MdftError
    with pytest.raises(MdftError):
        ws.guess(eng, bp, bx, ws.WS_BOLTZMANN)
    with pytest.raises(MdftError):
        ws.guess(eng, None, bx, 7)


def test_euler_lagrange_at_the_minimum():
    """At the minimum, ln(rho/rho0) = -beta vext + gamma wherever rho > 0."""
    from mdft import warmstart as ws
    from mdft.minimize import minimize
    w = workloads.Workload("el", 0, (15, 15, 15), (12.0, 12.0, 12.0), 2, 256)
    say("[ws] Euler-Lagrange identity at a converged minimum, 15^3 mmax=2")
    eng = _engine(w.n, w.length, w.mmax)
    eng.set_ck(ck.synthetic_ck(2, 256, 0.05, seed=21, n0=w.n0))
    sol = workloads.Solute(np.array([[6.0, 6.0, 6.0]]), np.array([0.5]), np.array([3.0]), np.array([0.8]))
    eng.compute_vext(sol, w.solvent)
    vext = eng.get_vext()
    x = eng.alloc_field()
    x.upload(workloads.guess_xi(vext, w.kT))
    res = minimize(eng, x, m=5, gtol=1e-9, maxeval=2000)
    xs = x.download()
    _, gf = ws.gamma(eng, x, want_field=True)
    gam = gf.download()
    rho = xs * xs
    sel = rho > 1e-3
    lhs = np.log(rho[sel])
    rhs = -vext[sel] / w.kT + gam[sel]
    err = np.abs(lhs - rhs).max()
    say(f"  |g|/|g0|={res.grad_ratio:.1e}; max |ln rho - (-beta v + gamma)| = {err:.2e} over {sel.sum()} points")
    assert err < 1e-3


def test_backward_from_output_gradient_matches_loss_grad():
    """mdft_wsnet_backward with the MSE output gradient gives loss_grad's
    parameter and input gradients."""
    from mdft.gpu import DeviceBuffer
    from mdft.warmstart import WarmStartNet
    say("[wsnet] backward from a given output gradient")
    n = (9, 8, 7)
    N = n[0] * n[1] * n[2]
    chans, ks, dils = (4, 16, 12, 6), (1, 3, 1), (1, 2, 1)
    net = WarmStartNet(chans, ks, dils, seed=9)
    eng = _engine((4, 4, 4), (8.0, 8.0, 8.0), 1)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(4, N)).astype(np.float32)
    t = rng.normal(size=(6, N)).astype(np.float32)
    dx, dt, dy = DeviceBuffer(eng, x.nbytes), DeviceBuffer(eng, t.nbytes), DeviceBuffer(eng, t.nbytes)
    g1, g2, dgo = DeviceBuffer(eng, x.nbytes), DeviceBuffer(eng, x.nbytes), DeviceBuffer(eng, t.nbytes)
    dx.upload(x)
    dt.upload(t)
    net.loss_grad(n, dx, dt, None, g1)
    ref = net.get_grads()
    net.forward(n, dx, dy)
    y = dy.download(np.float32).reshape(t.shape)
    dgo.upload((2.0 * (y - t) / t.size).astype(np.float32))
    net.backward(n, dx, dgo, g2)
    got = net.get_grads()
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    a, b = g1.download(np.float32), g2.download(np.float32)
    assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


def test_start_gradient_by_finite_differences():
    """dF/d(pred) from mdft_ws_start_gradient against central differences of F
    at guess(pred) (FP64 evaluations), including an imaginary channel of an
    (m, 0, 0) projection (no effect on a real field: gradient 0) and a point
    whose orientations are all below the density floor (gradient 0)."""
    from mdft import warmstart as ws
    from mdft.gpu import DeviceBuffer
    n, L, mmax = (6, 5, 8), (7.0, 6.0, 9.0), 3
    say(f"[ws] start gradient vs finite differences on {n} mmax={mmax}")
    eng = _engine(n, L, mmax)
    eng.set_ck(ck.synthetic_ck(mmax, 256, 0.05, seed=12, n0=0.0333))
    rng = np.random.default_rng(8)
    N, no, npj = eng.N, eng.no, eng.np_
    vext = rng.uniform(-3.0, 6.0, N * no)
    vext[5 * no:6 * no] = 700.0                       # point 5: excluded by the cold start
    eng.set_vext(vext)
    pred = (0.03 * rng.normal(size=(2 * npj, N))).astype(np.float32)
    pred[0, 9] = -6.0                                 # point 9: below the floor everywhere
    bp, bx, bg = DeviceBuffer(eng, pred.nbytes), eng.alloc_field(), eng.alloc_field()
    gp = DeviceBuffer(eng, pred.nbytes)

    def F(pv):
        bp.upload(pv)
        ws.guess(eng, bp, bx, ws.WS_DENSITY)
        return float(np.sum(eng.energy_and_gradient(bx, None)[0]))

    F(pred)
    eng.energy_and_gradient(bx, bg)
    ws.start_gradient(eng, bx, bg, gp)
    g = gp.download(np.float32).reshape(2 * npj, N)
    gmax = np.abs(g).max()
    probes = [(int(rng.integers(0, 2 * npj)), int(rng.integers(0, N))) for _ in range(10)]
    probes += [(0, 3), (2, 11), (5, 4), (1, 7), (0, 9), (3, 5)]    # ch 5: Im of (1, 0, 0); 1: Im of (0, 0, 0)
    h = 2e-3
    for ch, pt in probes:
        pp, pm = pred.copy(), pred.copy()
        pp[ch, pt] += h
        pm[ch, pt] -= h
        fd = (F(pp) - F(pm)) / (2 * h)
        say(f"  channel {ch} point {pt}: start gradient {g[ch, pt]:+.6e}, finite difference {fd:+.6e}")
        assert abs(fd - g[ch, pt]) <= 2e-4 * gmax
    assert g[5, 4] == 0.0 and g[1, 7] == 0.0 and g[0, 9] == 0.0 and g[3, 5] == 0.0


def _tiny_solute(rng, center):
    nat = int(rng.integers(2, 6))
    xyz = center + rng.uniform(-1.5, 1.5, (nat, 3))
    q = rng.normal(0.0, 0.5, nat)
    q -= q.mean()
    return workloads.Solute(xyz, q, rng.uniform(2.5, 3.5, nat), rng.uniform(0.2, 1.0, nat))


def test_training_lowers_density_loss_and_start_energy():
    """Both training phases on the device, on 16 solved solutes (32^3, 12 A box,
    mmax=3): the density epochs halve the loss, the epochs on F of the start
    bring F(start) below F(cold start), and so does the network's start on 4
    held-out solutes.  The evaluations to the 0.05 kJ/mol band are reported
    (16 solutes are too few for a network that saves evaluations; the gain of
    the trained network is the next test)."""
    from mdft import warmstart as ws
    from mdft.minimize import minimize
    w = workloads.Workload("ws_test", 3, (32, 32, 32), (12.0, 12.0, 12.0), 3, 512)
    say("[ws] two-phase training on 16 solutes, 4 held-out (32^3, mmax=3)")
    table = ck.synthetic_ck(3, 512, 0.05, seed=31, n0=w.n0)
    eng = _engine(w.n, w.length, w.mmax)
    eng.set_ck(table)
    rng = np.random.default_rng(11)
    sols = [_tiny_solute(rng, np.array(w.length) / 2) for _ in range(20)]
    xi, gb = eng.alloc_field(), eng.alloc_field()
    scale = ws.input_scale(eng, sols[:16], w.solvent)
    data = []
    for s in sols[:16]:
        eng.compute_vext(s, w.solvent)
        f = ws.features(eng, scale)
        ws.cold_start(eng, xi)
        minimize(eng, xi, g=gb, m=5, gtol=1e-5, maxeval=800)
        data.append((f, ws.density(eng, xi)))
    net = ws.WarmStartNet.for_engine(eng, seed=1)
    net.input_scale = scale
    hist = ws.train(net, w.n, data, epochs=40, lr=1e-3, seed=0)
    say(f"  density loss: first epoch {hist[0]:.4f}, last {hist[-1]:.4f}")
    assert hist[-1] < 0.5 * hist[0]
    ehist = ws.train_energy(net, eng, sols[:16], w.solvent, [d[0] for d in data], 40, lr=2e-4, seed=0)
    say(f"  mean F(start) - F(cold) on the training solutes: first energy epoch {ehist[0]:.2f}, last {ehist[-1]:.2f}")
    assert ehist[-1] < min(0.0, ehist[0])
    band = 0.05
    for s in sols[16:]:
        eng.compute_vext(s, w.solvent)
        ws.cold_start(eng, xi)
        ref = minimize(eng, xi, g=gb, m=5, gtol=1e-8, maxeval=2000)
        ws.cold_start(eng, xi)
        f_cold = ws._F(eng, xi)
        rc = minimize(eng, xi, g=gb, m=5, gtol=0.0, maxeval=800, ftarget=ref.F, fband=band)
        ws.warm_start(eng, net, xi)
        f_warm = ws._F(eng, xi)
        rw = minimize(eng, xi, g=gb, m=5, gtol=0.0, maxeval=800, ftarget=ref.F, fband=band)
        say(f"  F*={ref.F:.4f}: cold start F-F*={f_cold - ref.F:.3f}, {rc.evaluations} evaluations; "
            f"network start F-F*={f_warm - ref.F:.3f}, {rw.evaluations} evaluations")
        assert abs(rc.F - ref.F) <= band and abs(rw.F - ref.F) <= band   # both reached the band
        assert f_warm < f_cold


def test_trained_network_reduces_evaluations():
    """What the committed network (mdft/data/warmstart_solute20.npz, trained by
    mdft.train_warmstart on 160 solutes) gains over the method without it, on 8
    solutes outside its training and validation seeds, on its workload (64^3,
    20 A box, mmax=5, FP32 evaluations, the c(k) of the 128^3 workload): L-BFGS
    evaluations to the first iterate within 0.05 kJ/mol of the minimum (the
    stopping rule of the solvation free energy benchmark), cold start against
    the network's start.  The minimum F* comes from a tight run from the cold
    start.  (A relative-gradient rule such as |dF|/|dF0| < 1e-6 would be unfit
    here: a better start has a smaller dF0, so it is asked for a smaller
    absolute gradient.)"""
    from mdft import warmstart as ws
    from mdft.gpu import Engine
    from mdft.minimize import minimize
    net = ws.WarmStartNet.load(ws.DEFAULT_WEIGHTS)
    head = workloads.WORKLOADS["solute20_128"]
    w = workloads.Workload("warmstart_train", 3, (64,) * 3, (20.0,) * 3, head.mmax, head.nk, head.dk)
    say("[ws] committed network on 8 held-out solutes (64^3, 20 A, mmax=5, FP32)")
    eng = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, "fp32")
    eng.set_ck(head.ck_table())
    xi, gb = eng.alloc_field(), eng.alloc_field()
    band = 0.05
    cold, warm = [], []
    for seed in range(7000, 7008):
        sol = workloads.make_solute(w, seed)
        eng.compute_vext(sol, w.solvent)
        ws.cold_start(eng, xi)
        ref = minimize(eng, xi, g=gb, m=5, gtol=1e-6, maxeval=1000, fnoise=1e-6)
        ws.cold_start(eng, xi)
        f_cold = ws._F(eng, xi)
        rc = minimize(eng, xi, g=gb, m=5, gtol=0.0, maxeval=600, fnoise=1e-6, ftarget=ref.F, fband=band)
        ws.warm_start(eng, net, xi, start_mix=1.0)          # the prediction itself
        f_warm = ws._F(eng, xi)
        rw = minimize(eng, xi, g=gb, m=5, gtol=0.0, maxeval=600, fnoise=1e-6, ftarget=ref.F, fband=band)
        say(f"  solute {seed} ({len(sol.q)} atoms): F*={ref.F:.3f}; cold start F-F*={f_cold - ref.F:.1f}, "
            f"{rc.evaluations} evaluations; network start F-F*={f_warm - ref.F:.1f}, {rw.evaluations} evaluations")
        assert abs(rc.F - ref.F) <= band and abs(rw.F - ref.F) <= band
        assert f_warm < f_cold
        cold.append(rc.evaluations)
        warm.append(rw.evaluations)
    say(f"  mean evaluations to the band: cold start {np.mean(cold):.1f}, network start {np.mean(warm):.1f}")
    assert np.mean(warm) < np.mean(cold)

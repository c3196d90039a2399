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
    # features: projections of exp(-max(beta v, -8) / 2) - 1
    f = ws.features(eng).download(np.float32).reshape(2 * npj, N)
    fo = (np.exp(-0.5 * np.maximum(vext / kT, -8.0)) - 1.0).reshape(N, no)
    fp = np.stack([o.angl2proj(fo[i]) for i in range(N)])
    err = max(np.abs(f[0::2] - fp.real.T).max(), np.abs(f[1::2] - fp.imag.T).max()) / np.abs(fp).max()
    say(f"  features: max rel err {err:.2e}")
    assert err < 1e-6
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
    # density residual target: projections of xi^2 - c, c the cold start (0 where vext is excluding)
    vext[::11] = 700.0
    eng.set_vext(vext)
    c = workloads.guess_xi(vext, kT)
    bx.upload(xi)
    dt = ws.density(eng, bx).download(np.float32).reshape(2 * npj, N)
    P1 = np.stack([o.angl2proj(r) for r in (xi * xi - c).reshape(N, no)])
    err = max(np.abs(dt[0::2] - P1.real.T).max(), np.abs(dt[1::2] - P1.imag.T).max()) / np.abs(P1).max()
    say(f"  density residual projections: max rel err {err:.2e}")
    assert err < 1e-6
    # density guess: xi = 0 where c = 0, else sqrt(max(1 + proj2angl(pred), 0))
    pred2 = (rng.normal(0, 0.5, (2 * npj, N))).astype(np.float32)
    bp.upload(pred2)
    ws.guess(eng, bp, bx, ws.WS_DENSITY)
    pc = pred2[0::2].T.astype(np.float64) + 1j * pred2[1::2].T.astype(np.float64)
    rp = np.where(c > 0, np.maximum(1.0 + np.concatenate([o.proj2angl(pc[i]) for i in range(N)]), 0.0), 0.0)
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


def _tiny_solute(rng, center):
    nat = int(rng.integers(2, 6))
    xyz = center + rng.uniform(-1.5, 1.5, (nat, 3))
    q = rng.normal(0.0, 0.5, nat)
    q -= q.mean()
    return workloads.Solute(xyz, q, rng.uniform(2.5, 3.5, nat), rng.uniform(0.2, 1.0, nat))


def test_trained_network_reduces_iterations():
    """Train on 16 solved solutes (32^3, 12 A box, mmax=3), then compare L-BFGS
    evaluations to |dF|/|dF0| < 1e-6 from the cold start and from the
    network's guess (the F evaluations of the start selection counted) on 4
    held-out solutes; both reach the same F."""
    from mdft import warmstart as ws
    from mdft.minimize import minimize
    w = workloads.Workload("ws_test", 3, (32, 32, 32), (12.0, 12.0, 12.0), 3, 512)
    say("[ws] train on 16 solutes, test on 4 held-out (32^3, mmax=3)")
    table = ck.synthetic_ck(3, 512, 0.05, seed=31, n0=w.n0)
    eng = _engine(w.n, w.length, w.mmax)
    eng.set_ck(table)
    rng = np.random.default_rng(11)
    sols = [_tiny_solute(rng, np.array(w.length) / 2) for _ in range(20)]
    xi, gb, tmp = eng.alloc_field(), eng.alloc_field(), eng.alloc_field()
    data = []
    for s in sols[:16]:
        eng.compute_vext(s, w.solvent)
        f = ws.features(eng)
        ws.cold_start(eng, xi)
        minimize(eng, xi, g=gb, m=5, gtol=1e-5, maxeval=800)
        data.append((f, ws.density(eng, xi)))
    net = ws.WarmStartNet.for_engine(eng, seed=1)
    hist = ws.train(net, w.n, data, epochs=60, lr=1e-3, seed=0)
    say(f"  loss: first epoch {hist[0]:.4f}, last {hist[-1]:.4f}")
    assert hist[-1] < 0.5 * hist[0]
    cold, warm = [], []
    for s in sols[16:]:
        eng.compute_vext(s, w.solvent)
        ws.cold_start(eng, xi)
        rc = minimize(eng, xi, g=gb, m=5, gtol=1e-6, maxeval=800)
        info = ws.warm_start(eng, net, xi, fallback=tmp)
        rw = minimize(eng, xi, g=gb, m=5, gtol=1e-6, maxeval=800)
        say(f"  cold: {rc.evaluations} evaluations F={rc.F:.5f}; warm ({info['start']}, F at lam: {info['F']}): "
            f"{rw.evaluations}+{info['evaluations']} evaluations F={rw.F:.5f}")
        assert abs(rc.F - rw.F) < 0.05   # same minimum (kJ/mol)
        cold.append(rc.evaluations)
        warm.append(rw.evaluations + info["evaluations"])
    say(f"  mean evaluations cold {np.mean(cold):.1f}, warm {np.mean(warm):.1f}")
    assert np.mean(warm) < np.mean(cold)

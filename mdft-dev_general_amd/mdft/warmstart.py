"""Learned warm start of the minimisation (ctypes binding of include/mdft_wsnet.h).

The reference has no learned component; the minimiser it drives
(src/energy_minimization.f90) starts from the cold guess of
src/module_density.f90:267-284.  Here a periodic 3D convolutional network
(hand-written HIP forward/backward/Adam kernels, FP32) maps the angular
projections of a feature field of the external potential to the angular
projections of rho/rho0 - 1 of the converged density, and the minimiser
starts from xi = sqrt(max(1 + prediction, 0.01)), 0 where the reference's
cold start excludes the solvent.  A zero prediction is the cold start.  The
floor matters: at xi = 0 both gradient terms vanish, so a point predicted
empty would stay empty.  (The target is taken against 1 rather than against
the cold start: the cold start's exclusion boundary, beta vext = 36, is
invisible in the features, which saturate beyond beta vext = 8, and a target
with a jump there cannot be learned.)  With a scratch field the start is
the better (by F) of the prediction and the cold start, or of density
mixtures (1 - lam) rho_cold + lam rho_pred.  The minimisation afterwards is
the plain one: the network changes the number of iterations, not the answer,
and the cold start stays the fallback.

(A prediction of the OZ field gamma with the Euler-Lagrange start
xi = exp((-beta vext + gamma)/2) is also available, kind MDFT_WS_GAMMA; its
errors are exponentiated at high-density sites, which makes it the weaker
choice for a start.)

Training data are converged solutions of the plain method on small boxes with
the grid spacing of the target workload (the network is translation invariant
and sees about 10 A around each point).
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from .gpu import DeviceBuffer, Engine, load_library

HIDDEN = 32


def architecture(np_: int, hidden: int = HIDDEN):
    """(channels, kernel sizes, dilations): pointwise 2np -> hidden, four 3x3x3
    layers with dilations 1, 2, 4, 8 (receptive field 31 voxels), pointwise
    hidden -> 2np."""
    c = 2 * np_
    return (c, hidden, hidden, hidden, hidden, hidden, c), (1, 3, 3, 3, 3, 1), (1, 1, 2, 4, 8, 1)
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA, "warmstart_solute20.npz")

WSNET_EXPORTS = (
    "mdft_wsnet_create", "mdft_wsnet_destroy", "mdft_wsnet_last_error", "mdft_wsnet_num_params",
    "mdft_wsnet_get_params", "mdft_wsnet_set_params", "mdft_wsnet_get_grads", "mdft_wsnet_forward",
    "mdft_wsnet_loss_grad", "mdft_wsnet_adam", "mdft_ws_features", "mdft_ws_gamma", "mdft_ws_density",
    "mdft_ws_guess", "mdft_ws_mix", "mdft_wsnet_backward", "mdft_ws_start_gradient",
)
WS_GAMMA, WS_DENSITY, WS_BOLTZMANN = 0, 1, 2

_bound = False


def _lib():
    global _bound
    lib = load_library()
    if not _bound:
        vp, ip = C.c_void_p, C.POINTER(C.c_int)
        lib.mdft_wsnet_create.argtypes = [C.c_int, ip, ip, ip, C.c_ulonglong, C.c_int, C.POINTER(vp)]
        lib.mdft_wsnet_destroy.argtypes = [vp]
        lib.mdft_wsnet_destroy.restype = None
        lib.mdft_wsnet_last_error.argtypes = [vp]
        lib.mdft_wsnet_last_error.restype = C.c_char_p
        lib.mdft_wsnet_num_params.argtypes = [vp]
        lib.mdft_wsnet_num_params.restype = C.c_longlong
        lib.mdft_wsnet_get_params.argtypes = [vp, vp]
        lib.mdft_wsnet_set_params.argtypes = [vp, vp]
        lib.mdft_wsnet_get_grads.argtypes = [vp, vp]
        lib.mdft_wsnet_forward.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp]
        lib.mdft_wsnet_loss_grad.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, C.POINTER(C.c_double)]
        lib.mdft_wsnet_adam.argtypes = [vp, C.c_float]
        lib.mdft_ws_features.argtypes = [vp, vp, vp]
        lib.mdft_ws_gamma.argtypes = [vp, vp, vp, vp]
        lib.mdft_ws_density.argtypes = [vp, vp, vp]
        lib.mdft_ws_guess.argtypes = [vp, vp, C.c_int, vp]
        lib.mdft_ws_mix.argtypes = [vp, C.c_double, vp, vp, vp]
        lib.mdft_wsnet_backward.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp]
        lib.mdft_ws_start_gradient.argtypes = [vp, vp, vp, vp]
        for n in WSNET_EXPORTS:
            if n not in ("mdft_wsnet_destroy", "mdft_wsnet_last_error", "mdft_wsnet_num_params"):
                getattr(lib, n).restype = C.c_int
        _bound = True
    return lib


class WarmStartNet:
    """The network: create with random (seeded) weights or load saved ones."""

    def __init__(self, channels, ksize, dilations, seed: int = 0, device: int = 0):
        self.lib = _lib()
        self.input_scale = None   # per-channel feature scale [2 np] (float32) or None
        self.start_mix = 1.0      # lam of the start (1 - lam) rho_cold + lam rho_pred (calibrated)
        self.channels = tuple(int(c) for c in channels)
        self.ksize = tuple(int(k) for k in ksize)
        self.dilations = tuple(int(d) for d in dilations)
        ch = (C.c_int * len(self.channels))(*self.channels)
        ks = (C.c_int * len(self.ksize))(*self.ksize)
        dl = (C.c_int * len(self.dilations))(*self.dilations)
        h = C.c_void_p()
        rc = self.lib.mdft_wsnet_create(len(self.dilations), ch, ks, dl, int(seed), int(device), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"mdft_wsnet_create: {self.lib.mdft_wsnet_last_error(None).decode()}")
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"wsnet error {rc}: {self.lib.mdft_wsnet_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None) is not None:
            self.lib.mdft_wsnet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        return int(self.lib.mdft_wsnet_num_params(self._h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.mdft_wsnet_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._check(self.lib.mdft_wsnet_set_params(self._h, p.ctypes.data))

    def get_grads(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.mdft_wsnet_get_grads(self._h, out.ctypes.data))
        return out

    def layers(self, p=None):
        """Split a parameter vector into [(W [co][ci][k][k][k], b [co])]."""
        p = self.get_params() if p is None else p
        out, off = [], 0
        for ci, co, k in zip(self.channels[:-1], self.channels[1:], self.ksize):
            w = p[off:off + co * ci * k ** 3].reshape(co, ci, k, k, k)
            off += co * ci * k ** 3
            out.append((w, p[off:off + co]))
            off += co
        return out

    def forward(self, n, d_in, d_out) -> None:
        self._check(self.lib.mdft_wsnet_forward(self._h, n[0], n[1], n[2], _ptr(d_in), _ptr(d_out)))

    def loss_grad(self, n, d_in, d_target, d_mask=None, d_gin=None) -> float:
        loss = C.c_double()
        self._check(self.lib.mdft_wsnet_loss_grad(self._h, n[0], n[1], n[2], _ptr(d_in), _ptr(d_target),
                                                  _ptr(d_mask), _ptr(d_gin), C.byref(loss)))
        return loss.value

    def backward(self, n, d_in, d_gout, d_gin=None) -> None:
        """Parameter gradients (and dL/d(input) into d_gin) for the output gradient
        d_gout; follows forward() on the same d_in."""
        self._check(self.lib.mdft_wsnet_backward(self._h, n[0], n[1], n[2], _ptr(d_in), _ptr(d_gout), _ptr(d_gin)))

    def adam(self, lr: float) -> None:
        self._check(self.lib.mdft_wsnet_adam(self._h, float(lr)))

    def save(self, path: str, **meta) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        extra = {} if self.input_scale is None else {"input_scale": np.asarray(self.input_scale, np.float32)}
        extra["start_mix"] = np.asarray(float(self.start_mix))
        np.savez(path, params=self.get_params(), channels=np.array(self.channels), ksize=np.array(self.ksize),
                 dilations=np.array(self.dilations), **extra, **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, device: int = 0) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as z:
            net = cls(tuple(z["channels"]), tuple(z["ksize"]), tuple(z["dilations"]), device=device)
            net.set_params(z["params"])
            if "input_scale" in z.files:
                net.input_scale = np.ascontiguousarray(z["input_scale"], dtype=np.float32)
            if "start_mix" in z.files:
                net.start_mix = float(z["start_mix"])
        return net

    @classmethod
    def for_engine(cls, eng: Engine, seed: int = 0, hidden: int = HIDDEN) -> "WarmStartNet":
        return cls(*architecture(eng.np_, hidden), seed=seed)


def _ptr(b):
    if b is None:
        return None
    return b.ptr if isinstance(b, DeviceBuffer) else b


def features(eng: Engine, scale=None) -> DeviceBuffer:
    """[2 np][N] FP32: projections of the feature field clip(-beta vext / 2, -4, 4) / 4
    of the engine's vext, channel k times scale[k] (the network's input scale)."""
    out = DeviceBuffer(eng, 4 * 2 * eng.np_ * eng.N)
    sc = None
    if scale is not None:
        sc = np.ascontiguousarray(scale, dtype=np.float32)
        assert sc.size == 2 * eng.np_
    eng._check(_lib().mdft_ws_features(eng._h, out.ptr, None if sc is None else sc.ctypes.data))
    return out


def input_scale(eng: Engine, solutes, solvent) -> np.ndarray:
    """Per-channel input scale 1 / rms of the unscaled features over `solutes`
    (channels that are zero throughout get 1): every input channel of the
    network then has unit rms on the training distribution."""
    acc = np.zeros(2 * eng.np_)
    for sol in solutes:
        eng.compute_vext(sol, solvent)
        f = features(eng)
        a = f.download(np.float32).reshape(2 * eng.np_, eng.N).astype(np.float64)
        f.free()
        acc += (a * a).mean(axis=1)
    rms = np.sqrt(acc / max(len(solutes), 1))
    return np.where(rms > 1e-12, 1.0 / np.maximum(rms, 1e-12), 1.0).astype(np.float32)


def gamma(eng: Engine, xi: DeviceBuffer, want_field: bool = False):
    """Projections [2 np][N] (FP32) of gamma = c (*) drho for xi; with
    want_field also the field gamma(r, Omega) [N][no] (FP64)."""
    tgt = DeviceBuffer(eng, 4 * 2 * eng.np_ * eng.N)
    g = DeviceBuffer(eng, 8 * eng.N * eng.no) if want_field else None
    eng._check(_lib().mdft_ws_gamma(eng._h, xi.ptr, _ptr(g), tgt.ptr))
    return (tgt, g) if want_field else tgt


def density(eng: Engine, xi: DeviceBuffer) -> DeviceBuffer:
    """Projections [2 np][N] (FP32) of rho/rho0 - 1 for xi."""
    tgt = DeviceBuffer(eng, 4 * 2 * eng.np_ * eng.N)
    eng._check(_lib().mdft_ws_density(eng._h, xi.ptr, tgt.ptr))
    return tgt


def guess(eng: Engine, pred: DeviceBuffer | None, xi: DeviceBuffer, kind: int = WS_DENSITY) -> None:
    eng._check(_lib().mdft_ws_guess(eng._h, _ptr(pred), int(kind), xi.ptr))


def cold_start(eng: Engine, xi: DeviceBuffer) -> None:
    """xi <- the reference cold start (src/module_density.f90:267-284), on the device."""
    guess(eng, None, xi)


def boltzmann_start(eng: Engine, xi: DeviceBuffer) -> None:
    """xi <- exp(-max(beta vext, -3) / 2): the solvent density without correlations,
    capped at e^3 rho0, on the device."""
    guess(eng, None, xi, WS_BOLTZMANN)


def mix(eng: Engine, lam: float, xa: DeviceBuffer, xb: DeviceBuffer, out: DeviceBuffer) -> None:
    eng._check(_lib().mdft_ws_mix(eng._h, float(lam), xa.ptr, xb.ptr, out.ptr))


def start_gradient(eng: Engine, xi: DeviceBuffer, df: DeviceBuffer, out: DeviceBuffer | None = None) -> DeviceBuffer:
    """dF/d(pred) [2 np][N] (FP32) at the density guess xi = guess(pred), given
    dF at xi from energy_and_gradient."""
    out = out or DeviceBuffer(eng, 4 * 2 * eng.np_ * eng.N)
    eng._check(_lib().mdft_ws_start_gradient(eng._h, xi.ptr, df.ptr, out.ptr))
    return out


def _F(eng, xi):
    return float(np.sum(eng.energy_and_gradient(xi, None)[0]))


def warm_start(eng: Engine, net: WarmStartNet, xi: DeviceBuffer, fallback: DeviceBuffer | None = None,
               kind: int = WS_DENSITY, lams=(), start_mix: float | None = None) -> dict:
    """xi <- the network's starting density for the engine's vext: the density
    mixture (1 - lam) rho_cold + lam rho_pred of the reference cold start and
    the prediction, lam = start_mix if given, else net.start_mix (calibrated with
    the weights); a scratch field for the cold start is allocated when lam < 1
    and no `fallback` is given.  With `fallback` and `lams`, F is evaluated at the
    prediction (lam = 1), at the cold start (lam = 0) and at the mixtures `lams`
    instead, and the lowest is kept: that selection costs len(lams) + 2 F
    evaluations and never gives a start of higher F than the cold one."""
    assert net.channels[-1] == 2 * eng.np_ and net.channels[0] == 2 * eng.np_
    lam = float(net.start_mix if start_mix is None else start_mix)
    if not lams and lam <= 0.0 and kind == WS_DENSITY:
        # the calibration gives the prediction no weight: the cold start, without inference
        cold_start(eng, xi)
        return {"start": "cold (calibrated lam = 0, network not run)", "kind": "density", "lam": 0.0}
    f = features(eng, net.input_scale)
    pred = DeviceBuffer(eng, 4 * net.channels[-1] * eng.N)
    net.forward(eng.n, f, pred)
    f.free()
    guess(eng, pred, xi, kind)
    info = {"start": "learned", "kind": "density" if kind == WS_DENSITY else "gamma"}
    if not lams and lam < 1.0 and kind == WS_DENSITY:
        scratch = fallback if fallback is not None else eng.alloc_field()
        cold_start(eng, scratch)
        mix(eng, lam, scratch, xi, xi)
        if fallback is None:
            scratch.free()
        pred.free()
        info.update(start=f"learned, mixed with the cold start (lam={lam})", lam=lam)
        return info
    if fallback is not None:
        trials = {1.0: _F(eng, xi)}
        cold_start(eng, fallback)
        trials[0.0] = _F(eng, fallback)
        for lam in lams:
            if kind != WS_DENSITY:
                break
            guess(eng, pred, xi, kind)
            mix(eng, lam, fallback, xi, xi)
            trials[float(lam)] = _F(eng, xi)
        best = min(trials, key=lambda k: trials[k] if np.isfinite(trials[k]) else np.inf)
        info.update(F={str(k): v for k, v in trials.items()}, lam=best, evaluations=len(trials))
        if best == 0.0:
            eng._check(eng.lib.mdft_vec_lincomb(eng._h, eng.N * eng.no, 1.0, fallback.ptr, 0.0, None, xi.ptr))
            info["start"] = "cold (fallback)"
        else:
            guess(eng, pred, xi, kind)
            if best != 1.0:
                mix(eng, best, fallback, xi, xi)
                info["start"] = f"mixture lam={best}"
    pred.free()
    return info


def make_training_set(w, nsolutes: int, seed: int, ck_table, gtol: float = 1e-4, precision: str = "fp32",
                      maxeval: int = 600, log=None, nscale: int = 16):
    """Solve `nsolutes` seeded solutes of workload `w` (grid, box, mmax, c(k))
    with the plain method from the reference cold start; returns the engine, the
    input scale (from the first `nscale` solutes) and per-solute (features,
    density targets, evaluations, F) with device-resident fields."""
    from . import workloads
    from .minimize import minimize
    eng = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, precision)
    eng.set_ck(ck_table)
    xi = eng.alloc_field()
    g = eng.alloc_field()
    scale = input_scale(eng, [workloads.make_solute(w, seed + i) for i in range(min(nscale, nsolutes))], w.solvent)
    out = []
    for i in range(nsolutes):
        t0 = time.perf_counter()
        sol = workloads.make_solute(w, seed + i)
        eng.compute_vext(sol, w.solvent)
        feat = features(eng, scale)
        cold_start(eng, xi)
        res = minimize(eng, xi, g=g, m=5, gtol=gtol, maxeval=maxeval, fnoise=1e-6)
        tgt = density(eng, xi)
        out.append((feat, tgt, res.evaluations, res.F))
        if log:
            log(f"  solute {i}: {len(sol.q)} atoms, F={res.F:.3f}, evaluations={res.evaluations}, "
                f"status={res.status}, {time.perf_counter() - t0:.2f}s")
    return eng, scale, out


def train(net: WarmStartNet, n, data, epochs: int, lr: float = 1e-3, seed: int = 0, log=None):
    """data: [(features, targets, ...)] device fields of grid n."""
    """Adam over the samples in a seeded random order; returns per-epoch mean loss."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(data))
        tot = 0.0
        for i in order:
            feat, tgt = data[i][0], data[i][1]
            tot += net.loss_grad(n, feat, tgt)
            net.adam(lr)
        hist.append(tot / len(data))
        if log:
            log(f"  epoch {ep}: loss {hist[-1]:.5f}")
    return hist


def train_energy(net: WarmStartNet, eng: Engine, solutes, solvent, feats, epochs: int, lr: float = 1e-4,
                 seed: int = 0, log=None):
    """Train on the free energy of the start itself: for each solute, the loss
    is F at xi = guess(net(features)) and its gradient flows back through the
    guess (mdft_ws_start_gradient) into the network (mdft_wsnet_backward).
    feats[i] are the (scaled) features of solutes[i].  Returns the per-epoch
    mean of F(start) - F(cold start)."""
    rng = np.random.default_rng(seed)
    C_out = net.channels[-1]
    pred = DeviceBuffer(eng, 4 * C_out * eng.N)
    gpred = DeviceBuffer(eng, 4 * C_out * eng.N)
    xi, df = eng.alloc_field(), eng.alloc_field()
    fcold = []
    for sol in solutes:
        eng.compute_vext(sol, solvent)
        cold_start(eng, xi)
        fcold.append(_F(eng, xi))
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(len(solutes)):
            eng.compute_vext(solutes[i], solvent)
            net.forward(eng.n, feats[i], pred)
            guess(eng, pred, xi, WS_DENSITY)
            parts, _ = eng.energy_and_gradient(xi, df)
            tot += float(np.sum(parts)) - fcold[i]
            start_gradient(eng, xi, df, gpred)
            net.backward(eng.n, feats[i], gpred)
            net.adam(lr)
        hist.append(tot / len(solutes))
        if log:
            log(f"  energy epoch {ep}: mean F(start) - F(cold) {hist[-1]:.3f} kJ/mol, lr {lr:.1e}")
    for b in (pred, gpred, xi, df):
        b.free()
    return hist


def start_gap(net: WarmStartNet, eng: Engine, solutes, solvent, feats) -> list:
    """F(start) - F(cold start) (kJ/mol) of the network's start for each solute."""
    pred = DeviceBuffer(eng, 4 * net.channels[-1] * eng.N)
    xi = eng.alloc_field()
    out = []
    for sol, f in zip(solutes, feats):
        eng.compute_vext(sol, solvent)
        cold_start(eng, xi)
        fc = _F(eng, xi)
        net.forward(eng.n, f, pred)
        guess(eng, pred, xi, WS_DENSITY)
        out.append(_F(eng, xi) - fc)
    pred.free()
    xi.free()
    return out

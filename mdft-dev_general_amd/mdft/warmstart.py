"""Learned warm start of the minimisation (ctypes binding of include/mdft_wsnet.h).

The reference has no learned component; the minimiser it drives
(src/energy_minimization.f90) starts from the cold guess of
src/module_density.f90:267-284.  Here a periodic 3D convolutional network
(hand-written HIP forward/backward/Adam kernels, FP32) maps the angular
projections of a feature field of the external potential to the angular
projections of the part of the converged density the reference's cold start
misses, rho/rho0 - c (c = 0 where the cold start excludes the solvent, 1
elsewhere), and the minimiser starts from xi = sqrt(max(c + residual, 0))
(0 where c = 0) -- or from the best of a few density mixtures
(1 - lam) rho_cold + lam rho_pred, the cold start included, judged by F.  The minimisation afterwards is the plain one: the network changes the
number of iterations, not the answer, and the cold start stays the fallback.

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
    "mdft_ws_guess", "mdft_ws_mix",
)
WS_GAMMA, WS_DENSITY = 0, 1

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
        lib.mdft_ws_features.argtypes = [vp, vp]
        lib.mdft_ws_gamma.argtypes = [vp, vp, vp, vp]
        lib.mdft_ws_density.argtypes = [vp, vp, vp]
        lib.mdft_ws_guess.argtypes = [vp, vp, C.c_int, vp]
        lib.mdft_ws_mix.argtypes = [vp, C.c_double, vp, vp, vp]
        for n in WSNET_EXPORTS:
            if n not in ("mdft_wsnet_destroy", "mdft_wsnet_last_error", "mdft_wsnet_num_params"):
                getattr(lib, n).restype = C.c_int
        _bound = True
    return lib


class WarmStartNet:
    """The network: create with random (seeded) weights or load saved ones."""

    def __init__(self, channels, ksize, dilations, seed: int = 0, device: int = 0):
        self.lib = _lib()
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

    def adam(self, lr: float) -> None:
        self._check(self.lib.mdft_wsnet_adam(self._h, float(lr)))

    def save(self, path: str, **meta) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.get_params(), channels=np.array(self.channels), ksize=np.array(self.ksize),
                 dilations=np.array(self.dilations), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, device: int = 0) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as z:
            net = cls(tuple(z["channels"]), tuple(z["ksize"]), tuple(z["dilations"]), device=device)
            net.set_params(z["params"])
        return net

    @classmethod
    def for_engine(cls, eng: Engine, seed: int = 0, hidden: int = HIDDEN) -> "WarmStartNet":
        return cls(*architecture(eng.np_, hidden), seed=seed)


def _ptr(b):
    if b is None:
        return None
    return b.ptr if isinstance(b, DeviceBuffer) else b


def features(eng: Engine) -> DeviceBuffer:
    """[2 np][N] FP32: projections of the feature field of the engine's vext."""
    out = DeviceBuffer(eng, 4 * 2 * eng.np_ * eng.N)
    eng._check(_lib().mdft_ws_features(eng._h, out.ptr))
    return out


def gamma(eng: Engine, xi: DeviceBuffer, want_field: bool = False):
    """Projections [2 np][N] (FP32) of gamma = c (*) drho for xi; with
    want_field also the field gamma(r, Omega) [N][no] (FP64)."""
    tgt = DeviceBuffer(eng, 4 * 2 * eng.np_ * eng.N)
    g = DeviceBuffer(eng, 8 * eng.N * eng.no) if want_field else None
    eng._check(_lib().mdft_ws_gamma(eng._h, xi.ptr, _ptr(g), tgt.ptr))
    return (tgt, g) if want_field else tgt


def density(eng: Engine, xi: DeviceBuffer) -> DeviceBuffer:
    """Projections [2 np][N] (FP32) of rho/rho0 - c for xi (c: the cold start)."""
    tgt = DeviceBuffer(eng, 4 * 2 * eng.np_ * eng.N)
    eng._check(_lib().mdft_ws_density(eng._h, xi.ptr, tgt.ptr))
    return tgt


def guess(eng: Engine, pred: DeviceBuffer | None, xi: DeviceBuffer, kind: int = WS_DENSITY) -> None:
    eng._check(_lib().mdft_ws_guess(eng._h, _ptr(pred), int(kind), xi.ptr))


def cold_start(eng: Engine, xi: DeviceBuffer) -> None:
    """xi <- the reference cold start (src/module_density.f90:267-284), on the device."""
    guess(eng, None, xi)


def mix(eng: Engine, lam: float, xa: DeviceBuffer, xb: DeviceBuffer, out: DeviceBuffer) -> None:
    eng._check(_lib().mdft_ws_mix(eng._h, float(lam), xa.ptr, xb.ptr, out.ptr))


def _F(eng, xi):
    return float(np.sum(eng.energy_and_gradient(xi, None)[0]))


def warm_start(eng: Engine, net: WarmStartNet, xi: DeviceBuffer, fallback: DeviceBuffer | None = None,
               kind: int = WS_DENSITY, lams=(0.5,)) -> dict:
    """xi <- the network's starting density for the engine's vext.  With a
    scratch field `fallback`, F is evaluated at the prediction (lam = 1), at
    the cold start (lam = 0) and at the density mixtures `lams`, and the lowest
    is kept: a poor prediction costs len(lams) + 2 F evaluations and never a
    worse start than the cold one."""
    assert net.channels[-1] == 2 * eng.np_ and net.channels[0] == 2 * eng.np_
    f = features(eng)
    pred = DeviceBuffer(eng, 4 * net.channels[-1] * eng.N)
    net.forward(eng.n, f, pred)
    f.free()
    guess(eng, pred, xi, kind)
    info = {"start": "learned", "kind": "density" if kind == WS_DENSITY else "gamma"}
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
                      maxeval: int = 600, log=None):
    """Solve `nsolutes` seeded solutes of workload `w` (grid, box, mmax, c(k))
    with the plain method from the reference cold start; returns the engine and
    per-solute (features, density targets, evaluations, F) with device-resident fields."""
    from . import workloads
    from .minimize import minimize
    eng = Engine(w.n, w.length, w.mmax, w.mrso, w.kT, w.n0, precision)
    eng.set_ck(ck_table)
    xi = eng.alloc_field()
    g = eng.alloc_field()
    out = []
    for i in range(nsolutes):
        t0 = time.perf_counter()
        sol = workloads.make_solute(w, seed + i)
        eng.compute_vext(sol, w.solvent)
        feat = features(eng)
        cold_start(eng, xi)
        res = minimize(eng, xi, g=g, m=5, gtol=gtol, maxeval=maxeval, fnoise=1e-6)
        tgt = density(eng, xi)
        out.append((feat, tgt, res.evaluations, res.F))
        if log:
            log(f"  solute {i}: {len(sol.q)} atoms, F={res.F:.3f}, evaluations={res.evaluations}, "
                f"status={res.status}, {time.perf_counter() - t0:.2f}s")
    return eng, out


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

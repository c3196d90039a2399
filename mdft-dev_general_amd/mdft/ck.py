"""Direct correlation function tables c^{mn}_{mu nu, khi}(|q|).

* `CkTable` holds a table in the reference file order: np_file index tuples
  (m, n, mu, nu, khi) and rows c[iq, i] on a uniform |q| grid of step dq.
* `read_ck_file` / `write_ck_file` read and write the reference's text format
  (17-line header, then `q (re,im) (re,im) ...` rows), mirroring
  src/module_read_c_luc.f90:6-124 and the file layout of
  data/dcf/spce/spce-ck_nonzero_nmax1_ml.
* `synthetic_ck` builds the seeded synthetic c(k) of the workload
  definitions: real l-basis components A exp(-k^2 s^2/2) j_l(kR), mapped to the
  khi basis with 3j symbols (the l -> khi step of
  src/module_energy_luc_fast.f90:371-400), so that the symmetry laws stated in
  luc-files/ck_nmax1:5-7 hold, then scaled so that the spectral norm of every
  khi block of n0*C(k) is at most `bound`.
* `relayout` performs the reference's re-ordering, conjugation and sign
  folding (src/module_energy_cproj_mrso.f90:381-429) and is what the C-ABI
  entry mdft_set_ck does on the device side.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass
from fractions import Fraction

import numpy as np


@dataclass
class CkTable:
    m: np.ndarray
    n: np.ndarray
    mu: np.ndarray
    nu: np.ndarray
    khi: np.ndarray
    dq: float
    c: np.ndarray  # complex128 [nq_file, np_file]
    q: np.ndarray | None = None  # abscissae of the rows

    @property
    def np_file(self) -> int:
        return int(len(self.m))

    @property
    def nq_file(self) -> int:
        return int(self.c.shape[0])


def npluc(mmax: int, mrso: int = 2) -> int:
    """Number of (m, n, mu, nu, khi) tuples of a file (solvent%npluc)."""
    return len(tuples(mmax, mrso))


def tuples(mmax: int, mrso: int = 2):
    """All (m, n, mu, nu, khi): mu, nu multiples of mrso, |khi| <= min(m, n)."""
    out = []
    for khi in range(-mmax, mmax + 1):
        for m in range(mmax + 1):
            for n in range(mmax + 1):
                if abs(khi) > min(m, n):
                    continue
                for mu2 in range(-(m // mrso), m // mrso + 1):
                    for nu2 in range(-(n // mrso), n // mrso + 1):
                        out.append((m, n, mu2 * mrso, nu2 * mrso, khi))
    return out


def npc(mmax: int, mrso: int = 2) -> int:
    """Number of stored OZ coefficients per |q| (cproj_mrso.f90:381-383)."""
    s = 0
    for m in range(mmax + 1):
        for khi in range(-m, m + 1):
            for _mu2 in range(m // mrso + 1):
                for n in range(abs(khi), mmax + 1):
                    s += 2 * (n // mrso) + 1
    return s


def nq_needed(n, length, dq: float) -> int:
    """Rows of c the evaluation reads: read_c_luc.f90:101 with the q_max of
    cproj_mrso.f90:366 (largest kx, ky and kz over the first nz/2+1 planes)."""
    two_pi = 2.0 * math.acos(-1.0)

    def k(l, nn, ll):
        return (l - 1) * two_pi / ll if l <= nn // 2 + 1 else (l - 1 - nn) * two_pi / ll

    kx = max(k(l, n[0], length[0]) for l in range(1, n[0] + 1))
    ky = max(k(l, n[1], length[1]) for l in range(1, n[1] + 1))
    kz = max(k(l, n[2], length[2]) for l in range(1, n[2] // 2 + 2))
    qmax = math.sqrt(kx * kx + ky * ky + kz * kz)
    return int(qmax / dq + 0.5) + 2


# --------------------------------------------------------------------------- file format

_CPLX = re.compile(r"\(\s*([^,()]+?)\s*,\s*([^,()]+?)\s*\)")


def read_ck_file(path: str, mmax: int | None = None, mrso: int = 2) -> CkTable:
    """Parse a c-file in the reference format (read_c_luc.f90:65-109)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    head = lines[10:16]
    idx = {}
    for ln in head:
        tok = ln.split()
        if tok:
            idx[tok[0].lower()] = np.array([int(t) for t in tok[1:]], dtype=np.int32)
    m, n, mu, nu, khi = (idx[k] for k in ("m", "n", "mu", "nu", "khi"))
    rows, qs = [], []
    for ln in lines[17:]:
        if not ln.strip():
            continue
        qv = float(ln.split()[0].replace("D", "E"))
        vals = [complex(float(a.replace("D", "E")), float(b.replace("D", "E"))) for a, b in _CPLX.findall(ln)]
        rows.append(vals)
        qs.append(qv)
    c = np.array(rows, dtype=np.complex128)
    q = np.array(qs)
    # deltaAbscissa (module_input.f90:591-650): difference of the 4th and 3rd abscissae
    dq = float(q[3] - q[2])
    return CkTable(m, n, mu, nu, khi, dq, c, q)


def write_ck_file(path: str, t: CkTable) -> None:
    """Write `t` in the reference text format (17-line header + rows)."""
    def row(label, arr):
        return f" {label:<6}" + "".join(f"{int(v):12d}" for v in arr)
    with open(path, "w") as fh:
        fh.write("\n" * 10)
        fh.write(row("alpha", range(1, t.np_file + 1)) + "\n")
        for lab, arr in (("m", t.m), ("n", t.n), ("mu", t.mu), ("nu", t.nu), ("khi", t.khi)):
            fh.write(row(lab, arr) + "\n")
        fh.write("\n")
        q = t.q if t.q is not None else np.arange(t.nq_file) * t.dq
        for iq in range(t.nq_file):
            fh.write(f"{q[iq]:25.16E}" + "".join(f" ({v.real:24.16E},{v.imag:24.16E})" for v in t.c[iq]) + "\n")


# --------------------------------------------------------------------------- 3j symbols

def _fac(n: int) -> int:
    return math.factorial(n)


def threej(j1: int, j2: int, j3: int, m1: int, m2: int, m3: int) -> float:
    """Wigner 3j symbol by the Racah formula in exact rational arithmetic."""
    if m1 + m2 + m3 != 0 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    tri = Fraction(_fac(j1 + j2 - j3) * _fac(j1 - j2 + j3) * _fac(-j1 + j2 + j3), _fac(j1 + j2 + j3 + 1))
    pref = tri * _fac(j1 + m1) * _fac(j1 - m1) * _fac(j2 + m2) * _fac(j2 - m2) * _fac(j3 + m3) * _fac(j3 - m3)
    s = Fraction(0)
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    for k in range(kmin, kmax + 1):
        s += Fraction((-1) ** k, _fac(k) * _fac(j1 + j2 - j3 - k) * _fac(j1 - m1 - k) * _fac(j2 + m2 - k)
                      * _fac(j3 - j2 + m1 + k) * _fac(j3 - j1 - m2 + k))
    sign = -1 if (j1 - j2 - m3) % 2 else 1
    return sign * math.sqrt(float(pref)) * float(s)


# --------------------------------------------------------------------------- synthetic c(k)

def _lbasis_orbits(mmax: int, mrso: int):
    """Orbits of the l-basis index set under the two symmetries
    (a) c^{mnl}_{-mu,-nu} = (-1)^{m+n+l} c^{mnl}_{mu nu}  and
    (b) c^{nml}_{nu mu}   = (-1)^{m+n}   c^{mnl}_{mu nu}.
    Returns a list of orbits; each orbit is a list of (key, sign) relative to
    its first member, or None when the symmetries force the orbit to zero."""
    keys = []
    for m in range(mmax + 1):
        for n in range(mmax + 1):
            for l in range(abs(m - n), m + n + 1):
                for mu2 in range(-(m // mrso), m // mrso + 1):
                    for nu2 in range(-(n // mrso), n // mrso + 1):
                        keys.append((m, n, l, mu2 * mrso, nu2 * mrso))
    seen = set()
    orbits = []
    for k in keys:
        if k in seen:
            continue
        members = {k: 1}
        frontier = [k]
        ok = True
        while frontier:
            cur = frontier.pop()
            s = members[cur]
            m, n, l, mu, nu = cur
            for nk, ns in (((m, n, l, -mu, -nu), s * (-1) ** (m + n + l)), ((n, m, l, nu, mu), s * (-1) ** (m + n))):
                if nk in members:
                    if members[nk] != ns:
                        ok = False
                else:
                    members[nk] = ns
                    frontier.append(nk)
        seen.update(members)
        orbits.append(list(members.items()) if ok else [(kk, 0) for kk in members])
    return orbits


def synthetic_ck(mmax: int, nk: int, dk: float, seed: int, n0: float, mrso: int = 2,
                 bound: float = 0.8, a_range=(-2.0, 2.0), s_range=(0.5, 1.5), r_range=(2.0, 4.0),
                 self_term: float = 15.0, self_kc: float = 3.0) -> CkTable:
    """Seeded synthetic c(k) table in the reference file order.

    Each independent real l-basis component is n0*A*exp(-k^2 s^2/2)*j_l(kR)
    with A ~ U(a_range) [A^3], s ~ U(s_range) [A], R ~ U(r_range) [A]; j_l is
    the spherical Bessel function of the component's order l (j0 for l = 0).  The khi
    components are c^{mn}_{mu nu,khi} = sum_l 3j(m n l; khi -khi 0) c^{mnl}_{mu nu}
    (times i for odd l).  The whole table is then scaled by one factor so that
    the largest singular value of n0*C_khi(k) over all k and khi is <= bound.

    With self_term = s > 0 each row k then gains -(g(k) + s exp(-(k/kc)^8)) I,
    where g(k) is the largest singular value of C_khi(k) over khi and I is the
    identity operator of add_self_term.  The g(k) part makes the excess term
    convex (F_exc >= 0 for every drho), so F is bounded below next to charged
    sites; the s part is a compressibility term (rho0 c(k -> 0) = -s; s = 15 is
    about water's, S(0) = 1/(1 + s) ~ 0.06) that keeps the density next to
    charged sites finite.  Both parts vanish on the Nyquist planes of the
    grids used here (|k| >= 6.28 A^-1, where it is below 1e-150), where the reference's even-grid
    treatment is not symmetric."""
    from scipy.special import spherical_jn

    rng = np.random.default_rng(seed)
    k = np.arange(nk) * dk
    orbits = _lbasis_orbits(mmax, mrso)
    lval = {}
    for orb in orbits:
        a = rng.uniform(*a_range)
        s = rng.uniform(*s_range)
        r = rng.uniform(*r_range)
        l = orb[0][0][2]
        # j_l(kR): an order-l component must vanish like k^l at k = 0; with j0
        # for l > 0 the self-paired q = 0 point makes the functional non-symmetric
        base = n0 * a * np.exp(-0.5 * (k * s) ** 2) * spherical_jn(l, k * r)
        for key, sgn in orb:
            lval[key] = sgn * base
    tup = tuples(mmax, mrso)
    c = np.zeros((nk, len(tup)), dtype=np.complex128)
    for i, (m, n, mu, nu, khi) in enumerate(tup):
        acc = np.zeros(nk, dtype=np.complex128)
        for l in range(abs(m - n), m + n + 1):
            w3 = threej(m, n, l, khi, -khi, 0)
            if w3 == 0.0:
                continue
            acc += w3 * lval[(m, n, l, mu, nu)] * (1j if l % 2 else 1.0)
        c[:, i] = acc
    arr = np.array(tup, dtype=np.int32)
    t = CkTable(arr[:, 0].copy(), arr[:, 1].copy(), arr[:, 2].copy(), arr[:, 3].copy(), arr[:, 4].copy(),
                float(dk), c, k)
    norm = spectral_norm_max(t, mmax, mrso)
    if norm > bound:
        t.c *= bound / norm
    if self_term:
        add_self_term(t, -self_profile(t, mmax, mrso, self_term, self_kc))
    return t


def self_profile(t: CkTable, mmax: int, mrso: int, s: float, kc: float) -> np.ndarray:
    """Per-row strength of the self term of synthetic_ck."""
    k = np.arange(t.nq_file) * t.dq  # row i holds k = i dq
    return row_norms(t, mmax, mrso) + s * np.exp(-(k / kc) ** 8)


def row_norms(t: CkTable, mmax: int, mrso: int = 2) -> np.ndarray:
    """Largest singular value of C_khi(k) over khi, per row k."""
    g = np.zeros(t.nq_file)
    for _khi, (_keys, mat) in khi_blocks(t, mmax, mrso).items():
        if mat.shape[1]:
            g = np.maximum(g, np.linalg.norm(mat, ord=2, axis=(1, 2)))
    return g


def add_self_term(t: CkTable, lam) -> None:
    """Adds lam * I to the table (lam a scalar or one value per row k), where I
    is the identity operator of the excess functional in the reference's
    conventions: c(m, n, mu, nu; khi) = delta_mn delta_{mu,-nu} (-1)^khi.
    In the metric of the ideal term's curvature at rho0 (kT dv w / rho0), the
    excess Hessian of lam * I is -lam times the orthogonal projector on the
    represented projections (exact on odd grids; on even grids the reference's
    Nyquist-plane treatment is not symmetric and flips it there)."""
    lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), (t.nq_file,))
    for i in range(t.np_file):
        if t.m[i] == t.n[i] and t.mu[i] == -t.nu[i]:
            t.c[:, i] += lam * (-1.0) ** int(t.khi[i])


def khi_blocks(t: CkTable, mmax: int, mrso: int = 2):
    """Per khi: (row keys (m,mu), col keys (n,nu), matrix [nk, rows, cols])."""
    pos = {(int(t.m[i]), int(t.n[i]), int(t.mu[i]), int(t.nu[i]), int(t.khi[i])): i for i in range(t.np_file)}
    out = {}
    for khi in range(-mmax, mmax + 1):
        keys = [(m, mu2 * mrso) for m in range(abs(khi), mmax + 1) for mu2 in range(-(m // mrso), m // mrso + 1)]
        mat = np.zeros((t.nq_file, len(keys), len(keys)), dtype=np.complex128)
        for a, (m, mu) in enumerate(keys):
            for b, (n, nu) in enumerate(keys):
                i = pos.get((m, n, mu, nu, khi))
                if i is not None:
                    mat[:, a, b] = t.c[:, i]
        out[khi] = (keys, mat)
    return out


def spectral_norm_max(t: CkTable, mmax: int, mrso: int = 2) -> float:
    best = 0.0
    for _khi, (_keys, mat) in khi_blocks(t, mmax, mrso).items():
        if mat.shape[1] == 0:
            continue
        best = max(best, float(np.linalg.norm(mat, ord=2, axis=(1, 2)).max()))
    return best


def relayout(t: CkTable, mmax: int, mrso: int, nq: int) -> np.ndarray:
    """Stored OZ table [nq, npc]: cproj_mrso.f90:381-401 ordering (m, khi,
    mu2>=0, n>=|khi|, nu2), conj (:415) and the sign folding (:420-429)."""
    pos = {}
    for i in range(t.np_file):
        pos[(int(t.m[i]), int(t.n[i]), int(t.mu[i]) // mrso, int(t.nu[i]) // mrso, int(t.khi[i]))] = i
    cols, signs = [], []
    for m in range(mmax + 1):
        for khi in range(-m, m + 1):
            for mu2 in range(m // mrso + 1):
                for n in range(abs(khi), mmax + 1):
                    for nu2 in range(-(n // mrso), n // mrso + 1):
                        key = (m, n, mu2, nu2, khi)
                        if key not in pos:
                            raise KeyError(f"c-table lacks tuple m,n,mu2,nu2,khi={key}")
                        cols.append(pos[key])
                        nu = nu2 * mrso
                        signs.append((-1) ** abs(khi + nu) if nu < 0 else (-1) ** n)
    if t.nq_file < nq:
        raise ValueError(f"c-table has {t.nq_file} rows, {nq} needed")
    return np.conj(t.c[:nq, cols]) * np.array(signs, dtype=np.float64)[None, :]


def structure_factor_s0(t: CkTable) -> float:
    """S(0) = 1 / (1 - c_{00}^{00,0}(q = 0)) of a table in the reference's file
    convention, where the (m, n, mu, nu, khi) = (0, 0, 0, 0, 0) column is n0 times the
    k = 0 transform of the orientation-averaged direct correlation function (the file
    carries n0 c; the excess prefactor divides by n0, src/module_energy_cproj_mrso.f90:728)
    and row 0 is q = 0.  1 / S(0) is the factor by which the excess term raises the
    curvature of long-wavelength density changes over the ideal term's (SPC/E nmax1:
    c = -13.67, S(0) = 0.068; water's is about 0.06)."""
    idx = [i for i in range(t.np_file)
           if (int(t.m[i]), int(t.n[i]), int(t.mu[i]), int(t.nu[i]), int(t.khi[i])) == (0, 0, 0, 0, 0)]
    if not idx:
        raise ValueError("the table has no (0, 0, 0, 0, 0) column")
    q0 = float(t.q[0]) if t.q is not None else 0.0
    if q0 != 0.0:
        raise ValueError("row 0 of the table is not q = 0")
    c00 = float(np.real(t.c[0, idx[0]]))
    if not c00 < 1.0:
        raise ValueError(f"c_000(0) = {c00} >= 1: the homogeneous fluid is unstable (S(0) < 0)")
    return 1.0 / (1.0 - c00)

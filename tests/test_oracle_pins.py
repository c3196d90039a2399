"""Pins of the FP64 CPU oracle (oracle/mdft_oracle.c) against closed forms,
independent textbook routines, numpy, and -- where the reference tree is
present -- the reference's own Wigner-d and rotation-matrix code compiled from
its sources into oracle/_ref.  Each check is chosen so that a plausible slip
(wrong sign, conjugation, index order, dropped term) fails it."""
import math

import numpy as np
import pytest

from mdft import ck as ckmod


def jacobi_wigner_d(j, mp, m, beta):
    """Textbook d^j_{m'm}(beta) (sum form, standard phase convention)."""
    f = math.factorial
    s = 0.0
    for k in range(0, 2 * j + 1):
        a, b, c, d = j + m - k, k, mp - m + k, j - mp - k
        if min(a, b, c, d) < 0:
            continue
        s += (-1) ** (mp - m + k) * math.cos(beta / 2) ** (2 * j + m - mp - 2 * k) * \
            math.sin(beta / 2) ** (mp - m + 2 * k) / (f(a) * f(b) * f(c) * f(d))
    return math.sqrt(f(j + mp) * f(j - mp) * f(j + m) * f(j - m)) * s


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 7])
def test_gauss_legendre_matches_numpy(oracle_mod, n):
    x, w = oracle_mod.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    order = np.argsort(-xr)  # reference order: descending cos(theta)
    assert np.allclose(x, xr[order], atol=1e-15)
    assert np.allclose(w, wr[order] / 2.0, atol=1e-15)


def test_weights_normalisation(oracle_mod):
    for mmax in range(0, 6):
        o = oracle_mod.Oracle((4, 4, 4), (10.0, 10.0, 10.0), mmax, 2, 2.477, 0.0333)
        assert abs(o.w.sum() - 8 * np.pi ** 2 / 2) < 1e-12  # module_grid.f90:147-154


# ---------------------------------------------------------------- Wigner d
def test_wigner_known_answers(oracle_mod):
    """Closed forms of tools/test_function_harm_sph_de_luc.f90:24-36."""
    d = oracle_mod.wigner_small_d
    for th in np.linspace(0.05, np.pi - 0.05, 17):
        c, s = math.cos(th), math.sin(th)
        cases = {
            (1, 0, 0): c, (1, 1, 1): (1 + c) / 2, (1, 1, 0): -s / math.sqrt(2), (1, 1, -1): (1 - c) / 2,
            (2, 2, 2): ((1 + c) / 2) ** 2, (2, 2, 1): -(1 + c) / 2 * s, (2, 2, 0): math.sqrt(6) / 4 * s * s,
            (2, 2, -1): -(1 - c) / 2 * s, (2, 2, -2): ((1 - c) / 2) ** 2, (2, 1, 1): (1 + c) / 2 * (2 * c - 1),
            (2, 1, 0): -math.sqrt(1.5) * s * c, (2, 1, -1): (1 - c) / 2 * (2 * c + 1), (2, 0, 0): 1.5 * c * c - 0.5,
        }
        for (m, mu, mup), val in cases.items():
            assert abs(d(m, mu, mup, th) - val) < 1e-14, (m, mu, mup, th)


def test_wigner_matches_textbook_sum(oracle_mod):
    for m in range(0, 7):
        for mu in range(-m, m + 1):
            for mup in range(-m, m + 1):
                for th in (0.3, 1.1, 2.0, 2.9):
                    assert abs(oracle_mod.wigner_small_d(m, mu, mup, th) - jacobi_wigner_d(m, mu, mup, th)) < 1e-12


def test_wigner_orthogonality_on_gl_nodes(oracle_mod):
    for mmax in (2, 5):
        x, w = oracle_mod.gauss_legendre(mmax + 1)
        th = np.arccos(x)
        for a in (0, 1, 2):
            for b in (0, -2, 2):
                for m1 in range(max(abs(a), abs(b)), mmax + 1):
                    for m2 in range(max(abs(a), abs(b)), mmax + 1):
                        s = sum(w[i] * oracle_mod.wigner_small_d(m1, a, b, th[i]) *
                                oracle_mod.wigner_small_d(m2, a, b, th[i]) for i in range(len(th)))
                        assert abs(s - (1.0 / (2 * m1 + 1) if m1 == m2 else 0.0)) < 1e-13


# ---------------------------------------------------------------- R(q)
def test_rotation_closed_form(oracle_mod):
    """R^m_{a,b}(q) = exp(-i a phi_q) d^m_{a,b}(theta_q) for q off the z axis."""
    rng = np.random.default_rng(3)
    for _ in range(12):
        q = rng.normal(size=3)
        th = math.acos(q[2] / np.linalg.norm(q))
        ph = math.atan2(q[1], q[0])
        R = oracle_mod.rotation_matrix(6, q)
        for m in range(7):
            for a in range(-m, m + 1):
                for b in range(-m, m + 1):
                    ref = np.exp(-1j * a * ph) * jacobi_wigner_d(m, a, b, th)
                    assert abs(R[m, a + 6, b + 6] - ref) < 1e-12, (m, a, b)


def test_rotation_axis_cases(oracle_mod):
    L = 5
    W = 2 * L + 1
    for q in ([0.0, 0.0, 2.0], [0.0, 0.0, 0.0]):
        R = oracle_mod.rotation_matrix(L, np.array(q))
        for m in range(L + 1):
            blk = R[m, L - m:L + m + 1, L - m:L + m + 1]
            assert np.allclose(blk, np.eye(2 * m + 1), atol=1e-14)
    # q along -z: frame (x, -y, -z), i.e. Euler (0, pi, pi)
    R = oracle_mod.rotation_matrix(L, np.array([0.0, 0.0, -1.0]))
    for m in range(L + 1):
        for a in range(-m, m + 1):
            for b in range(-m, m + 1):
                ref = jacobi_wigner_d(m, a, b, math.pi) * np.exp(-1j * b * math.pi)
                assert abs(R[m, a + L, b + L] - ref) < 1e-12
    assert R.shape == (L + 1, W, W)


def test_rotation_blocks_unitary(oracle_mod):
    rng = np.random.default_rng(5)
    for _ in range(10):
        q = rng.normal(size=3) * 3
        R = oracle_mod.rotation_matrix(6, q)
        for m in range(7):
            blk = R[m, 6 - m:6 + m + 1, 6 - m:6 + m + 1]
            assert np.allclose(blk @ blk.conj().T, np.eye(2 * m + 1), atol=1e-12)


def test_oracle_matches_reference_pieces(oracle_mod):
    """Wigner-d and R(q) against the reference's own Fortran, compiled from its
    sources (src/module_wigner_d.f90, src/module_rotation.f90) into oracle/_ref."""
    import os
    if not os.path.isdir("/root/reference/src"):
        pytest.skip("reference tree not present")
    oracle_mod.build(ref=True)
    ref = oracle_mod.RefPieces()
    rng = np.random.default_rng(11)
    for m in range(0, 7):
        for mu in range(-m, m + 1):
            for mup in range(-m, m + 1):
                for th in rng.uniform(0, np.pi, 3):
                    assert abs(oracle_mod.wigner_small_d(m, mu, mup, th) - ref.wigner_small_d(m, mu, mup, th)) < 1e-13
    for mmax in range(0, 7):
        for t in range(30):
            q = rng.normal(size=3)
            if t % 5 == 0:
                q[:2] = 0
            if t % 7 == 0:
                q[:] = 0
            if t % 3 == 0:
                q = np.round(q * 3)
            assert np.abs(oracle_mod.rotation_matrix(mmax, q) - ref.rotation_matrix(mmax, q)).max() < 1e-13


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(4, 6, 8), (12, 10, 9), (7, 5, 3), (16, 16, 16)])
def test_fft3d_matches_numpy(oracle_mod, shape):
    rng = np.random.default_rng(0)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    plus = oracle_mod.fft3d(a, +1)
    minus = oracle_mod.fft3d(a, -1)
    n = a.size
    assert np.allclose(plus, np.fft.ifftn(a) * n, atol=1e-10)
    assert np.allclose(minus, np.fft.fftn(a), atol=1e-10)


# ---------------------------------------------------------------- angular transforms
def _ctx(oracle_mod, mmax, n=(6, 6, 6), length=(12.0, 12.0, 12.0)):
    return oracle_mod.Oracle(n, length, mmax, 2, 2.4777, 0.0333)


@pytest.mark.parametrize("mmax", [0, 1, 2, 3, 5])
def test_angular_projector(oracle_mod, mmax):
    o = _ctx(oracle_mod, mmax)
    rng = np.random.default_rng(mmax)
    f = rng.normal(size=o.no)
    p = o.angl2proj(f)
    fb = o.proj2angl(p)             # band-limited part
    assert np.allclose(o.angl2proj(fb), p, atol=1e-12)          # idempotent
    assert np.allclose(o.proj2angl(o.angl2proj(fb)), fb, atol=1e-12)
    const = o.angl2proj(np.full(o.no, 2.5))
    assert abs(const[0] - 2.5) < 1e-13 and np.all(np.abs(const[1:]) < 1e-13)


@pytest.mark.parametrize("mmax", [1, 3, 5])
def test_proj2angl_adjoint(oracle_mod, mmax):
    """The transpose of proj2angl (a real-linear map from the 2 np real and
    imaginary parts to the no orientations) is D angl2proj W^-1, W the
    quadrature weights and D = 8 pi^2 / mrso on mu = 0 channels, twice that on
    mu > 0 channels, 0 on the imaginary channel of (m, 0, 0) -- the factors the
    warm start's training gradient (mdft_ws_start_gradient) applies.  Checked
    against the transpose of the matrix built column by column."""
    from mdft.grid import angular_grid
    o = _ctx(oracle_mod, mmax)
    w = angular_grid(mmax, 2)["w"]
    A = np.zeros((o.no, 2 * o.np_))
    for k in range(2 * o.np_):
        pv = np.zeros(o.np_, complex)
        pv[k // 2] = 1.0 if k % 2 == 0 else 1j
        A[:, k] = o.proj2angl(pv)
    B = np.zeros((2 * o.np_, o.no))
    for i in range(o.no):
        e = np.zeros(o.no)
        e[i] = 1.0
        P = o.angl2proj(e)
        B[0::2, i], B[1::2, i] = P.real, P.imag
    s = 8.0 * np.pi ** 2 / 2
    D = np.zeros(2 * o.np_)
    for ip in range(o.np_):
        f = 2 * s if o.pmu[ip] > 0 else s
        D[2 * ip] = f
        D[2 * ip + 1] = 0.0 if (o.pmup[ip] == 0 and o.pmu[ip] == 0) else f
    assert np.abs(A.T - D[:, None] * B / w[None, :]).max() < 1e-12 * np.abs(A).max()


# ---------------------------------------------------------------- excess functional
def _setup(oracle_mod, mmax, n=8, length=10.0, seed=4, self_term=1.0):
    o = oracle_mod.Oracle((n, n, n), (length,) * 3, mmax, 2, 2.4777, 0.0333)
    t = ckmod.synthetic_ck(mmax, 256, 0.05, seed=seed, n0=0.0333, self_term=self_term)
    o.set_ck(t)
    return o, t


def _xi_from_drho(o, drho):
    return np.sqrt(1.0 + drho / o.rho0)


@pytest.mark.parametrize("mmax", [1, 2, 3])
def test_excess_zero_and_quadratic(oracle_mod, mmax):
    o, _ = _setup(oracle_mod, mmax)
    xi1 = np.ones(o.N * o.no)
    f0, df0 = o.excess(xi1)
    assert f0 == 0.0 and np.abs(df0).max() == 0.0
    rng = np.random.default_rng(1)
    d = 0.3 * o.rho0 * rng.uniform(-1, 1, o.N * o.no)
    f1, _ = o.excess(_xi_from_drho(o, d))
    f2, _ = o.excess(_xi_from_drho(o, 2.0 * d))
    assert abs(f2 - 4.0 * f1) < 1e-9 * abs(f1)


def _vexc(o, drho):
    xi = _xi_from_drho(o, drho)
    f, df = o.excess(xi)
    return f, df / (2.0 * o.rho0 * xi)


@pytest.mark.parametrize("mmax", [1, 2, 3])
def test_excess_gradient_is_exact_on_odd_grids(oracle_mod, mmax):
    """F_exc is a quadratic form in drho, so a central difference with a large
    step is exact; df = 2 rho0 xi vexc must match it.  The residual allowed
    (1e-5) covers the single-precision rounding inside the gamma chop
    (cproj_mrso.f90:618-623), which puts ~1e-8 relative noise on F."""
    o, _ = _setup(oracle_mod, mmax, n=5)
    rng = np.random.default_rng(2)
    d = 0.2 * o.rho0 * rng.uniform(-1, 1, o.N * o.no)
    b = 0.05 * o.rho0 * rng.normal(size=d.size)
    f0, vx = _vexc(o, d)
    assert abs(f0 - 0.5 * o.dv * float(d @ vx)) < 1e-12 * abs(f0)
    h = 0.1
    fp, _ = o.excess(_xi_from_drho(o, d + h * b), want_df=False)
    fm, _ = o.excess(_xi_from_drho(o, d - h * b), want_df=False)
    fd = (fp - fm) / (2 * h)
    an = o.dv * float(vx @ b)
    assert abs(fd - an) < 1e-5 * abs(an)


def _no_nyquist(o, d, n):
    g = np.fft.fftn(d.reshape(n, n, n, o.no), axes=(0, 1, 2))
    g[n // 2], g[:, n // 2], g[:, :, n // 2] = 0, 0, 0
    return np.fft.ifftn(g, axes=(0, 1, 2)).real.ravel()


def _asym(o, a, b):
    va = _vexc(o, a)[1]
    vb = _vexc(o, b)[1]
    return (a @ vb - b @ va) / abs(a @ vb)


def test_excess_operator_symmetry_and_nyquist_quirk(oracle_mod):
    """The map drho -> vexc is symmetric except through the even-grid Nyquist
    planes: the reference rotates the aliased +/-k_Nyquist modes with
    direction-dependent R(q) (cproj_mrso.f90:462-476,648-656), which breaks
    the symmetry for mmax >= 1.  Odd grids and Nyquist-free fields are
    symmetric to the single-precision chop level."""
    rng = np.random.default_rng(1)
    o5, _ = _setup(oracle_mod, 2, n=5, self_term=0.0)
    a = 0.2 * o5.rho0 * rng.uniform(-1, 1, o5.N * o5.no)
    b = 0.2 * o5.rho0 * rng.uniform(-1, 1, o5.N * o5.no)
    assert abs(_asym(o5, a, b)) < 1e-6
    o6, _ = _setup(oracle_mod, 2, n=6, self_term=0.0)
    a = 0.2 * o6.rho0 * rng.uniform(-1, 1, o6.N * o6.no)
    b = 0.2 * o6.rho0 * rng.uniform(-1, 1, o6.N * o6.no)
    assert abs(_asym(o6, _no_nyquist(o6, a, 6), _no_nyquist(o6, b, 6))) < 1e-6
    assert abs(_asym(o6, a, b)) > 1e-4


def test_ideal_gradient_finite_difference(oracle_mod):
    o, _ = _setup(oracle_mod, 2, n=4)
    rng = np.random.default_rng(2)
    vext = rng.uniform(-2, 5, o.N * o.no)
    xi = 1.0 + 0.3 * rng.uniform(-1, 1, o.N * o.no)
    xi[::7] = 1e-9   # near-zero branch of module_energy_ideal_and_external.f90:58-66
    fid, fext, df = o.ideal_ext(xi, vext)
    e = rng.normal(size=xi.size)
    e[::7] = 0.0
    h = 1e-6
    a = o.ideal_ext(xi + h * e, vext, want_df=False)
    b = o.ideal_ext(xi - h * e, vext, want_df=False)
    fd = ((a[0] + a[1]) - (b[0] + b[1])) / (2 * h)
    an = o.dv * float(df @ e)
    assert abs(fd - an) < 1e-7 * abs(an)


def test_ideal_closed_form_uniform(oracle_mod):
    """Uniform xi: F_id = kT V n0 (x ln x - x + 1) with x = xi^2, F_ext = V n0 x v."""
    o, _ = _setup(oracle_mod, 3, n=4, length=8.0)
    x, v = 1.3, 0.7
    fid, fext, _ = o.ideal_ext(np.full(o.N * o.no, np.sqrt(x)), np.full(o.N * o.no, v), want_df=False)
    V = 8.0 ** 3
    assert abs(fid - o.kT * V * o.n0 * (x * np.log(x) - x + 1)) < 1e-11 * abs(fid)
    assert abs(fext - V * o.n0 * x * v) < 1e-11 * abs(fext)
    fid0, _, _ = o.ideal_ext(np.zeros(o.N * o.no), np.zeros(o.N * o.no), want_df=False)
    assert abs(fid0 - o.kT * V * o.n0) < 1e-11 * fid0


def test_mmax0_is_scalar_convolution(oracle_mod):
    """At mmax = 0 the excess term is the scalar convolution
    F = 1/2 dv sum_r drho(r) pref w (c * drho)(r) computed with numpy FFTs."""
    n, Lb = 10, 11.0
    o, t = _setup(oracle_mod, 0, n=n, length=Lb)
    rng = np.random.default_rng(7)
    d = 0.4 * o.rho0 * rng.uniform(-1, 1, o.N)
    f, _ = o.excess(_xi_from_drho(o, d))
    # numpy restatement on the real field
    dr = d.reshape(n, n, n)  # [z, y, x]
    k = 2 * np.pi * np.fft.fftfreq(n, d=Lb / n)
    k = np.where(np.arange(n) == n // 2, abs(k), k)   # FFTW order with +Nyquist
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    kn = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    eff = kn / t.dq + 1.0
    iq = eff.astype(int)
    a = eff - iq
    cr = np.conj(t.c[:, 0])  # relayout conj; nu = n = 0 so the folded sign is +1
    ceff = a * cr[iq] + (1 - a) * cr[iq - 1]
    g = ceff * (np.fft.ifftn(dr) * dr.size)
    # DBL_EPSILON chop of cproj_mrso.f90:618-623: a component next to an exactly
    # zero partner is kept in single precision (default-kind CMPLX)
    eps = np.finfo(float).eps
    re, im = g.real.copy(), g.imag.copy()
    both = (np.abs(re) < eps) & (np.abs(im) < eps)
    re[both], im[both] = 0.0, 0.0
    m1 = np.abs(re) < eps
    re[m1], im[m1] = 0.0, im[m1].astype(np.float32)
    m2 = np.abs(im) < eps
    im[m2], re[m2] = 0.0, re[m2].astype(np.float32)
    gam = np.fft.fftn(re + 1j * im) / dr.size
    pref = -o.kT * 2.0 * 4 * np.pi ** 2 / 2 / o.n0
    w = 8 * np.pi ** 2 / 2
    fref = 0.5 * o.dv * np.sum(dr * pref * w * gam.real)
    assert abs(f - fref) < 1e-12 * abs(fref)


def test_excess_invariant_under_z_rotation(oracle_mod):
    """F_exc is unchanged when the whole orientational field is rotated by 90
    degrees about z: r -> R r on the cubic grid and projections pick up
    exp(i mup pi/2).  Pins the relative conventions of R(q), the projections
    and the -q pairing.  An odd grid is used: on even grids the reference's
    Nyquist-plane treatment depends on the direction of q, so the invariance
    only holds for Nyquist-free fields."""
    n = 9
    o, _ = _setup(oracle_mod, 3, n=n, length=9.0)
    rng = np.random.default_rng(9)
    raw = rng.uniform(-1, 1, (o.N, o.no)) * 0.2 * o.rho0
    P = np.array([o.angl2proj(raw[r]) for r in range(o.N)])
    drho = np.array([o.proj2angl(P[r]) for r in range(o.N)])
    f1, _ = o.excess(_xi_from_drho(o, drho.ravel()))
    # rotate: new(x, y) = old(R^-1 (x, y)) with R (x, y) = (-y, x)
    Pg = P.reshape(n, n, n, o.np_)
    ix = np.arange(n)
    Prot = np.empty_like(Pg)
    for x in range(n):
        for y in range(n):
            # R^-1 (x, y) = (y, -x)
            Prot[:, y, x, :] = Pg[:, (-x) % n, y, :]
    phase = np.exp(1j * o.pmup * np.pi / 2)
    Prot = Prot * phase[None, None, None, :]
    drho2 = np.array([o.proj2angl(p) for p in Prot.reshape(o.N, o.np_)])
    f2, _ = o.excess(_xi_from_drho(o, drho2.ravel()))
    assert abs(f2 - f1) < 1e-10 * abs(f1)
    assert ix.size == n


def test_stages_consistent(oracle_mod):
    o, _ = _setup(oracle_mod, 2, n=6)
    rng = np.random.default_rng(3)
    xi = 1.0 + 0.1 * rng.uniform(-1, 1, o.N * o.no)
    s1 = o.excess_stage(xi, 1)
    s2 = o.excess_stage(xi, 2)
    drho = o.rho0 * (xi.reshape(o.N, o.no) ** 2 - 1.0)
    for r in (0, 17, o.N - 1):
        assert np.allclose(s1[r], o.angl2proj(drho[r]), atol=1e-15)
    f = s1.reshape(6, 6, 6, o.np_)
    for ip in range(o.np_):
        ref = np.fft.ifftn(f[..., ip]) * f[..., ip].size
        assert np.allclose(s2[:, ip].reshape(6, 6, 6), ref, atol=1e-12)


def _excess_hessian_extremes(oracle_mod, mmax, n, table):
    """Extreme eigenvalues of the excess Hessian in drho, in the metric of the
    ideal term's curvature at rho0 (kT dv w_io / rho0), by Lanczos on
    F_exc's exact gradient (F_exc is quadratic in drho)."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    from mdft.grid import angular_grid
    o = oracle_mod.Oracle((n, n, n), (9.0, 9.0, 9.0), mmax, 2, 2.4777, 0.0333)
    o.set_ck(table)
    wq = angular_grid(mmax, 2)["w"]
    dm = 1.0 / np.sqrt(np.tile(2.4777 * o.dv * wq / o.rho0, o.N))

    def hv(v):
        u = dm * np.asarray(v).ravel()
        s = 1e-4 * o.rho0 / max(np.abs(u).max(), 1e-300)
        xi = np.sqrt(1.0 + s * u / o.rho0)
        _, df = o.excess(xi)
        return dm * (df * o.dv / (2.0 * o.rho0 * xi)) / s

    op = LinearOperator((o.N * o.no,) * 2, matvec=hv, dtype=np.float64)
    lo = eigsh(op, k=1, which="SA", tol=1e-7, return_eigenvectors=False)[0]
    hi = eigsh(op, k=1, which="LA", tol=1e-7, return_eigenvectors=False)[0]
    return lo, hi


def test_self_term_is_the_identity_and_makes_excess_convex(oracle_mod):
    """On an odd grid the table I = delta_mn delta_{mu,-nu} (-1)^khi gives an
    excess Hessian of exactly minus the projector (eigenvalues in [-1, 0] in the
    ideal-curvature metric); the synthetic table with its self term gives a
    positive semi-definite excess Hessian (F_exc >= 0 for every drho), while the
    random part alone is indefinite."""
    mmax, n = 2, 5
    ident = ckmod.synthetic_ck(mmax, 256, 0.05, seed=4, n0=0.0333, self_term=0.0)
    ident.c[:] = 0.0
    ckmod.add_self_term(ident, 1.0)
    lo, hi = _excess_hessian_extremes(oracle_mod, mmax, n, ident)
    assert abs(lo + 1.0) < 1e-6 and abs(hi) < 1e-6, (lo, hi)
    raw = ckmod.synthetic_ck(mmax, 256, 0.05, seed=4, n0=0.0333, self_term=0.0)
    lo, hi = _excess_hessian_extremes(oracle_mod, mmax, n, raw)
    assert lo < -1e-3 and hi > 1e-3, (lo, hi)
    full = ckmod.synthetic_ck(mmax, 256, 0.05, seed=4, n0=0.0333)
    lo, hi = _excess_hessian_extremes(oracle_mod, mmax, n, full)
    assert lo > -1e-7, (lo, hi)

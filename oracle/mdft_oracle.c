/*
 * mdft_oracle.c -- FP64 CPU restatement of the reference energy_and_gradient
 * path (dborgis/mdft-dev_general).  TEST INFRASTRUCTURE ONLY: it is the
 * checker the HIP library is compared against, never part of the product.
 *
 * Every function cites the reference file:line it restates.  The FFTs of the
 * reference (FFTW 3.3.4) are replaced by an own mixed-radix DFT that follows
 * the DFT definition; the 2D r2c/c2r over (psi,phi) in
 * module_orientation_projection_transform are written as the explicit sums
 * they compute.
 */
#include "mdft_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

struct or_ctx {
    int nx, ny, nz;
    double lx, ly, lz, dv;
    int mmax, mrso, ntheta, nphi, npsi, no, np, nmu2;
    double kT, n0, rho0;
    double *theta, *wtheta, *w;
    double *kx, *ky, *kz;
    int *pm, *pmup, *pmu, *pidx; /* pidx[(m*(2L+1)+mup+L)*nmu2 + mu2] */
    double *wig;                 /* wig[ip*ntheta + it] */
    double fm[8];
    /* c(k) after relayout */
    int npc, nq;
    double dq;
    cplx *ck; /* ck[iq*npc + ia] */
    /* orientation phase tables */
    double *cphi, *sphi; /* [nphi] cos/sin(2 pi k/nphi) */
    double *cpsi, *spsi; /* [npsi] */
};

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* module_grid.f90:262-301, Gauss-Legendre roots in cos(theta), weights summing to 1 */
void or_gauss_legendre(int n, double *x, double *w)
{
    const double pi = acos(-1.0);
    int m = (n + 1) / 2;
    for (int i = 1; i <= m; i++) {
        double xi = cos(pi * (i - 0.25) / (n + 0.5));
        double deltaxi = 1.0, pp = 1.0;
        int guard = 0;
        while (fabs(deltaxi) > DBL_EPSILON && guard++ < 200) {
            double p1 = 1.0, p2 = 0.0, p3;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((double)(2 * j - 1) * xi * p2 - (double)(j - 1) * p3) / (double)j;
            }
            pp = n * (xi * p1 - p2) / (xi * xi - 1.0);
            deltaxi = -p1 / pp;
            xi = xi + deltaxi;
        }
        x[i - 1] = xi;
        w[i - 1] = 1.0 / ((1.0 - xi * xi) * pp * pp);
        x[n - i] = -xi;
        w[n - i] = w[i - 1];
    }
}

/* module_mathematica.f90:10-44 (the table is exact up to 13!, which covers mmax <= 6) */
static double fact(int n)
{
    double f = 1.0;
    for (int i = 2; i <= n; i++) f *= (double)i;
    return f;
}

static double ipow_m1(int k) { return (k % 2 == 0) ? 1.0 : -1.0; }

/* module_wigner_d.f90:6-60: d^m_{mu,mup}(theta), Legendre recursion when mu or
 * mup is 0, explicit Wigner sum otherwise */
double or_wigner_small_d(int m, int mu, int mup, double theta)
{
    if (abs(mu) > m || abs(mup) > m) return 0.0;
    if (m == 0) return 1.0;
    if (mu == 0 || mup == 0) {
        int mu0 = mu;
        double theta0 = theta;
        if (mu == 0) {
            mu0 = mup;
            theta0 = -theta;
        }
        double x = 1.0;
        if (mu0 < 0) {
            x = ipow_m1(-mu0);
            mu0 = -mu0;
        }
        double cc = cos(theta0);
        double pm1 = 0.0, pm = 1.0, pm2;
        for (int l = mu0 + 1; l <= m; l++) {
            pm2 = pm1;
            pm1 = pm;
            pm = (cc * (double)(2 * l - 1) * pm1 - (double)(l + mu0 - 1) * pm2) / (double)(l - mu0);
        }
        return x * ipow_m1(mu0) * sqrt(fact(m - mu0) / fact(m + mu0)) * fact(2 * mu0) /
               (pow(2.0, mu0) * fact(mu0)) * pow(sin(theta0), mu0) * pm;
    }
    double s = 0.0;
    double cc = cos(0.5 * theta), ss = sin(0.5 * theta);
    int itmin = (mu - mup > 0) ? mu - mup : 0;
    int itmax = (m + mu < m - mup) ? m + mu : m - mup;
    for (int it = itmin; it <= itmax; it++) {
        s += ipow_m1(it) /
             (fact(m + mu - it) * fact(m - mup - it) * fact(it) * fact(it - mu + mup)) *
             pow(cc, 2 * m + mu - mup - 2 * it) * pow(ss, 2 * it - mu + mup);
    }
    return sqrt(fact(m + mu) * fact(m - mu) * fact(m + mup) * fact(m - mup)) * s;
}

/* module_rotation.f90:13-34 (coefficients a, b, c, d) and :119-247 (R^m_{mup,khi}(q)) */
static double sqrtof(int k) { return (k <= 0) ? 0.0 : sqrt((double)k); }

void or_rotation_matrix(int mmax, const double q[3], cplx *R)
{
    const int L = mmax, W = 2 * mmax + 1;
    const int total = (mmax + 1) * W * W;
#define IDX(l, m, m1) (((l) * W + (m) + L) * W + (m1) + L)
    if (mmax == 0) {
        for (int i = 0; i < total; i++) R[i] = 1.0;
        return;
    }
    double *f = calloc((size_t)total, sizeof(double));
    double *g = calloc((size_t)total, sizeof(double));
    f[IDX(0, 0, 0)] = 1.0;
    double r3[3], r2[3], r1[3];
    if (q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0) {
        r3[0] = 0.0; r3[1] = 0.0; r3[2] = 1.0;
    } else {
        r3[0] = q[0]; r3[1] = q[1]; r3[2] = q[2];
    }
    if (r3[0] != 0.0 || r3[1] != 0.0) { /* cross(z, r3) */
        r2[0] = 0.0 * r3[2] - 1.0 * r3[1];
        r2[1] = 1.0 * r3[0] - 0.0 * r3[2];
        r2[2] = 0.0 * r3[1] - 0.0 * r3[0];
    } else { /* cross(r3, x) */
        r2[0] = r3[1] * 0.0 - r3[2] * 0.0;
        r2[1] = r3[2] * 1.0 - r3[0] * 0.0;
        r2[2] = r3[0] * 0.0 - r3[1] * 1.0;
    }
    double n3 = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]);
    double n2 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
    for (int i = 0; i < 3; i++) { r3[i] /= n3; r2[i] /= n2; }
    r1[0] = r2[1] * r3[2] - r2[2] * r3[1];
    r1[1] = r2[2] * r3[0] - r2[0] * r3[2];
    r1[2] = r2[0] * r3[1] - r2[1] * r3[0];
    /* rmat(i,j): component i of vector j (columns r1, r2, r3) */
#define RM(i, j) ((j) == 1 ? r1[(i)-1] : ((j) == 2 ? r2[(i)-1] : r3[(i)-1]))
    const double s2 = sqrt(2.0);
    f[IDX(1, -1, -1)] = (RM(2, 2) + RM(1, 1)) * 0.5;
    f[IDX(1, -1, 0)] = RM(1, 3) / s2;
    f[IDX(1, -1, 1)] = (RM(2, 2) - RM(1, 1)) * 0.5;
    f[IDX(1, 0, -1)] = RM(3, 1) / s2;
    f[IDX(1, 0, 0)] = RM(3, 3);
    f[IDX(1, 0, 1)] = -RM(3, 1) / s2;
    f[IDX(1, 1, -1)] = (RM(2, 2) - RM(1, 1)) * 0.5;
    f[IDX(1, 1, 0)] = -RM(1, 3) / s2;
    f[IDX(1, 1, 1)] = (RM(2, 2) + RM(1, 1)) * 0.5;
    g[IDX(1, -1, -1)] = (RM(2, 1) - RM(1, 2)) * 0.5;
    g[IDX(1, -1, 0)] = RM(2, 3) / s2;
    g[IDX(1, -1, 1)] = (-RM(2, 1) - RM(1, 2)) * 0.5;
    g[IDX(1, 0, -1)] = -RM(3, 2) / s2;
    g[IDX(1, 0, 0)] = 0.0;
    g[IDX(1, 0, 1)] = -RM(3, 2) / s2;
    g[IDX(1, 1, -1)] = (RM(2, 1) + RM(1, 2)) * 0.5;
    g[IDX(1, 1, 0)] = RM(2, 3) / s2;
    g[IDX(1, 1, 1)] = (RM(1, 2) - RM(2, 1)) * 0.5;
#undef RM
    const double f10 = f[IDX(1, 0, 0)], fm0 = f[IDX(1, -1, 0)], fp0 = f[IDX(1, 1, 0)];
    const double gm0 = g[IDX(1, -1, 0)], gp0 = g[IDX(1, 1, 0)];
    const double f01 = f[IDX(1, 0, 1)], fm1 = f[IDX(1, -1, 1)], fp1 = f[IDX(1, 1, 1)];
    const double g01 = g[IDX(1, 0, 1)], gm1 = g[IDX(1, -1, 1)], gp1 = g[IDX(1, 1, 1)];
    for (int l = 2; l <= mmax; l++) {
        const int l1 = l - 1;
#define A(m, m1) (sqrtof(l + (m)) * sqrtof(l - (m)) / (sqrtof(l + (m1)) * sqrtof(l - (m1))))
#define B(m, m1) (sqrtof(l + (m)) * sqrtof(l + (m)-1) / (s2 * sqrtof(l + (m1)) * sqrtof(l - (m1))))
#define C(m) (s2 * sqrtof(l + (m)) * sqrtof(l - (m)) / (sqrtof(l + l) * sqrtof(l + l - 1)))
#define D(m) (sqrtof(l + (m)) * sqrtof(l + (m)-1) / (sqrtof(l + l) * sqrtof(l + l - 1)))
#define F(a, b, c) (((a) == l1 && abs(b) <= mmax) ? f[IDX(a, b, c)] : 0.0)
#define G(a, b, c) (((a) == l1 && abs(b) <= mmax) ? g[IDX(a, b, c)] : 0.0)
        for (int m = -l; m <= l; m++) {
            int m1min = (m > 0) ? 1 : 0;
            for (int m1 = m1min; m1 <= l - 1; m1++) {
                double fv, gv;
                if (m == -l) {
                    fv = B(-m, m1) * (fm0 * F(l1, m + 1, m1) - gm0 * G(l1, m + 1, m1));
                    gv = B(-m, m1) * (fm0 * G(l1, m + 1, m1) + gm0 * F(l1, m + 1, m1));
                } else if (m == l) {
                    fv = B(m, m1) * (fp0 * F(l1, m - 1, m1) - gp0 * G(l1, m - 1, m1));
                    gv = B(m, m1) * (fp0 * G(l1, m - 1, m1) + gp0 * F(l1, m - 1, m1));
                } else {
                    fv = A(m, m1) * (f10 * F(l1, m, m1)) +
                         B(m, m1) * (fp0 * F(l1, m - 1, m1) - gp0 * G(l1, m - 1, m1)) +
                         B(-m, m1) * (fm0 * F(l1, m + 1, m1) - gm0 * G(l1, m + 1, m1));
                    gv = A(m, m1) * (f10 * G(l1, m, m1)) +
                         B(m, m1) * (fp0 * G(l1, m - 1, m1) + gp0 * F(l1, m - 1, m1)) +
                         B(-m, m1) * (fm0 * G(l1, m + 1, m1) + gm0 * F(l1, m + 1, m1));
                }
                f[IDX(l, m, m1)] = fv;
                g[IDX(l, m, m1)] = gv;
                f[IDX(l, -m, -m1)] = ipow_m1(abs(m + m1)) * f[IDX(l, m, m1)];
                g[IDX(l, -m, -m1)] = -ipow_m1(abs(m + m1)) * g[IDX(l, m, m1)];
            }
            const int m1 = l;
            double fv, gv;
            if (m == -l) {
                fv = D(-m) * (fm1 * F(l1, m + 1, m1 - 1) - gm1 * G(l1, m + 1, m1 - 1));
                gv = D(-m) * (fm1 * G(l1, m + 1, m1 - 1) + gm1 * F(l1, m + 1, m1 - 1));
            } else if (m == l) {
                fv = D(m) * (fp1 * F(l1, m - 1, m1 - 1) - gp1 * G(l1, m - 1, m1 - 1));
                gv = D(m) * (fp1 * G(l1, m - 1, m1 - 1) + gp1 * F(l1, m - 1, m1 - 1));
            } else {
                fv = C(m) * (f01 * F(l1, m, m1 - 1) - g01 * G(l1, m, m1 - 1)) +
                     D(m) * (fp1 * F(l1, m - 1, m1 - 1) - gp1 * G(l1, m - 1, m1 - 1)) +
                     D(-m) * (fm1 * F(l1, m + 1, m1 - 1) - gm1 * G(l1, m + 1, m1 - 1));
                gv = C(m) * (f01 * G(l1, m, m1 - 1) + g01 * F(l1, m, m1 - 1)) +
                     D(m) * (fp1 * G(l1, m - 1, m1 - 1) + gp1 * F(l1, m - 1, m1 - 1)) +
                     D(-m) * (fm1 * G(l1, m + 1, m1 - 1) + gm1 * F(l1, m + 1, m1 - 1));
            }
            f[IDX(l, m, m1)] = fv;
            g[IDX(l, m, m1)] = gv;
            f[IDX(l, -m, -m1)] = ipow_m1(abs(m + m1)) * f[IDX(l, m, m1)];
            g[IDX(l, -m, -m1)] = -ipow_m1(abs(m + m1)) * g[IDX(l, m, m1)];
        }
#undef A
#undef B
#undef C
#undef D
#undef F
#undef G
    }
    for (int i = 0; i < total; i++) R[i] = CMPLX(f[i], g[i]);
    free(f);
    free(g);
#undef IDX
}

static int kidx_mq(int i, int n) { return (i == 0) ? 0 : n - i; } /* module_grid.f90:384-418 */

/* module_grid.f90:319-331 */
static double kproj(int l0, int n, double len)
{
    const double twopi = 2.0 * acos(-1.0);
    int l = l0 + 1;
    if (l <= n / 2 + 1) return (double)(l - 1) * twopi / len;
    return (double)(l - 1 - n) * twopi / len;
}

or_ctx *or_create(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso,
                  double kT, double n0)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || mmax < 0 || mmax > 6 || mrso <= 0) return NULL;
    or_ctx *c = calloc(1, sizeof(or_ctx));
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->lx = lx; c->ly = ly; c->lz = lz;
    c->dv = (lx / nx) * (ly / ny) * (lz / nz); /* grid.f90:84,90 product(dl) */
    c->mmax = mmax; c->mrso = mrso;
    c->ntheta = mmax + 1; /* grid.f90:96-99 */
    c->nphi = 2 * mmax + 1;
    c->npsi = 2 * (mmax / mrso) + 1;
    c->no = c->ntheta * c->nphi * c->npsi;
    c->nmu2 = mmax / mrso + 1;
    int np = 0;
    for (int m = 0; m <= mmax; m++) np += (2 * m + 1) * (m / mrso + 1); /* grid.f90:102 */
    c->np = np;
    c->kT = kT;
    c->n0 = n0;
    c->rho0 = n0 / (8.0 * acos(-1.0) * acos(-1.0) / mrso); /* solvent.f90:222 */
    /* angular quadrature, grid.f90:116-139 */
    c->theta = malloc(sizeof(double) * c->ntheta);
    c->wtheta = malloc(sizeof(double) * c->ntheta);
    or_gauss_legendre(c->ntheta, c->theta, c->wtheta);
    for (int i = 0; i < c->ntheta; i++) c->theta[i] = acos(c->theta[i]);
    c->w = malloc(sizeof(double) * c->no);
    const double eightpisq = 8.0 * acos(-1.0) * acos(-1.0);
    const double wphi = 1.0 / (double)c->nphi, wpsi = 1.0 / (double)(c->npsi * mrso);
    for (int it = 0, io = 0; it < c->ntheta; it++)
        for (int iphi = 0; iphi < c->nphi; iphi++)
            for (int ipsi = 0; ipsi < c->npsi; ipsi++, io++)
                c->w[io] = c->wtheta[it] * wphi * wpsi * eightpisq;
    /* k tables */
    c->kx = malloc(sizeof(double) * nx);
    c->ky = malloc(sizeof(double) * ny);
    c->kz = malloc(sizeof(double) * nz);
    for (int l = 0; l < nx; l++) c->kx[l] = kproj(l, nx, lx);
    for (int l = 0; l < ny; l++) c->ky[l] = kproj(l, ny, ly);
    for (int l = 0; l < nz; l++) c->kz[l] = kproj(l, nz, lz);
    /* projection tables, cproj_mrso.f90:151-175 */
    c->pm = malloc(sizeof(int) * np);
    c->pmup = malloc(sizeof(int) * np);
    c->pmu = malloc(sizeof(int) * np);
    const int W = 2 * mmax + 1;
    c->pidx = malloc(sizeof(int) * (mmax + 1) * W * c->nmu2);
    for (int i = 0; i < (mmax + 1) * W * c->nmu2; i++) c->pidx[i] = -1;
    int ip = 0;
    for (int m = 0; m <= mmax; m++)
        for (int mup = -m; mup <= m; mup++)
            for (int mu = 0; mu <= m; mu += mrso) {
                c->pidx[(m * W + mup + mmax) * c->nmu2 + mu / mrso] = ip;
                c->pm[ip] = m;
                c->pmup[ip] = mup;
                c->pmu[ip] = mu;
                ip++;
            }
    /* Wigner table, cproj_mrso.f90:224-235 (call order m, mup, mu) */
    c->wig = malloc(sizeof(double) * np * c->ntheta);
    for (int p = 0; p < np; p++)
        for (int it = 0; it < c->ntheta; it++)
            c->wig[p * c->ntheta + it] = or_wigner_small_d(c->pm[p], c->pmup[p], c->pmu[p], c->theta[it]);
    /* sqrt(2m+1), orientation_projection_transform.f90:35 */
    for (int m = 0; m < 7; m++) c->fm[m] = sqrt((double)(2 * m + 1));
    c->fm[7] = sqrt(14.0);
    c->cphi = malloc(sizeof(double) * c->nphi);
    c->sphi = malloc(sizeof(double) * c->nphi);
    c->cpsi = malloc(sizeof(double) * c->npsi);
    c->spsi = malloc(sizeof(double) * c->npsi);
    const double twopi = 2.0 * acos(-1.0);
    for (int k = 0; k < c->nphi; k++) {
        c->cphi[k] = cos(twopi * k / c->nphi);
        c->sphi[k] = sin(twopi * k / c->nphi);
    }
    for (int k = 0; k < c->npsi; k++) {
        c->cpsi[k] = cos(twopi * k / c->npsi);
        c->spsi[k] = sin(twopi * k / c->npsi);
    }
    return c;
}

void or_destroy(or_ctx *c)
{
    if (!c) return;
    free(c->theta); free(c->wtheta); free(c->w);
    free(c->kx); free(c->ky); free(c->kz);
    free(c->pm); free(c->pmup); free(c->pmu); free(c->pidx);
    free(c->wig); free(c->ck);
    free(c->cphi); free(c->sphi); free(c->cpsi); free(c->spsi);
    free(c);
}

void or_sizes(const or_ctx *c, int *out)
{
    out[0] = c->ntheta; out[1] = c->nphi; out[2] = c->npsi; out[3] = c->no;
    out[4] = c->np; out[5] = c->npc; out[6] = c->nq;
}

void or_tables(const or_ctx *c, double *theta, double *wtheta, double *w)
{
    memcpy(theta, c->theta, sizeof(double) * c->ntheta);
    memcpy(wtheta, c->wtheta, sizeof(double) * c->ntheta);
    memcpy(w, c->w, sizeof(double) * c->no);
}

void or_proj_tables(const or_ctx *c, int *m, int *mup, int *mu)
{
    memcpy(m, c->pm, sizeof(int) * c->np);
    memcpy(mup, c->pmup, sizeof(int) * c->np);
    memcpy(mu, c->pmu, sizeof(int) * c->np);
}

void or_k_tables(const or_ctx *c, double *kx, double *ky, double *kz)
{
    memcpy(kx, c->kx, sizeof(double) * c->nx);
    memcpy(ky, c->ky, sizeof(double) * c->ny);
    memcpy(kz, c->kz, sizeof(double) * c->nz);
}

static inline int P(const or_ctx *c, int m, int mup, int mu2)
{
    return c->pidx[(m * (2 * c->mmax + 1) + mup + c->mmax) * c->nmu2 + mu2];
}

/* read_c_luc.f90:101 (nq), cproj_mrso.f90:366 (qmax), :381-401 (relayout),
 * :415 (conj), :420-429 (sign folding) */
int or_set_ck(or_ctx *c, int np_file, const int *m, const int *n, const int *mu, const int *nu,
              const int *khi, int nq_file, double dq, const double complex *c_file)
{
    const int L = c->mmax, mrso = c->mrso, H = L / mrso;
    double kxm = c->kx[0], kym = c->ky[0], kzm = c->kz[0];
    for (int i = 0; i < c->nx; i++) if (c->kx[i] > kxm) kxm = c->kx[i];
    for (int i = 0; i < c->ny; i++) if (c->ky[i] > kym) kym = c->ky[i];
    for (int i = 0; i <= c->nz / 2; i++) if (c->kz[i] > kzm) kzm = c->kz[i];
    double qmax = sqrt(kxm * kxm + kym * kym + kzm * kzm);
    int nq = (int)(qmax / dq + 0.5) + 2;
    if (nq > nq_file) return -2;
    /* hash of the file tuples: (m, n, mu2, nu2, khi) -> i */
    const int W = 2 * L + 1, WH = 2 * H + 1;
    int *map = malloc(sizeof(int) * (L + 1) * (L + 1) * WH * WH * W);
    for (int i = 0; i < (L + 1) * (L + 1) * WH * WH * W; i++) map[i] = -1;
#define MAPI(a, b, c_, d, e) (((((a) * (L + 1) + (b)) * WH + (c_) + H) * WH + (d) + H) * W + (e) + L)
    for (int i = 0; i < np_file; i++) {
        if (m[i] < 0 || m[i] > L || n[i] < 0 || n[i] > L || abs(khi[i]) > L) continue;
        if (abs(mu[i] / mrso) > H || abs(nu[i] / mrso) > H) continue;
        map[MAPI(m[i], n[i], mu[i] / mrso, nu[i] / mrso, khi[i])] = i;
    }
    int npc = 0;
    for (int mm = 0; mm <= L; mm++)
        for (int k = -mm; k <= mm; k++)
            for (int mu2 = 0; mu2 <= mm / mrso; mu2++)
                for (int nn = abs(k); nn <= L; nn++)
                    npc += 2 * (nn / mrso) + 1;
    free(c->ck);
    c->ck = malloc(sizeof(cplx) * (size_t)npc * nq);
    int ia = 0;
    for (int mm = 0; mm <= L; mm++)
        for (int k = -mm; k <= mm; k++)
            for (int mu2 = 0; mu2 <= mm / mrso; mu2++)
                for (int nn = abs(k); nn <= L; nn++)
                    for (int nu2 = -nn / mrso; nu2 <= nn / mrso; nu2++) {
                        int src = map[MAPI(mm, nn, mu2, nu2, k)];
                        if (src < 0) {
                            free(map);
                            return -3;
                        }
                        int nuv = nu2 * mrso;
                        double sgn = (nuv < 0) ? ipow_m1(abs(k + nuv)) : ipow_m1(nn);
                        for (int iq = 0; iq < nq; iq++)
                            c->ck[(size_t)iq * npc + ia] = sgn * conj(c_file[(size_t)iq * np_file + src]);
                        ia++;
                    }
#undef MAPI
    free(map);
    c->npc = npc;
    c->nq = nq;
    c->dq = dq;
    return 0;
}

void or_get_ck(const or_ctx *c, double complex *out)
{
    memcpy(out, c->ck, sizeof(cplx) * (size_t)c->npc * c->nq);
}

/* orientation_projection_transform.f90:159-199 */
void or_angl2proj(const or_ctx *c, const double *o, double complex *p)
{
    const int nth = c->ntheta, nphi = c->nphi, npsi = c->npsi, L = c->mmax, H = L / c->mrso;
    const int W = 2 * L + 1;
    cplx f[7 * 13 * 7];
    const double norm = 1.0 / (double)(nphi * npsi);
    for (int it = 0; it < nth; it++)
        for (int mup = -L; mup <= L; mup++)
            for (int mu2 = 0; mu2 <= H; mu2++) {
                /* conj(r2c)/(nphi npsi) = sum o exp(+i(mup phi + mu psi)) / (nphi npsi) */
                double re = 0.0, im = 0.0;
                for (int iphi = 0; iphi < nphi; iphi++) {
                    int kph = ((mup * iphi) % nphi + nphi) % nphi;
                    for (int ipsi = 0; ipsi < npsi; ipsi++) {
                        int kps = (mu2 * ipsi) % npsi;
                        double cr = c->cphi[kph] * c->cpsi[kps] - c->sphi[kph] * c->spsi[kps];
                        double ci = c->cphi[kph] * c->spsi[kps] + c->sphi[kph] * c->cpsi[kps];
                        double v = o[(it * nphi + iphi) * npsi + ipsi];
                        re += v * cr;
                        im += v * ci;
                    }
                }
                f[(it * W + mup + L) * (H + 1) + mu2] = CMPLX(re * norm, im * norm);
            }
    for (int ip = 0; ip < c->np; ip++) {
        int m = c->pm[ip], mup = c->pmup[ip], mu2 = c->pmu[ip] / c->mrso;
        cplx s = 0.0;
        for (int it = 0; it < nth; it++)
            s += f[(it * W + mup + L) * (H + 1) + mu2] * c->wig[ip * nth + it] * c->wtheta[it];
        p[ip] = s * c->fm[m];
    }
}

/* orientation_projection_transform.f90:202-243; the c2r of conj(f) is the
 * real part of the Hermitian-weighted sum (weight 1 for mu=0, 2 otherwise) */
void or_proj2angl(const or_ctx *c, const double complex *p, double *o)
{
    const int nth = c->ntheta, nphi = c->nphi, npsi = c->npsi, L = c->mmax, mrso = c->mrso;
    const int H = L / mrso, W = 2 * L + 1;
    cplx f[7 * 13 * 7];
    for (int i = 0; i < nth * W * (H + 1); i++) f[i] = 0.0;
    for (int mup = -L; mup <= L; mup++)
        for (int mu2 = 0; mu2 <= H; mu2++) {
            int mlo = abs(mup) > mrso * mu2 ? abs(mup) : mrso * mu2;
            for (int m = mlo; m <= L; m++) {
                int ip = P(c, m, mup, mu2);
                for (int it = 0; it < nth; it++)
                    f[(it * W + mup + L) * (H + 1) + mu2] += p[ip] * c->wig[ip * nth + it] * c->fm[m];
            }
        }
    for (int it = 0; it < nth; it++)
        for (int iphi = 0; iphi < nphi; iphi++)
            for (int ipsi = 0; ipsi < npsi; ipsi++) {
                double s = 0.0;
                for (int mup = -L; mup <= L; mup++) {
                    int kph = ((mup * iphi) % nphi + nphi) % nphi;
                    for (int mu2 = 0; mu2 <= H; mu2++) {
                        int kps = (mu2 * ipsi) % npsi;
                        double cr = c->cphi[kph] * c->cpsi[kps] - c->sphi[kph] * c->spsi[kps];
                        double ci = c->cphi[kph] * c->spsi[kps] + c->sphi[kph] * c->cpsi[kps];
                        cplx x = conj(f[(it * W + mup + L) * (H + 1) + mu2]);
                        double wgt = (mu2 == 0) ? 1.0 : 2.0;
                        s += wgt * (creal(x) * cr - cimag(x) * ci);
                    }
                }
                o[(it * nphi + iphi) * npsi + ipsi] = s;
            }
}

/* ---- own FFT: recursive mixed radix on the DFT definition ----
 * tw[j] = exp(sign*2*pi*i*j/N) for the top-level length N; a sub-transform of
 * length n reads its twiddles with step N/n. */
static void dft1(cplx *x, int n, int stride, cplx *tmp, const cplx *tw, int tstep)
{
    if (n == 1) return;
    int r = 0;
    if (n % 4 == 0) r = 4;
    else if (n % 2 == 0) r = 2;
    else if (n % 3 == 0) r = 3;
    else if (n % 5 == 0) r = 5;
    if (r == 0) { /* plain DFT for a remaining prime factor */
        for (int k = 0; k < n; k++) {
            cplx s = 0.0;
            for (int j = 0; j < n; j++) s += x[j * stride] * tw[(int)(((long long)j * k) % n) * tstep];
            tmp[k] = s;
        }
        for (int k = 0; k < n; k++) x[k * stride] = tmp[k];
        return;
    }
    int m = n / r;
    cplx *sub = tmp;         /* n entries */
    cplx *scratch = tmp + n; /* scratch for the recursion */
    for (int s = 0; s < r; s++)
        for (int j = 0; j < m; j++) sub[s * m + j] = x[(j * r + s) * stride];
    for (int s = 0; s < r; s++) dft1(sub + s * m, m, 1, scratch, tw, tstep * r);
    for (int k = 0; k < n; k++) {
        cplx acc = 0.0;
        for (int s = 0; s < r; s++) acc += sub[s * m + (k % m)] * tw[(int)(((long long)s * k) % n) * tstep];
        x[k * stride] = acc;
    }
}

static cplx *twiddles(int n, int sign)
{
    const double twopi = 2.0 * acos(-1.0);
    cplx *t = malloc(sizeof(cplx) * n);
    for (int j = 0; j < n; j++) {
        double a = sign * twopi * (double)j / n;
        t[j] = CMPLX(cos(a), sin(a));
    }
    return t;
}

void or_fft3d(double complex *a, int nx, int ny, int nz, int sign)
{
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    cplx *twx = twiddles(nx, sign), *twy = twiddles(ny, sign), *twz = twiddles(nz, sign);
#pragma omp parallel
    {
        cplx *line = malloc(sizeof(cplx) * nmax);
        cplx *tmp = malloc(sizeof(cplx) * nmax * 4);
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; iz++)
            for (int iy = 0; iy < ny; iy++) {
                cplx *b = a + ((size_t)iz * ny + iy) * nx;
                for (int i = 0; i < nx; i++) line[i] = b[i];
                dft1(line, nx, 1, tmp, twx, 1);
                for (int i = 0; i < nx; i++) b[i] = line[i];
            }
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; iz++)
            for (int ix = 0; ix < nx; ix++) {
                cplx *b = a + (size_t)iz * ny * nx + ix;
                for (int i = 0; i < ny; i++) line[i] = b[(size_t)i * nx];
                dft1(line, ny, 1, tmp, twy, 1);
                for (int i = 0; i < ny; i++) b[(size_t)i * nx] = line[i];
            }
#pragma omp for collapse(2) schedule(static)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                cplx *b = a + (size_t)iy * nx + ix;
                for (int i = 0; i < nz; i++) line[i] = b[(size_t)i * nx * ny];
                dft1(line, nz, 1, tmp, twz, 1);
                for (int i = 0; i < nz; i++) b[(size_t)i * nx * ny] = line[i];
            }
        free(line);
        free(tmp);
    }
    free(twx); free(twy); free(twz);
}

/* cproj_mrso.f90:276-290: o = rho0 (xi^2 - 1), then angl2proj per point */
static void stage_angl2proj(const or_ctx *c, const double *xi, cplx *proj)
{
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    const int no = c->no, np = c->np;
#pragma omp parallel
    {
        double *o = malloc(sizeof(double) * no);
#pragma omp for schedule(static)
        for (size_t r = 0; r < N; r++) {
            for (int io = 0; io < no; io++) {
                double x = xi[r * no + io];
                o[io] = c->rho0 * (x * x - 1.0);
            }
            or_angl2proj(c, o, proj + r * np);
        }
        free(o);
    }
}

/* batched 3D FFT over the np projections stored as proj[r*np + ip]
 * (cproj_mrso.f90:306-329 sign +1; :690-719 sign -1 then *1/N) */
static void stage_fft(const or_ctx *c, cplx *proj, int sign)
{
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    const int np = c->np;
    cplx *buf = malloc(sizeof(cplx) * N);
    const double cst = 1.0 / (double)(c->nx * c->ny * c->nz);
    for (int ip = 0; ip < np; ip++) {
        for (size_t r = 0; r < N; r++) buf[r] = proj[r * np + ip];
        or_fft3d(buf, c->nx, c->ny, c->nz, sign);
        if (sign < 0)
            for (size_t r = 0; r < N; r++) proj[r * np + ip] = buf[r] * cst;
        else
            for (size_t r = 0; r < N; r++) proj[r * np + ip] = buf[r];
    }
    free(buf);
}

/* FLT_EPSILON chop of R, cproj_mrso.f90:491-492.  cmplx() without a KIND is a
 * default (single) precision complex, so the kept component is rounded to FP32. */
static inline cplx chop_R(cplx z)
{
    double re = creal(z), im = cimag(z);
    if (fabs(re) < (double)FLT_EPSILON) { re = 0.0; im = (double)(float)im; }
    if (fabs(im) < (double)FLT_EPSILON) { im = 0.0; re = (double)(float)re; }
    return CMPLX(re, im);
}

/* DBL_EPSILON chop of gamma, cproj_mrso.f90:618-623 (same default-kind rounding) */
static inline cplx chop_gamma(cplx z)
{
    double re = creal(z), im = cimag(z);
    if (fabs(re) < DBL_EPSILON && fabs(im) < DBL_EPSILON) { re = 0.0; im = 0.0; }
    if (fabs(re) < DBL_EPSILON) { re = 0.0; im = (double)(float)im; }
    if (fabs(im) < DBL_EPSILON) { im = 0.0; re = (double)(float)re; }
    return CMPLX(re, im);
}

/* One (q, -q) pair: cproj_mrso.f90:490-661 */
static void oz_pair(const or_ctx *c, cplx *proj, int ixq, int iyq, int izq, cplx *R, cplx *dq_,
                    cplx *dmq, cplx *gq, cplx *gmq, cplx *ceff)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, L = c->mmax, mrso = c->mrso, np = c->np;
    const int W = 2 * L + 1;
    const int ixm = kidx_mq(ixq, nx), iym = kidx_mq(iyq, ny), izm = kidx_mq(izq, nz);
    const int q_eq_mq = (ixm == ixq && iym == iyq && izm == izq);
    const double q[3] = {c->kx[ixq], c->ky[iyq], c->kz[izq]};
    cplx *Xq = proj + ((size_t)(izq * ny + iyq) * nx + ixq) * np;
    cplx *Xmq = proj + ((size_t)(izm * ny + iym) * nx + ixm) * np;
#define RR(m, a, b) R[((m) * W + (a) + L) * W + (b) + L]
    or_rotation_matrix(L, q, R);
    for (int i = 0; i < (L + 1) * W * W; i++) R[i] = chop_R(R[i]);
    /* rotation to the q frame, :505-523 */
    int ip = 0;
    for (int m = 0; m <= L; m++)
        for (int khi = -m; khi <= m; khi++)
            for (int mu2 = 0; mu2 <= m / mrso; mu2++) {
                cplx a = 0.0, b = 0.0;
                for (int mup = -m; mup <= m; mup++) {
                    a += Xq[P(c, m, mup, mu2)] * RR(m, mup, khi);
                    b += Xmq[P(c, m, mup, mu2)] * RR(m, mup, -khi);
                }
                dq_[ip] = a;
                dmq[ip] = b * ipow_m1(m);
                ip++;
            }
    /* linear interpolation of c in |q|, :534-539 */
    double effiq = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) / c->dq + 1.0;
    int iq = (int)effiq;
    double alpha = effiq - (double)iq;
    const cplx *c0 = c->ck + (size_t)(iq - 1) * c->npc;
    const cplx *c1 = c->ck + (size_t)iq * c->npc;
    for (int ia = 0; ia < c->npc; ia++) ceff[ia] = alpha * c1[ia] + (1.0 - alpha) * c0[ia];
    /* molecular OZ, :561-606 */
    int ia = 0;
    for (ip = 0; ip < np; ip++) {
        int khi = c->pmup[ip];
        cplx sq = 0.0, smq = 0.0;
        for (int n = abs(khi); n <= L; n++)
            for (int nu2 = -n / mrso; nu2 <= n / mrso; nu2++) {
                if (nu2 < 0) {
                    int j = P(c, n, khi, -nu2);
                    sq += ceff[ia] * dq_[j];
                    smq += ceff[ia] * dmq[j];
                } else {
                    int j = P(c, n, khi, nu2);
                    sq += ceff[ia] * conj(dmq[j]);
                    smq += ceff[ia] * conj(dq_[j]);
                }
                ia++;
            }
        gq[ip] = chop_gamma(sq);
        gmq[ip] = chop_gamma(smq);
    }
    /* rotation back with conj(R), :611-641 */
    ip = 0;
    for (int m = 0; m <= L; m++)
        for (int mup = -m; mup <= m; mup++)
            for (int mu2 = 0; mu2 <= m / mrso; mu2++) {
                cplx a = 0.0, b = 0.0;
                for (int khi = -m; khi <= m; khi++) {
                    a += gq[P(c, m, khi, mu2)] * conj(RR(m, mup, khi));
                    b += gmq[P(c, m, khi, mu2)] * conj(RR(m, mup, -khi));
                }
                dq_[ip] = a;
                dmq[ip] = b * ipow_m1(m);
                ip++;
            }
#undef RR
    /* store, :648-656 (for q == -q the second store wins) */
    for (ip = 0; ip < np; ip++) Xq[ip] = dq_[ip];
    /* 1-based ix_q == nx/2+1 etc. (:652) */
    const int nyq = (ixq == nx / 2) || (iyq == ny / 2) || (izq == nz / 2);
    if (q_eq_mq && nyq)
        for (ip = 0; ip < np; ip++) Xmq[ip] = conj(dmq[ip]);
    else
        for (ip = 0; ip < np; ip++) Xmq[ip] = dmq[ip];
}

/* cproj_mrso.f90:441-674.  The reference iterates iz = 1..nz/2+1, iy, ix and
 * skips a point once both it and its -q partner are done: the pair is
 * handled from the first member met in that order. */
static void stage_oz(const or_ctx *c, cplx *proj)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, L = c->mmax, np = c->np;
    const int W = 2 * L + 1;
#pragma omp parallel
    {
        cplx *R = malloc(sizeof(cplx) * (L + 1) * W * W);
        cplx *a = malloc(sizeof(cplx) * np), *b = malloc(sizeof(cplx) * np);
        cplx *g1 = malloc(sizeof(cplx) * np), *g2 = malloc(sizeof(cplx) * np);
        cplx *ceff = malloc(sizeof(cplx) * (c->npc > 0 ? c->npc : 1));
#pragma omp for schedule(dynamic, 1)
        for (int izq = 0; izq <= nz / 2; izq++) {
            for (int iyq = 0; iyq < ny; iyq++)
                for (int ixq = 0; ixq < nx; ixq++) {
                    int izm = kidx_mq(izq, nz), iym = kidx_mq(iyq, ny), ixm = kidx_mq(ixq, nx);
                    if (izm == izq) {
                        /* plane paired with itself: the first of the pair in (iy, ix) order */
                        long self = (long)iyq * nx + ixq, part = (long)iym * nx + ixm;
                        if (part < self) continue;
                    }
                    oz_pair(c, proj, ixq, iyq, izq, R, a, b, g1, g2, ceff);
                }
        }
        free(R); free(a); free(b); free(g1); free(g2); free(ceff);
    }
}

void or_excess_stage(const or_ctx *c, const double *xi, int stage, double complex *out)
{
    stage_angl2proj(c, xi, out);
    if (stage >= 2) stage_fft(c, out, +1);
    if (stage >= 3) stage_oz(c, out);
    if (stage >= 4) stage_fft(c, out, -1);
}

/* cproj_mrso.f90:61-782 */
void or_excess(const or_ctx *c, const double *xi, double *fexc, double *df)
{
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    const int no = c->no, np = c->np;
    cplx *proj = malloc(sizeof(cplx) * N * np);
    or_excess_stage(c, xi, 4, proj);
    const double fourpisq = 4.0 * acos(-1.0) * acos(-1.0);
    const double prefactor = -c->kT * 2.0 * fourpisq / c->mrso / c->n0; /* :728 */
    /* per-thread running sums (static schedule), added in thread order: the same result on
       every call with the same thread count */
    const int nt = omp_get_max_threads();
    double *part = calloc((size_t)nt, sizeof(double));
#pragma omp parallel
    {
        double *v = malloc(sizeof(double) * no);
        double ffp = 0.0;
#pragma omp for schedule(static)
        for (size_t r = 0; r < N; r++) {
            or_proj2angl(c, proj + r * np, v);
            double s = 0.0;
            for (int io = 0; io < no; io++) {
                double x = xi[r * no + io];
                double vexc = prefactor * c->w[io] * v[io];
                s += (x * x * c->rho0 - c->rho0) * vexc;
                if (df) df[r * no + io] += 2.0 * c->rho0 * x * vexc;
            }
            ffp += s;
        }
        part[omp_get_thread_num()] = ffp;
        free(v);
    }
    double ff = 0.0;
    for (int t = 0; t < nt; t++) ff += part[t];
    free(part);
    *fexc = ff * 0.5 * c->dv;
    free(proj);
}

/* module_energy_ideal_and_external.f90:14-74 */
void or_ideal_ext(const or_ctx *c, const double *xi, const double *vext, double *fid, double *fext,
                  double *df)
{
    const size_t N = (size_t)c->nx * c->ny * c->nz;
    const int no = c->no;
    const double kT = c->kT, dv = c->dv, rho0 = c->rho0;
    /* per-thread running sums (static schedule), added in thread order: the same result on
       every call with the same thread count */
    const int nt = omp_get_max_threads();
    double *pa = calloc(2 * (size_t)nt, sizeof(double)), *pb = pa + nt;
#pragma omp parallel
    {
        double a = 0.0, b = 0.0;
#pragma omp for schedule(static)
        for (size_t r = 0; r < N; r++)
            for (int io = 0; io < no; io++) {
                double x = xi[r * no + io];
                double rho = x * x * rho0;
                if (rho > DBL_EPSILON) {
                    a += kT * dv * c->w[io] * (rho * log(rho / rho0) - rho + rho0);
                    b += dv * c->w[io] * rho * vext[r * no + io];
                    if (df)
                        df[r * no + io] += c->w[io] * 2.0 * rho0 * x * (kT * log(rho / rho0) + vext[r * no + io]);
                } else {
                    a += kT * dv * c->w[io] * rho0;
                }
            }
        pa[omp_get_thread_num()] = a;
        pb[omp_get_thread_num()] = b;
    }
    double a = 0.0, b = 0.0;
    for (int t = 0; t < nt; t++) {
        a += pa[t];
        b += pb[t];
    }
    free(pa);
    *fid = a;
    *fext = b;
}

/* module_energy_and_gradient.f90:36-240 (bridge "no", neutral-solute path) */
void or_energy_and_gradient(const or_ctx *c, const double *xi, const double *vext, double *parts,
                            double *df)
{
    const size_t n = (size_t)c->nx * c->ny * c->nz * c->no;
    if (df) memset(df, 0, sizeof(double) * n);
    or_ideal_ext(c, xi, vext, &parts[0], &parts[1], df);
    or_excess(c, xi, &parts[2], df);
}

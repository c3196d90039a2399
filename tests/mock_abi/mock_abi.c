/*
 * mock_abi.c -- TEST INFRASTRUCTURE: a CPU implementation of the subset of the
 * C-ABI (include/mdft_gpu.h) that the Fortran host library calls, backed by the
 * FP64 oracle.  Linking the Fortran L-BFGS (mdft-dev_general_amd/fortran)
 * against it lets the host logic run and be checked without a GPU.  "Device"
 * pointers are plain host memory here.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mdft_gpu.h"
#include "mdft_oracle.h"

struct mdft_ctx {
    or_ctx *o;
    double *vext;
    size_t nvar;
    int no, np;
    char err[256];
};

static char g_err[256] = "";

int mdft_ctx_create(const mdft_desc *d, mdft_ctx **out)
{
    or_ctx *o = or_create(d->nx, d->ny, d->nz, d->lx, d->ly, d->lz, d->mmax, d->mrso, d->kT, d->n0);
    if (!o) { strcpy(g_err, "bad description"); return MDFT_EINVAL; }
    mdft_ctx *c = calloc(1, sizeof(mdft_ctx));
    int s[7];
    or_sizes(o, s);
    c->o = o;
    c->no = s[3];
    c->np = s[4];
    c->nvar = (size_t)d->nx * d->ny * d->nz * s[3];
    *out = c;
    return 0;
}

void mdft_ctx_destroy(mdft_ctx *c)
{
    if (!c) return;
    or_destroy(c->o);
    free(c->vext);
    free(c);
}

const char *mdft_last_error(const mdft_ctx *c) { return c ? c->err : g_err; }

int mdft_get_info(const mdft_ctx *c, int *info)
{
    int s[7];
    or_sizes(c->o, s);
    for (int i = 0; i < 7; i++) info[i] = s[i];
    info[7] = (int)(c->nvar / (size_t)c->no);
    info[8] = 0;
    info[9] = 2;
    return 0;
}

int mdft_set_ck(mdft_ctx *c, int np_file, const int *m, const int *n, const int *mu, const int *nu, const int *khi,
                int nq_file, double dq, const double *cf)
{
    return or_set_ck(c->o, np_file, m, n, mu, nu, khi, nq_file, dq, (const double complex *)cf) ? MDFT_ERANGE : 0;
}

int mdft_compute_vext(mdft_ctx *c, int natom, const double *atoms, int nsite, const double *sites, double rcut)
{
    (void)natom; (void)atoms; (void)nsite; (void)sites; (void)rcut;
    strcpy(c->err, "compute_vext is not available in the CPU mock");
    return MDFT_EINVAL;
}

int mdft_set_vext(mdft_ctx *c, const double *vext, int flags)
{
    (void)flags;
    free(c->vext);
    c->vext = malloc(sizeof(double) * c->nvar);
    memcpy(c->vext, vext, sizeof(double) * c->nvar);
    return 0;
}

/* s(io) = sqrt(w(io) / mean w), the scaled-variable factor */
static double *scale_table(mdft_ctx *c)
{
    int s[7];
    or_sizes(c->o, s);
    double *th = malloc(sizeof(double) * s[0]), *wt = malloc(sizeof(double) * s[0]);
    double *w = malloc(sizeof(double) * c->no);
    or_tables(c->o, th, wt, w);
    double wm = 0.0;
    for (int io = 0; io < c->no; io++) wm += w[io];
    wm /= c->no;
    for (int io = 0; io < c->no; io++) w[io] = sqrt(w[io] / wm);
    free(th);
    free(wt);
    return w;
}

int mdft_energy_and_gradient(mdft_ctx *c, const double *xi, double *parts, double *df, int flags)
{
    double *sc = (flags & MDFT_SCALED_VARS) ? scale_table(c) : NULL;
    const double *x = xi;
    double *xs = NULL;
    if (sc || (flags & MDFT_XI_FP32)) {
        xs = malloc(sizeof(double) * c->nvar);
        for (size_t i = 0; i < c->nvar; i++) {
            const double v = (flags & MDFT_XI_FP32) ? (double)((const float *)xi)[i] : xi[i];
            xs[i] = sc ? v / sc[i % c->no] : v;
        }
        x = xs;
    }
    double *tmp = df ? malloc(sizeof(double) * c->nvar) : NULL;
    or_energy_and_gradient(c->o, x, c->vext, parts, tmp);
    if (df)
        for (size_t i = 0; i < c->nvar; i++) {
            const double g = sc ? tmp[i] / sc[i % c->no] : tmp[i];
            if (flags & MDFT_DF_FP32) ((float *)df)[i] = (float)g;
            else df[i] = g;
        }
    free(tmp);
    free(xs);
    free(sc);
    return 0;
}

int mdft_energy_and_gradient_dots(mdft_ctx *c, const double *xi, double *parts, void *df, int flags, int k,
                                  const void *const *vecs, double *dots)
{
    int rc = mdft_energy_and_gradient(c, xi, parts, (double *)df, flags);
    if (rc) return rc;
    for (int j = 0; j < k; j++) {
        double s = 0.0;
        for (size_t i = 0; i < c->nvar; i++) {
            if (flags & MDFT_DF_FP32) s += (double)((const float *)df)[i] * ((const float *)vecs[j])[i];
            else s += ((const double *)df)[i] * ((const double *)vecs[j])[i];
        }
        dots[j] = s;
    }
    return 0;
}

int mdft_scale_variables(mdft_ctx *c, size_t n, int dir, int in_fp32, const void *in, double *out)
{
    double *sc = scale_table(c);
    for (size_t i = 0; i < n; i++) {
        const double v = in_fp32 ? (double)((const float *)in)[i] : ((const double *)in)[i];
        out[i] = dir > 0 ? v * sc[i % c->no] : v / sc[i % c->no];
    }
    free(sc);
    return 0;
}

int mdft_excess_energy(mdft_ctx *c, const double *xi, double *fexc, double *df, int flags)
{
    (void)flags;
    or_excess(c->o, xi, fexc, df);
    return 0;
}

int mdft_device_alloc(mdft_ctx *c, size_t bytes, void **p)
{
    (void)c;
    *p = aligned_alloc(256, (bytes + 255) / 256 * 256);
    return *p ? 0 : MDFT_ENOMEM;
}

int mdft_device_free(mdft_ctx *c, void *p) { (void)c; free(p); return 0; }
int mdft_memcpy_h2d(mdft_ctx *c, void *d, const void *s, size_t b) { (void)c; memcpy(d, s, b); return 0; }
int mdft_memcpy_d2h(mdft_ctx *c, void *d, const void *s, size_t b) { (void)c; memcpy(d, s, b); return 0; }

int mdft_vec_dot(mdft_ctx *c, size_t n, const double *x, const double *y, double *out)
{
    (void)c;
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += x[i] * y[i];
    *out = s;
    return 0;
}

int mdft_vec_axpy_dot(mdft_ctx *c, size_t n, double a, const double *x, double *y, const double *z, double *out)
{
    (void)c;
    double s = 0.0;
    for (size_t i = 0; i < n; i++) {
        y[i] += a * x[i];
        if (z) s += z[i] * y[i];
    }
    if (z && out) *out = s;
    return 0;
}

int mdft_vec_lincomb(mdft_ctx *c, size_t n, double a, const double *x, double b, const double *y, double *z)
{
    (void)c;
    for (size_t i = 0; i < n; i++) z[i] = a * x[i] + (y ? b * y[i] : 0.0);
    return 0;
}

int mdft_vec_multi_dot(mdft_ctx *c, size_t n, int k, const double *const *vecs, const double *x, double *out)
{
    (void)c;
    if (k < 1 || k > 16) return -1;
    for (int j = 0; j < k; j++) {
        double s = 0.0;
        for (size_t i = 0; i < n; i++) s += x[i] * vecs[j][i];
        out[j] = s;
    }
    return 0;
}

int mdft_vec_combine_step(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z,
                          double a, const double *x, double *y)
{
    (void)c;
    if (k < 1 || k > 16) return -1;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < k; j++) s += coef[j] * vecs[j][i];
        z[i] = s;
        y[i] = x[i] + a * s;
    }
    return 0;
}

int mdft_vec_combine(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z)
{
    (void)c;
    if (k < 1 || k > 16) return -1;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < k; j++) s += coef[j] * vecs[j][i];
        z[i] = s;
    }
    return 0;
}

/* present for the link of the Fortran module only; the host tests do not call it */
int mdft_charged_solute_corrections(const double *lengths, double n0, int natom, const double *atoms, int nsite,
                                    const double *sites, double relative_permittivity, double *out)
{
    (void)lengths; (void)n0; (void)natom; (void)atoms; (void)nsite; (void)sites; (void)relative_permittivity;
    out[0] = out[1] = 0.0;
    return 0;
}

/* FP32-stored directions and gradients */
int mdft_vec_multi_dot_f32(mdft_ctx *c, size_t n, int k, const float *const *vecs, const float *x, double *out)
{
    (void)c;
    for (int j = 0; j < k; j++) {
        double s = 0.0;
        for (size_t i = 0; i < n; i++) s += (double)x[i] * vecs[j][i];
        out[j] = s;
    }
    return 0;
}

int mdft_vec_combine_step_f32(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                              double a, const double *x, double *y)
{
    (void)c;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < k; j++) s += coef[j] * vecs[j][i];
        z[i] = (float)s;
        y[i] = x[i] + a * (double)z[i];
    }
    return 0;
}

int mdft_vec_axpy_f32(mdft_ctx *c, size_t n, double a, const float *z, const double *x, double *y)
{
    (void)c;
    for (size_t i = 0; i < n; i++) y[i] = x[i] + a * (double)z[i];
    return 0;
}

int mdft_vec_widen_f32(mdft_ctx *c, size_t n, const float *in, double *out)
{
    (void)c;
    for (size_t i = 0; i < n; i++) out[i] = (double)in[i];
    return 0;
}

int mdft_vec_combine_step_f32x(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                               double a, const float *x, float *y)
{
    (void)c;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < k; j++) s += coef[j] * vecs[j][i];
        z[i] = (float)s;
        y[i] = (float)((double)x[i] + a * (double)z[i]);
    }
    return 0;
}

int mdft_vec_axpy_f32x(mdft_ctx *c, size_t n, double a, const float *z, const float *x, float *y)
{
    (void)c;
    for (size_t i = 0; i < n; i++) y[i] = (float)((double)x[i] + a * (double)z[i]);
    return 0;
}

int mdft_vec_narrow_f64(mdft_ctx *c, size_t n, const double *in, float *out)
{
    (void)c;
    for (size_t i = 0; i < n; i++) out[i] = (float)in[i];
    return 0;
}

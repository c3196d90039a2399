/*
 * mdft_oracle.h -- FP64 CPU restatement of the reference MDFT functional
 * evaluation energy_and_gradient(f, df) (dborgis/mdft-dev_general).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Index conventions (0-based, Fortran column-major order of the reference):
 *   grid point   r  = ix + nx*(iy + ny*iz)
 *   orientation  io = ipsi + npsi*(iphi + nphi*itheta)       (grid.f90:128-135)
 *   projection   ip : m=0..mmax, mup=-m..m, mu=0..m step mrso (cproj_mrso.f90:156-175)
 *   fields       xi[r*no + io], vext[r*no + io], df[r*no + io],
 *                projections p[r*np + ip] (complex, reference deltarho_p(np,nx,ny,nz))
 *   c table      c[iq*npc + ia] after the relayout of cproj_mrso.f90:381-429
 */
#ifndef MDFT_ORACLE_H
#define MDFT_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(int nx, int ny, int nz, double lx, double ly, double lz,
                  int mmax, int mrso, double kT, double n0);
void or_destroy(or_ctx *c);

/* sizes: out[0..6] = ntheta, nphi, npsi, no, np, npc, nq */
void or_sizes(const or_ctx *c, int *out);
/* theta[ntheta], wtheta[ntheta], w[no] */
void or_tables(const or_ctx *c, double *theta, double *wtheta, double *w);
/* projection tables: m[np], mup[np], mu[np] */
void or_proj_tables(const or_ctx *c, int *m, int *mup, int *mu);
/* k tables kx[nx], ky[ny], kz[nz] */
void or_k_tables(const or_ctx *c, double *kx, double *ky, double *kz);

void or_gauss_legendre(int n, double *x, double *w);
double or_wigner_small_d(int m, int mu, int mup, double theta);
/* R[(m*(2L+1) + mup+L)*(2L+1) + khi+L], L = mmax, complex, no chop */
void or_rotation_matrix(int mmax, const double q[3], double complex *R);

/* c(k) in the file order of read_c_luc (np_file tuples, nq_file rows of
 * c_file[iq*np_file + i]).  Returns 0 or a negative error code. */
int or_set_ck(or_ctx *c, int np_file, const int *m, const int *n, const int *mu,
              const int *nu, const int *khi, int nq_file, double dq,
              const double complex *c_file);
/* npc*nq stored table (after relayout, conj and sign folding) */
void or_get_ck(const or_ctx *c, double complex *out);

void or_angl2proj(const or_ctx *c, const double *o, double complex *p);
void or_proj2angl(const or_ctx *c, const double complex *p, double *o);

/* in-place 3D C2C, x fastest; sign=+1 -> exp(+i k.r); unnormalised */
void or_fft3d(double complex *a, int nx, int ny, int nz, int sign);

/* Excess-term stages on xi: stage 1 = after angl2proj, 2 = after forward
 * FFT, 3 = after rotate/OZ/rotate-back, 4 = after inverse FFT.  out[N*np]. */
void or_excess_stage(const or_ctx *c, const double *xi, int stage, double complex *out);

/* F_exc and (if df != NULL) df += (1/dv) dF_exc/dxi. */
void or_excess(const or_ctx *c, const double *xi, double *fexc, double *df);
/* F_id, F_ext and (if df != NULL) df += ... */
void or_ideal_ext(const or_ctx *c, const double *xi, const double *vext,
                  double *fid, double *fext, double *df);
/* parts[0..2] = F_id, F_ext, F_exc; df zeroed then filled (nullable). */
void or_energy_and_gradient(const or_ctx *c, const double *xi, const double *vext,
                            double *parts, double *df);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif

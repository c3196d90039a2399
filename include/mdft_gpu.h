/*
 * mdft_gpu.h -- C-ABI of the MI355X MDFT functional engine.
 *
 * The drop-in boundary is the reference call
 *     subroutine energy_and_gradient(f, df)
 * (src/module_energy_and_gradient.f90:36,63-64), which sums the ideal+external
 * term (src/module_energy_ideal_and_external.f90:14-74) and the excess term
 * bound to the procedure pointer excess_energy (src/module_proc_ptrs.f90:3-21,
 * bound to energy_cproj_mrso at src/module_init_simu.f90:60-73).  Fortran
 * hosts bind these entry points with ISO_C_BINDING (module mdft_gpu in
 * mdft-dev_general_amd/fortran/mdft_gpu.f90).
 *
 * All entry points return 0 on success and a negative MDFT_E* code on error;
 * mdft_last_error() then describes the failure (the reference stops the
 * program with `error stop` in the same situations).  A context belongs to one
 * host thread and runs all its work on one HIP stream of its own.
 *
 * Array layouts follow the reference (Fortran column-major, leftmost fastest):
 *   xi, vext, df : double [nz][ny][nx][no]   (reference xi(no,nx,ny,nz))
 *   orientation index io = ipsi + npsi*(iphi + nphi*itheta)
 */
#ifndef MDFT_GPU_H
#define MDFT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdft_ctx mdft_ctx;

enum {
    MDFT_OK = 0,
    MDFT_EINVAL = -1,   /* invalid argument or inconsistent sizes */
    MDFT_ENOMEM = -2,   /* device allocation failed */
    MDFT_EHIP = -3,     /* HIP runtime error */
    MDFT_ESTATE = -4,   /* call order violated (e.g. evaluation before set_ck) */
    MDFT_ERANGE = -5,   /* c(k) table does not cover the grid's |q| range */
    MDFT_ENODEV = -6    /* no usable GPU */
};

/* Precision of the bulk arithmetic (transforms, FFTs, OZ contraction).  The
 * free-energy reductions, the ideal term and the gradient stay FP64. */
enum { MDFT_PREC_FP64 = 0, MDFT_PREC_FP32 = 1, MDFT_PREC_BF16 = 2 };

/* flags for evaluation calls */
enum {
    MDFT_DEVICE_PTRS = 1, /* xi / df / vext arguments are device pointers */
    MDFT_TIMERS = 2,      /* record per-phase HIP event timings (mdft_phase_times) */
    MDFT_DF_FP32 = 4,     /* mdft_energy_and_gradient only, with MDFT_DEVICE_PTRS and molecular
                             symmetry order 2, FP64 or FP32 contexts: df is a float array that
                             receives dF rounded to FP32 (the minimiser's FP32 vector storage) */
    MDFT_XI_FP32 = 16,    /* mdft_energy_and_gradient only, with MDFT_DEVICE_PTRS, molecular symmetry order 2,
                             mmax <= 5, FP64 or FP32 contexts: xi is a float array (the minimiser's iterate
                             stored in FP32; no reference counterpart) */
    MDFT_SCALED_VARS = 8  /* mdft_energy_and_gradient only, molecular symmetry order 2, FP64 or FP32
                             contexts: xi holds u(io, r) = s(io) xi(io, r) with s = sqrt(w / mean w)
                             (w the quadrature weights) and df receives dF/du = dF/dxi / s -- the
                             minimiser's change of variables, which takes the quadrature weight out
                             of the curvature of every term (no reference counterpart) */
};

typedef struct mdft_desc {
    int nx, ny, nz;        /* grid nodes, reference `boxnod` (module_grid.f90:74) */
    double lx, ly, lz;     /* box lengths in Angstrom, `boxlen` (module_grid.f90:63) */
    int mmax;              /* angular order, `mmax` (module_grid.f90:95), 0..6 */
    int mrso;              /* molecular rotational symmetry order, `molrotsymorder` */
    double kT;             /* kJ/mol (module_thermo.f90:26) */
    double n0;             /* bulk number density, 1/A^3 (module_solvent.f90:221) */
    int precision;         /* MDFT_PREC_* */
    int device;            /* HIP device ordinal */
} mdft_desc;

/* Creates the device state: quadrature, Wigner-d and k tables, the q/-q pair
 * schedule and all field buffers.  Replaces the one-time set-up done inside
 * energy_cproj_mrso (src/module_energy_cproj_mrso.f90:118-267) and grid%init
 * (src/module_grid.f90:51-212). */
int mdft_ctx_create(const mdft_desc *desc, mdft_ctx **out);
void mdft_ctx_destroy(mdft_ctx *ctx);

/* Last error text for ctx (or the last failed create when ctx == NULL). */
const char *mdft_last_error(const mdft_ctx *ctx);

/* info[0..9] = ntheta, nphi, npsi, no, np, npc, nq (0 before set_ck),
 *              nx*ny*nz (low 31 bits), precision, mrso */
int mdft_get_info(const mdft_ctx *ctx, int *info);
/* Angular quadrature: theta[ntheta] (radians), wtheta[ntheta], w[no]
 * (module_grid.f90:116-139). */
int mdft_get_quadrature(const mdft_ctx *ctx, double *theta, double *wtheta, double *w);

/* External potential vext(no,nx,ny,nz) in kJ/mol, already clipped by the
 * caller (module_vext.f90:638-640).  Replaces solvent(1)%vext. */
int mdft_set_vext(mdft_ctx *ctx, const double *vext, int flags);

/* Direct correlation function c^{mn}_{mu nu,khi}(|q|) exactly as read by
 * read_c_luc (src/module_read_c_luc.f90:6-124): np_file index tuples and
 * nq_file rows of complex values c[iq][i] (interleaved re, im).  The library
 * keeps the first nq = int(qmax/dq + 0.5) + 2 rows (read_c_luc.f90:101) and
 * applies the relayout, conjugation and sign folding of
 * src/module_energy_cproj_mrso.f90:381-429. */
int mdft_set_ck(mdft_ctx *ctx, int np_file, const int *m, const int *n, const int *mu,
                const int *nu, const int *khi, int nq_file, double dq, const double *c);

/* energy_and_gradient(f, df): f_parts[0..2] = F_id, F_ext, F_exc (kJ/mol).
 * df (nullable: F only, as in module_pressure_correction.f90:124) receives
 * the full gradient with the reference convention df = (1/dv) dF/dxi; it is
 * overwritten, not accumulated. */
int mdft_energy_and_gradient(mdft_ctx *ctx, const double *xi, double *f_parts, double *df, int flags);
/* The same with the products of the new gradient and k <= 3 device vectors:
 * dots[j] = df . vecs[j] (host), vecs[j] FP32 with MDFT_DF_FP32, else FP64;
 * vecs[j] == df gives |df|^2.  An FP32 context forms them in its final pass while
 * the gradient is written (the minimiser's multi-dot without its own pass over
 * the gradient); an FP64 context runs the multi-dot after the evaluation.  Needs
 * MDFT_DEVICE_PTRS, molecular symmetry order 2 and an FP64 or FP32 context.  No
 * reference counterpart. */
int mdft_energy_and_gradient_dots(mdft_ctx *ctx, const double *xi, double *f_parts, void *df, int flags, int k,
                                  const void *const *vecs, double *dots);

/* excess_energy_t plug-in (src/module_proc_ptrs.f90:3-21): F_exc only;
 * df (nullable) is accumulated into, as energy_cproj_mrso does (:739). */
int mdft_excess_energy(mdft_ctx *ctx, const double *xi, double *fexc, double *df, int flags);

/* energy_ideal_and_external(fid, fext, df) (module_energy_ideal_and_external.f90:14);
 * df (nullable) is accumulated into. */
int mdft_ideal_external(mdft_ctx *ctx, const double *xi, double *fid, double *fext, double *df, int flags);

/* Intermediate projection field of the excess term for parity checks:
 * stage 1 after angl2proj, 2 after the forward FFTs, 3 after rotate/OZ/rotate
 * back, 4 after the inverse FFTs.  out: complex double [nz][ny][nx][np]. */
int mdft_excess_stage(mdft_ctx *ctx, const double *xi, int stage, double *out, int flags);

/* HIP-event timings (ms) of the phases of the last evaluation run with
 * MDFT_TIMERS: [0] angl2proj, [1] forward FFT, [2] OZ, [3] inverse FFT,
 * [4] proj2angl+ideal+gradient, [5] reductions, [6] total; returns the
 * number of entries written (<= n). */
int mdft_phase_times(const mdft_ctx *ctx, double *ms, int n);

/* External potential of a solute computed on the device into the context
 * (replaces init_vext, src/module_vext.f90:14-115, for the Lennard-Jones part
 * of src/module_lennardjones.f90:9-226 and a minimum-image Coulomb term),
 * clipped at qfact/2 (src/module_vext.f90:638-640).
 * atoms[natom][6] = x, y, z (A, box frame), q (e), sigma (A), epsilon (kJ/mol);
 * sites[nsite][6] = x, y, z (A, molecular frame), q, sigma, epsilon. */
int mdft_compute_vext(mdft_ctx *ctx, int natom, const double *atoms, int nsite, const double *sites,
                      double rcut);
/* Electrostatics of mdft_compute_vext_ex. */
enum { MDFT_ELEC_POISSON = 0, MDFT_ELEC_DIRECT = 1, MDFT_ELEC_NONE = 2 };
/* The same with a choice of electrostatics:
 * MDFT_ELEC_POISSON -- the reference's default (direct_sum = F): the fast Poisson
 *   solver of src/module_fastpoissonsolver.f90:121-288 on the device -- cloud-in-cell
 *   solute charges (src/module_solute.f90:193-232), V(k) = 4 pi rho(k) / k^2, and per
 *   orientation the c2r of V(k) conj(sigma_io(k)) with the smeared solvent charge
 *   density sigma_io (src/module_solvent.f90:367-426), times qfact / N;
 * MDFT_ELEC_DIRECT -- minimum-image Coulomb sum (mdft_compute_vext);
 * MDFT_ELEC_NONE -- Lennard-Jones only.
 * The clip at qfact/2 follows the sum of the terms (src/module_vext.f90:600-669). */
int mdft_compute_vext_ex(mdft_ctx *ctx, int natom, const double *atoms, int nsite, const double *sites,
                         double rcut, int electrostatics);
/* Copies the context's vext(no,nx,ny,nz) out (host or, with MDFT_DEVICE_PTRS, device). */
int mdft_get_vext(mdft_ctx *ctx, double *out, int flags);

/* Solvation free energy corrections of src/module_pressure_correction.f90:7-156
 * (bridge-free branch) for the density xi (host, or device with MDFT_DEVICE_PTRS):
 * out[0] solvent molecules in the box, sum_r sum_io w rho0 xi^2 dv
 * out[1] the same without solute, n0 V;  out[2] dN = out[1] - out[0]
 * out[3] partial molar volume dN / n0 (A^3);  out[4] cavity volume, points with n(r)/n0 < 1e-3 (A^3)
 * out[5] the functional of the empty box (xi = 0, kJ/mol);  out[6] bulk pressure out[5] / V (kJ/mol/A^3)
 * out[7] PMV correction -dN/n0 * P (kJ/mol);  out[8] ideal-gas correction dN kT (kJ/mol).
 * The reference's estimate is F + out[7] + the LJ tail correction.  Needs c(k). */
int mdft_pressure_correction(mdft_ctx *ctx, const double *xi, int flags, double *out);
/* Long-range LJ correction of src/module_lennardjones.f90:145-149 (kJ/mol): atoms and
 * sites as for mdft_compute_vext, rcut the LJ cutoff (A). */
int mdft_lj_tail_correction(mdft_ctx *ctx, int natom, const double *atoms, int nsite, const double *sites,
                            double rcut, double *out);
/* Hunenberger's corrections for a solute with a net charge Q (|Q| > 1e-7; else both 0),
 * which the reference adds to F inside energy_and_gradient when the electrostatics are the
 * lattice (Poisson) ones (src/module_energy_and_gradient.f90:190-203): out[0] the type B
 * periodicity correction -1971.01 Q^2 / L (1 - 1 / eps_r), L the box side (the mean side
 * of a box that is not cubic) (:287-315); out[1] the type C P-scheme correction
 * -tr(Q_solvent) n0 2909.857 Q, tr(Q_solvent) = sum over axes of sum_s q_s r_s,i^2, each
 * term at or below 1e-10 taken as 0 (:321-344, src/module_solvent.f90:323-326).  kJ/mol.
 * lengths: the box (A); atoms and sites as for mdft_compute_vext.  Host arithmetic only:
 * no context, no device. */
int mdft_charged_solute_corrections(const double *lengths, double n0, int natom, const double *atoms, int nsite,
                                    const double *sites, double relative_permittivity, double *out);

/* FP64 vector algebra on device vectors of length n (16-byte aligned device
 * pointers), for a device-resident minimiser: the dot/axpy/copy/scal loops of
 * src/module_blas_etc.f90:8-184 over n = no*nx*ny*nz. */
int mdft_vec_dot(mdft_ctx *ctx, size_t n, const double *x, const double *y, double *out);
/* y += a*x; if z and out are non-null, *out = z.y (one pass) */
int mdft_vec_axpy_dot(mdft_ctx *ctx, size_t n, double a, const double *x, double *y, const double *z,
                      double *out);
/* z = a*x + b*y (y nullable: z = a*x); z may alias x or y */
int mdft_vec_lincomb(mdft_ctx *ctx, size_t n, double a, const double *x, double b, const double *y, double *z);
/* One pass over x and k <= 16 device vectors: out[j] = x.vecs[j] (out on the host). */
int mdft_vec_multi_dot(mdft_ctx *ctx, size_t n, int k, const double *const *vecs, const double *x, double *out);
/* mdft_vec_multi_dot without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  out[j] = x.vecs[j] for
 * 1 <= k <= 16 device vectors (out on the host), the same kernels and results as
 * mdft_vec_multi_dot; the per-block partial products live in a fixed 512 KiB device
 * array of the library (one per GPU, set up by mdft_ctx_create, held until the process
 * ends) instead of memory mdft_vec_multi_dot allocates on its first call.  k outside
 * 1..16 is rejected with MDFT_EINVAL.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (several contexts, streams or threads on one device) race. */
int mdft_vec_multi_dot_noalloc(mdft_ctx *ctx, size_t n, int k, const double *const *vecs, const double *x,
                               double *out);
/* One pass: z = sum_j coef[j] vecs[j], k <= 16 (z may be one of vecs). */
int mdft_vec_combine(mdft_ctx *ctx, size_t n, int k, const double *coef, const double *const *vecs, double *z);
/* One pass: z = sum_j coef[j] vecs[j] and y = x + a z, k <= 16 -- the new search
 * direction and the first trial point of the minimiser's line search (the
 * dcopy/daxpy pair of src/module_lbfgs_nocedal_mdft.f90:1207-1336 fused); y must
 * not alias x, z or any vecs[j]. */
int mdft_vec_combine_step(mdft_ctx *ctx, size_t n, int k, const double *coef, const double *const *vecs, double *z,
                          double a, const double *x, double *y);

/* The same passes with the minimiser's directions and gradients stored in FP32 (half
 * the bytes; FP64 accumulation; the iterate x, y stays FP64).  No reference
 * counterpart: the reference's L-BFGS vectors are FP64.  Device pointers, 16-byte
 * aligned. */
/* out[j] = x.vecs[j], FP64 sums of FP32 products (out on the host). */
int mdft_vec_multi_dot_f32(mdft_ctx *ctx, size_t n, int k, const float *const *vecs, const float *x, double *out);
/* z = (float) sum_j coef[j] vecs[j] and y = x + a z with z as stored. */
int mdft_vec_combine_step_f32(mdft_ctx *ctx, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                              double a, const double *x, double *y);
/* y = x + a z (z FP32; y may alias x). */
int mdft_vec_axpy_f32(mdft_ctx *ctx, size_t n, double a, const float *z, const double *x, double *y);
/* out = (double) in. */
int mdft_vec_widen_f32(mdft_ctx *ctx, size_t n, const float *in, double *out);
/* The minimiser's passes with the iterate itself stored in FP32 as well (MDFT_XI_FP32
 * evaluations): z = (float) sum_j coef[j] vecs[j] and y = (float)(x + a z), x and y FP32,
 * y not aliasing x; and y = (float)(x + a z), y may alias x.  No reference counterpart. */
int mdft_vec_combine_step_f32x(mdft_ctx *ctx, size_t n, int k, const double *coef, const float *const *vecs,
                               float *z, double a, const float *x, float *y);
int mdft_vec_axpy_f32x(mdft_ctx *ctx, size_t n, double a, const float *z, const float *x, float *y);
/* out = (float) in. */
int mdft_vec_narrow_f64(mdft_ctx *ctx, size_t n, const double *in, float *out);
/* The change of variables of MDFT_SCALED_VARS over n = (points) x no values: dir = 1,
 * out = s in (xi -> u, or dF/du -> dF/dxi); dir = -1, out = in / s (u -> xi).  in is
 * FP32 when in_fp32 (a gradient stored in FP32), else FP64; out may alias an FP64 in.
 * No reference counterpart. */
int mdft_scale_variables(mdft_ctx *ctx, size_t n, int dir, int in_fp32, const void *in, double *out);

/* ---- restart (density.bin) support, all device pointers ----
 * angl2proj / proj2angl of src/module_orientation_projection_transform.f90:159-199 and
 * :202-243 for npoints points (any count: a restart file's grid may differ from the
 * context's): d_f [npoints][no] real, d_proj [npoints][np] complex (interleaved FP64),
 * the context's orientation tables.  The reader of src/module_density.f90:61-261 rebuilds
 * xi with proj2angl from the projections the writer of src/module_postprocessing.f90:46-94
 * stored with angl2proj.  Molecular symmetry order 2. */
int mdft_angl2proj(mdft_ctx *ctx, const double *d_f, size_t npoints, double *d_proj);
int mdft_proj2angl(mdft_ctx *ctx, const double *d_proj, size_t npoints, double *d_f);
/* Trilinear regridding of src/module_density.f90:131-133,171-226: d_prev [n_prev][no]
 * (x fastest) on a grid of n_prev[3] nodes with spacing d_spacing[3], centred in the
 * context's box, onto d_xi [N][no]; points of the context's grid outside the previous
 * grid's nodes get `outside` (the reference leaves them at their initial xi = 1).  The
 * cell that writes a point is the reference's: the last of its loop order whose index
 * range ceiling(x0/d + 1) .. floor(x1/d + 1) contains it. */
int mdft_regrid_trilinear(mdft_ctx *ctx, const double *d_prev, const int *n_prev, const double *d_spacing,
                          double outside, double *d_xi);

/* Device memory owned by the caller, for device-resident fields
 * (MDFT_DEVICE_PTRS) and the device L-BFGS vectors. */
int mdft_device_alloc(mdft_ctx *ctx, size_t bytes, void **dptr);
int mdft_device_free(mdft_ctx *ctx, void *dptr);
int mdft_memcpy_h2d(mdft_ctx *ctx, void *dst, const void *src, size_t bytes);
int mdft_memcpy_d2h(mdft_ctx *ctx, void *dst, const void *src, size_t bytes);
int mdft_synchronize(mdft_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif

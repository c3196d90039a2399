// mdft_internal.h -- device-side context of the MI355X MDFT engine and the
// launchers shared between the translation units of libmdft_gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mdft_gpu.h"

#define MDFT_MAX_MMAX 6
#define MDFT_VEC_MAX 16  // vectors per multi-dot / combine pass
// Floor of rho/rho0 in the density guess of the warm start where the cold start is 1:
// at xi = 0 both gradient terms vanish (they carry the factor 2 rho0 xi), so a point
// predicted empty could never be refilled by the minimiser.
#define MDFT_WS_DENSITY_FLOOR 1e-2
// Lowest beta vext of the Boltzmann start, xi = exp(-max(beta vext, this) / 2): rho/rho0
// stays <= e^3 (next to charged sites beta vext reaches -90, whose Boltzmann factor
// overflows the FP32 excess term; the converged rho/rho0 rarely exceeds 10 there)
#define MDFT_WS_BOLTZMANN_MIN_BV (-3.0)
#define MDFT_FFT_MAX_STAGES 12
#define MDFT_OZ32_ROWS_PER_LANE 4  // OZ rows per lane in the FP32 q-space kernel (np <= 256)

// Angular tables of one context (all device pointers).
struct AngTables {
    int ntheta, nphi, npsi, no, np, L, H, mrso;
    const double *w;     // [no] quadrature weights (module_grid.f90:139)
    const double *a2p;   // [np][ntheta] wigner_d * wtheta
    const double *p2a;   // [np][ntheta] wigner_d * fm(m)
    const double *fm;    // [np] sqrt(2m+1) of each projection
    const double *cph;   // [nphi] cos(2 pi k / nphi)
    const double *sph;   // [nphi] sin(2 pi k / nphi)
    const double *cps;   // [npsi] cos(2 pi k / npsi)
    const double *sps;   // [npsi] sin(2 pi k / npsi)
    const int *pidx;     // [(L+1)][2L+1][H+1] projection index of (m, mup, mu2)
    const int *pm, *pmup, *pmu;  // [np]
};

struct FftAxisPlan {
    int n;                       // transform length
    int nst;                     // number of radix stages
    int radix[MDFT_FFT_MAX_STAGES];
    double2 *tw;                 // device [n]: exp(+2 pi i k / n)
};

// One (q, -q) pair of a q-space tile: grid indices of q and -q, flags (bit 0: q == -q,
// bit 1: Nyquist plane, bit 2: padding entry of a partial tile), the tile's interpolation
// row iq, the frame of q (cos, sin of theta and phi) and alpha = |q| / dq + 1 - iq.
struct TilePair {
    int gq, gmq, flag, iq;
    double ct, st, cp, sp, al, pad;
};

// Schedule of the (q, -q) pairs of the OZ step, grouped in chunks of pairs that
// share the same |q| interpolation interval.
struct PairSchedule {
    int npairs = 0, nchunks = 0;
    int *pq = nullptr;        // linear grid index of q, per pair
    int *pmq = nullptr;       // linear grid index of -q
    int *pflag = nullptr;     // bit0: q == -q, bit1: q on a Nyquist plane
    int *cstart = nullptr;    // first pair of each chunk
    int *ccount = nullptr;    // pairs in each chunk
    int *ciq = nullptr;       // 1-based lower interpolation row iq of each chunk
    int ntiles = 0;           // tiles of at most 8 pairs of one chunk (the matrix-core q-space kernel)
    int *tstart = nullptr;    // first pair of each tile
    int *tinfo = nullptr;     // pair count | iq << 4 of each tile
    TilePair *tpair = nullptr;  // [ntiles][8] the tiles' pairs, padded (flag bit 2 marks padding)
};

struct mdft_ctx {
    mdft_desc d{};
    int ntheta = 0, nphi = 0, npsi = 0, no = 0, np = 0, nmu2 = 0, L = 0, H = 0;
    size_t N = 0;
    double dv = 0, rho0 = 0;
    // host tables
    std::vector<double> theta, wtheta, w, wig;
    std::vector<int> pm, pmup, pmu, pidx;
    std::vector<double> kx, ky, kz;
    // device tables
    AngTables ang{};
    double *d_w = nullptr, *d_a2p = nullptr, *d_p2a = nullptr, *d_fm = nullptr;
    double *d_cph = nullptr, *d_sph = nullptr, *d_cps = nullptr, *d_sps = nullptr;
    int *d_pidx = nullptr, *d_pm = nullptr, *d_pmup = nullptr, *d_pmu = nullptr;
    int *d_ia0 = nullptr;         // [np] first OZ coefficient of each projection row
    double *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr;
    FftAxisPlan fft[3]{};
    // c(k)
    int npc = 0, nq = 0;
    double dq = 0;
    double2 *d_ck = nullptr;      // [nq][npc]
    float *d_oz32_ftab = nullptr; // Fourier coefficients of d^m(theta), FP32 q-space kernel
    int *d_oz32_lanerow = nullptr; // descriptor table of the FP32 q-space kernel (rotation offsets, lane rows, OZ columns)
    void *d_ozm_frag = nullptr;   // matrix-core q-space kernel: c(k) B fragments per interval (FP64 or FP32)
    void *d_ozm_frag16 = nullptr; // BF16 contexts: the same fragments in BF16, 32 k values per instruction
    int *d_ozm_col = nullptr;     // OZ column entries of each khi block
    void *d_ozm_ftab = nullptr;   // its Fourier coefficients of d^m(theta)
    int *d_ozm_units = nullptr;   // its rotation unit rounds [round][8] int4
    int *d_ozm_rlen = nullptr;    // iterations of each round
    int ozm_cr0[2][4][4] = {}, ozm_cnr[2][4][4] = {};  // first round and count per (direction, wave, class)
    std::vector<int> ia0_h;       // first OZ coefficient of each projection row (host)
    void *d_bf16_M = nullptr;     // angl2proj matrix, BF16 MFMA fragment order
    void *d_bf16_Q = nullptr;     // proj2angl matrix, BF16 MFMA fragment order
    PairSchedule pairs;
    // fields
    double *d_vext = nullptr;     // [N][no]
    void *d_proj = nullptr;       // [N][np] double2 (FP64) or float2 (FP32): reference deltarho_p(np,nx,ny,nz)
    double *d_xi_stage = nullptr; // staging for host xi
    double *d_df_stage = nullptr; // staging for host df
    double *d_partials = nullptr; // block partial sums
    double *d_slices = nullptr;   // [3 * 512] first-pass sums of d_partials
    double *d_sums = nullptr;     // [8] final reductions
    double *d_vpart = nullptr;    // vector-op block partials
    double *d_mpart = nullptr;    // multi-dot block partials [blocks][MDFT_VEC_MAX] + result
    float *df32 = nullptr;        // set during an evaluation with MDFT_DF_FP32: dF goes here, rounded to FP32
    double *d_vscale = nullptr;   // [2 no]: s(io) = sqrt(w(io) / mean w) and 1 / s(io) (scaled variables)
    const double *vsinv = nullptr;  // set during an evaluation with MDFT_SCALED_VARS: 1 / s
    const float *xi32 = nullptr;    // set during an evaluation with MDFT_XI_FP32: the FP32 iterate
    // set during mdft_energy_and_gradient_dots: products of dF with up to 3 vectors in the final pass
    int dot_k = 0, dot_self = 0;
    const void *dot_vec[3] = {nullptr, nullptr, nullptr};
    double *d_dpart = nullptr;    // [blocks][3] their block partials
    float *d_vext32 = nullptr;    // FP32 copy of vext read by the fast minimiser's final pass
    void *d_a2p_img[2] = {nullptr, nullptr};  // angl2proj_t's LDS table image, FP32 and FP64
    bool vext32_ok = false;       // d_vext32 holds the current vext
    size_t dpart_cap = 0;
    float *d_chscale = nullptr;   // per-channel input scale of the warm-start features [2 np]
    float *d_adjscale = nullptr;  // per-channel factor of the adjoint of proj2angl [2 np]
    double *h_sums = nullptr;     // pinned mirror of d_sums
    size_t partials_cap = 0;
    bool have_vext = false, have_ck = false;
    hipStream_t stream = nullptr;
    size_t lds_limit = 65536;      // max dynamic LDS per workgroup on this device
    int num_cus = 0;               // compute units of the device
    hipEvent_t ev[8]{};
    bool timers_ready = false;
    double phase_ms[8]{};
    int nphase = 0;
    std::string err;
};

// ---- launchers (return hipError_t of the launch) ----
hipError_t launch_angl2proj_f64(mdft_ctx *c, const double *xi, double2 *proj);
// mode: 0 = energy_and_gradient (df overwritten with ideal+exc gradient),
//       1 = excess only (df accumulated), 2 = ideal only (df accumulated)
hipError_t launch_final_f64(mdft_ctx *c, const double2 *proj, const double *xi, double *df, int mode,
                            int *nblocks_out);
hipError_t launch_ideal_f64(mdft_ctx *c, const double *xi, double *df, int *nblocks_out);
hipError_t launch_angl2proj_f32(mdft_ctx *c, const double *xi, float2 *proj);
hipError_t launch_final_f32(mdft_ctx *c, const float2 *proj, const double *xi, double *df, int mode, int *nblocks_out);
hipError_t launch_oz32(mdft_ctx *c, float2 *proj);
int oz32_tables(mdft_ctx *c, std::vector<float> &ftab, std::vector<int> &itab);
hipError_t launch_reduce3(mdft_ctx *c, int nblocks);
hipError_t launch_fft3d_f64(mdft_ctx *c, double2 *proj, int sign, int batch = 0);  // batch 0: np
hipError_t launch_fft3d_f32(mdft_ctx *c, float2 *proj, int sign, int batch = 0);
hipError_t launch_oz_f64(mdft_ctx *c, double2 *proj);
bool ozm_supported(const mdft_ctx *c);
int ozm_setup(mdft_ctx *c);
int ozm_frags(mdft_ctx *c, const std::vector<double2> &tab, int nq, int npc, const std::vector<int> &ia0);
hipError_t launch_ozm_f64(mdft_ctx *c, double2 *proj);
hipError_t launch_ozm_f32(mdft_ctx *c, float2 *proj);
hipError_t launch_widen_f32(mdft_ctx *c, const float2 *in, double2 *out, size_t n);
hipError_t launch_angl2proj_feat(mdft_ctx *c, const double *src, int feat, const double *aux = nullptr);
hipError_t launch_transform_f64(mdft_ctx *c, const double *f, double2 *proj, size_t npts, bool to_proj);
hipError_t launch_regrid(mdft_ctx *c, const double *prev, int pnx, int pny, const int *d_cell, const double *d_wgt,
                         double outside, double *xi);
hipError_t launch_proj_to_channels(mdft_ctx *c, float *out, float scale, const float *chscale = nullptr);
hipError_t launch_channels_to_proj(mdft_ctx *c, const float *in);
hipError_t launch_reduce3_buf(mdft_ctx *c, const double *part, int nblocks, double *out);
hipError_t launch_cold_start(mdft_ctx *c, double *xi, bool boltzmann = false);
hipError_t launch_density_mix(mdft_ctx *c, size_t total, double lam, const double *xa, const double *xb,
                              double *out);
int bf16_tables(mdft_ctx *c, const std::vector<double> &a2p, const std::vector<double> &p2a,
                const std::vector<double> &fm);
hipError_t launch_angl2proj_bf16(mdft_ctx *c, const double *xi, float2 *proj);
hipError_t launch_final_bf16(mdft_ctx *c, const float2 *proj, const double *xi, double *df, int mode, int *nblocks_out);
int fft_make_plan(int n, int *radix);
hipError_t vec_dot(mdft_ctx *c, size_t n, const double *x, const double *y, double *out);
hipError_t vec_axpy_dot(mdft_ctx *c, size_t n, double a, const double *x, double *y, const double *z, double *out);
hipError_t vec_lincomb(mdft_ctx *c, size_t n, double a, const double *x, double b, const double *y, double *z);
hipError_t launch_density_integrals(mdft_ctx *c, const double *xi, double cav, int *nblocks_out);
hipError_t vec_multi_dot(mdft_ctx *c, size_t n, int k, const double *const *vecs, const double *x, double *out);
hipError_t vec_combine(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z);
hipError_t vec_combine_step(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z,
                            double a, const double *x, double *y);
hipError_t vec_multi_dot32(mdft_ctx *c, size_t n, int k, const float *const *vecs, const float *x, double *out);
hipError_t vec_scratch_init(int device);
size_t vec_scratch_bytes();
hipError_t vec_multi_dot_noalloc(mdft_ctx *c, size_t n, int k, const double *const *vecs, const double *x,
                                 double *out);
hipError_t vec_combine_step32(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                              double a, const double *x, double *y);
hipError_t vec_axpy32(mdft_ctx *c, size_t n, double a, const float *z, const double *x, double *y);
hipError_t vec_widen(mdft_ctx *c, size_t n, const float *in, double *out);
hipError_t vec_combine_step32x(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                               double a, const float *x, float *y);
hipError_t vec_axpy32x(mdft_ctx *c, size_t n, double a, const float *z, const float *x, float *y);
hipError_t vec_narrow(mdft_ctx *c, size_t n, const double *in, float *out);
hipError_t vec_scale_orient(mdft_ctx *c, size_t n, int no, const double *fac, const void *in, bool in_f32, double *out);
hipError_t launch_vext(mdft_ctx *c, int natom, const double *d_atoms, int nsite, const double *d_site_p,
                       const double *d_site_off, double rcut, double qfact, double clip, bool coulomb = true);
hipError_t launch_poisson_vext(mdft_ctx *c, int natom, const double *h_atoms, int nsite, const double *h_site_q,
                               const double *d_site_off, double qfact);
hipError_t launch_vext_clip(mdft_ctx *c, double clip);

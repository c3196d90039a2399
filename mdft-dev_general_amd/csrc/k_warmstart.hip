// k_warmstart.hip -- per-point projections onto the low angular basis and the
// density guess of the learned warm start.
//
// The reference has no learned component.  Basis B_k(Omega), k = 0..3: 1 and
// the lab-frame components of the molecule's z axis (sin t cos p, sin t sin p,
// cos t) -- the m <= 1 projections for molecular symmetry order 2.  The
// quadrature weights and 1 / sum_io w_io B_k^2 are folded into the table Bw
// on the host, so a projection is a plain dot product over orientations.
#include <cfloat>

#include "mdft_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kPts = 16;  // grid points per workgroup (their [kPts][no] tile is staged in LDS)

// mode 0: features of vext (8 channels), mode 1: projections of a field (4 channels)
__global__ __launch_bounds__(kThreads) void ws_project(const double *__restrict__ f, const double *__restrict__ Bw,
                                                      float *__restrict__ out, size_t N, int no, double beta,
                                                      int mode) {
    extern __shared__ double tile[];  // [kPts][no]
    const size_t p0 = (size_t)blockIdx.x * kPts;
    const int npts = (int)((N - p0) < (size_t)kPts ? (N - p0) : (size_t)kPts);
    for (int i = threadIdx.x; i < npts * no; i += kThreads) tile[i] = f[p0 * no + i];
    __syncthreads();
    const int nch = mode == 0 ? 8 : 4;
    for (int o = threadIdx.x; o < npts * nch; o += kThreads) {
        const int pt = o / nch, ch = o % nch, k = ch & 3;
        const double *row = tile + pt * no, *b = Bw + (size_t)k * no;
        double s = 0.0;
        if (mode == 1) {
            for (int io = 0; io < no; io++) s += row[io] * b[io];
        } else if (ch < 4) {
            for (int io = 0; io < no; io++) s += (exp(-0.5 * fmax(beta * row[io], -8.0)) - 1.0) * b[io];
        } else {
            for (int io = 0; io < no; io++) s += tanh(-beta * row[io] / 8.0) * b[io];
        }
        out[(size_t)ch * N + p0 + pt] = (float)s;
    }
}

// xi = exp((-beta vext + sum_k pred_k B_k) / 2), exponent clipped to [-200, 7];
// without pred the reference cold start (src/module_density.f90:267-284):
// xi = 0 where vext >= (-ln(DBL_EPSILON) - 0.05) kT, 1 elsewhere
__global__ __launch_bounds__(kThreads) void ws_guess(const double *__restrict__ vext, const float *__restrict__ pred,
                                                    const double *__restrict__ B, double *__restrict__ xi, size_t N,
                                                    int no, double beta) {
    const size_t total = N * (size_t)no;
    const double vmax = -log(DBL_EPSILON) - 0.05;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        if (!pred) {
            xi[g] = (beta * vext[g] >= vmax) ? 0.0 : 1.0;
            continue;
        }
        const size_t p = g / no;
        const int io = (int)(g - p * no);
        double lg = -beta * vext[g];
#pragma unroll
        for (int k = 0; k < 4; k++) lg += (double)pred[(size_t)k * N + p] * B[(size_t)k * no + io];
        lg = fmin(fmax(lg, -200.0), 7.0);
        xi[g] = exp(0.5 * lg);
    }
}

}  // namespace

hipError_t launch_ws_project(mdft_ctx *c, const double *f, float *out, int mode) {
    const size_t nb = (c->N + kPts - 1) / kPts;
    const size_t lds = sizeof(double) * kPts * c->no;
    if (lds > c->lds_limit) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(ws_project, dim3((unsigned)nb), dim3(kThreads), lds, c->stream, f, c->d_wsBw, out, c->N, c->no,
                       1.0 / c->d.kT, mode);
    return hipGetLastError();
}

hipError_t launch_ws_guess(mdft_ctx *c, const float *pred, double *xi) {
    const size_t total = c->N * (size_t)c->no;
    size_t nb = (total + kThreads - 1) / kThreads;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(ws_guess, dim3((unsigned)nb), dim3(kThreads), 0, c->stream, c->d_vext, pred, c->d_wsB, xi, c->N,
                       c->no, 1.0 / c->d.kT);
    return hipGetLastError();
}

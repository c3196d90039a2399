// k_warmstart.hip -- layout changes between the projection field and the
// network's channel-major fields, and the reference cold start, for the
// learned warm start.
//
// The reference has no learned component.  The network sees the projections
// of a feature field of vext and predicts the projections of the converged
// density (rho/rho0 - 1) or of the OZ field gamma; both live in the context's
// projection layout [N][np] complex (FP64 or FP32) on the MDFT side and as
// 2 np real channels [2 np][N] (FP32, x fastest) on the network side: channel
// 2 ip + 0 holds the real part of projection ip, 2 ip + 1 the imaginary part.
#include <cfloat>

#include "mdft_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 32;  // grid points per workgroup in the transposes

// out[2 ip + c][p] = scale * chscale[2 ip + c] * proj[p][ip].c, chscale optional
// (via an LDS tile: both sides coalesced)
template <typename V>
__global__ __launch_bounds__(kThreads) void proj_to_channels(const V *__restrict__ proj, float *__restrict__ out,
                                                            size_t N, int np, float scale,
                                                            const float *__restrict__ chscale) {
    extern __shared__ float tile[];  // [kTile][2 np]
    const int w2 = 2 * np;
    const size_t p0 = (size_t)blockIdx.x * kTile;
    const int npts = (int)((N - p0) < (size_t)kTile ? (N - p0) : (size_t)kTile);
    const auto *src = reinterpret_cast<const decltype(V::x) *>(proj) + p0 * w2;
    for (int i = threadIdx.x; i < npts * w2; i += kThreads) tile[i] = (float)src[i];
    __syncthreads();
    for (int i = threadIdx.x; i < kTile * w2; i += kThreads) {
        const int ch = i / kTile, pt = i % kTile;
        if (pt < npts) out[(size_t)ch * N + p0 + pt] = (chscale ? scale * chscale[ch] : scale) * tile[pt * w2 + ch];
    }
}

// proj[p][ip] = (in[2 ip][p], in[2 ip + 1][p])
template <typename V>
__global__ __launch_bounds__(kThreads) void channels_to_proj(const float *__restrict__ in, V *__restrict__ proj,
                                                            size_t N, int np) {
    extern __shared__ float tile[];  // [kTile][2 np]
    using T = decltype(V::x);
    const int w2 = 2 * np;
    const size_t p0 = (size_t)blockIdx.x * kTile;
    const int npts = (int)((N - p0) < (size_t)kTile ? (N - p0) : (size_t)kTile);
    for (int i = threadIdx.x; i < kTile * w2; i += kThreads) {
        const int ch = i / kTile, pt = i % kTile;
        if (pt < npts) tile[pt * w2 + ch] = in[(size_t)ch * N + p0 + pt];
    }
    __syncthreads();
    T *dst = reinterpret_cast<T *>(proj) + p0 * w2;
    for (int i = threadIdx.x; i < npts * w2; i += kThreads) dst[i] = (T)tile[i];
}

// reference cold start (src/module_density.f90:267-284):
// xi = 0 where vext >= (-ln(DBL_EPSILON) - 0.05) kT, 1 elsewhere;
// with `boltzmann` set: xi = exp(-max(beta vext, MDFT_WS_BOLTZMANN_MIN_BV) / 2), the density
// of the solvent without its correlations, rho = rho0 exp(-beta vext), capped at e^3 rho0
__global__ __launch_bounds__(kThreads) void cold_start(const double *__restrict__ vext, double *__restrict__ xi,
                                                      size_t total, double beta, int boltzmann) {
    const double vmax = -log(DBL_EPSILON) - 0.05;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const double bv = beta * vext[g];
        xi[g] = boltzmann ? exp(-0.5 * fmax(bv, MDFT_WS_BOLTZMANN_MIN_BV)) : (bv >= vmax ? 0.0 : 1.0);
    }
}

// Trilinear regridding of a restart field (src/module_density.f90:171-226): new point
// (ix, iy, iz) takes the value interpolated in the previous grid's cell (cx, cy, cz) with
// weights (wx, wy, wz), per axis precomputed on the host (cell -1: the point lies outside
// the previous grid and gets `outside`).  Orientation fastest: one thread per (point, io).
struct RegridArgs {
    int nx, ny, nz, no, pnx, pny;
    double outside;
    const int *cell;      // [nx + ny + nz] 0-based cell of each new index along x, y, z, or -1
    const double *wgt;    // [nx + ny + nz] the fractional position inside that cell
};

__global__ __launch_bounds__(kThreads) void regrid_trilinear(const double *__restrict__ prev, double *__restrict__ xi,
                                                            RegridArgs a) {
    const size_t total = (size_t)a.nx * a.ny * a.nz * a.no;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const int io = (int)(g % a.no);
        const size_t r = g / a.no;
        const int ix = (int)(r % a.nx), iy = (int)((r / a.nx) % a.ny), iz = (int)(r / ((size_t)a.nx * a.ny));
        const int cx = a.cell[ix], cy = a.cell[a.nx + iy], cz = a.cell[a.nx + a.ny + iz];
        if (cx < 0 || cy < 0 || cz < 0) {
            xi[g] = a.outside;
            continue;
        }
        const double xd = a.wgt[ix], yd = a.wgt[a.nx + iy], zd = a.wgt[a.nx + a.ny + iz];
        auto p = [&](int dx, int dy, int dz) {
            return prev[(((size_t)(cz + dz) * a.pny + (cy + dy)) * a.pnx + (cx + dx)) * a.no + io];
        };
        const double x00 = p(0, 0, 0) * (1 - xd) + p(1, 0, 0) * xd;
        const double x01 = p(0, 0, 1) * (1 - xd) + p(1, 0, 1) * xd;
        const double x10 = p(0, 1, 0) * (1 - xd) + p(1, 1, 0) * xd;
        const double x11 = p(0, 1, 1) * (1 - xd) + p(1, 1, 1) * xd;
        const double x0 = x00 * (1 - yd) + x10 * yd, x1 = x01 * (1 - yd) + x11 * yd;
        xi[g] = x0 * (1 - zd) + x1 * zd;
    }
}

// out = sqrt((1 - lam) xa^2 + lam xb^2): the density mixture (1 - lam) rho_a + lam rho_b in xi
__global__ __launch_bounds__(kThreads) void density_mix(size_t total, double lam, const double *__restrict__ xa,
                                                       const double *__restrict__ xb, double *__restrict__ out) {
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const double a = xa[g], b = xb[g];
        out[g] = sqrt((1.0 - lam) * a * a + lam * b * b);
    }
}

}  // namespace

hipError_t launch_density_mix(mdft_ctx *c, size_t total, double lam, const double *xa, const double *xb,
                              double *out) {
    size_t nb = (total + kThreads - 1) / kThreads;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(density_mix, dim3((unsigned)nb), dim3(kThreads), 0, c->stream, total, lam, xa, xb, out);
    return hipGetLastError();
}

hipError_t launch_proj_to_channels(mdft_ctx *c, float *out, float scale, const float *chscale) {
    const size_t nb = (c->N + kTile - 1) / kTile;
    const size_t lds = sizeof(float) * kTile * 2 * c->np;
    if (c->d.precision == MDFT_PREC_FP64)
        hipLaunchKernelGGL(proj_to_channels<double2>, dim3((unsigned)nb), dim3(kThreads), lds, c->stream,
                           static_cast<const double2 *>(c->d_proj), out, c->N, c->np, scale, chscale);
    else
        hipLaunchKernelGGL(proj_to_channels<float2>, dim3((unsigned)nb), dim3(kThreads), lds, c->stream,
                           static_cast<const float2 *>(c->d_proj), out, c->N, c->np, scale, chscale);
    return hipGetLastError();
}

hipError_t launch_channels_to_proj(mdft_ctx *c, const float *in) {
    const size_t nb = (c->N + kTile - 1) / kTile;
    const size_t lds = sizeof(float) * kTile * 2 * c->np;
    if (c->d.precision == MDFT_PREC_FP64)
        hipLaunchKernelGGL(channels_to_proj<double2>, dim3((unsigned)nb), dim3(kThreads), lds, c->stream, in,
                           static_cast<double2 *>(c->d_proj), c->N, c->np);
    else
        hipLaunchKernelGGL(channels_to_proj<float2>, dim3((unsigned)nb), dim3(kThreads), lds, c->stream, in,
                           static_cast<float2 *>(c->d_proj), c->N, c->np);
    return hipGetLastError();
}

hipError_t launch_cold_start(mdft_ctx *c, double *xi, bool boltzmann) {
    const size_t total = c->N * (size_t)c->no;
    size_t nb = (total + kThreads - 1) / kThreads;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(cold_start, dim3((unsigned)nb), dim3(kThreads), 0, c->stream, c->d_vext, xi, total,
                       1.0 / c->d.kT, boltzmann ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_regrid(mdft_ctx *c, const double *prev, int pnx, int pny, const int *d_cell, const double *d_wgt,
                         double outside, double *xi) {
    RegridArgs a{c->d.nx, c->d.ny, c->d.nz, c->no, pnx, pny, outside, d_cell, d_wgt};
    const size_t total = c->N * (size_t)c->no;
    size_t nb = (total + kThreads - 1) / kThreads;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(regrid_trilinear, dim3((unsigned)nb), dim3(kThreads), 0, c->stream, prev, xi, a);
    return hipGetLastError();
}

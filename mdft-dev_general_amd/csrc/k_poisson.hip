// k_poisson.hip -- electrostatic part of vext by the reference's fast Poisson
// solver (src/module_fastpoissonsolver.f90:121-288, "new construction"), on
// the device.
//
//  1. the solute point charges are spread onto the 8 surrounding nodes with
//     trilinear (cloud-in-cell) weights, periodic, and divided by the volume
//     element (src/module_solute.f90:193-232 with floorNode / ceilingNode /
//     distToFloorNode of src/module_mathematica.f90:134-220);
//  2. rho(k) = sum_r rho(r) exp(-i k.r) (FFTW's forward r2c), and
//     V(k) = 4 pi rho(k) / k^2, 0 where k^2 <= DBL_EPSILON, with k from the
//     integer frequencies m = i - 1 for i <= n/2, else i - 1 - n (:201-219);
//  3. for each orientation io, vextq(io, r) = qfact / N * c2r[V(k) conj(sigma_io(k))]
//     (:239-249), sigma_io(k) = sum_sites q_s exp(-i k.R_io r_s) exp(-0.5^2 k^2 / 2),
//     the smeared charge density of a solvent molecule in orientation io, with k
//     from the grid's wave vectors (+pi/d on the Nyquist index;
//     src/module_solvent.f90:367-426);
//  4. vext = vext_LJ + vextq, then the clip at qfact/2 (src/module_vext.f90:600-669).
//
// The c2r of FFTW reads the half spectrum kx = 0 .. nx/2 and takes the real part
// of the kx = 0 (and, nx even, kx = nx/2) planes after the y and z transforms.
// Here the same result comes from the complex inverse transform of the half
// spectrum weighted 1 on those planes, 2 on the others and 0 on kx > nx/2,
// keeping the real part.  Orientations are transformed np at a time, each
// chunk a batch of the context's 3D FFT.
#include <cfloat>
#include <map>
#include <vector>

#include "mdft_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr double kSmootherRadius = 0.5;  // src/module_solvent.f90:381

struct PoissonArgs {
    int nx, ny, nz, no, nsite, io0, batch;
    double lx, ly, lz;
    const double *site_q;    // [nsite]
    const double *site_off;  // [no][nsite][3] rotated site positions
};

// X [N] holds (1/N) sum_r rho(r) exp(-i k.r) (the context's inverse-normalised
// transform of the charge density); Z [N][batch] receives the weighted half
// spectrum of V(k) conj(sigma_io(k)) for io = io0 .. io0 + batch - 1.
__global__ __launch_bounds__(kThreads) void poisson_spectrum(const double2 *__restrict__ X, double2 *__restrict__ Z,
                                                            PoissonArgs a) {
    const size_t N = (size_t)a.nx * a.ny * a.nz;
    const size_t total = N * a.batch;
    const double twopi = 2.0 * acos(-1.0), fourpi = 2.0 * twopi;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const int j = (int)(g % a.batch);
        const size_t r = g / a.batch;
        const int ix = (int)(r % a.nx), iy = (int)((r / a.nx) % a.ny), iz = (int)(r / ((size_t)a.nx * a.ny));
        const int io = a.io0 + j;
        double2 out = make_double2(0.0, 0.0);
        if (ix <= a.nx / 2 && io < a.no) {
            // integer frequencies of the Poisson kernel
            const int m1 = ix < a.nx / 2 ? ix : ix - a.nx;
            const int m2 = iy < a.ny / 2 ? iy : iy - a.ny;
            const int m3 = iz < a.nz / 2 ? iz : iz - a.nz;
            const double q1 = twopi * m1 / a.lx, q2 = twopi * m2 / a.ly, q3 = twopi * m3 / a.lz;
            const double k2 = q1 * q1 + q2 * q2 + q3 * q3;
            if (fabs(k2) > DBL_EPSILON) {
                const double2 x = X[r];
                const double f = (double)N * fourpi / k2;  // N undoes the 1/N of X
                const double vr = x.x * f, vi = x.y * f;
                // the grid's wave vectors: +pi/d on the Nyquist index
                const double kx = twopi * (ix <= a.nx / 2 ? ix : ix - a.nx) / a.lx;
                const double ky = twopi * (iy <= a.ny / 2 ? iy : iy - a.ny) / a.ly;
                const double kz = twopi * (iz <= a.nz / 2 ? iz : iz - a.nz) / a.lz;
                const double smooth = exp(-kSmootherRadius * kSmootherRadius * (kx * kx + ky * ky + kz * kz) / 2.0);
                double sr = 0.0, si = 0.0;  // sigma_io(k)
                for (int s = 0; s < a.nsite; s++) {
                    const double q = a.site_q[s];
                    if (fabs(q) <= DBL_EPSILON) continue;
                    const double *o = a.site_off + ((size_t)io * a.nsite + s) * 3;
                    const double kr = kx * o[0] + ky * o[1] + kz * o[2];
                    sr += q * cos(kr) * smooth;
                    si -= q * sin(kr) * smooth;
                }
                // V conj(sigma), weighted for the c2r
                const double w = (ix == 0 || (a.nx % 2 == 0 && ix == a.nx / 2)) ? 1.0 : 2.0;
                out = make_double2(w * (vr * sr + vi * si), w * (vi * sr - vr * si));
            }
        }
        Z[g] = out;
    }
}

// vext[r][io0 + j] += scale * Re Z[r][j]
__global__ __launch_bounds__(kThreads) void poisson_accumulate(const double2 *__restrict__ Z, double *__restrict__ vext,
                                                              size_t N, int no, int io0, int batch, double scale) {
    const size_t total = N * batch;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const int j = (int)(g % batch);
        if (io0 + j >= no) continue;
        const size_t r = g / batch;
        vext[r * no + io0 + j] += scale * Z[g].x;
    }
}

__global__ __launch_bounds__(kThreads) void zero_field(double2 *__restrict__ X, size_t N) {
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < N; g += (size_t)gridDim.x * kThreads)
        X[g] = make_double2(0.0, 0.0);
}

__global__ void scatter_set(double2 *__restrict__ X, const long long *idx, const double *val, int nval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nval) X[idx[i]] = make_double2(val[i], 0.0);
}

__global__ __launch_bounds__(kThreads) void clip_kernel(double *__restrict__ v, size_t total, double clip) {
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads)
        if (v[g] > clip) v[g] = clip;
}

unsigned blocks_for(size_t total) {
    size_t nb = (total + kThreads - 1) / kThreads;
    return (unsigned)(nb > 65536 ? 65536 : nb);
}

}  // namespace

hipError_t launch_vext_clip(mdft_ctx *c, double clip) {
    const size_t total = c->N * (size_t)c->no;
    hipLaunchKernelGGL(clip_kernel, dim3(blocks_for(total)), dim3(kThreads), 0, c->stream, c->d_vext, total, clip);
    return hipGetLastError();
}

// vext += the Poisson electrostatic energy of the solvent sites (steps 1-3 above).
// h_atoms [natom][6] (x, y, z, q, sigma, epsilon) and h_site_q [nsite] on the host,
// d_site_off [no][nsite][3] on the device.
hipError_t launch_poisson_vext(mdft_ctx *c, int natom, const double *h_atoms, int nsite, const double *h_site_q,
                               const double *d_site_off, double qfact) {
    const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
    const double len[3] = {c->d.lx, c->d.ly, c->d.lz};
    // 1. cloud-in-cell charges (the reference's do concurrent over sites, in site order)
    std::map<long long, double> rho;
    bool charged = false;
    for (int s = 0; s < natom; s++) {
        const double q = h_atoms[6 * s + 3];
        if (!(fabs(q) > DBL_EPSILON)) continue;
        charged = true;
        int m[3], p[3];
        double wp[3], wm[3];
        for (int a = 0; a < 3; a++) {
            const double x = h_atoms[6 * s + a], d = len[a] / n[a];
            m[a] = (int)floor(fmod(fmod(x, len[a]) + len[a], len[a]) / d);  // floorNode - 1 (0-based)
            if (m[a] >= n[a]) m[a] -= n[a];
            p[a] = (m[a] + 1) % n[a];                                     // ceilingNode - 1
            wp[a] = fabs(x / d - floor(x / d));                            // distToFloorNode
            wm[a] = 1.0 - wp[a];
        }
        for (int corner = 0; corner < 8; corner++) {
            const int cx = (corner & 1) ? p[0] : m[0], cy = (corner & 2) ? p[1] : m[1], cz = (corner & 4) ? p[2] : m[2];
            // q * w1 * w2 * w3, left to right as in the reference
            const double qw = q * ((corner & 1) ? wp[0] : wm[0]) * ((corner & 2) ? wp[1] : wm[1]) *
                              ((corner & 4) ? wp[2] : wm[2]);
            rho[((long long)cz * n[1] + cy) * n[0] + cx] += qw;
        }
    }
    if (!charged) return hipSuccess;  // no solute charge: no electrostatic potential (src/module_vext.f90:154-164)
    const double dvol = (len[0] / n[0]) * (len[1] / n[1]) * (len[2] / n[2]);
    std::vector<long long> idx;
    std::vector<double> val;
    bool nonzero = false;
    for (auto &kv : rho) {
        idx.push_back(kv.first);
        val.push_back(kv.second / dvol);
        nonzero = nonzero || fabs(kv.second / dvol) > DBL_EPSILON;
    }
    if (!nonzero) return hipSuccess;  // all |rho| <= DBL_EPSILON (:187-190)
    const size_t N = c->N;
    const int B = c->np;
    double2 *X = nullptr, *Z = nullptr;
    long long *d_idx = nullptr;
    double *d_val = nullptr, *d_q = nullptr;
    hipError_t e = hipMalloc(&X, sizeof(double2) * N);
    if (e == hipSuccess) e = hipMalloc(&Z, sizeof(double2) * N * B);
    if (e == hipSuccess) e = hipMalloc(&d_idx, sizeof(long long) * idx.size());
    if (e == hipSuccess) e = hipMalloc(&d_val, sizeof(double) * val.size());
    if (e == hipSuccess) e = hipMalloc(&d_q, sizeof(double) * nsite);
    if (e == hipSuccess) e = hipMemcpy(d_idx, idx.data(), sizeof(long long) * idx.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_val, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_q, h_site_q, sizeof(double) * nsite, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(zero_field, dim3(blocks_for(N)), dim3(kThreads), 0, c->stream, X, N);
        hipLaunchKernelGGL(scatter_set, dim3(((unsigned)val.size() + 255) / 256), dim3(256), 0, c->stream, X, d_idx,
                           d_val, (int)val.size());
        e = hipGetLastError();
    }
    // 2. forward transform exp(-i k.r), normalised by 1/N (undone in poisson_spectrum)
    if (e == hipSuccess) e = launch_fft3d_f64(c, X, -1, 1);
    // 3. per chunk of orientations: spectrum, unnormalised inverse exp(+i k.r), accumulate
    PoissonArgs a{n[0], n[1], n[2], c->no, nsite, 0, B, len[0], len[1], len[2], d_q, d_site_off};
    for (int io0 = 0; e == hipSuccess && io0 < c->no; io0 += B) {
        a.io0 = io0;
        hipLaunchKernelGGL(poisson_spectrum, dim3(blocks_for(N * B)), dim3(kThreads), 0, c->stream, X, Z, a);
        e = hipGetLastError();
        if (e == hipSuccess) e = launch_fft3d_f64(c, Z, +1, B);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(poisson_accumulate, dim3(blocks_for(N * B)), dim3(kThreads), 0, c->stream, Z, c->d_vext,
                               N, c->no, io0, B, qfact / (double)N);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(X);
    (void)hipFree(Z);
    (void)hipFree(d_idx);
    (void)hipFree(d_val);
    (void)hipFree(d_q);
    return e;
}

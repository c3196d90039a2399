// k_angular.hip -- orientation <-> projection transforms and the fused
// gradient/free-energy pass, FP64.
//
// angl2proj restates src/module_orientation_projection_transform.f90:159-199
// (a 2D DFT over (psi, phi) for each theta node, then a Gauss-Legendre
// contraction in theta with Wigner d and sqrt(2m+1)), driven as in
// src/module_energy_cproj_mrso.f90:276-290 with o = rho0 (xi^2 - 1).
// The final pass fuses proj2angl (:202-243), the F_exc reduction and dF update
// of src/module_energy_cproj_mrso.f90:728-762 with the ideal+external term of
// src/module_energy_ideal_and_external.f90:48-72, so xi is read once, vext once
// and dF written once per evaluation.
//
// Layout: points r are processed in tiles of PT consecutive points per
// workgroup; xi/vext/df tiles [PT][no] and projection tiles [PT][np] are
// contiguous in HBM, so every global access is a coalesced stream.
#include <cfloat>

#include "mdft_internal.h"

namespace {

constexpr int kThreads = 256;

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename T> struct C2;
template <> struct C2<double> { using t = double2; };
template <> struct C2<float> { using t = float2; };
template <typename V> __device__ inline V mk(decltype(V::x) a, decltype(V::x) b) { V v; v.x = a; v.y = b; return v; }

// An FP32 value held in the low word of an FP64 register pair.  The tile prefetches keep the
// loaded word as it is and convert it where it is used: a conversion next to each load would
// make the wave wait for every load before issuing the next, one HBM latency per element.
__device__ inline double hold_f32(float v) { return __longlong_as_double((long long)__float_as_uint(v)); }
__device__ inline double held_f32(double h) { return (double)__uint_as_float((unsigned)__double_as_longlong(h)); }

template <int PT>
__global__ __launch_bounds__(kThreads) void angl2proj_f64(const double *__restrict__ xi, double2 *__restrict__ proj,
                                                         size_t N, AngTables t, double rho0) {
    extern __shared__ double smem[];
    const int no = t.no, np = t.np, nth = t.ntheta, nphi = t.nphi, npsi = t.npsi, H1 = t.H + 1,
              W = 2 * t.L + 1;
    double *o = smem;                                           // [PT][no]
    double2 *G = reinterpret_cast<double2 *>(o + PT * no);      // [PT][nth][nphi][H1]
    double2 *F = G + PT * nth * nphi * H1;                      // [PT][nth][W][H1]
    const size_t r0 = (size_t)blockIdx.x * PT;
    const int npts = (int)((N - r0) < (size_t)PT ? (N - r0) : (size_t)PT);
    const int tid = threadIdx.x;
    const double *src = xi + r0 * no;
    for (int i = tid; i < npts * no; i += kThreads) {
        double x = src[i];
        o[i] = rho0 * (x * x - 1.0);
    }
    __syncthreads();
    // DFT over psi: G = sum_ipsi o exp(+i mu2 psi)
    const int nG = npts * nth * nphi * H1;
    for (int i = tid; i < nG; i += kThreads) {
        const int mu2 = i % H1;
        const double *line = o + (i / H1) * npsi;
        double re = 0.0, im = 0.0;
        for (int ipsi = 0; ipsi < npsi; ipsi++) {
            const int k = (mu2 * ipsi) % npsi;
            const double v = line[ipsi];
            re += v * t.cps[k];
            im += v * t.sps[k];
        }
        G[i] = make_double2(re, im);
    }
    __syncthreads();
    // DFT over phi: F = sum_iphi G exp(+i mup phi) / (nphi npsi)
    const double norm = 1.0 / (double)(nphi * npsi);
    const int nF = npts * nth * W * H1;
    for (int i = tid; i < nF; i += kThreads) {
        const int mu2 = i % H1;
        int rest = i / H1;
        const int mup = rest % W - t.L;
        rest /= W;  // pt*nth + it
        const double2 *g = G + (size_t)rest * nphi * H1 + mu2;
        double re = 0.0, im = 0.0;
        for (int iphi = 0; iphi < nphi; iphi++) {
            int k = (mup * iphi) % nphi;
            k += (k < 0) ? nphi : 0;
            const double2 v = g[iphi * H1];
            re += v.x * t.cph[k] - v.y * t.sph[k];
            im += v.x * t.sph[k] + v.y * t.cph[k];
        }
        F[i] = make_double2(re * norm, im * norm);
    }
    __syncthreads();
    // Gauss-Legendre contraction in theta with Wigner d
    double2 *dst = proj + r0 * np;
    for (int i = tid; i < npts * np; i += kThreads) {
        const int pt = i / np, ip = i - pt * np;
        const int mup = t.pmup[ip], mu2 = t.pmu[ip] / t.mrso;
        const double2 *f = F + ((size_t)pt * nth * W + (mup + t.L)) * H1 + mu2;
        const double *a = t.a2p + ip * nth;
        double re = 0.0, im = 0.0;
        for (int it = 0; it < nth; it++) {
            const double2 v = f[(size_t)it * W * H1];
            re += v.x * a[it];
            im += v.y * a[it];
        }
        dst[i] = make_double2(re * t.fm[ip], im * t.fm[ip]);
    }
}

// products of the gradient with up to 3 vectors (FP32 or FP64, bit j of self: the gradient
// itself), accumulated in the final pass (mdft_energy_and_gradient_dots)
struct DotArgs {
    int k, self, f32;
    const void *p[3];
};

__device__ inline double dot_operand(const DotArgs &da, int j, size_t g, double gv) {
    if ((da.self >> j) & 1) return gv;
    return da.f32 ? (double)static_cast<const float *>(da.p[j])[g] : static_cast<const double *>(da.p[j])[g];
}

__device__ inline void block_reduce3(double &a, double &b, double &c, double *partials) {
    __shared__ double red[3][kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wv] = a;
        red[1][wv] = b;
        red[2][wv] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int k = 0; k < kThreads / 64; k++) {
            s0 += red[0][k];
            s1 += red[1][k];
            s2 += red[2][k];
        }
        partials[blockIdx.x * 3 + 0] = s0;
        partials[blockIdx.x * 3 + 1] = s1;
        partials[blockIdx.x * 3 + 2] = s2;
    }
}

struct FinalConsts {
    double rho0, kT, dv, prefactor;
};

// mode 0: df = ideal+ext gradient + excess gradient (energy_and_gradient)
// mode 1: df += excess gradient (excess_energy plug-in)
// mode 2: df = gamma(r, Omega) = -(prefactor / kT) proj2angl(proj), the OZ field in units of ln(rho/rho0)
// mode 3: proj holds gamma's projections, df = exp((-beta vext + proj2angl(proj)) / 2) (xi not read)
// mode 4: proj holds the projections of rho/rho0 - 1: df = 0 where the reference cold start
// excludes the solvent (beta vext >= -ln(DBL_EPSILON) - 0.05), else
// sqrt(max(1 + proj2angl(proj), MDFT_WS_DENSITY_FLOOR))
// mode 5 (specialised kernels only): df = proj2angl(proj)
template <int PT>
__global__ __launch_bounds__(kThreads) void final_f64(const double2 *__restrict__ proj, const double *__restrict__ xi,
                                                     const double *__restrict__ vext, double *__restrict__ df,
                                                     double *__restrict__ partials, size_t N, AngTables t,
                                                     FinalConsts k, int mode) {
    extern __shared__ double smem[];
    const int no = t.no, np = t.np, nth = t.ntheta, nphi = t.nphi, npsi = t.npsi, H1 = t.H + 1,
              W = 2 * t.L + 1, L = t.L;
    double2 *Gm = reinterpret_cast<double2 *>(smem);   // [PT][np]
    double2 *f = Gm + PT * np;                         // [PT][nth][W][H1]
    double2 *h = f + PT * nth * W * H1;                // [PT][nth][nphi][H1]
    const size_t r0 = (size_t)blockIdx.x * PT;
    const int npts = (int)((N - r0) < (size_t)PT ? (N - r0) : (size_t)PT);
    const int tid = threadIdx.x;
    const double2 *src = proj + r0 * np;
    for (int i = tid; i < npts * np; i += kThreads) Gm[i] = src[i];
    __syncthreads();
    // f(theta, mu2, mup) = sum_m p(m, mup, mu2) d(theta) fm(m)   (:215-224)
    const int nf = npts * nth * W * H1;
    for (int i = tid; i < nf; i += kThreads) {
        const int mu2 = i % H1;
        int rest = i / H1;
        const int mup = rest % W - L;
        rest /= W;
        const int it = rest % nth, pt = rest / nth;
        const int a = abs(mup), b = t.mrso * mu2;
        const int mlo = a > b ? a : b;
        double re = 0.0, im = 0.0;
        for (int m = mlo; m <= L; m++) {
            const int ip = t.pidx[(m * W + mup + L) * H1 + mu2];
            const double2 v = Gm[pt * np + ip];
            const double c = t.p2a[ip * nth + it];
            re += v.x * c;
            im += v.y * c;
        }
        f[i] = make_double2(re, im);
    }
    __syncthreads();
    // h(theta, phi, mu2) = sum_mup conj(f) exp(+i mup phi)
    const int nh = npts * nth * nphi * H1;
    for (int i = tid; i < nh; i += kThreads) {
        const int mu2 = i % H1;
        int rest = i / H1;
        const int iphi = rest % nphi;
        rest /= nphi;  // pt*nth + it
        const double2 *fr = f + (size_t)rest * W * H1 + mu2;
        double re = 0.0, im = 0.0;
        for (int mup = -L; mup <= L; mup++) {
            int kk = (mup * iphi) % nphi;
            kk += (kk < 0) ? nphi : 0;
            const double2 v = fr[(mup + L) * H1];
            // conj(v) * (c + i s)
            re += v.x * t.cph[kk] + v.y * t.sph[kk];
            im += v.x * t.sph[kk] - v.y * t.cph[kk];
        }
        h[i] = make_double2(re, im);
    }
    __syncthreads();
    double fid = 0.0, fext = 0.0, fexc = 0.0;
    const double rho0 = k.rho0, kT = k.kT, dv = k.dv;
    for (int i = tid; i < npts * no; i += kThreads) {
        const int pt = i / no, io = i - pt * no;
        const int ipsi = io % npsi, iphi = (io / npsi) % nphi, it = io / (npsi * nphi);
        const double2 *hr = h + ((size_t)(pt * nth + it) * nphi + iphi) * H1;
        // c2r: real part of the Hermitian-weighted sum over mu2 (weight 2 for mu2 > 0)
        double v = hr[0].x;
        for (int mu2 = 1; mu2 < H1; mu2++) {
            const int kk = (mu2 * ipsi) % npsi;
            v += 2.0 * (hr[mu2].x * t.cps[kk] - hr[mu2].y * t.sps[kk]);
        }
        const double wio = t.w[io];
        const double vexc = k.prefactor * wio * v;
        const size_t g = (r0 + pt) * (size_t)no + io;
        const double x = (mode >= 3) ? 0.0 : xi[g];
        const double rho = x * x * rho0;
        fexc += (x * x * rho0 - rho0) * vexc;
        if (mode == 0) {
            const double ve = vext[g];
            double gid = 0.0;
            if (rho > DBL_EPSILON) {
                const double lg = log(rho / rho0);
                fid += kT * dv * wio * (rho * lg - rho + rho0);
                fext += dv * wio * rho * ve;
                gid = wio * 2.0 * rho0 * x * (kT * lg + ve);
            } else {
                fid += kT * dv * wio * rho0;
            }
            if (df) df[g] = gid + 2.0 * rho0 * x * vexc;
        } else if (mode == 1) {
            if (df) df[g] += 2.0 * rho0 * x * vexc;
        } else if (mode == 2) {  // gamma(r, Omega) in units of ln(rho/rho0)
            if (df) df[g] = -(k.prefactor / kT) * v;
        } else if (mode == 3) {  // proj holds gamma's projections; df = exp((-beta vext + gamma) / 2)
            df[g] = exp(0.5 * fmin(fmax(-vext[g] / kT + v, -200.0), 7.0));
        } else {  // mode 4: proj holds the projections of rho/rho0 - 1;
                  // df = sqrt(max(1 + v, floor)), 0 where the cold start excludes the solvent
            df[g] = (vext[g] / kT >= -log(DBL_EPSILON) - 0.05) ? 0.0 : sqrt(fmax(1.0 + v, MDFT_WS_DENSITY_FLOOR));
        }
    }
    block_reduce3(fid, fext, fexc, partials);
}

// Ideal + external term alone (module_energy_ideal_and_external.f90:48-72); df accumulated.
__global__ __launch_bounds__(kThreads) void ideal_f64(const double *__restrict__ xi, const double *__restrict__ vext,
                                                     double *__restrict__ df, double *__restrict__ partials,
                                                     size_t total, int no, const double *__restrict__ w,
                                                     FinalConsts k) {
    double fid = 0.0, fext = 0.0, zero = 0.0;
    const double rho0 = k.rho0, kT = k.kT, dv = k.dv;
    for (size_t g = (size_t)blockIdx.x * kThreads + threadIdx.x; g < total; g += (size_t)gridDim.x * kThreads) {
        const int io = (int)(g % (size_t)no);
        const double x = xi[g], wio = w[io];
        const double rho = x * x * rho0;
        if (rho > DBL_EPSILON) {
            const double lg = log(rho / rho0);
            const double ve = vext[g];
            fid += kT * dv * wio * (rho * lg - rho + rho0);
            fext += dv * wio * rho * ve;
            if (df) df[g] += wio * 2.0 * rho0 * x * (kT * lg + ve);
        } else {
            fid += kT * dv * wio * rho0;
        }
    }
    block_reduce3(fid, fext, zero, partials);
}

// Sums of the per-workgroup (F_id, F_ext, F_exc) partials in a fixed order (so
// repeated evaluations are bitwise identical): each workgroup of the first
// pass sums one contiguous slice, one workgroup sums the slices.
__global__ __launch_bounds__(kThreads) void reduce3(const double *__restrict__ partials, int nblocks,
                                                   double *__restrict__ out) {
    const int per = (nblocks + gridDim.x - 1) / gridDim.x;
    const int lo = blockIdx.x * per, hi = min(nblocks, lo + per);
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = lo + threadIdx.x; i < hi; i += kThreads) {
        a += partials[3 * i + 0];
        b += partials[3 * i + 1];
        c += partials[3 * i + 2];
    }
    __shared__ double red[3][kThreads];
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
            red[2][threadIdx.x] += red[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x + 0] = red[0][0];
        out[3 * blockIdx.x + 1] = red[1][0];
        out[3 * blockIdx.x + 2] = red[2][0];
    }
}

// ---- specialisations for mrso = 2 with compile-time angular sizes: all
// index arithmetic uses constant divisors, loops have fixed trip counts, and
// the phase / Wigner / weight tables are staged in LDS once per workgroup.
template <int L>
struct Dims {
    static constexpr int NTH = L + 1, NPHI = 2 * L + 1, H = L / 2, H1 = H + 1, NPSI = 2 * H + 1, W = 2 * L + 1;
    static constexpr int NO = NTH * NPHI * NPSI;
    static constexpr int np_of(int m) { return m < 0 ? 0 : np_of(m - 1) + (2 * m + 1) * (m / 2 + 1); }
    static constexpr int NP = np_of(L);
};

// x / D for 0 <= x < NMAX as (x * M) >> S with a 24-bit multiply (one full-rate instruction;
// the 32-bit multiply-high of the compiler's constant division issues at a quarter rate).  M
// and S are chosen at compile time and checked against x / D for every x in the range.
template <unsigned D, unsigned NMAX>
struct DivC {
    static constexpr bool exact(unsigned long long m, unsigned s) {
        for (unsigned x = 0; x < NMAX; x++)
            if ((unsigned)((x * m) >> s) != x / D) return false;
        return true;
    }
    static constexpr unsigned pick() {
        for (unsigned s = 0; s < 32; s++) {
            const unsigned long long m = ((1ull << s) + D - 1) / D;
            if (m < (1ull << 24) && (NMAX - 1) * m < (1ull << 32) && exact(m, s)) return s;
        }
        return 32;
    }
    static constexpr unsigned S = pick();
    static_assert(S < 32 && NMAX <= (1u << 24), "no 24-bit reciprocal for this divisor and range");
    static constexpr unsigned M = (unsigned)(((1ull << S) + D - 1) / D);
    __device__ static unsigned q(unsigned x) { return __umul24(x, M) >> S; }
};

// a * b for a, b < 2^24 and a product below 2^32, kept a 24-bit multiply: the empty asm stops
// the compiler from folding the product into a subtraction or a 64-bit address, where it would
// become a quarter-rate 32-bit multiply again
__device__ inline unsigned mul24(unsigned a, unsigned b) {
    unsigned r = __umul24(a, b);
    asm("" : "+v"(r));
    return r;
}

// The constant tables of angl2proj_t in the layout of its LDS.  The image is built once per
// context and precision (a2p_tab_build, the staging arithmetic run once) and each workgroup
// copies it with 16-byte loads: the index arithmetic and FP64 -> T conversions of the
// staging then run once instead of once per 4-point tile.
template <int L, typename T> struct alignas(16) A2pTab {
    using V = typename C2<T>::t;
    using D = Dims<L>;
    V twpsi[D::H1 * D::NPSI];  // (cos, sin)(2 pi mu2 ipsi / npsi) at [mu2][ipsi]
    V twphi[D::W * D::NPHI];   // (cos, sin)(2 pi mup iphi / nphi) / (nphi npsi) at [mup + L][iphi]
    T a2p[D::NP * D::NTH], fmv[D::NP];
    int pmup[D::NP], pmu2[D::NP];
};

template <int L, typename T>
__global__ __launch_bounds__(kThreads) void a2p_tab_build(A2pTab<L, T> *__restrict__ out, AngTables t) {
    using D = Dims<L>;
    using V = typename C2<T>::t;
    constexpr int NP = D::NP, NTH = D::NTH, NPHI = D::NPHI, NPSI = D::NPSI, H1 = D::H1, W = D::W;
    const int tid = threadIdx.x;
    for (int i = tid; i < H1 * NPSI; i += kThreads) {
        const int kk = ((i / NPSI) * (i % NPSI)) % NPSI;
        out->twpsi[i] = mk<V>((T)t.cps[kk], (T)t.sps[kk]);
    }
    for (int i = tid; i < W * NPHI; i += kThreads) {
        const int mup = i / NPHI - L, iphi = i % NPHI;
        int kk = (mup * iphi) % NPHI;
        kk += (kk < 0) ? NPHI : 0;
        const double nrm = 1.0 / (double)(NPHI * NPSI);
        out->twphi[i] = mk<V>((T)(t.cph[kk] * nrm), (T)(t.sph[kk] * nrm));
    }
    for (int i = tid; i < NP * NTH; i += kThreads) out->a2p[i] = (T)t.a2p[i];
    for (int i = tid; i < NP; i += kThreads) {
        out->fmv[i] = (T)t.fm[i];
        out->pmup[i] = t.pmup[i];
        out->pmu2[i] = t.pmu[i] / 2;
    }
}

// The psi, phi and theta steps of angl2proj over one tile (o in LDS and visible to every
// thread): G (psi DFT), F (phi DFT), then the theta contraction into the tile's projections.
template <int L, int PT, typename T>
__device__ inline void a2p_phases(const T *o, typename C2<T>::t *G, typename C2<T>::t *F, const A2pTab<L, T> &tb,
                                  typename C2<T>::t *__restrict__ dst, int npts, int tid) {
    using D = Dims<L>;
    using V = typename C2<T>::t;
    constexpr int NP = D::NP, NTH = D::NTH, NPHI = D::NPHI, NPSI = D::NPSI, H1 = D::H1, W = D::W;
    const V *twpsi = tb.twpsi, *twphi = tb.twphi;
    const T *a2p = tb.a2p, *fmv = tb.fmv;
    const int *pmup = tb.pmup, *pmu2 = tb.pmu2;
    for (int i = tid; i < npts * NTH * NPHI * H1; i += kThreads) {
        const int mu2 = i % H1;
        const T *line = o + (i / H1) * NPSI;
        const V *tw = twpsi + mu2 * NPSI;
        T re = 0, im = 0;
#pragma unroll
        for (int ipsi = 0; ipsi < NPSI; ipsi++) {
            const T v = line[ipsi];
            re += v * tw[ipsi].x;
            im += v * tw[ipsi].y;
        }
        G[i] = mk<V>(re, im);
    }
    __syncthreads();
    for (int i = tid; i < npts * NTH * W * H1; i += kThreads) {
        const int mu2 = i % H1;
        const int rest = i / H1;
        const int jm = rest % W;  // mup + L
        const V *g = G + (rest / W) * NPHI * H1 + mu2;
        const V *tw = twphi + jm * NPHI;
        T re = 0, im = 0;
#pragma unroll
        for (int iphi = 0; iphi < NPHI; iphi++) {
            const V v = g[iphi * H1], e = tw[iphi];
            if constexpr (sizeof(T) == 4) {
                // four multiply-adds per complex product (the plain expression compiles to six;
                // in FP64 the plain form keeps the kernel's register count)
                re = fma(v.x, e.x, re);
                re = fma(-v.y, e.y, re);
                im = fma(v.x, e.y, im);
                im = fma(v.y, e.x, im);
            } else {
                re += v.x * e.x - v.y * e.y;
                im += v.x * e.y + v.y * e.x;
            }
        }
        F[i] = mk<V>(re, im);
    }
    __syncthreads();
    auto theta = [&](int pt, int ip) {
        const V *f = F + (pt * NTH * W + (pmup[ip] + L)) * H1 + pmu2[ip];
        const T *a = a2p + ip * NTH;
        T re = 0, im = 0;
#pragma unroll
        for (int it = 0; it < NTH; it++) {
            const V v = f[it * W * H1];
            re += v.x * a[it];
            im += v.y * a[it];
        }
        return mk<V>(re * fmv[ip], im * fmv[ip]);
    };
    if constexpr (sizeof(T) == 4 && NP % 2 == 0) {
        // FP32: two projections of a point per thread, one 16-byte store
        for (int i = tid; i < npts * (NP / 2); i += kThreads) {
            const int pt = i / (NP / 2), ip = 2 * (i - pt * (NP / 2));
            const V u0 = theta(pt, ip), u1 = theta(pt, ip + 1);
            reinterpret_cast<float4 *>(dst)[i] = make_float4(u0.x, u0.y, u1.x, u1.y);
        }
    } else {
        for (int i = tid; i < npts * NP; i += kThreads) {
            const int pt = i / NP, ip = i - pt * NP;
            dst[i] = theta(pt, ip);
        }
    }
}

// feat = 0: o = rho0 (xi^2 - 1) (the excess term); feat = 1: o = clip(-beta xi / 2, -4, 4) / 4
// with xi the external potential -- the log of the square-root Boltzmann factor, bounded
// (features of the learned warm start); feat = 2: o = xi^2 - 1 = rho/rho0 - 1 (the
// warm start's density target; 0 in the bulk, -1 where the solvent is excluded); feat = 3:
// o = k0 aux / (xi w(io)) where xi^2 > MDFT_WS_DENSITY_FLOOR, else 0, with aux = dF at the
// density guess xi and k0 = dv / 2: dF/dr / w for r = rho/rho0, whose angl2proj is the
// adjoint of proj2angl up to a per-projection factor (the warm start's training gradient);
// feat = 4: o = xi (the plain transform). k0 is rho0 for feat 0.
template <int L, int PT, typename T>
__global__ __launch_bounds__(kThreads) void angl2proj_t(const double *__restrict__ xi, typename C2<T>::t *__restrict__ proj,
                                                       size_t N, AngTables t, double k0, int feat, double beta,
                                                       const double *__restrict__ aux,
                                                       const double *__restrict__ vsinv,
                                                       const float *__restrict__ xi32,
                                                       const A2pTab<L, T> *__restrict__ img) {
    using D = Dims<L>;
    using V = typename C2<T>::t;
    constexpr int NO = D::NO, NP = D::NP, NTH = D::NTH, NPHI = D::NPHI, H1 = D::H1, W = D::W;
    // o (read by the psi step) and F (written by the phi step, after the barrier that ends
    // the psi step) share their LDS: the smaller footprint lets more workgroups share a CU
    constexpr size_t kOB = sizeof(T) * PT * NO, kFB = sizeof(V) * PT * NTH * W * H1;
    __shared__ __align__(16) unsigned char oF[kOB > kFB ? kOB : kFB];
    __shared__ V G[PT * NTH * NPHI * H1];
    __shared__ A2pTab<L, T> tb;
    T *o = reinterpret_cast<T *>(oF);
    V *F = reinterpret_cast<V *>(oF);
    const int tid = threadIdx.x;
    const size_t r0 = (size_t)blockIdx.x * PT;
    const int npts = (int)((N - r0) < (size_t)PT ? (N - r0) : (size_t)PT);
    // the tile's FP64 input is loaded into registers before the tables are staged,
    // so its HBM latency overlaps the staging
    constexpr int NIT = (PT * NO + kThreads - 1) / kThreads;
    const double *src = xi + r0 * NO;
    double xr[NIT], sr[NIT];
#pragma unroll
    for (int q = 0; q < NIT; q++) {
        const int i = tid + q * kThreads;
        xr[q] = (i < npts * NO) ? (xi32 ? hold_f32(xi32[r0 * NO + i]) : src[i]) : 0.0;  // FP32 iterate: xi32
        sr[q] = vsinv ? vsinv[i % NO] : 1.0;  // scaled variables: xi = u / s(io), loaded with the tile
    }
    {
        constexpr int n16 = (int)(sizeof(A2pTab<L, T>) / 16);
        const uint4 *src16 = reinterpret_cast<const uint4 *>(img);
        uint4 *dst16 = reinterpret_cast<uint4 *>(&tb);
        for (int i = tid; i < n16; i += kThreads) dst16[i] = src16[i];
    }
#pragma unroll
    for (int q = 0; q < NIT; q++) {
        const int i = tid + q * kThreads;
        if (i >= npts * NO) break;
        const double xv = xi32 ? held_f32(xr[q]) : xr[q];
        const double x = vsinv ? xv * sr[q] : xv;
        if (feat == 0) o[i] = (T)(k0 * (x * x - 1.0));
        else if (feat == 1) o[i] = (T)(0.25 * fmin(fmax(-0.5 * beta * x, -4.0), 4.0));
        else if (feat == 2) o[i] = (T)(x * x - 1.0);
        else if (feat == 4) o[i] = (T)x;
        else o[i] = (x * x > MDFT_WS_DENSITY_FLOOR * (1.0 + 1e-12)) ? (T)(k0 * aux[r0 * NO + i] / (x * t.w[i % NO]))
                                                                   : (T)0;
    }
    __syncthreads();
    a2p_phases<L, PT, T>(o, G, F, tb, proj + r0 * NP, npts, tid);
}

// Persistent: each workgroup stages the tables once and then takes tiles blockIdx.x,
// blockIdx.x + gridDim.x, ...  A tile needs no barrier beyond its three: every LDS array a
// tile writes was last read before a barrier that all threads have passed by then.  The
// free-energy sums and gradient products accumulate over a workgroup's tiles and are
// written once, one partial per workgroup.
// DOT32: mode 0 with an FP32 iterate, forming the gradient products over FP32 vectors (the
// evaluations of the fast minimiser), on full tiles only (N a multiple of PT).  Every load
// of the tile is then issued on every path (the last partial round of a thread's elements
// with its index clamped into the tile), and dF is stored on every path: the wait before
// each use counts exactly the memory operations issued after it, so the product operands
// loaded ahead stay in flight.
template <int L, int PT, typename T, bool DOT32 = false>
// DOT32: at least 5 waves per SIMD (at most 96 registers): the tile's iterate, vext and product
// operands, loaded before the synthesis phases, stay in registers through them (measured faster
// than 6 or 7 waves with fewer of them in flight); the general variant, with every mode's
// branches, may take more registers rather than spill
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(DOT32 ? 5 : 4, 7))) void final_t(const typename C2<T>::t *__restrict__ proj, const double *__restrict__ xi,
                                                   const double *__restrict__ vext, double *__restrict__ df,
                                                   double *__restrict__ partials, size_t N, AngTables t,
                                                   FinalConsts k, int mode, float *__restrict__ df32,
                                                   const double *__restrict__ vsinv, DotArgs da,
                                                   double *__restrict__ dpart, const float *__restrict__ xi32,
                                                   const float *__restrict__ vext32) {
    using D = Dims<L>;
    using V = typename C2<T>::t;
    constexpr int NO = D::NO, NP = D::NP, NTH = D::NTH, NPHI = D::NPHI, NPSI = D::NPSI, H1 = D::H1, W = D::W;
    if constexpr (DOT32) mode = 0;
    __shared__ __align__(16) V Gm[PT * NP];
    __shared__ V f[PT * NTH * W * H1];
    __shared__ V h[PT * NTH * NPHI * H1];
    __shared__ V twphi[NPHI * W];   // (cos, sin)(2 pi mup iphi / nphi) at [iphi][mup + L]
    __shared__ V twpsi[NPSI * H1];  // 2 (cos, sin)(2 pi mu2 ipsi / npsi) at [ipsi][mu2]
    __shared__ T p2a[NP * NTH];
    __shared__ double wq[NO];
    __shared__ double sth[NTH];  // 1 / s of each theta (the weights, hence s, depend on theta only)
    // projection index of (m, mup, mu2), as bytes: the loaded index is then known to be small and
    // its products stay 24-bit multiplies (entries with m below the item's lowest m are never read)
    static_assert(NP <= 255, "projection indices are stored as bytes");
    __shared__ unsigned char pidx[(L + 1) * W * H1];
    __shared__ unsigned char ford[W * H1];  // the theta synthesis' (mup, mu2) items, longest m loop first
    const int tid = threadIdx.x;
    for (int i = tid; i < (L + 1) * W * H1; i += kThreads) pidx[i] = (unsigned char)t.pidx[i];
    {
        // the item (mup + L) * H1 + mu2 sums over m = max(|mup|, 2 mu2) .. L; ranked by that
        // length (ties by index), the lanes of a wave take loops of nearly one length
        auto len = [](int it) {
            const int a = it / H1 - L, b = 2 * (it % H1);
            const int aa = a < 0 ? -a : a;
            return L + 1 - (aa > b ? aa : b);
        };
        for (int i = tid; i < W * H1; i += kThreads) {
            const int li = len(i);
            int r = 0;
            for (int j = 0; j < W * H1; j++) {
                const int lj = len(j);
                r += (lj > li || (lj == li && j < i)) ? 1 : 0;
            }
            ford[r] = (unsigned char)i;
        }
    }
    for (int i = tid; i < NPHI * W; i += kThreads) {
        const int iphi = i / W, mup = i % W - L;
        int kk = (mup * iphi) % NPHI;
        kk += (kk < 0) ? NPHI : 0;
        twphi[i] = mk<V>((T)t.cph[kk], (T)t.sph[kk]);
    }
    for (int i = tid; i < NPSI * H1; i += kThreads) {
        const int kk = ((i % H1) * (i / H1)) % NPSI;
        twpsi[i] = mk<V>((T)(2.0 * t.cps[kk]), (T)(2.0 * t.sps[kk]));
    }
    for (int i = tid; i < NP * NTH; i += kThreads) p2a[i] = (T)t.p2a[i];
    for (int i = tid; i < NO; i += kThreads) wq[i] = t.w[i];
    for (int i = tid; i < NTH; i += kThreads) sth[i] = vsinv ? vsinv[i * NPHI * NPSI] : 1.0;
    double fid = 0.0, fext = 0.0, fexc = 0.0;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;  // products of dF with the DotArgs vectors
    double dss = 0.0;                     // DOT32: the product of dF with itself
    const double rho0 = k.rho0, kT = k.kT, dv = k.dv;
    const size_t ntiles = (N + PT - 1) / PT;
    for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        // the thread index, opaque to the compiler in each tile: the tile's indices and LDS
        // addresses depend on the thread only, and hoisted out of the tile loop they would
        // all stay in registers (and spill)
        int tl = tid;
        asm volatile("" : "+v"(tl));
        const size_t r0 = tile * PT;
        const int npts = DOT32 ? PT : (int)((N - r0) < (size_t)PT ? (N - r0) : (size_t)PT);
        // DOT32: the tile's base pointers (uniform), indexed by 32-bit element offsets
        const float *xi32t = DOT32 ? xi32 + r0 * NO : nullptr, *vx32t = DOT32 ? vext32 + r0 * NO : nullptr;
        float *df32t = DOT32 ? df32 + r0 * NO : nullptr;
        const V *src = proj + r0 * NP;
        if constexpr (sizeof(T) == 4 && NP % 2 == 0) {  // FP32: 16-byte loads, two projections each
            for (int i = tl; i < npts * (NP / 2); i += kThreads)
                reinterpret_cast<float4 *>(Gm)[i] = reinterpret_cast<const float4 *>(src)[i];
        } else {
            for (int i = tl; i < npts * NP; i += kThreads) Gm[i] = src[i];
        }
        // DOT32: the tile's FP32 iterate, FP32 vext and the FP32 operands of the first two
        // gradient products, loaded here so that their latency passes during the synthesis
        // phases; an operand that is not needed (the gradient itself, or beyond da.k) is read
        // from dF, which exists here
        constexpr int NIT = (PT * NO + kThreads - 1) / kThreads;
        float xr32[NIT], vr32[NIT];
        float dq[2][NIT];
        const float *dp0 = (da.k > 0 && !(da.self & 1)) ? static_cast<const float *>(da.p[0]) : df32;
        const float *dp1 = (da.k > 1 && !((da.self >> 1) & 1)) ? static_cast<const float *>(da.p[1]) : df32;
        const float *dp0t = DOT32 ? dp0 + r0 * NO : nullptr, *dp1t = DOT32 ? dp1 + r0 * NO : nullptr;
        if constexpr (DOT32) {
    #pragma unroll
            for (int q = 0; q < NIT; q++) {
                const int i = tl + q * kThreads;
                const int ii = i < npts * NO ? i : npts * NO - 1;
                xr32[q] = xi32t[ii];
                vr32[q] = vx32t[ii];
                dq[0][q] = dp0t[ii];
                dq[1][q] = dp1t[ii];
            }
        }
        __syncthreads();
        // f(theta, mu2, mup) = sum_m p(m, mup, mu2) d(theta) fm(m): one thread per
        // (point, mup, mu2), all theta at once (each projection read once); items by rank of
        // their loop length, the tile's points innermost
        for (int i = tl; i < PT * W * H1; i += kThreads) {
            const int rk = i / PT, pt = i - rk * PT;
            if (pt >= npts) continue;
            const int item = ford[rk];
            const int jm = (int)DivC<H1, W * H1>::q(item);
            const int mu2 = item - jm * H1, mup = jm - L;
            const int a = mup < 0 ? -mup : mup, b = 2 * mu2;
            const int mlo = a > b ? a : b;
            T re[NTH], im[NTH];
    #pragma unroll
            for (int it = 0; it < NTH; it++) re[it] = im[it] = (T)0;
            const unsigned char *pp = pidx + (mul24(jm, H1) + mu2);  // the item's entry of each m row
            const V *gp = Gm + mul24(pt, NP);
            for (int m = mlo; m <= L; m++) {
                const unsigned p = pp[m * (W * H1)];
                const V v = gp[p];
                const T *c = p2a + p * NTH;
    #pragma unroll
                for (int it = 0; it < NTH; it++) {
                    re[it] += v.x * c[it];
                    im[it] += v.y * c[it];
                }
            }
            V *fo = f + (mul24(mul24(pt, NTH * W) + jm, H1) + mu2);
    #pragma unroll
            for (int it = 0; it < NTH; it++) fo[it * (W * H1)] = mk<V>(re[it], im[it]);
        }
        __syncthreads();
        // h(theta, phi, mu2): the sum over mup of f(theta, mu2, mup) with the phi twiddles; items
        // mu2-major, so that whole waves take mu2 = 0, whose imaginary part the psi synthesis
        // below never reads
        constexpr int NL = PT * NTH * NPHI;  // (point, theta, phi) lines of a full tile
        for (int i = tl; i < NL * H1; i += kThreads) {
            const int mu2 = (int)DivC<NL, NL * H1>::q(i), line = i - mu2 * NL;
            if (line >= npts * NTH * NPHI) continue;
            const int rest = (int)DivC<NPHI, NL>::q(line), iphi = line - rest * NPHI;
            const V *fr = f + (mul24(rest, W * H1) + mu2);
            const V *tw = twphi + mul24(iphi, W);
            T re = 0, im = 0;
            if (mu2 == 0) {
    #pragma unroll
                for (int j = 0; j < W; j++) {
                    const V v = fr[j * H1], e = tw[j];
                    if constexpr (sizeof(T) == 4) {
                        re = fma(v.x, e.x, re);
                        re = fma(v.y, e.y, re);
                    } else {
                        re += v.x * e.x + v.y * e.y;
                    }
                }
            } else {
    #pragma unroll
                for (int j = 0; j < W; j++) {
                    const V v = fr[j * H1], e = tw[j];
                    if constexpr (sizeof(T) == 4) {  // four multiply-adds (FP32 only, as in angl2proj_t)
                        re = fma(v.x, e.x, re);
                        re = fma(v.y, e.y, re);
                        im = fma(v.x, e.y, im);
                        im = fma(-v.y, e.x, im);
                    } else {
                        re += v.x * e.x + v.y * e.y;
                        im += v.x * e.y - v.y * e.x;
                    }
                }
            }
            h[mul24(line, H1) + mu2] = mk<V>(re, im);
        }
        __syncthreads();
        // the general variant: the tile's xi and vext, all of a thread's loads issued together
        // for the loop below; the other workgroups on the compute unit cover their latency
        // (loaded before the synthesis phases, these FP64 values would hold twice the registers)
        double xr[NIT], vr[NIT];
    #pragma unroll
        for (int q = 0; q < NIT; q++) {
            if constexpr (DOT32) continue;
            const int i = tl + q * kThreads;
            xr[q] = 0.0;
            vr[q] = 0.0;
            if (i < npts * NO) {
                const size_t g = r0 * NO + i;
                if (mode < 3) xr[q] = xi32 ? hold_f32(xi32[g]) : xi[g];  // FP32 iterate: xi32
                if (mode == 0 || mode == 3 || mode == 4) vr[q] = vext[g];
            }
        }
        // FP32 path: the free-energy terms of a thread's elements in this tile, summed in FP32
        // and added to the FP64 sums once per tile
        float fexc_t = 0.0f, fid_t = 0.0f, fext_t = 0.0f;
    #pragma unroll
        for (int q = 0; q < NIT; q++) {
            const int i = tl + q * kThreads;
            if (i >= npts * NO) break;
            double xq;
            float ve32 = 0.0f;
            if constexpr (DOT32) {
                float xv = xr32[q];
                ve32 = vr32[q];
                asm volatile("" : "+v"(xv), "+v"(ve32));  // the conversion stays here, not next to the load
                xq = (double)xv;
            } else {
                xq = xi32 ? held_f32(xr[q]) : xr[q];
            }
            // LDS offsets by 24-bit products: every operand here is small and non-negative
            const unsigned pt = DivC<NO, PT * NO>::q(i);
            const int io = i - (int)mul24(pt, NO);
            const unsigned line = DivC<NPSI, NO>::q(io), ipsi = io - mul24(line, NPSI);  // line = it*NPHI + iphi
            const V *hr = h + mul24(mul24(pt, NTH * NPHI) + line, H1);
            const V *tp = twpsi + mul24(ipsi, H1);
            T vt = hr[0].x;
    #pragma unroll
            for (int mu2 = 1; mu2 < H1; mu2++) {
                if constexpr (sizeof(T) == 4) vt = fma(-hr[mu2].y, tp[mu2].y, fma(hr[mu2].x, tp[mu2].x, vt));
                else vt += hr[mu2].x * tp[mu2].x - hr[mu2].y * tp[mu2].y;
            }
            const size_t g = r0 * NO + i;
            if constexpr (sizeof(T) == 4) {
                if (mode == 0) {
                    // FP32 path: the per-orientation terms in FP32 (the transforms feeding them
                    // are FP32 already), summed in FP32 over the thread's elements of the tile
                    // and in FP64 over the tiles; rho/rho0 - 1 is formed in FP64 first so that
                    // the excess term keeps its small deviations from the bulk
                    const double si = vsinv ? sth[DivC<NPHI * NPSI, NO>::q(io)] : 1.0;  // scaled variables: xi = u / s, dF/du = dF/dxi / s
                    const double x = xq * si;
                    const double xx = x * x;
                    const float wf = (float)wq[io];
                    const float vexcf = (float)k.prefactor * wf * vt;
                    const float drf = (float)(xx - 1.0);
                    const float rho0f = (float)rho0, x2 = (float)xx, xf = (float)x;
                    fexc_t += rho0f * drf * vexcf;
                    float gid = 0.0f;
                    if (xx * rho0 > DBL_EPSILON) {
                        // ln(rho / rho0) from the hardware base-2 logarithm: x2 > 6e-15 here (rho > DBL_EPSILON),
                        // a normal FP32 number, so logf's denormal scaling and its extra-precise
                        // product with ln 2 change nothing at this path's FP32 accuracy
                        const float lg = __builtin_amdgcn_logf(x2) * 0.693147181f;
                        const float ve = DOT32 ? ve32 : (float)vr[q];
                        const float kTf = (float)kT, dvf = (float)dv;
                        fid_t += kTf * dvf * wf * rho0f * (x2 * lg - drf);
                        fext_t += dvf * wf * rho0f * x2 * ve;
                        gid = wf * 2.0f * rho0f * xf * (kTf * lg + ve);
                    } else {
                        fid_t += (float)(kT * dv * rho0) * wf;
                    }
                    const float gv = (gid + 2.0f * rho0f * xf * vexcf) * (float)si;
                    if constexpr (DOT32) {
                        df32t[i] = gv;
                        const double gd = (double)gv;
                        // the empty asm keeps the conversions here: hoisted next to the loads above,
                        // they would make each load wait
                        float o0 = dq[0][q], o1 = dq[1][q];
                        asm volatile("" : "+v"(o0), "+v"(o1));
                        // every product is formed (an operand that is not needed was read from dF
                        // and is discarded at the end); the gradient's own product separately
                        dss = fma(gd, gd, dss);
                        d0 = fma(gd, (double)o0, d0);
                        d1 = fma(gd, (double)o1, d1);
                        if (da.k > 2 && !((da.self >> 2) & 1)) d2 += gd * (double)static_cast<const float *>(da.p[2])[g];
                        continue;
                    }
                    if (df32) df32[g] = gv;
                    else if (df) df[g] = (double)gv;
                    if (da.k > 0) {
                        d0 += (double)gv * dot_operand(da, 0, g, (double)gv);
                        if (da.k > 1) d1 += (double)gv * dot_operand(da, 1, g, (double)gv);
                        if (da.k > 2) d2 += (double)gv * dot_operand(da, 2, g, (double)gv);
                    }
                    continue;
                }
            }
            // from here on FP64: free-energy reductions, rho ln rho and the gradient
            const double v = (double)vt;
            const double wio = wq[io];
            const double vexc = k.prefactor * wio * v;
            const double si = (mode == 0 && vsinv) ? sth[DivC<NPHI * NPSI, NO>::q(io)] : 1.0;  // scaled variables (mode 0)
            const double x = xq * si;
            const double rho = x * x * rho0;
            fexc += (x * x * rho0 - rho0) * vexc;
            if (mode == 0) {
                const double ve = vr[q];
                double gid = 0.0;
                if (rho > DBL_EPSILON) {
                    const double lg = log(x * x);  // ln(rho / rho0)
                    fid += kT * dv * wio * (rho * lg - rho + rho0);
                    fext += dv * wio * rho * ve;
                    gid = wio * 2.0 * rho0 * x * (kT * lg + ve);
                } else {
                    fid += kT * dv * wio * rho0;
                }
                const double gv = (gid + 2.0 * rho0 * x * vexc) * si;
                if (df32) df32[g] = (float)gv;
                else if (df) df[g] = gv;
                // (the FP64 kernel forms no gradient products: its context runs the separate
                // multi-dot, which keeps this kernel's registers where they were)
            } else if (mode == 1) {
                if (df) df[g] += 2.0 * rho0 * x * vexc;
            } else if (mode == 2) {  // gamma(r, Omega) in units of ln(rho/rho0)
                if (df) df[g] = -(k.prefactor / kT) * v;
            } else if (mode == 3) {  // proj holds gamma's projections; df = exp((-beta vext + gamma) / 2)
                df[g] = exp(0.5 * fmin(fmax(-vr[q] / kT + v, -200.0), 7.0));
            } else if (mode == 4) {  // proj holds the projections of rho/rho0 - 1;
                                     // df = sqrt(max(1 + v, floor)), 0 where the cold start excludes the solvent
                df[g] = (vr[q] / kT >= -log(DBL_EPSILON) - 0.05) ? 0.0 : sqrt(fmax(1.0 + v, MDFT_WS_DENSITY_FLOOR));
            } else {  // mode 5: df = proj2angl(proj), the plain transform
                df[g] = v;
            }
        }
        fexc += (double)fexc_t;
        fid += (double)fid_t;
        fext += (double)fext_t;
    }
    block_reduce3(fid, fext, fexc, partials);
    if constexpr (sizeof(T) == 4) {
        if (da.k > 0) {
            if constexpr (DOT32) {
                if (da.self & 1) d0 = dss;
                if ((da.self >> 1) & 1) d1 = dss;
                if ((da.self >> 2) & 1) d2 = dss;
            }
            __syncthreads();
            block_reduce3(d0, d1, d2, dpart);
        }
    }
}

constexpr int kPT = 4;  // points per workgroup tile (keeps the LDS tile under 64 KB at mmax = 5)

// Workgroups of a persistent kernel, never more than the tiles.  resident_grid: as many as
// the compute units hold at once (registers and LDS of that kernel); a larger grid leaves
// workgroups waiting for a slot until others have finished all their tiles.  The fast final
// pass runs faster so (3.85 -> 3.77 ms at 128^3, mmax 5); the general final pass, at 3 (FP64,
// the LDS limit) or 4 (FP32) resident workgroups per compute unit, runs faster with 8 per
// compute unit (FP64 6.4 against 6.9 ms), the later ones taking the slots of the first to finish.
template <typename K>
size_t resident_grid(const mdft_ctx *c, K kernel, size_t ntiles) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    const size_t g = (size_t)per_cu * (size_t)(c->num_cus > 0 ? c->num_cus : 256);
    return ntiles < g ? ntiles : g;
}
size_t oversubscribed_grid(const mdft_ctx *c, size_t ntiles) {
    const size_t g = 8 * (size_t)(c->num_cus > 0 ? c->num_cus : 256);
    return ntiles < g ? ntiles : g;
}

size_t a2p_lds(const AngTables &t) {
    return sizeof(double) * kPT * t.no + sizeof(double2) * kPT * t.ntheta * t.nphi * (t.H + 1) +
           sizeof(double2) * kPT * t.ntheta * (2 * t.L + 1) * (t.H + 1);
}
size_t final_lds(const AngTables &t) {
    return sizeof(double2) * kPT * t.np + sizeof(double2) * kPT * t.ntheta * (2 * t.L + 1) * (t.H + 1) +
           sizeof(double2) * kPT * t.ntheta * t.nphi * (t.H + 1);
}

hipError_t ensure_partials(mdft_ctx *c, size_t nblocks) {
    if (c->partials_cap >= nblocks * 3) return hipSuccess;
    if (c->d_partials) (void)hipFree(c->d_partials);
    c->d_partials = nullptr;
    hipError_t e = hipMalloc(&c->d_partials, sizeof(double) * nblocks * 3);
    if (e == hipSuccess) c->partials_cap = nblocks * 3;
    return e;
}

}  // namespace

template <int PT, typename T>
bool launch_a2p_t(mdft_ctx *c, const double *xi, typename C2<T>::t *proj, int feat = 0, const double *aux = nullptr,
                  size_t npts = 0) {
    if (!npts) npts = c->N;
    const size_t nb = (npts + PT - 1) / PT;
    const double beta = 1.0 / c->d.kT;
    const double k0 = (feat == 3) ? 0.5 * c->dv : c->rho0;
    void *&img = c->d_a2p_img[sizeof(T) == 8];
    switch (c->L) {
#define A2P_CASE(LL) \
    case LL: \
        if (!img) { \
            if (hipMalloc(&img, sizeof(A2pTab<LL, T>)) != hipSuccess) { img = nullptr; return false; } \
            hipLaunchKernelGGL((a2p_tab_build<LL, T>), dim3(1), dim3(kThreads), 0, c->stream, static_cast<A2pTab<LL, T> *>(img), c->ang); \
        } \
        hipLaunchKernelGGL((angl2proj_t<LL, PT, T>), dim3((unsigned)nb), dim3(kThreads), 0, c->stream, xi, proj, npts, c->ang, k0, feat, beta, aux, feat == 0 ? c->vsinv : nullptr, feat == 0 ? c->xi32 : nullptr, static_cast<const A2pTab<LL, T> *>(img)); \
        return true;
        A2P_CASE(0) A2P_CASE(1) A2P_CASE(2) A2P_CASE(3) A2P_CASE(4) A2P_CASE(5)
#undef A2P_CASE
        default: return false;
    }
}

template <int PT, typename T>
bool launch_final_t(mdft_ctx *c, const typename C2<T>::t *proj, const double *xi, double *df, int mode, FinalConsts k,
                    size_t npts = 0, size_t *grid_out = nullptr) {
    if (!npts) npts = c->N;
    const size_t ntiles = (npts + PT - 1) / PT;
    DotArgs da{};
    if (mode == 0 && c->dot_k > 0) {
        da.k = c->dot_k;
        da.self = c->dot_self;
        da.f32 = c->df32 != nullptr;
        for (int j = 0; j < 3; j++) da.p[j] = c->dot_vec[j];
    }
    // the FP32 kernel forming products over FP32 vectors (dF then FP32 too) has its own variant
    const bool dot32 = sizeof(T) == 4 && mode == 0 && da.k > 0 && da.f32 && c->xi32 && c->vext32_ok && npts % PT == 0;
    switch (c->L) {
#define FIN_CASE(LL) \
    case LL: \
        if (dot32) { \
            const size_t nb = resident_grid(c, final_t<LL, PT, T, sizeof(T) == 4>, ntiles); \
            if (grid_out) *grid_out = nb; \
            hipLaunchKernelGGL((final_t<LL, PT, T, sizeof(T) == 4>), dim3((unsigned)nb), dim3(kThreads), 0, c->stream, proj, xi, c->d_vext, df, c->d_partials, npts, c->ang, k, mode, c->df32, c->vsinv, da, c->d_dpart, c->xi32, c->d_vext32); \
        } else { \
            const size_t nb = oversubscribed_grid(c, ntiles); \
            if (grid_out) *grid_out = nb; \
            hipLaunchKernelGGL((final_t<LL, PT, T>), dim3((unsigned)nb), dim3(kThreads), 0, c->stream, proj, xi, c->d_vext, df, c->d_partials, npts, c->ang, k, mode, mode == 0 ? c->df32 : nullptr, mode == 0 ? c->vsinv : nullptr, da, c->d_dpart, mode == 0 ? c->xi32 : nullptr, nullptr); \
        } \
        return true;
        FIN_CASE(0) FIN_CASE(1) FIN_CASE(2) FIN_CASE(3) FIN_CASE(4) FIN_CASE(5)
#undef FIN_CASE
        default: return false;
    }
}

constexpr int kPTt = 4;   // points per workgroup of the specialised kernels, FP64
constexpr int kPT32 = 4;  // the same for FP32 (8 measured slower at 128^3, mmax = 5)

hipError_t launch_angl2proj_f64(mdft_ctx *c, const double *xi, double2 *proj) {
    if (c->d.mrso == 2 && launch_a2p_t<kPTt, double>(c, xi, proj)) return hipGetLastError();
    const size_t nb = (c->N + kPT - 1) / kPT;
    if (a2p_lds(c->ang) > c->lds_limit) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(angl2proj_f64<kPT>, dim3((unsigned)nb), dim3(kThreads), a2p_lds(c->ang), c->stream, xi, proj,
                       c->N, c->ang, c->rho0);
    return hipGetLastError();
}

hipError_t launch_final_f64(mdft_ctx *c, const double2 *proj, const double *xi, double *df, int mode,
                            int *nblocks_out) {
    const size_t nb = (c->N + kPT - 1) / kPT;
    if (final_lds(c->ang) > c->lds_limit) return hipErrorInvalidConfiguration;
    hipError_t e = ensure_partials(c, nb);
    if (e != hipSuccess) return e;
    FinalConsts k;
    k.rho0 = c->rho0;
    k.kT = c->d.kT;
    k.dv = c->dv;
    const double pi = acos(-1.0);
    const double fourpisq = 4.0 * pi * pi;
    k.prefactor = -c->d.kT * 2.0 * fourpisq / c->d.mrso / c->d.n0;  // cproj_mrso.f90:728
    if (c->d.mrso == 2) {
        const size_t nbt = (c->N + kPTt - 1) / kPTt;
        if ((e = ensure_partials(c, nbt)) != hipSuccess) return e;
        size_t grid = 0;
        if (launch_final_t<kPTt, double>(c, proj, xi, df, mode, k, 0, &grid)) {
            *nblocks_out = (int)grid;
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL(final_f64<kPT>, dim3((unsigned)nb), dim3(kThreads), final_lds(c->ang), c->stream, proj, xi,
                       c->d_vext, df, c->d_partials, c->N, c->ang, k, mode);
    *nblocks_out = (int)nb;
    return hipGetLastError();
}

hipError_t launch_ideal_f64(mdft_ctx *c, const double *xi, double *df, int *nblocks_out) {
    const size_t total = c->N * (size_t)c->no;
    size_t nb = (total + kThreads - 1) / kThreads;
    if (nb > 65536) nb = 65536;
    hipError_t e = ensure_partials(c, nb);
    if (e != hipSuccess) return e;
    FinalConsts k;
    k.rho0 = c->rho0;
    k.kT = c->d.kT;
    k.dv = c->dv;
    k.prefactor = 0.0;
    hipLaunchKernelGGL(ideal_f64, dim3((unsigned)nb), dim3(kThreads), 0, c->stream, xi, c->d_vext, df, c->d_partials,
                       total, c->no, c->d_w, k);
    *nblocks_out = (int)nb;
    return hipGetLastError();
}

hipError_t launch_reduce3_buf(mdft_ctx *c, const double *part, int nblocks, double *out) {
    constexpr int kSlices = 512;
    if (nblocks <= 4 * kThreads) {
        hipLaunchKernelGGL(reduce3, dim3(1), dim3(kThreads), 0, c->stream, part, nblocks, out);
        return hipGetLastError();
    }
    if (!c->d_slices) {
        hipError_t e = hipMalloc(&c->d_slices, sizeof(double) * 3 * kSlices);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(reduce3, dim3(kSlices), dim3(kThreads), 0, c->stream, part, nblocks, c->d_slices);
    hipLaunchKernelGGL(reduce3, dim3(1), dim3(kThreads), 0, c->stream, c->d_slices, kSlices, out);
    return hipGetLastError();
}

hipError_t launch_reduce3(mdft_ctx *c, int nblocks) { return launch_reduce3_buf(c, c->d_partials, nblocks, c->d_sums); }

FinalConsts final_consts(mdft_ctx *c) {
    FinalConsts k;
    k.rho0 = c->rho0;
    k.kT = c->d.kT;
    k.dv = c->dv;
    const double pi = acos(-1.0);
    k.prefactor = -c->d.kT * 2.0 * (4.0 * pi * pi) / c->d.mrso / c->d.n0;  // cproj_mrso.f90:728
    return k;
}

hipError_t launch_angl2proj_f32(mdft_ctx *c, const double *xi, float2 *proj) {
    if (c->d.mrso != 2 || !launch_a2p_t<kPT32, float>(c, xi, proj)) return hipErrorNotSupported;
    return hipGetLastError();
}

hipError_t launch_final_f32(mdft_ctx *c, const float2 *proj, const double *xi, double *df, int mode, int *nblocks_out) {
    const size_t nbt = (c->N + kPT32 - 1) / kPT32;
    hipError_t e = ensure_partials(c, nbt);
    if (e != hipSuccess) return e;
    size_t grid = 0;
    if (c->d.mrso != 2 || !launch_final_t<kPT32, float>(c, proj, xi, df, mode, final_consts(c), 0, &grid))
        return hipErrorNotSupported;
    *nblocks_out = (int)grid;
    return hipGetLastError();
}

// angl2proj of a warm-start field into the context's projection buffer (mrso = 2 kernels):
// feat 1: the feature field of src = vext; feat 2: the density target rho/rho0 - 1 of src = xi;
// feat 3: the start gradient field of src = xi with aux = dF
hipError_t launch_angl2proj_feat(mdft_ctx *c, const double *src, int feat, const double *aux) {
    if (c->d.mrso != 2) return hipErrorNotSupported;
    if (!aux) aux = c->d_vext;
    bool ok = (c->d.precision == MDFT_PREC_FP64)
                  ? launch_a2p_t<kPTt, double>(c, src, static_cast<double2 *>(c->d_proj), feat, aux)
                  : launch_a2p_t<kPT32, float>(c, src, static_cast<float2 *>(c->d_proj), feat, aux);
    return ok ? hipGetLastError() : hipErrorNotSupported;
}

// The plain FP64 transforms of npts points (any count, not only the context's grid):
// proj [npts][np] = angl2proj(f [npts][no]) and f = proj2angl(proj); mrso = 2 kernels.
hipError_t launch_transform_f64(mdft_ctx *c, const double *f, double2 *proj, size_t npts, bool to_proj) {
    if (c->d.mrso != 2) return hipErrorNotSupported;
    if (to_proj) return launch_a2p_t<kPTt, double>(c, f, proj, 4, nullptr, npts) ? hipGetLastError() : hipErrorNotSupported;
    hipError_t e = ensure_partials(c, (npts + kPTt - 1) / kPTt);
    if (e != hipSuccess) return e;
    return launch_final_t<kPTt, double>(c, proj, nullptr, const_cast<double *>(f), 5, final_consts(c), npts)
               ? hipGetLastError()
               : hipErrorNotSupported;
}

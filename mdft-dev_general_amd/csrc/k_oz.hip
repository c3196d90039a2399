// k_oz.hip -- rotation to the q frame, molecular Ornstein-Zernike contraction
// with c(|q|), rotation back, FP64.
//
// Restates the q-loop of src/module_energy_cproj_mrso.f90:441-674 for each
// (q, -q) pair: R^m_{mup,khi}(q) with the FLT_EPSILON chop of :491-492,
// rotation (:505-523), linear interpolation of c in |q| (:534-539), the OZ
// sums (:561-606), the DBL_EPSILON chop (:618-623), rotation back with conj(R)
// (:611-641) and the store rule for self-conjugate Nyquist points (:648-656).
// Both chops go through Fortran's default-kind CMPLX, so the component they
// keep is rounded to single precision; that rounding is reproduced here.
//
// One wave per pair, R by the recursion of src/module_rotation.f90:119-247.
// Pairs are pre-sorted by interpolation interval iq and cut into chunks; one
// workgroup takes one chunk, stages the two c rows it needs in LDS once, and
// its waves each take one pair at a time.  This is the kernel of any molecular
// symmetry order; order 2 (the workloads' water) runs the matrix-core kernel of
// k_ozm.hip.
#include <algorithm>
#include <array>
#include <cfloat>
#include <vector>

#include "mdft_internal.h"

namespace {


template <typename T> struct C2;
template <> struct C2<double> { using t = double2; };
template <> struct C2<float> { using t = float2; };
template <typename V> __device__ inline V mk(decltype(V::x) a, decltype(V::x) b) { V v; v.x = a; v.y = b; return v; }
template <typename V> __device__ inline V cmul(V a, V b) { return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename V> __device__ inline V cmulc(V a, V b) {  // a * conj(b)
    return mk<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

__device__ inline double2 chop_R(double2 z) {
    double re = z.x, im = z.y;
    if (fabs(re) < (double)FLT_EPSILON) { re = 0.0; im = (double)(float)im; }
    if (fabs(im) < (double)FLT_EPSILON) { im = 0.0; re = (double)(float)re; }
    return make_double2(re, im);
}

template <typename V>
__device__ inline V chop_gamma(V z) {
    using T = decltype(V::x);
    T re = z.x, im = z.y;
    if (fabs((double)re) < DBL_EPSILON && fabs((double)im) < DBL_EPSILON) { re = 0; im = 0; }
    if (fabs((double)re) < DBL_EPSILON) { re = 0; im = (T)(float)im; }
    if (fabs((double)im) < DBL_EPSILON) { im = 0; re = (T)(float)re; }
    return mk<V>(re, im);
}

__device__ inline double sqrtof(int k) { return k <= 0 ? 0.0 : sqrt((double)k); }

// compact index of R(m, mup, khi): sum_{m'<m} (2m'+1)^2 + (mup+m)(2m+1) + khi+m
__device__ inline int roff(int m) { return m * (4 * m * m - 1) / 3; }
__device__ inline int ridx(int m, int a, int b) { return roff(m) + (a + m) * (2 * m + 1) + (b + m); }

struct OzArgs {
    int nx, ny, nz, np, npc, L, H, mrso, W;
    double dq;
    const double *kx, *ky, *kz;
    const int *pidx, *pm, *pmup, *pmu, *ia0;
    const int *pq, *pmq, *pflag, *cstart, *ccount, *ciq;
    const double2 *ck;
};

// Per-wave LDS visibility: the lanes of one wave exchange data through LDS
// without a workgroup barrier (every pair is owned by exactly one wave).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// level-l entries of R read from level l-1 (compact storage), zero outside |m|,|m1| <= l-1
__device__ inline double2 Rl(const double2 *R, int l1, int m, int m1) {
    if (m < -l1 || m > l1 || m1 < -l1 || m1 > l1) return make_double2(0.0, 0.0);
    return R[ridx(l1, m, m1)];
}

// R(q) is built and kept in FP64 in both modes (it is small); the projection
// vectors, c rows and contractions use T.
template <int kWaves, typename T>
__global__ __launch_bounds__(64 * kWaves) void oz_kernel(typename C2<T>::t *__restrict__ proj, OzArgs a) {
    using V = typename C2<T>::t;
    constexpr int kThreads = 64 * kWaves;
    extern __shared__ __align__(16) unsigned char oz_smem[];
    const int np = a.np, npc = a.npc, L = a.L, mrso = a.mrso;
    const int nR = roff(L + 1);
    double2 *Rall = reinterpret_cast<double2 *>(oz_smem);
    V *c0 = reinterpret_cast<V *>(Rall + kWaves * nR), *c1 = c0 + npc;
    V *vbase = c1 + npc + (threadIdx.x >> 6) * (6 * np);
    int *tabs = reinterpret_cast<int *>(c1 + npc + kWaves * (6 * np));
    int *tm = tabs, *tmup = tm + np, *tmu2 = tmup + np, *tia0 = tmu2 + np, *poff = tia0 + np;
    double2 *R = Rall + (threadIdx.x >> 6) * nR;
    V *X0 = vbase, *X1 = X0 + np, *A = X1 + np, *B = A + np, *GQ = B + np, *GMQ = GQ + np;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int chunk = blockIdx.x;
    const int iq = a.ciq[chunk];  // 1-based lower row
    for (int i = threadIdx.x; i < npc; i += kThreads) {
        const double2 u = a.ck[(size_t)(iq - 1) * npc + i], w = a.ck[(size_t)iq * npc + i];
        c0[i] = mk<V>((T)u.x, (T)u.y);
        c1[i] = mk<V>((T)w.x, (T)w.y);
    }
    for (int i = threadIdx.x; i < np; i += kThreads) {
        tm[i] = a.pm[i];
        tmup[i] = a.pmup[i];
        tmu2[i] = a.pmu[i] / mrso;
        tia0[i] = a.ia0[i];
    }
    if (threadIdx.x == 0) {
        int o = 0;
        for (int m = 0; m <= L; m++) {
            poff[m] = o;
            o += (2 * m + 1) * (m / mrso + 1);
        }
    }
    __syncthreads();
    // projection index of (m, mup, mu2): cproj_mrso.f90:156-175 ordering
#define PIDX(m, mup, mu2) (poff[(m)] + ((mup) + (m)) * ((m) / mrso + 1) + (mu2))
    const int start = a.cstart[chunk], count = a.ccount[chunk];
    const double s2 = sqrt(2.0);
    for (int k = wv; k < count; k += kWaves) {
        const int pr = start + k;
        const int gq = a.pq[pr], gmq = a.pmq[pr], flag = a.pflag[pr];
        const int ix = gq % a.nx, iy = (gq / a.nx) % a.ny, iz = gq / (a.nx * a.ny);
        const double q0 = a.kx[ix], q1 = a.ky[iy], q2 = a.kz[iz];
        // alpha relative to the chunk's row so host and device agree on the interval
        const double alpha_d = sqrt(q0 * q0 + q1 * q1 + q2 * q2) / a.dq + 1.0 - (double)iq;
        const T alpha = (T)alpha_d, beta = (T)(1.0 - alpha_d);
        const V *src0 = proj + (size_t)gq * np, *src1 = proj + (size_t)gmq * np;
        for (int i = lane; i < np; i += 64) {
            X0[i] = src0[i];
            X1[i] = src1[i];
        }
        // ---- R(q): frame and m = 1 block (module_rotation.f90:151-193)
        double r3[3], r2[3], r1[3];
        if (q0 == 0.0 && q1 == 0.0 && q2 == 0.0) { r3[0] = 0.0; r3[1] = 0.0; r3[2] = 1.0; }
        else { r3[0] = q0; r3[1] = q1; r3[2] = q2; }
        if (r3[0] != 0.0 || r3[1] != 0.0) { r2[0] = -r3[1]; r2[1] = r3[0]; r2[2] = 0.0; }
        else { r2[0] = 0.0; r2[1] = r3[2]; r2[2] = -r3[1]; }
        const double n3 = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]);
        const double n2 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
        for (int i = 0; i < 3; i++) { r3[i] /= n3; r2[i] /= n2; }
        r1[0] = r2[1] * r3[2] - r2[2] * r3[1];
        r1[1] = r2[2] * r3[0] - r2[0] * r3[2];
        r1[2] = r2[0] * r3[1] - r2[1] * r3[0];
        // rmat(i,j) = component i of column j: (r1, r2, r3)
        const double m11 = r1[0], m21 = r1[1], m31 = r1[2];
        const double m12 = r2[0], m22 = r2[1], m32 = r2[2];
        const double m13 = r3[0], m23 = r3[1], m33 = r3[2];
        const double2 p00 = make_double2(m33, 0.0);
        const double2 pm0 = make_double2(m13 / s2, m23 / s2), pp0 = make_double2(-m13 / s2, m23 / s2);
        const double2 pm1 = make_double2((m22 - m11) * 0.5, (-m21 - m12) * 0.5);
        const double2 pp1 = make_double2((m22 + m11) * 0.5, (m12 - m21) * 0.5);
        const double2 p01 = make_double2(-m31 / s2, -m32 / s2);
        if (lane == 0) {
            R[0] = make_double2(1.0, 0.0);
            if (L >= 1) {
                R[ridx(1, -1, -1)] = make_double2((m22 + m11) * 0.5, (m21 - m12) * 0.5);
                R[ridx(1, -1, 0)] = pm0;
                R[ridx(1, -1, 1)] = pm1;
                R[ridx(1, 0, -1)] = make_double2(m31 / s2, -m32 / s2);
                R[ridx(1, 0, 0)] = p00;
                R[ridx(1, 0, 1)] = p01;
                R[ridx(1, 1, -1)] = make_double2((m22 - m11) * 0.5, (m21 + m12) * 0.5);
                R[ridx(1, 1, 0)] = pp0;
                R[ridx(1, 1, 1)] = pp1;
            }
        }
        wave_sync();
        // ---- R(q): recursion for l >= 2 (module_rotation.f90:202-245)
        for (int l = 2; l <= L; l++) {
            const int l1 = l - 1;
            const int ntask = (l + 1) * (l + 1) + l * l;
            for (int task = lane; task < ntask; task += 64) {
                // m <= 0 rows have m1 = 0..l, m > 0 rows m1 = 1..l
                int m, m1;
                const int nneg = (l + 1) * (l + 1);
                if (task < nneg) { m = -l + task / (l + 1); m1 = task % (l + 1); }
                else { const int t2 = task - nneg; m = 1 + t2 / l; m1 = 1 + t2 % l; }
                double fv, gv;
                if (m1 <= l - 1) {
                    const double am = sqrtof(l + m) * sqrtof(l - m) / (sqrtof(l + m1) * sqrtof(l - m1));
                    const double bp = sqrtof(l + m) * sqrtof(l + m - 1) / (s2 * sqrtof(l + m1) * sqrtof(l - m1));
                    const double bm = sqrtof(l - m) * sqrtof(l - m - 1) / (s2 * sqrtof(l + m1) * sqrtof(l - m1));
                    if (m == -l) {
                        const double2 u = Rl(R, l1, m + 1, m1);
                        fv = bm * (pm0.x * u.x - pm0.y * u.y);
                        gv = bm * (pm0.x * u.y + pm0.y * u.x);
                    } else if (m == l) {
                        const double2 u = Rl(R, l1, m - 1, m1);
                        fv = bp * (pp0.x * u.x - pp0.y * u.y);
                        gv = bp * (pp0.x * u.y + pp0.y * u.x);
                    } else {
                        const double2 u0 = Rl(R, l1, m, m1), um = Rl(R, l1, m - 1, m1), up = Rl(R, l1, m + 1, m1);
                        fv = am * (p00.x * u0.x) + bp * (pp0.x * um.x - pp0.y * um.y) + bm * (pm0.x * up.x - pm0.y * up.y);
                        gv = am * (p00.x * u0.y) + bp * (pp0.x * um.y + pp0.y * um.x) + bm * (pm0.x * up.y + pm0.y * up.x);
                    }
                } else {  // m1 == l
                    const double cm = s2 * sqrtof(l + m) * sqrtof(l - m) / (sqrtof(l + l) * sqrtof(l + l - 1));
                    const double dp = sqrtof(l + m) * sqrtof(l + m - 1) / (sqrtof(l + l) * sqrtof(l + l - 1));
                    const double dm = sqrtof(l - m) * sqrtof(l - m - 1) / (sqrtof(l + l) * sqrtof(l + l - 1));
                    if (m == -l) {
                        const double2 u = Rl(R, l1, m + 1, m1 - 1);
                        fv = dm * (pm1.x * u.x - pm1.y * u.y);
                        gv = dm * (pm1.x * u.y + pm1.y * u.x);
                    } else if (m == l) {
                        const double2 u = Rl(R, l1, m - 1, m1 - 1);
                        fv = dp * (pp1.x * u.x - pp1.y * u.y);
                        gv = dp * (pp1.x * u.y + pp1.y * u.x);
                    } else {
                        const double2 u0 = Rl(R, l1, m, m1 - 1), um = Rl(R, l1, m - 1, m1 - 1), up = Rl(R, l1, m + 1, m1 - 1);
                        fv = cm * (p01.x * u0.x - p01.y * u0.y) + dp * (pp1.x * um.x - pp1.y * um.y) + dm * (pm1.x * up.x - pm1.y * up.y);
                        gv = cm * (p01.x * u0.y + p01.y * u0.x) + dp * (pp1.x * um.y + pp1.y * um.x) + dm * (pm1.x * up.y + pm1.y * up.x);
                    }
                }
                const double sg = ((m + m1) & 1) ? -1.0 : 1.0;
                if (m == 0 && m1 == 0) {
                    R[ridx(l, 0, 0)] = make_double2(fv, -gv);  // the mirror statement negates g in place
                } else {
                    R[ridx(l, m, m1)] = make_double2(fv, gv);
                    R[ridx(l, -m, -m1)] = make_double2(sg * fv, -sg * gv);
                }
            }
            wave_sync();
        }
        for (int i = lane; i < nR; i += 64) R[i] = chop_R(R[i]);
        wave_sync();
        // ---- rotation to the q frame (:505-523)
        for (int j = lane; j < 2 * np; j += 64) {
            const bool first = j < np;
            const int ip = first ? j : j - np;
            const int m = tm[ip], khi = tmup[ip], mu2 = tmu2[ip];
            const V *X = first ? X0 : X1;
            const int kk = first ? khi : -khi;
            const int base = PIDX(m, -m, mu2), step = m / mrso + 1;
            const double2 *Rrow = R + ridx(m, -m, kk);
            const int rstep = 2 * m + 1;
            T sx = 0, sy = 0;
            for (int t = 0; t <= 2 * m; t++) {
                const double2 r = Rrow[t * rstep];
                const V v = cmul(X[base + t * step], mk<V>((T)r.x, (T)r.y));
                sx += v.x;
                sy += v.y;
            }
            if (first) A[ip] = mk<V>(sx, sy);
            else {
                const T sg = (m & 1) ? (T)-1 : (T)1;
                B[ip] = mk<V>(sx * sg, sy * sg);
            }
        }
        wave_sync();
        // ---- molecular OZ with interpolated c (:534-606)
        for (int j = lane; j < 2 * np; j += 64) {
            const bool first = j < np;
            const int ip = first ? j : j - np;
            const int khi = tmup[ip];
            const V *src_neg = first ? A : B;  // nu2 < 0 term
            const V *src_pos = first ? B : A;  // conj term
            int ia = tia0[ip];
            T sx = 0, sy = 0;
            for (int n = abs(khi); n <= L; n++) {
                const int h = n / mrso;
                const int base = PIDX(n, khi, 0);
                for (int nu2 = -h; nu2 <= h; nu2++, ia++) {
                    const V lo = c0[ia], hi = c1[ia];
                    const V ce = mk<V>(alpha * hi.x + beta * lo.x, alpha * hi.y + beta * lo.y);
                    V v;
                    if (nu2 < 0) {
                        v = cmul(ce, src_neg[base - nu2]);
                    } else {
                        const V x = src_pos[base + nu2];
                        v = cmul(ce, mk<V>(x.x, -x.y));
                    }
                    sx += v.x;
                    sy += v.y;
                }
            }
            if (first) GQ[ip] = chop_gamma(mk<V>(sx, sy));
            else GMQ[ip] = chop_gamma(mk<V>(sx, sy));
        }
        wave_sync();
        // ---- rotation back with conj(R) (:611-641) and store (:648-656)
        const bool eq = flag & 1, nyq = flag & 2;
        for (int j = lane; j < 2 * np; j += 64) {
            const bool first = j < np;
            if (first && eq) continue;  // for q == -q the -q store overwrites this one
            const int ip = first ? j : j - np;
            const int m = tm[ip], mup = tmup[ip], mu2 = tmu2[ip];
            const V *G = first ? GQ : GMQ;
            const int base = PIDX(m, -m, mu2), step = m / mrso + 1;
            const double2 *Rrow = R + ridx(m, mup, 0);
            T sx = 0, sy = 0;
            for (int t = 0; t <= 2 * m; t++) {
                const int khi = t - m;
                const double2 r = Rrow[first ? khi : -khi];
                const V v = cmulc(G[base + t * step], mk<V>((T)r.x, (T)r.y));
                sx += v.x;
                sy += v.y;
            }
            if (first) {
                proj[(size_t)gq * np + ip] = mk<V>(sx, sy);
            } else {
                const T sg = (m & 1) ? (T)-1 : (T)1;
                sx *= sg;
                sy *= sg;
                if (eq && nyq) sy = -sy;
                proj[(size_t)gmq * np + ip] = mk<V>(sx, sy);
            }
        }
        wave_sync();
    }
#undef PIDX
}

}  // namespace

template <typename T>
hipError_t launch_oz(mdft_ctx *c, typename C2<T>::t *proj) {
    using V = typename C2<T>::t;
    OzArgs a;
    a.nx = c->d.nx;
    a.ny = c->d.ny;
    a.nz = c->d.nz;
    a.np = c->np;
    a.npc = c->npc;
    a.L = c->L;
    a.H = c->H;
    a.mrso = c->d.mrso;
    a.W = 2 * c->L + 1;
    a.dq = c->dq;
    a.kx = c->d_kx;
    a.ky = c->d_ky;
    a.kz = c->d_kz;
    a.pidx = c->d_pidx;
    a.pm = c->d_pm;
    a.pmup = c->d_pmup;
    a.pmu = c->d_pmu;
    a.ia0 = c->d_ia0;
    a.pq = c->pairs.pq;
    a.pmq = c->pairs.pmq;
    a.pflag = c->pairs.pflag;
    a.cstart = c->pairs.cstart;
    a.ccount = c->pairs.ccount;
    a.ciq = c->pairs.ciq;
    a.ck = c->d_ck;
    const int nR = (c->L + 1) * (4 * (c->L + 1) * (c->L + 1) - 1) / 3;
    auto lds_for = [&](int waves) {
        return sizeof(double2) * (size_t)waves * nR + sizeof(V) * (2 * (size_t)c->npc + (size_t)waves * 6 * c->np) +
               sizeof(int) * (4 * (size_t)c->np + MDFT_MAX_MMAX + 2);
    };
    const dim3 grid((unsigned)c->pairs.nchunks);
    if (lds_for(8) <= c->lds_limit)
        hipLaunchKernelGGL((oz_kernel<8, T>), grid, dim3(512), lds_for(8), c->stream, proj, a);
    else if (lds_for(4) <= c->lds_limit)
        hipLaunchKernelGGL((oz_kernel<4, T>), grid, dim3(256), lds_for(4), c->stream, proj, a);
    else if (lds_for(2) <= c->lds_limit)
        hipLaunchKernelGGL((oz_kernel<2, T>), grid, dim3(128), lds_for(2), c->stream, proj, a);
    else if (lds_for(1) <= c->lds_limit)
        hipLaunchKernelGGL((oz_kernel<1, T>), grid, dim3(64), lds_for(1), c->stream, proj, a);
    else
        return hipErrorInvalidConfiguration;
    return hipGetLastError();
}

// any molecular symmetry order (the matrix-core kernel of k_ozm.hip takes order 2)
hipError_t launch_oz_f64(mdft_ctx *c, double2 *proj) { return launch_oz<double>(c, proj); }

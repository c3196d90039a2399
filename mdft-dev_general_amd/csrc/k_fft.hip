// k_fft.hip -- batched 3D complex FFTs of all projections, FP64.
//
// Replaces the FFTW plans of src/module_energy_cproj_mrso.f90:256-267 and the
// transforms at :306-329 (sign +1, unnormalised) and :690-719 (sign -1, then
// scaled by 1/(nx ny nz)).  The projection field is stored as the reference's
// deltarho_p(np,nx,ny,nz): projection index fastest.  Each pass transforms one
// axis: a workgroup takes one grid line along the axis for PB consecutive
// projections (a contiguous PB*16-byte run per grid point, so loads and stores
// are coalesced for every axis), runs a mixed-radix Stockham FFT in LDS
// (radix 8/4/2/3/5 stages, natural-order output, no bit reversal) and writes
// the line back in place.
#include "mdft_internal.h"

namespace {

constexpr int kThreads = 256;

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by sign*i
__device__ inline double2 mul_si(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }

template <int R>
__device__ inline void dft_small(double2 *v, double s);

template <>
__device__ inline void dft_small<2>(double2 *v, double) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
__device__ inline void dft_small<4>(double2 *v, double s) {
    double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    double2 b0 = cadd(v[1], v[3]), b1 = mul_si(csub(v[1], v[3]), s);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
}

template <>
__device__ inline void dft_small<3>(double2 *v, double s) {
    const double c1 = -0.5, s1 = 0.86602540378443864676;  // cos, sin of 2 pi / 3
    double2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    double2 y0 = cadd(v[0], t1);
    double2 m1 = make_double2(v[0].x + c1 * t1.x, v[0].y + c1 * t1.y);
    double2 m2 = mul_si(t2, s * s1);
    v[0] = y0;
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
}

template <>
__device__ inline void dft_small<5>(double2 *v, double s) {
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    double2 y0 = cadd(v[0], cadd(t1, t2));
    double2 a1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    double2 a2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    double2 b1 = make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    double2 b2 = make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    b1 = mul_si(b1, s);
    b2 = mul_si(b2, s);
    v[0] = y0;
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

template <>
__device__ inline void dft_small<8>(double2 *v, double s) {
    const double r2 = 0.70710678118654752440;
    // radix-2 over even/odd halves of a 4-point structure
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4>(e, s);
    dft_small<4>(o, s);
    // twiddles w8^k for k = 0..3 with sign s: exp(s i pi k / 4)
    double2 w1 = make_double2(r2, s * r2), w3 = make_double2(-r2, s * r2);
    o[1] = cmul(o[1], w1);
    o[2] = mul_si(o[2], s);
    o[3] = cmul(o[3], w3);
    for (int k = 0; k < 4; k++) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

// One Stockham stage of radix R over PB rows of length n held in LDS.
template <int R>
__device__ inline void stockham_stage(const double2 *src, double2 *dst, int n, int ld, int Ns, int rows,
                                      const double2 *__restrict__ tw, double s) {
    const int nb = n / R;
    const int tstep = n / (Ns * R);
    for (int task = threadIdx.x; task < rows * nb; task += kThreads) {
        const int row = task / nb, j = task - row * nb;
        const double2 *in = src + row * ld;
        double2 *out = dst + row * ld;
        const int k = j % Ns;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            double2 x = in[j + r * nb];
            if (r > 0 && k > 0) {
                double2 w = tw[(k * r * tstep) % n];
                w.y *= s;
                x = cmul(x, w);
            }
            v[r] = x;
        }
        dft_small<R>(v, s);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) out[base + r * Ns] = v[r];
    }
}

struct LineGeom {
    int n;          // transform length
    long long es;   // element stride (grid points)
    int na;         // size of the first other axis
    long long sa;   // stride of the first other axis (points)
    long long sb;   // stride of the second other axis (points)
};

template <int PB>
__global__ __launch_bounds__(kThreads) void fft_axis_f64(double2 *__restrict__ data, int np, LineGeom g, FftAxisPlan plan,
                                                        double sgn, double scale) {
    extern __shared__ double2 sbuf[];
    const int n = g.n, ld = n + 1;
    double2 *A = sbuf, *B = sbuf + PB * ld;
    const int line = blockIdx.x;
    const int a = line % g.na, b = line / g.na;
    const long long base = (long long)a * g.sa + (long long)b * g.sb;
    const int p0 = blockIdx.y * PB;
    const int pb = (np - p0) < PB ? (np - p0) : PB;
    for (int idx = threadIdx.x; idx < n * PB; idx += kThreads) {
        const int p = idx % PB, e = idx / PB;
        if (p < pb) A[p * ld + e] = data[(base + (long long)e * g.es) * np + p0 + p];
    }
    __syncthreads();
    double2 *src = A, *dst = B;
    int Ns = 1;
    for (int st = 0; st < plan.nst; st++) {
        const int R = plan.radix[st];
        switch (R) {
            case 8: stockham_stage<8>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            case 4: stockham_stage<4>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            case 2: stockham_stage<2>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            case 3: stockham_stage<3>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            default: stockham_stage<5>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
        }
        Ns *= R;
        __syncthreads();
        double2 *t = src;
        src = dst;
        dst = t;
    }
    for (int idx = threadIdx.x; idx < n * PB; idx += kThreads) {
        const int p = idx % PB, e = idx / PB;
        if (p < pb) {
            double2 v = src[p * ld + e];
            v.x *= scale;
            v.y *= scale;
            data[(base + (long long)e * g.es) * np + p0 + p] = v;
        }
    }
}

}  // namespace

// Radix plan: radix-8 stages while more than 8 remain, then 4, 2, 3, 5.
int fft_make_plan(int n, int *radix) {
    int k = 0;
    while (n % 8 == 0 && n > 8 && k < MDFT_FFT_MAX_STAGES) { radix[k++] = 8; n /= 8; }
    while (n % 4 == 0 && k < MDFT_FFT_MAX_STAGES) { radix[k++] = 4; n /= 4; }
    while (n % 2 == 0 && k < MDFT_FFT_MAX_STAGES) { radix[k++] = 2; n /= 2; }
    while (n % 3 == 0 && k < MDFT_FFT_MAX_STAGES) { radix[k++] = 3; n /= 3; }
    while (n % 5 == 0 && k < MDFT_FFT_MAX_STAGES) { radix[k++] = 5; n /= 5; }
    return (n == 1) ? k : -1;
}

hipError_t launch_fft3d_f64(mdft_ctx *c, double2 *proj, int sign) {
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz, np = c->np;
    const double scale_last = (sign < 0) ? 1.0 / (double)((long long)nx * ny * nz) : 1.0;
    LineGeom geo[3] = {
        {nx, 1, ny, (long long)nx, (long long)nx * ny},
        {ny, (long long)nx, nx, 1, (long long)nx * ny},
        {nz, (long long)nx * ny, nx, 1, (long long)nx},
    };
    const int PB = (np % 8 == 0) ? 8 : 4;
    for (int ax = 0; ax < 3; ax++) {
        const LineGeom &g = geo[ax];
        const size_t lines = c->N / (size_t)g.n;
        dim3 grid((unsigned)lines, (unsigned)((np + PB - 1) / PB));
        const double scale = (ax == 2) ? scale_last : 1.0;
        const size_t lds = sizeof(double2) * 2 * (size_t)PB * (g.n + 1);
        if (PB == 8)
            hipLaunchKernelGGL(fft_axis_f64<8>, grid, dim3(kThreads), lds, c->stream, proj, np, g, c->fft[ax],
                               (double)sign, scale);
        else
            hipLaunchKernelGGL(fft_axis_f64<4>, grid, dim3(kThreads), lds, c->stream, proj, np, g, c->fft[ax],
                               (double)sign, scale);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

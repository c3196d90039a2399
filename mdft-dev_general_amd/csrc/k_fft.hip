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

template <typename T> struct C2;
template <> struct C2<double> { using t = double2; };
template <> struct C2<float> { using t = float2; };

template <typename V> __device__ inline V mk(decltype(V::x) a, decltype(V::x) b) { V v; v.x = a; v.y = b; return v; }
template <typename V> __device__ inline V cmul(V a, V b) { return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename V> __device__ inline V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ inline V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
// multiply by s*i
template <typename V> __device__ inline V mul_si(V a, decltype(V::x) s) { return mk<V>(-s * a.y, s * a.x); }

template <int R, typename V>
struct Dft;

template <typename V>
struct Dft<2, V> {
    static __device__ inline void run(V *v, decltype(V::x)) {
        V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename V>
struct Dft<4, V> {
    static __device__ inline void run(V *v, decltype(V::x) s) {
        V a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        V b0 = cadd(v[1], v[3]), b1 = mul_si(csub(v[1], v[3]), s);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    }
};

template <typename V>
struct Dft<3, V> {
    static __device__ inline void run(V *v, decltype(V::x) s) {
        using T = decltype(V::x);
        const T c1 = (T)-0.5, s1 = (T)0.86602540378443864676;  // cos, sin of 2 pi / 3
        V t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        V y0 = cadd(v[0], t1);
        V m1 = mk<V>(v[0].x + c1 * t1.x, v[0].y + c1 * t1.y);
        V m2 = mul_si(t2, s * s1);
        v[0] = y0;
        v[1] = cadd(m1, m2);
        v[2] = csub(m1, m2);
    }
};

template <typename V>
struct Dft<5, V> {
    static __device__ inline void run(V *v, decltype(V::x) s) {
        using T = decltype(V::x);
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        V t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        V y0 = cadd(v[0], cadd(t1, t2));
        V a1 = mk<V>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        V a2 = mk<V>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        V b1 = mul_si(mk<V>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), s);
        V b2 = mul_si(mk<V>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), s);
        v[0] = y0;
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    }
};

template <typename V>
struct Dft<8, V> {
    static __device__ inline void run(V *v, decltype(V::x) s) {
        using T = decltype(V::x);
        const T r2 = (T)0.70710678118654752440;
        V e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, V>::run(e, s);
        Dft<4, V>::run(o, s);
        V w1 = mk<V>(r2, s * r2), w3 = mk<V>(-r2, s * r2);
        o[1] = cmul(o[1], w1);
        o[2] = mul_si(o[2], s);
        o[3] = cmul(o[3], w3);
        for (int k = 0; k < 4; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

// One Stockham stage of radix R over `rows` lines of length n held in LDS;
// twiddles read from the FP64 table and rounded to the working precision.
template <int R, typename V>
__device__ inline void stockham_stage(const V *src, V *dst, int n, int ld, int Ns, int rows,
                                      const double2 *__restrict__ tw, decltype(V::x) s) {
    using T = decltype(V::x);
    const int nb = n / R;
    const int tstep = n / (Ns * R);
    for (int task = threadIdx.x; task < rows * nb; task += kThreads) {
        const int row = task / nb, j = task - row * nb;
        const V *in = src + row * ld;
        V *out = dst + row * ld;
        const int k = j % Ns;
        V v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            V x = in[j + r * nb];
            if (r > 0 && k > 0) {
                const double2 w = tw[(k * r * tstep) % n];
                x = cmul(x, mk<V>((T)w.x, (T)w.y * s));
            }
            v[r] = x;
        }
        Dft<R, V>::run(v, s);
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
    int align;      // > 0: batches start where runs of `align` elements meet the cache-line grid
};

template <int PB, typename V>
__global__ __launch_bounds__(kThreads) void fft_axis(V *__restrict__ data, int np, LineGeom g, FftAxisPlan plan,
                                                    double sgn_d, double scale_d) {
    using T = decltype(V::x);
    extern __shared__ __align__(16) unsigned char fft_smem[];
    V *sbuf = reinterpret_cast<V *>(fft_smem);
    const T sgn = (T)sgn_d, scale = (T)scale_d;
    const int n = g.n, ld = n + 1;
    V *A = sbuf, *B = sbuf + PB * ld;
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
    V *src = A, *dst = B;
    int Ns = 1;
    for (int st = 0; st < plan.nst; st++) {
        const int R = plan.radix[st];
        switch (R) {
            case 8: stockham_stage<8, V>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            case 4: stockham_stage<4, V>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            case 2: stockham_stage<2, V>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            case 3: stockham_stage<3, V>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
            default: stockham_stage<5, V>(src, dst, n, ld, Ns, pb, plan.tw, sgn); break;
        }
        Ns *= R;
        __syncthreads();
        V *t = src;
        src = dst;
        dst = t;
    }
    for (int idx = threadIdx.x; idx < n * PB; idx += kThreads) {
        const int p = idx % PB, e = idx / PB;
        if (p < pb) {
            V v = src[p * ld + e];
            v.x *= scale;
            v.y *= scale;
            data[(base + (long long)e * g.es) * np + p0 + p] = v;
        }
    }
}


// ---- line lengths known at compile time: constant divisors and trip counts,
// twiddles staged in LDS once per workgroup.
constexpr int pick_radix(int rem) {
    return (rem % 8 == 0 && rem > 8) ? 8 : (rem % 4 == 0 ? 4 : (rem % 2 == 0 ? 2 : (rem % 3 == 0 ? 3 : 5)));
}

// In-place variant: one LDS line buffer per projection, the stages staged through
// registers.  Task t of a stage is (row = t % PB, butterfly j = t / PB), so a wave's
// consecutive lanes are consecutive projections of one grid point: the first stage
// reads its inputs straight from HBM and the last writes its outputs straight back,
// both in the 16 x 8 B (FP32) / 8 x 16 B (FP64) runs of a point record.
template <int N, int Ns, int REM, int PB, bool FIRST, typename V>
__device__ inline void run_stages(V *A, const V *twl, decltype(V::x) s, V *__restrict__ data, int np,
                                  const LineGeom &g, long long base, int p0, int pb, decltype(V::x) scale) {
    constexpr int R = pick_radix(REM), nb = N / R, ld = N + 1, tstep = N / (Ns * R);
    constexpr bool LAST = (REM == R);
    constexpr int NTASK = PB * nb, NT = (NTASK + kThreads - 1) / kThreads;
    V o[NT][R];
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int task = threadIdx.x + q * kThreads;
        if (NTASK % kThreads != 0 && task >= NTASK) continue;
        const int row = task % PB, j = task / PB, k = j % Ns;
        // FIRST: element e = j + r nb of the line at rowbase + r rstep; rows past the batch's
        // projections are transformed too but never stored: they read the last valid row, so
        // every load is issued unconditionally (a load under a branch made the wave wait for it
        // before issuing the next)
        const long long rowbase = (base + (long long)j * g.es) * np + p0 + (row < pb ? row : pb - 1);
        const long long rstep = (long long)nb * g.es * np;
        V v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int e = j + r * nb;
            V x;
            if constexpr (FIRST) {
                x = data[rowbase + r * rstep];
            } else {
                x = A[row * ld + e];
            }
            if (r > 0 && Ns > 1) {
                const V w = twl[(k * r * tstep) % N];
                x = cmul(x, mk<V>(w.x, w.y * s));
            }
            v[r] = x;
        }
        Dft<R, V>::run(v, s);
#pragma unroll
        for (int r = 0; r < R; r++) o[q][r] = v[r];
    }
    if constexpr (!FIRST) __syncthreads();  // every read of this stage precedes its in-place writes
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int task = threadIdx.x + q * kThreads;
        if (NTASK % kThreads != 0 && task >= NTASK) continue;
        const int row = task % PB, j = task / PB, k = j % Ns;
        const int ob = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if constexpr (LAST) {
                if (row < pb) {
                    V v = o[q][r];
                    v.x *= scale;
                    v.y *= scale;
                    data[(base + (long long)(ob + r * Ns) * g.es) * np + p0 + row] = v;
                }
            } else {
                A[row * ld + ob + r * Ns] = o[q][r];
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
        run_stages<N, Ns * R, REM / R, PB, false, V>(A, twl, s, data, np, g, base, p0, pb, scale);
    }
}

template <int N, int PB, typename V>
__global__ __launch_bounds__(kThreads) void fft_axis_r(V *__restrict__ data, int np, LineGeom g, const double2 *tw,
                                                      double sgn_d, double scale_d) {
    using T = decltype(V::x);
    extern __shared__ __align__(16) unsigned char fft_smem[];
    V *A = reinterpret_cast<V *>(fft_smem), *twl = A + PB * (N + 1);
    const T sgn = (T)sgn_d, scale = (T)scale_d;
    int line = blockIdx.x, by = blockIdx.y;
    if (((g.es == 1 && np % PB != 0) || g.align > 0) && ((gridDim.x * gridDim.y) & 7) == 0) {
        // cache lines shared between workgroups: along x the PB-runs of every other point
        // straddle two lines, whose other halves belong to the neighbouring batches of the
        // same line; with aligned batches (y, z) the partial lines at the ends of a point's
        // record belong to the neighbouring lines.  Workgroups go to the 8 XCDs round robin
        // in dispatch order, so each XCD (one L2) is given a contiguous range of lines with
        // the batches of each line in consecutive slots: a shared line is read from HBM once
        const unsigned id = blockIdx.x + gridDim.x * blockIdx.y, per = gridDim.x * gridDim.y / 8;
        const unsigned G = (id & 7) * per + (id >> 3);
        line = (int)(G / gridDim.y);
        by = (int)(G % gridDim.y);
    }
    const int a = line % g.na, b = line / g.na;
    const long long base = (long long)a * g.sa + (long long)b * g.sb;
    int p0, pb;
    const int s = g.align > 0 ? (int)((g.align - (base * np) % g.align) % g.align) : 0;
    if (s == 0) {
        p0 = by * PB;
        pb = (np - p0) < PB ? (np - p0) : PB;
    } else {
        p0 = by == 0 ? 0 : s + (by - 1) * PB;
        pb = by == 0 ? s : ((np - p0) < PB ? (np - p0) : PB);
    }
    if (pb <= 0) return;
    for (int i = threadIdx.x; i < N; i += kThreads) twl[i] = mk<V>((T)tw[i].x, (T)tw[i].y);
    __syncthreads();
    run_stages<N, 1, N, PB, true, V>(A, twl, sgn, data, np, g, base, p0, pb, scale);
}

template <int PB, typename V>
bool launch_axis_c(int n, dim3 grid, hipStream_t st, V *proj, int np, const LineGeom &g, const double2 *tw,
                   double sign, double scale) {
    const size_t lds = sizeof(V) * ((size_t)PB * (n + 1) + n);
    switch (n) {
#define FFT_C(NN) \
    case NN: hipLaunchKernelGGL((fft_axis_r<NN, PB, V>), grid, dim3(kThreads), lds, st, proj, np, g, tw, sign, scale); return true;
        FFT_C(16) FFT_C(32) FFT_C(48) FFT_C(64) FFT_C(96) FFT_C(128) FFT_C(192) FFT_C(256)
#undef FFT_C
        default: return false;
    }
}

template <typename V>
hipError_t launch_fft3d(mdft_ctx *c, V *proj, int sign, int batch) {
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz, np = batch > 0 ? batch : c->np;
    const double scale_last = (sign < 0) ? 1.0 / (double)((long long)nx * ny * nz) : 1.0;
    LineGeom geo[3] = {
        {nx, 1, ny, (long long)nx, (long long)nx * ny, 0},
        {ny, (long long)nx, nx, 1, (long long)nx * ny, 0},
        {nz, (long long)nx * ny, nx, 1, (long long)nx, 0},
    };
    // PB projections per workgroup: 8 (FP64) or 16 (FP32) keeps 128-byte runs per grid point
    constexpr int PBmax = sizeof(V) == 16 ? 8 : 16;
    const int PB = (np % PBmax == 0 || np > PBmax) ? PBmax : 4;
    // a point record of np elements that is not a whole number of cache lines puts the
    // runs of odd points (say) across two lines; along an axis whose point stride times
    // np is a whole number of lines, every point of a line shares the offset, and the
    // batches can start on line boundaries (one more, shorter, batch per line)
    for (int ax = 0; ax < 3; ax++)
        if (PB == PBmax && (np % PBmax) != 0 && (geo[ax].es * np) % PBmax == 0) geo[ax].align = PBmax;
    for (int ax = 0; ax < 3; ax++) {
        const LineGeom &g = geo[ax];
        const size_t lines = c->N / (size_t)g.n;
        dim3 grid((unsigned)lines, (unsigned)((np + PB - 1) / PB + (g.align > 0 ? 1 : 0)));
        const double scale = (ax == 2) ? scale_last : 1.0;
        const size_t lds = sizeof(V) * 2 * (size_t)PB * (g.n + 1);
        if (PB == PBmax && launch_axis_c<PBmax, V>(g.n, grid, c->stream, proj, np, g, c->fft[ax].tw, (double)sign, scale)) {
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            continue;
        }
        if (PB == PBmax)
            hipLaunchKernelGGL((fft_axis<PBmax, V>), grid, dim3(kThreads), lds, c->stream, proj, np, g, c->fft[ax],
                               (double)sign, scale);
        else
            hipLaunchKernelGGL((fft_axis<4, V>), grid, dim3(kThreads), lds, c->stream, proj, np, g, c->fft[ax],
                               (double)sign, scale);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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

hipError_t launch_fft3d_f64(mdft_ctx *c, double2 *proj, int sign, int batch) {
    return launch_fft3d<double2>(c, proj, sign, batch);
}

hipError_t launch_fft3d_f32(mdft_ctx *c, float2 *proj, int sign, int batch) {
    return launch_fft3d<float2>(c, proj, sign, batch);
}

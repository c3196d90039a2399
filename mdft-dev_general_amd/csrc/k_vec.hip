// k_vec.hip -- FP64 vector algebra over device-resident length-n vectors for
// the minimiser (replaces the dot/axpy/copy/scal loops of
// src/module_blas_etc.f90:8-184 that the L-BFGS of
// src/module_lbfgs_nocedal_mdft.f90 runs over n = no*nx*ny*nz).
//
// All kernels stream with 16-byte loads; dot products reduce per block in
// FP64 and then in one final block, so results are deterministic.
#include <mutex>

#include "mdft_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 4096;

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ inline void block_store_sum(double v, double *partials) {
    __shared__ double red[kThreads / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kThreads / 64; k++) s += red[k];
        partials[blockIdx.x] = s;
    }
}

// out[blockIdx] = sum x*y over the block's grid-stride share
__global__ __launch_bounds__(kThreads) void dot_partial(size_t n, const double *__restrict__ x,
                                                       const double *__restrict__ y, double *__restrict__ partials) {
    const size_t n2 = n / 2;
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (size_t)gridDim.x * kThreads) {
        const double2 a = x2[i], b = y2[i];
        s += a.x * b.x + a.y * b.y;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) s += x[n - 1] * y[n - 1];
    block_store_sum(s, partials);
}

// y += a x, then partial sums of z . y
__global__ __launch_bounds__(kThreads) void axpy_dot_partial(size_t n, double a, const double *__restrict__ x,
                                                            double *__restrict__ y, const double *__restrict__ z,
                                                            double *__restrict__ partials) {
    const size_t n2 = n / 2;
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    double2 *y2 = reinterpret_cast<double2 *>(y);
    const double2 *z2 = reinterpret_cast<const double2 *>(z);
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (size_t)gridDim.x * kThreads) {
        double2 v = y2[i];
        const double2 u = x2[i];
        v.x += a * u.x;
        v.y += a * u.y;
        y2[i] = v;
        if (z) {
            const double2 w = z2[i];
            s += w.x * v.x + w.y * v.y;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
        y[n - 1] += a * x[n - 1];
        if (z) s += z[n - 1] * y[n - 1];
    }
    block_store_sum(s, partials);
}

__global__ __launch_bounds__(kThreads) void sum_partials(const double *__restrict__ partials, int nb,
                                                        double *__restrict__ out) {
    __shared__ double red[kThreads];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kThreads) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = kThreads / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// z = a x + b y   (z may alias x or y); 16-byte accesses, two independent per trip
__global__ __launch_bounds__(kThreads) void lincomb(size_t n, double a, const double *x, double b, const double *y,
                                                   double *z) {
    const size_t n2 = n / 2, stride = (size_t)gridDim.x * kThreads;
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    double2 *z2 = reinterpret_cast<double2 *>(z);
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n2; i += 2 * stride) {
        const size_t j = i + stride;
        const double2 u = x2[i];
        double2 v = make_double2(a * u.x, a * u.y);
        double2 w = make_double2(0.0, 0.0);
        const bool two = j < n2;
        if (two) {
            const double2 u2 = x2[j];
            w = make_double2(a * u2.x, a * u2.y);
        }
        if (y) {
            const double2 t = y2[i];
            v.x += b * t.x;
            v.y += b * t.y;
            if (two) {
                const double2 t2 = y2[j];
                w.x += b * t2.x;
                w.y += b * t2.y;
            }
        }
        z2[i] = v;
        if (two) z2[j] = w;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) z[n - 1] = a * x[n - 1] + (y ? b * y[n - 1] : 0.0);
}

__global__ __launch_bounds__(kThreads) void widen(size_t n, const float2 *__restrict__ in, double2 *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float2 v = in[i];
        out[i] = make_double2((double)v.x, (double)v.y);
    }
}

// Up to kMaxVec vectors in one pass (vector-free L-BFGS): multi-dot and linear combination.
struct VecSet {
    const double *p[MDFT_VEC_MAX];
    double c[MDFT_VEC_MAX];
    int k;
    int self;  // multi-dot: bit j set when p[j] is x itself (its values are not loaded twice)
};

// part[blk][j] = sum_i x[i] v_j[i]   (16-byte accesses)
template <int K>
__global__ __launch_bounds__(kThreads) void multi_dot_partial(size_t n, const double *__restrict__ x, VecSet v,
                                                             double *__restrict__ part) {
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; j++) acc[j] = 0.0;
    const size_t n2 = n / 2;
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (size_t)gridDim.x * kThreads) {
        const double2 xv = x2[i];
#pragma unroll
        for (int j = 0; j < K; j++) {
                const double2 w = reinterpret_cast<const double2 *>(v.p[j])[i];  // x itself: a cache hit
                acc[j] += xv.x * w.x + xv.y * w.y;
            }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
#pragma unroll
        for (int j = 0; j < K; j++) acc[j] += x[n - 1] * v.p[j][n - 1];
    }
    __shared__ double red[MDFT_VEC_MAX][kThreads / 64];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const double s = wave_sum(acc[j]);
        if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; w++) s += red[threadIdx.x][w];
        part[(size_t)blockIdx.x * MDFT_VEC_MAX + threadIdx.x] = s;
    }
}

// out[j] = sum_b part[b][j], fixed order
__global__ __launch_bounds__(kThreads) void multi_dot_final(const double *__restrict__ part, int nb, int k,
                                                           double *__restrict__ out) {
    __shared__ double red[kThreads];
    for (int j = 0; j < k; j++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += kThreads) s += part[(size_t)b * MDFT_VEC_MAX + j];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int h = kThreads / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = red[0];
        __syncthreads();
    }
}

// z = sum_j c_j v_j   (z may be one of the v_j: each element is read before it is written)
template <int K>
__global__ __launch_bounds__(kThreads) void combine(size_t n, VecSet v, double *z) {
    const size_t n2 = n / 2;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (size_t)gridDim.x * kThreads) {
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int j = 0; j < K; j++) {
                const double2 w = reinterpret_cast<const double2 *>(v.p[j])[i];
                sx += v.c[j] * w.x;
                sy += v.c[j] * w.y;
            }
        reinterpret_cast<double2 *>(z)[i] = make_double2(sx, sy);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < K; j++) s += v.c[j] * v.p[j][n - 1];
        z[n - 1] = s;
    }
}

// z = sum_j c_j v_j and y = x + a z in one pass (the new search direction and the
// first trial point of the line search); y must not alias x, z or a v_j
template <int K>
__global__ __launch_bounds__(kThreads) void combine_step(size_t n, VecSet v, double *z, double a,
                                                        const double *__restrict__ x, double *__restrict__ y) {
    const size_t n2 = n / 2;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n2; i += (size_t)gridDim.x * kThreads) {
        const double2 xv = reinterpret_cast<const double2 *>(x)[i];
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int j = 0; j < K; j++) {
                const double2 w = reinterpret_cast<const double2 *>(v.p[j])[i];
                sx += v.c[j] * w.x;
                sy += v.c[j] * w.y;
            }
        reinterpret_cast<double2 *>(z)[i] = make_double2(sx, sy);
        reinterpret_cast<double2 *>(y)[i] = make_double2(xv.x + a * sx, xv.y + a * sy);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < K; j++) s += v.c[j] * v.p[j][n - 1];
        z[n - 1] = s;
        y[n - 1] = x[n - 1] + a * s;
    }
}

// ---- FP32 storage of the minimiser's directions and gradients (the mixed-precision
// path): the same passes over half the bytes, FP64 accumulation and an FP64 iterate.
// 16-byte accesses (four floats); n % 4 tail elements handled by one thread.
struct VecSet32 {
    const float *p[MDFT_VEC_MAX];
    double c[MDFT_VEC_MAX];
    int k;
    int self;
};

template <int K>
__global__ __launch_bounds__(kThreads) void multi_dot_partial32(size_t n, const float *__restrict__ x, VecSet32 v,
                                                               double *__restrict__ part) {
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; j++) acc[j] = 0.0;
    const size_t n4 = n / 4;
    const float4 *x4 = reinterpret_cast<const float4 *>(x);
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kThreads) {
        const float4 xv = x4[i];
#pragma unroll
        for (int j = 0; j < K; j++) {
                const float4 w = reinterpret_cast<const float4 *>(v.p[j])[i];  // x itself: a cache hit
                acc[j] += (double)xv.x * w.x + (double)xv.y * w.y + (double)xv.z * w.z + (double)xv.w * w.w;
            }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (size_t t = 4 * n4; t < n; t++)
#pragma unroll
            for (int j = 0; j < K; j++) acc[j] += (double)x[t] * v.p[j][t];
    __shared__ double red[MDFT_VEC_MAX][kThreads / 64];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const double s = wave_sum(acc[j]);
        if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; w++) s += red[threadIdx.x][w];
        part[(size_t)blockIdx.x * MDFT_VEC_MAX + threadIdx.x] = s;
    }
}

// z = (float) sum_j c_j v_j and y = x + a z (with z as stored, so that y - x is exactly
// a times the stored direction); y must not alias x, z or a v_j
template <int K>
__global__ __launch_bounds__(kThreads) void combine_step32(size_t n, VecSet32 v, float *z, double a,
                                                          const double *__restrict__ x, double *__restrict__ y) {
    const size_t n4 = n / 4;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kThreads) {
        const double2 x0 = reinterpret_cast<const double2 *>(x)[2 * i];
        const double2 x1 = reinterpret_cast<const double2 *>(x)[2 * i + 1];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int j = 0; j < K; j++) {
                const float4 w = reinterpret_cast<const float4 *>(v.p[j])[i];
                s0 += v.c[j] * w.x;
                s1 += v.c[j] * w.y;
                s2 += v.c[j] * w.z;
                s3 += v.c[j] * w.w;
            }
        const float4 zs = make_float4((float)s0, (float)s1, (float)s2, (float)s3);
        reinterpret_cast<float4 *>(z)[i] = zs;
        reinterpret_cast<double2 *>(y)[2 * i] = make_double2(x0.x + a * zs.x, x0.y + a * zs.y);
        reinterpret_cast<double2 *>(y)[2 * i + 1] = make_double2(x1.x + a * zs.z, x1.y + a * zs.w);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (size_t t = 4 * n4; t < n; t++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < K; j++) s += v.c[j] * v.p[j][t];
            z[t] = (float)s;
            y[t] = x[t] + a * (double)z[t];
        }
}

// y = x + a z, z FP32 (a line-search trial point along a stored direction)
__global__ __launch_bounds__(kThreads) void axpy32(size_t n, double a, const float *__restrict__ z,
                                                  const double *__restrict__ x, double *__restrict__ y) {
    const size_t n4 = n / 4;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kThreads) {
        const float4 w = reinterpret_cast<const float4 *>(z)[i];
        const double2 x0 = reinterpret_cast<const double2 *>(x)[2 * i];
        const double2 x1 = reinterpret_cast<const double2 *>(x)[2 * i + 1];
        reinterpret_cast<double2 *>(y)[2 * i] = make_double2(x0.x + a * w.x, x0.y + a * w.y);
        reinterpret_cast<double2 *>(y)[2 * i + 1] = make_double2(x1.x + a * w.z, x1.y + a * w.w);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (size_t t = 4 * n4; t < n; t++) y[t] = x[t] + a * (double)z[t];
}

// The same two passes with the iterate itself in FP32 (x, y float): y = x + a z
// formed in FP64 and rounded once.
template <int K>
__global__ __launch_bounds__(kThreads) void combine_step32x(size_t n, VecSet32 v, float *z, double a,
                                                           const float *__restrict__ x, float *__restrict__ y) {
    const size_t n4 = n / 4, stride = (size_t)gridDim.x * kThreads;
    const float4 *x4 = reinterpret_cast<const float4 *>(x);
    float4 *z4 = reinterpret_cast<float4 *>(z), *y4 = reinterpret_cast<float4 *>(y);
    auto finish = [&](size_t i, const float4 &xv, const float4 *w) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int j = 0; j < K; j++) {
            s0 += v.c[j] * w[j].x;
            s1 += v.c[j] * w[j].y;
            s2 += v.c[j] * w[j].z;
            s3 += v.c[j] * w[j].w;
        }
        const float4 zs = make_float4((float)s0, (float)s1, (float)s2, (float)s3);
        z4[i] = zs;
        y4[i] = make_float4((float)((double)xv.x + a * zs.x), (float)((double)xv.y + a * zs.y),
                            (float)((double)xv.z + a * zs.z), (float)((double)xv.w + a * zs.w));
    };
    // two element groups per iteration, all their loads issued before either is used
    size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    for (; i + stride < n4; i += 2 * stride) {
        const float4 xa = x4[i], xb = x4[i + stride];
        float4 wa[K], wb[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
            wa[j] = reinterpret_cast<const float4 *>(v.p[j])[i];
            wb[j] = reinterpret_cast<const float4 *>(v.p[j])[i + stride];
        }
        finish(i, xa, wa);
        finish(i + stride, xb, wb);
    }
    if (i < n4) {
        float4 wa[K];
#pragma unroll
        for (int j = 0; j < K; j++) wa[j] = reinterpret_cast<const float4 *>(v.p[j])[i];
        finish(i, x4[i], wa);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (size_t t = 4 * n4; t < n; t++) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < K; j++) s += v.c[j] * v.p[j][t];
            z[t] = (float)s;
            y[t] = (float)((double)x[t] + a * (double)z[t]);
        }
}

__global__ __launch_bounds__(kThreads) void axpy32x(size_t n, double a, const float *__restrict__ z, const float *x,
                                                   float *y) {
    const size_t n4 = n / 4;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kThreads) {
        const float4 w = reinterpret_cast<const float4 *>(z)[i];
        const float4 xv = reinterpret_cast<const float4 *>(x)[i];
        reinterpret_cast<float4 *>(y)[i] =
            make_float4((float)((double)xv.x + a * w.x), (float)((double)xv.y + a * w.y),
                        (float)((double)xv.z + a * w.z), (float)((double)xv.w + a * w.w));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (size_t t = 4 * n4; t < n; t++) y[t] = (float)((double)x[t] + a * (double)z[t]);
}

__global__ __launch_bounds__(kThreads) void narrow1(size_t n, const double *__restrict__ in, float *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        out[i] = (float)in[i];
}

__global__ __launch_bounds__(kThreads) void widen1(size_t n, const float *__restrict__ in, double *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        out[i] = (double)in[i];
}

// out[i] = in[i] fac[i % no] (in FP64 or FP32): the per-orientation change of variables
__global__ __launch_bounds__(kThreads) void scale_orient(size_t n, int no, const double *__restrict__ fac,
                                                        const double *__restrict__ in64, const float *__restrict__ in32,
                                                        double *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const double v = in32 ? (double)in32[i] : in64[i];
        out[i] = v * fac[i % (size_t)no];
    }
}

// Per grid point (one wave each): n(r) = rho0 sum_io w_io xi^2; block partials of
// sum_r n(r) and of the number of points with n(r) / n0 < cav (cavity points).
__global__ __launch_bounds__(kThreads) void density_integrals(const double *__restrict__ xi,
                                                             const double *__restrict__ w, size_t N, int no,
                                                             double rho0, double n0, double cav,
                                                             double *__restrict__ partials) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double sum = 0.0, ncav = 0.0;
    for (size_t p = (size_t)blockIdx.x * (kThreads / 64) + wv; p < N; p += (size_t)gridDim.x * (kThreads / 64)) {
        const double *row = xi + p * no;
        double s = 0.0;
        for (int io = lane; io < no; io += 64) s += w[io] * row[io] * row[io];
        s = wave_sum(s) * rho0;
        if (lane == 0) {
            sum += s;
            ncav += (s / n0 < cav) ? 1.0 : 0.0;
        }
    }
    __shared__ double red[2][kThreads / 64];
    if (lane == 0) {
        red[0][wv] = sum;
        red[1][wv] = ncav;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < kThreads / 64; k++) {
            a += red[0][k];
            b += red[1][k];
        }
        partials[3 * blockIdx.x + 0] = a;
        partials[3 * blockIdx.x + 1] = b;
        partials[3 * blockIdx.x + 2] = 0.0;
    }
}

// Launches kernel<K> for the runtime vector count k, 1 <= k <= MDFT_VEC_MAX (checked at the
// boundary): with K known the loop over the vectors has no branches, so a thread's loads of
// one element group are all in flight together instead of one wait per vector.
#define VEC_K_DISPATCH(k, LAUNCH)                                                        \
    switch (k) {                                                                          \
        case 1: LAUNCH(1); break;   case 2: LAUNCH(2); break;   case 3: LAUNCH(3); break;   \
        case 4: LAUNCH(4); break;   case 5: LAUNCH(5); break;   case 6: LAUNCH(6); break;   \
        case 7: LAUNCH(7); break;   case 8: LAUNCH(8); break;   case 9: LAUNCH(9); break;   \
        case 10: LAUNCH(10); break; case 11: LAUNCH(11); break; case 12: LAUNCH(12); break; \
        case 13: LAUNCH(13); break; case 14: LAUNCH(14); break; case 15: LAUNCH(15); break; \
        case 16: LAUNCH(16); break;                                                       \
        default: return hipErrorInvalidValue;                                             \
    }
static_assert(MDFT_VEC_MAX == 16, "VEC_K_DISPATCH lists 1..16");

unsigned blocks_for(size_t n) {
    size_t b = (n / 2 + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    if (b > kMaxBlocks) b = kMaxBlocks;
    return (unsigned)b;
}

// blocks_for capped at the workgroups the compute units hold at once for this kernel: a
// grid-stride kernel given more ends on a partial round of workgroups, at a fraction of the
// memory parallelism
template <typename K>
unsigned resident_blocks(const mdft_ctx *c, K kernel, size_t n) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    const size_t g = (size_t)per_cu * (size_t)(c->num_cus > 0 ? c->num_cus : 256);
    const unsigned b = blocks_for(n);
    return g < b ? (unsigned)g : b;
}

hipError_t ensure_vec_partials(mdft_ctx *c) {
    if (c->d_vpart) return hipSuccess;
    return hipMalloc(&c->d_vpart, sizeof(double) * (kMaxBlocks + 8));
}

hipError_t finish_sum(mdft_ctx *c, unsigned nb, double *out) {
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kThreads), 0, c->stream, c->d_vpart, (int)nb, c->d_vpart + kMaxBlocks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(c->h_sums + 4, c->d_vpart + kMaxBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) *out = c->h_sums[4];
    return e;
}

}  // namespace

hipError_t vec_dot(mdft_ctx *c, size_t n, const double *x, const double *y, double *out) {
    hipError_t e = ensure_vec_partials(c);
    if (e != hipSuccess) return e;
    const unsigned nb = resident_blocks(c, dot_partial, n);
    hipLaunchKernelGGL(dot_partial, dim3(nb), dim3(kThreads), 0, c->stream, n, x, y, c->d_vpart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return finish_sum(c, nb, out);
}

hipError_t vec_axpy_dot(mdft_ctx *c, size_t n, double a, const double *x, double *y, const double *z, double *out) {
    hipError_t e = ensure_vec_partials(c);
    if (e != hipSuccess) return e;
    const unsigned nb = resident_blocks(c, axpy_dot_partial, n);
    hipLaunchKernelGGL(axpy_dot_partial, dim3(nb), dim3(kThreads), 0, c->stream, n, a, x, y, z, c->d_vpart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (!z || !out) return hipStreamSynchronize(c->stream);
    return finish_sum(c, nb, out);
}

hipError_t vec_lincomb(mdft_ctx *c, size_t n, double a, const double *x, double b, const double *y, double *z) {
    const unsigned nb = resident_blocks(c, lincomb, 2 * n);
    hipLaunchKernelGGL(lincomb, dim3(nb), dim3(kThreads), 0, c->stream, n, a, x, b, y, z);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t launch_widen_f32(mdft_ctx *c, const float2 *in, double2 *out, size_t n) {
    hipLaunchKernelGGL(widen, dim3(resident_blocks(c, widen, 2 * n)), dim3(kThreads), 0, c->stream, n, in, out);
    return hipGetLastError();
}

hipError_t vec_multi_dot(mdft_ctx *c, size_t n, int k, const double *const *vecs, const double *x, double *out) {
    if (!c->d_mpart) {
        hipError_t e = hipMalloc(&c->d_mpart, sizeof(double) * ((size_t)kMaxBlocks * MDFT_VEC_MAX + MDFT_VEC_MAX));
        if (e != hipSuccess) return e;
    }
    VecSet v{};
    for (int j = 0; j < k; j++) {
        v.p[j] = vecs[j];
        if (vecs[j] == x) v.self |= 1 << j;
    }
    v.k = k;
    unsigned nb = 1;
#define L_(K) nb = resident_blocks(c, multi_dot_partial<K>, n); hipLaunchKernelGGL(multi_dot_partial<K>, dim3(nb), dim3(kThreads), 0, c->stream, n, x, v, c->d_mpart)
    VEC_K_DISPATCH(k, L_)
#undef L_
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    double *dout = c->d_mpart + (size_t)kMaxBlocks * MDFT_VEC_MAX;
    hipLaunchKernelGGL(multi_dot_final, dim3(1), dim3(kThreads), 0, c->stream, c->d_mpart, (int)nb, k, dout);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(out, dout, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream)) != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_combine(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z) {
    VecSet v{};
    for (int j = 0; j < k; j++) {
        v.p[j] = vecs[j];
        v.c[j] = coef[j];
    }
    v.k = k;
#define L_(K) hipLaunchKernelGGL(combine<K>, dim3(resident_blocks(c, combine<K>, 2 * n)), dim3(kThreads), 0, c->stream, n, v, z)
    VEC_K_DISPATCH(k, L_)
#undef L_
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_combine_step(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z,
                            double a, const double *x, double *y) {
    VecSet v{};
    for (int j = 0; j < k; j++) {
        v.p[j] = vecs[j];
        v.c[j] = coef[j];
    }
    v.k = k;
#define L_(K) hipLaunchKernelGGL(combine_step<K>, dim3(resident_blocks(c, combine_step<K>, 2 * n)), dim3(kThreads), 0, c->stream, n, v, z, a, x, y)
    VEC_K_DISPATCH(k, L_)
#undef L_
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

// ---- multi-dot without device allocation -------------------------------------------
// The per-block partial products [kMaxBlocks][MDFT_VEC_MAX] and the k results of
// vec_multi_dot_noalloc live in this fixed array (512 KiB) instead of memory the
// library allocates on first use.  Each device holds its own copy until the process
// ends; host code uses only the device address hipGetSymbolAddress returns for it.
constexpr size_t kScratchDoubles = (size_t)kMaxBlocks * MDFT_VEC_MAX + MDFT_VEC_MAX;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kScratchDevices] = {};
}  // namespace

hipError_t vec_scratch_init(int device) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current first
    double *p = nullptr;
    e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

size_t vec_scratch_bytes() { return sizeof(double) * kScratchDoubles; }

hipError_t vec_multi_dot_noalloc(mdft_ctx *c, size_t n, int k, const double *const *vecs, const double *x,
                                 double *out) {
    double *s = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c->d.device >= 0 && c->d.device < kScratchDevices) s = g_scratch_ptr[c->d.device];
    }
    if (!s) return hipErrorNotInitialized;
    VecSet v{};
    for (int j = 0; j < k; j++) {
        v.p[j] = vecs[j];
        if (vecs[j] == x) v.self |= 1 << j;
    }
    v.k = k;
    unsigned nb = 1;  // <= kMaxBlocks: every partial fits the array
#define L_(K) nb = resident_blocks(c, multi_dot_partial<K>, n); hipLaunchKernelGGL(multi_dot_partial<K>, dim3(nb), dim3(kThreads), 0, c->stream, n, x, v, s)
    VEC_K_DISPATCH(k, L_)
#undef L_
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    double *dout = s + (size_t)kMaxBlocks * MDFT_VEC_MAX;
    hipLaunchKernelGGL(multi_dot_final, dim3(1), dim3(kThreads), 0, c->stream, s, (int)nb, k, dout);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(out, dout, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream)) != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_multi_dot32(mdft_ctx *c, size_t n, int k, const float *const *vecs, const float *x, double *out) {
    if (!c->d_mpart) {
        hipError_t e = hipMalloc(&c->d_mpart, sizeof(double) * ((size_t)kMaxBlocks * MDFT_VEC_MAX + MDFT_VEC_MAX));
        if (e != hipSuccess) return e;
    }
    VecSet32 v{};
    for (int j = 0; j < k; j++) {
        v.p[j] = vecs[j];
        if (vecs[j] == x) v.self |= 1 << j;
    }
    v.k = k;
    unsigned nb = 1;
#define L_(K) nb = resident_blocks(c, multi_dot_partial32<K>, n / 2); hipLaunchKernelGGL(multi_dot_partial32<K>, dim3(nb), dim3(kThreads), 0, c->stream, n, x, v, c->d_mpart)
    VEC_K_DISPATCH(k, L_)
#undef L_
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    double *dout = c->d_mpart + (size_t)kMaxBlocks * MDFT_VEC_MAX;
    hipLaunchKernelGGL(multi_dot_final, dim3(1), dim3(kThreads), 0, c->stream, c->d_mpart, (int)nb, k, dout);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(out, dout, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream)) != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_combine_step32(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                              double a, const double *x, double *y) {
    VecSet32 v{};
    for (int j = 0; j < k; j++) {
        v.p[j] = vecs[j];
        v.c[j] = coef[j];
    }
    v.k = k;
#define L_(K) hipLaunchKernelGGL(combine_step32<K>, dim3(resident_blocks(c, combine_step32<K>, n / 2)), dim3(kThreads), 0, c->stream, n, v, z, a, x, y)
    VEC_K_DISPATCH(k, L_)
#undef L_
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_axpy32(mdft_ctx *c, size_t n, double a, const float *z, const double *x, double *y) {
    hipLaunchKernelGGL(axpy32, dim3(resident_blocks(c, axpy32, n / 2)), dim3(kThreads), 0, c->stream, n, a, z, x, y);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_combine_step32x(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                               double a, const float *x, float *y) {
    VecSet32 v{};
    for (int j = 0; j < k; j++) {
        v.p[j] = vecs[j];
        v.c[j] = coef[j];
    }
    v.k = k;
#define L_(K) hipLaunchKernelGGL(combine_step32x<K>, dim3(resident_blocks(c, combine_step32x<K>, n / 2)), dim3(kThreads), 0, c->stream, n, v, z, a, x, y)
    VEC_K_DISPATCH(k, L_)
#undef L_
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_axpy32x(mdft_ctx *c, size_t n, double a, const float *z, const float *x, float *y) {
    hipLaunchKernelGGL(axpy32x, dim3(resident_blocks(c, axpy32x, n / 2)), dim3(kThreads), 0, c->stream, n, a, z, x, y);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_narrow(mdft_ctx *c, size_t n, const double *in, float *out) {
    hipLaunchKernelGGL(narrow1, dim3(resident_blocks(c, narrow1, 2 * n)), dim3(kThreads), 0, c->stream, n, in, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_widen(mdft_ctx *c, size_t n, const float *in, double *out) {
    hipLaunchKernelGGL(widen1, dim3(resident_blocks(c, widen1, 2 * n)), dim3(kThreads), 0, c->stream, n, in, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t vec_scale_orient(mdft_ctx *c, size_t n, int no, const double *fac, const void *in, bool in_f32, double *out) {
    hipLaunchKernelGGL(scale_orient, dim3(resident_blocks(c, scale_orient, 2 * n)), dim3(kThreads), 0, c->stream, n, no, fac,
                       in_f32 ? nullptr : static_cast<const double *>(in), in_f32 ? static_cast<const float *>(in) : nullptr,
                       out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

hipError_t launch_density_integrals(mdft_ctx *c, const double *xi, double cav, int *nblocks_out) {
    const unsigned nb = 2048;
    if (c->partials_cap < 3 * (size_t)nb) {
        (void)hipFree(c->d_partials);
        c->d_partials = nullptr;
        c->partials_cap = 0;
        hipError_t e = hipMalloc(&c->d_partials, sizeof(double) * 3 * nb);
        if (e != hipSuccess) return e;
        c->partials_cap = 3 * (size_t)nb;
    }
    hipLaunchKernelGGL(density_integrals, dim3(nb), dim3(kThreads), 0, c->stream, xi, c->d_w, c->N, c->no, c->rho0,
                       c->d.n0, cav, c->d_partials);
    *nblocks_out = (int)nb;
    return hipGetLastError();
}

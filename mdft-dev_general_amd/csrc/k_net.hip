// k_net.hip -- the warm-start network: periodic dilated 3D convolutions
// (3x3x3, up to 32 channels) and pointwise (1x1x1, up to 256 channels) layers
// with leaky ReLU (slope 0.1), masked mean-squared loss, backward passes and
// Adam, all FP32.
//
// The reference has no learned component.  The network maps per-point
// features of the external potential to the low angular projections of the
// converged OZ field; the minimiser starts from the density that field
// implies.  Activations are stored channel-major [C][nz][ny][nx] (x fastest),
// so every channel of every tap is a contiguous, coalesced stream; the box is
// periodic, so the convolutions wrap.
//
//   forward      one thread per grid point, 32 output channels in registers, the
//                workgroup's weight slice [Cin][taps][32] in LDS read as 16-byte
//                groups (broadcast): 32 FMAs per gathered input value and 4 per
//                LDS read (KG = 8 channels and single-float reads where the slice
//                does not fit the device's LDS)
//   data grad    the same shape with input channels in registers
//   weight grad  one thread per (input channel, tap) with all output channels
//                in registers, over a chunk of points whose output gradients
//                sit in LDS; per-chunk partial sums, then a fixed-order sum
//   Adam         Kingma & Ba (2015), bias-corrected, one thread per parameter
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mdft_internal.h"
#include "mdft_wsnet.h"

namespace {

constexpr int kThreads = 256;
constexpr int KG = 8;            // output (forward) / input (data grad) channels per thread
constexpr int kMaxC3 = 32;       // channels of a 3x3x3 layer
constexpr int kMaxC1 = 256;      // channels of a pointwise layer
constexpr int kWChunk = 1024;    // points per workgroup in the weight gradient
constexpr int kWSub = 128;       // points staged in LDS at a time
constexpr float kLeak = 0.1f;    // slope of the leaky ReLU for negative inputs

struct Geo {
    int nx, ny, nz;
};

__device__ inline int wrap(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// neighbour offsets of point p for dilation d: idx[t] for t = (tz*3+ty)*3+tx, sign s
template <int TAPS>
__device__ inline void taps(int p, Geo g, int d, int s, int *idx) {
    if constexpr (TAPS == 1) {
        idx[0] = p;
        return;
    }
    const int x = p % g.nx, y = (p / g.nx) % g.ny, z = p / (g.nx * g.ny);
    int xs[3], ys[3], zs[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
        xs[t] = wrap(x + s * (t - 1) * d, g.nx);
        ys[t] = wrap(y + s * (t - 1) * d, g.ny) * g.nx;
        zs[t] = wrap(z + s * (t - 1) * d, g.nz) * g.nx * g.ny;
    }
#pragma unroll
    for (int tz = 0; tz < 3; tz++)
#pragma unroll
        for (int ty = 0; ty < 3; ty++)
#pragma unroll
            for (int tx = 0; tx < 3; tx++) idx[(tz * 3 + ty) * 3 + tx] = zs[tz] + ys[ty] + xs[tx];
}

// out[co][p] = b[co] + sum_ci sum_t W[co][ci][t] in[ci][p + (t-1) d], optional ReLU
template <int TAPS>
__global__ __launch_bounds__(kThreads) void conv_fwd(const float *__restrict__ in, const float *__restrict__ W,
                                                    const float *__restrict__ b, float *__restrict__ out, int cin,
                                                    int cout, Geo g, int d, int relu) {
    extern __shared__ float ws[];  // [KG][cin][TAPS]
    const int N = g.nx * g.ny * g.nz;
    const int co0 = blockIdx.y * KG;
    const int kg = min(KG, cout - co0);
    for (int i = threadIdx.x; i < kg * cin * TAPS; i += kThreads) ws[i] = W[(size_t)co0 * cin * TAPS + i];
    __syncthreads();
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= N) return;
    int idx[TAPS];
    taps<TAPS>(p, g, d, 1, idx);
    float acc[KG];
#pragma unroll
    for (int k = 0; k < KG; k++) acc[k] = 0.0f;
    for (int ci = 0; ci < cin; ci++) {
        const float *src = in + (size_t)ci * N;
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            const float v = src[idx[t]];
#pragma unroll
            for (int k = 0; k < KG; k++)
                if (k < kg) acc[k] += ws[(k * cin + ci) * TAPS + t] * v;
        }
    }
#pragma unroll
    for (int k = 0; k < KG; k++) {
        if (k < kg) {
            float v = acc[k] + b[co0 + k];
            if (relu && v < 0.0f) v *= kLeak;
            out[(size_t)(co0 + k) * N + p] = v;
        }
    }
}

// The same with KW output channels per thread and the weights stored [cin][TAPS][KW]
// (zero past cout), so each LDS read brings four weights of one (ci, t).
constexpr int KW = 32;
constexpr int kWideThreads = 1024;

template <int TAPS>
__global__ __launch_bounds__(kWideThreads) void conv_fwd_wide(const float *__restrict__ in, const float *__restrict__ W,
                                                             const float *__restrict__ b, float *__restrict__ out,
                                                             int cin, int cout, Geo g, int d, int relu) {
    extern __shared__ float4 ws4[];  // [cin][TAPS][KW / 4]
    float *ws = reinterpret_cast<float *>(ws4);
    const int N = g.nx * g.ny * g.nz;
    const int co0 = blockIdx.y * KW;
    const int kg = min(KW, cout - co0);
    for (int i = threadIdx.x; i < cin * TAPS * KW; i += kWideThreads) {
        const int k = i % KW, r = i / KW;  // r = ci * TAPS + t
        ws[i] = (k < kg) ? W[(size_t)(co0 + k) * cin * TAPS + r] : 0.0f;
    }
    __syncthreads();
    const int p = blockIdx.x * kWideThreads + threadIdx.x;
    if (p >= N) return;
    int idx[TAPS];
    taps<TAPS>(p, g, d, 1, idx);
    float acc[KW];
#pragma unroll
    for (int k = 0; k < KW; k++) acc[k] = 0.0f;
    for (int ci = 0; ci < cin; ci++) {
        const float *src = in + (size_t)ci * N;
        const float4 *wr = ws4 + (size_t)ci * TAPS * (KW / 4);
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            const float v = src[idx[t]];
#pragma unroll
            for (int q = 0; q < KW / 4; q++) {
                const float4 w = wr[t * (KW / 4) + q];
                acc[4 * q + 0] += w.x * v;
                acc[4 * q + 1] += w.y * v;
                acc[4 * q + 2] += w.z * v;
                acc[4 * q + 3] += w.w * v;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < KW; k++) {
        if (k < kg) {
            float v = acc[k] + b[co0 + k];
            if (relu && v < 0.0f) v *= kLeak;
            out[(size_t)(co0 + k) * N + p] = v;
        }
    }
}

// gradient through the leaky ReLU of a layer whose output is a: g *= (a > 0 ? 1 : kLeak)
__global__ __launch_bounds__(kThreads) void relu_bwd(float *__restrict__ gr, const float *__restrict__ a, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        if (!(a[i] > 0.0f)) gr[i] *= kLeak;
}

// din[ci][p] = sum_co sum_t W[co][ci][t] dz[co][p - (t-1) d]
template <int TAPS>
__global__ __launch_bounds__(kThreads) void conv_bwd_data(const float *__restrict__ dz, const float *__restrict__ W,
                                                         float *__restrict__ din, int cin, int cout, Geo g, int d) {
    extern __shared__ float ws[];  // [cout][KG][TAPS]
    const int N = g.nx * g.ny * g.nz;
    const int ci0 = blockIdx.y * KG;
    const int kg = min(KG, cin - ci0);
    for (int i = threadIdx.x; i < cout * KG * TAPS; i += kThreads) {
        const int co = i / (KG * TAPS), r = i % (KG * TAPS), k = r / TAPS, t = r % TAPS;
        ws[i] = (k < kg) ? W[((size_t)co * cin + ci0 + k) * TAPS + t] : 0.0f;
    }
    __syncthreads();
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= N) return;
    int idx[TAPS];
    taps<TAPS>(p, g, d, -1, idx);
    float acc[KG];
#pragma unroll
    for (int k = 0; k < KG; k++) acc[k] = 0.0f;
    for (int co = 0; co < cout; co++) {
        const float *src = dz + (size_t)co * N;
#pragma unroll
        for (int t = 0; t < TAPS; t++) {
            const float v = src[idx[t]];
#pragma unroll
            for (int k = 0; k < KG; k++) acc[k] += ws[(co * KG + k) * TAPS + t] * v;
        }
    }
#pragma unroll
    for (int k = 0; k < KG; k++)
        if (k < kg) din[(size_t)(ci0 + k) * N + p] = acc[k];
}

// Partial weight and bias gradients of one chunk of points:
// part[blk][(co * cin + ci) * 27 + t] = sum_p dz[co][p] a[ci][p + (t-1) d],
// part[blk][cout * cin * 27 + co] = sum_p dz[co][p].
template <int CO>
__global__ __launch_bounds__(1024) void conv_bwd_weight(const float *__restrict__ dz, const float *__restrict__ a,
                                                       float *__restrict__ part, int cin, int cout, Geo g, int d) {
    __shared__ float gs[CO * kWSub];
    const int N = g.nx * g.ny * g.nz;
    const int nW = cin * 27;
    const int tid = threadIdx.x;
    const bool wthread = tid < nW, bthread = tid == nW;
    const int ci = wthread ? tid / 27 : 0, t = wthread ? tid % 27 : 13;
    const int tz = t / 9, ty = (t / 3) % 3, tx = t % 3;
    float acc[CO];
#pragma unroll
    for (int k = 0; k < CO; k++) acc[k] = 0.0f;
    const int p0 = blockIdx.x * kWChunk;
    const int pend = min(N, p0 + kWChunk);
    for (int s0 = p0; s0 < pend; s0 += kWSub) {
        const int ns = min(kWSub, pend - s0);
        __syncthreads();
        for (int i = tid; i < CO * kWSub; i += blockDim.x) {
            const int co = i / kWSub, j = i % kWSub;
            gs[i] = (co < cout && j < ns) ? dz[(size_t)co * N + s0 + j] : 0.0f;
        }
        __syncthreads();
        if (wthread) {
            const float *src = a + (size_t)ci * N;
            for (int j = 0; j < ns; j++) {
                const int p = s0 + j;
                const int x = p % g.nx, y = (p / g.nx) % g.ny, z = p / (g.nx * g.ny);
                const int q = (wrap(z + (tz - 1) * d, g.nz) * g.ny + wrap(y + (ty - 1) * d, g.ny)) * g.nx +
                              wrap(x + (tx - 1) * d, g.nx);
                const float v = src[q];
#pragma unroll
                for (int k = 0; k < CO; k++) acc[k] += gs[k * kWSub + j] * v;
            }
        } else if (bthread) {
            for (int j = 0; j < ns; j++)
#pragma unroll
                for (int k = 0; k < CO; k++) acc[k] += gs[k * kWSub + j];
        }
    }
    float *dst = part + (size_t)blockIdx.x * (cout * nW + cout);
    if (wthread) {
#pragma unroll
        for (int k = 0; k < CO; k++)
            if (k < cout) dst[(size_t)k * nW + tid] = acc[k];
    } else if (bthread) {
#pragma unroll
        for (int k = 0; k < CO; k++)
            if (k < cout) dst[(size_t)cout * nW + k] = acc[k];
    }
}

// Pointwise layer: part[blk][co * cin + ci] = sum_p dz[co][p] a[ci][p],
// part[blk][cout * cin + co] = sum_p dz[co][p]; one thread per input channel
// (plus one for the bias), 32 output channels per workgroup (blockIdx.y).
__global__ __launch_bounds__(320) void pw_bwd_weight(const float *__restrict__ dz, const float *__restrict__ a,
                                                    float *__restrict__ part, int cin, int cout, int N) {
    constexpr int CO = 32;
    __shared__ float gs[CO * kWSub];
    const int tid = threadIdx.x;
    const int co0 = blockIdx.y * CO;
    const bool wthread = tid < cin, bthread = tid == cin;
    float acc[CO];
#pragma unroll
    for (int k = 0; k < CO; k++) acc[k] = 0.0f;
    const int p0 = blockIdx.x * kWChunk;
    const int pend = min(N, p0 + kWChunk);
    for (int s0 = p0; s0 < pend; s0 += kWSub) {
        const int ns = min(kWSub, pend - s0);
        __syncthreads();
        for (int i = tid; i < CO * kWSub; i += blockDim.x) {
            const int k = i / kWSub, j = i % kWSub;
            gs[i] = (co0 + k < cout && j < ns) ? dz[(size_t)(co0 + k) * N + s0 + j] : 0.0f;
        }
        __syncthreads();
        if (wthread) {
            const float *src = a + (size_t)tid * N + s0;
            for (int j = 0; j < ns; j++) {
                const float v = src[j];
#pragma unroll
                for (int k = 0; k < CO; k++) acc[k] += gs[k * kWSub + j] * v;
            }
        } else if (bthread) {
            for (int j = 0; j < ns; j++)
#pragma unroll
                for (int k = 0; k < CO; k++) acc[k] += gs[k * kWSub + j];
        }
    }
    float *dst = part + (size_t)blockIdx.x * (cout * cin + cout);
#pragma unroll
    for (int k = 0; k < CO; k++) {
        if (co0 + k >= cout) break;
        if (wthread) dst[(size_t)(co0 + k) * cin + tid] = acc[k];
        else if (bthread) dst[(size_t)cout * cin + co0 + k] = acc[k];
    }
}

// out[i] = sum_b part[b][i] (fixed order)
__global__ __launch_bounds__(kThreads) void sum_parts(const float *__restrict__ part, int nblk, int n,
                                                     float *__restrict__ out) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * n + i];
    out[i] = s;
}

// masked MSE: r = y - t, g = 2 m r / cnt; block partials of sum m r^2 (FP64)
__global__ __launch_bounds__(kThreads) void mse(const float *__restrict__ y, const float *__restrict__ tg,
                                               const float *__restrict__ mask, float *__restrict__ gout, int C, int N,
                                               double inv_cnt, double *__restrict__ part) {
    double s = 0.0;
    const size_t n = (size_t)C * N;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float m = mask ? mask[i % N] : 1.0f;
        const float r = y[i] - tg[i];
        s += (double)m * (double)r * (double)r;
        gout[i] = (float)(2.0 * inv_cnt) * m * r;
    }
    __shared__ double red[kThreads];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = kThreads / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void adam(float *__restrict__ p, const float *__restrict__ gr,
                                                float *__restrict__ m, float *__restrict__ v, size_t n, float lr,
                                                float b1, float b2, float eps, float c1, float c2) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float g = gr[i];
        const float mi = b1 * m[i] + (1.0f - b1) * g;
        const float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

unsigned grid1(size_t n) {
    size_t b = (n + kThreads - 1) / kThreads;
    return (unsigned)(b > 4096 ? 4096 : (b ? b : 1));
}

}  // namespace

struct mdft_wsnet {
    int device = 0;
    size_t lds_limit = 65536;  // the device's LDS per workgroup
    int nl = 0;
    std::vector<int> ch, dil, taps;  // taps: 27 (3x3x3) or 1 (pointwise)
    std::vector<size_t> woff, boff;  // offsets of W_l and b_l in the parameter vector
    size_t nparam = 0;
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    int nx = 0, ny = 0, nz = 0;
    std::vector<float *> act;  // act[0] unused (input is the caller's), act[l] = output of layer l
    float *gA = nullptr, *gB = nullptr;  // gradient ping-pong [maxC][N]
    float *part = nullptr;
    size_t part_cap = 0;
    double *dpart = nullptr, *hpart = nullptr;
    long long step = 0;
    const float *fwd_in = nullptr;  // input of the last forward pass (its activations are kept)
    hipStream_t stream = nullptr;
    std::string err;
};

namespace {

thread_local std::string g_ws_create_error;

int ws_fail(mdft_wsnet *n, int code, const std::string &msg) {
    n->err = msg;
    return code;
}

#define WSCHK(n, expr)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return ws_fail(n, MDFT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void free_acts(mdft_wsnet *n) {
    for (float *p : n->act) (void)hipFree(p);
    n->act.clear();
    (void)hipFree(n->gA);
    (void)hipFree(n->gB);
    n->gA = n->gB = nullptr;
    n->nx = n->ny = n->nz = 0;
}

int ensure_grid(mdft_wsnet *n, int nx, int ny, int nz) {
    if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * ny * nz >= (1LL << 31))
        return ws_fail(n, MDFT_EINVAL, "invalid grid");
    if (n->nx == nx && n->ny == ny && n->nz == nz) return MDFT_OK;
    free_acts(n);
    const size_t N = (size_t)nx * ny * nz;
    int cmax = 0;
    for (int c : n->ch) cmax = std::max(cmax, c);
    n->act.assign(n->nl + 1, nullptr);
    for (int l = 1; l <= n->nl; l++) WSCHK(n, hipMalloc(&n->act[l], sizeof(float) * N * n->ch[l]));
    WSCHK(n, hipMalloc(&n->gA, sizeof(float) * N * cmax));
    WSCHK(n, hipMalloc(&n->gB, sizeof(float) * N * cmax));
    n->nx = nx;
    n->ny = ny;
    n->nz = nz;
    return MDFT_OK;
}

int forward(mdft_wsnet *n, const float *in) {
    n->fwd_in = in;
    const Geo g{n->nx, n->ny, n->nz};
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    const float *src = in;
    for (int l = 1; l <= n->nl; l++) {
        const int cin = n->ch[l - 1], cout = n->ch[l], T = n->taps[l - 1];
        const size_t wide_lds = sizeof(float) * KW * cin * T;
        if (wide_lds <= n->lds_limit) {
            dim3 gw((unsigned)((N + kWideThreads - 1) / kWideThreads), (unsigned)((cout + KW - 1) / KW));
            if (T == 27)
                hipLaunchKernelGGL(conv_fwd_wide<27>, gw, dim3(kWideThreads), wide_lds, n->stream, src,
                                   n->params + n->woff[l - 1], n->params + n->boff[l - 1], n->act[l], cin, cout, g,
                                   n->dil[l - 1], l < n->nl ? 1 : 0);
            else
                hipLaunchKernelGGL(conv_fwd_wide<1>, gw, dim3(kWideThreads), wide_lds, n->stream, src,
                                   n->params + n->woff[l - 1], n->params + n->boff[l - 1], n->act[l], cin, cout, g, 1,
                                   l < n->nl ? 1 : 0);
            WSCHK(n, hipGetLastError());
            src = n->act[l];
            continue;
        }
        dim3 grid((unsigned)((N + kThreads - 1) / kThreads), (unsigned)((cout + KG - 1) / KG));
        const size_t lds = sizeof(float) * KG * cin * T;
        if (T == 27)
            hipLaunchKernelGGL(conv_fwd<27>, grid, dim3(kThreads), lds, n->stream, src, n->params + n->woff[l - 1],
                               n->params + n->boff[l - 1], n->act[l], cin, cout, g, n->dil[l - 1], l < n->nl ? 1 : 0);
        else
            hipLaunchKernelGGL(conv_fwd<1>, grid, dim3(kThreads), lds, n->stream, src, n->params + n->woff[l - 1],
                               n->params + n->boff[l - 1], n->act[l], cin, cout, g, 1, l < n->nl ? 1 : 0);
        WSCHK(n, hipGetLastError());
        src = n->act[l];
    }
    return MDFT_OK;
}

template <int CO>
void launch_wgrad(mdft_wsnet *n, const float *dz, const float *a, int cin, int cout, Geo g, int d, unsigned nblk) {
    const int thr = ((cin * 27 + 1 + 63) / 64) * 64;
    hipLaunchKernelGGL((conv_bwd_weight<CO>), dim3(nblk), dim3(thr), 0, n->stream, dz, a, n->part, cin, cout, g, d);
}

// Backward pass from the output gradient in n->gA, for the activations of the
// last forward pass on `d_in`: parameter gradients into n->grads, dL/d(input)
// into d_gin (when not null).
int backward(mdft_wsnet *n, const float *d_in, float *d_gin) {
    const Geo g{n->nx, n->ny, n->nz};
    const int N = g.nx * g.ny * g.nz;
    const int L = n->nl;
    // weight-gradient partials
    const unsigned nblk = (unsigned)((N + kWChunk - 1) / kWChunk);
    size_t need = 0;
    for (int l = 0; l < L; l++)
        need = std::max(need, (size_t)nblk * ((size_t)n->ch[l + 1] * n->ch[l] * n->taps[l] + n->ch[l + 1]));
    if (need > n->part_cap) {
        (void)hipFree(n->part);
        n->part = nullptr;
        WSCHK(n, hipMalloc(&n->part, sizeof(float) * need));
        n->part_cap = need;
    }
    float *gcur = n->gA, *gnext = n->gB;
    for (int l = L; l >= 1; l--) {
        const int cin = n->ch[l - 1], cout = n->ch[l], d = n->dil[l - 1];
        if (l < L) {
            hipLaunchKernelGGL(relu_bwd, dim3(grid1((size_t)N * cout)), dim3(kThreads), 0, n->stream, gcur, n->act[l],
                               (size_t)N * cout);
            WSCHK(n, hipGetLastError());
        }
        const float *a = (l == 1) ? d_in : n->act[l - 1];
        const int T = n->taps[l - 1];
        if (T == 27) {
            if (cout <= 8) launch_wgrad<8>(n, gcur, a, cin, cout, g, d, nblk);
            else if (cout <= 16) launch_wgrad<16>(n, gcur, a, cin, cout, g, d, nblk);
            else launch_wgrad<32>(n, gcur, a, cin, cout, g, d, nblk);
        } else {
            const int thr = ((cin + 1 + 63) / 64) * 64;
            hipLaunchKernelGGL(pw_bwd_weight, dim3(nblk, (unsigned)((cout + 31) / 32)), dim3(thr), 0, n->stream, gcur,
                               a, n->part, cin, cout, N);
        }
        WSCHK(n, hipGetLastError());
        const int nw = cout * cin * T + cout;  // W_l then b_l, contiguous in the parameter vector
        hipLaunchKernelGGL(sum_parts, dim3((nw + kThreads - 1) / kThreads), dim3(kThreads), 0, n->stream, n->part,
                           (int)nblk, nw, n->grads + n->woff[l - 1]);
        WSCHK(n, hipGetLastError());
        if (l > 1 || d_gin) {
            float *dst = (l == 1) ? d_gin : gnext;
            dim3 grid((unsigned)((N + kThreads - 1) / kThreads), (unsigned)((cin + KG - 1) / KG));
            const size_t lds = sizeof(float) * cout * KG * T;
            if (T == 27)
                hipLaunchKernelGGL(conv_bwd_data<27>, grid, dim3(kThreads), lds, n->stream, gcur,
                                   n->params + n->woff[l - 1], dst, cin, cout, g, d);
            else
                hipLaunchKernelGGL(conv_bwd_data<1>, grid, dim3(kThreads), lds, n->stream, gcur,
                                   n->params + n->woff[l - 1], dst, cin, cout, g, 1);
            WSCHK(n, hipGetLastError());
            std::swap(gcur, gnext);
        }
    }
    return MDFT_OK;
}

}  // namespace

extern "C" {

int mdft_wsnet_create(int nlayers, const int *channels, const int *ksize, const int *dilations,
                      unsigned long long seed, int device, mdft_wsnet **out) {
    if (!out || !channels || !ksize || !dilations || nlayers < 1) {
        g_ws_create_error = "bad arguments";
        return MDFT_EINVAL;
    }
    *out = nullptr;
    for (int l = 0; l < nlayers; l++) {
        const int cmax = ksize[l] == 3 ? kMaxC3 : kMaxC1;
        if ((ksize[l] != 1 && ksize[l] != 3) || dilations[l] < 1 || channels[l] < 1 || channels[l] > cmax ||
            channels[l + 1] < 1 || channels[l + 1] > cmax) {
            g_ws_create_error = "layers: kernel size 3 (<= 32 channels) or 1 (<= 256 channels), dilation >= 1";
            return MDFT_EINVAL;
        }
    }
    if (hipSetDevice(device) != hipSuccess) {
        g_ws_create_error = "hipSetDevice failed";
        return MDFT_ENODEV;
    }
    mdft_wsnet *n = new mdft_wsnet();
    n->device = device;
    {
        int lim = 0;
        if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lim > 0)
            n->lds_limit = (size_t)lim;
    }
    n->nl = nlayers;
    n->ch.assign(channels, channels + nlayers + 1);
    n->dil.assign(dilations, dilations + nlayers);
    for (int l = 0; l < nlayers; l++) n->taps.push_back(ksize[l] == 3 ? 27 : 1);
    size_t off = 0;
    for (int l = 0; l < nlayers; l++) {
        n->woff.push_back(off);
        off += (size_t)n->ch[l + 1] * n->ch[l] * n->taps[l];
        n->boff.push_back(off);
        off += n->ch[l + 1];
    }
    n->nparam = off;
    // He-normal initialisation, zero biases (splitmix64 + Box-Muller)
    std::vector<float> h(off, 0.0f);
    unsigned long long s = seed;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ULL;
        unsigned long long z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < nlayers; l++) {
        const double sd = sqrt(2.0 / ((double)n->taps[l] * n->ch[l]));
        for (size_t i = 0; i < (size_t)n->ch[l + 1] * n->ch[l] * n->taps[l]; i++) {
            const double u1 = next(), u2 = next();
            h[n->woff[l] + i] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
        }
    }
    bool ok = hipStreamCreate(&n->stream) == hipSuccess;
    ok = ok && hipMalloc(&n->params, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMalloc(&n->grads, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMalloc(&n->m1, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMalloc(&n->m2, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMalloc(&n->dpart, sizeof(double) * 4096) == hipSuccess;
    ok = ok && hipHostMalloc(&n->hpart, sizeof(double) * 4096, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipMemcpy(n->params, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemset(n->grads, 0, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMemset(n->m1, 0, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMemset(n->m2, 0, sizeof(float) * off) == hipSuccess;
    if (!ok) {
        mdft_wsnet_destroy(n);
        g_ws_create_error = "device allocation failed";
        return MDFT_ENOMEM;
    }
    *out = n;
    return MDFT_OK;
}

void mdft_wsnet_destroy(mdft_wsnet *n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    free_acts(n);
    (void)hipFree(n->params);
    (void)hipFree(n->grads);
    (void)hipFree(n->m1);
    (void)hipFree(n->m2);
    (void)hipFree(n->part);
    (void)hipFree(n->dpart);
    if (n->hpart) (void)hipHostFree(n->hpart);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

const char *mdft_wsnet_last_error(const mdft_wsnet *n) { return n ? n->err.c_str() : g_ws_create_error.c_str(); }

long long mdft_wsnet_num_params(const mdft_wsnet *n) { return n ? (long long)n->nparam : -1; }

int mdft_wsnet_get_params(mdft_wsnet *n, float *host) {
    if (!n || !host) return MDFT_EINVAL;
    WSCHK(n, hipMemcpy(host, n->params, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    return MDFT_OK;
}

int mdft_wsnet_set_params(mdft_wsnet *n, const float *host) {
    if (!n || !host) return MDFT_EINVAL;
    WSCHK(n, hipMemcpy(n->params, host, sizeof(float) * n->nparam, hipMemcpyHostToDevice));
    WSCHK(n, hipMemset(n->m1, 0, sizeof(float) * n->nparam));
    WSCHK(n, hipMemset(n->m2, 0, sizeof(float) * n->nparam));
    n->step = 0;
    return MDFT_OK;
}

int mdft_wsnet_get_grads(mdft_wsnet *n, float *host) {
    if (!n || !host) return MDFT_EINVAL;
    WSCHK(n, hipMemcpy(host, n->grads, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
    return MDFT_OK;
}

int mdft_wsnet_forward(mdft_wsnet *n, int nx, int ny, int nz, const float *d_in, float *d_out) {
    if (!n || !d_in || !d_out) return MDFT_EINVAL;
    int rc = ensure_grid(n, nx, ny, nz);
    if (rc) return rc;
    if ((rc = forward(n, d_in))) return rc;
    const size_t N = (size_t)nx * ny * nz;
    WSCHK(n, hipMemcpyAsync(d_out, n->act[n->nl], sizeof(float) * N * n->ch[n->nl], hipMemcpyDeviceToDevice,
                            n->stream));
    WSCHK(n, hipStreamSynchronize(n->stream));
    return MDFT_OK;
}

int mdft_wsnet_loss_grad(mdft_wsnet *n, int nx, int ny, int nz, const float *d_in, const float *d_target,
                         const float *d_mask, float *d_gin, double *loss) {
    if (!n || !d_in || !d_target || !loss) return MDFT_EINVAL;
    int rc = ensure_grid(n, nx, ny, nz);
    if (rc) return rc;
    if ((rc = forward(n, d_in))) return rc;
    const Geo g{nx, ny, nz};
    const int N = nx * ny * nz;
    const int L = n->nl;
    // loss and its gradient w.r.t. the output
    double cnt = (double)N * n->ch[L];
    if (d_mask) {
        std::vector<float> hm(N);
        WSCHK(n, hipMemcpy(hm.data(), d_mask, sizeof(float) * N, hipMemcpyDeviceToHost));
        double s = 0.0;
        for (float v : hm) s += v;
        cnt = s * n->ch[L];
        if (!(cnt > 0.0)) return ws_fail(n, MDFT_EINVAL, "empty mask");
    }
    const unsigned nb = grid1((size_t)N * n->ch[L]);
    hipLaunchKernelGGL(mse, dim3(nb), dim3(kThreads), 0, n->stream, n->act[L], d_target, d_mask, n->gA, n->ch[L], N,
                       1.0 / cnt, n->dpart);
    WSCHK(n, hipGetLastError());
    if ((rc = backward(n, d_in, d_gin))) return rc;
    WSCHK(n, hipMemcpyAsync(n->hpart, n->dpart, sizeof(double) * nb, hipMemcpyDeviceToHost, n->stream));
    WSCHK(n, hipStreamSynchronize(n->stream));
    double s = 0.0;
    for (unsigned i = 0; i < nb; i++) s += n->hpart[i];
    *loss = s / cnt;
    return MDFT_OK;
}

int mdft_wsnet_backward(mdft_wsnet *n, int nx, int ny, int nz, const float *d_in, const float *d_gout,
                        float *d_gin) {
    if (!n || !d_in || !d_gout) return MDFT_EINVAL;
    if (n->nx != nx || n->ny != ny || n->nz != nz || n->fwd_in != d_in)
        return ws_fail(n, MDFT_ESTATE, "backward must follow forward on the same input and grid");
    const size_t N = (size_t)nx * ny * nz;
    WSCHK(n, hipMemcpyAsync(n->gA, d_gout, sizeof(float) * N * n->ch[n->nl], hipMemcpyDeviceToDevice, n->stream));
    int rc = backward(n, d_in, d_gin);
    if (rc) return rc;
    WSCHK(n, hipStreamSynchronize(n->stream));
    return MDFT_OK;
}

int mdft_wsnet_adam(mdft_wsnet *n, float lr) {
    if (!n || !(lr > 0.0f)) return MDFT_EINVAL;
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 - pow(0.9, (double)n->step)), c2 = (float)(1.0 - pow(0.999, (double)n->step));
    hipLaunchKernelGGL(adam, dim3(grid1(n->nparam)), dim3(kThreads), 0, n->stream, n->params, n->grads, n->m1, n->m2,
                       n->nparam, lr, b1, b2, 1e-8f, c1, c2);
    WSCHK(n, hipGetLastError());
    WSCHK(n, hipStreamSynchronize(n->stream));
    return MDFT_OK;
}

}  // extern "C"

// k_angular_bf16.hip -- the orientation <-> projection transforms as dense
// GEMMs over grid points on the matrix cores (BF16 operands, FP32 accumulate).
//
// Same maps as k_angular.hip (src/module_orientation_projection_transform.f90:
// 159-243 as driven by src/module_energy_cproj_mrso.f90:276-290,728-762), but
// instead of the separable psi/phi DFTs and the theta contraction each map is
// one real matrix:
//   angl2proj:  P[r][c] = sum_io o[r][io] M[io][c],   c = 2 ip + (re, im)
//   proj2angl:  v[r][io] = sum_c P[r][c] Q[c][io]
// built on the host from the same quadrature, phase and Wigner tables.  A wave
// takes 16 grid points at a time and runs v_mfma_f32_16x16x32_bf16 over the
// K dimension for every 16-column tile of the output; the constant matrix sits
// in LDS already in the per-lane fragment order (16 contiguous bytes per lane
// per MFMA, conflict-free).  The free-energy reductions, rho ln rho and dF stay
// FP64 (the final pass reads xi and vext in FP64 as in the other paths).
//
// Fragment layouts of v_mfma_f32_16x16x32_bf16 (64 lanes): A (16 x 32): lane l
// holds row l % 16, k = 8 (l / 16) + j, j = 0..7; B (32 x 16): lane l holds
// column l % 16, the same k; D (16 x 16): lane l holds column l % 16, rows
// 4 (l / 16) + i, i = 0..3.
#include <cfloat>
#include <vector>

#include "mdft_internal.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kWaves = 8;
constexpr int kThreads = 64 * kWaves;

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }

template <int L>
struct BDims {
    static constexpr int NO = (L + 1) * (2 * L + 1) * (2 * (L / 2) + 1);
    static constexpr int np_of(int m) { return m < 0 ? 0 : np_of(m - 1) + (2 * m + 1) * (m / 2 + 1); }
    static constexpr int NP2 = 2 * np_of(L);
    // angl2proj GEMM: K = NO (padded to 32), N = NP2 (padded to 16)
    static constexpr int A_KS = (NO + 31) / 32, A_NT = (NP2 + 15) / 16;
    // proj2angl GEMM: K = NP2 (padded to 32), N = NO (padded to 16)
    static constexpr int F_KS = (NP2 + 31) / 32, F_NT = (NO + 15) / 16;
};

template <int L>
__global__ __launch_bounds__(kThreads) void a2p_bf16(const double *__restrict__ xi, const bf16x8 *__restrict__ Mf,
                                                    float *__restrict__ proj, size_t N, double rho0) {
    using D = BDims<L>;
    constexpr int NO = D::NO, NP2 = D::NP2, KS = D::A_KS, NT = D::A_NT;
    extern __shared__ __align__(16) unsigned char smem[];
    bf16x8 *Ms = reinterpret_cast<bf16x8 *>(smem);  // [KS][NT][64]
    for (int i = threadIdx.x; i < KS * NT * 64; i += kThreads) Ms[i] = Mf[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const size_t ntiles = (N + 15) / 16;
    for (size_t tile = (size_t)blockIdx.x * kWaves + w; tile < ntiles; tile += (size_t)gridDim.x * kWaves) {
        const size_t p = tile * 16 + row;
        const bool live = p < N;
        const double *src = xi + p * NO;
        floatx4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            bf16x8 a;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = ks * 32 + kq * 8 + j;
                float o = 0.0f;
                if (live && k < NO) {
                    const double x = src[k];
                    o = (float)(rho0 * (x * x - 1.0));
                }
                a[j] = to_bf16(o);
            }
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, Ms[(ks * NT + t) * 64 + lane], acc[t], 0, 0, 0);
        }
        // D: column t*16 + lane%16, rows 4*(lane/16) + i
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int c = t * 16 + row;
            if (c < NP2) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const size_t pr = tile * 16 + kq * 4 + i;
                    if (pr < N) proj[pr * NP2 + c] = acc[t][i];
                }
            }
        }
    }
}

struct FinalConstsB {
    double rho0, kT, dv, prefactor;
};

__device__ inline void block_sum3(double a, double b, double c, double *partials) {
    __shared__ double red[3][kWaves];
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
        for (int k = 0; k < kWaves; k++) {
            s0 += red[0][k];
            s1 += red[1][k];
            s2 += red[2][k];
        }
        partials[blockIdx.x * 3 + 0] = s0;
        partials[blockIdx.x * 3 + 1] = s1;
        partials[blockIdx.x * 3 + 2] = s2;
    }
}

// proj2angl GEMM + the fused FP64 final pass (modes as final_t in k_angular.hip)
template <int L>
__global__ __launch_bounds__(kThreads) void final_bf16(const float *__restrict__ proj, const bf16x8 *__restrict__ Qf,
                                                      const double *__restrict__ xi, const double *__restrict__ vext,
                                                      const double *__restrict__ wq, double *__restrict__ df,
                                                      double *__restrict__ partials, size_t N, FinalConstsB k,
                                                      int mode) {
    using D = BDims<L>;
    constexpr int NO = D::NO, NP2 = D::NP2, KS = D::F_KS, NT = D::F_NT;
    extern __shared__ __align__(16) unsigned char smem[];
    bf16x8 *Qs = reinterpret_cast<bf16x8 *>(smem);  // [KS][NT][64]
    for (int i = threadIdx.x; i < KS * NT * 64; i += kThreads) Qs[i] = Qf[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const double rho0 = k.rho0, kT = k.kT, dv = k.dv;
    double fid = 0.0, fext = 0.0, fexc = 0.0;
    const size_t ntiles = (N + 15) / 16;
    for (size_t tile = (size_t)blockIdx.x * kWaves + w; tile < ntiles; tile += (size_t)gridDim.x * kWaves) {
        const size_t p = tile * 16 + row;
        const bool live = p < N;
        const float *src = proj + p * NP2;
        floatx4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            bf16x8 a;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int c = ks * 32 + kq * 8 + j;
                a[j] = to_bf16((live && c < NP2) ? src[c] : 0.0f);
            }
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, Qs[(ks * NT + t) * 64 + lane], acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int io = t * 16 + row;
            if (io >= NO) continue;
            const double wio = wq[io];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const size_t pr = tile * 16 + kq * 4 + i;
                if (pr >= N) continue;
                const double v = (double)acc[t][i];
                const double vexc = k.prefactor * wio * v;
                const size_t g = pr * NO + io;
                const double x = (mode >= 3) ? 0.0 : xi[g];
                const double rho = x * x * rho0;
                fexc += (rho - rho0) * vexc;
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
                } else if (mode == 2) {
                    if (df) df[g] = -(k.prefactor / kT) * v;
                } else if (mode == 3) {
                    df[g] = exp(0.5 * fmin(fmax(-vext[g] / kT + v, -200.0), 7.0));
                } else {
                    df[g] = (vext[g] / kT >= -log(DBL_EPSILON) - 0.05) ? 0.0 : sqrt(fmax(1.0 + v, MDFT_WS_DENSITY_FLOOR));
                }
            }
        }
    }
    block_sum3(fid, fext, fexc, partials);
}

template <int L>
size_t a2p_lds() { return sizeof(bf16x8) * BDims<L>::A_KS * BDims<L>::A_NT * 64; }
template <int L>
size_t fin_lds() { return sizeof(bf16x8) * BDims<L>::F_KS * BDims<L>::F_NT * 64; }

// Rearranges a dense K x Nc matrix (row-major, FP64) into the B-fragment order
// [KS][NT][64 lanes][8] of v_mfma_f32_16x16x32_bf16, zero-padded.
std::vector<__bf16> to_fragments(const std::vector<double> &m, int K, int Nc) {
    const int KS = (K + 31) / 32, NT = (Nc + 15) / 16;
    std::vector<__bf16> out((size_t)KS * NT * 64 * 8);
    for (int ks = 0; ks < KS; ks++)
        for (int t = 0; t < NT; t++)
            for (int lane = 0; lane < 64; lane++)
                for (int j = 0; j < 8; j++) {
                    const int kk = ks * 32 + (lane >> 4) * 8 + j, n = t * 16 + (lane & 15);
                    const double v = (kk < K && n < Nc) ? m[(size_t)kk * Nc + n] : 0.0;
                    out[(((size_t)ks * NT + t) * 64 + lane) * 8 + j] = (__bf16)(float)v;
                }
    return out;
}

int blocks_for(mdft_ctx *c) { return c->num_cus > 0 ? c->num_cus : 256; }

}  // namespace

// Host matrices of the two maps from the context's tables (same arithmetic as
// the separable kernels: phases exp(+i 2 pi (mu2 ipsi / npsi + mup iphi / nphi)),
// Gauss-Legendre-weighted Wigner d and sqrt(2m+1) for angl2proj; the
// Hermitian-folded real part for proj2angl).
int bf16_tables(mdft_ctx *c, const std::vector<double> &a2p, const std::vector<double> &p2a,
                const std::vector<double> &fm) {
    const int no = c->no, np = c->np, npsi = c->npsi, nphi = c->nphi, nth = c->ntheta, H1 = c->H + 1;
    const double pi = acos(-1.0);
    std::vector<double> M((size_t)no * 2 * np), Q((size_t)2 * np * no);
    for (int it = 0; it < nth; it++)
        for (int iphi = 0; iphi < nphi; iphi++)
            for (int ipsi = 0; ipsi < npsi; ipsi++) {
                const int io = ipsi + npsi * (iphi + nphi * it);
                for (int ip = 0; ip < np; ip++) {
                    const int mup = c->pmup[ip], mu2 = c->pmu[ip] / c->d.mrso;
                    const int kpsi = (mu2 * ipsi) % npsi, kphi = ((mup * iphi) % nphi + nphi) % nphi;
                    const double ang = 2.0 * pi * ((double)kpsi / npsi + (double)kphi / nphi);
                    const double am = fm[ip] * a2p[(size_t)ip * nth + it] / (double)(nphi * npsi);
                    M[(size_t)io * 2 * np + 2 * ip] = am * cos(ang);
                    M[(size_t)io * 2 * np + 2 * ip + 1] = am * sin(ang);
                    const double wmu = (mu2 == 0) ? 1.0 : 2.0;
                    const double qm = wmu * p2a[(size_t)ip * nth + it];
                    Q[(size_t)(2 * ip) * no + io] = qm * cos(ang);
                    Q[(size_t)(2 * ip + 1) * no + io] = qm * sin(ang);
                }
            }
    (void)H1;
    std::vector<__bf16> Mf = to_fragments(M, no, 2 * np), Qf = to_fragments(Q, 2 * np, no);
    if (hipMalloc(&c->d_bf16_M, Mf.size() * sizeof(__bf16)) != hipSuccess ||
        hipMalloc(&c->d_bf16_Q, Qf.size() * sizeof(__bf16)) != hipSuccess)
        return MDFT_ENOMEM;
    if (hipMemcpy(c->d_bf16_M, Mf.data(), Mf.size() * sizeof(__bf16), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_bf16_Q, Qf.data(), Qf.size() * sizeof(__bf16), hipMemcpyHostToDevice) != hipSuccess)
        return MDFT_EHIP;
    return MDFT_OK;
}

hipError_t launch_angl2proj_bf16(mdft_ctx *c, const double *xi, float2 *proj) {
    const dim3 grid((unsigned)blocks_for(c));
    const bf16x8 *M = static_cast<const bf16x8 *>(c->d_bf16_M);
    float *P = reinterpret_cast<float *>(proj);
    switch (c->L) {
#define A_CASE(LL) \
    case LL: hipLaunchKernelGGL(a2p_bf16<LL>, grid, dim3(kThreads), a2p_lds<LL>(), c->stream, xi, M, P, c->N, c->rho0); break;
        A_CASE(0) A_CASE(1) A_CASE(2) A_CASE(3) A_CASE(4) A_CASE(5)
#undef A_CASE
        default: return hipErrorNotSupported;
    }
    return hipGetLastError();
}

hipError_t launch_final_bf16(mdft_ctx *c, const float2 *proj, const double *xi, double *df, int mode,
                             int *nblocks_out) {
    const int nb = blocks_for(c);
    if (c->partials_cap < (size_t)nb * 3) {
        (void)hipFree(c->d_partials);
        c->d_partials = nullptr;
        c->partials_cap = 0;
        hipError_t e = hipMalloc(&c->d_partials, sizeof(double) * 3 * nb);
        if (e != hipSuccess) return e;
        c->partials_cap = 3 * (size_t)nb;
    }
    FinalConstsB k;
    k.rho0 = c->rho0;
    k.kT = c->d.kT;
    k.dv = c->dv;
    k.prefactor = -c->d.kT * 2.0 * (4.0 * acos(-1.0) * acos(-1.0)) / c->d.mrso / c->d.n0;  // cproj_mrso.f90:728
    const bf16x8 *Q = static_cast<const bf16x8 *>(c->d_bf16_Q);
    const float *P = reinterpret_cast<const float *>(proj);
    switch (c->L) {
#define F_CASE(LL) \
    case LL: hipLaunchKernelGGL(final_bf16<LL>, dim3((unsigned)nb), dim3(kThreads), fin_lds<LL>(), c->stream, P, Q, xi, c->d_vext, c->d_w, df, c->d_partials, c->N, k, mode); break;
        F_CASE(0) F_CASE(1) F_CASE(2) F_CASE(3) F_CASE(4) F_CASE(5)
#undef F_CASE
        default: return hipErrorNotSupported;
    }
    *nblocks_out = nb;
    return hipGetLastError();
}

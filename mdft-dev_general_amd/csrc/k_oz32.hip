// k_oz32.hip -- the q-space step of the excess term in FP32, factored form.
//
// Same contraction as k_oz.hip (src/module_energy_cproj_mrso.f90:441-674 for
// each (q, -q) pair), evaluated with the factorisation
//     R^m_{a,b}(q) = exp(-i a phi_q) d^m_{a,b}(theta_q)
// of the reference's rotation matrices (src/module_rotation.f90:119-247 builds
// the same matrices from the frame of q by recursion in m).  Rotation to the q
// frame is then a phase multiply followed by a REAL matrix product, and the
// rotation back a real product followed by the conjugate phase: half the
// multiplies of the complex form and no sequential recursion.  d^m(theta) is
// built for every pair, all entries in parallel, from its Fourier series
//     d^m_{a,b}(theta) = i^{a-b} sum_k D^m_{k,a} D^m_{k,b} exp(-i k theta),
// D^m = d^m(pi/2), folded on the host into real coefficients of cos(k theta)
// (a-b even) or sin(k theta) (a-b odd), k = 0..m.
//
// Differences from the FP64 kernel, all below the FP32 path's 1e-5 tolerance:
// the FLT_EPSILON / DBL_EPSILON chops of :491-492 and :618-623 (single
// precision rounding of one component, zeroing below 1.2e-7 resp. 2.2e-16)
// are not applied.  Nyquist k-vectors are +pi/dl as in the reference
// (src/module_grid.f90:319-331), the self-conjugate Nyquist store rule
// (:648-656) and the pair roles are those of the FP64 kernel.
//
// Work layout: one workgroup per chunk of pairs sharing the interpolation
// interval iq (c rows staged once in LDS as FP32); one wave per pair, with
// per-wave LDS buffers and wave-local synchronisation only.  In the OZ sums
// every lane owns fixed projection rows (balanced on the host), so each c
// coefficient is read once per pair and shared by the q and -q sums.
#include <algorithm>
#include <vector>

#include "mdft_internal.h"

namespace {

constexpr int kRowsPerLane = MDFT_OZ32_ROWS_PER_LANE;
constexpr int kMaxL = MDFT_MAX_MMAX;

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int roff(int m) { return m * (4 * m * m - 1) / 3; }  // sum_{l<m} (2l+1)^2

__device__ inline float2 cmulf(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

struct Oz32Args {
    int nx, ny, np, npc, L, mrso;
    double dq;
    const double *kx, *ky, *kz;
    const int *pm, *pmup, *pmu, *ia0;
    const int *pq, *pmq, *pflag, *cstart, *ccount, *ciq;
    const double2 *ck;
    const float *ftab;   // Fourier coefficients of d^m(theta)
    const int *itab;     // descriptor table (see oz32_itab_size)
};

// Integer descriptor table (host-built by oz32_tables, staged in LDS), in order:
//   rin  [2 np][4]  rotation to the q frame, output o (o >= np: the -q side):
//                   {Y offset, step, d offset, W | side << 8 | (m odd) << 9}
//   rout [2 np][4]  rotation back, output o: {G offset, step, d row, W | side << 8 |
//                   (m odd) << 9 | |mup| << 10 | (mup < 0) << 14}
//   lrow [kRowsPerLane][64]  OZ rows of each lane (-1: none)
//   dd   [nR]       entry e of the d^m(theta) block: Fourier-table offset << 4 | m << 1 | (a - b odd)
//   ncol [kNKhi]    OZ columns of each khi;  ctab [kNKhi][kCMax]: A/B offset << 1 | (nu < 0)
constexpr int kNKhi = 2 * kMaxL + 1;
constexpr int kCMax = 32;

__host__ __device__ inline int oz32_itab_size(int np, int nR) {
    return 16 * np + kRowsPerLane * 64 + nR + kNKhi + kNKhi * kCMax;
}

template <int kWaves>
__global__ __launch_bounds__(64 * kWaves) void oz32(float2 *__restrict__ proj, Oz32Args a) {
    constexpr int kThreads = 64 * kWaves;
    extern __shared__ __align__(16) unsigned char sm[];
    const int np = a.np, npc = a.npc, L = a.L;
    const int nR = roff(L + 1);
    int nTab = 0;
    for (int m = 0; m <= L; m++) nTab += (2 * m + 1) * (2 * m + 1) * (m + 1);
    const int nF = (nTab + kWaves * nR + 3) & ~3;  // floats, padded to 16 bytes
    const int nI = oz32_itab_size(np, nR);
    // ---- LDS carve-up: float2, then floats, then the integer table
    float2 *c0 = reinterpret_cast<float2 *>(sm), *c1 = c0 + npc;
    float2 *wv2 = c1 + npc;  // per wave: Y [2 np] (q, -q), A [2 np], G [2 np], phase [kMaxL + 1]
    const int per2 = (6 * np + kMaxL + 2) & ~1;  // even: every region below stays 16-byte aligned
    float *ftab = reinterpret_cast<float *>(wv2 + kWaves * per2);
    float *wvd = ftab + nTab;  // per wave: d [nR]
    int *itab = reinterpret_cast<int *>(ftab + nF);
    int *tmup = itab + nI, *tia0 = tmup + np;
    const int4 *rin = reinterpret_cast<const int4 *>(itab), *rout = rin + 2 * np;
    const int *lrow = itab + 16 * np, *dd = lrow + kRowsPerLane * 64, *ncol = dd + nR, *ctab = ncol + kNKhi;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float2 *Y = wv2 + w * per2, *A = Y + 2 * np, *G = A + 2 * np, *ph = G + 2 * np;
    float *d = wvd + w * nR;

    const int chunk = blockIdx.x;
    const int iq = a.ciq[chunk];
    for (int i = threadIdx.x; i < npc; i += kThreads) {
        const double2 u = a.ck[(size_t)(iq - 1) * npc + i], v = a.ck[(size_t)iq * npc + i];
        c0[i] = make_float2((float)u.x, (float)u.y);
        c1[i] = make_float2((float)v.x, (float)v.y);
    }
    for (int i = threadIdx.x; i < nTab; i += kThreads) ftab[i] = a.ftab[i];
    for (int i = threadIdx.x; i < nI; i += kThreads) itab[i] = a.itab[i];
    for (int i = threadIdx.x; i < np; i += kThreads) {
        tmup[i] = a.pmup[i];
        tia0[i] = a.ia0[i];
    }
    __syncthreads();

    const int start = a.cstart[chunk], count = a.ccount[chunk];
    for (int k = w; k < count; k += kWaves) {
        const int pr = start + k;
        const int gq = a.pq[pr], gmq = a.pmq[pr], flag = a.pflag[pr];
        const bool eq = flag & 1, nyq = flag & 2;
        const int ix = gq % a.nx, iy = (gq / a.nx) % a.ny, iz = gq / (a.nx * a.ny);
        const double q0 = a.kx[ix], q1 = a.ky[iy], q2 = a.kz[iz];
        const double qn = sqrt(q0 * q0 + q1 * q1 + q2 * q2), rxy = sqrt(q0 * q0 + q1 * q1);
        const float alpha = (float)(qn / a.dq + 1.0 - (double)iq);
        // frame angles: theta from the z axis, phi in the xy plane; q along -z
        // is the Euler rotation (pi, pi, 0) of the reference's frame (0, pi, pi)
        const float ct = qn > 0.0 ? (float)(q2 / qn) : 1.0f, st = qn > 0.0 ? (float)(rxy / qn) : 0.0f;
        const float cp = rxy > 0.0 ? (float)(q0 / rxy) : (q2 < 0.0 ? -1.0f : 1.0f);
        const float sp = rxy > 0.0 ? (float)(q1 / rxy) : 0.0f;
        float ckt[kMaxL + 1], skt[kMaxL + 1];
        ckt[0] = 1.0f;
        skt[0] = 0.0f;
        float cpk = 1.0f, spk = 0.0f;
#pragma unroll
        for (int j = 1; j <= kMaxL; j++) {
            ckt[j] = ckt[j - 1] * ct - skt[j - 1] * st;
            skt[j] = skt[j - 1] * ct + ckt[j - 1] * st;
        }
#pragma unroll
        for (int j = 0; j <= kMaxL; j++) {
            if (lane == j) ph[j] = make_float2(cpk, spk);  // exp(+i j phi)
            const float nc = cpk * cp - spk * sp;
            spk = spk * cp + cpk * sp;
            cpk = nc;
        }
        // d^m_{a,b}(theta), every entry from its Fourier coefficients
        for (int e = lane; e < nR; e += 64) {
            const int ds = dd[e];
            const float *T = ftab + (ds >> 4);
            const int m = (ds >> 1) & 7;
            const bool odd = ds & 1;
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j <= kMaxL; j++)
                if (j <= m) s += T[j] * (odd ? skt[j] : ckt[j]);
            d[e] = s;
        }
        wave_sync();
        // phase multiply: Y = X exp(-i mup phi), q side then -q side
        const float2 *src0 = proj + (size_t)gq * np, *src1 = proj + (size_t)gmq * np;
        for (int i = lane; i < np; i += 64) {
            const int mup = tmup[i];
            const float2 e = ph[mup < 0 ? -mup : mup];
            const float2 em = make_float2(e.x, mup < 0 ? e.y : -e.y);
            Y[i] = cmulf(src0[i], em);
            Y[np + i] = cmulf(src1[i], em);
        }
        wave_sync();
        // rotation to the q frame: A(m,khi,mu2) = sum_mup Y_q(m,mup,mu2) d(mup,khi),
        // B(m,khi,mu2) = (-1)^m sum_mup Y_-q(m,mup,mu2) d(mup,-khi)   (B = A + np)
        for (int o = lane; o < 2 * np; o += 64) {
            const int4 r = rin[o];
            const int W = r.w & 255;
            const float2 *y = Y + r.x;
            const float *dc = d + r.z;
            float sx = 0.0f, sy = 0.0f;
#pragma unroll 4
            for (int t = 0; t < W; t++) {
                const float dv = dc[t * W];
                const float2 v = y[t * r.y];
                sx += dv * v.x;
                sy += dv * v.y;
            }
            const float sg = ((r.w >> 8) & (r.w >> 9) & 1) ? -1.0f : 1.0f;
            A[o] = make_float2(sg * sx, sg * sy);
        }
        wave_sync();
        // molecular OZ (:561-606): nu < 0 from the same side, nu >= 0 conjugated
        // from the other side; both sides share each interpolated coefficient
#pragma unroll
        for (int rr = 0; rr < kRowsPerLane; rr++) {
            const int ip = lrow[rr * 64 + lane];
            if (ip < 0) continue;
            const int kk = tmup[ip] + kMaxL;
            const int nc = ncol[kk];
            const int *ctk = ctab + kk * kCMax;
            const float2 *lo = c0 + tia0[ip], *hi = c1 + tia0[ip];
            float qx = 0.0f, qy = 0.0f, mx = 0.0f, my = 0.0f;
#pragma unroll 4
            for (int c = 0; c < nc; c++) {
                const int ent = ctk[c];
                const int off = ent >> 1;
                const float2 l = lo[c], h = hi[c];
                const float cx = l.x + alpha * (h.x - l.x), cy = l.y + alpha * (h.y - l.y);
                const float2 aa = A[off], bb = A[np + off];
                float2 u, v;  // sources of the q and -q sums
                if (ent & 1) {
                    u = aa;
                    v = bb;
                } else {
                    u = make_float2(bb.x, -bb.y);
                    v = make_float2(aa.x, -aa.y);
                }
                // four multiply-adds per complex product (the plain expression compiles to six)
                qx = fmaf(cx, u.x, qx);
                qx = fmaf(-cy, u.y, qx);
                qy = fmaf(cx, u.y, qy);
                qy = fmaf(cy, u.x, qy);
                mx = fmaf(cx, v.x, mx);
                mx = fmaf(-cy, v.y, mx);
                my = fmaf(cx, v.y, my);
                my = fmaf(cy, v.x, my);
            }
            G[ip] = make_float2(qx, qy);
            G[np + ip] = make_float2(mx, my);
        }
        wave_sync();
        // rotation back (:611-641) and store (:648-656)
        for (int o = lane; o < 2 * np; o += 64) {
            const int4 r = rout[o];
            const bool second = (r.w >> 8) & 1;
            if (!second && eq) continue;  // q == -q: the second store wins
            const int W = r.w & 255;
            const float2 *g = G + r.x;
            const float *dr = d + r.z;
            float sx = 0.0f, sy = 0.0f;
#pragma unroll 4
            for (int t = 0; t < W; t++) {
                const float dv = dr[second ? (W - 1 - t) : t];  // d(mup, -khi) for the -q side
                const float2 v = g[t * r.y];
                sx += dv * v.x;
                sy += dv * v.y;
            }
            const float2 e = ph[(r.w >> 10) & 15];
            const float2 ep = make_float2(e.x, ((r.w >> 14) & 1) ? -e.y : e.y);  // exp(+i mup phi)
            float2 out = cmulf(make_float2(sx, sy), ep);
            if (second) {
                if ((r.w >> 9) & 1) {
                    out.x = -out.x;
                    out.y = -out.y;
                }
                if (eq && nyq) out.y = -out.y;
                proj[(size_t)gmq * np + (o - np)] = out;
            } else {
                proj[(size_t)gq * np + o] = out;
            }
        }
        wave_sync();
    }
}

double textbook_d(int j, int mp, int m, double beta) {
    auto f = [](int n) {
        double r = 1.0;
        for (int i = 2; i <= n; i++) r *= i;
        return r;
    };
    double s = 0.0;
    for (int k = 0; k <= 2 * j; k++) {
        const int a = j + m - k, b = k, c = mp - m + k, e = j - mp - k;
        if (a < 0 || b < 0 || c < 0 || e < 0) continue;
        s += (((mp - m + k) & 1) ? -1.0 : 1.0) * pow(cos(beta / 2), 2 * j + m - mp - 2 * k) *
             pow(sin(beta / 2), mp - m + 2 * k) / (f(a) * f(b) * f(c) * f(e));
    }
    return sqrt(f(j + mp) * f(j - mp) * f(j + m) * f(j - m)) * s;
}

}  // namespace

// Host tables of the FP32 q-space kernel: Fourier coefficients of d^m(theta)
// and the integer descriptor table (layout at oz32_itab_size): per-output
// offsets of the two rotations, the balanced assignment of OZ rows to lanes,
// the d-block entries and the OZ columns of each khi.
int oz32_tables(mdft_ctx *c, std::vector<float> &ftab, std::vector<int> &itab) {
    const int L = c->L;
    const double hp = 0.5 * acos(-1.0);
    ftab.clear();
    for (int m = 0; m <= L; m++) {
        const int W = 2 * m + 1;
        std::vector<double> D(W * W);
        for (int k = -m; k <= m; k++)
            for (int a = -m; a <= m; a++) D[(k + m) * W + (a + m)] = textbook_d(m, k, a, hp);
        for (int a = -m; a <= m; a++)
            for (int b = -m; b <= m; b++) {
                const int ab = ((a - b) % 4 + 4) % 4;  // i^{a-b} = 1, i, -1, -i
                for (int k = 0; k <= m; k++) {
                    const double sk = D[(k + m) * W + (a + m)] * D[(k + m) * W + (b + m)];
                    const double smk = D[(-k + m) * W + (a + m)] * D[(-k + m) * W + (b + m)];
                    double v;
                    if (k == 0) v = (ab == 0) ? sk : (ab == 2 ? -sk : 0.0);
                    // even a-b: i^{a-b} (S_k + S_-k) cos; odd: Re(i^{a-b} (-i)) (S_k - S_-k) sin
                    else if (ab == 0) v = sk + smk;
                    else if (ab == 2) v = -(sk + smk);
                    else if (ab == 1) v = sk - smk;
                    else v = -(sk - smk);
                    ftab.push_back((float)v);
                }
            }
    }
    // longest-processing-time assignment of the np OZ rows to 64 lanes
    std::vector<int> len(c->np), order(c->np);
    for (int ip = 0; ip < c->np; ip++) {
        const int khi = std::abs(c->pmup[ip]);
        int n_ = 0;
        for (int n = khi; n <= L; n++) n_ += 2 * (n / c->d.mrso) + 1;
        len[ip] = n_;
        order[ip] = ip;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return len[x] > len[y]; });
    std::vector<int> lanerow(MDFT_OZ32_ROWS_PER_LANE * 64, -1);
    std::vector<int> load(64, 0), cnt(64, 0);
    for (int ip : order) {
        int best = -1;
        for (int l = 0; l < 64; l++)
            if (cnt[l] < MDFT_OZ32_ROWS_PER_LANE && (best < 0 || load[l] < load[best])) best = l;
        if (best < 0) return MDFT_EINVAL;
        lanerow[cnt[best] * 64 + best] = ip;
        cnt[best]++;
        load[best] += len[ip];
    }
    const int np = c->np, mrso = c->d.mrso;
    if (L > kMaxL) return MDFT_EINVAL;
    std::vector<int> poff(L + 2, 0), foff(L + 2, 0);
    for (int m = 0; m <= L; m++) {
        poff[m + 1] = poff[m] + (2 * m + 1) * (m / mrso + 1);
        foff[m + 1] = foff[m] + (2 * m + 1) * (2 * m + 1) * (m + 1);
    }
    auto pidx = [&](int m, int mup, int mu2) { return poff[m] + (mup + m) * (m / mrso + 1) + mu2; };
    auto roffh = [](int m) { return m * (4 * m * m - 1) / 3; };
    const int nR = roffh(L + 1);
    itab.assign(oz32_itab_size(np, nR), 0);
    int *rin = itab.data(), *rout = rin + 8 * np, *lrow = itab.data() + 16 * np, *dd = lrow + MDFT_OZ32_ROWS_PER_LANE * 64;
    int *ncol = dd + nR, *ctab = ncol + kNKhi;
    for (int o = 0; o < 2 * np; o++) {
        const int second = o >= np ? 1 : 0, ip = o - second * np;
        const int m = c->pm[ip], mup = c->pmup[ip], mu2 = c->pmu[ip] / mrso, W = 2 * m + 1;
        const int base = second * np + pidx(m, -m, mu2), step = m / mrso + 1;
        rin[4 * o] = base;
        rin[4 * o + 1] = step;
        rin[4 * o + 2] = roffh(m) + (second ? -mup : mup) + m;  // mup plays khi here
        rin[4 * o + 3] = W | second << 8 | (m & 1) << 9;
        rout[4 * o] = base;
        rout[4 * o + 1] = step;
        rout[4 * o + 2] = roffh(m) + (mup + m) * W;
        rout[4 * o + 3] = W | second << 8 | (m & 1) << 9 | std::abs(mup) << 10 | (mup < 0 ? 1 : 0) << 14;
    }
    std::copy(lanerow.begin(), lanerow.end(), lrow);
    for (int e = 0, m = 0; e < nR; e++) {
        while (m < L && roffh(m + 1) <= e) m++;
        const int rem = e - roffh(m), W = 2 * m + 1, ab = rem / W - rem % W;
        dd[e] = (foff[m] + rem * (m + 1)) << 4 | m << 1 | (ab & 1);
    }
    for (int khi = -L; khi <= L; khi++) {
        int n_ = 0;
        for (int n = std::abs(khi); n <= L; n++) {
            const int h = n / mrso;
            for (int nu2 = -h; nu2 <= h; nu2++) {
                if (n_ >= kCMax) return MDFT_EINVAL;
                ctab[(khi + kMaxL) * kCMax + n_++] = (pidx(n, khi, 0) + std::abs(nu2)) << 1 | (nu2 < 0 ? 1 : 0);
            }
        }
        ncol[khi + kMaxL] = n_;
    }
    return MDFT_OK;
}

hipError_t launch_oz32(mdft_ctx *c, float2 *proj) {
    Oz32Args a;
    a.nx = c->d.nx;
    a.ny = c->d.ny;
    a.np = c->np;
    a.npc = c->npc;
    a.L = c->L;
    a.mrso = c->d.mrso;
    a.dq = c->dq;
    a.kx = c->d_kx;
    a.ky = c->d_ky;
    a.kz = c->d_kz;
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
    a.ftab = c->d_oz32_ftab;
    a.itab = c->d_oz32_lanerow;
    int nTab = 0;
    for (int m = 0; m <= c->L; m++) nTab += (2 * m + 1) * (2 * m + 1) * (m + 1);
    const int nR = (c->L + 1) * (4 * (c->L + 1) * (c->L + 1) - 1) / 3;
    auto lds_for = [&](int waves) {
        const size_t nF = ((size_t)nTab + (size_t)waves * nR + 3) & ~(size_t)3;
        return sizeof(float2) * (2 * (size_t)c->npc + (size_t)waves * ((6 * c->np + MDFT_MAX_MMAX + 2) & ~1)) +
               sizeof(float) * nF + sizeof(int) * ((size_t)oz32_itab_size(c->np, nR) + 2 * (size_t)c->np);
    };
    const dim3 grid((unsigned)c->pairs.nchunks);
    if (lds_for(8) <= c->lds_limit)
        hipLaunchKernelGGL((oz32<8>), grid, dim3(512), lds_for(8), c->stream, proj, a);
    else if (lds_for(4) <= c->lds_limit)
        hipLaunchKernelGGL((oz32<4>), grid, dim3(256), lds_for(4), c->stream, proj, a);
    else if (lds_for(1) <= c->lds_limit)
        hipLaunchKernelGGL((oz32<1>), grid, dim3(64), lds_for(1), c->stream, proj, a);
    else
        return hipErrorInvalidConfiguration;
    return hipGetLastError();
}

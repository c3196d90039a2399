// k_ozm.hip -- the q-space step of the excess term (rotation to the q frame,
// molecular Ornstein-Zernike contraction with c(|q|), rotation back) with the
// contraction on the matrix cores, FP64 (v_mfma_f64_16x16x4) and FP32
// (v_mfma_f32_16x16x4), for molecular symmetry order 2 and mmax <= 6.
//
// Restates the q-loop of src/module_energy_cproj_mrso.f90:441-674 for each
// (q, -q) pair: R^m_{mup,khi}(q) chopped at FLT_EPSILON (:491-492), rotation
// (:505-523), c interpolated linearly in |q| (:534-539), the OZ sums
// (:561-606), the DBL_EPSILON chop of gamma (:618-623), rotation back with
// conj(R) (:611-641), the -q side with R(., -khi) and (-1)^m (:517-520,
// 636-638) and the store rule of self-conjugate Nyquist points (:648-656).
// The chops apply to the FP64 kernel; the FP32 kernel skips them (their effect,
// single-precision rounding of one component, is below FP32's own rounding).
//
// Work layout.  A workgroup of four waves takes a tile of 8 pairs that share the
// interpolation interval iq (pairs are sorted by iq on the host, one host-built
// record per pair: grid indices of q and -q, flags, the angles of q and alpha);
// the tile's 16 (side, pair) rows are the M dimension of the matrix products.
// Workgroup b takes tile (b & 7) * per + (b >> 3), so the tiles of one
// accelerator die (workgroups go to the 8 dies round robin) are contiguous in
// |q|, and their c(k) fragments share that die's L2.  Per tile:
//  0. the 16 projection records X (np complex each) are copied into LDS with
//     whole-record coalesced loads;
//  1. rotation to the q frame on the vector ALUs, lanes over pairs: lane
//     (pair p = l & 7, slot j = l >> 3) takes rotation units (m, khi) from a
//     host-built round table (units of equal m/2 + 1 in a round, so the lanes
//     run one instruction stream; rounds balanced over the four waves), forms
//     R^m(mup, khi) = exp(-i mup phi_q) d^m(mup, khi)(theta_q) once per unit
//     entry (d^m from its Fourier series in theta, coefficients host-built and
//     read one iteration ahead) and uses it for both sides and every mu:
//       A_q(m, khi, mu) = sum_mup X_q(m, mup, mu) R(mup, khi)
//       A_-q(m, -khi, mu) = (-1)^m sum_mup X_-q(m, mup, mu) R(mup, khi);
//     in FP32 the phase e^{-i mup phi} is applied once per element of X as it
//     enters LDS and once per output, so the unit loops multiply by real d^m;
//  2. OZ contraction (OZ blocks khi spread over the four waves): for each khi
//     the rows (m, mu) of both sides are
//       gamma_s(m, khi, mu) = sum_{(n, nu)} c_{m mu; n nu}^{khi}(|q|) V_s(n, nu)
//     with V_s(n, nu) = A_s(n, khi, |nu|) for nu < 0 and conj(A_-s(n, khi, nu))
//     otherwise.  All pairs of the tile share c[iq] and c[iq+1], so with
//     c = (1 - alpha) c[iq] + alpha c[iq+1] the sums are two matrix products,
//     [16 rows x 2 ncol] (re/im of V) times a [2 ncol x 4 nrow] real form of
//     (c[iq], c[iq+1]) whose fragments are precomputed per interval in the
//     order the matrix instruction reads them (ozm_frags, at c(k) upload);
//     alpha and 1 - alpha are applied per row after the product, the result
//     transposed through LDS to the (side, pair) rows;
//  3. rotation back on the vector ALUs, units (m, mup):
//       out_q(m, mup, mu) = sum_khi gamma_q(m, khi, mu) conj(R(mup, khi))
//       out_-q(m, mup, mu) = (-1)^m sum_khi gamma_-q(m, khi, mu) conj(R(mup, -khi))
//     into LDS, then copied out record by record (coalesced).
// R is formed from the factorisation with d^m from its Fourier series (both
// exact in FP64 to rounding), then chopped as the reference chops its
// recursion's result; the OZ sums are summed in the matrix unit's order (a
// rounding-level difference from the reference's order).
// BF16 contexts run the FP32 kernel with the contraction in BF16 operands and
// FP32 accumulation (v_mfma_f32_16x16x32_bf16: 16 complex columns per
// instruction), the c(k) fragments stored in BF16.
#include <algorithm>
#include <cfloat>
#include <utility>
#include <vector>

#include "mdft_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;

template <typename T> struct Ty;
template <> struct Ty<double> {
    using c = double2;
    using acc = d4;
};
template <> struct Ty<float> {
    using c = float2;
    using acc = f4;
};

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// Operand layout of both (lane l): A[i][k] with i = l & 15, k = l >> 4; B[k][j] with
// k = l >> 4, j = l & 15; result register r of lane l holds D[drow(l >> 4, r)][l & 15],
// rows 4 g + r for f32 and 4 r + g for f64 (measured on gfx950).
template <typename T> __device__ __forceinline__ constexpr int drow(int g, int r);
template <> __device__ __forceinline__ constexpr int drow<float>(int g, int r) { return 4 * g + r; }
template <> __device__ __forceinline__ constexpr int drow<double>(int g, int r) { return 4 * r + g; }

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FLT_EPSILON chop of an R entry, keeping the surviving component in single
// precision (Fortran's default-kind CMPLX of :491-492)
__device__ inline double2 chop_R(double2 z) {
    double re = z.x, im = z.y;
    if (fabs(re) < (double)FLT_EPSILON) {
        re = 0.0;
        im = (double)(float)im;
    }
    if (fabs(im) < (double)FLT_EPSILON) {
        im = 0.0;
        re = (double)(float)re;
    }
    return make_double2(re, im);
}
__device__ inline float2 chop_R(float2 z) { return z; }

// DBL_EPSILON chop of gamma (:618-623)
__device__ inline void chop_gamma(double &re, double &im) {
    if (fabs(re) < DBL_EPSILON && fabs(im) < DBL_EPSILON) {
        re = 0.0;
        im = 0.0;
    }
    if (fabs(re) < DBL_EPSILON) {
        re = 0.0;
        im = (double)(float)im;
    }
    if (fabs(im) < DBL_EPSILON) {
        im = 0.0;
        re = (double)(float)re;
    }
}
__device__ inline void chop_gamma(float &, float &) {}

// ---- the angular plan of one mmax (molecular symmetry order 2) -------------------
// Projection index p(m, mup, mu2) = poff(m) + (mup + m) (m/2 + 1) + mu2 (the
// context's projection order).  OZ block khi: rows (m, mu2), m = |khi|..L, and
// columns (n, nu2), n = |khi|..L, nu2 = -n/2..n/2, both ascending (the c(k)
// relayout order of mdft_set_ck).
template <int L> struct Plan {
    static constexpr int iabs(int x) { return x < 0 ? -x : x; }
    static constexpr int nmu(int m) { return m / 2 + 1; }
    static constexpr int poff(int m) {
        int o = 0;
        for (int l = 0; l < m; l++) o += (2 * l + 1) * nmu(l);
        return o;
    }
    static constexpr int np = poff(L + 1);
    static constexpr int nps = np | 1;  // LDS row stride (odd: rows fall in different banks)
    static constexpr int kp = ((L + 2) / 2) * 2;  // Fourier coefficients stored per d entry (even)
    static constexpr int nclass = L / 2 + 1;      // rotation classes (m/2 + 1 = 1 .. nclass)
    static constexpr int pidx(int m, int mup, int mu2) { return poff(m) + (mup + m) * nmu(m) + mu2; }
    static constexpr int roff(int m) { return m * (4 * m * m - 1) / 3; }
    static constexpr int ncols(int khi) {
        int c = 0;
        for (int n = iabs(khi); n <= L; n++) c += 2 * (n / 2) + 1;
        return c;
    }
    static constexpr int nrows(int khi) {
        int r = 0;
        for (int m = iabs(khi); m <= L; m++) r += nmu(m);
        return r;
    }
    // column c of block khi: projection entry << 1 | (nu2 < 0); padding columns read entry 0
    static constexpr int col(int khi, int c) {
        for (int n = iabs(khi); n <= L; n++)
            for (int nu2 = -(n / 2); nu2 <= n / 2; nu2++) {
                if (c == 0) return pidx(n, khi, iabs(nu2)) << 1 | (nu2 < 0 ? 1 : 0);
                c--;
            }
        return 1;
    }
    // row r of block khi: its projection entry, -1 for padding rows
    static constexpr int row(int khi, int r) {
        for (int m = iabs(khi); m <= L; m++)
            for (int mu2 = 0; mu2 < nmu(m); mu2++) {
                if (r == 0) return pidx(m, khi, mu2);
                r--;
            }
        return -1;
    }
    static constexpr int nkb(int khi) { return (ncols(khi) + 1) / 2; }
    static constexpr int nnb(int khi) { return (nrows(khi) + 3) / 4; }
    static constexpr int foff(int khi) {
        int f = 0;
        for (int k = -L; k < khi; k++) f += nkb(k) * nnb(k);
        return f;
    }
    static constexpr int nfrag = foff(L + 1);
    // BF16 contraction: k blocks of 32 (16 complex columns)
    static constexpr int nkb32(int khi) { return (ncols(khi) + 15) / 16; }
    static constexpr int foff32(int khi) {
        int f = 0;
        for (int k = -L; k < khi; k++) f += nkb32(k) * nnb(k);
        return f;
    }
    static constexpr int nfrag32 = foff32(L + 1);
};

constexpr int kWaves = 4;  // waves per tile

template <typename T> struct OzmArgs {
    int ntiles, per;
    const TilePair *tpair;      // [tile][8] the tiles' pairs
    const int *pmup;            // mup of each projection
    const T *frag;              // [interval][nfrag][64] B fragments of (c[iq], c[iq+1])
    const bf16x8 *frag16;       // BF16 contexts: [interval][nfrag32][64] B fragments, 8 values per lane
    const int *colent;          // [2 L + 1][32] OZ column entries (entry << 1 | own side) of each block
    const T *ftab;              // [entry][kp] Fourier coefficients of d^m(theta)
    const int4 *units;          // [round][8] rotation units
    const int *rlen;            // [round] iterations of each round
    int cr0[2][kWaves][4], cnr[2][kWaves][4];  // first round and count per (direction, wave, class)
};

// transpose scratch of the OZ epilogue: per wave a 16 x 16 block of the products, rows
// padded so that a lane's four values are one aligned 16-byte-multiple read
template <typename T> constexpr int scr_stride() { return sizeof(T) == 8 ? 18 : 20; }

// LDS of one tile: X (then the rotated-back outputs; during the OZ step the transpose
// scratch) and A (then gamma) per (side, pair) row, the pairs' angles, interpolation
// weights and grid indices
template <typename T, int L> struct Lds {
    using C = typename Ty<T>::c;
    static constexpr int kScr = kWaves * 16 * scr_stride<T>();  // scratch values (T)
    static constexpr int xrows() {
        const int rowb = Plan<L>::nps * (int)sizeof(C), need = (kScr * (int)sizeof(T) + rowb - 1) / rowb;
        return need > 16 ? need : 16;
    }
    static constexpr int kXRows = xrows();
    C xs[kXRows][Plan<L>::nps];
    C buf[16][Plan<L>::nps];
    T trig[8][2][Plan<L>::kp];  // cos k theta, sin k theta of each pair
    C ph[8][Plan<L>::kp];       // exp(+i k phi) of each pair
    C ab[8];                    // (alpha, 1 - alpha) of each pair
    int g[16];                  // grid index of each row (q side rows 0-7, -q side rows 8-15)
    int flag[8];                // pair flags; bit 2: padding pair of a partial tile
};

// ---- rotations: one class of units (m/2 + 1 = NMU) ----------------------------------
// rotation in: X from s.xs, A to s.buf; rotation back: gamma from s.buf, the outputs
// (sign, Nyquist conjugate applied) to s.xs
template <typename T, int L, int NMU, bool BACK>
__device__ __forceinline__ void rot_class(const OzmArgs<T> &a, Lds<T, L> &s, int r0, int nr, int p, int j) {
    using C = typename Ty<T>::c;
    constexpr int KP = Plan<L>::kp;
    constexpr int KC = (2 * NMU < KP) ? 2 * NMU : KP;  // coefficients that can be nonzero for m <= 2 NMU - 1
    // FP32 (no chop to honour): R = exp(-i mup phi) d factored, X carries the phase from its
    // load into LDS, the rotation back applies it once per output; real d products in between
    constexpr bool FACT = sizeof(T) == 4;
    const T *ct = s.trig[p][0], *st = s.trig[p][1];
    const C *xq = s.xs[p], *xmq = s.xs[8 + p];
    const C *gq = s.buf[p], *gmq = s.buf[8 + p];
    const bool conj_mq = (s.flag[p] & 3) == 3;  // q == -q on a Nyquist plane
    for (int r = 0; r < nr; r++) {
        const int4 u = a.units[(r0 + r) * 8 + j];
        const int len = a.rlen[r0 + r];
        const int x0 = u.x & 0xffff, e1 = u.x >> 16, e2 = u.y & 0xffff, f0 = u.y >> 16;
        const int m = u.z & 255, w = (u.z >> 8) & 255, p0 = (u.z >> 16) & 1;
        const int sub = u.w - 64;  // khi (rotation in) or mup (rotation back)
        T ta[KC], tb[KC];          // trig of the even / odd iterations: sin when mup - khi is odd
#pragma unroll
        for (int k = 0; k < KC; k++) {
            ta[k] = p0 ? st[k] : ct[k];
            tb[k] = p0 ? ct[k] : st[k];
        }
        C acq[NMU], acm[NMU];
#pragma unroll
        for (int i = 0; i < NMU; i++) {
            acq[i].x = acq[i].y = (T)0;
            acm[i].x = acm[i].y = (T)0;
        }
        C eph;  // rotation back: exp(-i mup phi) of the unit
        {
            const C z = s.ph[p][sub < 0 ? -sub : sub];
            eph.x = z.x;
            eph.y = sub < 0 ? z.y : -z.y;
        }
        const int fstep = BACK ? KP : w * KP;
        // Fourier coefficients of one iteration, loaded one iteration ahead of their use
        struct Ops {
            C f[KC / 2];
        };
        auto load = [&](int t, Ops &o) {
            if (t >= w) return;
            const C *F = reinterpret_cast<const C *>(a.ftab + f0 + t * fstep);
#pragma unroll
            for (int k = 0; k < KC / 2; k++) o.f[k] = F[k];
        };
        auto step = [&](int t, const Ops &o, const T(&tg)[KC]) {
            if (t >= w) return;
            T v = (T)0;
#pragma unroll
            for (int k = 0; k < KC / 2; k++) {
                v += o.f[k].x * tg[2 * k];
                v += o.f[k].y * tg[2 * k + 1];
            }
            if constexpr (FACT) {
#pragma unroll
                for (int i = 0; i < NMU; i++) {
                    const C y1 = BACK ? gq[x0 + t * NMU + i] : xq[x0 + t * NMU + i];
                    const C y2 = BACK ? gmq[x0 + (2 * m - t) * NMU + i] : xmq[x0 + t * NMU + i];
                    acq[i].x = fma(y1.x, v, acq[i].x);
                    acq[i].y = fma(y1.y, v, acq[i].y);
                    acm[i].x = fma(y2.x, v, acm[i].x);
                    acm[i].y = fma(y2.y, v, acm[i].y);
                }
                return;
            }
            C e;
            if (BACK) {
                e = eph;
            } else {
                const int mup = t - m;
                const C z = s.ph[p][mup < 0 ? -mup : mup];
                e.x = z.x;
                e.y = mup < 0 ? z.y : -z.y;
            }
            C R;
            R.x = e.x * v;
            R.y = e.y * v;
            R = chop_R(R);
#pragma unroll
            for (int i = 0; i < NMU; i++) {
                if (!BACK) {  // acc += x R, four multiply-adds
                    const C x1 = xq[x0 + t * NMU + i], x2 = xmq[x0 + t * NMU + i];
                    acq[i].x = fma(x1.x, R.x, acq[i].x);
                    acq[i].x = fma(-x1.y, R.y, acq[i].x);
                    acq[i].y = fma(x1.x, R.y, acq[i].y);
                    acq[i].y = fma(x1.y, R.x, acq[i].y);
                    acm[i].x = fma(x2.x, R.x, acm[i].x);
                    acm[i].x = fma(-x2.y, R.y, acm[i].x);
                    acm[i].y = fma(x2.x, R.y, acm[i].y);
                    acm[i].y = fma(x2.y, R.x, acm[i].y);
                } else {  // acc += g conj(R)
                    const C g1 = gq[x0 + t * NMU + i], g2 = gmq[x0 + (2 * m - t) * NMU + i];
                    acq[i].x = fma(g1.x, R.x, acq[i].x);
                    acq[i].x = fma(g1.y, R.y, acq[i].x);
                    acq[i].y = fma(g1.y, R.x, acq[i].y);
                    acq[i].y = fma(-g1.x, R.y, acq[i].y);
                    acm[i].x = fma(g2.x, R.x, acm[i].x);
                    acm[i].x = fma(g2.y, R.y, acm[i].x);
                    acm[i].y = fma(g2.y, R.x, acm[i].y);
                    acm[i].y = fma(-g2.x, R.y, acm[i].y);
                }
            }
        };
        Ops o0, o1;
        load(0, o0);
        for (int t = 0; t < len; t += 2) {
            load(t + 1, o1);
            step(t, o0, ta);
            load(t + 2, o0);
            step(t + 1, o1, tb);
        }
        if (w > 0) {
            const T sg = (m & 1) ? (T)-1 : (T)1;
            if (FACT && BACK) {  // exp(+i mup phi) = conj(eph)
#pragma unroll
                for (int i = 0; i < NMU; i++) {
                    const C u = acq[i], v = acm[i];
                    acq[i].x = u.x * eph.x + u.y * eph.y;
                    acq[i].y = u.y * eph.x - u.x * eph.y;
                    acm[i].x = v.x * eph.x + v.y * eph.y;
                    acm[i].y = v.y * eph.x - v.x * eph.y;
                }
            }
#pragma unroll
            for (int i = 0; i < NMU; i++) {
                C o;
                o.x = sg * acm[i].x;
                o.y = sg * acm[i].y;
                if (!BACK) {
                    s.buf[p][e1 + i] = acq[i];
                    s.buf[8 + p][e2 + i] = o;
                } else {
                    if (conj_mq) o.y = -o.y;
                    s.xs[p][e1 + i] = acq[i];
                    s.xs[8 + p][e1 + i] = o;
                }
            }
        }
        (void)e2;
        (void)sub;
    }
}

template <typename T, int L, int NMU, bool BACK>
__device__ __forceinline__ void rot_all(const OzmArgs<T> &a, Lds<T, L> &s, int w, int p, int j) {
    rot_class<T, L, NMU, BACK>(a, s, a.cr0[BACK][w][NMU - 1], a.cnr[BACK][w][NMU - 1], p, j);
    if constexpr (NMU > 1) rot_all<T, L, NMU - 1, BACK>(a, s, w, p, j);
}

// ---- OZ block khi on the matrix cores -------------------------------------------------
// A operand (16 x 4 per instruction): lane l holds row l & 15 (side (l >> 3) & 1, pair
// l & 7), k = l >> 4 of the block: column c = 2 kb + (l >> 5), re (k even) or im.
template <typename T, int L, int KHI, int KB>
__device__ __forceinline__ T xfrag(const typename Ty<T>::c *own, const typename Ty<T>::c *oth, bool hi, bool pk) {
    constexpr int c0 = Plan<L>::col(KHI, 2 * KB), c1 = Plan<L>::col(KHI, 2 * KB + 1);
    const int e = hi ? (c1 >> 1) : (c0 >> 1);
    const bool isown = hi ? (c1 & 1) : (c0 & 1);
    const typename Ty<T>::c v = (isown ? own : oth)[e];
    return pk ? (isown ? v.y : -v.y) : v.x;
}

template <typename T, int L, int KHI, int... KB>
__device__ __forceinline__ void load_x(T *x, const typename Ty<T>::c *own, const typename Ty<T>::c *oth, bool hi,
                                       bool pk, std::integer_sequence<int, KB...>) {
    ((x[KB] = xfrag<T, L, KHI, KB>(own, oth, hi, pk)), ...);
}

// D (16 x 16) of n-block NB: lane l holds rows drow(l >> 4, i), i = 0..3, column l & 15 =
// 4 rr + q (OZ row 4 NB + rr; q = re lo, im lo, re hi, im hi).  Through the wave's scratch
// lane l then takes (side, pair) row ps = l & 15 and OZ row 4 NB + (l >> 4):
// gamma = (1 - alpha) lo + alpha hi, chopped, into the row's entry of that OZ row.
template <typename T, int L, int KHI, int NB>
__device__ __forceinline__ void gamma_store(typename Ty<T>::acc d, Lds<T, L> &s, T *scr, T al, T be, int l) {
    using C = typename Ty<T>::c;
    constexpr int SS = scr_stride<T>();
    constexpr int r0 = Plan<L>::row(KHI, 4 * NB), r1 = Plan<L>::row(KHI, 4 * NB + 1);
    constexpr int r2 = Plan<L>::row(KHI, 4 * NB + 2), r3 = Plan<L>::row(KHI, 4 * NB + 3);
    const int g = l >> 4, n = l & 15;
#pragma unroll
    for (int i = 0; i < 4; i++) scr[drow<T>(g, i) * SS + n] = d[i];
    wave_sync();
    const int ps = l & 15, rr = l >> 4;
    const C *v = reinterpret_cast<const C *>(scr + ps * SS + 4 * rr);
    const C lo = v[0], hi = v[1];
    T re = be * lo.x + al * hi.x, im = be * lo.y + al * hi.y;
    chop_gamma(re, im);
    const int row = rr == 0 ? r0 : (rr == 1 ? r1 : (rr == 2 ? r2 : r3));
    if (row >= 0) {
        C o;
        o.x = re;
        o.y = im;
        s.buf[ps][row] = o;
    }
    wave_sync();  // the scratch is read before the next block writes it
}

template <typename T, int L, int KHI, int... NB>
__device__ __forceinline__ void store_all(typename Ty<T>::acc *acc, Lds<T, L> &s, T *scr, T al, T be, int l,
                                          std::integer_sequence<int, NB...>) {
    (gamma_store<T, L, KHI, NB>(acc[NB], s, scr, al, be, l), ...);
}

// blocks khi = KHI, KHI + kWaves, ... (the blocks of one wave; blocks are independent: block
// khi reads and writes only the entries (., khi, .) of the rows)
template <typename T, int L, int KHI>
__device__ __forceinline__ void oz_blocks(Lds<T, L> &s, const T *__restrict__ fr, T *scr, T al, T be, int l) {
    using P = Plan<L>;
    using A = typename Ty<T>::acc;
    constexpr int NKB = P::nkb(KHI), NNB = P::nnb(KHI), F0 = P::foff(KHI);
    const int row16 = l & 15;
    const bool hi = (l >> 5) & 1, pk = (l >> 4) & 1;
    T b[NKB * NNB];
#pragma unroll
    for (int f = 0; f < NKB * NNB; f++) b[f] = fr[(F0 + f) * 64 + l];
    T x[NKB];
    load_x<T, L, KHI>(x, s.buf[row16], s.buf[row16 ^ 8], hi, pk, std::make_integer_sequence<int, NKB>{});
    A acc[NNB];
#pragma unroll
    for (int nb = 0; nb < NNB; nb++) acc[nb] = A{0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < NKB; kb++)
#pragma unroll
        for (int nb = 0; nb < NNB; nb++) acc[nb] = mfma16(x[kb], b[kb * NNB + nb], acc[nb]);
    // every read of block khi precedes its writes in the wave's LDS order
    store_all<T, L, KHI>(acc, s, scr, al, be, l, std::make_integer_sequence<int, NNB>{});
    if constexpr (KHI + kWaves <= L) oz_blocks<T, L, KHI + kWaves>(s, fr, scr, al, be, l);
}

// BF16 contraction of block khi (FP32 rotations and epilogue): A (16 x 32) lane l holds row
// l & 15, k = 8 (l >> 4) + j: column 16 kb + 4 (l >> 4) + (j >> 1), re (j even) or im; B the
// same k, column l & 15; D as in the FP32 form.
template <int L, int KHI>
__device__ __forceinline__ void oz_blocks_bf(Lds<float, L> &s, const bf16x8 *__restrict__ fr,
                                             const int *__restrict__ colent, float *scr, float al, float be, int l) {
    using P = Plan<L>;
    constexpr int NKB = P::nkb32(KHI), NNB = P::nnb(KHI), F0 = P::foff32(KHI);
    const int row16 = l & 15, g = l >> 4;
    const float2 *own = s.buf[row16], *oth = s.buf[row16 ^ 8];
    const int *ce = colent + (KHI + L) * 32 + 4 * g;
    f4 acc[NNB];
#pragma unroll
    for (int nb = 0; nb < NNB; nb++) acc[nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
        bf16x8 a;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = ce[16 * kb + u];
            const bool isown = e & 1;
            const float2 v = (isown ? own : oth)[e >> 1];
            a[2 * u] = (__bf16)v.x;
            a[2 * u + 1] = (__bf16)(isown ? v.y : -v.y);
        }
#pragma unroll
        for (int nb = 0; nb < NNB; nb++)
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, fr[(F0 + kb * NNB + nb) * 64 + l], acc[nb], 0, 0, 0);
    }
    store_all<float, L, KHI>(acc, s, scr, al, be, l, std::make_integer_sequence<int, NNB>{});
    if constexpr (KHI + kWaves <= L) oz_blocks_bf<L, KHI + kWaves>(s, fr, colent, scr, al, be, l);
}

template <typename T, int L, bool BF, int W0>
__device__ __forceinline__ void oz_wave(Lds<T, L> &s, const OzmArgs<T> &a, size_t iv, T *scr, T al, T be, int w,
                                        int l) {
    if constexpr (W0 < kWaves && -L + W0 <= L) {
        if (w == W0) {
            if constexpr (BF) oz_blocks_bf<L, -L + W0>(s, a.frag16 + iv * Plan<L>::nfrag32 * 64, a.colent, scr, al, be, l);
            else oz_blocks<T, L, -L + W0>(s, a.frag + iv * Plan<L>::nfrag * 64, scr, al, be, l);
        } else {
            oz_wave<T, L, BF, W0 + 1>(s, a, iv, scr, al, be, w, l);
        }
    }
}

// FP32: at least 7 waves per SIMD (72 registers, no spills; at the compiler's own choice of 88
// registers, 5 waves and 2.13 against 2.01 ms at 128^3, mmax 5)
template <typename T, int L, bool BF>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 7 : 1, 8))) void ozm(typename Ty<T>::c *__restrict__ proj, OzmArgs<T> a) {
    using P = Plan<L>;
    using C = typename Ty<T>::c;
    constexpr int KP = P::kp, NP = P::np;
    extern __shared__ __align__(16) unsigned char smem[];
    Lds<T, L> &s = *reinterpret_cast<Lds<T, L> *>(smem);
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    const int p = l & 7, j = l >> 3;
    // XCD-aware order: consecutive workgroups go to different XCDs, so each XCD takes a
    // contiguous range of tiles (tiles of one interval share its c fragments in that L2)
    const int b = blockIdx.x;
    const int tile = (b & 7) * a.per + (b >> 3);
    if (tile >= a.ntiles) return;
    const int iq = a.tpair[(size_t)tile * 8].iq;  // the same for the tile's pairs
    if (tid < 8) {
        const TilePair tp = a.tpair[(size_t)tile * 8 + tid];
        s.g[tid] = tp.gq;
        s.g[8 + tid] = tp.gmq;
        s.flag[tid] = tp.flag;
        C abv;
        abv.x = (T)tp.al;
        abv.y = (T)(1.0 - tp.al);
        s.ab[tid] = abv;
        double ck = 1.0, sk = 0.0, cpk = 1.0, spk = 0.0;
        for (int k = 0; k < KP; k++) {
            s.trig[tid][0][k] = (T)ck;
            s.trig[tid][1][k] = (T)sk;
            C e;
            e.x = (T)cpk;
            e.y = (T)spk;
            s.ph[tid][k] = e;
            const double nc = ck * tp.ct - sk * tp.st;
            sk = sk * tp.ct + ck * tp.st;
            ck = nc;
            const double nq = cpk * tp.cp - spk * tp.sp;
            spk = spk * tp.cp + cpk * tp.sp;
            cpk = nq;
        }
    }
    __syncthreads();
    // X of the 16 (side, pair) rows, each row one contiguous record; in FP32 times
    // exp(-i mup phi) for the factored rotation, two elements per 16-byte access when the
    // records are a whole number of 16-byte words
    if constexpr (sizeof(T) == 4 && NP % 2 == 0) {
        constexpr int NP2 = NP / 2;
        for (int i = tid; i < 16 * NP2; i += 64 * kWaves) {
            const int r = i / NP2, e = 2 * (i - r * NP2);
            const float4 x2 = reinterpret_cast<const float4 *>(proj + (size_t)s.g[r] * NP)[e / 2];
            C out[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int mup = a.pmup[e + h];
                const C z = s.ph[r & 7][mup < 0 ? -mup : mup];
                const T ex = z.x, ey = mup < 0 ? z.y : -z.y;
                const T ux = h ? x2.z : x2.x, uy = h ? x2.w : x2.y;
                out[h].x = ux * ex - uy * ey;
                out[h].y = ux * ey + uy * ex;
            }
            s.xs[r][e] = out[0];
            s.xs[r][e + 1] = out[1];
        }
    } else {
        for (int i = tid; i < 16 * NP; i += 64 * kWaves) {
            const int r = i / NP, e = i - r * NP;
            C x = proj[(size_t)s.g[r] * NP + e];
            if constexpr (sizeof(T) == 4) {
                const int mup = a.pmup[e];
                const C z = s.ph[r & 7][mup < 0 ? -mup : mup];
                const T ex = z.x, ey = mup < 0 ? z.y : -z.y;
                const C u = x;
                x.x = u.x * ex - u.y * ey;
                x.y = u.x * ey + u.y * ex;
            }
            s.xs[r][e] = x;
        }
    }
    __syncthreads();
    rot_all<T, L, P::nclass, false>(a, s, w, p, j);
    __syncthreads();
    {
        const C abv = s.ab[l & 7];  // the pair of row l & 15
        T *scr = reinterpret_cast<T *>(&s.xs[0][0]) + w * 16 * scr_stride<T>();
        oz_wave<T, L, BF, 0>(s, a, (size_t)(iq - 1), scr, abv.x, abv.y, w, l);
    }
    __syncthreads();
    rot_all<T, L, P::nclass, true>(a, s, w, p, j);
    __syncthreads();
    // store rule (:648-656): both sides of real pairs; for q == -q only the -q side
    if constexpr (sizeof(T) == 4 && NP % 2 == 0) {
        constexpr int NP2 = NP / 2;
        for (int i = tid; i < 16 * NP2; i += 64 * kWaves) {
            const int r = i / NP2, e = 2 * (i - r * NP2);
            const int f = s.flag[r & 7];
            if ((f & 4) || (r < 8 && (f & 1))) continue;
            const C u0 = s.xs[r][e], u1 = s.xs[r][e + 1];
            reinterpret_cast<float4 *>(proj + (size_t)s.g[r] * NP)[e / 2] = make_float4(u0.x, u0.y, u1.x, u1.y);
        }
    } else {
        for (int i = tid; i < 16 * NP; i += 64 * kWaves) {
            const int r = i / NP, e = i - r * NP;
            const int f = s.flag[r & 7];
            if ((f & 4) || (r < 8 && (f & 1))) continue;
            proj[(size_t)s.g[r] * NP + e] = s.xs[r][e];
        }
    }
}

// ---- host side ---------------------------------------------------------------------
double textbook_d_h(int j, int mp, int m, double beta) {
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

struct HostPlan {
    int L, np, kp, nclass, nfrag, nfrag32;
    std::vector<int> col[2 * MDFT_MAX_MMAX + 1], row[2 * MDFT_MAX_MMAX + 1];
    std::vector<int> nkb, nnb, foff, ncol, nkb32, foff32;
};

template <int L> HostPlan host_plan() {
    using P = Plan<L>;
    HostPlan h;
    h.L = L;
    h.np = P::np;
    h.kp = P::kp;
    h.nclass = P::nclass;
    h.nfrag = P::nfrag;
    h.nfrag32 = P::nfrag32;
    for (int khi = -L; khi <= L; khi++) {
        for (int c = 0; c < 32; c++) h.col[khi + L].push_back(P::col(khi, c));
        for (int r = 0; r < 4 * P::nnb(khi); r++) h.row[khi + L].push_back(P::row(khi, r));
        h.nkb.push_back(P::nkb(khi));
        h.nnb.push_back(P::nnb(khi));
        h.foff.push_back(P::foff(khi));
        h.ncol.push_back(P::ncols(khi));
        h.nkb32.push_back(P::nkb32(khi));
        h.foff32.push_back(P::foff32(khi));
    }
    return h;
}

HostPlan host_plan_of(int L) {
    switch (L) {
        case 0: return host_plan<0>();
        case 1: return host_plan<1>();
        case 2: return host_plan<2>();
        case 3: return host_plan<3>();
        case 4: return host_plan<4>();
        case 5: return host_plan<5>();
        default: return host_plan<6>();
    }
}

template <typename T, int L, bool BF> hipError_t launch_ozm_t(mdft_ctx *c, typename Ty<T>::c *proj) {
    OzmArgs<T> a;
    a.ntiles = c->pairs.ntiles;
    a.per = (c->pairs.ntiles + 7) / 8;
    a.tpair = c->pairs.tpair;
    a.pmup = c->d_pmup;
    a.frag = static_cast<const T *>(c->d_ozm_frag);
    a.frag16 = static_cast<const bf16x8 *>(c->d_ozm_frag16);
    a.colent = c->d_ozm_col;
    a.ftab = static_cast<const T *>(c->d_ozm_ftab);
    a.units = reinterpret_cast<const int4 *>(c->d_ozm_units);
    a.rlen = c->d_ozm_rlen;
    for (int d = 0; d < 2; d++)
        for (int wv = 0; wv < kWaves; wv++)
            for (int k = 0; k < 4; k++) {
                a.cr0[d][wv][k] = c->ozm_cr0[d][wv][k];
                a.cnr[d][wv][k] = c->ozm_cnr[d][wv][k];
            }
    if (a.ntiles == 0) return hipSuccess;
    const size_t lds = sizeof(Lds<T, L>);
    if (lds > c->lds_limit) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((ozm<T, L, BF>), dim3((unsigned)(8 * a.per)), dim3(64 * kWaves), lds, c->stream, proj, a);
    return hipGetLastError();
}

template <typename T, bool BF> hipError_t launch_ozm_any(mdft_ctx *c, typename Ty<T>::c *proj) {
    switch (c->L) {
        case 0: return launch_ozm_t<T, 0, BF>(c, proj);
        case 1: return launch_ozm_t<T, 1, BF>(c, proj);
        case 2: return launch_ozm_t<T, 2, BF>(c, proj);
        case 3: return launch_ozm_t<T, 3, BF>(c, proj);
        case 4: return launch_ozm_t<T, 4, BF>(c, proj);
        case 5: return launch_ozm_t<T, 5, BF>(c, proj);
        case 6: return launch_ozm_t<T, 6, BF>(c, proj);
        default: return hipErrorInvalidConfiguration;
    }
}

template <typename T> int upload_t(mdft_ctx *c, void **dst, const std::vector<double> &src) {
    std::vector<T> v(src.begin(), src.end());
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    if (hipMalloc(dst, sizeof(T) * (v.size() ? v.size() : 1)) != hipSuccess) return MDFT_ENOMEM;
    if (!v.empty() && hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
        return MDFT_EHIP;
    return MDFT_OK;
}

}  // namespace

static size_t ozm_lds(int L, bool fp64) {
    switch (L) {
        case 0: return fp64 ? sizeof(Lds<double, 0>) : sizeof(Lds<float, 0>);
        case 1: return fp64 ? sizeof(Lds<double, 1>) : sizeof(Lds<float, 1>);
        case 2: return fp64 ? sizeof(Lds<double, 2>) : sizeof(Lds<float, 2>);
        case 3: return fp64 ? sizeof(Lds<double, 3>) : sizeof(Lds<float, 3>);
        case 4: return fp64 ? sizeof(Lds<double, 4>) : sizeof(Lds<float, 4>);
        case 5: return fp64 ? sizeof(Lds<double, 5>) : sizeof(Lds<float, 5>);
        default: return fp64 ? sizeof(Lds<double, 6>) : sizeof(Lds<float, 6>);
    }
}

bool ozm_supported(const mdft_ctx *c) {
    return c->d.mrso == 2 && c->L >= 0 && c->L <= 6 &&
           ozm_lds(c->L, c->d.precision == MDFT_PREC_FP64) <= c->lds_limit;
}

// Fourier coefficients of d^m(theta) per entry (m, a, b) (kp per entry, zero beyond m) and
// the rotation round tables of the context's mmax.
int ozm_setup(mdft_ctx *c) {
    if (!ozm_supported(c)) return MDFT_OK;
    const HostPlan h = host_plan_of(c->L);
    const int L = c->L, kp = h.kp;
    auto roff = [](int m) { return m * (4 * m * m - 1) / 3; };
    const double hp = 0.5 * acos(-1.0);
    std::vector<double> ftab((size_t)roff(L + 1) * kp, 0.0);
    for (int m = 0; m <= L; m++) {
        const int W = 2 * m + 1;
        std::vector<double> D(W * W);
        for (int k = -m; k <= m; k++)
            for (int x = -m; x <= m; x++) D[(k + m) * W + (x + m)] = textbook_d_h(m, k, x, hp);
        // d^m_{a,b}(theta) = i^{a-b} sum_k D_{k,a} D_{k,b} exp(-i k theta) with D = d^m(pi/2):
        // cos(k theta) coefficients for even a - b, sin(k theta) for odd
        for (int x = -m; x <= m; x++)
            for (int y = -m; y <= m; y++) {
                const int ab = ((x - y) % 4 + 4) % 4;
                const size_t e = (size_t)roff(m) + (x + m) * W + (y + m);
                for (int k = 0; k <= m; k++) {
                    const double sk = D[(k + m) * W + (x + m)] * D[(k + m) * W + (y + m)];
                    const double smk = D[(-k + m) * W + (x + m)] * D[(-k + m) * W + (y + m)];
                    double v;
                    if (k == 0) v = (ab == 0) ? sk : (ab == 2 ? -sk : 0.0);
                    else if (ab == 0) v = sk + smk;
                    else if (ab == 2) v = -(sk + smk);
                    else if (ab == 1) v = sk - smk;
                    else v = -(sk - smk);
                    ftab[e * kp + k] = v;
                }
            }
    }
    auto nmu = [](int m) { return m / 2 + 1; };
    std::vector<int> poff(L + 2, 0);
    for (int m = 0; m <= L; m++) poff[m + 1] = poff[m] + (2 * m + 1) * nmu(m);
    auto pidx = [&](int m, int mup, int mu2) { return poff[m] + (mup + m) * nmu(m) + mu2; };
    // rounds of 8 units (m, khi) (rotation in) or (m, mup) (rotation back) of one class
    // m/2 + 1 = k, m descending; a round runs as long as its longest unit (2 m + 1)
    struct Round {
        int back, k, len;
        int4 u[8];
    };
    std::vector<Round> rounds;
    for (int back = 0; back < 2; back++)
        for (int k = 1; k <= h.nclass; k++) {
            std::vector<std::pair<int, int>> us;
            for (int m = std::min(L, 2 * k - 1); m >= 2 * k - 2; m--)
                for (int x = -m; x <= m; x++) us.push_back({m, x});
            for (size_t u0 = 0; u0 < us.size(); u0 += 8) {
                Round r{back, k, 2 * us[u0].first + 1, {}};
                for (int jj = 0; jj < 8; jj++) {
                    int4 e{0, 0, 0, 64};
                    if (u0 + jj < us.size()) {
                        const int m = us[u0 + jj].first, x = us[u0 + jj].second, W = 2 * m + 1;
                        if (!back) {  // rotation in, unit (m, khi = x): iterate mup
                            e.x = poff[m] | pidx(m, x, 0) << 16;
                            e.y = pidx(m, -x, 0) | ((roff(m) + (x + m)) * kp) << 16;
                        } else {  // rotation back, unit (m, mup = x): iterate khi
                            e.x = poff[m] | pidx(m, x, 0) << 16;
                            e.y = ((roff(m) + (x + m) * W) * kp) << 16;
                        }
                        e.z = m | W << 8 | ((m + x) & 1) << 16;
                        e.w = x + 64;
                    }
                    r.u[jj] = e;
                }
                rounds.push_back(r);
            }
        }
    // rounds to waves, longest first onto the least loaded wave (cost ~ iterations x (m/2 + 1))
    std::vector<int> owner(rounds.size());
    for (int back = 0; back < 2; back++) {
        std::vector<int> idx;
        for (int i = 0; i < (int)rounds.size(); i++)
            if (rounds[i].back == back) idx.push_back(i);
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            return rounds[x].len * rounds[x].k > rounds[y].len * rounds[y].k;
        });
        std::vector<int> load(kWaves, 0);
        for (int i : idx) {
            const int wv = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            owner[i] = wv;
            load[wv] += rounds[i].len * rounds[i].k;
        }
    }
    std::vector<int> units, rlen;
    for (int back = 0; back < 2; back++)
        for (int wv = 0; wv < kWaves; wv++)
            for (int k = 1; k <= 4; k++) {
                c->ozm_cr0[back][wv][k - 1] = (int)rlen.size();
                c->ozm_cnr[back][wv][k - 1] = 0;
                for (size_t i = 0; i < rounds.size(); i++) {
                    const Round &r = rounds[i];
                    if (r.back != back || r.k != k || owner[i] != wv) continue;
                    rlen.push_back(r.len);
                    c->ozm_cnr[back][wv][k - 1]++;
                    for (int jj = 0; jj < 8; jj++) {
                        units.push_back(r.u[jj].x);
                        units.push_back(r.u[jj].y);
                        units.push_back(r.u[jj].z);
                        units.push_back(r.u[jj].w);
                    }
                }
            }
    int rc = c->d.precision == MDFT_PREC_FP64 ? upload_t<double>(c, &c->d_ozm_ftab, ftab)
                                                : upload_t<float>(c, &c->d_ozm_ftab, ftab);
    if (rc != MDFT_OK) return rc;
    if (hipMalloc((void **)&c->d_ozm_units, sizeof(int) * units.size()) != hipSuccess ||
        hipMalloc((void **)&c->d_ozm_rlen, sizeof(int) * rlen.size()) != hipSuccess)
        return MDFT_ENOMEM;
    if (hipMemcpy(c->d_ozm_units, units.data(), sizeof(int) * units.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_ozm_rlen, rlen.data(), sizeof(int) * rlen.size(), hipMemcpyHostToDevice) != hipSuccess)
        return MDFT_EHIP;
    // OZ column entries of each block (entry << 1 | own side; padding columns read entry 0), 32 per block
    std::vector<int> colent;
    for (int khi = -L; khi <= L; khi++)
        for (int cc = 0; cc < 32; cc++) colent.push_back(cc < h.ncol[khi + L] ? h.col[khi + L][cc] : 1);
    if (hipMalloc((void **)&c->d_ozm_col, sizeof(int) * colent.size()) != hipSuccess) return MDFT_ENOMEM;
    if (hipMemcpy(c->d_ozm_col, colent.data(), sizeof(int) * colent.size(), hipMemcpyHostToDevice) != hipSuccess)
        return MDFT_EHIP;
    return MDFT_OK;
}

// B fragments of every interpolation interval (rows iq - 1 and iq of the relayouted c(k)
// table tab[nq][npc], ia0 = first coefficient of each projection row).  Fragment f of block
// khi (kb, nb), lane l: column c = 2 kb + (l >> 5), re/im of V = (l >> 4) & 1, OZ row
// r = 4 nb + ((l & 15) >> 2), q = l & 3 (re lo, im lo, re hi, im hi); the real form of the
// complex product V c: re <- (re V re c - im V im c), im <- (re V im c + im V re c).
int ozm_frags(mdft_ctx *c, const std::vector<double2> &tab, int nq, int npc, const std::vector<int> &ia0) {
    if (!ozm_supported(c)) return MDFT_OK;
    const HostPlan h = host_plan_of(c->L);
    const int L = c->L, nint = nq - 1;
    std::vector<double> fr((size_t)nint * h.nfrag * 64, 0.0);
    for (int ii = 0; ii < nint; ii++)
        for (int khi = -L; khi <= L; khi++) {
            const int kk = khi + L, nkb = h.nkb[kk], nnb = h.nnb[kk];
            int ncols = 0;
            for (int n = std::abs(khi); n <= L; n++) ncols += 2 * (n / 2) + 1;
            for (int kb = 0; kb < nkb; kb++)
                for (int nb = 0; nb < nnb; nb++) {
                    const size_t f = (size_t)ii * h.nfrag + h.foff[kk] + kb * nnb + nb;
                    for (int l = 0; l < 64; l++) {
                        const int cc = 2 * kb + (l >> 5), pk = (l >> 4) & 1, r = 4 * nb + ((l & 15) >> 2), q = l & 3;
                        const int rowent = h.row[kk][r];
                        if (cc >= ncols || rowent < 0) continue;
                        const double2 v = tab[(size_t)(ii + (q >> 1)) * npc + ia0[rowent] + cc];
                        const int qc = q & 1;
                        fr[f * 64 + l] = pk == 0 ? (qc == 0 ? v.x : v.y) : (qc == 0 ? -v.y : v.x);
                    }
                }
        }
    int rc = c->d.precision == MDFT_PREC_FP64 ? upload_t<double>(c, &c->d_ozm_frag, fr)
                                                : upload_t<float>(c, &c->d_ozm_frag, fr);
    if (rc != MDFT_OK || c->d.precision != MDFT_PREC_BF16) return rc;
    // BF16 fragments of v_mfma_f32_16x16x32_bf16: lane l, value j: column 16 kb + 4 (l >> 4) + (j >> 1),
    // re/im of V = j & 1, OZ row 4 nb + ((l & 15) >> 2), q = l & 3
    std::vector<__bf16> f16((size_t)nint * h.nfrag32 * 64 * 8, (__bf16)0.0f);
    for (int ii = 0; ii < nint; ii++)
        for (int khi = -L; khi <= L; khi++) {
            const int kk = khi + L, nkb = h.nkb32[kk], nnb = h.nnb[kk], ncols = h.ncol[kk];
            for (int kb = 0; kb < nkb; kb++)
                for (int nb = 0; nb < nnb; nb++) {
                    const size_t f = (size_t)ii * h.nfrag32 + h.foff32[kk] + kb * nnb + nb;
                    for (int l = 0; l < 64; l++)
                        for (int j = 0; j < 8; j++) {
                            const int cc = 16 * kb + 4 * (l >> 4) + (j >> 1), pk = j & 1;
                            const int r = 4 * nb + ((l & 15) >> 2), q = l & 3;
                            const int rowent = h.row[kk][r];
                            if (cc >= ncols || rowent < 0) continue;
                            const double2 v = tab[(size_t)(ii + (q >> 1)) * npc + ia0[rowent] + cc];
                            const int qc = q & 1;
                            const double x = pk == 0 ? (qc == 0 ? v.x : v.y) : (qc == 0 ? -v.y : v.x);
                            f16[(f * 64 + l) * 8 + j] = (__bf16)(float)x;
                        }
                }
        }
    if (c->d_ozm_frag16) (void)hipFree(c->d_ozm_frag16);
    c->d_ozm_frag16 = nullptr;
    if (hipMalloc(&c->d_ozm_frag16, sizeof(__bf16) * f16.size()) != hipSuccess) return MDFT_ENOMEM;
    if (hipMemcpy(c->d_ozm_frag16, f16.data(), sizeof(__bf16) * f16.size(), hipMemcpyHostToDevice) != hipSuccess)
        return MDFT_EHIP;
    return MDFT_OK;
}

hipError_t launch_ozm_f64(mdft_ctx *c, double2 *proj) { return launch_ozm_any<double, false>(c, proj); }
hipError_t launch_ozm_f32(mdft_ctx *c, float2 *proj) {
    return c->d.precision == MDFT_PREC_BF16 ? launch_ozm_any<float, true>(c, proj) : launch_ozm_any<float, false>(c, proj);
}

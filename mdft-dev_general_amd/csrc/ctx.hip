// ctx.hip -- context set-up and the C-ABI entry points of libmdft_gpu.so.
//
// Host-side tables restate the reference's one-time set-up:
//   angular quadrature        src/module_grid.f90:95-139, 262-301
//   k vectors and -q map      src/module_grid.f90:319-331, 384-418
//   projection index tables   src/module_energy_cproj_mrso.f90:151-175
//   Wigner-d table            src/module_energy_cproj_mrso.f90:224-235, src/module_wigner_d.f90:6-60
//   c(k) relayout             src/module_read_c_luc.f90:101, src/module_energy_cproj_mrso.f90:366-429
// and the evaluation order of energy_and_gradient
// (src/module_energy_and_gradient.f90:36-240).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>

#include "mdft_internal.h"
#include "mdft_wsnet.h"

static std::string g_create_error;

#define HIPCHK(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
            return e_ == hipErrorOutOfMemory ? MDFT_ENOMEM : MDFT_EHIP;                    \
        }                                                                                  \
    } while (0)

namespace {

int fail(mdft_ctx *c, int code, const std::string &msg) {
    c->err = msg;
    return code;
}

// module_grid.f90:262-301
void gauss_legendre(int n, double *x, double *w) {
    const double pi = acos(-1.0);
    const int m = (n + 1) / 2;
    for (int i = 1; i <= m; i++) {
        double xi = cos(pi * (i - 0.25) / (n + 0.5));
        double deltaxi = 1.0, pp = 1.0;
        int guard = 0;
        while (fabs(deltaxi) > DBL_EPSILON && guard++ < 200) {
            double p1 = 1.0, p2 = 0.0, p3;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((double)(2 * j - 1) * xi * p2 - (double)(j - 1) * p3) / (double)j;
            }
            pp = n * (xi * p1 - p2) / (xi * xi - 1.0);
            deltaxi = -p1 / pp;
            xi += deltaxi;
        }
        x[i - 1] = xi;
        w[i - 1] = 1.0 / ((1.0 - xi * xi) * pp * pp);
        x[n - i] = -xi;
        w[n - i] = w[i - 1];
    }
}

double factorial(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
}

// module_wigner_d.f90:6-60
double wigner_small_d(int m, int mu, int mup, double theta) {
    if (std::abs(mu) > m || std::abs(mup) > m) return 0.0;
    if (m == 0) return 1.0;
    if (mu == 0 || mup == 0) {
        int mu0 = mu;
        double th = theta;
        if (mu == 0) {
            mu0 = mup;
            th = -theta;
        }
        double x = 1.0;
        if (mu0 < 0) {
            x = ((-mu0) % 2) ? -1.0 : 1.0;
            mu0 = -mu0;
        }
        const double cc = cos(th);
        double pm1 = 0.0, pm = 1.0, pm2;
        for (int l = mu0 + 1; l <= m; l++) {
            pm2 = pm1;
            pm1 = pm;
            pm = (cc * (double)(2 * l - 1) * pm1 - (double)(l + mu0 - 1) * pm2) / (double)(l - mu0);
        }
        const double sgn = (mu0 % 2) ? -1.0 : 1.0;
        return x * sgn * sqrt(factorial(m - mu0) / factorial(m + mu0)) * factorial(2 * mu0) /
               (pow(2.0, mu0) * factorial(mu0)) * pow(sin(th), mu0) * pm;
    }
    double s = 0.0;
    const double cc = cos(0.5 * theta), ss = sin(0.5 * theta);
    const int itmin = std::max(0, mu - mup), itmax = std::min(m + mu, m - mup);
    for (int it = itmin; it <= itmax; it++)
        s += ((it % 2) ? -1.0 : 1.0) /
             (factorial(m + mu - it) * factorial(m - mup - it) * factorial(it) * factorial(it - mu + mup)) *
             pow(cc, 2 * m + mu - mup - 2 * it) * pow(ss, 2 * it - mu + mup);
    return sqrt(factorial(m + mu) * factorial(m - mu) * factorial(m + mup) * factorial(m - mup)) * s;
}

// module_grid.f90:319-331
double kproj(int l0, int n, double len) {
    const double twopi = 2.0 * acos(-1.0);
    const int l = l0 + 1;
    if (l <= n / 2 + 1) return (double)(l - 1) * twopi / len;
    return (double)(l - 1 - n) * twopi / len;
}

int mq(int i, int n) { return i == 0 ? 0 : n - i; }  // module_grid.f90:384-418

template <typename T>
hipError_t upload(mdft_ctx *c, T **dst, const T *src, size_t n) {
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) return e;
    if (n) e = hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice);
    return e;
}

void free_pairs(PairSchedule &p) {
    (void)hipFree(p.pq);
    (void)hipFree(p.pmq);
    (void)hipFree(p.pflag);
    (void)hipFree(p.cstart);
    (void)hipFree(p.ccount);
    (void)hipFree(p.ciq);
    (void)hipFree(p.tstart);
    (void)hipFree(p.tinfo);
    (void)hipFree(p.tpair);
    p = PairSchedule{};
}

// (q, -q) pairs in the order the reference visits them (cproj_mrso.f90:462-476):
// planes iz = 0..nz/2; in the planes that are their own -q partner the pair is
// handled from its first member in (iy, ix) order.  Pairs are then grouped by
// interpolation row iq (stable) and cut into chunks.
int build_pairs(mdft_ctx *c) {
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    std::vector<int> pq, pmq, pflag, piq;
    pq.reserve(c->N / 2 + 8);
    for (int iz = 0; iz <= nz / 2; iz++) {
        const int izm = mq(iz, nz);
        for (int iy = 0; iy < ny; iy++) {
            const int iym = mq(iy, ny);
            for (int ix = 0; ix < nx; ix++) {
                const int ixm = mq(ix, nx);
                if (izm == iz) {
                    const long self = (long)iy * nx + ix, part = (long)iym * nx + ixm;
                    if (part < self) continue;
                }
                const double q0 = c->kx[ix], q1 = c->ky[iy], q2 = c->kz[iz];
                const double eff = sqrt(q0 * q0 + q1 * q1 + q2 * q2) / c->dq + 1.0;
                const int iq = (int)eff;
                if (iq < 1 || iq + 1 > c->nq) return fail(c, MDFT_ERANGE, "c(k) table too short for |q| of the grid");
                const int g = (iz * ny + iy) * nx + ix, gm = (izm * ny + iym) * nx + ixm;
                int flag = (g == gm) ? 1 : 0;
                if (ix == nx / 2 || iy == ny / 2 || iz == nz / 2) flag |= 2;
                pq.push_back(g);
                pmq.push_back(gm);
                pflag.push_back(flag);
                piq.push_back(iq);
            }
        }
    }
    const int npairs = (int)pq.size();
    std::vector<int> order(npairs);
    for (int i = 0; i < npairs; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return piq[a] < piq[b]; });
    std::vector<int> spq(npairs), spmq(npairs), sflag(npairs), cstart, ccount, ciq;
    const int CH = 256;
    for (int i = 0; i < npairs; i++) {
        const int k = order[i];
        spq[i] = pq[k];
        spmq[i] = pmq[k];
        sflag[i] = pflag[k];
        if (cstart.empty() || ciq.back() != piq[k] || ccount.back() == CH) {
            cstart.push_back(i);
            ccount.push_back(0);
            ciq.push_back(piq[k]);
        }
        ccount.back()++;
    }
    free_pairs(c->pairs);
    PairSchedule &p = c->pairs;
    p.npairs = npairs;
    p.nchunks = (int)cstart.size();
    HIPCHK(c, upload(c, &p.pq, spq.data(), spq.size()));
    HIPCHK(c, upload(c, &p.pmq, spmq.data(), spmq.size()));
    HIPCHK(c, upload(c, &p.pflag, sflag.data(), sflag.size()));
    HIPCHK(c, upload(c, &p.cstart, cstart.data(), cstart.size()));
    HIPCHK(c, upload(c, &p.ccount, ccount.data(), ccount.size()));
    HIPCHK(c, upload(c, &p.ciq, ciq.data(), ciq.size()));
    // tiles of at most 8 pairs of one chunk (one wave each in the matrix-core kernel)
    std::vector<int> tstart, tinfo;
    std::vector<TilePair> tpair;
    for (size_t k = 0; k < cstart.size(); k++)
        for (int o = 0; o < ccount[k]; o += 8) {
            const int n8 = std::min(8, ccount[k] - o);
            tstart.push_back(cstart[k] + o);
            tinfo.push_back(n8 | ciq[k] << 4);
            for (int j = 0; j < 8; j++) {
                // padding entries repeat the tile's first pair (valid memory, never stored)
                const int i = cstart[k] + o + (j < n8 ? j : 0);
                TilePair t{};
                t.gq = spq[i];
                t.gmq = spmq[i];
                t.flag = sflag[i] | (j < n8 ? 0 : 4);
                t.iq = ciq[k];
                const int ix = t.gq % nx, iy = (t.gq / nx) % ny, iz = t.gq / (nx * ny);
                const double q0 = c->kx[ix], q1 = c->ky[iy], q2 = c->kz[iz];
                const double qn = sqrt(q0 * q0 + q1 * q1 + q2 * q2), rxy = sqrt(q0 * q0 + q1 * q1);
                t.al = qn / c->dq + 1.0 - (double)ciq[k];
                // frame of q: theta from z, phi in the xy plane; q along -z: theta = pi, phi = pi;
                // q = 0: the z frame
                t.ct = qn > 0.0 ? q2 / qn : 1.0;
                t.st = qn > 0.0 ? rxy / qn : 0.0;
                t.cp = rxy > 0.0 ? q0 / rxy : (q2 < 0.0 ? -1.0 : 1.0);
                t.sp = rxy > 0.0 ? q1 / rxy : 0.0;
                tpair.push_back(t);
            }
        }
    p.ntiles = (int)tstart.size();
    HIPCHK(c, upload(c, &p.tstart, tstart.data(), tstart.size()));
    HIPCHK(c, upload(c, &p.tinfo, tinfo.data(), tinfo.size()));
    HIPCHK(c, upload(c, &p.tpair, tpair.data(), tpair.size()));
    return MDFT_OK;
}

int ensure_events(mdft_ctx *c) {
    if (c->timers_ready) return MDFT_OK;
    for (auto &e : c->ev) HIPCHK(c, hipEventCreate(&e));
    c->timers_ready = true;
    return MDFT_OK;
}

int ready(mdft_ctx *c) {
    if (!c) return MDFT_EINVAL;
    if (!c->have_vext) return fail(c, MDFT_ESTATE, "vext not set (call mdft_set_vext first)");
    if (!c->have_ck) return fail(c, MDFT_ESTATE, "c(k) not set (call mdft_set_ck first)");
    return MDFT_OK;
}

size_t proj_elem_bytes(const mdft_ctx *c) {
    return c->d.precision == MDFT_PREC_FP64 ? sizeof(double2) : sizeof(float2);
}

// final pass (proj2angl + F + dF) of the context's precision
hipError_t launch_final(mdft_ctx *c, const double *xi, double *df, int mode, int *nb) {
    switch (c->d.precision) {
        case MDFT_PREC_FP32: return launch_final_f32(c, static_cast<const float2 *>(c->d_proj), xi, df, mode, nb);
        case MDFT_PREC_BF16: return launch_final_bf16(c, static_cast<const float2 *>(c->d_proj), xi, df, mode, nb);
        default: return launch_final_f64(c, static_cast<const double2 *>(c->d_proj), xi, df, mode, nb);
    }
}

hipError_t ensure_stage(mdft_ctx *c, double **buf) {
    if (*buf) return hipSuccess;
    return hipMalloc((void **)buf, sizeof(double) * c->N * c->no);
}

// Runs the excess-term projection pipeline up to `stage` (1..4) on device xi.
int run_excess_pipeline(mdft_ctx *c, const double *xi, int stage, bool timers) {
    const bool f32 = c->d.precision != MDFT_PREC_FP64;
    const bool bf16 = c->d.precision == MDFT_PREC_BF16;
    double2 *p64 = static_cast<double2 *>(c->d_proj);
    float2 *p32 = static_cast<float2 *>(c->d_proj);
    if (timers) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, bf16 ? launch_angl2proj_bf16(c, xi, p32)
                   : (f32 ? launch_angl2proj_f32(c, xi, p32) : launch_angl2proj_f64(c, xi, p64)));
    if (timers) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    if (stage >= 2) HIPCHK(c, f32 ? launch_fft3d_f32(c, p32, +1) : launch_fft3d_f64(c, p64, +1));
    if (timers) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    if (stage >= 3) {
        if (ozm_supported(c)) HIPCHK(c, f32 ? launch_ozm_f32(c, p32) : launch_ozm_f64(c, p64));
        else HIPCHK(c, f32 ? launch_oz32(c, p32) : launch_oz_f64(c, p64));
    }
    if (timers) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    if (stage >= 4) HIPCHK(c, f32 ? launch_fft3d_f32(c, p32, -1) : launch_fft3d_f64(c, p64, -1));
    if (timers) HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    return MDFT_OK;
}

int read_sums(mdft_ctx *c, int nblocks, double *out3) {
    HIPCHK(c, launch_reduce3(c, nblocks));
    HIPCHK(c, hipMemcpyAsync(c->h_sums, c->d_sums, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out3[0] = c->h_sums[0];
    out3[1] = c->h_sums[1];
    out3[2] = c->h_sums[2];
    return MDFT_OK;
}

int finish_timers(mdft_ctx *c) {
    float t;
    const int pairs[6][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {5, 6}};
    for (int i = 0; i < 6; i++) {
        HIPCHK(c, hipEventElapsedTime(&t, c->ev[pairs[i][0]], c->ev[pairs[i][1]]));
        c->phase_ms[i] = t;
    }
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[6]));
    c->phase_ms[6] = t;
    c->nphase = 7;
    return MDFT_OK;
}

}  // namespace

extern "C" {

int mdft_ctx_create(const mdft_desc *desc, mdft_ctx **out) {
    if (!desc || !out) {
        g_create_error = "null argument";
        return MDFT_EINVAL;
    }
    *out = nullptr;
    const mdft_desc &d = *desc;
    if (d.nx < 1 || d.ny < 1 || d.nz < 1 || !(d.lx > 0) || !(d.ly > 0) || !(d.lz > 0) || d.mmax < 0 ||
        d.mmax > MDFT_MAX_MMAX || d.mrso < 1 || !(d.kT > 0) || !(d.n0 > 0)) {
        g_create_error = "invalid grid/solvent description";
        return MDFT_EINVAL;
    }
    if (d.precision != MDFT_PREC_FP64 && d.precision != MDFT_PREC_FP32 && d.precision != MDFT_PREC_BF16) {
        g_create_error = "precision must be MDFT_PREC_FP64, MDFT_PREC_FP32 or MDFT_PREC_BF16";
        return MDFT_EINVAL;
    }
    if (d.precision != MDFT_PREC_FP64 && (d.mrso != 2 || d.mmax > 5)) {
        g_create_error = "MDFT_PREC_FP32 and MDFT_PREC_BF16 need molecular symmetry order 2 and mmax <= 5";
        return MDFT_EINVAL;
    }
    if ((long long)d.nx * d.ny * d.nz >= (1LL << 31)) {
        g_create_error = "grid too large (nx*ny*nz must fit in 31 bits)";
        return MDFT_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device";
        return MDFT_ENODEV;
    }
    if (d.device < 0 || d.device >= ndev) {
        g_create_error = "device ordinal out of range";
        return MDFT_EINVAL;
    }
    if (hipSetDevice(d.device) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return MDFT_EHIP;
    }
    mdft_ctx *c = new mdft_ctx();
    c->d = d;
    c->L = d.mmax;
    c->H = d.mmax / d.mrso;
    c->ntheta = d.mmax + 1;
    c->nphi = 2 * d.mmax + 1;
    c->npsi = 2 * (d.mmax / d.mrso) + 1;
    c->no = c->ntheta * c->nphi * c->npsi;
    c->nmu2 = c->H + 1;
    c->N = (size_t)d.nx * d.ny * d.nz;
    c->dv = (d.lx / d.nx) * (d.ly / d.ny) * (d.lz / d.nz);
    const double pi = acos(-1.0);
    c->rho0 = d.n0 / (8.0 * pi * pi / d.mrso);
    int np = 0;
    for (int m = 0; m <= d.mmax; m++) np += (2 * m + 1) * (m / d.mrso + 1);
    c->np = np;
    // quadrature
    c->theta.resize(c->ntheta);
    c->wtheta.resize(c->ntheta);
    gauss_legendre(c->ntheta, c->theta.data(), c->wtheta.data());
    for (auto &t : c->theta) t = acos(t);
    c->w.resize(c->no);
    const double eightpisq = 8.0 * pi * pi;
    for (int it = 0, io = 0; it < c->ntheta; it++)
        for (int iphi = 0; iphi < c->nphi; iphi++)
            for (int ipsi = 0; ipsi < c->npsi; ipsi++, io++)
                c->w[io] = c->wtheta[it] * (1.0 / c->nphi) * (1.0 / (double)(c->npsi * d.mrso)) * eightpisq;
    // projection tables
    const int W = 2 * c->L + 1;
    c->pidx.assign((size_t)(c->L + 1) * W * c->nmu2, -1);
    for (int m = 0, ip = 0; m <= c->L; m++)
        for (int mup = -m; mup <= m; mup++)
            for (int mu = 0; mu <= m; mu += d.mrso, ip++) {
                c->pidx[(m * W + mup + c->L) * c->nmu2 + mu / d.mrso] = ip;
                c->pm.push_back(m);
                c->pmup.push_back(mup);
                c->pmu.push_back(mu);
            }
    c->wig.resize((size_t)np * c->ntheta);
    std::vector<double> a2p(np * c->ntheta), p2a(np * c->ntheta), fmv(np);
    for (int ip = 0; ip < np; ip++) {
        const int m = c->pm[ip];
        const double fm = (m == 7) ? sqrt(14.0) : sqrt((double)(2 * m + 1));
        fmv[ip] = fm;
        for (int it = 0; it < c->ntheta; it++) {
            const double dd = wigner_small_d(m, c->pmup[ip], c->pmu[ip], c->theta[it]);
            c->wig[ip * c->ntheta + it] = dd;
            a2p[ip * c->ntheta + it] = dd * c->wtheta[it];
            p2a[ip * c->ntheta + it] = dd * fm;
        }
    }
    std::vector<double> cph(c->nphi), sph(c->nphi), cps(c->npsi), sps(c->npsi);
    for (int k = 0; k < c->nphi; k++) {
        cph[k] = cos(2.0 * pi * k / c->nphi);
        sph[k] = sin(2.0 * pi * k / c->nphi);
    }
    for (int k = 0; k < c->npsi; k++) {
        cps[k] = cos(2.0 * pi * k / c->npsi);
        sps[k] = sin(2.0 * pi * k / c->npsi);
    }
    // OZ coefficient offsets per projection row (cproj_mrso.f90:381 ordering)
    std::vector<int> ia0(np);
    for (int ip = 0, ia = 0; ip < np; ip++) {
        ia0[ip] = ia;
        for (int n = std::abs(c->pmup[ip]); n <= c->L; n++) ia += 2 * (n / d.mrso) + 1;
    }
    c->kx.resize(d.nx);
    c->ky.resize(d.ny);
    c->kz.resize(d.nz);
    for (int l = 0; l < d.nx; l++) c->kx[l] = kproj(l, d.nx, d.lx);
    for (int l = 0; l < d.ny; l++) c->ky[l] = kproj(l, d.ny, d.ly);
    for (int l = 0; l < d.nz; l++) c->kz[l] = kproj(l, d.nz, d.lz);

    auto bad = [&](int code, const std::string &msg) {
        g_create_error = msg;
        mdft_ctx_destroy(c);
        return code;
    };
    const int ns[3] = {d.nx, d.ny, d.nz};
    for (int ax = 0; ax < 3; ax++) {
        FftAxisPlan &p = c->fft[ax];
        p.n = ns[ax];
        p.nst = fft_make_plan(ns[ax], p.radix);
        if (p.nst < 0) return bad(MDFT_EINVAL, "grid sizes must factor into 2, 3 and 5");
        std::vector<double2> tw(ns[ax]);
        for (int k = 0; k < ns[ax]; k++) tw[k] = make_double2(cos(2.0 * pi * k / ns[ax]), sin(2.0 * pi * k / ns[ax]));
        if (upload(c, &p.tw, tw.data(), tw.size()) != hipSuccess) return bad(MDFT_ENOMEM, "device allocation failed");
    }
    {
        int lim = 0;
        if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, d.device) == hipSuccess && lim > 0)
            c->lds_limit = (size_t)lim;
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d.device) == hipSuccess && ncu > 0)
            c->num_cus = ncu;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(MDFT_EHIP, "hipStreamCreate failed");
    if (vec_scratch_init(d.device) != hipSuccess)
        return bad(MDFT_EHIP, "device address of the no-allocation scratch array (hipGetSymbolAddress) failed");
    bool ok = upload(c, &c->d_w, c->w.data(), c->w.size()) == hipSuccess &&
              upload(c, &c->d_a2p, a2p.data(), a2p.size()) == hipSuccess &&
              upload(c, &c->d_p2a, p2a.data(), p2a.size()) == hipSuccess &&
              upload(c, &c->d_fm, fmv.data(), fmv.size()) == hipSuccess &&
              upload(c, &c->d_cph, cph.data(), cph.size()) == hipSuccess &&
              upload(c, &c->d_sph, sph.data(), sph.size()) == hipSuccess &&
              upload(c, &c->d_cps, cps.data(), cps.size()) == hipSuccess &&
              upload(c, &c->d_sps, sps.data(), sps.size()) == hipSuccess &&
              upload(c, &c->d_pidx, c->pidx.data(), c->pidx.size()) == hipSuccess &&
              upload(c, &c->d_pm, c->pm.data(), c->pm.size()) == hipSuccess &&
              upload(c, &c->d_pmup, c->pmup.data(), c->pmup.size()) == hipSuccess &&
              upload(c, &c->d_pmu, c->pmu.data(), c->pmu.size()) == hipSuccess &&
              upload(c, &c->d_ia0, ia0.data(), ia0.size()) == hipSuccess &&
              upload(c, &c->d_kx, c->kx.data(), c->kx.size()) == hipSuccess &&
              upload(c, &c->d_ky, c->ky.data(), c->ky.size()) == hipSuccess &&
              upload(c, &c->d_kz, c->kz.data(), c->kz.size()) == hipSuccess;
    c->ia0_h = ia0;
    ok = ok && hipMalloc(&c->d_proj, proj_elem_bytes(c) * c->N * np) == hipSuccess;
    ok = ok && hipMalloc(&c->d_vext, sizeof(double) * c->N * c->no) == hipSuccess;
    ok = ok && hipMalloc(&c->d_sums, sizeof(double) * 8) == hipSuccess;
    ok = ok && hipHostMalloc(&c->h_sums, sizeof(double) * 8, hipHostMallocDefault) == hipSuccess;
    if (!ok) return bad(MDFT_ENOMEM, "device allocation failed");
    if (d.precision == MDFT_PREC_BF16) {
        const int rc = bf16_tables(c, a2p, p2a, fmv);
        if (rc != MDFT_OK) return bad(rc, "BF16 transform tables");
    }
    if (d.precision != MDFT_PREC_FP64) {
        std::vector<float> ftab;
        std::vector<int> lanerow;
        if (oz32_tables(c, ftab, lanerow) != MDFT_OK) return bad(MDFT_EINVAL, "too many projections per lane");
        if (upload(c, &c->d_oz32_ftab, ftab.data(), ftab.size()) != hipSuccess ||
            upload(c, &c->d_oz32_lanerow, lanerow.data(), lanerow.size()) != hipSuccess)
            return bad(MDFT_ENOMEM, "device allocation failed");
    }
    if (ozm_setup(c) != MDFT_OK) return bad(MDFT_ENOMEM, "matrix-core q-space tables");
    AngTables &t = c->ang;
    t.ntheta = c->ntheta;
    t.nphi = c->nphi;
    t.npsi = c->npsi;
    t.no = c->no;
    t.np = np;
    t.L = c->L;
    t.H = c->H;
    t.mrso = d.mrso;
    t.w = c->d_w;
    t.a2p = c->d_a2p;
    t.p2a = c->d_p2a;
    t.fm = c->d_fm;
    t.cph = c->d_cph;
    t.sph = c->d_sph;
    t.cps = c->d_cps;
    t.sps = c->d_sps;
    t.pidx = c->d_pidx;
    t.pm = c->d_pm;
    t.pmup = c->d_pmup;
    t.pmu = c->d_pmu;
    *out = c;
    return MDFT_OK;
}

void mdft_ctx_destroy(mdft_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &f : c->fft) (void)hipFree(f.tw);
    void *bufs[] = {c->d_w,   c->d_a2p,  c->d_p2a,  c->d_fm,   c->d_cph,      c->d_sph,        c->d_cps,
                    c->d_sps, c->d_pidx, c->d_pm,   c->d_pmup, c->d_pmu,      c->d_ia0,        c->d_kx,
                    c->d_ky,  c->d_kz,   c->d_ck,   c->d_vext, c->d_proj,     c->d_xi_stage,   c->d_df_stage,
                    c->d_partials, c->d_slices, c->d_sums, c->d_vpart, c->d_mpart, c->d_oz32_ftab, c->d_oz32_lanerow,
                    c->d_ozm_frag, c->d_ozm_frag16, c->d_ozm_col, c->d_ozm_ftab, c->d_ozm_units, c->d_ozm_rlen,
                    c->d_bf16_M, c->d_bf16_Q, c->d_chscale, c->d_adjscale, c->d_vscale, c->d_dpart, c->d_vext32,
                    c->d_a2p_img[0], c->d_a2p_img[1]};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    free_pairs(c->pairs);
    if (c->h_sums) (void)hipHostFree(c->h_sums);
    if (c->timers_ready)
        for (auto &e : c->ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *mdft_last_error(const mdft_ctx *c) {
    if (!c) return g_create_error.c_str();
    return c->err.c_str();
}

int mdft_get_info(const mdft_ctx *c, int *info) {
    if (!c || !info) return MDFT_EINVAL;
    info[0] = c->ntheta;
    info[1] = c->nphi;
    info[2] = c->npsi;
    info[3] = c->no;
    info[4] = c->np;
    info[5] = c->npc;
    info[6] = c->nq;
    info[7] = (int)(c->N & 0x7fffffff);
    info[8] = c->d.precision;
    info[9] = c->d.mrso;
    return MDFT_OK;
}

int mdft_get_quadrature(const mdft_ctx *c, double *theta, double *wtheta, double *w) {
    if (!c) return MDFT_EINVAL;
    if (theta) memcpy(theta, c->theta.data(), sizeof(double) * c->ntheta);
    if (wtheta) memcpy(wtheta, c->wtheta.data(), sizeof(double) * c->ntheta);
    if (w) memcpy(w, c->w.data(), sizeof(double) * c->no);
    return MDFT_OK;
}

int mdft_set_vext(mdft_ctx *c, const double *vext, int flags) {
    if (!c || !vext) return MDFT_EINVAL;
    HIPCHK(c, hipMemcpyAsync(c->d_vext, vext, sizeof(double) * c->N * c->no,
                             (flags & MDFT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_vext = true;
    c->vext32_ok = false;
    return MDFT_OK;
}

int mdft_set_ck(mdft_ctx *c, int np_file, const int *m, const int *n, const int *mu, const int *nu, const int *khi,
                int nq_file, double dq, const double *cf) {
    if (!c || np_file < 1 || !m || !n || !mu || !nu || !khi || nq_file < 2 || !(dq > 0) || !cf)
        return MDFT_EINVAL;
    const int L = c->L, mrso = c->d.mrso;
    // q_max of cproj_mrso.f90:366 and nq of read_c_luc.f90:101
    double kxm = c->kx[0], kym = c->ky[0], kzm = c->kz[0];
    for (double v : c->kx) kxm = std::max(kxm, v);
    for (double v : c->ky) kym = std::max(kym, v);
    for (int i = 0; i <= c->d.nz / 2; i++) kzm = std::max(kzm, c->kz[i]);
    const double qmax = sqrt(kxm * kxm + kym * kym + kzm * kzm);
    const int nq = (int)(qmax / dq + 0.5) + 2;
    if (nq > nq_file) return fail(c, MDFT_ERANGE, "c(k) file has fewer |q| rows than the grid needs");
    std::map<std::tuple<int, int, int, int, int>, int> pos;
    for (int i = 0; i < np_file; i++) {
        if (mu[i] % mrso || nu[i] % mrso) continue;
        pos[std::make_tuple(m[i], n[i], mu[i] / mrso, nu[i] / mrso, khi[i])] = i;
    }
    std::vector<int> src;
    std::vector<double> sgn;
    for (int mm = 0; mm <= L; mm++)
        for (int k = -mm; k <= mm; k++)
            for (int mu2 = 0; mu2 <= mm / mrso; mu2++)
                for (int nn = std::abs(k); nn <= L; nn++)
                    for (int nu2 = -nn / mrso; nu2 <= nn / mrso; nu2++) {
                        auto it = pos.find(std::make_tuple(mm, nn, mu2, nu2, k));
                        if (it == pos.end()) {
                            char buf[160];
                            snprintf(buf, sizeof buf, "c(k) table lacks (m,n,mu,nu,khi)=(%d,%d,%d,%d,%d)", mm, nn,
                                     mu2 * mrso, nu2 * mrso, k);
                            return fail(c, MDFT_EINVAL, buf);
                        }
                        src.push_back(it->second);
                        const int nuv = nu2 * mrso;
                        sgn.push_back(nuv < 0 ? ((std::abs(k + nuv) % 2) ? -1.0 : 1.0) : ((nn % 2) ? -1.0 : 1.0));
                    }
    const int npc = (int)src.size();
    std::vector<double2> tab((size_t)npc * nq);
    for (int iq = 0; iq < nq; iq++)
        for (int ia = 0; ia < npc; ia++) {
            const double *v = cf + 2 * ((size_t)iq * np_file + src[ia]);
            // conj (cproj_mrso.f90:415) then the sign folding (:420-429)
            tab[(size_t)iq * npc + ia] = make_double2(sgn[ia] * v[0], sgn[ia] * -v[1]);
        }
    if (c->d_ck) (void)hipFree(c->d_ck);
    c->d_ck = nullptr;
    HIPCHK(c, upload(c, &c->d_ck, tab.data(), tab.size()));
    {
        const int frc = ozm_frags(c, tab, nq, npc, c->ia0_h);
        if (frc != MDFT_OK) return fail(c, frc, "c(k) fragments of the matrix-core q-space kernel");
    }
    c->npc = npc;
    c->nq = nq;
    c->dq = dq;
    const int rc = build_pairs(c);
    if (rc != MDFT_OK) return rc;
    c->have_ck = true;
    return MDFT_OK;
}

int mdft_excess_stage(mdft_ctx *c, const double *xi, int stage, double *out, int flags) {
    int rc = ready(c);
    if (rc) return rc;
    if (stage < 1 || stage > 4 || !xi || !out) return fail(c, MDFT_EINVAL, "bad stage arguments");
    const bool dev = flags & MDFT_DEVICE_PTRS;
    const double *xd = xi;
    if (!dev) {
        HIPCHK(c, ensure_stage(c, &c->d_xi_stage));
        HIPCHK(c, hipMemcpyAsync(c->d_xi_stage, xi, sizeof(double) * c->N * c->no, hipMemcpyHostToDevice, c->stream));
        xd = c->d_xi_stage;
    }
    rc = run_excess_pipeline(c, xd, stage, false);
    if (rc) return rc;
    const size_t ne = c->N * c->np;
    if (c->d.precision == MDFT_PREC_FP64) {
        HIPCHK(c, hipMemcpyAsync(out, c->d_proj, sizeof(double2) * ne,
                                 dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else if (dev) {
        HIPCHK(c, launch_widen_f32(c, static_cast<const float2 *>(c->d_proj), reinterpret_cast<double2 *>(out), ne));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<float2> tmp(ne);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), c->d_proj, sizeof(float2) * ne, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < ne; i++) {
            out[2 * i] = tmp[i].x;
            out[2 * i + 1] = tmp[i].y;
        }
    }
    return MDFT_OK;
}

// mode 0: full energy_and_gradient, 1: excess only
// s(io) = sqrt(w(io) / mean w) and 1 / s(io), on the device (scaled variables u = s xi)
static int ensure_vscale(mdft_ctx *c) {
    if (c->d_vscale) return MDFT_OK;
    const int no = c->no;
    double wm = 0.0;
    for (int io = 0; io < no; io++) wm += c->w[io];
    wm /= no;
    std::vector<double> h(2 * (size_t)no);
    for (int io = 0; io < no; io++) {
        h[io] = sqrt(c->w[io] / wm);
        h[no + io] = 1.0 / h[io];
    }
    HIPCHK(c, hipMalloc(&c->d_vscale, sizeof(double) * h.size()));
    HIPCHK(c, hipMemcpy(c->d_vscale, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return MDFT_OK;
}

static int evaluate(mdft_ctx *c, const double *xi, double *df, int flags, int mode, double *sums, int dot_k = 0,
                    const void *const *dot_vecs = nullptr, double *dots = nullptr) {
    int rc = ready(c);
    if (rc) return rc;
    if (!xi) return fail(c, MDFT_EINVAL, "xi is null");
    const bool dev = flags & MDFT_DEVICE_PTRS;
    const bool timers = flags & MDFT_TIMERS;
    const bool df32 = flags & MDFT_DF_FP32;
    if (df32 && (!dev || mode != 0 || !df || c->d.mrso != 2 || c->d.precision == MDFT_PREC_BF16))
        return fail(c, MDFT_EINVAL, "MDFT_DF_FP32 needs device pointers, a gradient, molecular symmetry order 2 "
                                    "and an FP64 or FP32 context");
    const bool scaled = flags & MDFT_SCALED_VARS;
    if (scaled && (mode != 0 || c->d.mrso != 2 || c->d.precision == MDFT_PREC_BF16))
        return fail(c, MDFT_EINVAL, "MDFT_SCALED_VARS needs mdft_energy_and_gradient, molecular symmetry order 2 "
                                    "and an FP64 or FP32 context");
    const bool x32 = flags & MDFT_XI_FP32;
    if (x32 && (!dev || mode != 0 || c->d.mrso != 2 || c->d.precision == MDFT_PREC_BF16))
        return fail(c, MDFT_EINVAL, "MDFT_XI_FP32 needs device pointers, mdft_energy_and_gradient, molecular "
                                    "symmetry order 2 and an FP64 or FP32 context");
    // the FP32 gradient and iterate, the scaled variables and the fused products exist only in
    // the transforms specialised for mmax <= 5; the generic kernels of larger orders have none
    if ((df32 || scaled || x32 || dot_k > 0) && c->L > 5)
        return fail(c, MDFT_EINVAL,
                    "MDFT_DF_FP32, MDFT_XI_FP32, MDFT_SCALED_VARS and gradient products need mmax <= 5");
    if (scaled && (rc = ensure_vscale(c))) return rc;
    if (dot_k > 0) {
        if (dot_k > 3 || !dot_vecs || !dots || !dev || !df || mode != 0 || c->d.mrso != 2 ||
            c->d.precision == MDFT_PREC_BF16)
            return fail(c, MDFT_EINVAL, "gradient products need 1-3 vectors, device pointers, a gradient, "
                                        "molecular symmetry order 2 and an FP64 or FP32 context");
        const size_t need = 3 * ((c->N + 3) / 4);
        if (c->dpart_cap < need) {
            if (c->d_dpart) HIPCHK(c, hipFree(c->d_dpart));
            c->d_dpart = nullptr;
            c->dpart_cap = 0;
            HIPCHK(c, hipMalloc(&c->d_dpart, sizeof(double) * need));
            c->dpart_cap = need;
        }
        c->dot_self = 0;
        for (int j = 0; j < dot_k; j++) {
            if (!dot_vecs[j]) return fail(c, MDFT_EINVAL, "null vector");
            c->dot_vec[j] = dot_vecs[j];
            if (dot_vecs[j] == static_cast<const void *>(df)) c->dot_self |= 1 << j;
        }
    }
    if (timers && (rc = ensure_events(c))) return rc;
    const size_t nbytes = sizeof(double) * c->N * c->no;
    const double *xd = xi;
    double *dd = df32 ? nullptr : df;
    if (!dev) {
        HIPCHK(c, ensure_stage(c, &c->d_xi_stage));
        HIPCHK(c, hipMemcpyAsync(c->d_xi_stage, xi, nbytes, hipMemcpyHostToDevice, c->stream));
        xd = c->d_xi_stage;
        if (df) {
            HIPCHK(c, ensure_stage(c, &c->d_df_stage));
            if (mode == 1)
                HIPCHK(c, hipMemcpyAsync(c->d_df_stage, df, nbytes, hipMemcpyHostToDevice, c->stream));
            dd = c->d_df_stage;
        }
    }
    if (x32 && df32 && dot_k > 0 && c->d.precision == MDFT_PREC_FP32 && !c->vext32_ok) {
        // the fast minimiser's final pass reads vext in FP32 (its FP32 terms round vext to FP32
        // first, so the copy changes no result): made once per vext
        if (!c->d_vext32) HIPCHK(c, hipMalloc(&c->d_vext32, sizeof(float) * c->N * c->no));
        HIPCHK(c, vec_narrow(c, c->N * c->no, c->d_vext, c->d_vext32));
        c->vext32_ok = true;
    }
    c->vsinv = scaled ? c->d_vscale + c->no : nullptr;
    c->xi32 = x32 ? reinterpret_cast<const float *>(xi) : nullptr;
    rc = run_excess_pipeline(c, xd, 4, timers);
    if (rc) {
        c->vsinv = nullptr;
        c->xi32 = nullptr;
        return rc;
    }
    int nb = 0;
    c->df32 = df32 ? reinterpret_cast<float *>(df) : nullptr;
    // the FP32 final pass forms the products; an FP64 context runs the multi-dot after it
    const bool fused_dots = dot_k > 0 && c->d.precision == MDFT_PREC_FP32;
    c->dot_k = fused_dots ? dot_k : 0;
    const hipError_t fe = launch_final(c, xd, dd, mode, &nb);
    c->df32 = nullptr;
    c->vsinv = nullptr;
    c->xi32 = nullptr;
    c->dot_k = 0;
    HIPCHK(c, fe);
    if (fused_dots) {
        HIPCHK(c, launch_reduce3_buf(c, c->d_dpart, nb, c->d_sums + 5));
        HIPCHK(c, hipMemcpyAsync(c->h_sums + 5, c->d_sums + 5, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    }
    if (timers) HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
    rc = read_sums(c, nb, sums);
    if (rc) return rc;
    if (timers) {
        HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev[6]));
        if ((rc = finish_timers(c))) return rc;
    }
    if (!dev && df) {
        HIPCHK(c, hipMemcpyAsync(df, dd, nbytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    sums[2] *= 0.5 * c->dv;  // ff = ff * 0.5 * dv (cproj_mrso.f90:745)
    if (fused_dots) {
        for (int j = 0; j < dot_k; j++) dots[j] = c->h_sums[5 + j];  // read_sums synchronised the stream
    } else if (dot_k > 0) {
        const size_t n = c->N * (size_t)c->no;
        if (df32)
            HIPCHK(c, vec_multi_dot32(c, n, dot_k, reinterpret_cast<const float *const *>(dot_vecs),
                                      reinterpret_cast<const float *>(df), dots));
        else
            HIPCHK(c, vec_multi_dot(c, n, dot_k, reinterpret_cast<const double *const *>(dot_vecs), df, dots));
    }
    return MDFT_OK;
}

int mdft_energy_and_gradient(mdft_ctx *c, const double *xi, double *f_parts, double *df, int flags) {
    if (!f_parts) return c ? fail(c, MDFT_EINVAL, "f_parts is null") : MDFT_EINVAL;
    double s[3];
    const int rc = evaluate(c, xi, df, flags, 0, s);
    if (rc) return rc;
    f_parts[0] = s[0];
    f_parts[1] = s[1];
    f_parts[2] = s[2];
    return MDFT_OK;
}

int mdft_energy_and_gradient_dots(mdft_ctx *c, const double *xi, double *f_parts, void *df, int flags, int k,
                                  const void *const *vecs, double *dots) {
    if (!f_parts || k < 1) return c ? fail(c, MDFT_EINVAL, "f_parts is null or no vectors") : MDFT_EINVAL;
    double s[3];
    const int rc = evaluate(c, xi, static_cast<double *>(df), flags, 0, s, k, vecs, dots);
    if (rc) return rc;
    f_parts[0] = s[0];
    f_parts[1] = s[1];
    f_parts[2] = s[2];
    return MDFT_OK;
}

int mdft_excess_energy(mdft_ctx *c, const double *xi, double *fexc, double *df, int flags) {
    if (!fexc) return c ? fail(c, MDFT_EINVAL, "fexc is null") : MDFT_EINVAL;
    double s[3];
    const int rc = evaluate(c, xi, df, flags, 1, s);
    if (rc) return rc;
    *fexc = s[2];
    return MDFT_OK;
}

int mdft_ideal_external(mdft_ctx *c, const double *xi, double *fid, double *fext, double *df, int flags) {
    if (!c) return MDFT_EINVAL;
    if (!c->have_vext) return fail(c, MDFT_ESTATE, "vext not set (call mdft_set_vext first)");
    if (!xi || !fid || !fext) return fail(c, MDFT_EINVAL, "null argument");
    const bool dev = flags & MDFT_DEVICE_PTRS;
    const size_t nbytes = sizeof(double) * c->N * c->no;
    const double *xd = xi;
    double *dd = df;
    if (!dev) {
        HIPCHK(c, ensure_stage(c, &c->d_xi_stage));
        HIPCHK(c, hipMemcpyAsync(c->d_xi_stage, xi, nbytes, hipMemcpyHostToDevice, c->stream));
        xd = c->d_xi_stage;
        if (df) {
            HIPCHK(c, ensure_stage(c, &c->d_df_stage));
            HIPCHK(c, hipMemcpyAsync(c->d_df_stage, df, nbytes, hipMemcpyHostToDevice, c->stream));
            dd = c->d_df_stage;
        }
    }
    int nb = 0;
    HIPCHK(c, launch_ideal_f64(c, xd, dd, &nb));
    double s[3];
    const int rc = read_sums(c, nb, s);
    if (rc) return rc;
    if (!dev && df) {
        HIPCHK(c, hipMemcpyAsync(df, dd, nbytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    *fid = s[0];
    *fext = s[1];
    return MDFT_OK;
}

int mdft_compute_vext(mdft_ctx *c, int natom, const double *atoms, int nsite, const double *sites, double rcut) {
    return mdft_compute_vext_ex(c, natom, atoms, nsite, sites, rcut, MDFT_ELEC_DIRECT);
}

int mdft_compute_vext_ex(mdft_ctx *c, int natom, const double *atoms, int nsite, const double *sites, double rcut,
                         int electrostatics) {
    if (!c || natom < 0 || nsite < 1 || (natom > 0 && !atoms) || !sites || !(rcut > 0))
        return c ? fail(c, MDFT_EINVAL, "bad vext arguments") : MDFT_EINVAL;
    if (electrostatics != MDFT_ELEC_POISSON && electrostatics != MDFT_ELEC_DIRECT && electrostatics != MDFT_ELEC_NONE)
        return fail(c, MDFT_EINVAL, "unknown electrostatics");
    // physical constants of src/module_constants.f90:54-57
    const double pi = acos(-1.0);
    const double eps0 = 8.854187817e-12, qunit = 1.602176565e-19, navo = 6.02214129e23;
    const double qfact = qunit * qunit * 1.0e-3 * navo / (4.0 * pi * eps0 * 1.0e-10);
    // site offsets per orientation: rotation matrix of src/module_grid.f90:187-206
    std::vector<double> off((size_t)c->no * nsite * 3), sp((size_t)nsite * 3), at((size_t)(natom > 0 ? natom : 1) * 6);
    for (int it = 0, io = 0; it < c->ntheta; it++)
        for (int iphi = 0; iphi < c->nphi; iphi++)
            for (int ipsi = 0; ipsi < c->npsi; ipsi++, io++) {
                const double th = c->theta[it], ph = iphi * 2.0 * pi / c->nphi;
                const double ps = ipsi * 2.0 * pi / c->d.mrso / c->npsi;
                const double ct = cos(th), st = sin(th), cp = cos(ph), sph = sin(ph), cs = cos(ps), ss = sin(ps);
                const double R[3][3] = {{ct * cp * cs - sph * ss, -ct * cp * ss - sph * cs, st * cp},
                                        {ct * sph * cs + cp * ss, -ct * sph * ss + cp * cs, st * sph},
                                        {-st * cs, st * ss, ct}};
                for (int s = 0; s < nsite; s++)
                    for (int k = 0; k < 3; k++)
                        off[((size_t)io * nsite + s) * 3 + k] =
                            R[k][0] * sites[6 * s] + R[k][1] * sites[6 * s + 1] + R[k][2] * sites[6 * s + 2];
            }
    for (int s = 0; s < nsite; s++) {
        sp[3 * s] = sites[6 * s + 3];
        sp[3 * s + 1] = sites[6 * s + 4];
        sp[3 * s + 2] = sites[6 * s + 5];
    }
    for (int i = 0; i < natom * 6; i++) at[i] = atoms[i];
    double *d_off = nullptr, *d_sp = nullptr, *d_at = nullptr;
    HIPCHK(c, upload(c, &d_off, off.data(), off.size()));
    HIPCHK(c, upload(c, &d_sp, sp.data(), sp.size()));
    HIPCHK(c, upload(c, &d_at, at.data(), at.size()));
    // the clip at qfact/2 (src/module_vext.f90:638-640) follows the sum of all terms
    const bool direct = electrostatics == MDFT_ELEC_DIRECT, poisson = electrostatics == MDFT_ELEC_POISSON;
    hipError_t e = launch_vext(c, natom, d_at, nsite, d_sp, d_off, rcut, qfact, poisson ? HUGE_VAL : qfact / 2.0, direct);
    if (e == hipSuccess && poisson) {
        std::vector<double> sq(nsite);
        for (int s = 0; s < nsite; s++) sq[s] = sites[6 * s + 3];
        e = launch_poisson_vext(c, natom, at.data(), nsite, sq.data(), d_off, qfact);
        if (e == hipSuccess) e = launch_vext_clip(c, qfact / 2.0);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_off);
    (void)hipFree(d_sp);
    (void)hipFree(d_at);
    HIPCHK(c, e);
    c->have_vext = true;
    c->vext32_ok = false;
    return MDFT_OK;
}

int mdft_get_vext(mdft_ctx *c, double *out, int flags) {
    if (!c || !out) return MDFT_EINVAL;
    HIPCHK(c, hipMemcpyAsync(out, c->d_vext, sizeof(double) * c->N * c->no,
                             (flags & MDFT_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

int mdft_vec_dot(mdft_ctx *c, size_t n, const double *x, const double *y, double *out) {
    if (!c || !x || !y || !out) return MDFT_EINVAL;
    if (!aligned16(x) || !aligned16(y)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    HIPCHK(c, vec_dot(c, n, x, y, out));
    return MDFT_OK;
}

int mdft_vec_axpy_dot(mdft_ctx *c, size_t n, double a, const double *x, double *y, const double *z, double *out) {
    if (!c || !x || !y) return MDFT_EINVAL;
    if (!aligned16(x) || !aligned16(y) || (z && !aligned16(z)))
        return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    HIPCHK(c, vec_axpy_dot(c, n, a, x, y, z, out));
    return MDFT_OK;
}

int mdft_vec_lincomb(mdft_ctx *c, size_t n, double a, const double *x, double b, const double *y, double *z) {
    if (!c || !x || !z) return MDFT_EINVAL;
    if (!aligned16(x) || !aligned16(z) || (y && !aligned16(y)))
        return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    HIPCHK(c, vec_lincomb(c, n, a, x, b, y, z));
    return MDFT_OK;
}

int mdft_vec_multi_dot(mdft_ctx *c, size_t n, int k, const double *const *vecs, const double *x, double *out) {
    if (!c || !vecs || !x || !out || k < 1 || k > MDFT_VEC_MAX) return MDFT_EINVAL;
    if (!aligned16(x)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    for (int j = 0; j < k; j++)
        if (!vecs[j] || !aligned16(vecs[j])) return fail(c, MDFT_EINVAL, "null or unaligned vector");
    HIPCHK(c, vec_multi_dot(c, n, k, vecs, x, out));
    return MDFT_OK;
}

int mdft_vec_multi_dot_noalloc(mdft_ctx *c, size_t n, int k, const double *const *vecs, const double *x,
                               double *out) {
    if (!c || !vecs || !x || !out) return MDFT_EINVAL;
    if (k < 1 || k > MDFT_VEC_MAX)
        return fail(c, MDFT_EINVAL, "mdft_vec_multi_dot_noalloc: k = " + std::to_string(k) +
                                        " vectors; its fixed scratch array holds the partial products of 1.." +
                                        std::to_string(MDFT_VEC_MAX) + " vectors");
    if (!aligned16(x)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    for (int j = 0; j < k; j++)
        if (!vecs[j] || !aligned16(vecs[j])) return fail(c, MDFT_EINVAL, "null or unaligned vector");
    HIPCHK(c, vec_multi_dot_noalloc(c, n, k, vecs, x, out));
    return MDFT_OK;
}

int mdft_vec_combine(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z) {
    if (!c || !coef || !vecs || !z || k < 1 || k > MDFT_VEC_MAX) return MDFT_EINVAL;
    if (!aligned16(z)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    for (int j = 0; j < k; j++)
        if (!vecs[j] || !aligned16(vecs[j])) return fail(c, MDFT_EINVAL, "null or unaligned vector");
    HIPCHK(c, vec_combine(c, n, k, coef, vecs, z));
    return MDFT_OK;
}

int mdft_vec_combine_step(mdft_ctx *c, size_t n, int k, const double *coef, const double *const *vecs, double *z,
                          double a, const double *x, double *y) {
    if (!c || !coef || !vecs || !z || !x || !y || k < 1 || k > MDFT_VEC_MAX) return MDFT_EINVAL;
    if (!aligned16(z) || !aligned16(x) || !aligned16(y)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    for (int j = 0; j < k; j++) {
        if (!vecs[j] || !aligned16(vecs[j])) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
        if (vecs[j] == y) return fail(c, MDFT_EINVAL, "y must not alias an input vector");
    }
    if (y == x || y == z) return fail(c, MDFT_EINVAL, "y must not alias x or z");
    HIPCHK(c, vec_combine_step(c, n, k, coef, vecs, z, a, x, y));
    return MDFT_OK;
}

int mdft_vec_multi_dot_f32(mdft_ctx *c, size_t n, int k, const float *const *vecs, const float *x, double *out) {
    if (!c || !vecs || !x || !out || k < 1 || k > MDFT_VEC_MAX) return MDFT_EINVAL;
    if (!aligned16(x)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    for (int j = 0; j < k; j++)
        if (!vecs[j] || !aligned16(vecs[j])) return fail(c, MDFT_EINVAL, "null or unaligned vector");
    HIPCHK(c, vec_multi_dot32(c, n, k, vecs, x, out));
    return MDFT_OK;
}

int mdft_vec_combine_step_f32(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs, float *z,
                              double a, const double *x, double *y) {
    if (!c || !coef || !vecs || !z || !x || !y || k < 1 || k > MDFT_VEC_MAX) return MDFT_EINVAL;
    if (!aligned16(z) || !aligned16(x) || !aligned16(y)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    for (int j = 0; j < k; j++)
        if (!vecs[j] || !aligned16(vecs[j])) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    if (y == x) return fail(c, MDFT_EINVAL, "y must not alias x");
    HIPCHK(c, vec_combine_step32(c, n, k, coef, vecs, z, a, x, y));
    return MDFT_OK;
}

int mdft_vec_axpy_f32(mdft_ctx *c, size_t n, double a, const float *z, const double *x, double *y) {
    if (!c || !z || !x || !y) return MDFT_EINVAL;
    if (!aligned16(z) || !aligned16(x) || !aligned16(y)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    HIPCHK(c, vec_axpy32(c, n, a, z, x, y));
    return MDFT_OK;
}

int mdft_vec_combine_step_f32x(mdft_ctx *c, size_t n, int k, const double *coef, const float *const *vecs,
                               float *z, double a, const float *x, float *y) {
    if (!c || !coef || !vecs || !z || !x || !y || k < 1 || k > MDFT_VEC_MAX) return MDFT_EINVAL;
    if (!aligned16(z) || !aligned16(x) || !aligned16(y)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    for (int j = 0; j < k; j++)
        if (!vecs[j] || !aligned16(vecs[j])) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    if (y == x) return fail(c, MDFT_EINVAL, "y must not alias x");
    HIPCHK(c, vec_combine_step32x(c, n, k, coef, vecs, z, a, x, y));
    return MDFT_OK;
}

int mdft_vec_axpy_f32x(mdft_ctx *c, size_t n, double a, const float *z, const float *x, float *y) {
    if (!c || !z || !x || !y) return MDFT_EINVAL;
    if (!aligned16(z) || !aligned16(x) || !aligned16(y)) return fail(c, MDFT_EINVAL, "vector pointers must be 16-byte aligned");
    HIPCHK(c, vec_axpy32x(c, n, a, z, x, y));
    return MDFT_OK;
}

int mdft_vec_narrow_f64(mdft_ctx *c, size_t n, const double *in, float *out) {
    if (!c || !in || !out) return MDFT_EINVAL;
    HIPCHK(c, vec_narrow(c, n, in, out));
    return MDFT_OK;
}

int mdft_scale_variables(mdft_ctx *c, size_t n, int dir, int in_fp32, const void *in, double *out) {
    if (!c || !in || !out || (dir != 1 && dir != -1)) return MDFT_EINVAL;
    if (n % (size_t)c->no) return fail(c, MDFT_EINVAL, "n must be a whole number of grid points times no");
    int rc = ensure_vscale(c);
    if (rc) return rc;
    HIPCHK(c, vec_scale_orient(c, n, c->no, dir > 0 ? c->d_vscale : c->d_vscale + c->no, in, in_fp32 != 0, out));
    return MDFT_OK;
}

int mdft_vec_widen_f32(mdft_ctx *c, size_t n, const float *in, double *out) {
    if (!c || !in || !out) return MDFT_EINVAL;
    HIPCHK(c, vec_widen(c, n, in, out));
    return MDFT_OK;
}

// src/module_pressure_correction.f90:7-156 (the bridge-free branch)
int mdft_pressure_correction(mdft_ctx *c, const double *xi, int flags, double *out) {
    int rc = ready(c);
    if (rc) return rc;
    if (!xi || !out) return fail(c, MDFT_EINVAL, "null argument");
    const size_t nbytes = sizeof(double) * c->N * c->no;
    const double *xd = xi;
    if (!(flags & MDFT_DEVICE_PTRS)) {
        HIPCHK(c, ensure_stage(c, &c->d_xi_stage));
        HIPCHK(c, hipMemcpyAsync(c->d_xi_stage, xi, nbytes, hipMemcpyHostToDevice, c->stream));
        xd = c->d_xi_stage;
    }
    // molecules in the box: n(r) = sum_io w rho0 xi^2 integrated over the box (:46-47, :74);
    // cavity: points with n(r)/n0 < 1e-3 (:50-62)
    int nb = 0;
    double s3[3];
    HIPCHK(c, launch_density_integrals(c, xd, 1.0e-3, &nb));
    if ((rc = read_sums(c, nb, s3))) return rc;
    const double volume = c->d.lx * c->d.ly * c->d.lz;
    const double n_with = s3[0] * c->dv, n_without = c->d.n0 * volume;
    const double dN = n_without - n_with;
    // bulk grand potential: the functional of the empty box, xi = 0 (:115-119)
    double *zero = nullptr;
    HIPCHK(c, hipMalloc(&zero, nbytes));
    hipError_t e = hipMemsetAsync(zero, 0, nbytes, c->stream);
    double sums[3] = {0.0, 0.0, 0.0};
    if (e == hipSuccess) rc = evaluate(c, zero, nullptr, MDFT_DEVICE_PTRS, 0, sums);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(zero);
    if (e != hipSuccess) return fail(c, MDFT_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    if (rc) return rc;
    const double f_empty = sums[0] + sums[1] + sums[2];
    const double pbulk = f_empty / volume;  // kJ/mol/A^3 (:120)
    out[0] = n_with;
    out[1] = n_without;
    out[2] = dN;
    out[3] = dN / c->d.n0;          // partial molar volume, A^3 (:85)
    out[4] = s3[1] * c->dv;         // cavity volume, A^3
    out[5] = f_empty;
    out[6] = pbulk;
    out[7] = -dN / c->d.n0 * pbulk;  // PMV correction (:128)
    out[8] = dN * c->d.kT;           // ideal-gas ("Pid") correction (:129)
    return MDFT_OK;
}

// src/module_lennardjones.f90:145-149: for each solute atom and each solvent site with eps > 0,
// -(16 pi / 3) eps_uv sigma_uv^6 n0 / rcut^3 with Lorentz-Berthelot mixing; pi as the reference writes it
int mdft_lj_tail_correction(mdft_ctx *c, int natom, const double *atoms, int nsite, const double *sites, double rcut,
                            double *out) {
    if (!c || !atoms || !sites || !out || natom < 0 || nsite < 0 || !(rcut > 0.0)) return MDFT_EINVAL;
    const double pi_ref = 3.14159;
    double e = 0.0;
    for (int u = 0; u < natom; u++)
        for (int ss = 0; ss < nsite; ss++) {
            const double eps_s = sites[6 * ss + 5], sig_s = sites[6 * ss + 4];
            if (eps_s <= DBL_EPSILON) continue;
            const double epsuv = sqrt(atoms[6 * u + 5] * eps_s);
            const double siguv = 0.5 * (atoms[6 * u + 4] + sig_s);
            const double s6 = siguv * siguv * siguv * siguv * siguv * siguv;
            e -= (16.0 * pi_ref / 3.0) * epsuv * s6 * c->d.n0 / (rcut * rcut * rcut);
        }
    *out = e;
    return MDFT_OK;
}

int mdft_charged_solute_corrections(const double *lengths, double n0, int natom, const double *atoms, int nsite,
                                    const double *sites, double relative_permittivity, double *out) {
    if (!lengths || !atoms || !sites || !out || natom < 0 || nsite < 0) return MDFT_EINVAL;
    out[0] = out[1] = 0.0;
    double charge = 0.0;
    for (int u = 0; u < natom; u++) charge += atoms[6 * u + 3];
    // only a solute with a net charge is corrected (src/module_energy_and_gradient.f90:192-203)
    if (fabs(charge) <= 1e-7) return MDFT_OK;
    // type B, periodicity (:287-315): the lattice sum of a point charge in a cubic box of side L,
    // screened by the solvent's permittivity; the mean side of a box that is not cubic
    const bool cubic = lengths[1] == lengths[0] && lengths[2] == lengths[0];
    const double L = cubic ? lengths[0] : (lengths[0] + lengths[1] + lengths[2]) / 3.0;
    if (!(L > DBL_EPSILON) || !(relative_permittivity != 0.0)) return MDFT_EINVAL;
    out[0] = -1971.01 * charge * charge / L * (1.0 - 1.0 / relative_permittivity);
    // type C, P-scheme (:321-344): the trace of the solvent molecule's quadrupole
    // sum_s q_s r_s r_s^T, each diagonal component set to 0 at or below 1e-10
    // (src/module_solvent.f90:323-326, src/module_mathematica.f90:48-70)
    double trace = 0.0;
    for (int i = 0; i < 3; i++) {
        double qii = 0.0;
        for (int s = 0; s < nsite; s++) qii += sites[6 * s + 3] * sites[6 * s + i] * sites[6 * s + i];
        trace += (fabs(qii) <= 1e-10) ? 0.0 : qii;
    }
    out[1] = -trace * n0 * 2909.857 * charge;
    return MDFT_OK;
}

int mdft_phase_times(const mdft_ctx *c, double *ms, int n) {
    if (!c || !ms) return MDFT_EINVAL;
    const int k = std::min(n, c->nphase);
    for (int i = 0; i < k; i++) ms[i] = c->phase_ms[i];
    return k;
}

int mdft_device_alloc(mdft_ctx *c, size_t bytes, void **dptr) {
    if (!c || !dptr) return MDFT_EINVAL;
    HIPCHK(c, hipMalloc(dptr, bytes ? bytes : 1));
    return MDFT_OK;
}

int mdft_device_free(mdft_ctx *c, void *dptr) {
    if (!c) return MDFT_EINVAL;
    if (dptr) HIPCHK(c, hipFree(dptr));
    return MDFT_OK;
}

int mdft_memcpy_h2d(mdft_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return MDFT_EINVAL;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_memcpy_d2h(mdft_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return MDFT_EINVAL;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_synchronize(mdft_ctx *c) {
    if (!c) return MDFT_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

// ---- learned warm start: coupling to the context (no reference counterpart)

int mdft_ws_features(mdft_ctx *c, float *d_feat, const float *h_scale) {
    if (!c || !d_feat) return MDFT_EINVAL;
    if (!c->have_vext) return fail(c, MDFT_ESTATE, "vext not set");
    if (c->d.mrso != 2) return fail(c, MDFT_EINVAL, "the warm start needs molecular symmetry order 2");
    if (h_scale) {
        const size_t bytes = sizeof(float) * 2 * c->np;
        if (!c->d_chscale) HIPCHK(c, hipMalloc(&c->d_chscale, bytes));
        HIPCHK(c, hipMemcpyAsync(c->d_chscale, h_scale, bytes, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, launch_angl2proj_feat(c, c->d_vext, 1));
    HIPCHK(c, launch_proj_to_channels(c, d_feat, 1.0f, h_scale ? c->d_chscale : nullptr));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_ws_gamma(mdft_ctx *c, const double *d_xi, double *d_gamma, float *d_target) {
    int rc = ready(c);
    if (rc) return rc;
    if (!d_xi) return fail(c, MDFT_EINVAL, "null argument");
    if ((rc = run_excess_pipeline(c, d_xi, 4, false))) return rc;
    // gamma = -(prefactor / kT) proj2angl(P4), prefactor of cproj_mrso.f90:728
    const double pi = acos(-1.0);
    const float scale = (float)(2.0 * 4.0 * pi * pi / c->d.mrso / c->d.n0);
    if (d_target) HIPCHK(c, launch_proj_to_channels(c, d_target, scale));
    if (d_gamma) {
        int nb = 0;
        HIPCHK(c, launch_final(c, d_xi, d_gamma, 2, &nb));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_ws_density(mdft_ctx *c, const double *d_xi, float *d_target) {
    int rc = ready(c);
    if (rc) return rc;
    if (!d_xi || !d_target) return fail(c, MDFT_EINVAL, "null argument");
    if (c->d.mrso != 2) return fail(c, MDFT_EINVAL, "the warm start needs molecular symmetry order 2");
    if (!c->have_vext) return fail(c, MDFT_ESTATE, "vext not set");
    HIPCHK(c, launch_angl2proj_feat(c, d_xi, 2));  // projections of rho/rho0 - 1
    HIPCHK(c, launch_proj_to_channels(c, d_target, 1.0f));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_ws_guess(mdft_ctx *c, const float *d_pred, int kind, double *d_xi) {
    if (!c || !d_xi) return MDFT_EINVAL;
    if (!c->have_vext) return fail(c, MDFT_ESTATE, "vext not set");
    if (kind != MDFT_WS_GAMMA && kind != MDFT_WS_DENSITY && kind != MDFT_WS_BOLTZMANN)
        return fail(c, MDFT_EINVAL, "unknown guess kind");
    if (!d_pred) {
        HIPCHK(c, launch_cold_start(c, d_xi, kind == MDFT_WS_BOLTZMANN));
    } else {
        if (kind == MDFT_WS_BOLTZMANN) return fail(c, MDFT_EINVAL, "the Boltzmann start takes no prediction");
        HIPCHK(c, launch_channels_to_proj(c, d_pred));
        int nb = 0;
        HIPCHK(c, launch_final(c, nullptr, d_xi, kind == MDFT_WS_GAMMA ? 3 : 4, &nb));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_ws_start_gradient(mdft_ctx *c, const double *d_xi, const double *d_df, float *d_gpred) {
    if (!c || !d_xi || !d_df || !d_gpred) return MDFT_EINVAL;
    if (!c->have_vext) return fail(c, MDFT_ESTATE, "vext not set");
    if (c->d.mrso != 2) return fail(c, MDFT_EINVAL, "the warm start needs molecular symmetry order 2");
    if (!c->d_adjscale) {
        // proj2angl^T = D angl2proj W^-1 (W: the quadrature weights), D = 8 pi^2 / mrso on the
        // channels of mu = 0, twice that for mu > 0 (each stands for its mu < 0 partner too),
        // and 0 on the imaginary channel of (m, mup = 0, mu = 0), which a real field cannot carry
        const double s = 8.0 * acos(-1.0) * acos(-1.0) / c->d.mrso;
        std::vector<float> h(2 * (size_t)c->np);
        size_t ip = 0;
        for (int m = 0; m <= c->L; m++)
            for (int mup = -m; mup <= m; mup++)
                for (int mu = 0; mu <= m; mu += c->d.mrso, ip++) {
                    const double f = (mu > 0) ? 2.0 * s : s;
                    h[2 * ip] = (float)f;
                    h[2 * ip + 1] = (mup == 0 && mu == 0) ? 0.0f : (float)f;
                }
        if (ip != (size_t)c->np) return fail(c, MDFT_EINVAL, "projection count mismatch");
        HIPCHK(c, hipMalloc(&c->d_adjscale, sizeof(float) * h.size()));
        HIPCHK(c, hipMemcpy(c->d_adjscale, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(c, launch_angl2proj_feat(c, d_xi, 3, d_df));
    HIPCHK(c, launch_proj_to_channels(c, d_gpred, 1.0f, c->d_adjscale));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

// ---- restart (src/module_density.f90:61-261, src/module_postprocessing.f90:46-94)

int mdft_angl2proj(mdft_ctx *c, const double *d_f, size_t npoints, double *d_proj) {
    if (!c || !d_f || !d_proj || npoints == 0) return MDFT_EINVAL;
    if (c->d.mrso != 2) return fail(c, MDFT_EINVAL, "the restart transforms need molecular symmetry order 2");
    HIPCHK(c, launch_transform_f64(c, d_f, reinterpret_cast<double2 *>(d_proj), npoints, true));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_proj2angl(mdft_ctx *c, const double *d_proj, size_t npoints, double *d_f) {
    if (!c || !d_f || !d_proj || npoints == 0) return MDFT_EINVAL;
    if (c->d.mrso != 2) return fail(c, MDFT_EINVAL, "the restart transforms need molecular symmetry order 2");
    HIPCHK(c, launch_transform_f64(c, d_f, reinterpret_cast<double2 *>(const_cast<double *>(d_proj)), npoints, false));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

int mdft_regrid_trilinear(mdft_ctx *c, const double *d_prev, const int *n_prev, const double *d_spacing,
                          double outside, double *d_xi) {
    if (!c || !d_prev || !n_prev || !d_spacing || !d_xi) return MDFT_EINVAL;
    for (int a = 0; a < 3; a++)
        if (n_prev[a] < 2 || !(d_spacing[a] > 0.0)) return fail(c, MDFT_EINVAL, "previous grid needs >= 2 nodes and a positive spacing per axis");
    const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
    const double len[3] = {c->d.lx, c->d.ly, c->d.lz};
    std::vector<int> cell;
    std::vector<double> wgt;
    for (int a = 0; a < 3; a++) {
        // the previous grid is centred in the box (:131-133); a new index i (1-based) is written
        // by every previous cell cp with ceiling(x0/d + 1) <= i <= floor(x1/d + 1) (:191-197),
        // the last one in the loop order winning -- the largest such cell
        const double d = len[a] / n[a], dp = d_spacing[a];
        const double off = (len[a] - dp * n_prev[a]) / 2;
        for (int i = 1; i <= n[a]; i++) {
            const double x = (i - 1) * d;
            int best = -1;
            double w = 0.0;
            const int est = (int)floor((x - off) / dp) + 1;
            for (int cp = std::min(n_prev[a] - 1, est + 1); cp >= std::max(1, est - 1); cp--) {
                const double x0 = (cp - 1) * dp + off, x1 = cp * dp + off;
                const int lo = std::max(1, (int)ceil(x0 / d + 1)), hi = std::min(n[a], (int)floor(x1 / d + 1));
                if (i >= lo && i <= hi) {
                    best = cp - 1;
                    w = (x - x0) / (x1 - x0);
                    break;
                }
            }
            cell.push_back(best);
            wgt.push_back(w);
        }
    }
    int *d_cell = nullptr;
    double *d_wgt = nullptr;
    HIPCHK(c, hipMalloc(&d_cell, sizeof(int) * cell.size()));
    hipError_t e = hipMalloc(&d_wgt, sizeof(double) * wgt.size());
    if (e == hipSuccess) e = hipMemcpy(d_cell, cell.data(), sizeof(int) * cell.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_wgt, wgt.data(), sizeof(double) * wgt.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_regrid(c, d_prev, n_prev[0], n_prev[1], d_cell, d_wgt, outside, d_xi);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_cell);
    (void)hipFree(d_wgt);
    if (e != hipSuccess) return fail(c, MDFT_EHIP, std::string("regrid: ") + hipGetErrorString(e));
    return MDFT_OK;
}

int mdft_ws_mix(mdft_ctx *c, double lam, const double *d_xa, const double *d_xb, double *d_out) {
    if (!c || !d_xa || !d_xb || !d_out) return MDFT_EINVAL;
    HIPCHK(c, launch_density_mix(c, c->N * (size_t)c->no, lam, d_xa, d_xb, d_out));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDFT_OK;
}

}  // extern "C"

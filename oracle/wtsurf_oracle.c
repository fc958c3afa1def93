/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see wtsurf_oracle.h for the scope
 * rule and how each function is pinned).  FP64 throughout, like the
 * reference (complex(dp), src/module.f90:11).
 */
#include "wtsurf_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* Pi as the reference truncates it (src/module.f90:719). */
static const double REF_PI = 3.14159265358979;

static inline double cabs1(cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* src/ham_qlayer2qlayer.f90:45-97.  Hij(:,:,ic) accumulates H(R) e^{2 pi i k.R||}
 * for every R with |R3| <= ijmax; H00 takes block (i,j) = Hij(j-i) for
 * i,j in 1..np, H01 takes block (i, j-np) = Hij(j-i) for j in np+1..2np. */
int wto_ham_qlayer2qlayer(int nw, int np, int nr, int ijmax, const int *irvec,
                          const double *hr_d, const double *k, double *h00_d,
                          double *h01_d)
{
    if (nw <= 0 || np <= 0 || nr < 0 || ijmax < 0) return -1;
    const cplx *hr = (const cplx *)hr_d;
    cplx *h00 = (cplx *)h00_d, *h01 = (cplx *)h01_d;
    const int n = nw * np;
    const size_t nw2 = (size_t)nw * nw;
    const int nslot = 2 * ijmax + 1;
    cplx *hij = calloc(nw2 * (size_t)nslot, sizeof(cplx));
    if (!hij) return -2;

    for (int ir = 0; ir < nr; ++ir) {
        const int ia = irvec[3 * ir], ib = irvec[3 * ir + 1], ic = irvec[3 * ir + 2];
        if (abs(ic) > ijmax) continue;
        const double kdotr = k[0] * (double)ia + k[1] * (double)ib;
        const double ph = 2.0 * REF_PI * kdotr;
        const cplx ratio = cos(ph) + I * sin(ph);
        cplx *dst = hij + nw2 * (size_t)(ic + ijmax);
        const cplx *src = hr + nw2 * (size_t)ir;
        for (size_t e = 0; e < nw2; ++e) dst[e] += src[e] * ratio; /* ndegen = 1 */
    }

    memset(h00, 0, sizeof(cplx) * (size_t)n * n);
    memset(h01, 0, sizeof(cplx) * (size_t)n * n);
    for (int i = 1; i <= np; ++i) {
        for (int j = 1; j <= 2 * np; ++j) {
            const int d = j - i;
            if (abs(d) > ijmax) continue;
            cplx *dst = (j <= np) ? h00 : h01;
            const int jb = (j <= np) ? j : j - np;
            const cplx *blk = hij + nw2 * (size_t)(d + ijmax);
            for (int c = 0; c < nw; ++c)
                for (int r = 0; r < nw; ++r)
                    dst[(size_t)(nw * (i - 1) + r) + (size_t)n * (nw * (jb - 1) + c)] =
                        blk[r + (size_t)nw * c];
        }
    }
    free(hij);
    return 0;
}

/* LU with partial pivoting on |re|+|im| (first maximum wins), as zgetrf,
 * then X = U^{-1} L^{-1} P I, as zgetrs, written back into a
 * (src/inverse.f90:40-55).  Work array b is n*n. */
static int lu_inverse(int n, cplx *a, cplx *b, int *ipiv)
{
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = cabs1(a[j + (size_t)n * j]);
        for (int i = j + 1; i < n; ++i) {
            const double v = cabs1(a[i + (size_t)n * j]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[j] = p;
        if (best == 0.0) { if (!info) info = j + 1; continue; }
        if (p != j)
            for (int c = 0; c < n; ++c) {
                const cplx t = a[j + (size_t)n * c];
                a[j + (size_t)n * c] = a[p + (size_t)n * c];
                a[p + (size_t)n * c] = t;
            }
        const cplx rp = 1.0 / a[j + (size_t)n * j];
        for (int i = j + 1; i < n; ++i) a[i + (size_t)n * j] *= rp;
        for (int c = j + 1; c < n; ++c) {
            const cplx u = a[j + (size_t)n * c];
            if (u == 0.0) continue;
            cplx *col = a + (size_t)n * c;
            const cplx *l = a + (size_t)n * j;
            for (int i = j + 1; i < n; ++i) col[i] -= l[i] * u;
        }
    }
    if (info) {
        /* zgesv skips the solve on a zero pivot and inv copies back its
         * identity right-hand side (src/inverse.f90:44-55). */
        memset(a, 0, sizeof(cplx) * (size_t)n * n);
        for (int i = 0; i < n; ++i) a[i + (size_t)n * i] = 1.0;
        return info;
    }
    /* B = P * I */
    memset(b, 0, sizeof(cplx) * (size_t)n * n);
    for (int i = 0; i < n; ++i) b[i + (size_t)n * i] = 1.0;
    for (int j = 0; j < n; ++j)
        if (ipiv[j] != j)
            for (int c = 0; c < n; ++c) {
                const cplx t = b[j + (size_t)n * c];
                b[j + (size_t)n * c] = b[ipiv[j] + (size_t)n * c];
                b[ipiv[j] + (size_t)n * c] = t;
            }
    for (int c = 0; c < n; ++c) {
        cplx *x = b + (size_t)n * c;
        for (int kk = 0; kk < n; ++kk) { /* unit lower */
            const cplx v = x[kk];
            if (v == 0.0) continue;
            const cplx *l = a + (size_t)n * kk;
            for (int i = kk + 1; i < n; ++i) x[i] -= v * l[i];
        }
        for (int kk = n - 1; kk >= 0; --kk) { /* upper */
            if (x[kk] == 0.0) continue;
            x[kk] /= a[kk + (size_t)n * kk];
            const cplx v = x[kk];
            const cplx *u = a + (size_t)n * kk;
            for (int i = 0; i < kk; ++i) x[i] -= v * u[i];
        }
    }
    memcpy(a, b, sizeof(cplx) * (size_t)n * n);
    return 0;
}

int wto_inv(int n, double *a_d)
{
    if (n <= 0) return -1;
    cplx *b = malloc(sizeof(cplx) * (size_t)n * n);
    int *ipiv = malloc(sizeof(int) * (size_t)n);
    if (!b || !ipiv) { free(b); free(ipiv); return -2; }
    const int info = lu_inverse(n, (cplx *)a_d, b, ipiv);
    free(b);
    free(ipiv);
    return info;
}

static void gemm(int n, const cplx *a, const cplx *b, cplx *c)
{
    memset(c, 0, sizeof(cplx) * (size_t)n * n);
    for (int j = 0; j < n; ++j) {
        cplx *cj = c + (size_t)n * j;
        for (int kk = 0; kk < n; ++kk) {
            const cplx bkj = b[kk + (size_t)n * j];
            const cplx *ak = a + (size_t)n * kk;
            for (int i = 0; i < n; ++i) cj[i] += ak[i] * bkj;
        }
    }
}

void wto_mat_mul(int n, const double *a, const double *b, double *c)
{
    gemm(n, (const cplx *)a, (const cplx *)b, (cplx *)c);
}

/* One decimation on caller-provided scratch (9 n*n matrices + ipiv).
 * src/surfgreen.f90:74-134. */
static int decimate(int n, double omega, double eta, const cplx *h00,
                    const cplx *h01, cplx *gll, cplx *grr, cplx *gb, int itermax,
                    double accuracy, int *iters, cplx *w, int *ipiv)
{
    const size_t nn = (size_t)n * n;
    cplx *ai = w, *bi = w + nn, *ei = w + 2 * nn, *es = w + 3 * nn,
         *est = w + 4 * nn, *m1 = w + 5 * nn, *m2 = w + 6 * nn, *g0 = w + 7 * nn,
         *scr = w + 8 * nn;
    const cplx z = omega + I * eta;
    memcpy(ei, h00, sizeof(cplx) * nn);
    memcpy(es, h00, sizeof(cplx) * nn);
    memcpy(est, h00, sizeof(cplx) * nn);
    memcpy(ai, h01, sizeof(cplx) * nn);
    for (int j = 0; j < n; ++j) /* beta = H01^dagger */
        for (int i = 0; i < n; ++i) bi[i + (size_t)n * j] = conj(h01[j + (size_t)n * i]);

    int info = 0, it;
    for (it = 1; it <= itermax; ++it) {
        for (size_t e = 0; e < nn; ++e) g0[e] = -ei[e];
        for (int i = 0; i < n; ++i) g0[i + (size_t)n * i] += z;
        const int inf = lu_inverse(n, g0, scr, ipiv);
        if (inf && !info) info = inf;
        gemm(n, ai, g0, m1);            /* alpha g */
        gemm(n, bi, g0, m2);            /* beta g */
        gemm(n, m1, bi, g0);            /* alpha g beta */
        for (size_t e = 0; e < nn; ++e) { ei[e] += g0[e]; es[e] += g0[e]; }
        gemm(n, m2, ai, g0);            /* beta g alpha */
        for (size_t e = 0; e < nn; ++e) { ei[e] += g0[e]; est[e] += g0[e]; }
        gemm(n, m1, ai, g0);            /* alpha g alpha */
        memcpy(ai, g0, sizeof(cplx) * nn);
        gemm(n, m2, bi, g0);            /* beta g beta */
        memcpy(bi, g0, sizeof(cplx) * nn);
        double s = 0.0;
        for (size_t e = 0; e < nn; ++e) s += cabs(ai[e]);
        if (s <= accuracy) break;
    }
    if (iters) *iters = it > itermax ? itermax : it;

    cplx *outs[3] = {gll, grr, gb};
    const cplx *eps[3] = {es, est, ei};
    for (int q = 0; q < 3; ++q) {
        cplx *o = outs[q];
        for (size_t e = 0; e < nn; ++e) o[e] = -eps[q][e];
        for (int i = 0; i < n; ++i) o[i + (size_t)n * i] += z;
        const int inf = lu_inverse(n, o, scr, ipiv);
        if (inf && !info) info = inf;
    }
    return info;
}

int wto_surfgreen_1985(int n, double omega, double eta, const double *h00,
                       const double *h01, double *gll, double *grr, double *gb,
                       int itermax, double accuracy, int *iters)
{
    if (n <= 0 || itermax < 1) return -1;
    cplx *w = malloc(sizeof(cplx) * 9 * (size_t)n * n);
    int *ipiv = malloc(sizeof(int) * (size_t)n);
    if (!w || !ipiv) { free(w); free(ipiv); return -2; }
    const int info = decimate(n, omega, eta, (const cplx *)h00, (const cplx *)h01,
                              (cplx *)gll, (cplx *)grr, (cplx *)gb, itermax,
                              accuracy, iters, w, ipiv);
    free(w);
    free(ipiv);
    return info;
}

/* Spin operators of src/surfstat.f90:102-115 (the same as
 * src/fermiarc.f90:85-101): in principal layer i (0-based here) and for
 * j < nw/2, sigma_x(a, b) = sigma_x(b, a) = 1, sigma_y(a, b) = -i,
 * sigma_y(b, a) = i, sigma_z(a, a) = 1, sigma_z(b, b) = -1 with a = nw i + j,
 * b = a + nw/2.  s is three n x n matrices (x, y, z), zeroed here. */
static void spin_operators(int nw, int np, cplx *s)
{
    const int n = nw * np, nh = nw / 2;
    const size_t nn = (size_t)n * n;
    memset(s, 0, sizeof(cplx) * 3 * nn);
    cplx *sx = s, *sy = s + nn, *sz = s + 2 * nn;
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < nh; ++j) {
            const size_t a = (size_t)nw * i + j, b = a + nh;
            sx[a + n * b] = 1.0;
            sx[b + n * a] = 1.0;
            sy[a + n * b] = -I;
            sy[b + n * a] = I;
            sz[a + n * a] = 1.0;
            sz[b + n * b] = -1.0;
        }
}

/* -sum_{e} Im (G sigma_c)(io_e, io_e) / divisor for c = x, y, z, with the full
 * product G sigma_c formed as the reference does (mat_mul, src/surfstat.f90:163-192;
 * src/fermiarc.f90:270-327). */
static void spin_sums(int n, const cplx *g, const cplx *sig, cplx *ctemp, const int *list,
                      int cnt, int offset, double divisor, double out[3])
{
    const size_t nn = (size_t)n * n;
    for (int c = 0; c < 3; ++c) {
        gemm(n, g, sig + c * nn, ctemp);
        double s = 0.0;
        for (int e = 0; e < cnt; ++e) {
            const int io = offset + list[e] - 1;
            s -= cimag(ctemp[io + (size_t)n * io]) / divisor;
        }
        out[c] = s;
    }
}

/* src/surfstat.f90:123-192 (divisor = 1) and src/fermiarc.f90:206-328
 * (divisor = pi): dos_l = -sum_{io in top} Im GLL(io,io),
 * dos_r = -sum_{b in bot} Im GRR(N-nw+b, N-nw+b), dos_bulk = -sum_i Im GB(i,i),
 * and with spin_mode != 0 the spin-resolved sums of spin_sums: left on GLL
 * over top; right on GRR over top (spin_mode 1, surfstat) or over the bottom
 * orbitals of the last layer (spin_mode 2, SurfaceDOSkk). */
static int spectral_impl(int nw, int np, int nr, int ijmax, const int *irvec,
                         const double *hr, int nk, const double *kpts, int nomega,
                         const double *omega, double eta, int ntop, const int *top,
                         int nbot, const int *bot, int itermax, double accuracy,
                         double divisor, int nthreads, double *dos_l, double *dos_r,
                         double *dos_bulk, int *iters, int spin_mode, double *spin_l,
                         double *spin_r)
{
    const int n = nw * np;
    if (nw <= 0 || np <= 0 || nk < 0 || nomega < 0 || itermax < 1) return -1;
    if (spin_mode < 0 || spin_mode > 2 || (spin_mode && (!spin_l || !spin_r))) return -1;
    for (int i = 0; i < ntop; ++i) if (top[i] < 1 || top[i] > n) return -1;
    for (int i = 0; i < nbot; ++i) if (bot[i] < 1 || bot[i] > nw) return -1;
    const size_t nn = (size_t)n * n;
    cplx *hk = malloc(sizeof(cplx) * 2 * nn * (size_t)(nk > 0 ? nk : 1));
    cplx *sig = spin_mode ? malloc(sizeof(cplx) * 3 * nn) : NULL;
    if (!hk || (spin_mode && !sig)) { free(hk); free(sig); return -2; }
    if (spin_mode) spin_operators(nw, np, sig);
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < nk; ++ik)
        wto_ham_qlayer2qlayer(nw, np, nr, ijmax, irvec, hr, kpts + 2 * ik,
                              (double *)(hk + 2 * nn * ik),
                              (double *)(hk + 2 * nn * ik + nn));

    const size_t total_sz = (size_t)nk * nomega;
#pragma omp parallel reduction(| : err)
    {
        cplx *w = malloc(sizeof(cplx) * 13 * nn);
        int *ipiv = malloc(sizeof(int) * (size_t)n);
        if (!w || !ipiv) {
            err |= 2;
        } else {
            cplx *gll = w + 9 * nn, *grr = w + 10 * nn, *gb = w + 11 * nn, *ct = w + 12 * nn;
            const long total = (long)nk * nomega;
#pragma omp for schedule(dynamic, 1)
            for (long t = 0; t < total; ++t) {
                const int ik = (int)(t / nomega), iw = (int)(t % nomega);
                int it = 0;
                const int inf = decimate(n, omega[iw], eta, hk + 2 * nn * ik,
                                         hk + 2 * nn * ik + nn, gll, grr, gb,
                                         itermax, accuracy, &it, w, ipiv);
                if (inf) err |= 1;
                double dl = 0.0, dr = 0.0, db = 0.0;
                for (int i = 0; i < ntop; ++i) {
                    const int io = top[i] - 1;
                    dl -= cimag(gll[io + (size_t)n * io]) / divisor;
                }
                for (int i = 0; i < nbot; ++i) {
                    const int io = n - nw + bot[i] - 1;
                    dr -= cimag(grr[io + (size_t)n * io]) / divisor;
                }
                for (int i = 0; i < n; ++i) db -= cimag(gb[i + (size_t)n * i]) / divisor;
                const size_t o = (size_t)ik + (size_t)nk * iw;
                dos_l[o] = dl;
                dos_r[o] = dr;
                dos_bulk[o] = db;
                if (iters) iters[o] = it;
                if (spin_mode) {
                    double sl[3], sr[3];
                    spin_sums(n, gll, sig, ct, top, ntop, 0, divisor, sl);
                    if (spin_mode == 1)
                        spin_sums(n, grr, sig, ct, top, ntop, 0, divisor, sr);
                    else
                        spin_sums(n, grr, sig, ct, bot, nbot, n - nw, divisor, sr);
                    for (int c = 0; c < 3; ++c) {
                        spin_l[o + c * total_sz] = sl[c];
                        spin_r[o + c * total_sz] = sr[c];
                    }
                }
            }
        }
        free(w);
        free(ipiv);
    }
    free(hk);
    free(sig);
    return err ? (err & 2 ? -2 : 1) : 0;
}

int wto_spectral(int nw, int np, int nr, int ijmax, const int *irvec,
                 const double *hr, int nk, const double *kpts, int nomega,
                 const double *omega, double eta, int ntop, const int *top,
                 int nbot, const int *bot, int itermax, double accuracy,
                 double divisor, int nthreads, double *dos_l, double *dos_r,
                 double *dos_bulk, int *iters)
{
    return spectral_impl(nw, np, nr, ijmax, irvec, hr, nk, kpts, nomega, omega, eta, ntop, top,
                         nbot, bot, itermax, accuracy, divisor, nthreads, dos_l, dos_r,
                         dos_bulk, iters, 0, NULL, NULL);
}

int wto_spectral_spin(int nw, int np, int nr, int ijmax, const int *irvec,
                      const double *hr, int nk, const double *kpts, int nomega,
                      const double *omega, double eta, int ntop, const int *top,
                      int nbot, const int *bot, int itermax, double accuracy,
                      double divisor, int nthreads, double *dos_l, double *dos_r,
                      double *dos_bulk, int *iters, int spin_mode, double *spin_l,
                      double *spin_r)
{
    if (spin_mode < 1) return -1;
    return spectral_impl(nw, np, nr, ijmax, irvec, hr, nk, kpts, nomega, omega, eta, ntop, top,
                         nbot, bot, itermax, accuracy, divisor, nthreads, dos_l, dos_r,
                         dos_bulk, iters, spin_mode, spin_l, spin_r);
}

/* The QPI autocorrelation of SurfaceDOSkk (src/fermiarc.f90:496-560), as the
 * reference loops it: for every q = (iq1, iq2) of the k-plane,
 *   out(iq1, iq2) = sum_{ik2} sum_{ik1} sum_f a_f(ik1, ik2) a_f(ik1q, ik2q),
 *   ik1q = mod(ik1 + iq1 - 2, m1) + 1,  ik2q = mod(ik2 + iq2 - 2, m2) + 1
 * (1-based), the field terms of one (ik1, ik2) added in order before the
 * running sum (jsdos, :545-549).  fields is [nfield][nky][nkx] (ik1 fastest,
 * the reference's (nkx, nky) arrays), out is [nky][nkx]. */
int wto_kplane_autocorrelation(int nkx, int nky, int m1, int m2, int nfield,
                               const double *fields, double *out)
{
    if (nkx < 1 || nky < 1 || m1 < 1 || m2 < 1 || nfield < 1) return -1;
    if (m1 > nkx || m2 > nky) return -1;
    const size_t plane = (size_t)nkx * nky;
#pragma omp parallel for schedule(dynamic, 1)
    for (long iq = 0; iq < (long)plane; ++iq) {
        const int iq1 = (int)(iq % nkx) + 1, iq2 = (int)(iq / nkx) + 1;
        double acc = 0.0;
        for (int ik2 = 1; ik2 <= nky; ++ik2)
            for (int ik1 = 1; ik1 <= nkx; ++ik1) {
                const int ik1q = (ik1 + iq1 - 2) % m1 + 1, ik2q = (ik2 + iq2 - 2) % m2 + 1;
                const size_t p = (size_t)(ik1 - 1) + (size_t)nkx * (ik2 - 1);
                const size_t pq = (size_t)(ik1q - 1) + (size_t)nkx * (ik2q - 1);
                double term = fields[p] * fields[pq];
                for (int f = 1; f < nfield; ++f)
                    term += fields[f * plane + p] * fields[f * plane + pq];
                acc += term;
            }
        out[iq] = acc;
    }
    return 0;
}

/* The k-path joint density of states of surfstat_jdos (src/surfstat.f90:891-907),
 * as the reference loops it, for each field a and energy j separately:
 *   out(iq, j) = sum_{ik1 = imin1}^{imax1} a(ik1, j) a(ik1 + iq1, j),
 *   Nk_half = (nk - 1)/2, iq1 = iq - Nk_half,
 *   imin1 = max(-Nk_half - iq1, -Nk_half) + Nk_half + 1,
 *   imax1 = min(Nk_half - iq1, Nk_half) + Nk_half + 1      (1-based).
 * fields and out are [nfield][nomega][nk] (k fastest, the reference's
 * dos_l(knv2, omeganum)). */
int wto_kpath_jdos(int nk, int nomega, int nfield, const double *fields, double *out)
{
    if (nk < 1 || nomega < 1 || nfield < 1) return -1;
    const int nkh = (nk - 1) / 2;
#pragma omp parallel for schedule(static)
    for (long fj = 0; fj < (long)nfield * nomega; ++fj) {
        const double *a = fields + (size_t)fj * nk;
        double *o = out + (size_t)fj * nk;
        for (int iq = 1; iq <= nk; ++iq) {
            const int iq1 = iq - nkh;
            const int lo = (-nkh - iq1 > -nkh ? -nkh - iq1 : -nkh) + nkh + 1;
            const int hi = (nkh - iq1 < nkh ? nkh - iq1 : nkh) + nkh + 1;
            double s = 0.0;
            for (int ik1 = lo; ik1 <= hi; ++ik1) s += a[ik1 - 1] * a[ik1 + iq1 - 1];
            o[iq - 1] = s;
        }
    }
    return 0;
}

/*
 * ps_oracle.c -- CPU FP64 restatement of the polyspectrum FFT estimator.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; see ps_oracle.h).
 * Built with -ffp-contract=off so that, like the reference (common.mk:10,
 * -std=c99, no FMA contraction), every mask comparison and corner norm is
 * evaluated with separately rounded operations.
 *
 * What pins each function (the tests under tests/ do the checking):
 *   or_fft1d / or_fft3d      DFT definition by brute force on small n, numpy.fft
 *                            golden spectra, Parseval, impulses, plane waves.
 *   or_kfilter (all kinds)   integer brute force of each criterion, the closed
 *                            form of the corner extremes sum((|m_i| -+ e)^2),
 *                            popcounts of shells against lattice-point counts.
 *   or_calc_k3_min_and_max   closed form of the 8-candidate extremes.
 *   or_polyspectrum          the direct lattice polygon sum (independent of
 *                            any FFT), exact integer polygon counts, scaling
 *                            B(l*d) = l^3 B, shift invariance, phase-only P.
 *   or_num_polygons          exact lattice polygon counts.
 *   or_polygon_gradient      central differences of or_polyspectrum (exact for
 *                            the cubic amplitude estimator with Richardson's
 *                            two-step formula), an independent numpy gradient.
 */
#include "ps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI acos(-1.0) /* the reference's PI (filter.c:19, kvectors.c:20) */

typedef struct { double re, im; } cx;

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time (radix 4, 2, odd primes).   */
/* The GPU path uses a Stockham LDS formulation instead; this one is kept     */
/* structurally different on purpose.                                         */
/* ------------------------------------------------------------------------ */

static int factorize(int n, int *f)
{
    int nf = 0;
    while (n % 4 == 0) { f[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { f[nf++] = 2; n /= 2; }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) { f[nf++] = p; n /= p; }
    if (n > 1) f[nf++] = n;
    return nf;
}

/* tw[t] = exp(sign * 2 pi i t / n), from a long double evaluation of the
 * exactly reduced angle. */
static cx *make_twiddles(int n, int sign)
{
    cx *tw = (cx *)malloc(sizeof(cx) * (size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; t++) {
        long double a = two_pi * (long double)t / (long double)n;
        tw[t].re = (double)cosl(a);
        tw[t].im = (double)((long double)sign * sinl(a));
    }
    return tw;
}

static inline cx cmul(cx a, cx b)
{
    cx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}

static void fft_rec(cx *out, const cx *in, int n, long is, const int *fac,
                    const cx *tw, int ntot, int sign)
{
    const int p = fac[0];
    const int m = n / p;
    if (m == 1) {
        for (int q = 0; q < p; q++) out[q] = in[(long)q * is];
    } else {
        for (int q = 0; q < p; q++)
            fft_rec(out + (long)q * m, in + (long)q * is, m, is * p, fac + 1, tw, ntot, sign);
    }
    const int step = ntot / n; /* W_n^j = tw[j * step] */
    cx t[64];
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            cx v = out[(long)q * m + k];
            t[q] = (q == 0 || k == 0) ? v : cmul(v, tw[((long)q * k * step) % ntot]);
        }
        if (p == 2) {
            out[k].re = t[0].re + t[1].re;           out[k].im = t[0].im + t[1].im;
            out[m + k].re = t[0].re - t[1].re;       out[m + k].im = t[0].im - t[1].im;
        } else if (p == 4) {
            cx y0 = { t[0].re + t[2].re, t[0].im + t[2].im };
            cx y1 = { t[0].re - t[2].re, t[0].im - t[2].im };
            cx y2 = { t[1].re + t[3].re, t[1].im + t[3].im };
            cx d  = { t[1].re - t[3].re, t[1].im - t[3].im };
            /* times W_4 = sign * i */
            cx y3 = { -sign * d.im, sign * d.re };
            out[k].re = y0.re + y2.re;           out[k].im = y0.im + y2.im;
            out[2 * m + k].re = y0.re - y2.re;   out[2 * m + k].im = y0.im - y2.im;
            out[m + k].re = y1.re + y3.re;       out[m + k].im = y1.im + y3.im;
            out[3 * m + k].re = y1.re - y3.re;   out[3 * m + k].im = y1.im - y3.im;
        } else {
            const int pstep = ntot / p;
            for (int r = 0; r < p; r++) {
                cx acc = t[0];
                for (int q = 1; q < p; q++) {
                    cx w = tw[((long)((q * r) % p)) * pstep];
                    cx v = cmul(t[q], w);
                    acc.re += v.re; acc.im += v.im;
                }
                out[(long)r * m + k] = acc;
            }
        }
    }
}

int or_fft1d(int n, double *data, int sign)
{
    int fac[64];
    if (n < 1) return -1;
    if (n == 1) return 0;
    int nf = factorize(n, fac);
    for (int i = 0; i < nf; i++) if (fac[i] > 64) return -2;
    cx *tw = make_twiddles(n, sign);
    cx *tmp = (cx *)malloc(sizeof(cx) * (size_t)n);
    fft_rec(tmp, (const cx *)data, n, 1, fac, tw, n, sign);
    memcpy(data, tmp, sizeof(cx) * (size_t)n);
    free(tmp);
    free(tw);
    return 0;
}

int or_fft3d(int n, double *data_, int sign)
{
    int fac[64];
    if (n < 1) return -1;
    if (n == 1) return 0;
    int nf = factorize(n, fac);
    for (int i = 0; i < nf; i++) if (fac[i] > 64) return -2;
    cx *data = (cx *)data_;
    cx *tw = make_twiddles(n, sign);
    const int B = 16; /* lines gathered per block on the strided axes */
    cx *buf = (cx *)malloc(sizeof(cx) * (size_t)n * B);
    cx *tmp = (cx *)malloc(sizeof(cx) * (size_t)n);
    const long n2 = (long)n * n;

    /* axis 2 (contiguous) */
    for (long r = 0; r < n2; r++) {
        cx *row = data + r * n;
        fft_rec(tmp, row, n, 1, fac, tw, n, sign);
        memcpy(row, tmp, sizeof(cx) * (size_t)n);
    }
    /* axis 1 (stride n) and axis 0 (stride n^2) */
    for (int axis = 1; axis >= 0; axis--) {
        const long stride = (axis == 1) ? n : n2;
        for (long o = 0; o < n; o++) {            /* the other strided index */
            for (int l0 = 0; l0 < n; l0 += B) {
                int nb = (n - l0 < B) ? n - l0 : B;
                long base = (axis == 1) ? o * n2 + l0 : o * n + l0;
                for (int e = 0; e < n; e++)
                    for (int b = 0; b < nb; b++)
                        buf[(long)b * n + e] = data[base + e * stride + b];
                for (int b = 0; b < nb; b++) {
                    fft_rec(tmp, buf + (long)b * n, n, 1, fac, tw, n, sign);
                    memcpy(buf + (long)b * n, tmp, sizeof(cx) * (size_t)n);
                }
                for (int e = 0; e < n; e++)
                    for (int b = 0; b < nb; b++)
                        data[base + e * stride + b] = buf[(long)b * n + e];
            }
        }
    }
    free(tmp);
    free(buf);
    free(tw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* input promotion and phases                                                */
/* ------------------------------------------------------------------------ */

void or_promote_f32(int64_t count, const float *in, double *out)
{
    /* grid.c:168: (double)tmparray[...] + 0*I */
    for (int64_t i = 0; i < count; i++) { out[2 * i] = (double)in[i]; out[2 * i + 1] = 0.0; }
}

void or_promote_f64(int64_t count, const double *in, double *out)
{
    /* grid.c:223 */
    for (int64_t i = 0; i < count; i++) { out[2 * i] = in[i]; out[2 * i + 1] = 0.0; }
}

void or_calc_phase(int nbins, double *array)
{
    /* utils_fftw.c:78-93: value / cabs(value); an exact zero becomes NaN */
    int64_t cnt = (int64_t)nbins * nbins * nbins;
    for (int64_t i = 0; i < cnt; i++) {
        double re = array[2 * i], im = array[2 * i + 1];
        double a = hypot(re, im);
        array[2 * i] = re / a;
        array[2 * i + 1] = im / a;
    }
}

/* ------------------------------------------------------------------------ */
/* masks (filter.c)                                                          */
/* ------------------------------------------------------------------------ */

/* index -> signed lattice coordinate, filter.c:44-55 */
static inline int lat(int i, int nbins) { return (i < nbins / 2) ? i : -nbins + i; }

/* filter.c:257-289, including the reference's running value passing through
 * sqrt and back through squaring on every call */
void or_get_corners(int mx, int my, int mz, double binwidth, double *mmin, double *mmax)
{
    double x = (double)mx, y = (double)my, z = (double)mz;
    double e = 0.5 * binwidth;
    double minValue = (*mmin) * (*mmin);
    double maxValue = (*mmax) * (*mmax);
    double c[8];
    c[0] = (x - e) * (x - e) + (y - e) * (y - e) + (z - e) * (z - e);
    c[1] = (x + e) * (x + e) + (y - e) * (y - e) + (z - e) * (z - e);
    c[2] = (x - e) * (x - e) + (y + e) * (y + e) + (z - e) * (z - e);
    c[3] = (x - e) * (x - e) + (y - e) * (y - e) + (z + e) * (z + e);
    c[4] = (x + e) * (x + e) + (y + e) * (y + e) + (z - e) * (z - e);
    c[5] = (x + e) * (x + e) + (y - e) * (y - e) + (z + e) * (z + e);
    c[6] = (x - e) * (x - e) + (y + e) * (y + e) + (z + e) * (z + e);
    c[7] = (x + e) * (x + e) + (y + e) * (y + e) + (z + e) * (z + e);
    for (int i = 0; i < 8; i++) {
        if (c[i] > maxValue) maxValue = c[i];
        if (c[i] < minValue) minValue = c[i];
    }
    *mmin = sqrt(minValue);
    *mmax = sqrt(maxValue);
}

void or_kfilter(int nbins, int kind, double k, double *kmin, double *kmax,
                double binwidth, double boxsize, uint8_t *mask)
{
    const double k_in_kf = k / (2. * OR_PI) * boxsize;          /* filter.c:36 */
    double mmin_r = 0., mmax_r = 0.;
    if (kind == OR_MASK_RANGE || kind == OR_MASK_WATKINSON_KN) {
        mmin_r = *kmin / (2. * OR_PI) * boxsize;                  /* filter.c:137-138 */
        mmax_r = *kmax / (2. * OR_PI) * boxsize;
    }
    double dmin = 1.e30, dmax = 0.;                                /* filter.c:85-86 */
    for (int i = 0; i < nbins; i++) {
        int mx = lat(i, nbins);
        for (int j = 0; j < nbins; j++) {
            int my = lat(j, nbins);
            for (int l = 0; l < nbins; l++) {
                int mz = lat(l, nbins);
                int64_t idx = ((int64_t)i * nbins + j) * nbins + l;
                double m = sqrt((double)(mx * mx + my * my + mz * mz));
                int in = 0;
                switch (kind) {
                case OR_MASK_DETERMINING: {
                    double expr = fabs(k_in_kf - m);
                    in = expr <= 0.5 * binwidth;
                    if (in) or_get_corners(mx, my, mz, binwidth, &dmin, &dmax);
                    break;
                }
                case OR_MASK_RANGE: {
                    double emin = mmin_r - m, emax = m - mmax_r;
                    in = (emin <= 0.5 * binwidth && emax <= 0.5 * binwidth);
                    break;
                }
                case OR_MASK_WATKINSON: {
                    double pmin = 1.e30, pmax = 0.;
                    or_get_corners(mx, my, mz, binwidth, &pmin, &pmax);
                    in = (pmin <= k_in_kf && k_in_kf <= pmax);
                    break;
                }
                case OR_MASK_WATKINSON_KN: {
                    double pmin = 1.e30, pmax = 0.;
                    or_get_corners(mx, my, mz, binwidth, &pmin, &pmax);
                    in = ((pmax >= mmax_r && pmin <= mmax_r) || (pmax <= mmax_r && pmin >= mmin_r) ||
                          (pmax >= mmin_r && pmin <= mmin_r));
                    break;
                }
                default: {
                    double expr = fabs(k_in_kf - m);
                    in = expr <= 0.5 * binwidth;
                }
                }
                mask[idx] = (uint8_t)(in ? 1 : 0);
            }
        }
    }
    if (kind == OR_MASK_DETERMINING) {                             /* filter.c:112-113 */
        *kmin = dmin * 2. * OR_PI / boxsize;
        *kmax = dmax * 2. * OR_PI / boxsize;
    }
}

/* filter.c:291-319 */
void or_calc_k3_min_and_max(const double *k, double *kmin, double *kmax)
{
    double c = (k[0] * k[0] + k[1] * k[1] - k[2] * k[2]) / (2. * k[0] * k[1]);
    double t[8];
    t[0] = kmin[0] * kmin[0] + kmin[1] * kmin[1] - 2. * kmin[0] * kmin[1] * c;
    t[1] = kmin[0] * kmin[0] + kmax[1] * kmax[1] - 2. * kmin[0] * kmax[1] * c;
    t[2] = kmax[0] * kmax[0] + kmin[1] * kmin[1] - 2. * kmax[0] * kmin[1] * c;
    t[3] = kmax[0] * kmax[0] + kmax[1] * kmax[1] - 2. * kmax[0] * kmax[1] * c;
    t[4] = k[0] * k[0] + kmax[1] * kmax[1] - 2. * k[0] * kmax[1] * c;
    t[5] = kmax[0] * kmax[0] + k[1] * k[1] - 2. * kmax[0] * k[1] * c;
    t[6] = k[0] * k[0] + kmin[1] * kmin[1] - 2. * k[0] * kmin[1] * c;
    t[7] = kmin[0] * kmin[0] + k[1] * k[1] - 2. * kmin[0] * k[1] * c;
    double lo = 1.e30, hi = 0.;
    for (int i = 0; i < 8; i++) {
        if (t[i] < lo) lo = t[i];
        if (t[i] > hi) hi = t[i];
    }
    kmin[2] = sqrt(lo);
    kmax[2] = sqrt(hi);
}

/* kvectors.c:226-247 */
double or_calc_k3(double k1, double k2, double cos_theta)
{
    return sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * cos_theta);
}
double or_calc_k3min(double k1, double k2, double theta, double dcos)
{
    return or_calc_k3(k1, k2, fmin(1., cos(theta) + 0.5 * dcos));
}
double or_calc_k3max(double k1, double k2, double theta, double dcos)
{
    return or_calc_k3(k1, k2, fmax(-1., cos(theta) - 0.5 * dcos));
}

/* ------------------------------------------------------------------------ */
/* grid operations (utils_fftw.c:32-76)                                      */
/* ------------------------------------------------------------------------ */

void or_product(int64_t count, const double *a, const double *b, double *out)
{
    for (int64_t i = 0; i < count; i++) {
        double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        out[2 * i] = ar * br - ai * bi;
        out[2 * i + 1] = ar * bi + ai * br;
    }
}

void or_sum(int64_t count, const double *a, double *out2)
{
    double re = 0., im = 0.;
    for (int64_t i = 0; i < count; i++) { re += a[2 * i]; im += a[2 * i + 1]; }
    out2[0] = re;
    out2[1] = im;
}

/* the per-leg mask choice of polyspectrum.c:79-118 (and :186-222) */
static void leg_mask(int nbins, int i, int n, const double *k, double *kmin, double *kmax,
                     double bw, int kcase, double L, uint8_t *mask)
{
    switch (kcase) {
    case 1:
        if (i < n - 1) {
            or_kfilter(nbins, OR_MASK_DETERMINING, k[i], &kmin[i], &kmax[i], bw, L, mask);
        } else {
            if (n == 3) {
                or_calc_k3_min_and_max(k, kmin, kmax);
            } else {
                /* filter.c:321-325 writes through a mis-indexed double**;
                 * this is its intended meaning */
                kmin[n - 1] = k[n - 1];
                kmax[n - 1] = k[n - 1];
            }
            or_kfilter(nbins, OR_MASK_RANGE, 0., &kmin[i], &kmax[i], bw, L, mask);
        }
        break;
    case 2:
        if (i < n - 1) or_kfilter(nbins, OR_MASK_DEFAULT, k[i], NULL, NULL, bw, L, mask);
        else or_kfilter(nbins, OR_MASK_RANGE, 0., &kmin[i], &kmax[i], bw, L, mask);
        break;
    case 3:
        if (i < n - 1) or_kfilter(nbins, OR_MASK_WATKINSON, k[i], NULL, NULL, bw, L, mask);
        else or_kfilter(nbins, OR_MASK_WATKINSON_KN, 0., &kmin[i], &kmax[i], bw, L, mask);
        break;
    default:
        or_kfilter(nbins, OR_MASK_DEFAULT, k[i], NULL, NULL, bw, L, mask);
    }
}

static void mask_to_complex(int64_t cnt, const uint8_t *m, double *c)
{
    for (int64_t i = 0; i < cnt; i++) { c[2 * i] = m[i] ? 1. : 0.; c[2 * i + 1] = 0.; }
}

/* polyspectrum.c:51-161 */
double or_polyspectrum(int nbins, const double *fft_array, int n, const double *k,
                       double *kmin, double *kmax, double kbinwidth, int kcase,
                       double boxsize, int diagnostic_sums, double *sums_out)
{
    const int64_t cnt = (int64_t)nbins * nbins * nbins;
    double *accI = (double *)malloc(sizeof(double) * 2 * (size_t)cnt);
    double *accD = (double *)malloc(sizeof(double) * 2 * (size_t)cnt);
    double *filt = (double *)malloc(sizeof(double) * 2 * (size_t)cnt);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)cnt);
    uint8_t *mask = (uint8_t *)malloc((size_t)cnt);
    double volume = boxsize * boxsize * boxsize;
    for (int64_t i = 0; i < cnt; i++) {
        accI[2 * i] = 1.; accI[2 * i + 1] = 0.;
        accD[2 * i] = 1.; accD[2 * i + 1] = 0.;
    }
    for (int i = 0; i < n; i++) {
        leg_mask(nbins, i, n, k, kmin, kmax, kbinwidth, kcase, boxsize, mask);
        mask_to_complex(cnt, mask, filt);
        /* I = IFFT(filter)  (polyspectrum.c:41-49) */
        memcpy(tmp, filt, sizeof(double) * 2 * (size_t)cnt);
        or_fft3d(nbins, tmp, +1);
        or_product(cnt, accI, tmp, accI);
        /* D = IFFT(filter * field)  (polyspectrum.c:24-38) */
        or_product(cnt, filt, fft_array, tmp);
        or_fft3d(nbins, tmp, +1);
        or_product(cnt, accD, tmp, accD);
    }
    double sI[2], sD[2];
    or_sum(cnt, accI, sI);
    or_sum(cnt, accD, sD);
    if (diagnostic_sums) {
        /* polyspectrum.c:142 sums each accumulator twice more for a printf */
        double t[2];
        or_sum(cnt, accI, t); or_sum(cnt, accI, t);
        or_sum(cnt, accD, t); or_sum(cnt, accD, t);
    }
    if (sums_out) { sums_out[0] = sD[0]; sums_out[1] = sD[1]; sums_out[2] = sI[0]; sums_out[3] = sI[1]; }
    double result = pow(volume, n - 1) * pow(nbins, -n * 3) * sD[0] / sI[0];  /* :155 */
    free(mask); free(tmp); free(filt); free(accD); free(accI);
    return result;
}

/* polyspectrum.c:163-252 */
double or_num_polygons(int nbins, int n, const double *k, double *kmin, double *kmax,
                       double kbinwidth, int kcase, double boxsize)
{
    const int64_t cnt = (int64_t)nbins * nbins * nbins;
    double *accI = (double *)malloc(sizeof(double) * 2 * (size_t)cnt);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)cnt);
    uint8_t *mask = (uint8_t *)malloc((size_t)cnt);
    for (int64_t i = 0; i < cnt; i++) { accI[2 * i] = 1.; accI[2 * i + 1] = 0.; }
    for (int i = 0; i < n; i++) {
        leg_mask(nbins, i, n, k, kmin, kmax, kbinwidth, kcase, boxsize, mask);
        mask_to_complex(cnt, mask, tmp);
        or_fft3d(nbins, tmp, +1);
        or_product(cnt, accI, tmp, accI);
    }
    double sI[2];
    or_sum(cnt, accI, sI);
    free(mask); free(tmp); free(accI);
    return sI[0] / (double)((int64_t)nbins * nbins * nbins);          /* :247 */
}

/* ------------------------------------------------------------------------ */
/* gradient of one polygon value with respect to the real field              */
/* ------------------------------------------------------------------------ */

/* value = V^(n-1) N^(-3n) Re S / Re C with S = sum_x prod_i D_i(x) (the
 * estimator above, polyspectrum.c:51-161), D_i = B(m_i s), s = F(delta) or,
 * with phase, p = s/|s| (utils_fftw.c:78-93); F/B are the unnormalised
 * forward/backward transforms of fft.c.  C does not depend on delta.  With
 * Q_i = prod_{j != i} D_j,
 *   dS = sum_k R(k) ds(k),  R = sum_i m_i B(Q_i)   (sum_x Q B(f) = sum_k f B(Q)),
 * and ds(k) = sum_y d delta(y) e^{-iky}, so d Re S / d delta = Re F(R).  With
 * phase, dp = (ds - p Re(conj(p) ds)) / |s| turns R into
 * R' = (R - conj(p) Re(R p)) / |s| (0 where s = 0).  grad += weight * value'
 * (N^3 doubles); kmin/kmax are updated as by or_polyspectrum. */
double or_polygon_gradient(int nbins, const double *field, int n, const double *k, double *kmin, double *kmax,
                           double kbinwidth, int kcase, double boxsize, int phase, double weight, double *grad)
{
    const int64_t cnt = (int64_t)nbins * nbins * nbins;
    const size_t cb = sizeof(double) * 2 * (size_t)cnt;
    double *s = (double *)malloc(cb), *src = (double *)malloc(cb), *R = (double *)calloc(1, cb);
    double *Q = (double *)malloc(cb), *accI = (double *)malloc(cb), *tmp = (double *)malloc(cb);
    double *D[3] = {NULL, NULL, NULL};
    uint8_t *mask[3] = {NULL, NULL, NULL};
    for (int64_t i = 0; i < cnt; i++) { s[2 * i] = field[i]; s[2 * i + 1] = 0.; }
    or_fft3d(nbins, s, -1);
    memcpy(src, s, cb);
    if (phase) or_calc_phase(nbins, src);
    for (int64_t i = 0; i < cnt; i++) { accI[2 * i] = 1.; accI[2 * i + 1] = 0.; }
    for (int i = 0; i < n; i++) {
        mask[i] = (uint8_t *)malloc((size_t)cnt);
        D[i] = (double *)malloc(cb);
        leg_mask(nbins, i, n, k, kmin, kmax, kbinwidth, kcase, boxsize, mask[i]);
        mask_to_complex(cnt, mask[i], tmp);
        or_fft3d(nbins, tmp, +1);
        or_product(cnt, accI, tmp, accI);
        for (int64_t c = 0; c < cnt; c++) {
            D[i][2 * c] = mask[i][c] ? src[2 * c] : 0.;
            D[i][2 * c + 1] = mask[i][c] ? src[2 * c + 1] : 0.;
        }
        or_fft3d(nbins, D[i], +1);
    }
    /* S and the R of every leg */
    for (int64_t c = 0; c < cnt; c++) { Q[2 * c] = 1.; Q[2 * c + 1] = 0.; }
    for (int i = 0; i < n; i++) or_product(cnt, Q, D[i], Q);
    double sD[2], sI[2];
    or_sum(cnt, Q, sD);
    or_sum(cnt, accI, sI);
    for (int i = 0; i < n; i++) {
        for (int64_t c = 0; c < cnt; c++) { Q[2 * c] = 1.; Q[2 * c + 1] = 0.; }
        for (int j = 0; j < n; j++)
            if (j != i) or_product(cnt, Q, D[j], Q);
        or_fft3d(nbins, Q, +1);
        for (int64_t c = 0; c < cnt; c++)
            if (mask[i][c]) { R[2 * c] += Q[2 * c]; R[2 * c + 1] += Q[2 * c + 1]; }
    }
    if (phase) {
        for (int64_t c = 0; c < cnt; c++) {
            const double a = hypot(s[2 * c], s[2 * c + 1]);
            if (a > 0.) {
                const double pr = s[2 * c] / a, pi = s[2 * c + 1] / a;
                const double re = R[2 * c] * pr - R[2 * c + 1] * pi;  /* Re(R p) */
                R[2 * c] = (R[2 * c] - pr * re) / a;
                R[2 * c + 1] = (R[2 * c + 1] + pi * re) / a;
            } else {
                R[2 * c] = R[2 * c + 1] = 0.;
            }
        }
    }
    or_fft3d(nbins, R, -1);
    const double volume = boxsize * boxsize * boxsize;
    const double cn = pow(volume, n - 1) * pow(nbins, -n * 3);
    const double coef = weight * cn / sI[0];
    for (int64_t c = 0; c < cnt; c++) grad[c] += coef * R[2 * c];
    for (int i = 0; i < n; i++) { free(mask[i]); free(D[i]); }
    free(s); free(src); free(R); free(Q); free(accI); free(tmp);
    return cn * sD[0] / sI[0];
}

/* ------------------------------------------------------------------------ */
/* direct lattice polygon sum (independent of any FFT)                       */
/* ------------------------------------------------------------------------ */

int64_t or_direct_polygon_sum(int nbins, int n, const double *spec,
                              const uint8_t *mask1, const uint8_t *mask2,
                              const uint8_t *mask3, double *sum_out)
{
    const int64_t cnt = (int64_t)nbins * nbins * nbins;
    int64_t *l1 = (int64_t *)malloc(sizeof(int64_t) * (size_t)cnt);
    int64_t *l2 = (int64_t *)malloc(sizeof(int64_t) * (size_t)cnt);
    int64_t n1 = 0, n2 = 0, count = 0;
    for (int64_t i = 0; i < cnt; i++) {
        if (mask1[i]) l1[n1++] = i;
        if (mask2[i]) l2[n2++] = i;
    }
    long double sre = 0.L, sim = 0.L;
    const int64_t N = nbins;
    for (int64_t a = 0; a < n1; a++) {
        int64_t i1 = l1[a];
        int64_t x1 = i1 / (N * N), y1 = (i1 / N) % N, z1 = i1 % N;
        double ar = spec[2 * i1], ai = spec[2 * i1 + 1];
        if (n == 2) {
            int64_t x2 = (N - x1) % N, y2 = (N - y1) % N, z2 = (N - z1) % N;
            int64_t i2 = (x2 * N + y2) * N + z2;
            if (!mask2[i2]) continue;
            double br = spec[2 * i2], bi = spec[2 * i2 + 1];
            sre += (long double)ar * br - (long double)ai * bi;
            sim += (long double)ar * bi + (long double)ai * br;
            count++;
            continue;
        }
        for (int64_t b = 0; b < n2; b++) {
            int64_t i2 = l2[b];
            int64_t x2 = i2 / (N * N), y2 = (i2 / N) % N, z2 = i2 % N;
            int64_t x3 = (2 * N - x1 - x2) % N, y3 = (2 * N - y1 - y2) % N, z3 = (2 * N - z1 - z2) % N;
            int64_t i3 = (x3 * N + y3) * N + z3;
            if (!mask3[i3]) continue;
            long double br = spec[2 * i2], bi = spec[2 * i2 + 1];
            long double cr = spec[2 * i3], ci = spec[2 * i3 + 1];
            long double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
            sre += pr * cr - pi * ci;
            sim += pr * ci + pi * cr;
            count++;
        }
    }
    sum_out[0] = (double)sre;
    sum_out[1] = (double)sim;
    free(l2); free(l1);
    return count;
}

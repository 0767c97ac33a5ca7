/*
 * ps_oracle.h -- CPU FP64 restatement of the polyspectrum FFT estimator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * No product code links or calls it.
 *
 * Conventions follow the reference (annehutter/polyspectrum):
 *   - grids are C-ordered N^3, index i*N*N + j*N + l (grid.c:168);
 *   - complex arrays are interleaved (re, im) doubles, the layout of
 *     fftw_complex (utils_fftw.c:18-30);
 *   - forward FFT uses e^{-i}, backward e^{+i}, both unnormalised
 *     (fft.c:19-57, Documentation.tex:103-111).
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mask kinds, one per reference mask builder */
enum {
    OR_MASK_DEFAULT = 0,        /* construct_kfilter               filter.c:34-63   */
    OR_MASK_DETERMINING = 1,    /* ..._with_determining_kbinwidth  filter.c:78-114  */
    OR_MASK_RANGE = 2,          /* ..._with_kbinwidth              filter.c:130-162 */
    OR_MASK_WATKINSON = 3,      /* construct_kfilter_Watkinson     filter.c:178-208 */
    OR_MASK_WATKINSON_KN = 4    /* construct_kfilter_Watkinson_kn  filter.c:219-250 */
};

/* 1D / 3D complex FFT, in place, unnormalised.  sign = -1 forward, +1 backward. */
int  or_fft1d(int n, double *data, int sign);
int  or_fft3d(int n, double *data, int sign);

/* promote a real float32/float64 grid to complex double (grid.c:121-228) */
void or_promote_f32(int64_t count, const float *in, double *out);
void or_promote_f64(int64_t count, const double *in, double *out);

/* z <- z/|z| over the whole grid (utils_fftw.c:78-93) */
void or_calc_phase(int nbins, double *array);

/* shell masks (filter.c); mask is N^3 bytes of 0/1.  kmin/kmax are read for
 * RANGE and WATKINSON_KN, written for DETERMINING. */
void or_kfilter(int nbins, int kind, double k, double *kmin, double *kmax,
                double binwidth, double boxsize, uint8_t *mask);
void or_get_corners(int mx, int my, int mz, double binwidth, double *mmin, double *mmax);
void or_calc_k3_min_and_max(const double *k, double *kmin, double *kmax);

/* k-configuration helpers (kvectors.c:172-247) */
double or_calc_k3(double k1, double k2, double cos_theta);
double or_calc_k3min(double k1, double k2, double theta, double dcos);
double or_calc_k3max(double k1, double k2, double theta, double dcos);

/* The estimator as polyspectrum.c:51-161 computes it: C2C FP64, complex
 * products, sequential complex sums.  fft_array is the full complex spectrum
 * (2*N^3 doubles).  k, kmin, kmax have n entries; kmin/kmax are updated in
 * place as the reference does for DERIVED binning.  sums_out (optional, 4
 * doubles) receives (re, im) of sum(prod D) and sum(prod I).
 * quirk_nan_kn: the reference's n=2 DERIVED leg goes through undefined
 * behaviour (filter.c:321-325); the oracle follows the intended meaning
 * kmin[n-1] = kmax[n-1] = k[n-1]. */
double or_polyspectrum(int nbins, const double *fft_array, int n, const double *k,
                       double *kmin, double *kmax, double kbinwidth, int kbinning_case,
                       double boxsize, int diagnostic_sums, double *sums_out);
double or_num_polygons(int nbins, int n, const double *k, double *kmin, double *kmax,
                       double kbinwidth, int kbinning_case, double boxsize);

/* Gradient of one polygon's value with respect to the real N^3 field:
 * grad += weight * d value / d field (value of or_polyspectrum on the
 * field's forward transform; phase = useOnlyPhase).  Returns the value;
 * kmin/kmax are updated as by or_polyspectrum. */
double or_polygon_gradient(int nbins, const double *field, int n, const double *k, double *kmin, double *kmax,
                           double kbinwidth, int kbinning_case, double boxsize, int phase, double weight,
                           double *grad);

/* Independent direct estimator: sum over lattice polygons m_1+..+m_n = 0 (mod N)
 * with m_i in mask_i of prod spec(m_i).  n = 2 or 3.  Returns the count of
 * closed polygons; sum_out receives (re, im) of the spectral product sum. */
int64_t or_direct_polygon_sum(int nbins, int n, const double *spec,
                              const uint8_t *mask1, const uint8_t *mask2,
                              const uint8_t *mask3, double *sum_out);

/* pieces of the per-triangle cost, for timing the reference algorithm */
void   or_product(int64_t count, const double *a, const double *b, double *out);
void   or_sum(int64_t count, const double *a, double *out2);

#ifdef __cplusplus
}
#endif
#endif

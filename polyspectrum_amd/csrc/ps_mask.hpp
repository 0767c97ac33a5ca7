// ps_mask.hpp -- k-shell membership tests, evaluated per k-cell on the fly.
//
// Each test restates one reference mask builder (filter.c) with the same FP64
// expression and operation order; the pragma keeps the compiler from fusing
// multiplies and adds so that membership is bit-identical to a -std=c99 gcc
// build of the reference.  The parameters are precomputed on the host with the
// reference's own expressions (k_in_kf = k / (2*PI) * boxsize, filter.c:36).
#pragma once
#include <hip/hip_runtime.h>

namespace ps {

enum MaskKind : int {
    MASK_DEFAULT = 0,       // construct_kfilter                     filter.c:34-63
    MASK_DETERMINING = 1,   // construct_kfilter_with_determining..  filter.c:78-114 (same membership)
    MASK_RANGE = 2,         // construct_kfilter_with_kbinwidth      filter.c:130-162
    MASK_WATKINSON = 3,     // construct_kfilter_Watkinson           filter.c:178-208
    MASK_WATKINSON_KN = 4   // construct_kfilter_Watkinson_kn        filter.c:219-250
};

struct MaskParams {
    int kind;
    int pad_;
    double kkf;   // k in units of k_F (DEFAULT, DETERMINING, WATKINSON)
    double mmin;  // kmin in units of k_F (RANGE, WATKINSON_KN)
    double mmax;  // kmax in units of k_F (RANGE, WATKINSON_KN)
    double bw;    // binwidth in cells
};

// index -> signed lattice coordinate (filter.c:44-55)
__host__ __device__ __forceinline__ int lattice(int i, int n) { return (i < n / 2) ? i : i - n; }

// the 8 squared cube-corner norms of filter.c:257-289, started from
// (mmin, mmax) = (1e30, 0) as the WATKINSON builders do (filter.c:201-203)
__host__ __device__ __forceinline__ void corner_extremes(int mx, int my, int mz, double bw, double &lo2, double &hi2)
{
#pragma clang fp contract(off)
    const double x = (double)mx, y = (double)my, z = (double)mz;
    const double e = 0.5 * bw;
    const double xm = (x - e) * (x - e), xp = (x + e) * (x + e);
    const double ym = (y - e) * (y - e), yp = (y + e) * (y + e);
    const double zm = (z - e) * (z - e), zp = (z + e) * (z + e);
    double c[8];
    c[0] = xm + ym + zm;
    c[1] = xp + ym + zm;
    c[2] = xm + yp + zm;
    c[3] = xm + ym + zp;
    c[4] = xp + yp + zm;
    c[5] = xp + ym + zp;
    c[6] = xm + yp + zp;
    c[7] = xp + yp + zp;
    double lo = 1.e30 * 1.e30, hi = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (c[i] > hi) hi = c[i];
        if (c[i] < lo) lo = c[i];
    }
    lo2 = lo;
    hi2 = hi;
}

__host__ __device__ __forceinline__ bool mask_eval(const MaskParams &p, int mx, int my, int mz)
{
#pragma clang fp contract(off)
    switch (p.kind) {
    case MASK_RANGE: {
        const double m = sqrt((double)(mx * mx + my * my + mz * mz));
        const double emin = p.mmin - m, emax = m - p.mmax;
        return (emin <= 0.5 * p.bw && emax <= 0.5 * p.bw);
    }
    case MASK_WATKINSON: {
        double lo2, hi2;
        corner_extremes(mx, my, mz, p.bw, lo2, hi2);
        const double pmin = sqrt(lo2), pmax = sqrt(hi2);
        return (pmin <= p.kkf && p.kkf <= pmax);
    }
    case MASK_WATKINSON_KN: {
        double lo2, hi2;
        corner_extremes(mx, my, mz, p.bw, lo2, hi2);
        const double pmin = sqrt(lo2), pmax = sqrt(hi2);
        return ((pmax >= p.mmax && pmin <= p.mmax) || (pmax <= p.mmax && pmin >= p.mmin) ||
                (pmax >= p.mmin && pmin <= p.mmin));
    }
    default: {
        const double m = sqrt((double)(mx * mx + my * my + mz * mz));
        return fabs(p.kkf - m) <= 0.5 * p.bw;
    }
    }
}

// mask_eval from the cell's norm m = sqrt(mx^2 + my^2 + mz^2) (int sum) and
// its cube-corner extremes pmin, pmax (corner_extremes, then sqrt; read by
// the WATKINSON kinds only), taken once for several masks of one binwidth
__device__ __forceinline__ bool mask_eval_norms(const MaskParams &p, double m, double pmin, double pmax)
{
#pragma clang fp contract(off)
    switch (p.kind) {
    case MASK_RANGE: {
        const double emin = p.mmin - m, emax = m - p.mmax;
        return (emin <= 0.5 * p.bw && emax <= 0.5 * p.bw);
    }
    case MASK_WATKINSON: return (pmin <= p.kkf && p.kkf <= pmax);
    case MASK_WATKINSON_KN:
        return ((pmax >= p.mmax && pmin <= p.mmax) || (pmax <= p.mmax && pmin >= p.mmin) ||
                (pmax >= p.mmin && pmin <= p.mmin));
    default: return fabs(p.kkf - m) <= 0.5 * p.bw;
    }
}

// DEFAULT, DETERMINING and RANGE membership depend on the cell only through
// the integer m2 = mx^2 + my^2 + mz^2 (the same int sum and FP64 square root
// as mask_eval), and monotonically: a shell is an interval [lo, hi] of m2.
__host__ __device__ __forceinline__ bool mask_is_radial(int kind)
{
    return kind != MASK_WATKINSON && kind != MASK_WATKINSON_KN;
}

__device__ __forceinline__ bool mask_eval_m2(const MaskParams &p, int m2)
{
#pragma clang fp contract(off)
    const double m = sqrt((double)m2);
    if (p.kind == MASK_RANGE) {
        const double emin = p.mmin - m, emax = m - p.mmax;
        return (emin <= 0.5 * p.bw && emax <= 0.5 * p.bw);
    }
    return fabs(p.kkf - m) <= 0.5 * p.bw;
}

}  // namespace ps

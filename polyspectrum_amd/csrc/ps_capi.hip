// ps_capi.hip -- host orchestration behind the C ABI (include/polyspectrum_gpu.h).
//
// Two execution shapes:
//  * ps_polyspectrum / ps_num_polygons: the reference's per-configuration
//    shape (polyspectrum.c:51-252): for each leg a full-grid mask and inverse
//    transform of the masked spectrum and of the mask, products into N^3 FP64
//    accumulators, then FP64 sums.  Used in FP64 as the parity path.
//  * ps_batch_*: every configuration of a run at once.  Shells are
//    de-duplicated into legs, the first inverse passes see only the planes
//    and band a shell can touch, and the last pass multiplies the legs of
//    every polygon of a group and reduces in FP64 without writing fields.
//    The launch sequence is captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "ps_internal.h"

using namespace ps;

// ---------------------------------------------------------------- errors
#define HIPCK(call)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "polyspectrum_gpu: %s failed: %s (%s:%d)\n", #call,              \
                    hipGetErrorString(e_), __FILE__, __LINE__);                              \
            return (e_ == hipErrorOutOfMemory) ? PS_ERR_NOMEM : PS_ERR_HIP;                  \
        }                                                                                    \
    } while (0)
#define PSCK(call)                       \
    do {                                 \
        int s_ = (call);                 \
        if (s_ != PS_OK) {               \
            if (s_ > 0) {                \
                fprintf(stderr, "polyspectrum_gpu: launch failed: %s (%s:%d)\n", \
                        hipGetErrorString((hipError_t)s_), __FILE__, __LINE__);  \
                return PS_ERR_HIP;       \
            }                            \
            return s_;                   \
        }                                \
    } while (0)

static double ref_pi() { return acos(-1.0); }  // the reference's PI (filter.c:19)

// ---------------------------------------------------------------- context
struct ps_ctx {
    int N = 0;
    double L = 0.0;
    int fp64 = 1;
    int nslots = 1;
    int device = 0;
    int cus = 256;
    hipStream_t st = nullptr;
    void *tw = nullptr;                  // forward twiddles exp(-2 pi i t/N), N entries
    std::map<int, void *> twm;           // twiddle tables of smaller transform grids
    std::vector<hipStream_t> br;         // branch streams: independent waves of a batch side by side
    std::vector<void *> spec;            // half spectra per slot
    std::vector<int> loaded;
    void *staging = nullptr;             // field upload buffer (N^3 doubles)
    int staging_is_double = 0;
    int staging_slot = -1;               // slot whose last load went through the upload buffer (-1: none)
    int fwd_variant = 0;                 // radix plan / tile width of the forward transform (ps_forward_autotune)
    void *fscratch = nullptr;            // half spectrum the forward profiler transforms into
    int phase_only = 0;
    int method = PS_METHOD_DIRECT;       // evaluation of batches created without explicit parameters
    // reference-shaped path buffers (lazily allocated)
    void *gd = nullptr, *gi = nullptr;   // half-spectrum sized work buffers
    double *accD = nullptr, *accI = nullptr;
    double2 *partials = nullptr;
    double *dsum = nullptr;              // 2 doubles
    LegDev *dleg = nullptr;
    double2 *dcorner = nullptr;          // corner-extreme partials
    // mixed precision: polygons whose error estimate exceeds refine_tol are
    // evaluated again from r64, an FP64 context holding the same fields
    double refine_tol = 0.0;
    ps_ctx *r64 = nullptr;
    size_t csize() const { return fp64 ? 16 : 8; }
    int64_t half_elems() const { return (int64_t)(N / 2 + 1) * N * N; }
    int64_t cells() const { return (int64_t)N * N * N; }
};

extern "C" const char *ps_strerror(int s)
{
    switch (s) {
    case PS_OK: return "ok";
    case PS_ERR_INVALID: return "invalid argument";
    case PS_ERR_UNSUPPORTED_N: return "unsupported grid size (N must be 2^a*3^b in 16..512)";
    case PS_ERR_HIP: return "HIP runtime error";
    case PS_ERR_NOMEM: return "out of memory";
    case PS_ERR_STATE: return "no field loaded in this slot";
    case PS_ERR_NO_DEVICE: return "no GPU available";
    case PS_ERR_SCRATCH: return "the shells' cells exceed the static scratch buffer (PS_DIRECT_SCRATCH_BYTES)";
    }
    return "unknown status";
}

extern "C" int ps_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int ps_grid_supported(int nbins) { return supported_grid(nbins); }

template <typename T> static int make_twiddles(int N, void **out)
{
    std::vector<T> h(2 * (size_t)N);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < N; ++t) {
        long double a = two_pi * (long double)t / (long double)N;
        h[2 * t] = (T)cosl(a);
        h[2 * t + 1] = (T)(-sinl(a));
    }
    HIPCK(hipMalloc(out, sizeof(T) * 2 * (size_t)N));
    HIPCK(hipMemcpy(*out, h.data(), sizeof(T) * 2 * (size_t)N, hipMemcpyHostToDevice));
    return PS_OK;
}

extern "C" int ps_ctx_destroy(ps_ctx *c)
{
    if (!c) return PS_ERR_INVALID;
    if (c->r64) ps_ctx_destroy(c->r64);
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void *p : c->spec) if (p) (void)hipFree(p);
    for (auto &kv : c->twm) (void)hipFree(kv.second);
    void *bufs[] = {c->tw,       c->staging,  c->gd,   c->gi,   c->accD,    c->accI,
                    c->partials, c->dsum,     c->dleg, c->dcorner, c->fscratch};
    for (void *p : bufs) if (p) (void)hipFree(p);
    for (hipStream_t b : c->br) {
        (void)hipStreamSynchronize(b);
        (void)hipStreamDestroy(b);
    }
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return PS_OK;
}

static void tune_load();
static int forward_variant_tuned(const ps_ctx *c);

extern "C" int ps_ctx_create(int nbins, double boxsize, int precision, int nslots, ps_ctx **out)
{
    if (!out || nslots < 1 || !(boxsize > 0.0)) return PS_ERR_INVALID;
    *out = nullptr;
    if (!supported_grid(nbins)) return PS_ERR_UNSUPPORTED_N;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PS_ERR_NO_DEVICE;
    ps_ctx *c = new ps_ctx();
    c->N = nbins;
    c->L = boxsize;
    c->fp64 = (precision == PS_PREC_FP64) ? 1 : 0;
    c->nslots = nslots;
    c->spec.assign(nslots, nullptr);
    c->loaded.assign(nslots, 0);
    int st = PS_OK;
    hipDeviceProp_t prop;
    if (hipGetDevice(&c->device) != hipSuccess || hipGetDeviceProperties(&prop, c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        ps_ctx_destroy(c);
        return PS_ERR_HIP;
    }
    c->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    c->fwd_variant = forward_variant_tuned(c);
    st = c->fp64 ? make_twiddles<double>(nbins, &c->tw) : make_twiddles<float>(nbins, &c->tw);
    if (st != PS_OK) {
        ps_ctx_destroy(c);
        return st;
    }
    *out = c;
    return PS_OK;
}

static int ensure_spec(ps_ctx *c, int slot)
{
    if (!c->spec[slot]) HIPCK(hipMalloc(&c->spec[slot], c->csize() * (size_t)c->half_elems()));
    return PS_OK;
}

static int forward(ps_ctx *c, int slot, const void *dfield, int is_double)
{
    PSCK(ensure_spec(c, slot));
    PSCK(run_forward(c->fp64, c->N, dfield, is_double, c->spec[slot], c->tw, c->st, c->fwd_variant));
    return PS_OK;
}

extern "C" int ps_load_field(ps_ctx *c, int slot, const void *field, int field_is_double, int on_device)
{
    if (!c || !field || slot < 0 || slot >= c->nslots) return PS_ERR_INVALID;
    HIPCK(hipSetDevice(c->device));
    const void *src = field;
    if (!on_device) {
        if (!c->staging) HIPCK(hipMalloc(&c->staging, sizeof(double) * (size_t)c->cells()));
        const size_t bytes = (field_is_double ? 8 : 4) * (size_t)c->cells();
        HIPCK(hipMemcpyAsync(c->staging, field, bytes, hipMemcpyHostToDevice, c->st));
        src = c->staging;
        c->staging_is_double = field_is_double ? 1 : 0;
    }
    // the upload buffer holds this slot's field only after a host upload
    c->staging_slot = on_device ? (c->staging_slot == slot ? -1 : c->staging_slot) : -1;
    PSCK(forward(c, slot, src, field_is_double));
    HIPCK(hipStreamSynchronize(c->st));
    c->loaded[slot] = 1;
    if (!on_device) c->staging_slot = slot;
    // the FP64 spectrum of the same field for refinement (src is on the device)
    if (c->r64) PSCK(ps_load_field(c->r64, slot, src, field_is_double, 1));
    return PS_OK;
}

extern "C" int ps_forward_again(ps_ctx *c, int slot)
{
    if (!c || slot < 0 || slot >= c->nslots) return PS_ERR_INVALID;
    if (!c->staging || !c->loaded[slot] || c->staging_slot != slot) return PS_ERR_STATE;
    HIPCK(hipSetDevice(c->device));
    PSCK(forward(c, slot, c->staging, c->staging_is_double));
    if (c->r64) {
        HIPCK(hipStreamSynchronize(c->st));
        PSCK(ps_load_field(c->r64, slot, c->staging, c->staging_is_double, 1));
    }
    return PS_OK;
}

extern "C" int ps_get_spectrum(ps_ctx *c, int slot, void *out)
{
    if (!c || !out || slot < 0 || slot >= c->nslots) return PS_ERR_INVALID;
    if (!c->loaded[slot]) return PS_ERR_STATE;
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipMemcpyAsync(out, c->spec[slot], c->csize() * (size_t)c->half_elems(), hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    return PS_OK;
}

extern "C" int ps_set_method(ps_ctx *c, int method)
{
    if (!c || (method != PS_METHOD_DIRECT && method != PS_METHOD_FOURIER)) return PS_ERR_INVALID;
    c->method = method;
    return PS_OK;
}

extern "C" int ps_set_phase_only(ps_ctx *c, int phase_only)
{
    if (!c) return PS_ERR_INVALID;
    c->phase_only = (phase_only == 1) ? 1 : 0;  // run.c:48: only the value 1 switches it on
    if (c->r64) c->r64->phase_only = c->phase_only;
    return PS_OK;
}

extern "C" int ps_set_refinement(ps_ctx *c, double tol)
{
    if (!c || !(tol >= 0.0)) return PS_ERR_INVALID;
    if (c->fp64) return (tol > 0.0) ? PS_ERR_INVALID : PS_OK;  // FP64 contexts need none
    HIPCK(hipSetDevice(c->device));
    if (tol > 0.0 && !c->r64) {
        PSCK(ps_ctx_create(c->N, c->L, PS_PREC_FP64, c->nslots, &c->r64));
        c->r64->phase_only = c->phase_only;
        // the field still in the upload buffer; other slots once they are loaded again
        if (c->staging && c->staging_slot >= 0 && c->loaded[c->staging_slot])
            PSCK(ps_load_field(c->r64, c->staging_slot, c->staging, c->staging_is_double, 1));
    }
    c->refine_tol = tol;
    return PS_OK;
}

// ---------------------------------------------------------------- masks (host side)
static MaskParams mp_shell(int kind, double k, double bw, double L)
{
    MaskParams p{};
    p.kind = kind;
    p.kkf = k / (2. * ref_pi()) * L;  // filter.c:36
    p.bw = bw;
    return p;
}
static MaskParams mp_range(int kind, double kmin, double kmax, double bw, double L)
{
    MaskParams p{};
    p.kind = kind;
    p.mmin = kmin / (2. * ref_pi()) * L;  // filter.c:137-138
    p.mmax = kmax / (2. * ref_pi()) * L;
    p.bw = bw;
    return p;
}

// largest |m_i| a cell selected by p can have, capped at N/2: every test
// implies |m_i| <= r (a cube-corner norm is at least max_i |m_i| - bw/2); the
// 1e-9 relative margin covers the rounding of the tests' FP64 expressions
static int shell_band(const MaskParams &p, int N)
{
    double r;
    switch (p.kind) {
    case MASK_RANGE: r = p.mmax; break;
    case MASK_WATKINSON_KN: r = std::max(p.mmin, p.mmax); break;
    default: r = p.kkf;
    }
    r += 0.5 * fabs(p.bw);
    if (!(r < (double)N)) return N / 2;
    if (r < 0.0) return 0;
    const int K = (int)floor(r + 1e-9 * std::max(1.0, r));
    return std::min(K, N / 2);
}

// DERIVED binning: kmin/kmax of a shell from its cells' cube corners
// (construct_kfilter_with_determining_kbinwidth filter.c:78-114)
constexpr int kCornerBlocks = 256;  // partial extremes of the corner reduction

// corner: kCornerBlocks device partials, or nullptr for the context's own buffer
static int derived_kminmax(ps_ctx *c, double k, double bw, double *kmin, double *kmax, double2 *corner = nullptr)
{
    MaskParams mp = mp_shell(MASK_DETERMINING, k, bw, c->L);
    const int K = shell_band(mp, c->N);
    const int nblocks = kCornerBlocks;
    if (!corner) {
        if (!c->dcorner) HIPCK(hipMalloc(&c->dcorner, sizeof(double2) * nblocks));
        corner = c->dcorner;
    }
    PSCK(run_corner_extremes(mp, c->N, K, bw, corner, nblocks, c->st));
    double2 h[kCornerBlocks];
    HIPCK(hipMemcpyAsync(h, corner, sizeof(double2) * nblocks, hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    double lo = 1.e30 * 1.e30, hi = 0.0;
    for (auto &v : h) {
        lo = std::min(lo, v.x);
        hi = std::max(hi, v.y);
    }
    double mmin = 1.e30, mmax = 0.;  // filter.c:85-86 (unchanged when no cell is selected)
    if (hi > 0.0) {
        mmin = sqrt(lo);
        mmax = sqrt(hi);
    }
    *kmin = mmin * 2. * ref_pi() / c->L;  // filter.c:112-113
    *kmax = mmax * 2. * ref_pi() / c->L;
    return PS_OK;
}

// filter.c:291-319
static void k3_min_and_max(const double *k, double *kmin, double *kmax)
{
    double ct = (k[0] * k[0] + k[1] * k[1] - k[2] * k[2]) / (2. * k[0] * k[1]);
    double t[8];
    t[0] = kmin[0] * kmin[0] + kmin[1] * kmin[1] - 2. * kmin[0] * kmin[1] * ct;
    t[1] = kmin[0] * kmin[0] + kmax[1] * kmax[1] - 2. * kmin[0] * kmax[1] * ct;
    t[2] = kmax[0] * kmax[0] + kmin[1] * kmin[1] - 2. * kmax[0] * kmin[1] * ct;
    t[3] = kmax[0] * kmax[0] + kmax[1] * kmax[1] - 2. * kmax[0] * kmax[1] * ct;
    t[4] = k[0] * k[0] + kmax[1] * kmax[1] - 2. * k[0] * kmax[1] * ct;
    t[5] = kmax[0] * kmax[0] + k[1] * k[1] - 2. * kmax[0] * k[1] * ct;
    t[6] = k[0] * k[0] + kmin[1] * kmin[1] - 2. * k[0] * kmin[1] * ct;
    t[7] = kmin[0] * kmin[0] + k[1] * k[1] - 2. * kmin[0] * k[1] * ct;
    double lo = 1.e30, hi = 0.;
    for (int i = 0; i < 8; i++) {
        if (t[i] < lo) lo = t[i];
        if (t[i] > hi) hi = t[i];
    }
    kmin[2] = sqrt(lo);
    kmax[2] = sqrt(hi);
}

// the per-leg mask choice of polyspectrum.c:79-118; may update kmin/kmax
static int leg_mask(ps_ctx *c, int i, int n, const double *k, double *kmin, double *kmax, double bw, int kcase,
                    MaskParams *out, double2 *corner = nullptr)
{
    const double L = c->L;
    switch (kcase) {
    case PS_BINNING_DERIVED:
        if (i < n - 1) {
            PSCK(derived_kminmax(c, k[i], bw, &kmin[i], &kmax[i], corner));
            *out = mp_shell(MASK_DETERMINING, k[i], bw, L);
        } else {
            if (n == 3) {
                k3_min_and_max(k, kmin, kmax);
            } else {
                // filter.c:321-325 writes through a mis-indexed pointer; its
                // intended effect is kmin[n-1] = kmax[n-1] = k[n-1]
                kmin[n - 1] = k[n - 1];
                kmax[n - 1] = k[n - 1];
            }
            *out = mp_range(MASK_RANGE, kmin[i], kmax[i], bw, L);
        }
        break;
    case PS_BINNING_GIVEN:
        *out = (i < n - 1) ? mp_shell(MASK_DEFAULT, k[i], bw, L) : mp_range(MASK_RANGE, kmin[i], kmax[i], bw, L);
        break;
    case PS_BINNING_WATKINSON:
        *out = (i < n - 1) ? mp_shell(MASK_WATKINSON, k[i], bw, L)
                           : mp_range(MASK_WATKINSON_KN, kmin[i], kmax[i], bw, L);
        break;
    default:
        *out = mp_shell(MASK_DEFAULT, k[i], bw, L);
    }
    return PS_OK;
}

// ---------------------------------------------------------------- reference-shaped path
static int ensure_plain_buffers(ps_ctx *c)
{
    const size_t hb = c->csize() * (size_t)c->half_elems();
    if (!c->gd) HIPCK(hipMalloc(&c->gd, hb));
    if (!c->gi) HIPCK(hipMalloc(&c->gi, hb));
    if (!c->accD) HIPCK(hipMalloc(&c->accD, sizeof(double) * (size_t)c->cells()));
    if (!c->accI) HIPCK(hipMalloc(&c->accI, sizeof(double) * (size_t)c->cells()));
    const int64_t nb = (int64_t)c->N * c->N / tile_rows(c->N, c->fp64);
    if (!c->partials) HIPCK(hipMalloc(&c->partials, sizeof(double2) * 2 * (size_t)nb));
    if (!c->dsum) HIPCK(hipMalloc(&c->dsum, sizeof(double) * 2));
    if (!c->dleg) HIPCK(hipMalloc(&c->dleg, sizeof(LegDev)));
    return PS_OK;
}

// one leg of the reference-shaped path: full (unpruned) inverse transforms
static int plain_leg(ps_ctx *c, const MaskParams &mp, const void *spec, int mode)
{
    const int N = c->N;
    LegDev L{};
    L.mask = mp;
    L.spec = spec;
    L.phase_only = c->phase_only;
    L.Kz = N / 2;
    L.Kx = N / 2;
    L.Bx = N;
    L.ns = N;
    L.g1_begin = 0;
    L.g2_begin = 0;
    L.g1d = spec ? c->gd : nullptr;
    L.g1i = c->gi;
    L.g2d = L.g1d;
    L.g2i = c->gi;
    HIPCK(hipMemcpyAsync(c->dleg, &L, sizeof(LegDev), hipMemcpyHostToDevice, c->st));
    const int tl = g_tile_lines(N, c->fp64);
    const int64_t g1 = (int64_t)(N / 2 + 1) * ((N + tl - 1) / tl);
    const int64_t g2 = 2 * (int64_t)(N / 2 + 1) * (N / g2_tile_cols(N, c->fp64, 1));
    PSCK(run_inverse_g(c->fp64, N, c->dleg, 1, 0, g1, 0, g2, 1 << 30, 1, 0, c->tw, c->st));
    PSCK(run_zc_plain(c->fp64, N, L.g2d, L.g2i, N / 2, spec ? c->accD : nullptr, c->accI, mode, c->partials, c->tw,
                      c->st));
    // the next leg overwrites the host-side descriptor buffer: keep order
    HIPCK(hipStreamSynchronize(c->st));
    return PS_OK;
}

// sumI receives sum prod I, or with want_abs (D parts only) sum |prod D|
static int plain_run(ps_ctx *c, int slot, int n, const double *k, double *kmin, double *kmax, double bw, int kcase,
                     double *sumD, double *sumI, int want_abs = 0)
{
    PSCK(ensure_plain_buffers(c));
    const void *spec = (slot >= 0) ? c->spec[slot] : nullptr;
    // accumulators start at 1 (polyspectrum.c:66-70)
    PSCK(run_fill(c->accI, c->cells(), 1.0, c->st));
    if (spec) PSCK(run_fill(c->accD, c->cells(), 1.0, c->st));
    for (int i = 0; i < n; ++i) {
        MaskParams mp;
        PSCK(leg_mask(c, i, n, k, kmin, kmax, bw, kcase, &mp));
        PSCK(plain_leg(c, mp, spec, (i < n - 1) ? 1 : (want_abs ? 3 : 2)));
    }
    const int64_t nb = (int64_t)c->N * c->N / tile_rows(c->N, c->fp64);
    PSCK(run_reduce_partials(c->partials, 0, nullptr, nb, 1, c->dsum, c->dsum + 1, c->st));
    double h[2];
    HIPCK(hipMemcpyAsync(h, c->dsum, sizeof(h), hipMemcpyDeviceToHost, c->st));
    HIPCK(hipStreamSynchronize(c->st));
    if (sumD) *sumD = h[0];
    *sumI = h[1];
    return PS_OK;
}

extern "C" int ps_polyspectrum(ps_ctx *c, int slot, int n, const double *k, double *kmin, double *kmax,
                               double kbinwidth, int kbinning_case, double *result, double *sums)
{
    if (!c || !k || !kmin || !kmax || !result || n < 2 || n > 3) return PS_ERR_INVALID;
    if (slot < 0 || slot >= c->nslots) return PS_ERR_INVALID;
    if (!c->loaded[slot]) return PS_ERR_STATE;
    HIPCK(hipSetDevice(c->device));
    double sD = 0, sI = 0;
    PSCK(plain_run(c, slot, n, k, kmin, kmax, kbinwidth, kbinning_case, &sD, &sI));
    const double volume = c->L * c->L * c->L;
    *result = pow(volume, n - 1) * pow(c->N, -n * 3) * sD / sI;  // polyspectrum.c:155
    if (sums) {
        sums[0] = sD;
        sums[1] = sI;
    }
    return PS_OK;
}

extern "C" int ps_condition_numbers(ps_ctx *c, const ps_polygon *polys, int npoly, double kbinwidth,
                                    int kbinning_case, double *kappa)
{
    if (!c || !polys || npoly < 1 || !kappa) return PS_ERR_INVALID;
    for (int p = 0; p < npoly; ++p) {
        if (polys[p].n < 2 || polys[p].n > 3 || polys[p].slot < 0 || polys[p].slot >= c->nslots)
            return PS_ERR_INVALID;
        if (!c->loaded[polys[p].slot]) return PS_ERR_STATE;
    }
    HIPCK(hipSetDevice(c->device));
    for (int p = 0; p < npoly; ++p) {
        ps_polygon q = polys[p];  // DERIVED binning updates kmin/kmax of the copy
        double sD = 0.0, sA = 0.0;
        PSCK(plain_run(c, q.slot, q.n, q.k, q.kmin, q.kmax, kbinwidth, kbinning_case, &sD, &sA, 1));
        kappa[p] = sA / fabs(sD);
    }
    return PS_OK;
}

extern "C" int ps_num_polygons(ps_ctx *c, int n, const double *k, double *kmin, double *kmax, double kbinwidth,
                               int kbinning_case, double *result)
{
    if (!c || !k || !kmin || !kmax || !result || n < 2 || n > 3) return PS_ERR_INVALID;
    HIPCK(hipSetDevice(c->device));
    double sI = 0;
    PSCK(plain_run(c, -1, n, k, kmin, kmax, kbinwidth, kbinning_case, nullptr, &sI));
    *result = sI / (double)((int64_t)c->N * c->N * c->N);  // polyspectrum.c:247
    return PS_OK;
}

// ---------------------------------------------------------------- batched path
// The groups of a batch run in waves: one transform grid M per wave (the
// smallest alias-free grid of its groups), its legs transformed at M, its
// buffers in the batch's arena (disjoint per wave when they all fit, so that
// the waves run side by side on branch streams; else one region reused in turn).
struct Wave {
    int M = 0;
    const void *tw = nullptr;            // twiddle table of M
    std::vector<LegDev> legs;            // legs transformed in this wave (Fourier: prefix legs only)
    std::vector<size_t> leg_off;         // arena offsets of their G1/G2 buffers
    LegDev *dlegs = nullptr;
    int64_t g1_tiles = 0, g2_tiles = 0;
    // direct method, D parts only (dmode): the legs with D-only tiles, and the
    // legs without their spectrum (I parts) for the sums taken at planning
    int dmode = 0;
    LegDev *dlegs_d = nullptr, *dlegs_i = nullptr;
    int64_t g1_tiles_d = 0, g2_tiles_d = 0;
    std::vector<int64_t> g1b_d, g2b_d;   // per leg: first D-only G1 / G2 tile
    int kz_cap = 1;                      // 1 + largest Kz of the wave's legs
    size_t bytes = 0;
    // direct method
    std::vector<GroupDev> groups;        // leg indices local to the wave
    std::vector<int> suffix;
    GroupDev *dgroups = nullptr;
    int *dsuffix = nullptr;
    int64_t items = 0;                   // Zc work items
    // Fourier method
    std::vector<SpecGroupDev> sgroups;
    std::vector<size_t> grp_off;         // arena offsets of the groups' Zf/Fx buffers
    SpecGroupDev *dsgroups = nullptr;
    int64_t zf_tiles = 0, fx_tiles = 0, dot_items = 0;
    int kinds = 0;                       // dot pass instantiations needed: 1 radial, 2 cube-corner masks
    double cost = 0.0;                   // model bytes of one execution (balances the branch streams)
    int branch = 0;                      // branch stream when waves run side by side
};

struct ps_batch {
    ps_ctx *c = nullptr;
    int npoly = 0;
    int method = PS_METHOD_DIRECT;
    std::vector<int> poly_n;
    std::vector<int> poly_out;           // output slot of each polygon
    int nout = 0;
    std::vector<LegDev> legs;            // host table of the de-duplicated legs (masks, bands)
    std::vector<GroupDev> groups;        // host table of the groups (global leg ids)
    std::vector<int> suffix;
    int max_chunks = 1;
    int grid_cap = 2048;
    int g2_div = 1;                      // G2 (and Fx) tile width divisor (1 or 2)
    int chunk_tiles = 64;                // Zc tiles (direct) or dot steps (Fourier) per work item
    int overlap = 1;                     // independent waves may run side by side
    std::vector<void *> allocs;
    int *dcounts = nullptr;
    double *dscale = nullptr;            // per wave-leg amplitude scales
    double2 *dpartials = nullptr;
    double *dsum = nullptr;              // 3*nout: D sums, I sums, |prod D| sums
    double *hsum = nullptr;              // pinned host copy of dsum
    std::vector<Wave> waves;
    MaskParams *dmasks = nullptr;        // suffix masks in suffix order (Fourier method)
    int2 *divl = nullptr;                // their m^2 intervals and counts
    void *arena = nullptr;
    size_t arena_bytes = 0;
    int concurrent = 0;                  // waves have disjoint buffers and run on branch streams
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> wend;        // per branch stream: its waves are done
    std::vector<int> rec_order;          // waves in capture order when side by side
    void *adj[9] = {};                   // adjoint work buffers, kept after the first ps_batch_adjoint
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    double model_bytes = 0.0;
    std::vector<int> leg_slot;           // field slot of every leg
    std::vector<double> last_sumI;       // sum prod I of every output (last execute)
    int have_sums = 0;
    // mixed-precision refinement
    double *dabs = nullptr;              // sum |prod D| of every output: dsum + 2 nout
    int abs_rec = 0;                     // the captured sequence takes those sums
    std::vector<ps_polygon> polys;       // the polygons as planned
    double bw = 0.0;
    int kcase = 0;
    std::map<std::vector<int>, ps_batch *> ref;  // FP64 batches of refined polygon sets (planned on first use)
    std::vector<std::vector<int>> ref_order;      // their keys, oldest first (at most kRefineSets kept)
    std::vector<int> last_flagged;                // the set the last results refined
    std::vector<int> refined;            // per polygon: evaluated again in FP64 by the last results
    std::vector<double> err_est;         // per polygon: error estimate of the mixed value (last results)
    int nstages = 4;
    double stage_bytes[PS_MAX_STAGES] = {};
};

static const char *const kStageDirect[] = {"G1 masked gather + ky inverse", "G2 kx inverse (band -> all x)",
                                           "Zc fused C2R + product-reduce", "final reduce"};
static const char *const kStageFourier[] = {"G1 masked gather + ky inverse (prefix legs)",
                                            "G2 kx inverse (band -> all x)",
                                            "Zf C2R + product + forward z",
                                            "Fx forward x (all x -> band)",
                                            "dot forward y + shell sums",
                                            "final reduce"};

extern "C" int ps_batch_destroy(ps_batch *b)
{
    if (!b) return PS_ERR_INVALID;
    for (auto &kv : b->ref)
        if (kv.second) ps_batch_destroy(kv.second);
    if (b->c && b->c->st) (void)hipStreamSynchronize(b->c->st);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (hipEvent_t e : b->wend) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (void *p : b->allocs) (void)hipFree(p);
    for (Wave &w : b->waves) {
        void *wb[] = {w.dlegs, w.dgroups, w.dsuffix, w.dsgroups, w.dlegs_d, w.dlegs_i};
        for (void *p : wb) if (p) (void)hipFree(p);
    }
    void *bufs[] = {b->dcounts, b->dpartials, b->dsum, b->dscale, b->dmasks, b->divl, b->arena};
    if (b->hsum) (void)hipHostFree(b->hsum);
    for (void *p : bufs) if (p) (void)hipFree(p);
    for (void *p : b->adj) if (p) (void)hipFree(p);
    delete b;
    return PS_OK;
}

static bool same_mask(const MaskParams &a, const MaskParams &b)
{
    return a.kind == b.kind && a.kkf == b.kkf && a.mmin == b.mmin && a.mmax == b.mmax && a.bw == b.bw;
}

// The passes of one wave on stream ws: the inverse transforms of its legs,
// then (direct) the fused product-reduce, or (Fourier) Zf, Fx and dot.
// Direct waves in dmode transform the D parts only.
static int record_wave(ps_batch *b, Wave &w, hipStream_t ws)
{
    ps_ctx *c = b->c;
    const int M = w.M, nl = (int)w.legs.size();
    if (b->method == PS_METHOD_DIRECT && w.dmode) {
        PSCK(run_inverse_g(c->fp64, M, w.dlegs_d, nl, 0, w.g1_tiles_d, 0, w.g2_tiles_d, b->grid_cap, b->g2_div, 1,
                           w.tw, ws));
        // with the magnitude sums of the products when the batch is refined
        PSCK(run_zc_fused(c->fp64, M, w.dgroups, (int)w.groups.size(), w.dlegs_d, w.dsuffix, 0, w.items,
                          b->dpartials, b->max_chunks, b->grid_cap, w.kz_cap, b->abs_rec ? 3 : 1, w.tw, ws));
        return PS_OK;
    }
    if (nl > 0) {
        if (b->dscale) PSCK(run_leg_scale(c->fp64, M, w.dlegs, nl, ws));
        PSCK(run_inverse_g(c->fp64, M, w.dlegs, nl, 0, w.g1_tiles, 0, w.g2_tiles, b->grid_cap, b->g2_div, 0, w.tw,
                           ws));
    }
    if (b->method == PS_METHOD_FOURIER) {
        const int ng = (int)w.sgroups.size();
        PSCK(run_spec_zf(c->fp64, M, w.dsgroups, ng, w.dlegs, 0, w.zf_tiles, b->grid_cap, w.kz_cap, w.tw, ws));
        PSCK(run_spec_fx(c->fp64, M, w.dsgroups, ng, 0, w.fx_tiles, b->grid_cap, b->g2_div, w.tw, ws));
        PSCK(run_spec_dot(c->fp64, M, w.dsgroups, ng, b->dmasks, b->divl, 0, w.dot_items, b->dpartials,
                          b->max_chunks, b->grid_cap, w.kinds, w.tw, ws));
    } else {
        PSCK(run_zc_fused(c->fp64, M, w.dgroups, (int)w.groups.size(), w.dlegs, w.dsuffix, 0, w.items,
                          b->dpartials, b->max_chunks, b->grid_cap, w.kz_cap, 0, w.tw, ws));
    }
    return PS_OK;
}

// Mixed precision: the I sums taken with FP64 transforms of the bare masks in
// the batch's arena (before its first use by an execution), wave by wave, so
// that the polygon counts are exact integers as the reference's (FP32
// transforms leave a relative error of about 1e-7: a few polygons in 3.5e7).  The FP64 tiles and work items are numbered
// here; the I sums go to dsum[nout ..].
static int batch_i_sums_fp64(ps_batch *b)
{
    ps_ctx *c = b->c;
    const size_t cs = sizeof(double) * 2;
    int st = PS_OK;
    void *partials = nullptr, *dcounts = nullptr, *dsumD = nullptr;
    // work items per group in the FP64 numbering, and per output its count
    std::vector<int> counts(b->nout, 0);
    int max_chunks = 1;
    for (Wave &w : b->waves) {
        const int M = w.M;
        const int64_t tiles = (int64_t)M * (M / fused_tile_rows(M, 1));
        const int nchunks = (int)((tiles + b->chunk_tiles - 1) / b->chunk_tiles);
        for (const GroupDev &G : w.groups)
            for (int o = 0; o < G.nsuf; ++o) counts[G.out_begin + o] = nchunks;
        max_chunks = std::max(max_chunks, nchunks);
    }
    if (hipMalloc(&partials, sizeof(double2) * 2 * (size_t)b->nout * max_chunks) != hipSuccess ||
        hipMalloc(&dcounts, sizeof(int) * counts.size()) != hipSuccess ||
        hipMalloc(&dsumD, sizeof(double) * (size_t)b->nout) != hipSuccess ||
        hipMemcpy(dcounts, counts.data(), sizeof(int) * counts.size(), hipMemcpyHostToDevice) != hipSuccess)
        st = PS_ERR_NOMEM;
    for (size_t wi = 0; wi < b->waves.size() && st == PS_OK; ++wi) {
        const Wave &w = b->waves[wi];
        const int M = w.M, nl = (int)w.legs.size(), ng = (int)w.groups.size();
        void *tw = nullptr, *arena = nullptr, *own = nullptr, *dlegs = nullptr, *dgroups = nullptr;
        std::vector<LegDev> legs(w.legs);
        std::vector<GroupDev> groups(w.groups);
        const int tl = g_tile_lines(M, 1);
        const int64_t g2_per_kz = M / g2_tile_cols(M, 1, 1);
        int64_t g1 = 0, g2 = 0, items = 0;
        size_t bytes = 0;
        std::vector<size_t> off(nl);
        for (int i = 0; i < nl; ++i) {
            LegDev &L = legs[i];
            const int K = L.Kz;
            L.g1_begin = g1;
            L.g2_begin = g2;
            g1 += (int64_t)(K + 1) * ((L.Bx + tl - 1) / tl);
            g2 += 2 * (int64_t)(K + 1) * g2_per_kz;
            off[i] = bytes;
            bytes += cs * ((size_t)(K + 1) * L.Bx * M + (size_t)(K + 1) * M * M);
        }
        const int64_t tiles = (int64_t)M * (M / fused_tile_rows(M, 1));
        const int nchunks = (int)((tiles + b->chunk_tiles - 1) / b->chunk_tiles);
        for (GroupDev &G : groups) {
            G.item_begin = items;
            items += nchunks;
        }
        // the wave's FP64 I buffers take as many bytes as its mixed D and I
        // buffers: they go into the batch's arena, unused until the first
        // execution (freeing a temporary buffer of 1.9 GB took up to 51 ms)
        if (bytes <= b->arena_bytes) arena = b->arena;
        else if (hipMalloc(&own, bytes) == hipSuccess) arena = own;
        if (make_twiddles<double>(M, &tw) != PS_OK || !arena ||
            hipMalloc(&dlegs, sizeof(LegDev) * std::max(nl, 1)) != hipSuccess ||
            hipMalloc(&dgroups, sizeof(GroupDev) * std::max(ng, 1)) != hipSuccess) {
            st = PS_ERR_NOMEM;
        } else {
            for (int i = 0; i < nl; ++i) {
                LegDev &L = legs[i];
                const size_t n1 = (size_t)(L.Kz + 1) * L.Bx * M;
                char *p = (char *)arena + off[i];
                L.spec = nullptr;
                L.g1d = L.g2d = nullptr;
                L.g1i = p;
                L.g2i = p + cs * n1;
                L.scale = nullptr;
            }
            if (hipMemcpy(dlegs, legs.data(), sizeof(LegDev) * nl, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dgroups, groups.data(), sizeof(GroupDev) * ng, hipMemcpyHostToDevice) != hipSuccess)
                st = PS_ERR_HIP;
            if (st == PS_OK)
                st = run_inverse_g(1, M, (const LegDev *)dlegs, nl, 0, g1, 0, g2, b->grid_cap, 1, 0, tw, c->st);
            if (st == PS_OK)
                st = run_zc_fused(1, M, (const GroupDev *)dgroups, ng, (const LegDev *)dlegs, w.dsuffix, 0, items,
                                  (double2 *)partials, max_chunks, b->grid_cap, w.kz_cap, 0, tw, c->st);
            if (st > 0) st = PS_ERR_HIP;
            if (hipStreamSynchronize(c->st) != hipSuccess && st == PS_OK) st = PS_ERR_HIP;
        }
        void *bufs[] = {tw, own, dlegs, dgroups};
        for (void *q : bufs)
            if (q) (void)hipFree(q);
    }
    if (st == PS_OK) {
        st = run_reduce_partials((const double2 *)partials, max_chunks, (const int *)dcounts, 0, b->nout,
                                 (double *)dsumD, b->dsum + b->nout, c->st);
        if (st > 0) st = PS_ERR_HIP;
        if (hipStreamSynchronize(c->st) != hipSuccess && st == PS_OK) st = PS_ERR_HIP;
    }
    void *bufs[] = {partials, dcounts, dsumD};
    for (void *q : bufs)
        if (q) (void)hipFree(q);
    return st;
}

// Direct method: the I sums (polygon counts, field independent) of every
// group, from the legs' bare masks, once when the batch is planned (in FP64
// for mixed-precision contexts too).
static int batch_i_sums(ps_batch *b)
{
    ps_ctx *c = b->c;
    if (!c->fp64) return batch_i_sums_fp64(b);
    for (Wave &w : b->waves) {
        const int nl = (int)w.legs.size();
        PSCK(run_inverse_g(c->fp64, w.M, w.dlegs_i, nl, 0, w.g1_tiles, 0, w.g2_tiles, b->grid_cap, b->g2_div, 0,
                           w.tw, c->st));
        PSCK(run_zc_fused(c->fp64, w.M, w.dgroups, (int)w.groups.size(), w.dlegs_i, w.dsuffix, 0, w.items,
                          b->dpartials, b->max_chunks, b->grid_cap, w.kz_cap, 0, w.tw, c->st));
    }
    PSCK(run_reduce_partials(b->dpartials, b->max_chunks, b->dcounts, 0, b->nout, b->dsum, b->dsum + b->nout,
                             c->st));
    HIPCK(hipStreamSynchronize(c->st));
    return PS_OK;
}

// The launch sequence of one batch (captured into graphs): the waves one
// after another on s, or forked onto the branch streams (each wave on the
// branch the planner gave it, in the planner's capture order) and joined,
// then the final reduction on s.
static int batch_record_on(ps_batch *b, hipStream_t s)
{
    ps_ctx *c = b->c;
    const int nb = (int)b->wend.size();
    if (b->concurrent && nb > 1) {
        HIPCK(hipEventRecord(b->fork, s));
        for (int k = 0; k < nb; ++k) HIPCK(hipStreamWaitEvent(c->br[k], b->fork, 0));
        for (int i : b->rec_order) PSCK(record_wave(b, b->waves[i], c->br[b->waves[i].branch]));
        for (int k = 0; k < nb; ++k) {
            HIPCK(hipEventRecord(b->wend[k], c->br[k]));
            HIPCK(hipStreamWaitEvent(s, b->wend[k], 0));
        }
    } else {
        for (Wave &w : b->waves) PSCK(record_wave(b, w, s));
    }
    // direct method: the I sums were taken when the batch was planned (the
    // second partial of each output is sum |prod D| when refined, else unused)
    double *outI = (b->method == PS_METHOD_DIRECT) ? (b->abs_rec ? b->dabs : nullptr) : b->dsum + b->nout;
    PSCK(run_reduce_partials(b->dpartials, b->max_chunks, b->dcounts, 0, b->nout, b->dsum, outI, s));
    return PS_OK;
}
static int batch_record(ps_batch *b) { return batch_record_on(b, b->c->st); }

// (re)capture the launch sequence into the batch's graph (PS_NO_GRAPH: none,
// executions record the sequence each time)
static int batch_capture(ps_batch *b)
{
    ps_ctx *c = b->c;
    if (b->exec) HIPCK(hipGraphExecDestroy(b->exec));
    if (b->graph) HIPCK(hipGraphDestroy(b->graph));
    b->exec = nullptr;
    b->graph = nullptr;
    if (getenv("PS_NO_GRAPH")) return PS_OK;
    HIPCK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    const int st = batch_record(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->st, &g);
    if (st != PS_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return st != PS_OK ? st : PS_ERR_HIP;
    }
    b->graph = g;
    HIPCK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return PS_OK;
}

// refinement needs the magnitude sums: switch them on (or off) with the
// context's setting, re-capturing the sequence
static int batch_sync_refinement(ps_batch *b)
{
    ps_ctx *c = b->c;
    const int want = (c->refine_tol > 0.0 && !c->fp64 && b->method == PS_METHOD_DIRECT) ? 1 : 0;
    if (want == b->abs_rec) return PS_OK;
    if (want) b->dabs = b->dsum + 2 * (size_t)b->nout;
    b->abs_rec = want;
    return batch_capture(b);
}

// ---------------------------------------------------------------- launch-shape tuning
// Best launch parameters found by ps_batch_autotune, per batch shape; kept for
// the process and, when PS_TUNE_FILE names a file, appended to / read from it.
static std::map<std::string, ps_launch_params> g_tuned;
static std::map<std::string, int> g_fwd_tuned;  // forward-transform variant per (N, precision, CUs)
static bool g_tune_loaded = false;

static std::string fwd_key(const ps_ctx *c)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "fwd_N%d_fp%d_cu%d", c->N, c->fp64, c->cus);
    return buf;
}

static std::string tune_key(const ps_ctx *c, const ps_polygon *polys, int npoly, double bw, int kcase)
{
    double ks = 0.0;
    int slots = 0;
    for (int p = 0; p < npoly; ++p) {
        for (int i = 0; i < polys[p].n; ++i) ks += polys[p].k[i] + polys[p].kmin[i] + polys[p].kmax[i];
        slots = std::max(slots, polys[p].slot + 1);
    }
    char buf[256];
    snprintf(buf, sizeof(buf), "N%d_fp%d_m%d_cu%d_np%d_kc%d_sl%d_bw%.6g_ks%.9g", c->N, c->fp64, c->method, c->cus,
             npoly, kcase, slots, bw, ks);
    return buf;
}

static void tune_load()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *f = getenv("PS_TUNE_FILE");
    if (!f) return;
    FILE *fp = fopen(f, "r");
    if (!fp) return;
    char key[256];
    ps_launch_params lp;
    while (fscanf(fp, "%255s %d %d %d %d %d", key, &lp.grid_cap, &lp.chunk_tiles, &lp.g2_div, &lp.overlap,
                  &lp.method) == 6) {
        if (strncmp(key, "fwd_", 4) == 0) g_fwd_tuned[key] = lp.grid_cap;  // a forward variant
        else g_tuned[key] = lp;
    }
    fclose(fp);
}

static int forward_variant_tuned(const ps_ctx *c)
{
    tune_load();
    auto it = g_fwd_tuned.find(fwd_key(c));
    const int v = (it == g_fwd_tuned.end()) ? 0 : it->second;
    return (v >= 0 && v < forward_variants(c->N, c->fp64)) ? v : 0;
}

extern "C" int ps_set_forward_variant(ps_ctx *c, int variant)
{
    if (!c || variant < 0 || variant >= forward_variants(c->N, c->fp64)) return PS_ERR_INVALID;
    c->fwd_variant = variant;
    return PS_OK;
}

extern "C" int ps_forward_variant_count(const ps_ctx *c) { return c ? forward_variants(c->N, c->fp64) : 0; }

extern "C" int ps_forward_autotune(ps_ctx *c, int slot, int reps, int *best_variant, double *best_ms)
{
    if (!c || reps < 1 || slot < 0 || slot >= c->nslots) return PS_ERR_INVALID;
    const int nv = forward_variants(c->N, c->fp64);
    const int keep = c->fwd_variant;
    int best = keep;
    double bms = 1e300;
    for (int v = 0; v < nv; ++v) {
        c->fwd_variant = v;
        double ms = 0.0;
        ps_forward_profile(c, slot, 1, &ms, nullptr);  // warm-up
        const int s = ps_forward_profile(c, slot, reps, &ms, nullptr);
        if (s != PS_OK) {
            c->fwd_variant = keep;
            return s;
        }
        if (ms < bms) {
            bms = ms;
            best = v;
        }
    }
    c->fwd_variant = best;
    g_fwd_tuned[fwd_key(c)] = best;
    if (const char *f = getenv("PS_TUNE_FILE")) {
        if (FILE *fp = fopen(f, "a")) {
            fprintf(fp, "%s %d 0 0 0 0\n", fwd_key(c).c_str(), best);
            fclose(fp);
        }
    }
    if (best_variant) *best_variant = best;
    if (best_ms) *best_ms = bms;
    return PS_OK;
}

static ps_launch_params default_params(const ps_ctx *c)
{
    ps_launch_params lp;
    lp.grid_cap = 16 * c->cus;
    lp.method = c->method;
    lp.chunk_tiles = 1;  // dot steps (Fourier) or Zc tiles (direct) per work item
    // G2 tiles of full width where the band fits the prefetch (half-width
    // tiles everywhere took the realisation graph of 128-256^3 fields from
    // 2.35-2.39 to 2.61-2.63 ms; the autotuner picks them where they pay);
    // the waves on at most 3 branch streams (the 512^3 sweep's refined step
    // 0.955-0.957 ms against 0.974-0.981 ms with 4, at these tile widths and
    // 4096 workgroups)
    lp.g2_div = 1;
    lp.overlap = 3;
    if (const char *e = getenv("PS_CHUNK_TILES")) lp.chunk_tiles = std::max(1, atoi(e));
    if (const char *e = getenv("PS_GRID_CAP")) lp.grid_cap = std::max(1, atoi(e));
    return lp;
}

extern "C" int ps_batch_create(ps_ctx *c, ps_polygon *polys, int npoly, double bw, int kcase, ps_batch **out)
{
    return ps_batch_create_ex(c, polys, npoly, bw, kcase, nullptr, out);
}

#define PCK(call)                                                                            \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return (e_ == hipErrorOutOfMemory) ? PS_ERR_NOMEM : PS_ERR_HIP; \
    } while (0)

// smallest supported transform grid m >= need, at most the spectrum's N
// (216 is supported but not used as a reduced grid: its fused pass holds one
// line per wavefront and is slower than 256 on the 512^3 sweep)
// Euclidean bound of a shell: |m| <= r for every cell its test selects (the
// cube-corner tests reach half a cell diagonal beyond their radius, since a
// corner norm is at least |m| - sqrt(3)/2 bw)
static double shell_radius(const MaskParams &p)
{
    const double e = 0.5 * fabs(p.bw), ec = 0.8660254037844387 * fabs(p.bw);
    switch (p.kind) {
    case MASK_RANGE: return p.mmax + e;
    case MASK_WATKINSON: return p.kkf + ec;
    case MASK_WATKINSON_KN: return std::max(p.mmin, p.mmax) + ec;
    default: return p.kkf + e;
    }
}

// Smallest supported grid M <= N on which the sums of a group do not alias.
// A closed combination S = m_1 + ... + m_n of the legs' cells with S = 0
// (mod M) but S != 0 has a component of size >= M, so |S| >= M; it cannot
// occur when M exceeds the sum of the legs' per-axis bands (ksum) or of their
// Euclidean radii (rsum, with a margin for the rounding of the tests).  A leg
// whose band exceeds M/2 is cut to the grid: its cells of a closed polygon
// have |m_i| <= (sum of the other radii) < M - r_i, hence |m_i| < M/2, and
// lie inside the cut.
static int alias_free_grid(int ksum, double rsum, int N)
{
    static const int kGrids[] = {16, 24, 32, 48, 64, 96, 128, 144, 192, 256, 384, 512};
    for (int m : kGrids) {
        if (m > N) break;
        if (m > ksum || (double)m > rsum + 1e-6) return m;
    }
    return N;
}

// Plans the waves of a batch (either method): per group its transform grid,
// the waves (one grid each, within the arena budget), per wave the legs it
// transforms, the tiles and work items of every pass, the model bytes, the
// buffers and the device tables.
static int plan_waves(ps_batch *b, std::vector<int> &counts)
{
    ps_ctx *c = b->c;
    const int N = c->N, fp64 = c->fp64;
    const size_t cs = c->csize();
    const bool fourier = (b->method == PS_METHOD_FOURIER);
    const int ng = (int)b->groups.size();
    // ---- Fourier: suffix masks and their m^2 intervals
    const size_t nsf = b->suffix.size();
    std::vector<char> radial_ok(nsf, 0);
    if (fourier) {
        std::vector<MaskParams> sm(nsf);
        std::vector<int> skmax(nsf);
        for (size_t s = 0; s < nsf; ++s) {
            sm[s] = b->legs[b->suffix[s]].mask;
            skmax[s] = b->legs[b->suffix[s]].Kz;
        }
        PCK(hipMalloc(&b->dmasks, sizeof(MaskParams) * nsf));
        PCK(hipMalloc(&b->divl, sizeof(int2) * 2 * nsf));
        int *dk = nullptr;
        PCK(hipMalloc(&dk, sizeof(int) * nsf));
        b->allocs.push_back(dk);
        PCK(hipMemcpy(b->dmasks, sm.data(), sizeof(MaskParams) * nsf, hipMemcpyHostToDevice));
        PCK(hipMemcpy(dk, skmax.data(), sizeof(int) * nsf, hipMemcpyHostToDevice));
        PSCK(run_shell_intervals(b->dmasks, dk, (int)nsf, b->divl, c->st));
        std::vector<int2> hv(2 * nsf);
        PCK(hipMemcpyAsync(hv.data(), b->divl, sizeof(int2) * 2 * nsf, hipMemcpyDeviceToHost, c->st));
        PCK(hipStreamSynchronize(c->st));
        bool patched = false;
        for (size_t s = 0; s < nsf; ++s) {
            // a radial shell is one m^2 interval; anything else is tested per cell
            const int2 iv = hv[2 * s];
            const int cnt = hv[2 * s + 1].x;
            const bool ok = mask_is_radial(sm[s].kind) && (iv.y < iv.x ? cnt == 0 : cnt == iv.y - iv.x + 1);
            if (!ok && mask_is_radial(sm[s].kind)) {
                hv[2 * s] = make_int2(-1, -1);
                patched = true;
            }
            radial_ok[s] = mask_is_radial(sm[s].kind) && hv[2 * s].x >= 0;
        }
        if (patched) PCK(hipMemcpy(b->divl, hv.data(), sizeof(int2) * 2 * nsf, hipMemcpyHostToDevice));
    }
    // ---- every leg's mask as an m^2 interval where it is radial (the G1
    // gather then tests an integer range instead of a square root per cell)
    {
        const size_t nl = b->legs.size();
        std::vector<MaskParams> lm(nl);
        std::vector<int> lk(nl);
        for (size_t i = 0; i < nl; ++i) {
            lm[i] = b->legs[i].mask;
            lk[i] = b->legs[i].Kz;
        }
        MaskParams *dm = nullptr;
        int *dk = nullptr;
        int2 *dv = nullptr;
        PCK(hipMalloc(&dm, sizeof(MaskParams) * nl));
        b->allocs.push_back(dm);
        PCK(hipMalloc(&dk, sizeof(int) * nl));
        b->allocs.push_back(dk);
        PCK(hipMalloc(&dv, sizeof(int2) * 2 * nl));
        b->allocs.push_back(dv);
        PCK(hipMemcpy(dm, lm.data(), sizeof(MaskParams) * nl, hipMemcpyHostToDevice));
        PCK(hipMemcpy(dk, lk.data(), sizeof(int) * nl, hipMemcpyHostToDevice));
        PSCK(run_shell_intervals(dm, dk, (int)nl, dv, c->st));
        std::vector<int2> hv(2 * nl);
        PCK(hipMemcpyAsync(hv.data(), dv, sizeof(int2) * 2 * nl, hipMemcpyDeviceToHost, c->st));
        PCK(hipStreamSynchronize(c->st));
        for (size_t i = 0; i < nl; ++i) {
            const int2 iv = hv[2 * i];
            const int cnt = hv[2 * i + 1].x;
            const bool ok = mask_is_radial(lm[i].kind) && (iv.y < iv.x ? cnt == 0 : cnt == iv.y - iv.x + 1);
            b->legs[i].ivlo = ok ? iv.x : -1;
            b->legs[i].ivhi = ok ? iv.y : -1;
        }
    }
    // ---- per group: grid and (Fourier) descriptor
    std::vector<int> gM(ng);
    std::vector<SpecGroupDev> sg(fourier ? ng : 0);
    for (int gi = 0; gi < ng; ++gi) {
        const GroupDev &G = b->groups[gi];
        int Kt = 0;
        for (int o = 0; o < G.nsuf; ++o) Kt = std::max(Kt, b->legs[b->suffix[G.suf_begin + o]].Kz);
        const int KA = (G.npre >= 1) ? b->legs[G.preA].Kz : 0;
        const int KB = (G.npre == 2) ? b->legs[G.preB].Kz : (G.pre_sq ? KA : 0);
        if (!fourier) {
            // the product of all legs must not alias onto m = 0
            double rt = 0.0;
            for (int o = 0; o < G.nsuf; ++o) rt = std::max(rt, shell_radius(b->legs[b->suffix[G.suf_begin + o]].mask));
            const double rA = (G.npre >= 1) ? shell_radius(b->legs[G.preA].mask) : 0.0;
            const double rB = (G.npre == 2) ? shell_radius(b->legs[G.preB].mask) : (G.pre_sq ? rA : 0.0);
            gM[gi] = (G.npre == 0) ? alias_free_grid(2 * Kt, 2.0 * rt, N) : alias_free_grid(KA + KB + Kt, rA + rB + rt, N);
            continue;
        }
        SpecGroupDev S{};
        S.fft = (G.npre == 2 || G.pre_sq) ? 1 : 0;
        int Kq = Kt, need = 0;
        if (S.fft) {
            // the product of the prefix legs vanishes beyond |m_i| = K_A + K_B
            Kq = std::min(Kq, KA + KB);
            need = KA + KB + Kq + 1;
        } else {
            if (G.npre == 1) Kq = std::min(Kq, KA);
            need = 2 * Kq + 1;
        }
        Kq = std::min(Kq, N / 2);
        const int M = alias_free_grid(need - 1, need - 1.0, N);
        gM[gi] = M;
        S.Kq = Kq;
        S.Kqz = std::min(Kq, M / 2);
        S.Bq = std::min(2 * Kq + 1, M);
        S.npre = S.fft ? G.npre : 0;
        S.pre_sq = G.pre_sq;
        S.preA = G.preA;  // global leg ids here, wave-local below
        S.preB = (G.npre == 2) ? G.preB : G.preA;
        S.nsuf = G.nsuf;
        S.suf_begin = G.suf_begin;
        S.out_begin = G.out_begin;
        S.has_A = (!S.fft && G.npre == 1) ? 1 : 0;
        if (S.has_A) S.mA = b->legs[G.preA].mask;
        const LegDev &L0 = b->legs[b->suffix[G.suf_begin]];
        S.spec = L0.spec;
        S.phase_only = L0.phase_only;
        S.ns = N;
        S.fscale = S.fft ? pow((double)N / (double)M, 3) : 1.0;
        S.radial = 1;
        for (int o = 0; o < G.nsuf; ++o) S.radial &= radial_ok[G.suf_begin + o] ? 1 : 0;
        sg[gi] = S;
    }
    // ---- waves: one grid each, within the arena budget; grids in increasing
    // order, groups in their order within a grid
    size_t budget = (size_t)16 << 30;
    size_t mfree = 0, mtot = 0;
    if (hipMemGetInfo(&mfree, &mtot) == hipSuccess && mfree / 2 < budget) budget = mfree / 2;
    if (const char *e = getenv("PS_ARENA_MB")) budget = (size_t)std::max(1, atoi(e)) << 20;
    std::vector<Wave> &W = b->waves;
    std::vector<int> grids(gM.begin(), gM.end());
    std::sort(grids.begin(), grids.end());
    grids.erase(std::unique(grids.begin(), grids.end()), grids.end());
    for (int M : grids) {
        auto leg_bytes = [&](const LegDev &L) -> size_t {
            const int Bx = std::min(2 * L.Kz + 1, M);
            return 2 * cs * ((size_t)(L.Kz + 1) * Bx * M + (size_t)(L.Kz + 1) * M * M);
        };
        std::map<int, int> wave_leg;  // leg id -> index in the current wave
        bool fresh = true;
        for (int gi = 0; gi < ng; ++gi) {
            if (gM[gi] != M) continue;
            const GroupDev &G = b->groups[gi];
            // legs the group transforms (Fourier: the prefix of triangle groups only)
            std::vector<int> use;
            if (!fourier || sg[gi].fft) {
                if (G.npre >= 1) use.push_back(G.preA);
                if (G.npre == 2) use.push_back(G.preB);
            }
            if (!fourier)
                for (int o = 0; o < G.nsuf; ++o) use.push_back(b->suffix[G.suf_begin + o]);
            const size_t gbytes =
                (fourier && sg[gi].fft)
                    ? 2 * cs * ((size_t)(sg[gi].Kqz + 1) * M * M + (size_t)(sg[gi].Kqz + 1) * sg[gi].Bq * M)
                    : 0;
            size_t need = gbytes;
            for (int p : use)
                if (fresh || !wave_leg.count(p)) need += leg_bytes(b->legs[p]);
            if (fresh || (W.back().bytes + need > budget)) {
                W.emplace_back();
                W.back().M = M;
                wave_leg.clear();
                fresh = false;
            }
            Wave &w = W.back();
            for (int p : use) {
                if (wave_leg.count(p)) continue;
                wave_leg[p] = (int)w.legs.size();
                LegDev L = b->legs[p];
                // On grid M > K_A + K_B + K_t the modes of a leg that meet the
                // others in a closed product have |m| <= min(K_i, K_j + K_l)
                // < M / 2, so a shell wider than the grid (a polygon that
                // cannot close) keeps only its planes and rows below M / 2.
                L.Kz = L.Kx = std::min(L.Kz, M / 2);
                L.Bx = std::min(2 * L.Kz + 1, M);
                w.leg_off.push_back(w.bytes);
                w.bytes += leg_bytes(L);
                w.legs.push_back(L);
            }
            if (fourier) {
                SpecGroupDev S = sg[gi];
                if (S.fft) {
                    S.preA = wave_leg[S.preA];
                    S.preB = wave_leg[S.preB];
                } else {
                    S.preA = S.preB = 0;
                }
                w.grp_off.push_back(S.fft ? w.bytes : 0);
                w.bytes += gbytes;
                w.sgroups.push_back(S);
            } else {
                GroupDev D = G;
                if (D.npre >= 1) D.preA = wave_leg[G.preA];
                if (D.npre == 2) D.preB = wave_leg[G.preB];
                D.suf_begin = (int)w.suffix.size();
                for (int o = 0; o < G.nsuf; ++o) w.suffix.push_back(wave_leg[b->suffix[G.suf_begin + o]]);
                D.fscale = pow((double)N / (double)M, 3);
                w.groups.push_back(D);
            }
        }
    }
    // ---- tiles, work items and model bytes
    size_t arena = 0, nscale = 0, total_bytes = 0;
    for (Wave &w : W) {
        const int M = w.M;
        const int tl = g_tile_lines(M, fp64);
        double wave_bytes0 = 0.0;
        for (int i = 0; i < PS_MAX_STAGES; ++i) wave_bytes0 += b->stage_bytes[i];
        if (M == N) {
            w.tw = c->tw;
        } else {
            auto it = c->twm.find(M);
            if (it == c->twm.end()) {
                void *t = nullptr;
                PSCK(fp64 ? make_twiddles<double>(M, &t) : make_twiddles<float>(M, &t));
                it = c->twm.emplace(M, t).first;
            }
            w.tw = it->second;
        }
        // direct waves transform the D parts only where the fused pass can
        // pair rows (an even number of rows per wavefront)
        w.dmode = (!fourier && zc_dmode_supported(M, fp64)) ? 1 : 0;
        const int parts = w.dmode ? 1 : 2;
        const int64_t g2_per_kz = M / g2_tile_cols(M, fp64, b->g2_div);
        for (LegDev &L : w.legs) {
            const int K = L.Kz;
            w.kz_cap = std::max(w.kz_cap, K + 1);
            L.g1_begin = w.g1_tiles;
            L.g2_begin = w.g2_tiles;
            w.g1_tiles += (int64_t)(K + 1) * ((L.Bx + tl - 1) / tl);
            w.g2_tiles += 2 * (int64_t)(K + 1) * g2_per_kz;
            if (w.dmode) {
                w.g1b_d.push_back(w.g1_tiles_d);
                w.g2b_d.push_back(w.g2_tiles_d);
                const int ld = g1_dmode_lines(M, fp64);
                w.g1_tiles_d += (int64_t)(K + 1) * ((L.Bx + ld - 1) / ld);
                w.g2_tiles_d += (int64_t)(K + 1) * g2_per_kz;
            }
            // G1 reads the spectrum cells of the band box (an upper bound: the
            // gather reads the shell's cells) and writes the band lines (D, I
            // or D only) that meet the shell's disk mx^2 + mz^2 <= r^2; G2
            // reads those and writes the planes
            double disk = 0.0;
            for (int kz = 0; kz <= K; ++kz) {
                const int mz = lattice(kz, M), lim = (L.r2 < 0) ? M * M : L.r2 - mz * mz;
                for (int j = 0; j < L.Bx; ++j) {
                    const int mx = lattice((L.Bx == M || j <= L.Kx) ? j : j + M - L.Bx, M);
                    if (mx * mx <= lim) disk += 1.0;
                }
            }
            const double n1 = disk * M, n2 = (double)(K + 1) * M * M;
            b->stage_bytes[0] += (double)cs * ((double)(K + 1) * L.Bx * L.Bx + parts * n1);
            b->stage_bytes[1] += (double)cs * (parts * n1 + parts * n2);
        }
        if (fourier) {
            const int zrows = spec_zf_rows(M, fp64);
            const int fxcols = tile_rows(M, fp64) / b->g2_div;
            const int dot_lines = b->chunk_tiles * spec_dot_lines(M, fp64);
            for (SpecGroupDev &S : w.sgroups) {
                w.kinds |= S.radial ? 1 : 2;
                S.dot_lines = dot_lines;
                S.zf_begin = w.zf_tiles;
                S.fx_begin = w.fx_tiles;
                S.dot_begin = w.dot_items;
                const double planes = (double)(S.Kqz + 1) * M * M, band = (double)(S.Kqz + 1) * S.Bq * M;
                if (S.fft) {
                    w.zf_tiles += (int64_t)M * (M / zrows);
                    w.fx_tiles += 2 * (int64_t)(S.Kqz + 1) * (M / fxcols);
                    double pre = (double)(w.legs[S.preA].Kz + 1);
                    if (S.npre == 2) pre += (double)(w.legs[S.preB].Kz + 1);
                    b->stage_bytes[2] += 2.0 * cs * (pre * M * M + planes);
                    b->stage_bytes[3] += 2.0 * cs * (planes + band);
                    b->stage_bytes[4] += 2.0 * cs * band;
                }
                b->stage_bytes[4] += (double)cs * (S.Kqz + 1) * (double)S.Bq * S.Bq;  // spectrum cells of the band
                const int64_t nlines = 2 * (int64_t)(S.Kqz + 1) * S.Bq;
                const int items = (int)((nlines + dot_lines - 1) / dot_lines);
                w.dot_items += items;
                for (int o = 0; o < S.nsuf; ++o) counts[S.out_begin + o] = items;
                b->max_chunks = std::max(b->max_chunks, items);
            }
        } else {
            const int rows = fused_tile_rows(M, fp64);
            const int64_t tiles_per_group = (int64_t)M * (M / rows);
            const int nchunks = (int)((tiles_per_group + b->chunk_tiles - 1) / b->chunk_tiles);
            for (GroupDev &G : w.groups) {
                G.chunk_tiles = b->chunk_tiles;
                G.item_begin = w.items;
                w.items += nchunks;
                for (int o = 0; o < G.nsuf; ++o) counts[G.out_begin + o] = nchunks;
                b->max_chunks = std::max(b->max_chunks, nchunks);
                // model bytes of the fused pass: every leg's D,I planes read once per group
                int64_t planes = 0;
                if (G.npre >= 1) planes += w.legs[G.preA].Kz + 1;
                if (G.npre == 2) planes += w.legs[G.preB].Kz + 1;
                for (int o = 0; o < G.nsuf; ++o) planes += w.legs[w.suffix[G.suf_begin + o]].Kz + 1;
                b->stage_bytes[2] += parts * cs * (double)planes * M * M;
            }
        }
        w.cost = -wave_bytes0;
        for (int i = 0; i < PS_MAX_STAGES; ++i) w.cost += b->stage_bytes[i];
        arena = std::max(arena, w.bytes);
        total_bytes += w.bytes;
        nscale += w.legs.size();
    }
    b->nstages = fourier ? 6 : 4;
    b->stage_bytes[b->nstages - 1] = 32.0 * (double)b->nout * b->max_chunks;
    b->stage_bytes[b->nstages - 2] += b->stage_bytes[b->nstages - 1];  // the last pass writes the partial sums
    // waves side by side (disjoint buffers, branch streams) when allowed and they all fit
    b->concurrent = (b->overlap && W.size() > 1 && total_bytes <= budget) ? 1 : 0;
    if (b->concurrent) arena = total_bytes;
    // ---- buffers and device tables
    if (arena > 0) PCK(hipMalloc(&b->arena, arena));
    b->arena_bytes = arena;
    // amplitude scales: the direct method in mixed precision (its last pass
    // packs D + i I of one leg); the Fourier method in both precisions (its
    // forward transform packs parts of the prefix product, P_D ~ |H|^2 P_I)
    if (nscale > 0 && (fourier || !fp64)) PCK(hipMalloc(&b->dscale, sizeof(double) * nscale));
    size_t si = 0;
    char *base = (char *)b->arena;
    for (size_t wi = 0; wi < W.size(); ++wi) {
        Wave &w = W[wi];
        const int M = w.M;
        if (b->concurrent && wi > 0) base += W[wi - 1].bytes;  // disjoint regions
        for (size_t i = 0; i < w.legs.size(); ++i) {
            LegDev &L = w.legs[i];
            const size_t n1 = (size_t)(L.Kz + 1) * L.Bx * M, n2 = (size_t)(L.Kz + 1) * M * M;
            char *p = base + w.leg_off[i];
            L.g1d = L.spec ? p : nullptr;
            L.g1i = p + cs * n1;
            L.g2d = L.spec ? p + 2 * cs * n1 : nullptr;
            L.g2i = p + 2 * cs * n1 + cs * n2;
            L.scale = b->dscale ? b->dscale + si++ : nullptr;
        }
        if (!w.legs.empty()) {
            PCK(hipMalloc(&w.dlegs, sizeof(LegDev) * w.legs.size()));
            PCK(hipMemcpy(w.dlegs, w.legs.data(), sizeof(LegDev) * w.legs.size(), hipMemcpyHostToDevice));
        }
        if (!fourier && !w.legs.empty()) {
            // I parts only (bare masks), for the sums taken at planning
            std::vector<LegDev> li(w.legs);
            for (LegDev &L : li) {
                L.spec = nullptr;
                L.g1d = L.g2d = nullptr;
                L.scale = nullptr;
            }
            PCK(hipMalloc(&w.dlegs_i, sizeof(LegDev) * li.size()));
            PCK(hipMemcpy(w.dlegs_i, li.data(), sizeof(LegDev) * li.size(), hipMemcpyHostToDevice));
            if (w.dmode) {
                // D parts only, with their own tiles
                std::vector<LegDev> ld(w.legs);
                for (size_t i = 0; i < ld.size(); ++i) {
                    ld[i].g1_begin = w.g1b_d[i];
                    ld[i].g2_begin = w.g2b_d[i];
                    ld[i].scale = nullptr;
                }
                PCK(hipMalloc(&w.dlegs_d, sizeof(LegDev) * ld.size()));
                PCK(hipMemcpy(w.dlegs_d, ld.data(), sizeof(LegDev) * ld.size(), hipMemcpyHostToDevice));
            }
        }
        if (fourier) {
            for (size_t g = 0; g < w.sgroups.size(); ++g) {
                SpecGroupDev &S = w.sgroups[g];
                if (!S.fft) continue;
                const size_t np = (size_t)(S.Kqz + 1) * M * M, nb = (size_t)(S.Kqz + 1) * S.Bq * M;
                char *p = base + w.grp_off[g];
                S.qd = p;
                S.qi = p + cs * np;
                S.bd = p + 2 * cs * np;
                S.bi = p + 2 * cs * np + cs * nb;
                S.scaleA = w.legs[S.preA].scale;
                S.scaleB = (S.npre == 2) ? w.legs[S.preB].scale : nullptr;
            }
            PCK(hipMalloc(&w.dsgroups, sizeof(SpecGroupDev) * w.sgroups.size()));
            PCK(hipMemcpy(w.dsgroups, w.sgroups.data(), sizeof(SpecGroupDev) * w.sgroups.size(),
                          hipMemcpyHostToDevice));
        } else {
            PCK(hipMalloc(&w.dgroups, sizeof(GroupDev) * w.groups.size()));
            PCK(hipMemcpy(w.dgroups, w.groups.data(), sizeof(GroupDev) * w.groups.size(), hipMemcpyHostToDevice));
            PCK(hipMalloc(&w.dsuffix, sizeof(int) * w.suffix.size()));
            PCK(hipMemcpy(w.dsuffix, w.suffix.data(), sizeof(int) * w.suffix.size(), hipMemcpyHostToDevice));
        }
    }
    if (b->concurrent) {
        // overlap 1: up to 4 branch streams; 2 .. 4: at most that many
        const int nb = std::min((int)W.size(), b->overlap >= 2 ? b->overlap : 4);
        // waves to branches, largest first onto the least loaded branch; each
        // branch runs its waves smallest first (the small grids are latency
        // bound and overlap the large ones), and the capture takes the
        // branches' first waves first, so every branch starts at once
        std::vector<int> idx(W.size());
        for (size_t i = 0; i < W.size(); ++i) idx[i] = (int)i;
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return W[x].cost > W[y].cost; });
        std::vector<double> load(nb, 0.0);
        std::vector<std::vector<int>> per(nb);
        for (int i : idx) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            W[i].branch = k;
            load[k] += W[i].cost;
            per[k].push_back(i);
        }
        for (auto &v : per) std::reverse(v.begin(), v.end());
        b->rec_order.clear();
        for (size_t pos = 0; b->rec_order.size() < W.size(); ++pos)
            for (int k = 0; k < nb; ++k)
                if (pos < per[k].size()) b->rec_order.push_back(per[k][pos]);
        while ((int)c->br.size() < nb) {
            hipStream_t t = nullptr;
            PCK(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));  // default priority
            c->br.push_back(t);
        }
        PCK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        b->wend.assign(nb, nullptr);
        for (auto &e : b->wend) PCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    return PS_OK;
}
#undef PCK


// Direct method: the polygons of one prefix key split into classes by the
// transform grid their product needs.  grid[q] is the smallest alias-free
// grid of polygon q; a class runs on the largest grid among its members and
// transforms the npre prefix legs once more.  The classes minimise
// sum M^3 * (npre + members) (every leg costs about M^3 on grid M) over every
// subset of the grids that occur (at most the 11 supported ones); members
// keep their input order.
static std::vector<std::vector<int>> split_by_grid(const std::vector<int> &ps_, const std::vector<int> &grid, int npre)
{
    std::vector<int> g(grid.begin(), grid.end());
    std::sort(g.begin(), g.end());
    g.erase(std::unique(g.begin(), g.end()), g.end());
    const int k = (int)g.size();
    std::vector<double> cnt(k, 0.0);
    for (int m : grid) cnt[std::lower_bound(g.begin(), g.end(), m) - g.begin()] += 1.0;
    unsigned best = 0;
    double best_cost = -1.0;
    for (unsigned sub = 0; sub < (1u << (k - 1)); ++sub) {
        const unsigned chosen = sub | (1u << (k - 1));  // the largest grid is always used
        double cost = 0.0, members = 0.0;
        for (int i = 0; i < k; ++i) {
            members += cnt[i];
            if (chosen & (1u << i)) {
                cost += std::pow((double)g[i], 3) * (npre + members);
                members = 0.0;
            }
        }
        if (best_cost < 0.0 || cost < best_cost) {
            best_cost = cost;
            best = chosen;
        }
    }
    std::vector<std::vector<int>> out;
    std::vector<int> cls(k, -1);
    for (int i = k - 1, cur = -1; i >= 0; --i) {
        if (best & (1u << i)) cur = i;
        cls[i] = cur;
    }
    std::vector<int> slot(k, -1);
    for (int i = 0; i < k; ++i)
        if (best & (1u << i)) {
            slot[i] = (int)out.size();
            out.emplace_back();
        }
    for (size_t q = 0; q < ps_.size(); ++q) {
        const int i = (int)(std::lower_bound(g.begin(), g.end(), grid[q]) - g.begin());
        out[slot[cls[i]]].push_back(ps_[q]);
    }
    return out;
}

extern "C" int ps_batch_create_ex(ps_ctx *c, ps_polygon *polys, int npoly, double bw, int kcase,
                                  const ps_launch_params *params, ps_batch **out)
{
    if (!c || !polys || npoly < 1 || !out) return PS_ERR_INVALID;
    *out = nullptr;
    HIPCK(hipSetDevice(c->device));
    ps_launch_params lp = default_params(c);
    if (params) {
        lp = *params;
    } else {
        tune_load();
        auto it = g_tuned.find(tune_key(c, polys, npoly, bw, kcase));
        if (it != g_tuned.end()) lp = it->second;
    }
    if (lp.grid_cap < 1 || lp.chunk_tiles < 1 || (lp.g2_div != 1 && lp.g2_div != 2) ||
        lp.overlap < 0 || lp.overlap > 4 || (lp.method != PS_METHOD_FOURIER && lp.method != PS_METHOD_DIRECT))
        return PS_ERR_INVALID;
    for (int p = 0; p < npoly; ++p) {
        if (polys[p].n < 2 || polys[p].n > 3) return PS_ERR_INVALID;
        if (polys[p].slot < 0 || polys[p].slot >= c->nslots) return PS_ERR_INVALID;
        if (!c->loaded[polys[p].slot]) return PS_ERR_STATE;
    }
    const int N = c->N;
    ps_batch *b = new ps_batch();
    b->c = c;
    b->npoly = npoly;
    b->method = lp.method;
    b->grid_cap = lp.grid_cap;
    b->chunk_tiles = lp.chunk_tiles;
    b->overlap = lp.overlap;
    b->g2_div = (tile_rows(c->N, c->fp64) / lp.g2_div >= 4) ? lp.g2_div : 1;
    // 1. legs of every polygon, de-duplicated by (slot, mask)
    std::vector<std::vector<int>> pleg(npoly);
    std::vector<std::pair<int, MaskParams>> keys;
    for (int p = 0; p < npoly; ++p) {
        ps_polygon &P = polys[p];
        for (int i = 0; i < P.n; ++i) {
            MaskParams mp;
            int s = leg_mask(c, i, P.n, P.k, P.kmin, P.kmax, bw, kcase, &mp);
            if (s != PS_OK) {
                ps_batch_destroy(b);
                return s;
            }
            int id = -1;
            for (size_t q = 0; q < keys.size(); ++q)
                if (keys[q].first == P.slot && same_mask(keys[q].second, mp)) { id = (int)q; break; }
            if (id < 0) {
                id = (int)keys.size();
                keys.push_back({P.slot, mp});
            }
            pleg[p].push_back(id);
        }
    }
    // groups: polygons of one field slot sharing their prefix legs, in a fixed order
    typedef std::tuple<int, int, int, int> GKey;  // (kind, leg a, leg b, slot)
    std::map<GKey, std::vector<int>> bykey;
    for (int p = 0; p < npoly; ++p) {
        const std::vector<int> &l = pleg[p];
        const int sl = polys[p].slot;
        if (polys[p].n == 3) bykey[GKey{3, l[0], l[1], sl}].push_back(p);
        else if (l[0] == l[1]) bykey[GKey{1, -1, -1, sl}].push_back(p);
        else bykey[GKey{2, l[0], -1, sl}].push_back(p);
    }
    std::vector<std::pair<GKey, std::vector<int>>> glist(bykey.begin(), bykey.end());
    // order the legs by the first group that uses them, so that the legs a
    // group introduces form one contiguous range of tiles
    {
        std::vector<int> newid(keys.size(), -1);
        int nxt = 0;
        auto take = [&](int id) {
            if (id >= 0 && newid[id] < 0) newid[id] = nxt++;
        };
        for (auto &gkv : glist) {
            const int kind = std::get<0>(gkv.first);
            const std::vector<int> &ps_ = gkv.second;
            for (size_t s0 = 0; s0 < ps_.size(); s0 += MAX_SUFFIX) {
                if (kind == 3) {
                    take(std::get<1>(gkv.first));
                    take(std::get<2>(gkv.first));
                } else if (kind == 2) {
                    take(std::get<1>(gkv.first));
                }
                for (size_t q = s0; q < std::min(ps_.size(), s0 + MAX_SUFFIX); ++q) {
                    const int pp = ps_[q];
                    take(kind == 1 ? pleg[pp][0] : pleg[pp][polys[pp].n - 1]);
                }
            }
        }
        for (size_t i = 0; i < keys.size(); ++i) take((int)i);  // unused legs (none expected)
        std::vector<std::pair<int, MaskParams>> k2(keys.size());
        for (size_t i = 0; i < keys.size(); ++i) k2[newid[i]] = keys[i];
        keys.swap(k2);
        for (auto &l : pleg)
            for (int &id : l) id = newid[id];
        for (auto &gkv : glist) {
            auto &t = gkv.first;
            if (std::get<1>(t) >= 0) std::get<1>(t) = newid[std::get<1>(t)];
            if (std::get<2>(t) >= 0) std::get<2>(t) = newid[std::get<2>(t)];
        }
    }
    // 2. host leg table (buffers are the method's)
    for (auto &kv : keys) {
        b->leg_slot.push_back(kv.first);
        LegDev L{};
        L.mask = kv.second;
        L.spec = c->spec[kv.first];
        L.phase_only = c->phase_only;
        const int K = shell_band(kv.second, N);
        L.Kz = K;
        L.Kx = K;
        L.Bx = std::min(2 * K + 1, N);
        L.ns = N;
        // every selected cell has |m| <= r, so a (kz, kx) line with
        // mz^2 + mx^2 > r^2 holds none (the first passes skip it)
        const double r = shell_radius(kv.second);
        if (r >= 0.0 && r < (double)N) L.r2 = (int)floor(r * r * (1.0 + 1e-9) + 1e-9);
        b->legs.push_back(L);
    }
    // 3. group descriptors
    b->poly_out.assign(npoly, -1);
    b->poly_n.assign(npoly, 0);
    for (auto &kv : glist) {
        const int kind = std::get<0>(kv.first);
        std::vector<std::vector<int>> classes{kv.second};
        std::vector<int> class_chunk{MAX_SUFFIX};
        if (b->method == PS_METHOD_DIRECT) {
            const int a = std::get<1>(kv.first), a2 = std::get<2>(kv.first);
            const int KA = (a >= 0) ? b->legs[a].Kz : 0;
            const int KB = (a2 >= 0) ? b->legs[a2].Kz : 0;
            const double rA = (a >= 0) ? shell_radius(b->legs[a].mask) : 0.0;
            const double rB = (a2 >= 0) ? shell_radius(b->legs[a2].mask) : 0.0;
            std::vector<int> grid;
            std::map<int, int> grid_of;
            for (int p : kv.second) {
                const LegDev &Lt = b->legs[kind == 1 ? pleg[p][0] : pleg[p][polys[p].n - 1]];
                const int Kt = Lt.Kz;
                const double rt = shell_radius(Lt.mask);
                grid.push_back(kind == 1 ? alias_free_grid(2 * Kt, 2.0 * rt, N)
                                         : alias_free_grid(KA + KB + Kt, rA + rB + rt, N));
                grid_of[p] = grid.back();
            }
            const int npre = (kind == 3) ? (a == a2 ? 1 : 2) : (kind == 2 ? 1 : 0);
            classes = split_by_grid(kv.second, grid, npre);
            // a small grid has few fused-pass tiles (work items), each running
            // through every leg of its group: there its suffixes are split
            // into more groups (each recomputing the prefix) so that about
            // 256 items run side by side
            class_chunk.clear();
            for (const std::vector<int> &cl : classes) {
                int M = 0;
                for (int p : cl) M = std::max(M, grid_of[p]);
                const int64_t tiles = (int64_t)M * (M / fused_tile_rows(M, c->fp64));
                const int split = (int)std::min<int64_t>((int64_t)cl.size(), std::max<int64_t>(1, (256 + tiles - 1) / tiles));
                class_chunk.push_back(std::min<int>(MAX_SUFFIX, (int)((cl.size() + split - 1) / split)));
            }
        }
        std::vector<std::vector<int>> chunks;  // at most MAX_SUFFIX polygons each
        for (size_t ci = 0; ci < classes.size(); ++ci) {
            const std::vector<int> &cl = classes[ci];
            const size_t cs = (size_t)class_chunk[ci];
            for (size_t s0 = 0; s0 < cl.size(); s0 += cs)
                chunks.emplace_back(cl.begin() + s0, cl.begin() + std::min(cl.size(), s0 + cs));
        }
        for (const std::vector<int> &ps_ : chunks) {
            GroupDev G{};
            if (kind == 3) {
                G.preA = std::get<1>(kv.first);
                G.preB = std::get<2>(kv.first);
                G.pre_sq = (G.preA == G.preB) ? 1 : 0;
                G.npre = G.pre_sq ? 1 : 2;
            } else if (kind == 2) {
                G.npre = 1;
                G.preA = std::get<1>(kv.first);
            } else {
                G.npre = 0;
                G.suf_sq = 1;
            }
            G.nsuf = (int)ps_.size();
            G.suf_begin = (int)b->suffix.size();
            G.out_begin = b->nout;
            G.chunk_tiles = b->chunk_tiles;
            G.fscale = 1.0;
            for (const int p : ps_) {
                const std::vector<int> &l = pleg[p];
                b->suffix.push_back(kind == 1 ? l[0] : l[polys[p].n - 1]);
                b->poly_out[p] = b->nout;
                b->poly_n[p] = polys[p].n;
                b->nout++;
            }
            b->groups.push_back(G);
        }
    }
    // 4. the method's buffers, tiles and device tables
    std::vector<int> counts(b->nout, 0);
    int st = plan_waves(b, counts);
    if (st != PS_OK) {
        ps_batch_destroy(b);
        return st;
    }
    b->model_bytes = 0.0;
    for (int i = 0; i < b->nstages; ++i) b->model_bytes += b->stage_bytes[i];
#define BCK(call)                                       \
    do {                                                \
        if ((call) != hipSuccess) { st = PS_ERR_HIP; goto fail; } \
    } while (0)
    BCK(hipMalloc(&b->dcounts, sizeof(int) * counts.size()));
    BCK(hipMalloc(&b->dpartials, sizeof(double2) * 2 * (size_t)b->nout * b->max_chunks));
    BCK(hipMalloc(&b->dsum, sizeof(double) * 3 * (size_t)b->nout));  // D sums, I sums, |prod D| sums
    BCK(hipHostMalloc(&b->hsum, sizeof(double) * 3 * (size_t)b->nout, hipHostMallocDefault));
    BCK(hipMemcpy(b->dcounts, counts.data(), sizeof(int) * counts.size(), hipMemcpyHostToDevice));
    BCK(hipEventCreate(&b->ev0));
    BCK(hipEventCreate(&b->ev1));
    if (b->method == PS_METHOD_DIRECT) {
        st = batch_i_sums(b);
        if (st != PS_OK) goto fail;
    }
    // 5. the polygons as planned (DERIVED ranges filled in), for refinement
    b->polys.assign(polys, polys + npoly);
    b->bw = bw;
    b->kcase = kcase;
    b->refined.assign(npoly, 0);
    b->err_est.assign(npoly, 0.0);
    // 6. capture the launch sequence once (with the magnitude sums when the
    // context refines mixed-precision values)
    if (c->refine_tol > 0.0 && !c->fp64 && b->method == PS_METHOD_DIRECT) {
        b->dabs = b->dsum + 2 * (size_t)b->nout;
        b->abs_rec = 1;
    }
    st = batch_capture(b);
    if (st != PS_OK) goto fail;
#undef BCK
    *out = b;
    return PS_OK;
fail:
    ps_batch_destroy(b);
    return st;
}

// device sums of the last execution -> normalised values (polyspectrum.c:155, :247)
// Error estimate of a mixed-precision value: the FP32 data (spectrum,
// filtered fields) carry each term of S = sum prod D to about u32 = 2^-24 of
// its size; over the #polygons closed lattice polygons of the sum these
// errors add up like a random walk, and a floor of a few u32 remains (the
// rounding of the normalised terms themselves), so
//   |dS| / |S| <= u32 (C0 + C kappa / sqrt(#polygons)),
//   kappa = sum |prod D| / |sum prod D|.
// C0 = 16, C = 32 bound the measured errors of every polygon of the 512^3
// sweep with a factor 1.4 to spare.
static constexpr double kRefineC0 = 16.0, kRefineC = 32.0;
static constexpr size_t kRefineSets = 8;  // FP64 batches kept per batch (refined polygon sets)

// one copy of the D and I sums (and the magnitude sums when refining) into
// the pinned host buffer, queued on the context's stream behind the passes
static int batch_queue_sums(ps_batch *b, int with_abs)
{
    HIPCK(hipMemcpyAsync(b->hsum, b->dsum, sizeof(double) * (2 + with_abs) * (size_t)b->nout,
                         hipMemcpyDeviceToHost, b->c->st));
    return PS_OK;
}

// spec: the FP64 batch of the last refined set, launched (on its own
// context's stream, its sums' copy queued behind it) beside this execution;
// its values are used when the set is the same again, else the new set is
// evaluated after the results.  queued: the caller has queued this batch's
// copy (batch_queue_sums) and waited for it.
static int batch_results(ps_batch *b, double *values, double *npolygons, double *sums, int refine,
                         ps_batch *spec = nullptr, int queued = 0)
{
    ps_ctx *c = b->c;
    const int with_abs = (refine && b->abs_rec) ? 1 : 0;
    const double *h = b->hsum;
    if (!queued) {
        PSCK(batch_queue_sums(b, with_abs));
        HIPCK(hipStreamSynchronize(c->st));
    }
    b->last_sumI.assign(h + b->nout, h + 2 * b->nout);
    b->have_sums = 1;
    const double *ha = with_abs ? h + 2 * b->nout : nullptr;
    const double volume = c->L * c->L * c->L;
    const double cells = (double)((int64_t)c->N * c->N * c->N);
    std::vector<int> flagged;
    for (int p = 0; p < b->npoly; ++p) {
        const int o = b->poly_out[p], n = b->poly_n[p];
        const double sD = h[o], sI = h[b->nout + o];
        values[p] = pow(volume, n - 1) * pow(c->N, -n * 3) * sD / sI;  // polyspectrum.c:155
        if (npolygons) npolygons[p] = sI / cells;                      // polyspectrum.c:247
        if (sums) {
            sums[2 * p] = sD;
            sums[2 * p + 1] = sI;
        }
        b->refined[p] = 0;
        b->err_est[p] = 0.0;
        if (!ha) continue;
        const double npoly = std::max(1.0, sI / cells);
        const double kappa = ha[o] / fabs(sD);
        b->err_est[p] = (sD != 0.0) ? 0x1p-24 * (kRefineC0 + kRefineC * kappa / sqrt(npoly)) : HUGE_VAL;
        if (!(b->err_est[p] <= c->refine_tol)) flagged.push_back(p);
    }
    if (spec) HIPCK(hipStreamSynchronize(spec->c->st));  // its passes and sums' copy
    b->last_flagged = flagged;
    if (flagged.empty()) return PS_OK;
    // the flagged polygons again, together, with the FP64 batched path of the
    // companion context (one batch per flagged set, planned on first use)
    for (int p : flagged) {
        const int sl = b->polys[p].slot;
        if (sl < 0 || sl >= c->nslots || !c->r64 || !c->r64->loaded[sl]) return PS_ERR_STATE;
    }
    ps_batch *rb = nullptr;
    auto it = b->ref.find(flagged);
    if (it != b->ref.end()) {
        rb = it->second;
    } else {
        std::vector<ps_polygon> q;
        for (int p : flagged) q.push_back(b->polys[p]);
        // its waves one after another on the FP64 context's stream: it runs
        // beside the next execution's graph, whose branches hold the other
        // streams (measured 0.978-0.985 ms per 512^3 step with two refined
        // triangles, against 1.003-1.006 with its waves on branch streams)
        ps_launch_params lp = default_params(c->r64);
        lp.overlap = 0;
        PSCK(ps_batch_create_ex(c->r64, q.data(), (int)q.size(), b->bw, b->kcase, &lp, &rb));
        if (b->ref_order.size() >= kRefineSets) {
            ps_batch_destroy(b->ref[b->ref_order.front()]);
            b->ref.erase(b->ref_order.front());
            b->ref_order.erase(b->ref_order.begin());
        }
        b->ref[flagged] = rb;
        b->ref_order.push_back(flagged);
    }
    const size_t nf = flagged.size();
    std::vector<double> v(nf), np(nf), s2(2 * nf);
    if (rb == spec) PSCK(batch_results(rb, v.data(), np.data(), s2.data(), 0, nullptr, 1));  // ran beside this execution
    else PSCK(ps_batch_execute(rb, v.data(), np.data(), s2.data()));
    HIPCK(hipSetDevice(c->device));
    for (size_t i = 0; i < nf; ++i) {
        const int p = flagged[i];
        values[p] = v[i];
        if (npolygons) npolygons[p] = np[i];
        if (sums) {
            sums[2 * p] = s2[2 * i];
            sums[2 * p + 1] = s2[2 * i + 1];
        }
        b->refined[p] = 1;
    }
    return PS_OK;
}

extern "C" int ps_batch_execute(ps_batch *b, double *values, double *npolygons, double *sums)
{
    if (!b || !values) return PS_ERR_INVALID;
    ps_ctx *c = b->c;
    HIPCK(hipSetDevice(c->device));
    PSCK(batch_sync_refinement(b));
    HIPCK(hipEventRecord(b->ev0, c->st));
    if (b->exec) HIPCK(hipGraphLaunch(b->exec, c->st));
    else PSCK(batch_record(b));
    HIPCK(hipEventRecord(b->ev1, c->st));
    // the sums' copy queued behind the passes: one wait for both
    PSCK(batch_queue_sums(b, b->abs_rec));
    // refinement: the FP64 batch of the set the last execution refined runs
    // beside this one, on the FP64 context's stream, launched after it so
    // that this execution's passes start first (measured 1.03-1.06 ms per
    // 512^3 step with it launched before, 0.985-0.989 after; recomputed every
    // time, used if the set comes out the same)
    ps_batch *spec = nullptr;
    if (b->abs_rec && c->r64 && !b->last_flagged.empty()) {
        auto it = b->ref.find(b->last_flagged);
        if (it != b->ref.end() && it->second->exec) {
            spec = it->second;
            HIPCK(hipGraphLaunch(spec->exec, spec->c->st));
            PSCK(batch_queue_sums(spec, 0));
        }
    }
    HIPCK(hipStreamSynchronize(c->st));
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    return batch_results(b, values, npolygons, sums, 1, spec, 1);
}

extern "C" int ps_batch_results(ps_batch *b, double *values, double *npolygons, double *sums)
{
    if (!b || !values) return PS_ERR_INVALID;
    HIPCK(hipSetDevice(b->c->device));
    // after ps_multi_execute the magnitude sums exist only if the batch took
    // them when the multi-batch was captured: no refinement then
    return batch_results(b, values, npolygons, sums, 0);
}

extern "C" int ps_batch_refinement(const ps_batch *b, int *refined, double *error_estimates)
{
    if (!b) return PS_ERR_INVALID;
    int n = 0;
    for (int p = 0; p < b->npoly; ++p) {
        if (refined) refined[p] = b->refined[p];
        if (error_estimates) error_estimates[p] = b->err_est[p];
        n += b->refined[p];
    }
    return n;
}

// ---------------------------------------------------------------- multi-batch
// Several batches (contexts of different grid sizes, many field slots) in
// one captured launch sequence on the stream of the first batch's context.
struct ps_multi {
    std::vector<ps_batch *> b;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};

extern "C" int ps_multi_destroy(ps_multi *m)
{
    if (!m) return PS_ERR_INVALID;
    if (m->st) (void)hipStreamSynchronize(m->st);
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    delete m;
    return PS_OK;
}

extern "C" int ps_multi_create(ps_batch **batches, int nbatch, ps_multi **out)
{
    if (!batches || nbatch < 1 || !out) return PS_ERR_INVALID;
    *out = nullptr;
    for (int i = 0; i < nbatch; ++i)
        if (!batches[i] || batches[i]->c->device != batches[0]->c->device) return PS_ERR_INVALID;
    HIPCK(hipSetDevice(batches[0]->c->device));
    ps_multi *m = new ps_multi();
    m->b.assign(batches, batches + nbatch);
    m->st = batches[0]->c->st;
    int st = PS_OK;
    // the other contexts' pending work (field uploads) must be complete
    for (int i = 0; i < nbatch; ++i)
        if (hipStreamSynchronize(batches[i]->c->st) != hipSuccess) st = PS_ERR_HIP;
    if (st == PS_OK && (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess))
        st = PS_ERR_HIP;
    if (st == PS_OK && hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal) != hipSuccess) st = PS_ERR_HIP;
    if (st == PS_OK) {
        for (int i = 0; i < nbatch && st == PS_OK; ++i) st = batch_record_on(batches[i], m->st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(m->st, &g);
        if (e != hipSuccess && st == PS_OK) st = PS_ERR_HIP;
        m->graph = g;
        if (st == PS_OK && hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0) != hipSuccess)
            st = PS_ERR_HIP;
    }
    if (st != PS_OK) {
        ps_multi_destroy(m);
        return st;
    }
    *out = m;
    return PS_OK;
}

extern "C" int ps_multi_execute(ps_multi *m)
{
    if (!m) return PS_ERR_INVALID;
    HIPCK(hipSetDevice(m->b[0]->c->device));
    HIPCK(hipEventRecord(m->ev0, m->st));
    HIPCK(hipGraphLaunch(m->exec, m->st));
    HIPCK(hipEventRecord(m->ev1, m->st));
    HIPCK(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_ms = ms;
    return PS_OK;
}

extern "C" double ps_multi_last_ms(const ps_multi *m) { return m ? m->last_ms : -1.0; }

// ---------------------------------------------------------------- adjoint
__global__ void k_to_float(float *__restrict__ out, const double *__restrict__ in, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

// gradient of sum_p weights[p] * values[p] with respect to the field of a slot
extern "C" int ps_batch_adjoint(ps_batch *b, int slot, const double *weights, void *grad, int grad_is_double,
                                int on_device)
{
    if (!b || !weights || !grad) return PS_ERR_INVALID;
    ps_ctx *c = b->c;
    if (slot < 0 || slot >= c->nslots || !c->loaded[slot]) return PS_ERR_STATE;
    if (!b->have_sums) return PS_ERR_STATE;
    // useOnlyPhase: the filtered fields are taken from p = z/|z| (utils_fftw.c:78-93)
    int phase = 0;
    for (size_t i = 0; i < b->legs.size(); ++i)
        if (b->leg_slot[i] == slot && b->legs[i].phase_only) phase = 1;
    HIPCK(hipSetDevice(c->device));
    const int N = c->N;
    const int64_t cells = c->cells(), half = c->half_elems();
    const size_t hb = c->csize() * (size_t)half, fb = sizeof(double) * (size_t)cells;
    // coefficient of every output: dL/dS_o = sum_p w_p * c_n / C_o
    std::vector<double> coef(b->nout, 0.0);
    const double volume = c->L * c->L * c->L;
    for (int p = 0; p < b->npoly; ++p) {
        const int o = b->poly_out[p], n = b->poly_n[p];
        const double cn = pow(volume, n - 1) * pow(c->N, -n * 3);
        coef[o] += weights[p] * cn / b->last_sumI[o];
    }
    int st = PS_OK;
    void *H1 = nullptr, *H2 = nullptr;
    double *DA = nullptr, *DB = nullptr, *E = nullptr, *Q = nullptr, *G = nullptr, *dw = nullptr;
    MaskParams *dm = nullptr;
#define ACK(call)                                                   \
    do {                                                            \
        hipError_t e_ = (call);                                     \
        if (e_ != hipSuccess) { st = (e_ == hipErrorOutOfMemory) ? PS_ERR_NOMEM : PS_ERR_HIP; goto done; } \
    } while (0)
#define RCK(call)                                                   \
    do {                                                            \
        int r_ = (call);                                            \
        if (r_ != PS_OK) { st = (r_ > 0) ? PS_ERR_HIP : r_; goto done; } \
    } while (0)
    {
        // two half spectra, five FP64 fields and the mask lists, allocated by
        // the first call and kept with the batch
        const size_t nml = b->suffix.size() + 2;  // leg A, leg B, every suffix leg of the batch
        const size_t sz[9] = {hb, hb, fb, fb, fb, fb, fb, sizeof(MaskParams) * nml, sizeof(double) * nml};
        for (int i = 0; i < 9; ++i)
            if (!b->adj[i]) ACK(hipMalloc(&b->adj[i], sz[i]));
    }
    H1 = b->adj[0];
    H2 = b->adj[1];
    DA = (double *)b->adj[2];
    DB = (double *)b->adj[3];
    E = (double *)b->adj[4];
    Q = (double *)b->adj[5];
    G = (double *)b->adj[6];
    dm = (MaskParams *)b->adj[7];
    dw = (double *)b->adj[8];
    RCK(run_fill(G, cells, 0.0, c->st));
    {
    // groups with the same prefix legs (split by transform grid or by the
    // suffix limit when the batch was planned) form one weighted mask W
    std::map<std::tuple<int, int, int, int>, std::vector<int>> prefix_groups;
    for (int gi = 0; gi < (int)b->groups.size(); ++gi) {
        const GroupDev &g = b->groups[gi];
        if (b->leg_slot[b->suffix[g.suf_begin]] != slot) continue;
        prefix_groups[std::make_tuple(g.npre, g.pre_sq, g.npre >= 1 ? g.preA : -1, g.npre == 2 ? g.preB : -1)]
            .push_back(gi);
    }
    for (const auto &kv : prefix_groups) {
        const GroupDev &g = b->groups[kv.second[0]];
        // host mask lists: [0] leg A, [1] leg B, [2..] the weighted suffix masks W
        std::vector<MaskParams> hm;
        std::vector<double> hw;
        // power spectra: dS/d delta = 2 L_A L_t delta = 2 N^3 IFFT(F_A F_t FFT delta)
        // (the unnormalised transforms give IFFT FFT = N^3)
        const double two = (g.npre == 0 || (g.npre == 1 && !g.pre_sq)) ? 2.0 * (double)cells : 1.0;
        hm.push_back(g.npre >= 1 ? b->legs[g.preA].mask : MaskParams{});
        hw.push_back(1.0);
        hm.push_back(g.npre == 2 ? b->legs[g.preB].mask : MaskParams{});
        hw.push_back(1.0);
        for (int gi : kv.second) {
            const GroupDev &h = b->groups[gi];
            for (int q = 0; q < h.nsuf; ++q) {
                hm.push_back(b->legs[b->suffix[h.suf_begin + q]].mask);
                hw.push_back(two * coef[h.out_begin + q]);
            }
        }
        const int nsw = (int)hm.size() - 2;
        // device lists: [0] A, [1] B, [2..] W (nsw masks)
        MaskParams *mA = dm, *mB = dm + 1, *mW = dm + 2;
        double *wA = dw, *wB = dw + 1, *wW = dw + 2;
        ACK(hipMemcpyAsync(mA, &hm[0], sizeof(MaskParams), hipMemcpyHostToDevice, c->st));
        ACK(hipMemcpyAsync(wA, &hw[0], sizeof(double), hipMemcpyHostToDevice, c->st));
        ACK(hipMemcpyAsync(mB, &hm[1], sizeof(MaskParams), hipMemcpyHostToDevice, c->st));
        ACK(hipMemcpyAsync(wB, &hw[1], sizeof(double), hipMemcpyHostToDevice, c->st));
        ACK(hipMemcpyAsync(mW, &hm[2], sizeof(MaskParams) * nsw, hipMemcpyHostToDevice, c->st));
        ACK(hipMemcpyAsync(wW, &hw[2], sizeof(double) * nsw, hipMemcpyHostToDevice, c->st));
        const void *spec = c->spec[slot];
        // H1 = the spectrum the filtered fields are taken from: z, or z/|z|
        auto load_spec = [&]() -> int {
            if (hipMemcpyAsync(H1, spec, hb, hipMemcpyDeviceToDevice, c->st) != hipSuccess) return PS_ERR_HIP;
            return phase ? run_phase_normalize(c->fp64, H1, half, c->st) : PS_OK;
        };
        // G += IFFT(H1), H1 the gradient in spectrum space; with useOnlyPhase
        // first pulled back through p = z/|z| (the gradient with respect to z)
        auto finish = [&]() -> int {
            if (phase) {
                const int r = run_phase_pullback(c->fp64, H1, spec, half, c->st);
                if (r != PS_OK) return r;
            }
            return run_inverse_real(c->fp64, N, H1, G, 1, c->tw, c->st);
        };
        auto filtered = [&](const MaskParams *m, const double *w, int nm, double *out) -> int {
            // out = IFFT(mask * spectrum)
            int r = load_spec();
            if (r != PS_OK) return r;
            r = run_apply_masks(c->fp64, N, H1, m, w, nm, c->st);
            if (r != PS_OK) return r;
            return run_inverse_real(c->fp64, N, H1, out, 0, c->tw, c->st);
        };
        auto project = [&](const double *f, void *H, const MaskParams *m, const double *w, int nm) -> int {
            // H = mask * FFT(f)
            int r = run_forward(c->fp64, N, f, 1, H, c->tw, c->st, c->fwd_variant);
            if (r != PS_OK) return r;
            return run_apply_masks(c->fp64, N, H, m, w, nm, c->st);
        };
        if (g.npre == 0) {
            // S_t = sum D_t^2: dS_t = 2 L_t delta (masks are idempotent), the
            // factor 2 is in the weights
            RCK(load_spec());
            RCK(run_apply_masks(c->fp64, N, H1, mW, wW, nsw, c->st));
            RCK(finish());
        } else if (g.npre == 1 && !g.pre_sq) {
            // S_t = sum D_A D_t: dS = L_A[E] + L_W[D_A] = 2 IFFT(F_A W delta)
            RCK(load_spec());
            RCK(run_apply_masks(c->fp64, N, H1, mA, wA, 1, c->st));
            RCK(run_apply_masks(c->fp64, N, H1, mW, wW, nsw, c->st));
            RCK(finish());
        } else if (g.pre_sq) {
            // S_t = sum D_A^2 D_t: grad = L_A[2 D_A E] + L_W[D_A^2]
            RCK(filtered(mA, wA, 1, DA));
            RCK(filtered(mW, wW, nsw, E));
            RCK(run_mul_fields(Q, DA, E, 2.0, cells, c->st));
            RCK(project(Q, H2, mA, wA, 1));
            RCK(run_mul_fields(Q, DA, nullptr, 1.0, cells, c->st));
            RCK(project(Q, H1, mW, wW, nsw));
            RCK(run_add_spec(c->fp64, H1, H2, half, c->st));
            RCK(finish());
        } else {
            // S_t = sum D_A D_B D_t: grad = L_A[D_B E] + L_B[D_A E] + L_W[D_A D_B]
            RCK(filtered(mA, wA, 1, DA));
            RCK(filtered(mB, wB, 1, DB));
            RCK(filtered(mW, wW, nsw, E));
            RCK(run_mul_fields(Q, DB, E, 1.0, cells, c->st));
            RCK(project(Q, H2, mA, wA, 1));
            RCK(run_mul_fields(Q, DA, E, 1.0, cells, c->st));
            RCK(project(Q, H1, mB, wB, 1));
            RCK(run_add_spec(c->fp64, H2, H1, half, c->st));
            RCK(run_mul_fields(Q, DA, DB, 1.0, cells, c->st));
            RCK(project(Q, H1, mW, wW, nsw));
            RCK(run_add_spec(c->fp64, H1, H2, half, c->st));
            RCK(finish());
        }
        ACK(hipStreamSynchronize(c->st));  // the host mask lists are reused
    }
    }
    if (on_device) {
        if (grad_is_double) {
            ACK(hipMemcpyAsync(grad, G, fb, hipMemcpyDeviceToDevice, c->st));
        } else {
            hipLaunchKernelGGL(k_to_float, dim3(4096), dim3(256), 0, c->st, (float *)grad, G, cells);
            ACK(hipGetLastError());
        }
    } else if (grad_is_double) {
        ACK(hipMemcpyAsync(grad, G, fb, hipMemcpyDeviceToHost, c->st));
    } else {
        hipLaunchKernelGGL(k_to_float, dim3(4096), dim3(256), 0, c->st, (float *)Q, G, cells);
        ACK(hipGetLastError());
        ACK(hipMemcpyAsync(grad, Q, sizeof(float) * (size_t)cells, hipMemcpyDeviceToHost, c->st));
    }
    ACK(hipStreamSynchronize(c->st));
done:
#undef ACK
#undef RCK
    (void)hipStreamSynchronize(c->st);
    return st;
}

extern "C" double ps_batch_last_ms(const ps_batch *b) { return b ? b->last_ms : -1.0; }

extern "C" int ps_batch_autotune(ps_ctx *c, ps_polygon *polys, int npoly, double bw, int kcase, int reps,
                                 ps_launch_params *best, double *best_ms)
{
    if (!c || !polys || npoly < 1 || reps < 1) return PS_ERR_INVALID;
    const int caps[] = {2 * c->cus, 4 * c->cus, 8 * c->cus, 16 * c->cus};
    // work-item sizes: Zc tiles (direct) or dot steps (Fourier) per item
    const int chunks_direct[] = {1, 4, 16, 64};
    const int chunks_fourier[] = {1, 2, 4};
    const int divs[] = {1, 2};
    std::vector<ps_polygon> work(polys, polys + npoly);
    std::vector<double> v(npoly);
    // one candidate, after a warm-up execution: the fastest of n executions
    // (the least disturbed by other work on the device), or their median;
    // each timed on the host from the call to its results, as a caller sees
    // it (with the FP64 refinement beside the step, where the context refines)
    auto trial = [&](const ps_launch_params &lp, double *ms, int n = 0, bool median = false) -> int {
        ps_batch *b = nullptr;
        std::copy(polys, polys + npoly, work.begin());
        int s = ps_batch_create_ex(c, work.data(), npoly, bw, kcase, &lp, &b);
        if (s != PS_OK) return s;
        if (n < 1) n = reps;
        std::vector<double> ts;
        s = ps_batch_execute(b, v.data(), nullptr, nullptr);
        for (int r = 0; r < n && s == PS_OK; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            s = ps_batch_execute(b, v.data(), nullptr, nullptr);
            ts.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        ps_batch_destroy(b);
        std::sort(ts.begin(), ts.end());
        *ms = ts.empty() ? 1e300 : (median ? ts[ts.size() / 2] : ts.front());
        return s;
    };
    ps_launch_params bp = default_params(c);
    double bms = 1e300;
    const int method = c->method;
    std::vector<std::pair<double, ps_launch_params>> ranked;
    for (int cap : caps)
        for (int ch : (method == PS_METHOD_FOURIER
                           ? std::vector<int>(std::begin(chunks_fourier), std::end(chunks_fourier))
                           : std::vector<int>(std::begin(chunks_direct), std::end(chunks_direct))))
            for (int dv : divs) {
                const ps_launch_params lp{cap, ch, dv, 1, method};
                double t = 0.0;
                const int s = trial(lp, &t);
                if (s != PS_OK) return s;
                ranked.push_back({t, lp});
            }
    // the four fastest again, by the median of 3 reps replays: a shape whose
    // fastest replay was a lucky one does not win
    std::stable_sort(ranked.begin(), ranked.end(),
                     [](const std::pair<double, ps_launch_params> &a, const std::pair<double, ps_launch_params> &b) {
                         return a.first < b.first;
                     });
    for (size_t i = 0; i < ranked.size() && i < 4; ++i) {
        double t = 0.0;
        const int s = trial(ranked[i].second, &t, 3 * reps, true);
        if (s != PS_OK) return s;
        if (t < bms) {
            bms = t;
            bp = ranked[i].second;
        }
    }
    // at the best shape: the waves on at most 3 or 2 branch streams, and one
    // after another (fewer kernels side by side starve the ones with the
    // largest LDS and register footprint less: 3 measured 3% faster than 4
    // on the 512^3 sweep)
    const ps_launch_params shape = bp;
    for (int ov : {3, 2, 0}) {
        ps_launch_params lp = shape;
        lp.overlap = ov;
        double t = 0.0;
        const int s = trial(lp, &t, 3 * reps, true);
        if (s != PS_OK) return s;
        if (t < bms) {
            bms = t;
            bp = lp;
        }
    }
    const std::string key = tune_key(c, polys, npoly, bw, kcase);
    tune_load();
    g_tuned[key] = bp;
    if (const char *f = getenv("PS_TUNE_FILE")) {
        if (FILE *fp = fopen(f, "a")) {
            fprintf(fp, "%s %d %d %d %d %d\n", key.c_str(), bp.grid_cap, bp.chunk_tiles, bp.g2_div, bp.overlap,
                    bp.method);
            fclose(fp);
        }
    }
    if (best) *best = bp;
    if (best_ms) *best_ms = bms;
    return PS_OK;
}

extern "C" int ps_batch_params(const ps_batch *b, ps_launch_params *out)
{
    if (!b || !out) return PS_ERR_INVALID;
    out->grid_cap = b->grid_cap;
    out->chunk_tiles = b->chunk_tiles;
    out->g2_div = b->g2_div;
    out->overlap = b->overlap;
    out->method = b->method;
    return PS_OK;
}

extern "C" int ps_batch_profile(ps_batch *b, int reps, double *stage_ms, double *stage_bytes)
{
    if (!b || reps < 1 || !stage_ms) return PS_ERR_INVALID;
    ps_ctx *c = b->c;
    HIPCK(hipSetDevice(c->device));
    PSCK(batch_sync_refinement(b));  // the passes as the captured step runs them
    const int fp = c->fp64;
    const hipStream_t s = c->st;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    int st = PS_OK;
    // one pass (over the whole batch, or one wave of it) between two events
    auto timed = [&](int stage, auto &&launch) {
        if (st != PS_OK) return;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            if (e0) (void)hipEventDestroy(e0);
            st = PS_ERR_HIP;
            return;
        }
        marks.push_back({stage, {e0, e1}});
        (void)hipEventRecord(e0, s);
        st = launch();
        (void)hipEventRecord(e1, s);
    };
    for (int r = 0; r < reps && st == PS_OK; ++r) {
        for (Wave &w : b->waves) {
            const int nl = (int)w.legs.size(), M = w.M;
            if (b->method == PS_METHOD_DIRECT && w.dmode) {
                timed(0, [&] {
                    return run_inverse_g(fp, M, w.dlegs_d, nl, 0, w.g1_tiles_d, 0, 0, b->grid_cap, b->g2_div, 1, w.tw, s);
                });
                timed(1, [&] {
                    return run_inverse_g(fp, M, w.dlegs_d, nl, 0, 0, 0, w.g2_tiles_d, b->grid_cap, b->g2_div, 1, w.tw, s);
                });
                timed(2, [&] {
                    return run_zc_fused(fp, M, w.dgroups, (int)w.groups.size(), w.dlegs_d, w.dsuffix, 0, w.items,
                                        b->dpartials, b->max_chunks, b->grid_cap, w.kz_cap, b->abs_rec ? 3 : 1, w.tw,
                                        s);
                });
                continue;
            }
            if (nl > 0) {
                timed(0, [&] {
                    const int x = b->dscale ? run_leg_scale(fp, M, w.dlegs, nl, s) : PS_OK;
                    return x != PS_OK ? x
                                      : run_inverse_g(fp, M, w.dlegs, nl, 0, w.g1_tiles, 0, 0, b->grid_cap, b->g2_div,
                                                      0, w.tw, s);
                });
                timed(1, [&] {
                    return run_inverse_g(fp, M, w.dlegs, nl, 0, 0, 0, w.g2_tiles, b->grid_cap, b->g2_div, 0, w.tw, s);
                });
            }
            if (b->method == PS_METHOD_FOURIER) {
                const int ng = (int)w.sgroups.size();
                timed(2, [&] {
                    return run_spec_zf(fp, M, w.dsgroups, ng, w.dlegs, 0, w.zf_tiles, b->grid_cap, w.kz_cap, w.tw, s);
                });
                timed(3, [&] {
                    return run_spec_fx(fp, M, w.dsgroups, ng, 0, w.fx_tiles, b->grid_cap, b->g2_div, w.tw, s);
                });
                timed(4, [&] {
                    return run_spec_dot(fp, M, w.dsgroups, ng, b->dmasks, b->divl, 0, w.dot_items, b->dpartials,
                                        b->max_chunks, b->grid_cap, w.kinds, w.tw, s);
                });
            } else {
                timed(2, [&] {
                    return run_zc_fused(fp, M, w.dgroups, (int)w.groups.size(), w.dlegs, w.dsuffix, 0, w.items,
                                        b->dpartials, b->max_chunks, b->grid_cap, w.kz_cap, 0, w.tw, s);
                });
            }
        }
        timed(b->nstages - 1, [&] {
            double *outI = (b->method == PS_METHOD_DIRECT) ? (b->abs_rec ? b->dabs : nullptr) : b->dsum + b->nout;
            return run_reduce_partials(b->dpartials, b->max_chunks, b->dcounts, 0, b->nout, b->dsum, outI, s);
        });
    }
    if (hipStreamSynchronize(s) != hipSuccess && st == PS_OK) st = PS_ERR_HIP;
    double acc[PS_MAX_STAGES] = {};
    for (auto &m : marks) {
        float ms = 0.f;
        if (st == PS_OK && hipEventElapsedTime(&ms, m.second.first, m.second.second) == hipSuccess) acc[m.first] += ms;
        (void)hipEventDestroy(m.second.first);
        (void)hipEventDestroy(m.second.second);
    }
    if (st != PS_OK) return st > 0 ? PS_ERR_HIP : st;
    for (int i = 0; i < PS_MAX_STAGES; ++i) {
        stage_ms[i] = (i < b->nstages) ? acc[i] / reps : 0.0;
        if (stage_bytes) stage_bytes[i] = (i < b->nstages) ? b->stage_bytes[i] : 0.0;
    }
    return PS_OK;
}

extern "C" int ps_batch_stage_count(const ps_batch *b) { return b ? b->nstages : 0; }

extern "C" const char *ps_batch_stage_name(const ps_batch *b, int stage)
{
    if (!b || stage < 0 || stage >= b->nstages) return nullptr;
    return b->method == PS_METHOD_FOURIER ? kStageFourier[stage] : kStageDirect[stage];
}

extern "C" int ps_batch_wave_count(const ps_batch *b) { return b ? (int)b->waves.size() : 0; }

extern "C" int ps_batch_wave_info(const ps_batch *b, int w, int *grid, int *nlegs, int *ngroups, double *bytes)
{
    if (!b || w < 0 || w >= (int)b->waves.size()) return PS_ERR_INVALID;
    const Wave &v = b->waves[w];
    if (grid) *grid = v.M;
    if (nlegs) *nlegs = (int)v.legs.size();
    if (ngroups) *ngroups = (int)(b->method == PS_METHOD_FOURIER ? v.sgroups.size() : v.groups.size());
    if (bytes) *bytes = (double)v.bytes;
    return PS_OK;
}

extern "C" int ps_forward_profile(ps_ctx *c, int slot, int reps, double *ms, double *bytes)
{
    if (!c || reps < 1 || !ms || slot < 0 || slot >= c->nslots) return PS_ERR_INVALID;
    if (!c->staging || !c->loaded[slot] || c->staging_slot != slot) return PS_ERR_STATE;
    HIPCK(hipSetDevice(c->device));
    // the transforms go to a scratch half spectrum: the resident spectra stay as they are
    if (!c->fscratch) HIPCK(hipMalloc(&c->fscratch, c->csize() * (size_t)c->half_elems()));
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    (void)hipEventRecord(e0, c->st);
    int st = PS_OK;
    for (int r = 0; r < reps && st == PS_OK; ++r)
        st = run_forward(c->fp64, c->N, c->staging, c->staging_is_double, c->fscratch, c->tw, c->st, c->fwd_variant);
    (void)hipEventRecord(e1, c->st);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != PS_OK) return st > 0 ? PS_ERR_HIP : st;
    *ms = t / reps;
    if (bytes) {
        // R2C model (three passes, each reading and writing its array once):
        // s*N^3 read + cs*NH*N^2 written, then 2 x (read + write) of the half spectrum
        const double n3 = (double)c->cells(), half = (double)c->half_elems();
        *bytes = (c->staging_is_double ? 8.0 : 4.0) * n3 + c->csize() * half * 5.0;
    }
    return PS_OK;
}
extern "C" double ps_batch_model_bytes(const ps_batch *b) { return b ? b->model_bytes : -1.0; }

// ---------------------------------------------------------------- transform-free check
// Cells of one shell, packed 10 bits per coordinate (offset N/2): the band
// |m_i| <= K of the shell scanned with the estimator's membership test.  K
// bounds every selected cell: each test implies a cube-corner norm or |m| of
// at most r = shell_band's radius, and a cube-corner norm is at least
// max_i |m_i| - bw/2.
static void shell_cells(const MaskParams &mp, int N, std::vector<int> &out)
{
    out.clear();
    const int h = N / 2;
    const int K = std::min(h, shell_band(mp, N) + 1);
    const int lo = -K, hi = std::min(K, h - 1);
    for (int mx = lo; mx <= hi; ++mx)
        for (int my = lo; my <= hi; ++my)
            for (int mz = lo; mz <= hi; ++mz)
                if (mask_eval(mp, mx, my, mz)) out.push_back((mx + h) | ((my + h) << 10) | ((mz + h) << 20));
}

// device buffers of one call, freed on every return path
struct DevTemps {
    std::vector<void *> p;
    template <typename T> hipError_t alloc(T **q, size_t count)
    {
        const hipError_t e = hipMalloc((void **)q, sizeof(T) * std::max<size_t>(count, 1));
        if (e == hipSuccess) p.push_back(*q);
        return e;
    }
    void release()
    {
        for (void *q : p) (void)hipFree(q);
        p.clear();
    }
    ~DevTemps() { release(); }
};

// Device buffers of one shell pairing of the transform-free check: taken
// from a caller's scratch region (no allocation) or allocated per call.
struct DirectBufs {
    unsigned char *base = nullptr;  // scratch region, or nullptr: allocate
    size_t size = 0;
    size_t used = 0;
    DevTemps tmp;
    template <typename T> int take(T **q, size_t count)
    {
        if (!base) return tmp.alloc(q, count) == hipSuccess ? PS_OK : PS_ERR_NOMEM;
        const size_t bytes = (sizeof(T) * std::max<size_t>(count, 1) + 255) & ~(size_t)255;
        if (used + bytes > size) return PS_ERR_SCRATCH;
        *q = reinterpret_cast<T *>(base + used);
        used += bytes;
        return PS_OK;
    }
    void release()
    {
        tmp.release();
        used = 0;
    }
};

// scratch: a device region of scratch_bytes for every buffer (first
// kCornerBlocks double2 for DERIVED corner extremes), or nullptr to allocate
static int direct_check_impl(ps_ctx *c, ps_polygon *polys, int npoly, double bw, int kcase, double *values,
                             double *npolygons, double *sums, unsigned char *scratch, size_t scratch_bytes)
{
    if (!c || !polys || npoly <= 0 || !values) return PS_ERR_INVALID;
    if (c->N > 1024) return PS_ERR_UNSUPPORTED_N;  // cells are packed 10 bits per coordinate
    HIPCK(hipSetDevice(c->device));
    double2 *corner = scratch ? reinterpret_cast<double2 *>(scratch) : nullptr;
    const size_t corner_bytes = scratch ? (sizeof(double2) * kCornerBlocks + 255) & ~(size_t)255 : 0;
    // polygons sharing their enumerated shells (bispectra: legs 1 and 2,
    // power spectra: leg 1) share one pass over the cells
    struct Pairing {
        int n = 0, slot = 0;
        MaskParams m1{}, m2{};
        std::vector<int> poly;          // member polygons
        std::vector<MaskParams> last;   // their closing shells
    };
    std::vector<Pairing> pg;
    for (int p = 0; p < npoly; ++p) {
        ps_polygon &P = polys[p];
        if (P.n != 2 && P.n != 3) return PS_ERR_INVALID;
        if (P.slot < 0 || P.slot >= c->nslots) return PS_ERR_INVALID;
        if (!c->loaded[P.slot]) return PS_ERR_STATE;
        MaskParams mp[3];
        for (int i = 0; i < P.n; ++i) PSCK(leg_mask(c, i, P.n, P.k, P.kmin, P.kmax, bw, kcase, &mp[i], corner));
        size_t g = 0;
        while (g < pg.size() && !(pg[g].n == P.n && pg[g].slot == P.slot && same_mask(pg[g].m1, mp[0]) &&
                                  (P.n == 2 || same_mask(pg[g].m2, mp[1]))))
            ++g;
        if (g == pg.size()) {
            Pairing q;
            q.n = P.n;
            q.slot = P.slot;
            q.m1 = mp[0];
            if (P.n == 3) q.m2 = mp[1];
            pg.push_back(q);
        }
        pg[g].poly.push_back(p);
        pg[g].last.push_back(mp[P.n - 1]);
    }
    const int N = c->N;
    std::vector<double> S(npoly, 0.0), C(npoly, 0.0);
    std::vector<int> cells1, cells2;
    DirectBufs buf;
    buf.base = scratch ? scratch + corner_bytes : nullptr;
    buf.size = scratch_bytes > corner_bytes ? scratch_bytes - corner_bytes : 0;
    for (const Pairing &q : pg) {
        shell_cells(q.m1, N, cells1);
        if (q.n == 3) shell_cells(q.m2, N, cells2);
        const int n1 = (int)cells1.size(), n2 = (q.n == 3) ? (int)cells2.size() : 0;
        if (n1 == 0 || (q.n == 3 && n2 == 0)) continue;  // no polygon: both sums 0
        const void *H = c->spec[q.slot];
        const int nb = direct_sum_blocks(n1, n2);
        int *dc1 = nullptr, *dc2 = nullptr;
        double2 *dv1 = nullptr, *dv2 = nullptr, *dpart = nullptr;
        // the previous pairing's reads of these buffers are complete (its
        // partial sums were copied back synchronously)
        PSCK(buf.take(&dc1, n1));
        PSCK(buf.take(&dv1, n1));
        PSCK(buf.take(&dpart, (size_t)nb * DIRECT_MEMBERS));
        if (n2 > 0) {
            PSCK(buf.take(&dc2, n2));
            PSCK(buf.take(&dv2, n2));
        }
        HIPCK(hipMemcpyAsync(dc1, cells1.data(), sizeof(int) * n1, hipMemcpyHostToDevice, c->st));
        PSCK(run_direct_values(c->fp64, H, N, c->phase_only, dc1, n1, dv1, c->st));
        if (n2 > 0) {
            HIPCK(hipMemcpyAsync(dc2, cells2.data(), sizeof(int) * n2, hipMemcpyHostToDevice, c->st));
            PSCK(run_direct_values(c->fp64, H, N, c->phase_only, dc2, n2, dv2, c->st));
        }
        std::vector<double2> hp((size_t)nb * DIRECT_MEMBERS);
        for (size_t m0 = 0; m0 < q.last.size(); m0 += DIRECT_MEMBERS) {
            DirectMembers mem{};
            mem.nm = (int)std::min<size_t>(DIRECT_MEMBERS, q.last.size() - m0);
            mem.bw = bw;
            for (int j = 0; j < mem.nm; ++j) {
                mem.mask[j] = q.last[m0 + j];
                if (!mask_is_radial(mem.mask[j].kind)) mem.corners = 1;
            }
            PSCK(run_direct_sums(c->fp64, H, N, c->phase_only, dc1, dv1, n1, dc2, dv2, n2, mem, dpart, c->st));
            HIPCK(hipMemcpyAsync(hp.data(), dpart, sizeof(double2) * hp.size(), hipMemcpyDeviceToHost, c->st));
            HIPCK(hipStreamSynchronize(c->st));
            for (int j = 0; j < mem.nm; ++j) {
                long double s = 0.L, cnt = 0.L;  // fixed order over the workgroups
                for (int k = 0; k < nb; ++k) {
                    s += hp[(size_t)k * DIRECT_MEMBERS + j].x;
                    cnt += hp[(size_t)k * DIRECT_MEMBERS + j].y;
                }
                S[q.poly[m0 + j]] = (double)s;
                C[q.poly[m0 + j]] = (double)cnt;
            }
        }
        HIPCK(hipStreamSynchronize(c->st));  // the copies of the cell lists have been read
        buf.release();
    }
    const double volume = c->L * c->L * c->L;
    const double cells = (double)c->cells();
    for (int p = 0; p < npoly; ++p) {
        const int n = polys[p].n;
        const double sD = cells * S[p], sI = cells * C[p];
        values[p] = pow(volume, n - 1) * pow(c->N, -n * 3) * sD / sI;  // polyspectrum.c:155
        if (npolygons) npolygons[p] = C[p];
        if (sums) {
            sums[2 * p] = sD;
            sums[2 * p + 1] = sI;
        }
    }
    return PS_OK;
}

extern "C" int ps_direct_check(ps_ctx *c, ps_polygon *polys, int npoly, double bw, int kcase, double *values,
                               double *npolygons, double *sums)
{
    return direct_check_impl(c, polys, npoly, bw, kcase, values, npolygons, sums, nullptr, 0);
}

// ---------------------------------------------------------------- allocation-free check
// The buffers of ps_direct_check_static: one file-scope device array per GPU
// (the code object's copy on each device), held until the process ends.  Its
// host-side name is not a device address; every use goes through the address
// hipGetSymbolAddress returns on that device, kept per device.
static __device__ __align__(256) unsigned char scratch[PS_DIRECT_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static unsigned char *g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static int scratch_on(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices) return PS_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        HIPCK(hipSetDevice(device));
        HIPCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_dev[device];
    return PS_OK;
}

extern "C" int ps_direct_check_static(ps_ctx *c, ps_polygon *polys, int npoly, double bw, int kcase,
                                      double *values, double *npolygons, double *sums)
{
    if (!c) return PS_ERR_INVALID;
    unsigned char *p = nullptr;
    PSCK(scratch_on(c->device, &p));
    return direct_check_impl(c, polys, npoly, bw, kcase, values, npolygons, sums, p, PS_DIRECT_SCRATCH_BYTES);
}

// ---------------------------------------------------------------- drop-in symbols
// A process-wide FP64 context per (nbins, boxsize), as the reference's
// functions carry no handle.
static ps_ctx *g_legacy = nullptr;
// The spectrum resident in the drop-in context: the caller's array and a hash
// of its contents when it was uploaded.  run.c passes one array to every call
// (run.c:47, :87), so each call after the first hashes the array in host
// memory (several threads, about memory speed) instead of uploading it again.
static const void *g_legacy_src = nullptr;
static uint64_t g_legacy_hash = 0;

static int legacy_ctx(int nbins, double boxsize)
{
    if (g_legacy && (g_legacy->N != nbins || g_legacy->L != boxsize)) {
        ps_ctx_destroy(g_legacy);
        g_legacy = nullptr;
        g_legacy_src = nullptr;
    }
    if (!g_legacy) return ps_ctx_create(nbins, boxsize, PS_PREC_FP64, 1, &g_legacy);
    return PS_OK;
}

// 64-bit hash of n words: four interleaved multiply-xor chains per thread,
// the threads' results combined in a fixed order
static uint64_t hash_words(const uint64_t *w, size_t n)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nth = std::max<size_t>(1, std::min<size_t>(hw ? hw : 1, 16));
    const size_t per = (n + nth - 1) / nth;
    std::vector<uint64_t> part(nth, 0);
    auto work = [&](size_t t) {
        const size_t a = std::min(n, t * per), e = std::min(n, a + per);
        uint64_t h[4] = {0x9E3779B97F4A7C15ull, 0xC2B2AE3D27D4EB4Full, 0x165667B19E3779F9ull, 0x27D4EB2F165667C5ull};
        size_t i = a;
        for (; i + 4 <= e; i += 4)
            for (int j = 0; j < 4; ++j) h[j] = (h[j] ^ w[i + j]) * 0x100000001B3ull + (h[j] >> 29);
        for (; i < e; ++i) h[0] = (h[0] ^ w[i]) * 0x100000001B3ull + (h[0] >> 29);
        part[t] = h[0] ^ (h[1] * 3) ^ (h[2] * 5) ^ (h[3] * 7) ^ (uint64_t)(e - a);
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nth; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    uint64_t r = 0;
    for (size_t t = 0; t < nth; ++t) r = (r ^ part[t]) * 0x9E3779B97F4A7C15ull + t;
    return r;
}

// full spectrum rows F[kx][ky][0..N/2] (pitch N/2+1) -> H[kz][kx][ky]
__global__ void k_full_to_half(const double2 *__restrict__ rows, double2 *__restrict__ H, int N)
{
    __shared__ double2 t[32][33];
    const int NH = N / 2 + 1;
    const int kx = blockIdx.z;
    const int ky0 = blockIdx.y * 32, kz0 = blockIdx.x * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int ky = ky0 + r, kz = kz0 + tx;
        if (ky < N && kz < NH) t[r][tx] = rows[((int64_t)kx * N + ky) * NH + kz];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int kz = kz0 + r, ky = ky0 + tx;
        if (ky < N && kz < NH) H[((int64_t)kz * N + kx) * N + ky] = t[tx][r];
    }
}

extern "C" double polyspectrum(int nbins, int local_n0, int local_n0_start, ps_fftw_complex *fft_array, int n,
                               double *k, double *kmin, double *kmax, double kbinwidth, int kbinningCase,
                               double boxsize)
{
    const double nan = std::nan("");
    if (local_n0 != nbins || local_n0_start != 0 || !fft_array) return nan;
    if (legacy_ctx(nbins, boxsize) != PS_OK) return nan;
    ps_ctx *c = g_legacy;
    const int N = nbins, NH = N / 2 + 1;
    if (ensure_spec(c, 0) != PS_OK || ensure_plain_buffers(c) != PS_OK) return nan;
    const uint64_t h = hash_words(reinterpret_cast<const uint64_t *>(fft_array), 2 * (size_t)N * N * N);
    if (!(c->loaded[0] && g_legacy_src == fft_array && g_legacy_hash == h)) {
        g_legacy_src = nullptr;
        // the work buffer gd holds exactly N^2 rows of N/2+1 complex doubles
        double2 *rows = (double2 *)c->gd;
        if (hipMemcpy2DAsync(rows, NH * 16, fft_array, (size_t)N * 16, NH * 16, (size_t)N * N,
                             hipMemcpyHostToDevice, c->st) != hipSuccess)
            return nan;
        hipLaunchKernelGGL(k_full_to_half, dim3((NH + 31) / 32, (N + 31) / 32, N), dim3(32, 8), 0, c->st, rows,
                           (double2 *)c->spec[0], N);
        if (hipStreamSynchronize(c->st) != hipSuccess) return nan;
        c->loaded[0] = 1;
        g_legacy_src = fft_array;
        g_legacy_hash = h;
    }
    c->phase_only = 0;
    double r = nan, sums[2] = {0.0, 0.0};
    if (ps_polyspectrum(c, 0, n, k, kmin, kmax, kbinwidth, kbinningCase, &r, sums) != PS_OK) return nan;
    // polyspectrum.c:142 (rank 0): Im/Re of both sums; the imaginary parts of
    // sums of real filtered fields are exactly zero here
    printf("rank %d: cimag(sumNumPolygons)/creal(sumNumPolygons) = %e\n\tcimag(sumValuesPolygons)/creal(sumValuesPolygons) = %e\n",
           0, 0.0 / sums[1], 0.0 / sums[0]);
    return r;
}

extern "C" double num_polygons(int nbins, int local_n0, int local_n0_start, int n, double *k, double *kmin,
                               double *kmax, double kbinwidth, int kbinningCase, double boxsize)
{
    const double nan = std::nan("");
    if (local_n0 != nbins || local_n0_start != 0) return nan;
    if (legacy_ctx(nbins, boxsize) != PS_OK) return nan;
    double r = nan;
    if (ps_num_polygons(g_legacy, n, k, kmin, kmax, kbinwidth, kbinningCase, &r) != PS_OK) return nan;
    return r;
}

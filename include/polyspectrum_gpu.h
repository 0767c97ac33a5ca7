/*
 * polyspectrum_gpu.h -- C ABI of the MI355X (gfx950) polyspectrum estimator.
 *
 * The library replaces the estimator layer of annehutter/polyspectrum
 * (src/polyspectrum.[ch], src/filter.[ch], src/fft.[ch], src/utils_fftw.[ch])
 * behind plain C entry points: pointers, sizes and int status codes, no
 * language runtime types.  A field is uploaded once, transformed on the GPU
 * and kept resident; estimates are returned as doubles.
 *
 * Two forms are exported:
 *   1. a context API (ps_*) that keeps the spectrum in HBM and evaluates
 *      single configurations (the reference's per-triangle shape) or whole
 *      batches of configurations (one captured hipGraph);
 *   2. drop-in symbols with the reference's exact signatures
 *      (polyspectrum.h:7-8), so an unmodified run.c can link against them.
 *
 * Conventions are the reference's: N^3 grids in C order (grid.c:168), FFTW
 * sign conventions and no normalisation (fft.c:19-57), the normalisation
 * V^(n-1) N^(-3n) sum(prod D) / sum(prod I) (polyspectrum.c:155) and
 * Npolygons = sum(prod I) / N^3 (polyspectrum.c:247).
 *
 * Threading: a context is used by one host thread at a time.
 */
#ifndef POLYSPECTRUM_GPU_H
#define POLYSPECTRUM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PS_OK 0
#define PS_ERR_INVALID (-1)        /* bad argument (null pointer, n < 2, ...) */
#define PS_ERR_UNSUPPORTED_N (-2)  /* grid size without a compiled FFT plan */
#define PS_ERR_HIP (-3)            /* a HIP runtime call failed */
#define PS_ERR_NOMEM (-4)          /* device or host allocation failed */
#define PS_ERR_STATE (-5)          /* no field loaded, or wrong slot */
#define PS_ERR_NO_DEVICE (-6)      /* no usable GPU */
#define PS_ERR_SCRATCH (-7)        /* work exceeds the static scratch buffer */

/* ---- precision of a context ------------------------------------------ */
#define PS_PREC_FP64 1  /* FP64 FFTs and products: the plain parity path */
#define PS_PREC_MIXED 0 /* FP32 FFTs, FP64 compensated reductions and normalisation */

/* ---- kbinningCase values (confObj.c:98-105) ---------------------------- */
#define PS_BINNING_DEFAULT 0
#define PS_BINNING_DERIVED 1
#define PS_BINNING_GIVEN 2
#define PS_BINNING_WATKINSON 3

/* Evaluation of a batch (ps_set_method, ps_launch_params::method).  Both
 * give the reference's sum over the grid of the product of the filtered
 * fields, each group of polygons on the smallest grid on which that product
 * does not alias (the sum over the spectrum's grid is (N/M)^3 times the sum
 * over such a grid M):
 *   DIRECT (default): every leg inverse transformed, the pointwise product
 *            of the filtered fields and its sum, as the reference does.
 *   FOURIER: per group of polygons sharing their first legs, the product of
 *            those legs is forward transformed once and each last leg is a
 *            sum over its shell in Fourier space (Parseval); power spectra
 *            need no transform.  Same sums in another order. */
#define PS_METHOD_FOURIER 0
#define PS_METHOD_DIRECT 1
#define PS_MAX_STAGES 6 /* passes ps_batch_profile reports at most */

/* fftw_complex without <complex.h> (FFTW's own fallback typedef) */
typedef double ps_fftw_complex[2];

typedef struct ps_ctx ps_ctx;
typedef struct ps_batch ps_batch;

/* One polygon configuration as polyspectrum() receives it: n legs with
 * k[i], kmin[i], kmax[i] (h/Mpc).  kmin/kmax are read for the last leg in
 * GIVEN/WATKINSON binning (run.c:59-83) and written for DERIVED binning
 * (filter.c:112-113, :317-318). */
typedef struct ps_polygon {
    int n;      /* 2 (power spectrum) or 3 (bispectrum) */
    int slot;   /* field slot of the context */
    double k[3];
    double kmin[3];
    double kmax[3];
} ps_polygon;

/* Human-readable text of a status code. */
const char *ps_strerror(int status);

/* Number of GPUs visible to HIP (0 when none). */
int ps_device_count(void);

/* Whether an N^3 grid has a compiled FFT plan (N = 16, 24, 32, 48, 64, 96,
 * 128, 144, 192, 216, 256, 384 or 512). */
int ps_grid_supported(int nbins);

/* Create a context for N^3 grids of side boxsize (h^-1 Mpc) on the current
 * device.  nslots fields can be resident at once (realisations).
 * Replaces: initGrid_with_values grid.c:47-84, allocate_3D_array_fftw_complex
 * utils_fftw.c:18-30. */
int ps_ctx_create(int nbins, double boxsize, int precision, int nslots, ps_ctx **out);
/* Batches and multi-batches refer to their context: destroy them first. */
int ps_ctx_destroy(ps_ctx *ctx);

/* Upload a real N^3 field (float32 if field_is_double == 0, float64
 * otherwise; host memory unless on_device) into a slot and forward transform
 * it; the spectrum stays resident.
 * Replaces: get_FT_field input_field.c:40-81 (read_grid grid.c:121-228 +
 * fft_real_to_kspace fft.c:19-37). */
int ps_load_field(ps_ctx *ctx, int slot, const void *field, int field_is_double, int on_device);

/* Use only the phases z/|z| of the spectrum (useOnlyPhase=1).
 * Replaces: calc_phase_3D_fftw_array utils_fftw.c:78-93 (run.c:48-49); the
 * normalisation is applied to the cells a shell selects, when they are read. */
int ps_set_phase_only(ps_ctx *ctx, int phase_only);

/* Method (PS_METHOD_DIRECT by default) of the batches created afterwards
 * without explicit launch parameters, and of ps_batch_autotune.
 * Replaces: nothing in the reference (it evaluates every polygon the DIRECT
 * way, polyspectrum.c:51-161); the choice changes the order of the sums only. */
int ps_set_method(ps_ctx *ctx, int method);

/* Mixed-precision contexts: refine ill-conditioned values in FP64.  With tol
 * > 0 the context keeps a second, FP64 spectrum of every field loaded
 * afterwards (and of the field still in its upload buffer), the batches'
 * product passes also sum |prod D| (kappa = sum |prod D| / |sum prod D|;
 * every x plane up to grid 192, every fourth above, scaled by 4),
 * and ps_batch_execute evaluates every polygon whose error estimate
 * u32 (16 + 32 kappa / sqrt(#polygons)) (u32 = 2^-24) exceeds tol again
 * with the FP64 batched path (one FP64 batch per set of flagged polygons,
 * planned on first use, the last 8 sets kept) and returns those values.  tol = 0 switches it off.  PS_ERR_INVALID for FP64
 * contexts with tol > 0.
 * Replaces: nothing in the reference (FP64 throughout); the FP32 data of the
 * mixed path carry each term of a sum to about u32, which near a zero of the
 * polyspectrum (large kappa) limits the relative accuracy. */
int ps_set_refinement(ps_ctx *ctx, double tol);

/* One configuration, shaped like the reference: per leg a mask, an inverse
 * FFT of the masked spectrum and of the mask, products into N^3
 * accumulators, then the sums.  result = polyspectrum value, sums (optional,
 * 2 doubles) = (Re sum prod D, Re sum prod I).
 * Replaces: polyspectrum polyspectrum.c:51-161. */
int ps_polyspectrum(ps_ctx *ctx, int slot, int n, const double *k, double *kmin, double *kmax,
                    double kbinwidth, int kbinning_case, double *result, double *sums);

/* Replaces: num_polygons polyspectrum.c:163-252. */
int ps_num_polygons(ps_ctx *ctx, int n, const double *k, double *kmin, double *kmax, double kbinwidth,
                    int kbinning_case, double *result);

/* Condition number of each polygon's sum, kappa[p] = sum_x |prod_i D_i(x)| /
 * |sum_x prod_i D_i(x)|: the factor by which rounding errors of the filtered
 * fields are amplified in the polyspectrum value (large where B crosses
 * zero).  Evaluated through the reference-shaped path in the context's
 * precision (one configuration at a time, full grid); kmin/kmax of polys are
 * not changed.  No counterpart in the reference (polyspectrum.c:139-140 takes
 * the sums only). */
int ps_condition_numbers(ps_ctx *ctx, const ps_polygon *polys, int npoly, double kbinwidth, int kbinning_case,
                         double *kappa);

/* Batched evaluation of many configurations: shells shared by several
 * polygons are filtered and transformed once, only the k-cells a shell can
 * touch enter the first inverse passes, and the last inverse pass computes
 * the products and FP64 compensated sums of all polygons of a group without
 * writing the filtered fields.  The launch sequence is captured once in a
 * hipGraph and replayed by ps_batch_execute.
 * values[i]: polyspectrum value; npolygons[i] (optional): Npolygons;
 * sums (optional, 2*npoly): (sum prod D, sum prod I).
 * For DERIVED binning the planner writes the derived kmin/kmax back into
 * polys (as polyspectrum() does).
 * Replaces: the loop run.c:56-107 over polyspectrum()/num_polygons(). */
int ps_batch_create(ps_ctx *ctx, ps_polygon *polys, int npoly, double kbinwidth, int kbinning_case,
                    ps_batch **out);

/* Launch shape of a batch: persistent grid size of the pass kernels, Zc tiles
 * (DIRECT) or dot-pass steps (FOURIER) per work item (sets how many
 * compensated partial sums each polygon has), the divisor of the G2/Fx tile
 * width (1: 16 FP32 / 8 FP64 columns, 2: half), the schedule of the waves
 * (groups sharing one transform grid; overlap 1: waves side by side on up to
 * 4 branch streams when their buffers fit together, 2 .. 4: on at most that
 * many, 0: one after another, reusing one buffer region) and the method
 * (PS_METHOD_*). */
typedef struct ps_launch_params {
    int grid_cap;
    int chunk_tiles;
    int g2_div;
    int overlap;
    int method;
} ps_launch_params;

/* ps_batch_create with explicit launch parameters (NULL: the tuned ones for
 * this batch shape if ps_batch_autotune ran, else defaults). */
int ps_batch_create_ex(ps_ctx *ctx, ps_polygon *polys, int npoly, double kbinwidth, int kbinning_case,
                       const ps_launch_params *params, ps_batch **out);

/* Autotuner: times candidate launch shapes on this batch (the fastest of reps
 * executions each, on the host clock from the call to its results, with the
 * refinement where the context refines), times the four fastest again (the
 * median of 3 reps executions), then the best of those with the waves on at
 * most 3 and 2 branch streams and one after another, and keeps the best for
 * later ps_batch_create calls of the same batch shape (and appends it to the
 * file PS_TUNE_FILE names, if set); best_ms receives its median time. */
int ps_batch_autotune(ps_ctx *ctx, ps_polygon *polys, int npoly, double kbinwidth, int kbinning_case, int reps,
                      ps_launch_params *best, double *best_ms);
int ps_batch_params(const ps_batch *b, ps_launch_params *out);
int ps_batch_execute(ps_batch *b, double *values, double *npolygons, double *sums);
/* Values of the batch's last execution (by ps_batch_execute or as part of a
 * ps_multi), without launching anything. */
int ps_batch_results(ps_batch *b, double *values, double *npolygons, double *sums);
/* Refinement of the last ps_batch_execute: per polygon (optional arrays of
 * npoly) whether its value came from the FP64 path and the error estimate of
 * its mixed-precision value (0 without refinement).  Returns the number of
 * refined polygons (or a negative PS_ERR_*).
 * Replaces: nothing in the reference. */
int ps_batch_refinement(const ps_batch *b, int *refined, double *error_estimates);

/* Several batches -- contexts of different grid sizes, each with any number
 * of field slots (realisations) -- captured into one hipGraph on the first
 * batch's stream.  ps_multi_execute replays it and waits; read each batch's
 * values with ps_batch_results.  All contexts must be on one device.
 * Replaces: one reference run (main.c) per field and grid size. */
typedef struct ps_multi ps_multi;
int ps_multi_create(ps_batch **batches, int nbatch, ps_multi **out);
int ps_multi_execute(ps_multi *m);
double ps_multi_last_ms(const ps_multi *m);
int ps_multi_destroy(ps_multi *m);

/* Time of the last ps_batch_execute on the GPU (ms), from HIP events. */
double ps_batch_last_ms(const ps_batch *b);
/* HBM bytes the batch touches per execution by the algorithmic model
 * (reads + writes of every pass), for roofline reporting. */
double ps_batch_model_bytes(const ps_batch *b);
int ps_batch_destroy(ps_batch *b);

/* Run the batch pass by pass (not from the graph) reps times with HIP events
 * around every launch; stage_ms[PS_MAX_STAGES] receives the mean time of each
 * pass (ps_batch_stage_count of them: DIRECT G1, G2, Zc, reduce; FOURIER G1,
 * G2, Zf, Fx, dot, reduce), stage_bytes[PS_MAX_STAGES] (optional) the
 * algorithmic HBM bytes of each. */
int ps_batch_profile(ps_batch *b, int reps, double *stage_ms, double *stage_bytes);
int ps_batch_stage_count(const ps_batch *b);
const char *ps_batch_stage_name(const ps_batch *b, int stage);

/* The batch's plan (no counterpart in the reference; for tuning): its waves,
 * each on one transform grid.  ps_batch_wave_info gives wave w's grid size,
 * the legs it transforms, its polygon groups and the arena bytes it uses
 * (any output pointer may be NULL). */
int ps_batch_wave_count(const ps_batch *b);
int ps_batch_wave_info(const ps_batch *b, int w, int *grid, int *nlegs, int *ngroups, double *bytes);

/* Mean time (ms) of reps forward transforms of the field last uploaded from
 * host memory into this slot, and the algorithmic bytes of one (3 passes,
 * each reading and writing its array).  The transforms write a scratch half
 * spectrum: the resident spectra do not change.  PS_ERR_STATE unless the
 * slot's field is the last host upload of the context (a later host upload
 * to another slot, or a load from device memory, replaces or bypasses it). */
int ps_forward_profile(ps_ctx *ctx, int slot, int reps, double *ms, double *bytes);

/* Adjoint: the gradient of sum_p weights[p] * values[p] (values of the last
 * ps_batch_execute) with respect to the real field of a slot, dL/d delta(x)
 * for weights[p] = dL/dvalues[p].  For a group of polygons sharing legs A, B
 * and suffix legs t (coefficients a_t = weights * V^(n-1) N^(-3n) / sum prod I)
 * it evaluates L_A[D_B E] + L_B[D_A E] + L_W[D_A D_B] with L_X = IFFT * mask_X * FFT,
 * D_X = L_X delta, W = sum_t a_t mask_t and E = L_W delta: every filtered field
 * is recomputed, none is stored per polygon.  grad: N^3 values (float64 if
 * grad_is_double, else float32), host memory unless on_device.  The first call
 * allocates its work buffers (two half spectra and five float64 N^3 fields,
 * 6.4 GB at 512^3 in mixed precision) and the batch keeps them until
 * ps_batch_destroy.  With useOnlyPhase the filtered fields are taken from
 * p = z/|z| (utils_fftw.c:78-93) and the gradient in spectrum space Y is
 * pulled back through the normalisation, Y' = (Y - p Re(conj(p) Y)) / |z|,
 * before the final inverse transform.  No counterpart in the reference. */
int ps_batch_adjoint(ps_batch *b, int slot, const double *weights, void *grad, int grad_is_double, int on_device);

/* Transform-free check of polygon values: the same sums taken directly over
 * the closed lattice polygons m_1 + ... + m_n = 0 (mod N) with every m_i in
 * its leg's shell (the masks ps_batch_create uses; DERIVED kmin/kmax are
 * written back), from the resident spectra:
 *   sum prod D = N^3 * sum over polygons of prod H(m_i),
 *   sum prod I = N^3 * number of polygons.
 * values, npolygons and sums as ps_batch_execute; the counts are exact.
 * Cost: |shell 1| x |shell 2| cell pairs per pair of first shells
 * (bispectra), |shell 1| cells (power spectra); temporary device buffers.
 * Replaces: nothing in the reference; an independent check of
 * polyspectrum()/num_polygons() (polyspectrum.c:51-252). */
int ps_direct_check(ps_ctx *ctx, ps_polygon *polys, int npoly, double kbinwidth, int kbinning_case, double *values,
                    double *npolygons, double *sums);

/* ps_direct_check that allocates no device memory, for callers that cannot
 * call hipMalloc (code under stream capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  The cell lists,
 * cell values, partial sums and DERIVED corner extremes live in one static
 * device array of PS_DIRECT_SCRATCH_BYTES per GPU (held until the process
 * ends); a polygon set whose two first shells need more returns
 * PS_ERR_SCRATCH (about 1.4 million cells in both shells together).  Every
 * call on a GPU shares that array: calls that overlap on one GPU (from
 * several host threads) race.  Results equal ps_direct_check's bit for bit.
 * Host-side work lists use ordinary host memory. */
#define PS_DIRECT_SCRATCH_BYTES ((size_t)32 << 20)
int ps_direct_check_static(ps_ctx *ctx, ps_polygon *polys, int npoly, double kbinwidth, int kbinning_case,
                           double *values, double *npolygons, double *sums);

/* Forward-transform variants: radix plan (variant % 2: 0 = radix-3 stages,
 * radix-8 stages, then a radix-4/2 stage; 1 = the radix-4/2 stage first and
 * the radix-3 stages last) and tile width (variant < 2: the default lines per
 * LDS tile, >= 2: half of them), where the grid has them
 * (ps_forward_variant_count: 1 below N = 128).  ps_forward_autotune times
 * every variant (the fastest of reps transforms of the slot's last host
 * upload, into the scratch spectrum), keeps the fastest for this context,
 * for contexts of the same N and precision created later in the process and
 * in the file PS_TUNE_FILE names (if set).  All variants give the same
 * transform up to rounding.  No counterpart in the reference (fft.c:30 plans
 * with FFTW_ESTIMATE). */
int ps_forward_variant_count(const ps_ctx *ctx);
int ps_set_forward_variant(ps_ctx *ctx, int variant);
int ps_forward_autotune(ps_ctx *ctx, int slot, int reps, int *best_variant, double *best_ms);

/* Re-run the forward transform of a slot from its field in the upload
 * buffer (timing of the FFT passes alone); PS_ERR_STATE under the same
 * condition as ps_forward_profile. */
int ps_forward_again(ps_ctx *ctx, int slot);

/* Copy the resident half spectrum of a slot to host memory: (N/2+1)*N*N
 * complex values (complex128 for FP64 contexts, complex64 for mixed), laid
 * out [kz][kx][ky] with kz in [0, N/2]; element (kz, kx, ky) equals the
 * reference's fft_array[(kx*N + ky)*N + kz] (input_field.c:77). */
int ps_get_spectrum(ps_ctx *ctx, int slot, void *out);

/* ---- drop-in symbols (reference signatures) ---------------------------- */
/* polyspectrum.h:7.  fft_array is the full complex N^3 spectrum of a real
 * field (as get_FT_field returns it); local_n0 must equal nbins and
 * local_n0_start 0 (serial build).  Runs the FP64 path; returns NaN on error.
 * Prints polyspectrum.c:142's line (rank 0, Im/Re of both sums: 0 here, the
 * sums of real fields).  The spectrum is uploaded on the first call and kept
 * while later calls pass the same array with the same contents (a hash of the
 * whole array, taken in host memory on every call). */
double polyspectrum(int nbins, int local_n0, int local_n0_start, ps_fftw_complex *fft_array, int n, double *k,
                    double *kmin, double *kmax, double kbinwidth, int kbinningCase, double boxsize);
/* polyspectrum.h:8 */
double num_polygons(int nbins, int local_n0, int local_n0_start, int n, double *k, double *kmin, double *kmax,
                    double kbinwidth, int kbinningCase, double boxsize);

#ifdef __cplusplus
}
#endif
#endif

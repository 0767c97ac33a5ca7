// ps_internal.h -- structures shared by the kernels and the host orchestration.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/polyspectrum_gpu.h"
#include "ps_mask.hpp"

namespace ps {

constexpr int MAX_SUFFIX = 32;  // suffix legs per fused group (per-wave accumulators in LDS)

// lines (or rows) per LDS tile: 16 FP32 / 8 FP64 lines of N <= 512 fit 64 KB
__host__ __device__ constexpr int tile_lines_for(int N, bool fp64)
{
    return (N <= 64) ? ((N % 16 == 0) ? 16 : 8) : (fp64 ? 8 : 16);
}

// One filtered field ("leg"): a shell mask applied to a spectrum (D part) and
// to unit amplitude (I part), inverse transformed through the G1/G2 passes.
struct LegDev {
    MaskParams mask;
    const void *spec;  // half spectrum H (nullptr: I part only)
    int phase_only;    // normalise selected cells to z/|z|
    int Kz;            // planes kz in [0, Kz] can be non-zero
    int Kx;            // |mx| <= Kx can be non-zero
    int Bx;            // band width min(2*Kx+1, N)
    int64_t g1_begin;  // first tile of this leg in the G1 pass
    int64_t g2_begin;  // first tile of this leg in the G2 pass
    void *g1d, *g1i;   // band buffers [Kz+1][Bx][N] (D part may be null)
    void *g2d, *g2i;   // full buffers [Kz+1][N][N]
    double *scale;     // device: rms |spectrum| over the shell (D part enters G1 divided by it)
};

// A batch of polygons sharing a prefix product: output s of the group is
// sum_x P(x) * D_{suffix[s]}(x) (and the same for I), P = prod of prefix legs.
struct GroupDev {
    int npre;            // 0, 1 or 2 prefix legs
    int pre_sq;          // prefix leg preA enters squared (k1 == k2)
    int preA, preB;      // leg indices
    int nsuf;            // suffix legs, <= MAX_SUFFIX
    int suf_begin;       // index into the suffix leg list
    int suf_sq;          // suffix legs enter squared (power spectra, npre == 0)
    int out_begin;       // first output slot
    int chunk_tiles;     // Zc tiles per work item
    int pad_;
    int64_t item_begin;  // first work item of this group
};

int supported_grid(int N);
int tile_rows(int N, int fp64);
int fused_tile_rows(int N, int fp64);
int g_tile_lines(int N, int fp64);

int run_forward(int fp64, int N, const void *field, int field_is_double, void *H, const void *tw, hipStream_t st);
// G1 over tiles [g1a, g1b), G2 over tiles [g2a, g2b) of the leg table
int run_inverse_g(int fp64, int N, const LegDev *legs, int nlegs, int64_t g1a, int64_t g1b, int64_t g2a, int64_t g2b,
                  int grid_cap, int g2_div, const void *tw, hipStream_t st);
int run_zc_plain(int fp64, int N, const void *gd, const void *gi, int Kz, double *accD, double *accI, int mode,
                 double2 *partials, const void *tw, hipStream_t st);
// Zc over work items [i0, i1)
int run_zc_fused(int fp64, int N, const GroupDev *groups, int ngroups, const LegDev *legs, const int *suffix,
                 int64_t i0, int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kz_cap, const void *tw,
                 hipStream_t st);
int run_reduce_partials(const double2 *partials, int64_t stride, const int *counts, int64_t uniform_count,
                        int64_t nout, double *outD, double *outI, hipStream_t st);
int run_fill(double *p, int64_t n, double v, hipStream_t st);
int run_leg_scale(int fp64, int N, const LegDev *legs, int nlegs, hipStream_t st);
// adjoint helpers: inverse C2R of a half spectrum (destroyed) into a real
// double field; H *= sum_j w_j mask_j; a += b; out = alpha a b
int run_inverse_real(int fp64, int N, void *H, double *out, int accumulate, const void *tw, hipStream_t st);
int run_apply_masks(int fp64, int N, void *H, const MaskParams *masks, const double *w, int nm, hipStream_t st);
int run_add_spec(int fp64, void *a, const void *b, int64_t n, hipStream_t st);
int run_mul_fields(double *out, const double *a, const double *b, double alpha, int64_t n, hipStream_t st);
int run_corner_extremes(const MaskParams &mp, int N, int K, double bw, double2 *partials, int nblocks, hipStream_t st);

}  // namespace ps

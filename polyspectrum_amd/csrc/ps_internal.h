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
    return (N % 16 != 0) ? 8 : ((N <= 64) ? 16 : (fp64 ? 8 : 16));
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
    int ns;            // grid size of the spectrum (the transforms may run on a smaller grid)
    int64_t g1_begin;  // first tile of this leg in the G1 pass
    int64_t g2_begin;  // first tile of this leg in the G2 pass
    void *g1d, *g1i;   // band buffers [Kz+1][Bx][N] (D part may be null)
    void *g2d, *g2i;   // full buffers [Kz+1][N][N]
    double *scale;     // device: rms |spectrum| over the shell (D part enters G1 divided by it)
    int ivlo = -1;     // radial mask: the cells with ivlo <= m^2 <= ivhi ((1, 0): none);
    int ivhi = -1;     // ivlo < 0: test every cell with mask_eval
    int r2 = -1;       // the shell's cells have mx^2 + mz^2 <= r2: (kz, kx) lines beyond are empty (-1: no bound)
};

// A batch of polygons sharing a prefix product: output s of the group is
// sum_x P(x) * D_{suffix[s]}(x) (and the same for I), P = prod of prefix legs.
// Every leg is band limited to |m_i| <= K, so the sum over the spectrum's
// grid equals (N_spectrum / N)^3 times the sum over any grid N > K_A + K_B +
// K_t (no product term aliases onto m = 0): each group runs on the smallest
// supported such grid.
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
    double fscale;       // (N_spectrum / N)^3: the group runs on an alias-free grid N <= N_spectrum
};

// A polygon group evaluated in Fourier space.  Output o of the group is the
// real-space sum  sum_x P(x) D_o(x)  written as a sum over the half spectrum
//   S_o = sum_k w(kz) Re(conj Q(k) V(k)) m_o(k),   w = 1 on kz = 0, N/2, else 2,
// with V the field's spectrum (D sums) or 1 (I sums), m_o the suffix shell and
// Q = F[P] the forward transform of the prefix product (Parseval).
//   fft = 1 (triangles): P = product of the prefix legs in real space, taken
//     through G1, G2 and the Zf pass (inverse), then Zf, Fx and the dot pass
//     (forward), kept on the band |mx|, |my| <= Kq, kz <= Kqz.  P is band
//     limited to K_A + K_B, so a grid M > K_A + K_B + Kq carries these sums
//     without aliasing; the passes run on the smallest supported such M.
//   fft = 0 (power spectra): Q = N^3 m_A H read from the spectrum (has_A = 0:
//     Q = N^3 H, the suffix leg is its own partner).
struct SpecGroupDev {
    int fft;
    int npre, pre_sq;        // prefix legs (fft = 1)
    int preA, preB;          // indices into the pass's leg table
    int nsuf, suf_begin;     // suffix masks: masks[suf_begin + o], o < nsuf
    int out_begin;           // first output slot
    int Kq, Kqz, Bq;         // band of Q; Bq = min(2 Kq + 1, N)
    int has_A;               // fft = 0: Q carries the mask mA
    int phase_only;          // V = H / |H| on the selected cells
    int dot_lines;           // lines (D and I lines counted apart) per dot work item
    int radial;              // every suffix mask is one m^2 interval (per-line |my| ranges)
    int ns;                  // grid size of the spectrum; the passes run on the kernels' grid M <= ns
    double fscale;           // (ns / M)^3 for fft groups (sum over the ns grid = (ns/M)^3 x the M grid's), else 1
    MaskParams mA;
    const void *spec;        // half spectrum of the group's field slot
    void *qd, *qi;           // Zf output planes [Kqz+1][N][N] (D, I)
    void *bd, *bi;           // Fx output band [Kqz+1][Bq][N]
    const double *scaleA;    // amplitude scales of the prefix legs (mixed precision), may be null
    const double *scaleB;
    int64_t zf_begin;        // first Zf tile of the group
    int64_t fx_begin;        // first Fx tile
    int64_t dot_begin;       // first dot work item
};

// Transform-free check: third-leg (power spectra: second-leg) shells tested
// together against the cells that close the polygons of one pair of shells.
constexpr int DIRECT_MEMBERS = 16;
struct DirectMembers {
    MaskParams mask[DIRECT_MEMBERS];
    int nm;       // members in use
    int corners;  // some member is a cube-corner (WATKINSON) mask
    double bw;    // binwidth of every member
};

int supported_grid(int N);
int tile_rows(int N, int fp64);
int g2_tile_cols(int N, int fp64, int g2_div);  // columns of a G2 tile
int g1_dmode_lines(int N, int fp64);            // kx lines of a D-only G1 tile
int fused_tile_rows(int N, int fp64);
int g_tile_lines(int N, int fp64);

// variant: radix plan and tile width of the forward transform (0 .. forward_variants(N, fp64) - 1)
int run_forward(int fp64, int N, const void *field, int field_is_double, void *H, const void *tw, hipStream_t st,
                int variant = 0);
int forward_variants(int N, int fp64);
// G1 over tiles [g1a, g1b), G2 over tiles [g2a, g2b) of the leg table;
// dmode: G1 tiles of 2 TLC D lines and no I lines (the leg table's tile
// counts must match: D-part G2 tiles only)
int run_inverse_g(int fp64, int N, const LegDev *legs, int nlegs, int64_t g1a, int64_t g1b, int64_t g2a, int64_t g2b,
                  int grid_cap, int g2_div, int dmode, const void *tw, hipStream_t st);
int run_zc_plain(int fp64, int N, const void *gd, const void *gi, int Kz, double *accD, double *accI, int mode,
                 double2 *partials, const void *tw, hipStream_t st);
// Zc over work items [i0, i1); dmode: D parts only, two rows per line (the I
// sums are taken once per batch), where zc_dmode_supported(N, fp64)
int run_zc_fused(int fp64, int N, const GroupDev *groups, int ngroups, const LegDev *legs, const int *suffix,
                 int64_t i0, int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kz_cap, int dmode,
                 const void *tw, hipStream_t st);
int zc_dmode_supported(int N, int fp64);
int run_reduce_partials(const double2 *partials, int64_t stride, const int *counts, int64_t uniform_count,
                        int64_t nout, double *outD, double *outI, hipStream_t st);
int run_fill(double *p, int64_t n, double v, hipStream_t st);
int run_direct_values(int fp64, const void *H, int N, int phase, const int *cells, int n, double2 *out, hipStream_t st);
int direct_sum_blocks(int n1, int n2);  // partials (of DIRECT_MEMBERS each) run_direct_sums writes
int run_direct_sums(int fp64, const void *H, int N, int phase, const int *c1, const double2 *v1, int n1,
                    const int *c2, const double2 *v2, int n2, const DirectMembers &mem, double2 *partials,
                    hipStream_t st);
int run_leg_scale(int fp64, int N, const LegDev *legs, int nlegs, hipStream_t st);
// adjoint helpers: inverse C2R of a half spectrum (destroyed) into a real
// double field; H *= sum_j w_j mask_j; a += b; out = alpha a b
int run_inverse_real(int fp64, int N, void *H, double *out, int accumulate, const void *tw, hipStream_t st);
int run_apply_masks(int fp64, int N, void *H, const MaskParams *masks, const double *w, int nm, hipStream_t st);
int run_add_spec(int fp64, void *a, const void *b, int64_t n, hipStream_t st);
int run_mul_fields(double *out, const double *a, const double *b, double alpha, int64_t n, hipStream_t st);
// phase-only adjoint: H = z/|z| in place; H = (H - p Re(conj(p) H)) / |z| with p = z/|z|, z = spec
int run_phase_normalize(int fp64, void *H, int64_t n, hipStream_t st);
int run_phase_pullback(int fp64, void *H, const void *spec, int64_t n, hipStream_t st);
int run_corner_extremes(const MaskParams &mp, int N, int K, double bw, double2 *partials, int nblocks, hipStream_t st);

// Fourier-space groups: tiles [t0, t1) of the Zf and Fx passes, work items
// [i0, i1) of the dot pass.  run_shell_intervals writes, per mask s,
// ivl[2s] = (lo, hi), the m^2 interval the mask selects ((1, 0): empty or not
// radial), and ivl[2s+1].x = the number of m^2 values it selects.
int spec_zf_rows(int N, int fp64);
int spec_dot_lines(int N, int fp64);  // lines one dot workgroup transforms per step
int run_spec_zf(int fp64, int N, const SpecGroupDev *groups, int ngroups, const LegDev *legs, int64_t t0, int64_t t1,
                int grid_cap, int kz_cap, const void *tw, hipStream_t st);
int run_spec_fx(int fp64, int N, const SpecGroupDev *groups, int ngroups, int64_t t0, int64_t t1, int grid_cap,
                int div, const void *tw, hipStream_t st);
// kinds: bit 0 = the wave has radial groups, bit 1 = groups with cube-corner masks
int run_spec_dot(int fp64, int N, const SpecGroupDev *groups, int ngroups, const MaskParams *masks, const int2 *ivl,
                 int64_t i0, int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kinds, const void *tw,
                 hipStream_t st);
int run_shell_intervals(const MaskParams *masks, const int *kmax, int n, int2 *out, hipStream_t st);

}  // namespace ps

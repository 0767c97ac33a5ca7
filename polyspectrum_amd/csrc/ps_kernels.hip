// ps_kernels.hip -- gfx950 kernels of the polyspectrum estimator.
//
// Spectrum layout in HBM ("half spectrum"): H[kz][kx][ky], kz in [0, N/2],
// kx, ky in [0, N), ky fastest.  The real field is C-ordered [x][y][z]
// (grid.c:168).  Because the field is real and every shell mask is symmetric
// under m -> -m, every filtered spectrum is Hermitian and its inverse
// transform is real; the reference's complex results have imaginary parts at
// roundoff (polyspectrum.c:142 prints that ratio), so half spectra and real
// outputs carry all of the estimator.
//
// Passes of one 3D transform (each a batch of line FFTs in LDS, ps_fft_core):
//   forward : Z  (R2C along z, rows read contiguously, transposed tile write)
//             Y  (C2C along ky, contiguous lines)
//             X  (C2C along kx, stride N, tiles of adjacent ky columns)
//   inverse : G1 (C2C along ky of the masked spectrum, gathered on the fly;
//                 only planes kz <= Kz and the kx band |mx| <= Kx that a shell
//                 can touch)
//             G2 (C2C along kx from the band to all x)
//             Zc (C2R along kz with a fused epilogue: store, multiply into an
//                 accumulator, dot with an accumulator, or the multi-leg
//                 product-and-reduce of a batch of polygons)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdlib>
#include <type_traits>

#include "ps_device.hpp"
#include "ps_fft_core.hpp"
#include "ps_internal.h"
#include "ps_mask.hpp"
#include "ps_wave_fft.hpp"

namespace ps {

typedef float f32x4 __attribute__((ext_vector_type(4)));  // a pair of FP32 complex values, in registers

// ---------------------------------------------------------------- helpers
template <typename T> __device__ __forceinline__ T ld_real(const float *p, int64_t i) { return (T)p[i]; }
template <typename T> __device__ __forceinline__ T ld_real(const double *p, int64_t i) { return (T)p[i]; }

// ================================================================ forward
// Z pass: R2C along z.  Two real rows share one complex FFT (z = a + i b);
// A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / (2i).
template <typename T, typename TI, int N, int TL = tile_lines<T, N>(), int PLAN = 0>
__global__ void __launch_bounds__(NT) k_fwd_z(const TI *__restrict__ in, cplx<T> *__restrict__ H,
                                              const cplx<T> *__restrict__ tw)
{
    // TL: real rows per tile
    constexpr int TLC = TL / 2;
    __shared__ cplx<T> lds[N * TLC];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * TL, x = blockIdx.y;
    const int64_t base = ((int64_t)x * N + y0) * N;
    stage_twiddles<T, N>(stw, tw, tid);
    tile_loop<TL * N>(tid, [&](int idx) {
        const int row = idx / N, z = idx % N;
        T v = ld_real<T>(in, base + idx);
        T *p = reinterpret_cast<T *>(&lds[lidx<TLC, T>(z, row >> 1)]);
        p[row & 1] = v;
    });
    __syncthreads();
    fft_lds<T, N, TLC, -1, 0, PLAN>(lds, stw, tid);
    constexpr int NH = N / 2 + 1;
    tile_loop<NH * TLC>(tid, [&](int idx) {
        const int c = idx % TLC, kz = idx / TLC;
        const cplx<T> z1 = lds[lidx<TLC, T>(kz, c)];
        const cplx<T> z2 = lds[lidx<TLC, T>((N - kz) % N, c)];
        const cplx<T> A = mk<T>((z1.x + z2.x) * (T)0.5, (z1.y - z2.y) * (T)0.5);
        const cplx<T> B = mk<T>((z1.y + z2.y) * (T)0.5, (z2.x - z1.x) * (T)0.5);
        const int64_t o = ((int64_t)kz * N + x) * N + y0 + 2 * c;
        H[o] = A;
        H[o + 1] = B;
    });
}

// Y pass: C2C along the contiguous axis for nlines consecutive lines.
template <typename T, int N, int S, int TLC = tile_lines<T, N>(), int PLAN = 0>
__global__ void __launch_bounds__(NT) k_c2c_contig(cplx<T> *__restrict__ data, const cplx<T> *__restrict__ tw)
{
    __shared__ cplx<T> lds[N * TLC];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x;
    cplx<T> *base = data + (int64_t)blockIdx.x * TLC * N;
    stage_twiddles<T, N>(stw, tw, tid);
    tile_loop<TLC * N>(tid, [&](int idx) { lds[lidx<TLC, T>(idx % N, idx / N)] = base[idx]; });
    __syncthreads();
    fft_lds<T, N, TLC, S, 0, PLAN>(lds, stw, tid);
    tile_loop<TLC * N>(tid, [&](int idx) { base[idx] = lds[lidx<TLC, T>(idx % N, idx / N)]; });
}

// X pass: C2C along the stride-N axis; block = (ky tile, plane).
template <typename T, int N, int S, int TLC = tile_lines<T, N>(), int PLAN = 0>
__global__ void __launch_bounds__(NT) k_c2c_strided(cplx<T> *__restrict__ data, const cplx<T> *__restrict__ tw)
{
    __shared__ cplx<T> lds[N * TLC];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x;
    cplx<T> *base = data + (int64_t)blockIdx.y * N * N + blockIdx.x * TLC;
    stage_twiddles<T, N>(stw, tw, tid);
    tile_loop<TLC * N>(tid, [&](int idx) {
        const int l = idx % TLC, e = idx / TLC;
        lds[lidx<TLC, T>(e, l)] = base[(int64_t)e * N + l];
    });
    __syncthreads();
    fft_lds<T, N, TLC, S, 0, PLAN>(lds, stw, tid);
    tile_loop<TLC * N>(tid, [&](int idx) {
        const int l = idx % TLC, e = idx / TLC;
        base[(int64_t)e * N + l] = lds[lidx<TLC, T>(e, l)];
    });
}

// ================================================================ inverse

// G1: masked gather + C2C along ky.  Tile = (leg, kz, band tile of TLC kx
// lines); LDS holds TLC D-lines (mask * spectrum) and TLC I-lines (mask).
// (D parts only: k_inv_g1d below.)
// 4 wavefronts per SIMD where the transform's registers allow (N < 384): the
// gather and band passes are latency bound (measured: G1 0.57 -> 0.44 ms, G2
// 0.89 -> 0.77 ms on the 512^3 sweep)
template <typename T, int N> constexpr int band_pass_waves() { return (N >= 384) ? 2 : 4; }

template <typename T, int N>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(band_pass_waves<T, N>())))
k_inv_g1(const LegDev *__restrict__ legs, int nlegs, int64_t t0, int64_t t1, const cplx<T> *__restrict__ tw)
{
    constexpr int TLC = tile_lines<T, N>() / 2;
    constexpr int TL2 = 2 * TLC;
    __shared__ cplx<T> lds[N * TL2];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x;
    constexpr int lpt = TLC;  // kx lines per tile
    stage_twiddles<T, N>(stw, tw, tid);
    for (int64_t tile = t0 + blockIdx.x; tile < t1; tile += gridDim.x) {
        const int li = find_owner(legs, nlegs, tile, &LegDev::g1_begin);
        const LegDev &L = legs[li];
        const int nbt = (L.Bx + lpt - 1) / lpt;
        const int64_t local = tile - L.g1_begin;
        const int kz = (int)(local / nbt);
        const int j0 = (int)(local % nbt) * lpt;
        const int mz = lattice(kz, N);
        const cplx<T> *spec = reinterpret_cast<const cplx<T> *>(L.spec);
        const T inv_scale = L.scale ? (T)(1.0 / *(const PS_GLOBAL double *)L.scale) : (T)1;
        auto gather = [&](int idx) {
            const int l = idx / N, ky = idx % N;
            const int j = j0 + l;
            cplx<T> d = mk<T>(0, 0), u = mk<T>(0, 0);
            if (j < L.Bx) {
                const int mx = lattice(band_to_k(j, L.Kx, L.Bx, N), N), my = lattice(ky, N);
                const int m2 = mx * mx + my * my + mz * mz;
                if (L.ivlo >= 0 ? (m2 >= L.ivlo && m2 <= L.ivhi) : mask_eval(L.mask, mx, my, mz)) {
                    u.x = (T)1;
                    if (spec) {
                        // the cell's index on the spectrum's grid (ns >= N)
                        const int ns = L.ns;
                        d = gload<T>(spec, ((int64_t)kz * ns + (mx < 0 ? mx + ns : mx)) * ns + (my < 0 ? my + ns : my));
                        if (L.phase_only) {
                            // z / |z| (utils_fftw.c:78-93), applied to the cells a shell selects
                            const T a = hypot(d.x, d.y);
                            d = mk<T>(d.x / a, d.y / a);
                        }
                        d = mk<T>(d.x * inv_scale, d.y * inv_scale);
                    }
                }
            }
            lds[lidx<TL2, T>(ky, l)] = d;
            lds[lidx<TL2, T>(ky, l + TLC)] = u;
        };
        tile_loop_u<TLC * N, 4>(tid, gather);
        __syncthreads();
        fft_lds<T, N, TL2, +1>(lds, stw, tid);
        cplx<T> *gd = reinterpret_cast<cplx<T> *>(L.g1d);
        cplx<T> *gi = reinterpret_cast<cplx<T> *>(L.g1i);
        auto put = [&](int idx) {
            const int l = idx / N, y = idx % N;
            const int j = j0 + l;
            if (j < L.Bx) {
                const int64_t o = ((int64_t)kz * L.Bx + j) * N + y;
                if (gd) gstore<T>(gd, o, lds[lidx<TL2, T>(y, l)]);
                gstore<T>(gi, o, lds[lidx<TL2, T>(y, l + TLC)]);
            }
        };
        tile_loop_u<TLC * N, 4>(tid, put);
        __syncthreads();
    }
}

// G1 for D parts only (D-only waves): tile = (leg, kz, LPD kx lines).  The
// next tile's masked cells are gathered into registers while the current
// tile is transformed and stored, where they fit in 8 values per thread (the
// tile is halved where only the half tile fits).
template <typename T, int N> constexpr int g1d_lines_t()
{
    constexpr int full = 2 * (tile_lines<T, N>() / 2);
    constexpr int per = full * N / NT;
    return (per > 8 && per / 2 <= 8) ? full / 2 : full;
}

template <typename T, int N>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(band_pass_waves<T, N>())))
k_inv_g1d(const LegDev *__restrict__ legs, int nlegs, int64_t t0, int64_t t1, const cplx<T> *__restrict__ tw)
{
    constexpr int LPD = g1d_lines_t<T, N>();
    constexpr int TOT = LPD * N;
    constexpr int NL = (TOT + NT - 1) / NT;
    constexpr bool PF = NL <= 8;
    __shared__ cplx<T> lds[N * LPD];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x;
    stage_twiddles<T, N>(stw, tw, tid);
    struct Tile {
        int li, kz, j0;
        int empty;  // no line of the tile meets the shell's disk: nothing to transform or store
    };
    auto decode = [&](int64_t tile) -> Tile {
        const int li = find_owner(legs, nlegs, tile, &LegDev::g1_begin);
        const LegDev &L = legs[li];
        const int nbt = (L.Bx + LPD - 1) / LPD;
        const int64_t local = tile - L.g1_begin;
        Tile t{li, (int)(local / nbt), (int)(local % nbt) * LPD, 0};
        if (L.r2 >= 0) {
            // |mx| rises with the band line j up to the split s, then falls
            // as P - j (the band's negative half): its least value over the
            // tile's lines [j0, j1] is at one of the two ends
            const int j1 = min(t.j0 + LPD, L.Bx) - 1;
            const int split = (L.Bx == N) ? N / 2 : L.Kx + 1;
            int mn = N;
            if (t.j0 < split) mn = t.j0;
            if (j1 >= split) mn = min(mn, L.Bx - j1);
            const int mz = lattice(t.kz, N);
            t.empty = mn * mn > L.r2 - mz * mz;
        }
        return t;
    };
    // masked spectrum cell idx = (line l, ky) of a tile, as in k_inv_g1
    auto gather = [&](const Tile &t, int idx) -> cplx<T> {
        const LegDev &L = legs[t.li];
        const int l = idx / N, ky = idx % N;
        const int j = t.j0 + l;
        cplx<T> d = mk<T>(0, 0);
        if (j < L.Bx) {
            const int mx = lattice(band_to_k(j, L.Kx, L.Bx, N), N), my = lattice(ky, N), mz = lattice(t.kz, N);
            const int m2 = mx * mx + my * my + mz * mz;
            if (L.ivlo >= 0 ? (m2 >= L.ivlo && m2 <= L.ivhi) : mask_eval(L.mask, mx, my, mz)) {
                const cplx<T> *spec = reinterpret_cast<const cplx<T> *>(L.spec);
                const int ns = L.ns;
                d = gload<T>(spec, ((int64_t)t.kz * ns + (mx < 0 ? mx + ns : mx)) * ns + (my < 0 ? my + ns : my));
                if (L.phase_only) {
                    // z / |z| (utils_fftw.c:78-93), applied to the cells a shell selects
                    const T a = hypot(d.x, d.y);
                    d = mk<T>(d.x / a, d.y / a);
                }
                if (L.scale) {
                    const T inv = (T)(1.0 / *L.scale);
                    d = mk<T>(d.x * inv, d.y * inv);
                }
            }
        }
        return d;
    };
    auto put = [&](const Tile &t) {
        const LegDev &L = legs[t.li];
        cplx<T> *gd = reinterpret_cast<cplx<T> *>(L.g1d);
        if constexpr (sizeof(T) == 4) {
            // FP32: two adjacent y of one line per 16-byte store
            PS_GLOBAL f32x4 *g4 = (PS_GLOBAL f32x4 *)gd;
            tile_loop_u<TOT / 2, 4>(tid, [&](int c) {
                const int l = c / (N / 2), y = 2 * (c % (N / 2));
                const int j = t.j0 + l;
                if (j < L.Bx) {
                    const cplx<T> a = lds[lidx<LPD, T>(y, l)], b = lds[lidx<LPD, T>(y + 1, l)];
                    f32x4 v;
                    v.x = a.x;
                    v.y = a.y;
                    v.z = b.x;
                    v.w = b.y;
                    g4[(((int64_t)t.kz * L.Bx + j) * N + y) / 2] = v;
                }
            });
        } else {
            tile_loop_u<TOT, 4>(tid, [&](int idx) {
                const int l = idx / N, y = idx % N;
                const int j = t.j0 + l;
                if (j < L.Bx) gstore<T>(gd, ((int64_t)t.kz * L.Bx + j) * N + y, lds[lidx<LPD, T>(y, l)]);
            });
        }
    };
    if constexpr (PF) {
        int64_t tile = t0 + blockIdx.x;
        if (tile >= t1) return;
        Tile cur = decode(tile);
        cplx<T> r[NL];
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const int idx = tid + m * NT;
            r[m] = (TOT % NT == 0 || idx < TOT) ? gather(cur, idx) : mk<T>(0, 0);
        }
        while (true) {
#pragma unroll
            for (int m = 0; m < NL; ++m) {
                const int idx = tid + m * NT;
                if (TOT % NT == 0 || idx < TOT) lds[lidx<LPD, T>(idx % N, idx / N)] = r[m];
            }
            __syncthreads();
            const int64_t nt = tile + gridDim.x;
            Tile nx = cur;
            if (nt < t1) {
                nx = decode(nt);
#pragma unroll
                for (int m = 0; m < NL; ++m) {
                    const int idx = tid + m * NT;
                    r[m] = (TOT % NT == 0 || idx < TOT) ? gather(nx, idx) : mk<T>(0, 0);
                }
            }
            if (!cur.empty) {  // uniform over the workgroup
                fft_lds<T, N, LPD, +1>(lds, stw, tid);
                put(cur);
            }
            __syncthreads();
            if (nt >= t1) break;
            tile = nt;
            cur = nx;
        }
    } else {
        for (int64_t tile = t0 + blockIdx.x; tile < t1; tile += gridDim.x) {
            const Tile cur = decode(tile);
            if (cur.empty) continue;  // uniform over the workgroup
            tile_loop_u<TOT, 4>(tid, [&](int idx) { lds[lidx<LPD, T>(idx % N, idx / N)] = gather(cur, idx); });
            __syncthreads();
            fft_lds<T, N, LPD, +1>(lds, stw, tid);
            put(cur);
            __syncthreads();
        }
    }
}

// G2: C2C along kx from the band to every x.  Tile = (leg, part, kz, y tile
// of TLC adjacent columns), part 0 = D, 1 = I; stores are TLC-value segments.
template <typename T, int N, int WD>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(band_pass_waves<T, N>()))) k_inv_g2(const LegDev *__restrict__ legs, int nlegs, int64_t t0, int64_t t1, const cplx<T> *__restrict__ tw)
{
    constexpr int TLC = tile_lines<T, N>() / WD;  // WD: tile width divisor picked by the autotuner
    constexpr int TOT = TLC * N;
    constexpr int NL = (TOT + NT - 1) / NT;  // band values per thread of a tile
    // where they fit, the next tile's band values are loaded into registers
    // while the current tile is transformed and stored
    constexpr bool PF = NL <= 8;
    __shared__ cplx<T> lds[N * TLC];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x;
    constexpr int nyt = N / TLC;
    stage_twiddles<T, N>(stw, tw, tid);
    struct Tile {
        const cplx<T> *g;
        cplx<T> *o;
        int kz, y0, Kx, Bx;
        int kx2;  // band lines with mx^2 > kx2 hold no cell of the shell (read as zero)
    };
    // Tiles narrower than a 128-byte line share each line with the next y
    // tile.  Workgroups are dispatched to the 8 XCDs (8 separate L2 caches)
    // in turn, so workgroups b and b + 8 share an L2: within every complete
    // block of 16 tiles, tile 2r + h is processed as tile r + 8h, and both
    // halves of a line are read through one L2 at about the same time.
    constexpr bool PAIR = TLC * (int)sizeof(cplx<T>) < 128;
    auto decode = [&](int64_t tile) -> Tile {
        if (PAIR && (tile & ~(int64_t)15) >= t0 && (tile | 15) < t1) tile = (tile & ~(int64_t)15) | ((tile & 7) << 1) | ((tile >> 3) & 1);
        const int li = find_owner(legs, nlegs, tile, &LegDev::g2_begin);
        const LegDev &L = legs[li];
        const int64_t local = tile - L.g2_begin;
        const int64_t per_part = (int64_t)(L.Kz + 1) * nyt;
        const int part = (int)(local / per_part);
        const int64_t rest = local - part * per_part;
        Tile t;
        t.g = reinterpret_cast<const cplx<T> *>(part ? L.g1i : L.g1d);
        t.o = reinterpret_cast<cplx<T> *>(part ? L.g2i : L.g2d);
        t.kz = (int)(rest / nyt);
        t.y0 = (int)(rest % nyt) * TLC;
        t.Kx = L.Kx;
        t.Bx = L.Bx;
        const int mz = lattice(t.kz, N);
        t.kx2 = (L.r2 < 0) ? N * N : L.r2 - mz * mz;
        return t;
    };
    auto band = [&](const Tile &t, int idx) -> cplx<T> {
        const int l = idx % TLC, kx = idx / TLC;
        const int j = k_to_band(kx, t.Kx, t.Bx, N);
        const int mx = lattice(kx, N);
        return (j >= 0 && t.g && mx * mx <= t.kx2) ? gload<T>(t.g, ((int64_t)t.kz * t.Bx + j) * N + t.y0 + l)
                                                   : mk<T>(0, 0);
    };
    auto put = [&](const Tile &t) {
        if (t.o) {
            tile_loop_u<TOT, 8>(tid, [&](int idx) {
                const int l = idx % TLC, x = idx / TLC;
                gstore<T>(t.o, ((int64_t)t.kz * N + x) * N + t.y0 + l, lds[lidx<TLC, T>(x, l)]);
            });
        }
    };
    if constexpr (PF) {
        int64_t tile = t0 + blockIdx.x;
        if (tile >= t1) return;
        Tile cur = decode(tile);
        cplx<T> r[NL];
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const int idx = tid + m * NT;
            r[m] = (TOT % NT == 0 || idx < TOT) ? band(cur, idx) : mk<T>(0, 0);
        }
        while (true) {
#pragma unroll
            for (int m = 0; m < NL; ++m) {
                const int idx = tid + m * NT;
                if (TOT % NT == 0 || idx < TOT) lds[lidx<TLC, T>(idx / TLC, idx % TLC)] = r[m];
            }
            __syncthreads();
            const int64_t nt = tile + gridDim.x;
            Tile nx = cur;
            if (nt < t1) {
                nx = decode(nt);
#pragma unroll
                for (int m = 0; m < NL; ++m) {
                    const int idx = tid + m * NT;
                    r[m] = (TOT % NT == 0 || idx < TOT) ? band(nx, idx) : mk<T>(0, 0);
                }
            }
            fft_lds<T, N, TLC, +1>(lds, stw, tid);
            put(cur);
            __syncthreads();
            if (nt >= t1) break;
            tile = nt;
            cur = nx;
        }
    } else {
        for (int64_t tile = t0 + blockIdx.x; tile < t1; tile += gridDim.x) {
            const Tile cur = decode(tile);
            tile_loop_u<TOT, 8>(tid, [&](int idx) { lds[lidx<TLC, T>(idx / TLC, idx % TLC)] = band(cur, idx); });
            __syncthreads();
            fft_lds<T, N, TLC, +1>(lds, stw, tid);
            put(cur);
            __syncthreads();
        }
    }
}

// Build the complex z-line Z = XD + i XI of one y row from the two half
// spectra (Hermitian extension; only real parts at kz = 0 and N/2, as a C2R
// transform uses them).  Line l of the tile = row y0 + l.
template <typename T, int N, int TLC>
__device__ __forceinline__ void load_zline_pair(cplx<T> *lds, const cplx<T> *__restrict__ gd,
                                                const cplx<T> *__restrict__ gi, int Kz, int x, int y0, int tid)
{
    constexpr int NH = N / 2 + 1;
    for (int idx = tid; idx < NH * TLC; idx += NT) {
        const int l = idx % TLC, kz = idx / TLC;
        cplx<T> X = mk<T>(0, 0), Y = mk<T>(0, 0);
        if (kz <= Kz) {
            const int64_t o = ((int64_t)kz * N + x) * N + y0 + l;
            if (gd) X = gload<T>(gd, o);
            if (gi) Y = gload<T>(gi, o);
        }
        if (kz == 0 || kz == N / 2) {
            lds[lidx<TLC, T>(kz, l)] = mk<T>(X.x, Y.x);
        } else {
            lds[lidx<TLC, T>(kz, l)] = mk<T>(X.x - Y.y, X.y + Y.x);
            lds[lidx<TLC, T>(N - kz, l)] = mk<T>(X.x + Y.y, Y.x - X.y);
        }
    }
}

// Zc for the reference-shaped path: per tile (x, y0..y0+TLC), C2R of the
// D and I half spectra of one leg, then
//   mode 0: accD = D, accI = I
//   mode 1: accD *= D, accI *= I                   (utils_fftw.c:32-44)
//   mode 2: partial sums of accD*D and accI*I       (utils_fftw.c:60-76)
//   mode 3: partial sums of accD*D and |accD*D|      (condition number)
// Accumulators are FP64 real N^3 grids.
template <typename T, int N>
__global__ void __launch_bounds__(NT) k_inv_zc_plain(const cplx<T> *__restrict__ gd, const cplx<T> *__restrict__ gi,
                                                     int Kz, double *__restrict__ accD, double *__restrict__ accI,
                                                     int mode, double2 *__restrict__ partials,
                                                     const cplx<T> *__restrict__ tw)
{
    constexpr int TLC = tile_lines<T, N>();
    __shared__ cplx<T> lds[N * TLC];
    __shared__ double red[4 * NT / 64];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * TLC, x = blockIdx.y;
    load_zline_pair<T, N, TLC>(lds, gd, gi, Kz, x, y0, tid);
    __syncthreads();
    fft_lds<T, N, TLC, +1>(lds, tw, tid);
    const int64_t base = ((int64_t)x * N + y0) * N;
    KSum sd, si;
    sd.init();
    si.init();
    for (int idx = tid; idx < TLC * N; idx += NT) {
        const int l = idx / N, z = idx % N;
        const cplx<T> v = lds[lidx<TLC, T>(z, l)];
        const int64_t o = base + idx;
        if (mode == 0) {
            if (accD) accD[o] = (double)v.x;
            accI[o] = (double)v.y;
        } else if (mode == 1) {
            if (accD) accD[o] *= (double)v.x;
            accI[o] *= (double)v.y;
        } else if (mode == 2) {
            if (accD) sd.add(accD[o] * (double)v.x);
            si.add(accI[o] * (double)v.y);
        }
    }
    if (mode == 3) {
        // the sum of the D products and the sum of their magnitudes (the
        // numerator of the sum's condition number); the I part is not summed
        for (int idx = tid; idx < TLC * N; idx += NT) {
            const int l = idx / N, z = idx % N;
            const double p = accD[base + idx] * (double)lds[lidx<TLC, T>(z, l)].x;
            sd.add(p);
            si.add(fabs(p));
        }
    }
    if (mode >= 2) {
        double a = sd.s, b = sd.c, c = si.s, d = si.c;
        wave_sum2(a, b);
        wave_sum2(c, d);
        const int w = tid / 64;
        if ((tid & 63) == 0) {
            red[4 * w + 0] = a; red[4 * w + 1] = b; red[4 * w + 2] = c; red[4 * w + 3] = d;
        }
        __syncthreads();
        if (tid == 0) {
            KSum A, B;
            A.init();
            B.init();
            for (int i = 0; i < NT / 64; ++i) {
                A.add(red[4 * i + 0]); A.add(red[4 * i + 1]);
                B.add(red[4 * i + 2]); B.add(red[4 * i + 3]);
            }
            const int64_t bi = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
            partials[2 * bi + 0] = make_double2(A.s, A.c);
            partials[2 * bi + 1] = make_double2(B.s, B.c);
        }
    }
}

// Zc for batches of polygons: tile = (group, x, ROWS adjacent y rows).  For
// every leg the workgroup stages the leg's planes kz <= Kz of the tile
// (coalesced, ROWS adjacent values per plane) in one of two LDS buffers
// while the wavefronts transform the previous leg from the other buffer:
// each wavefront runs the C2R of (D + i I) of its LW rows with
// wavefront-private exchanges.  The prefix product P of the shared legs stays
// in registers in FP64 (the product of two FP32 values is exact there); every
// suffix leg t adds sum_z P*D_t and sum_z P_I*I_t, tree-summed over the
// wavefront and accumulated with compensation per wavefront.  A work item is
// a chunk of consecutive tiles of one group; its partial sums go to
// partials[out][chunk] as (sum, compensation).  Dynamic LDS: two staging
// buffers of kz_cap planes (kz_cap = 1 + the largest Kz of the batch).
//
// D-only mode (DM): the I parts (the inverse transforms of the bare masks)
// do not depend on the field, so their sums are taken once when a batch is
// planned; every execution then transforms the D parts only, two adjacent y
// rows per complex line (D(a) + i D(b)), so half as many lines per tile.
// y rows per wavefront of the fused pass: wave_lines, except FP32 at 144
// (6 rows of 144 = 3 paired lines held 212 VGPRs, two waves per SIMD)
template <typename T, int N> constexpr int zc_wave_rows() { return (sizeof(T) == 4 && N == 144) ? 4 : wave_lines<T, N>(); }

// D-only passes with magnitude sums (ABS): sum |prod D|, the condition
// number's numerator, is taken on every x plane of grids up to 192 (exact),
// and on the planes x = 0 mod 4 of larger ones, scaled by 4 (an estimate;
// every supported grid is a multiple of 4)
template <int N> constexpr int abs_stride() { return N <= 192 ? 1 : 4; }

template <typename T, int N, bool DM = false> struct ZcCfg {
    static constexpr int LW0 = zc_wave_rows<T, N>();  // y rows per wavefront
    static constexpr int LW = DM ? LW0 / 2 : LW0;   // transformed lines per wavefront
    // 4 wavefronts: measured faster on the 512^3 sweep than 8 (whose 16-row
    // tiles read whole 128 B lines per plane but synchronise twice as many
    // wavefronts), 2 (Zc 628 -> 676 us) or 1 (853 us)
    static constexpr int NWV = 4;
    static constexpr int BT = 64 * NWV;  // threads per workgroup
    static constexpr int ROWS = NWV * LW0;  // y rows of a tile
    static constexpr int SL = NWV * LW;     // staged lines of a tile
    static constexpr int SP = SL + 1;       // staging pitch
    static constexpr int NH = N / 2 + 1;
    // staged values per thread when prefetching (legs with more planes are
    // staged without overlap)
    static constexpr int MPT = ((NH * SL + BT - 1) / BT < 4) ? (NH * SL + BT - 1) / BT : 4;
};
template <typename T, int N> constexpr bool zc_dmode_ok() { return zc_wave_rows<T, N>() % 2 == 0; }
// Compact form of the fused pass: twiddles read from an LDS table and one
// staging buffer (the loads of the next step still overlap the transforms, in
// registers), which leaves registers and LDS for four workgroups per CU where
// the transforms fit in 128 VGPRs (FP32 D-only at power-of-two lengths 64 ..
// 256: measured 128: 74 -> 56 us, 256: 141 -> 121 us on the 512^3 sweep) and
// helps FP64 (1.68 -> 1.47 ms); FP32 at 96 and 192 keep register twiddles and
// two staging buffers (spills otherwise).
template <typename T, int N> constexpr bool zc_pow2_mid() { return N >= 64 && N <= 256 && (N & (N - 1)) == 0; }
template <typename T, int N> constexpr bool zc_compact() { return sizeof(T) == 8 || zc_pow2_mid<T, N>(); }
template <typename T, int N, bool DM> constexpr int zc_waves()
{
    return (DM && sizeof(T) == 4 && zc_pow2_mid<T, N>()) ? 4 : 1;
}

// ABS (D-only mode): partial slot 1 carries sum |prod D| per output, the
// numerator of the sum's condition number (an estimate of its rounding error)
template <typename T, int N, bool DM, bool ABS = false>
__global__ void __launch_bounds__((ZcCfg<T, N, DM>::BT)) __attribute__((amdgpu_waves_per_eu(zc_waves<T, N, DM>()))) k_inv_zc_fused(const GroupDev *__restrict__ groups,
                                                         int ngroups, const LegDev *__restrict__ legs,
                                                         const int *__restrict__ suffix, int64_t i0, int64_t i1,
                                                         double2 *__restrict__ partials, int max_chunks,
                                                         const cplx<T> *__restrict__ twg, int kz_cap, int nbuf)
{
    using C = ZcCfg<T, N, DM>;
    constexpr int LW = C::LW, NWV = C::NWV, ROWS = C::ROWS, SL = C::SL, SP = C::SP, MPT = C::MPT, BT = C::BT;
    constexpr int LP = wave_fft_pitch<T, N>();
    using WO = WaveOut<T, N, LW>;
    extern __shared__ unsigned char dyn_lds[];
    __shared__ cplx<T> wbuf[NWV * LW * LP];
    __shared__ double acc[NWV][MAX_SUFFIX][4];
    cplx<T> *stage_buf = reinterpret_cast<cplx<T> *>(dyn_lds);  // [nbuf buffers][2 parts][kz_cap * SP]
    const int part = kz_cap * SP;
    const int tid = threadIdx.x, w = tid / 64, lane = tid & 63;
    constexpr bool CMP = zc_compact<T, N>();
    __shared__ cplx<T> stw[CMP ? N : 1];
    WaveTw<T, N, LW> twr;
    if constexpr (CMP) stage_twiddles<T, N, BT>(stw, twg, tid);
    else twr.template load<+1>(twg, lane);
    cplx<T> *wb = wbuf + w * LW * LP;
    constexpr int tiles_per_x = N / ROWS;
    const int64_t ntile = (int64_t)N * tiles_per_x;
    for (int64_t item = i0 + blockIdx.x; item < i1; item += gridDim.x) {
        const int gi = find_owner(groups, ngroups, item, &GroupDev::item_begin);
        const GroupDev &G = groups[gi];
        const int chunk = (int)(item - G.item_begin);
        const int64_t t_begin = (int64_t)chunk * G.chunk_tiles;
        const int64_t t_end = (t_begin + G.chunk_tiles < ntile) ? t_begin + G.chunk_tiles : ntile;
        const int nleg = G.npre + G.nsuf;
        const int64_t nsteps = (t_end - t_begin) * nleg;  // (tile, leg) steps of this item
        for (int i = tid; i < NWV * MAX_SUFFIX * 4; i += BT) (&acc[0][0][0])[i] = 0.0;
        // staged copy of step s = (tile, leg) into buffer b
        auto leg_of = [&](int q) -> const LegDev & {
            return legs[q < G.npre ? (q == 0 ? G.preA : G.preB) : suffix[G.suf_begin + q - G.npre]];
        };
        cplx<T> pfD[MPT], pfI[MPT];
        // step (t, q) = (tile, leg): registers for up to MPT*NT staged values,
        // the rest (legs with many planes) copied directly at store time
        // staged line l of a tile: X = D and Y = I of row y0 + l, or (DM)
        // X = D of row y0 + 2l and Y = D of row y0 + 2l + 1
        auto load_xy = [&](const LegDev &L, int kz, int x, int y0, int l, cplx<T> &X, cplx<T> &Y) {
            const cplx<T> *gd = reinterpret_cast<const cplx<T> *>(L.g2d);
            if constexpr (DM) {
                const int64_t o = ((int64_t)kz * N + x) * N + y0 + 2 * l;
                X = gload<T>(gd, o);
                Y = gload<T>(gd, o + 1);
            } else {
                const cplx<T> *gi = reinterpret_cast<const cplx<T> *>(L.g2i);
                const int64_t o = ((int64_t)kz * N + x) * N + y0 + l;
                X = gd ? gload<T>(gd, o) : mk<T>(0, 0);
                Y = gi ? gload<T>(gi, o) : mk<T>(0, 0);
            }
        };
        auto fetch = [&](int64_t t, int q) {
            const LegDev &L = leg_of(q);
            const int x = (int)(t / tiles_per_x), y0 = (int)(t % tiles_per_x) * ROWS;
            const int cnt = (L.Kz + 1) * SL;
#pragma unroll
            for (int m = 0; m < MPT; ++m) {
                const int idx = tid + m * BT;
                pfD[m] = mk<T>(0, 0);
                pfI[m] = mk<T>(0, 0);
                if (idx < cnt) load_xy(L, idx / SL, x, y0, idx % SL, pfD[m], pfI[m]);
            }
        };
        auto store = [&](int64_t t, int q, int b) {
            const LegDev &L = leg_of(q);
            const int x = (int)(t / tiles_per_x), y0 = (int)(t % tiles_per_x) * ROWS;
            const int cnt = (L.Kz + 1) * SL;
            cplx<T> *sd = stage_buf + (2 * b) * part, *si = stage_buf + (2 * b + 1) * part;
            // the z line of row y is Z[k] = X[k] + i Y[k] for k <= N/2 and
            // Z[N-k] = conj X[k] + i conj Y[k]; only real parts at k = 0, N/2
            auto put = [&](int kz, int row, cplx<T> X, cplx<T> Y) {
                if (kz == 0 || kz == N / 2) {
                    sd[kz * SP + row] = mk<T>(X.x, Y.x);
                } else {
                    sd[kz * SP + row] = mk<T>(X.x - Y.y, X.y + Y.x);
                    si[kz * SP + row] = mk<T>(X.x + Y.y, Y.x - X.y);
                }
            };
#pragma unroll
            for (int m = 0; m < MPT; ++m) {
                const int idx = tid + m * BT;
                if (idx < cnt) put(idx / SL, idx % SL, pfD[m], pfI[m]);
            }
            for (int idx = tid + MPT * BT; idx < cnt; idx += BT) {
                cplx<T> X, Y;
                load_xy(L, idx / SL, x, y0, idx % SL, X, Y);
                put(idx / SL, idx % SL, X, Y);
            }
        };
        fetch(t_begin, 0);
        store(t_begin, 0, 0);
        __syncthreads();
        int q = 0;
        int64_t tcur = t_begin;
        // prefix product of the shared legs per output cell, in the precision
        // of the transforms; each suffix product is then formed in FP64
        T pd[WO::IT][WO::R], pi[WO::IT][WO::R];
        for (int64_t s = 0; s < nsteps; ++s) {
            const int b = (nbuf == 2) ? (int)(s & 1) : 0;
            const int qn = (q + 1 == nleg) ? 0 : q + 1;
            const int64_t tn = (qn == 0) ? tcur + 1 : tcur;
            if (s + 1 < nsteps) fetch(tn, qn);
            if (q == 0) {
#pragma unroll
                for (int i = 0; i < WO::IT; ++i)
#pragma unroll
                    for (int r = 0; r < WO::R; ++r) { pd[i][r] = (T)1; pi[i][r] = (T)1; }
            }
            const int Kz = leg_of(q).Kz;
            const cplx<T> *sD = stage_buf + (2 * b) * part, *sI = stage_buf + (2 * b + 1) * part;
            auto src = [&](int l, int e) -> cplx<T> {
                const int row = w * LW + l;
                const bool lo = e <= N / 2;
                const int k = lo ? e : N - e;
                if (k > Kz) return mk<T>(0, 0);
                return (lo ? sD : sI)[k * SP + row];
            };
            // the squared-leg flags are uniform: each choice is its own
            // instantiation, so no per-element selects enter the transforms
            const bool edge = Kz < N / stage_radix(N, 0);
            auto run = [&](auto &sink) {
                if constexpr (CMP) {
                    const LdsTw<T, N, LW> tw{stw};
                    if (edge) wave_fft<T, N, LW, +1, 0, true, LdsTw<T, N, LW>>(wb, tw, lane, src, sink);
                    else wave_fft<T, N, LW, +1, 0, false, LdsTw<T, N, LW>>(wb, tw, lane, src, sink);
                } else {
                    const RegTw<T, N, LW> tw{twr};
                    if (edge) wave_fft<T, N, LW, +1, 0, true, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
                    else wave_fft<T, N, LW, +1, 0, false, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
                }
            };
            if (q < G.npre) {
                auto sink1 = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] *= v.x;
                    pi[i][r] *= v.y;
                };
                auto sink2 = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] *= v.x * v.x;
                    pi[i][r] *= v.y * v.y;
                };
                if (G.pre_sq) run(sink2);
                else run(sink1);
            } else {
                // DM: both parts of a line are D of two rows, one sum (and
                // with ABS the sum of the magnitudes of its terms in si)
                double sd = 0.0, si = 0.0;
                auto sink1 = [&](auto with_abs) {
                    return [&](int i, int r, int, int, cplx<T> v) {
                        sd += (double)pd[i][r] * (double)v.x;
                        if constexpr (DM) sd += (double)pi[i][r] * (double)v.y;
                        else si += (double)pi[i][r] * (double)v.y;
                        if constexpr (DM && decltype(with_abs)::value) {
                            si += fabs((double)pd[i][r]) * fabs((double)v.x);
                            si += fabs((double)pi[i][r]) * fabs((double)v.y);
                        }
                    };
                };
                auto sink2 = [&](auto with_abs) {
                    return [&](int i, int r, int, int, cplx<T> v) {
                        const double dv = (double)v.x, iv = (double)v.y;
                        sd += (double)pd[i][r] * (dv * dv);
                        if constexpr (DM) sd += (double)pi[i][r] * (iv * iv);
                        else si += (double)pi[i][r] * (iv * iv);
                        if constexpr (DM && decltype(with_abs)::value) {
                            si += fabs((double)pd[i][r]) * (dv * dv);
                            si += fabs((double)pi[i][r]) * (iv * iv);
                        }
                    };
                };
                // the magnitude sums on every abs_stride-th x plane
                const bool ab = DM && ABS && ((int)(tcur / tiles_per_x) % abs_stride<N>() == 0);
                if (ab) {
                    auto s1 = sink1(std::true_type{});
                    auto s2 = sink2(std::true_type{});
                    if (G.suf_sq) run(s2);
                    else run(s1);
                } else {
                    auto s1 = sink1(std::false_type{});
                    auto s2 = sink2(std::false_type{});
                    if (G.suf_sq) run(s2);
                    else run(s1);
                }
                sd = wave_sum_dpp(sd);
                if (!DM || ab) si = wave_sum_dpp(si);
                if (lane == 0) {
                    const int o = q - G.npre;
                    KSum a; a.s = acc[w][o][0]; a.c = acc[w][o][1]; a.add(sd);
                    KSum c; c.s = acc[w][o][2]; c.c = acc[w][o][3]; c.add(si);
                    acc[w][o][0] = a.s; acc[w][o][1] = a.c; acc[w][o][2] = c.s; acc[w][o][3] = c.c;
                }
            }
            if (nbuf == 1) __syncthreads();  // one buffer: every wavefront has read it
            if (s + 1 < nsteps) store(tn, qn, (nbuf == 2) ? (b ^ 1) : 0);
            __syncthreads();
            q = qn;
            tcur = tn;
        }
        for (int o = tid; o < G.nsuf; o += BT) {
            KSum a, c;
            a.init();
            c.init();
            for (int ww = 0; ww < NWV; ++ww) {
                a.add(acc[ww][o][0]); a.add(acc[ww][o][1]);
                c.add(acc[ww][o][2]); c.add(acc[ww][o][3]);
            }
            // undo the legs' amplitude scales (D parts entered G1 divided by them)
            double f = 1.0;
            if (G.npre >= 1) {
                const double sa = legs[G.preA].scale ? *legs[G.preA].scale : 1.0;
                f *= G.pre_sq ? sa * sa : sa;
            }
            if (G.npre == 2 && legs[G.preB].scale) f *= *legs[G.preB].scale;
            const LegDev &Ls = legs[suffix[G.suf_begin + o]];
            const double st = Ls.scale ? *Ls.scale : 1.0;
            f *= G.suf_sq ? st * st : st;
            // the grid ratio (N_spectrum / N)^3 of a group run on a reduced grid
            const double g = G.fscale;
            f *= g;
            const int64_t po = (int64_t)(G.out_begin + o) * max_chunks + chunk;
            partials[2 * po + 0] = make_double2(a.s * f, a.c * f);
            // I parts: the grid ratio only; |prod D| (ABS): the D part's scale
            const double fi = (DM && ABS) ? f * abs_stride<N>() : g;
            partials[2 * po + 1] = make_double2(c.s * fi, c.c * fi);
        }
        __syncthreads();
    }
}

// Zc, D parts only (FP32), with the planes staged as raw pairs.  Tiles,
// work items, steps and partial sums are those of k_inv_zc_fused<T, N, true>
// (the workgroup loads the next step's planes of the tile into registers
// while its wavefronts transform the current step, then stores them into the
// other staging buffer).  The staging buffers hold the raw pairs
// (D(a), D(b)) of rows a = y0 + 2l, b = a + 1, one line l of SL per tile,
// with consecutive kz adjacent; the Hermitian extension Z[k] = D(a)[k] +
// i D(b)[k], Z[N-k] = conj D(a)[k] + i conj D(b)[k] (real parts only at k =
// 0, N/2) is formed as the first butterfly stage reads them.  One 16-byte
// store per pair, and a line pitch of 16/SL (mod 16) 16-byte words: the
// stores of a quarter wavefront (16/SL planes of all SL lines) and the reads
// of the first stage (consecutive kz of one line) fill distinct banks.
// where it measured faster than k_inv_zc_fused on the 512^3 sweep (192: 260 ->
// 241 us, 144: 96 -> 91 us, 96: 62 -> 55 us); the compact form of the latter
// keeps four wavefronts per SIMD at 64, 128 and 256 (128: 59 against 76 us)
template <typename T, int N> constexpr bool zc_pair_ok()
{
    return sizeof(T) == 4 && zc_dmode_ok<T, N>() && !zc_pow2_mid<T, N>();
}
template <int NH, int SL> constexpr int zcp_pitch()
{
    const int r = (SL >= 16) ? 1 : 16 / SL;
    int p = NH;
    while (p % 16 != r % 16) ++p;
    return p;
}

template <typename T, int N, bool ABS = false>
__global__ void __launch_bounds__((ZcCfg<T, N, true>::BT)) __attribute__((amdgpu_waves_per_eu(zc_waves<T, N, true>())))
k_zc_pairs(const GroupDev *__restrict__ groups, int ngroups, const LegDev *__restrict__ legs,
           const int *__restrict__ suffix, int64_t i0, int64_t i1, double2 *__restrict__ partials, int max_chunks,
           const cplx<T> *__restrict__ twg)
{
    using C = ZcCfg<T, N, true>;
    constexpr int LW = C::LW, NWV = C::NWV, ROWS = C::ROWS, BT = C::BT, SL = C::SL;
    constexpr int NH = N / 2 + 1;
    constexpr int SP = zcp_pitch<NH, SL>();   // 16-byte words per staged line
    constexpr int LP = wave_fft_pitch<T, N>();
    constexpr int MPT = (NH * SL + BT - 1) / BT;  // staged pairs per thread
    using WO = WaveOut<T, N, LW>;
    static_assert(sizeof(T) == 4, "FP32 pairs of complex values are 16 bytes");
    __shared__ f32x4 stage[2][SL * SP];
    __shared__ cplx<T> wbuf[NWV * LW * LP];
    __shared__ double acc[NWV][MAX_SUFFIX][2];
    __shared__ float acca[ABS ? NWV : 1][ABS ? MAX_SUFFIX : 1];  // sum |prod D| (an estimate: FP32, no compensation)
    const int tid = threadIdx.x, w = tid / 64, lane = tid & 63;
    constexpr bool CMP = zc_compact<T, N>();
    __shared__ cplx<T> stw[CMP ? N : 1];
    WaveTw<T, N, LW> twr;
    if constexpr (CMP) stage_twiddles<T, N, BT>(stw, twg, tid);
    else twr.template load<+1>(twg, lane);
    cplx<T> *wb = wbuf + w * LW * LP;
    constexpr int tiles_per_x = N / ROWS;
    const int64_t ntile = (int64_t)N * tiles_per_x;
    for (int64_t item = i0 + blockIdx.x; item < i1; item += gridDim.x) {
        const int gi = find_owner(groups, ngroups, item, &GroupDev::item_begin);
        const GroupDev &G = groups[gi];
        const int chunk = (int)(item - G.item_begin);
        const int64_t t_begin = (int64_t)chunk * G.chunk_tiles;
        const int64_t t_end = (t_begin + G.chunk_tiles < ntile) ? t_begin + G.chunk_tiles : ntile;
        const int nleg = G.npre + G.nsuf;
        const int64_t nsteps = (t_end - t_begin) * nleg;  // (tile, leg) steps of this item
        for (int i = tid; i < NWV * MAX_SUFFIX * 2; i += BT) (&acc[0][0][0])[i] = 0.0;
        if constexpr (ABS)
            for (int i = tid; i < NWV * MAX_SUFFIX; i += BT) (&acca[0][0])[i] = 0.f;
        auto leg_of = [&](int q) -> const LegDev & {
            return legs[q < G.npre ? (q == 0 ? G.preA : G.preB) : suffix[G.suf_begin + q - G.npre]];
        };
        f32x4 pf[MPT];
        // the tile's rows y0 .. y0 + 2 SL - 1 of plane kz at x, as SL pairs of
        // complex values (16 bytes each; y0 is even)
        auto fetch = [&](int64_t t, int q) {
            const LegDev &L = leg_of(q);
            const int x = (int)(t / tiles_per_x), y0 = (int)(t % tiles_per_x) * ROWS;
            const int cnt = (L.Kz + 1) * SL;
            const PS_GLOBAL f32x4 *gd = (const PS_GLOBAL f32x4 *)L.g2d;
#pragma unroll
            for (int m = 0; m < MPT; ++m) {
                const int idx = tid + BT * m;
                if (idx < cnt) pf[m] = gd[((int64_t)(idx / SL) * N + x) * (N / 2) + y0 / 2 + idx % SL];
            }
        };
        auto store = [&](int q, int b) {
            const int cnt = (leg_of(q).Kz + 1) * SL;
#pragma unroll
            for (int m = 0; m < MPT; ++m) {
                const int idx = tid + BT * m;
                if (idx < cnt) stage[b][(idx % SL) * SP + idx / SL] = pf[m];
            }
        };
        fetch(t_begin, 0);
        store(0, 0);
        __syncthreads();
        int q = 0;
        int64_t tcur = t_begin;
        T pd[WO::IT][WO::R], pi[WO::IT][WO::R];
        for (int64_t s = 0; s < nsteps; ++s) {
            const int b = (int)(s & 1);
            const int qn = (q + 1 == nleg) ? 0 : q + 1;
            const int64_t tn = (qn == 0) ? tcur + 1 : tcur;
            if (s + 1 < nsteps) fetch(tn, qn);
            if (q == 0) {
#pragma unroll
                for (int i = 0; i < WO::IT; ++i)
#pragma unroll
                    for (int r = 0; r < WO::R; ++r) { pd[i][r] = (T)1; pi[i][r] = (T)1; }
            }
            const int Kz = leg_of(q).Kz;
            const f32x4 *sb = stage[b] + (w * LW) * SP;
            auto src = [&](int l, int e) -> cplx<T> {
                const bool lo = e <= N / 2;
                const int k = lo ? e : N - e;
                if (k > Kz) return mk<T>(0, 0);
                const f32x4 v = sb[l * SP + k];
                const bool edge_k = (k == 0) || (2 * k == N);
                const T xi = edge_k ? (T)0 : v.y, yi = edge_k ? (T)0 : v.w;
                return lo ? mk<T>(v.x - yi, xi + v.z) : mk<T>(v.x + yi, v.z - xi);
            };
            const bool edge = Kz < N / stage_radix(N, 0);
            auto run = [&](auto &sink) {
                if constexpr (CMP) {
                    const LdsTw<T, N, LW> tw{stw};
                    if (edge) wave_fft<T, N, LW, +1, 0, true, LdsTw<T, N, LW>>(wb, tw, lane, src, sink);
                    else wave_fft<T, N, LW, +1, 0, false, LdsTw<T, N, LW>>(wb, tw, lane, src, sink);
                } else {
                    const RegTw<T, N, LW> tw{twr};
                    if (edge) wave_fft<T, N, LW, +1, 0, true, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
                    else wave_fft<T, N, LW, +1, 0, false, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
                }
            };
            if (q < G.npre) {
                auto sink1 = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] *= v.x;
                    pi[i][r] *= v.y;
                };
                auto sink2 = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] *= v.x * v.x;
                    pi[i][r] *= v.y * v.y;
                };
                if (G.pre_sq) run(sink2);
                else run(sink1);
            } else {
                // both parts of a line are D of two rows: one sum
                double sd = 0.0, sa = 0.0;
                auto sink1 = [&](auto with_abs) {
                    return [&](int i, int r, int, int, cplx<T> v) {
                        sd += (double)pd[i][r] * (double)v.x;
                        sd += (double)pi[i][r] * (double)v.y;
                        if constexpr (decltype(with_abs)::value) {
                            sa += fabs((double)pd[i][r]) * fabs((double)v.x);
                            sa += fabs((double)pi[i][r]) * fabs((double)v.y);
                        }
                    };
                };
                auto sink2 = [&](auto with_abs) {
                    return [&](int i, int r, int, int, cplx<T> v) {
                        const double dv = (double)v.x, iv = (double)v.y;
                        sd += (double)pd[i][r] * (dv * dv);
                        sd += (double)pi[i][r] * (iv * iv);
                        if constexpr (decltype(with_abs)::value) {
                            sa += fabs((double)pd[i][r]) * (dv * dv);
                            sa += fabs((double)pi[i][r]) * (iv * iv);
                        }
                    };
                };
                // the magnitude sums on every abs_stride-th x plane
                const bool ab = ABS && ((int)(tcur / tiles_per_x) % abs_stride<N>() == 0);
                if (ab) {
                    auto s1 = sink1(std::true_type{});
                    auto s2 = sink2(std::true_type{});
                    if (G.suf_sq) run(s2);
                    else run(s1);
                } else {
                    auto s1 = sink1(std::false_type{});
                    auto s2 = sink2(std::false_type{});
                    if (G.suf_sq) run(s2);
                    else run(s1);
                }
                sd = wave_sum_dpp(sd);
                if (ab) sa = wave_sum_dpp(sa);
                if (lane == 0) {
                    const int o = q - G.npre;
                    if constexpr (ABS) acca[w][o] += (float)sa;
                    KSum a;
                    a.s = acc[w][o][0];
                    a.c = acc[w][o][1];
                    a.add(sd);
                    acc[w][o][0] = a.s;
                    acc[w][o][1] = a.c;
                }
            }
            if (s + 1 < nsteps) store(qn, b ^ 1);
            __syncthreads();
            q = qn;
            tcur = tn;
        }
        for (int o = tid; o < G.nsuf; o += BT) {
            KSum a;
            a.init();
            for (int ww = 0; ww < NWV; ++ww) {
                a.add(acc[ww][o][0]);
                a.add(acc[ww][o][1]);
            }
            // undo the legs' amplitude scales (D parts entered G1 divided by them)
            double f = 1.0;
            if (G.npre >= 1) {
                const double sa = legs[G.preA].scale ? *legs[G.preA].scale : 1.0;
                f *= G.pre_sq ? sa * sa : sa;
            }
            if (G.npre == 2 && legs[G.preB].scale) f *= *legs[G.preB].scale;
            const LegDev &Ls = legs[suffix[G.suf_begin + o]];
            const double sc = Ls.scale ? *Ls.scale : 1.0;
            f *= G.suf_sq ? sc * sc : sc;
            // the grid ratio (N_spectrum / N)^3 of a group run on a reduced grid
            f *= G.fscale;
            const int64_t po = (int64_t)(G.out_begin + o) * max_chunks + chunk;
            partials[2 * po + 0] = make_double2(a.s * f, a.c * f);
            double sabs = 0.0;
            if constexpr (ABS)
                for (int ww = 0; ww < NWV; ++ww) sabs += (double)acca[ww][o];
            partials[2 * po + 1] = make_double2(sabs * f * abs_stride<N>(), 0.0);
        }
        __syncthreads();
    }
}

// Zc, D parts only (FP32), with the planes copied into LDS by the memory
// system (global_load_lds_dwordx4: no registers hold data in flight) and
// three staging buffers, so that each step's planes are requested two steps
// before they are transformed.  Tiles, work items, steps, transforms, sinks
// and partial sums are those of k_zc_pairs.
//
// Step s of a work item runs as: wait for this wavefront's copies of step s
// (s_waitcnt vmcnt with the number of copy instructions it has issued since),
// workgroup barrier (every wavefront's copies of step s are in LDS, and every
// wavefront has finished reading the buffer of step s - 1), issue the copies
// of step s + 2 into that buffer, transform step s.  The copies are issued in
// inline assembly, out of the compiler's sight: its own wait insertion would
// otherwise wait for every copy in flight before each read of the staging
// LDS.  The kernel waits for them itself, with the count of copy
// instructions each wavefront issued; any other vector memory instruction
// the compiler places in between only makes the wait longer.
//
// One copy instruction moves 64 lanes x 16 bytes to consecutive LDS slots
// (M0 + 16 lane), so the LDS order is fixed by which global pair each lane
// reads.  Slot s of a step holds plane kz = s / SL of staged line l = ((s mod
// SL) - kz) mod SL: the SL slots of a plane are one 16 SL-byte segment of
// global memory (whole 128-byte lines), and the first butterfly stage, which
// reads consecutive kz of one line, finds them in distinct 16-byte bank
// groups of every 256 bytes.
//
// Used where k_zc_pairs is, and at 128, where it measured faster than the
// compact form of k_inv_zc_fused (Zc 0.410 -> 0.400 ms on the 512^3 sweep,
// with three workgroups per CU against four).
template <typename T, int N> constexpr bool zc_dma_ok()
{
    return (zc_pair_ok<T, N>() || (sizeof(T) == 4 && zc_dmode_ok<T, N>() && N == 128)) &&
           (ZcCfg<T, N, true>::SL % 8 == 0) && N >= 96;
}

// copy 16 bytes per active lane from g to LDS byte offset lds_base + 16 lane
// (lds_base uniform over the wavefront); M0 is restored afterwards
__device__ __forceinline__ void lds_copy16(const PS_GLOBAL void *g, unsigned lds_base)
{
    unsigned saved;
    asm volatile("s_mov_b32 %0, m0\n\t"
                 "s_mov_b32 m0, %2\n\t"
                 "s_nop 0\n\t"
                 "global_load_lds_dwordx4 %1, off\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(saved)
                 : "v"(g), "s"(lds_base)
                 : "memory");
}

// wait until at most n vector memory instructions of this wavefront are
// outstanding (n uniform; counts above 7 wait for 7)
__device__ __forceinline__ void wait_vm_at_most(int n)
{
    switch (n) {
    case 0: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    case 1: asm volatile("s_waitcnt vmcnt(1)" ::: "memory"); break;
    case 2: asm volatile("s_waitcnt vmcnt(2)" ::: "memory"); break;
    case 3: asm volatile("s_waitcnt vmcnt(3)" ::: "memory"); break;
    case 4: asm volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
    case 5: asm volatile("s_waitcnt vmcnt(5)" ::: "memory"); break;
    case 6: asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(7)" ::: "memory"); break;
    }
}

// the value of lane 0 (all lanes active), through two scalar registers
__device__ __forceinline__ double lane0_f64(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// NBUF staging buffers: copies are requested NBUF - 1 steps ahead (3: two
// steps; 2: one step), with the registers capped for four waves per SIMD
// where the LDS allows four workgroups per CU (two buffers, or three at 144)
template <typename T, int N, bool ABS = false, int NBUF = 3>
__global__ void __launch_bounds__((ZcCfg<T, N, true>::BT))
__attribute__((amdgpu_waves_per_eu((NBUF == 2 || N == 144) ? 4 : zc_waves<T, N, true>())))
k_zc_dma(const GroupDev *__restrict__ groups, int ngroups, const LegDev *__restrict__ legs,
         const int *__restrict__ suffix, int64_t i0, int64_t i1, double2 *__restrict__ partials, int max_chunks,
         const cplx<T> *__restrict__ twg)
{
    using C = ZcCfg<T, N, true>;
    constexpr int LW = C::LW, NWV = C::NWV, ROWS = C::ROWS, BT = C::BT, SL = C::SL;
    constexpr int NH = N / 2 + 1;
    constexpr int LP = wave_fft_pitch<T, N>();
    constexpr int SLOTS = NH * SL;                    // 16-byte slots of a staging buffer
    constexpr int MPT = (SLOTS + BT - 1) / BT;        // copy instructions per wavefront and step, at most
    static_assert(NBUF == 2 || NBUF == 3, "one or two steps ahead");
    using WO = WaveOut<T, N, LW>;
    static_assert(sizeof(T) == 4, "FP32 pairs of complex values are 16 bytes");
    static_assert(SL % 8 == 0 && (SL & (SL - 1)) == 0, "slot rotation: whole 128-byte segments per plane, SL a power of two");
    static_assert(MPT <= 7, "waits count at most 7 copy instructions (one step's)");
    static_assert(MAX_SUFFIX <= 64, "one lane per suffix");
    __shared__ f32x4 stage[NBUF][SLOTS];
    __shared__ cplx<T> wbuf[NWV * LW * LP];
    __shared__ double acc[NWV][MAX_SUFFIX][2];
    __shared__ float acca[ABS ? NWV : 1][ABS ? MAX_SUFFIX : 1];  // sum |prod D| (an estimate: FP32, no compensation)
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);
    WaveTw<T, N, LW> twr;
    twr.template load<+1>(twg, lane);
    // the twiddles have arrived before the first copy is issued (an empty
    // use of each register makes the compiler wait for its load here): no
    // waits for them are left inside the step loop, where they would also
    // wait for copies in flight
#pragma unroll
    for (int i = 0; i < WaveTw<T, N, LW>::total(); ++i) asm volatile("" ::"v"(twr.w[i].x), "v"(twr.w[i].y));
    cplx<T> *wb = wbuf + w * LW * LP;
    constexpr int tiles_per_x = N / ROWS;
    const int64_t ntile = (int64_t)N * tiles_per_x;
    const unsigned stage_lds = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void *)&stage[0][0];
    // the group, leg and suffix tables through scalar loads (a vector load's
    // wait would also wait for the copies in flight)
    const PS_CONST GroupDev *cgroups = const_tab(groups);
    const PS_CONST LegDev *clegs = const_tab(legs);
    const PS_CONST int *csuffix = const_tab(suffix);
    for (int64_t item = i0 + blockIdx.x; item < i1; item += gridDim.x) {
        int gi = 0;
        {
            int hi = ngroups - 1;
            while (gi < hi) {
                const int mid = (gi + hi + 1) >> 1;
                if (cgroups[mid].item_begin <= item) gi = mid;
                else hi = mid - 1;
            }
        }
        const PS_CONST GroupDev &G = cgroups[gi];
        const int chunk = (int)(item - G.item_begin);
        const int64_t t_begin = (int64_t)chunk * G.chunk_tiles;
        const int64_t t_end = (t_begin + G.chunk_tiles < ntile) ? t_begin + G.chunk_tiles : ntile;
        const int nleg = G.npre + G.nsuf;
        const int64_t nsteps = (t_end - t_begin) * nleg;  // (tile, leg) steps of this item
        // this wavefront's sums of suffix o in lane o's registers (the
        // compensated sum of lane 0's wave totals, step by step)
        KSum ks;
        ks.init();
        float ka = 0.f;
        auto leg_of = [&](int q) -> const PS_CONST LegDev & {
            return clegs[q < G.npre ? (q == 0 ? G.preA : G.preB) : csuffix[G.suf_begin + q - G.npre]];
        };
        // issue the copies of step (t, q) into buffer b; returns their count
        auto issue = [&](int64_t t, int q, int b) -> int {
            const PS_CONST LegDev &L = leg_of(q);
            const int x = (int)(t / tiles_per_x), y0 = (int)(t % tiles_per_x) * ROWS;
            const int cnt = (L.Kz + 1) * SL;
            const PS_GLOBAL f32x4 *gd = (const PS_GLOBAL f32x4 *)L.g2d;
            int n = 0;
#pragma unroll
            for (int m = 0; m < MPT; ++m) {
                const int base = BT * m + 64 * w;  // uniform over the wavefront
                if (base < cnt) {
                    const int s = base + lane;
                    if (s < cnt) {
                        const int kz = s / SL;
                        const int l = ((s % SL) - kz % SL + SL) % SL;
                        lds_copy16(gd + ((int64_t)kz * N + x) * (N / 2) + y0 / 2 + l,
                                   stage_lds + 16u * (unsigned)(b * SLOTS + base));
                    }
                    ++n;
                }
            }
            return n;
        };
        // steps s, s + 1, s + 2 as (tile, leg) and the copy counts of s + 1, s + 2
        int q0 = 0;
        int64_t t0s = t_begin;
        auto next_of = [&](int64_t t, int q, int64_t &tn, int &qn) {
            qn = (q + 1 == nleg) ? 0 : q + 1;
            tn = (qn == 0) ? t + 1 : t;
        };
        int64_t t1s, t2s;
        int q1, q2;
        next_of(t0s, q0, t1s, q1);
        next_of(t1s, q1, t2s, q2);
        int n1 = 0, n2 = 0;
        issue(t0s, q0, 0);
        if (NBUF == 3 && 1 < nsteps) n1 = issue(t1s, q1, 1);
        T pd[WO::IT][WO::R], pi[WO::IT][WO::R];
        for (int64_t s = 0; s < nsteps; ++s) {
            const int b = (int)(s % NBUF);
            // this wavefront's copies of step s are in LDS once at most the
            // n1 instructions issued after them are outstanding
            wait_vm_at_most(n1);
            __syncthreads();
            if constexpr (NBUF == 3) {
                if (s + 2 < nsteps) n2 = issue(t2s, q2, (int)((s + 2) % NBUF));
                else n2 = 0;
            } else {
                if (s + 1 < nsteps) issue(t1s, q1, (int)((s + 1) % NBUF));
                n2 = 0;  // nothing is issued between a step's copies and its wait
            }
            const int q = q0;
            if (q == 0) {
#pragma unroll
                for (int i = 0; i < WO::IT; ++i)
#pragma unroll
                    for (int r = 0; r < WO::R; ++r) { pd[i][r] = (T)1; pi[i][r] = (T)1; }
            }
            const int Kz = leg_of(q).Kz;
            const f32x4 *sb = stage[b];
            auto src = [&](int l, int e) -> cplx<T> {
                const bool lo = e <= N / 2;
                const int k = lo ? e : N - e;
                if (k > Kz) return mk<T>(0, 0);
                const int line = w * LW + l;
                const f32x4 v = sb[k * SL + ((line + k) & (SL - 1))];
                const bool edge_k = (k == 0) || (2 * k == N);
                const T xi = edge_k ? (T)0 : v.y, yi = edge_k ? (T)0 : v.w;
                return lo ? mk<T>(v.x - yi, xi + v.z) : mk<T>(v.x + yi, v.z - xi);
            };
            const bool edge = Kz < N / stage_radix(N, 0);
            auto run = [&](auto &sink) {
                const RegTw<T, N, LW> tw{twr};
                if (edge) wave_fft<T, N, LW, +1, 0, true, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
                else wave_fft<T, N, LW, +1, 0, false, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
            };
            if (q < G.npre) {
                auto sink1 = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] *= v.x;
                    pi[i][r] *= v.y;
                };
                auto sink2 = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] *= v.x * v.x;
                    pi[i][r] *= v.y * v.y;
                };
                if (G.pre_sq) run(sink2);
                else run(sink1);
            } else {
                double sd = 0.0;
                T sa = 0;  // magnitude sums: an estimate, FP32 products and per-lane sums suffice
                auto sink1 = [&](auto with_abs) {
                    return [&](int i, int r, int, int, cplx<T> v) {
                        sd += (double)pd[i][r] * (double)v.x;
                        sd += (double)pi[i][r] * (double)v.y;
                        if constexpr (decltype(with_abs)::value) {
                            sa += fabs(pd[i][r]) * fabs(v.x);
                            sa += fabs(pi[i][r]) * fabs(v.y);
                        }
                    };
                };
                auto sink2 = [&](auto with_abs) {
                    return [&](int i, int r, int, int, cplx<T> v) {
                        const double dv = (double)v.x, iv = (double)v.y;
                        sd += (double)pd[i][r] * (dv * dv);
                        sd += (double)pi[i][r] * (iv * iv);
                        if constexpr (decltype(with_abs)::value) {
                            sa += fabs(pd[i][r]) * (v.x * v.x);
                            sa += fabs(pi[i][r]) * (v.y * v.y);
                        }
                    };
                };
                // the magnitude sums on every abs_stride-th x plane
                const bool ab = ABS && ((int)(t0s / tiles_per_x) % abs_stride<N>() == 0);
                if (ab) {
                    auto s1 = sink1(std::true_type{});
                    auto s2 = sink2(std::true_type{});
                    if (G.suf_sq) run(s2);
                    else run(s1);
                } else {
                    auto s1 = sink1(std::false_type{});
                    auto s2 = sink2(std::false_type{});
                    if (G.suf_sq) run(s2);
                    else run(s1);
                }
                sd = lane0_f64(wave_sum_dpp(sd));
                double sat = 0.0;
                if (ab) sat = lane0_f64(wave_sum_dpp((double)sa));
                if (lane == q - G.npre) {
                    if constexpr (ABS) ka += (float)sat;
                    ks.add(sd);
                }
            }
            t0s = t1s; q0 = q1;
            t1s = t2s; q1 = q2;
            next_of(t1s, q1, t2s, q2);
            n1 = n2;
        }
        if (lane < MAX_SUFFIX) {
            acc[w][lane][0] = ks.s;
            acc[w][lane][1] = ks.c;
            if constexpr (ABS) acca[w][lane] = ka;
        }
        __syncthreads();
        for (int o = tid; o < G.nsuf; o += BT) {
            KSum a;
            a.init();
            for (int ww = 0; ww < NWV; ++ww) {
                a.add(acc[ww][o][0]);
                a.add(acc[ww][o][1]);
            }
            // undo the legs' amplitude scales (D parts entered G1 divided by them)
            double f = 1.0;
            if (G.npre >= 1) {
                const double sa = legs[G.preA].scale ? *legs[G.preA].scale : 1.0;
                f *= G.pre_sq ? sa * sa : sa;
            }
            if (G.npre == 2 && legs[G.preB].scale) f *= *legs[G.preB].scale;
            const LegDev &Ls = legs[suffix[G.suf_begin + o]];
            const double sc = Ls.scale ? *Ls.scale : 1.0;
            f *= G.suf_sq ? sc * sc : sc;
            // the grid ratio (N_spectrum / N)^3 of a group run on a reduced grid
            f *= G.fscale;
            const int64_t po = (int64_t)(G.out_begin + o) * max_chunks + chunk;
            partials[2 * po + 0] = make_double2(a.s * f, a.c * f);
            double sabs = 0.0;
            if constexpr (ABS)
                for (int ww = 0; ww < NWV; ++ww) sabs += (double)acca[ww][o];
            partials[2 * po + 1] = make_double2(sabs * f * abs_stride<N>(), 0.0);
        }
        __syncthreads();
    }
}

// G2 for the D-only waves in FP32, with the band values copied into LDS by
// the memory system (global_load_lds_dwordx4, 16 bytes = two adjacent
// columns per lane) one tile ahead, into three tile buffers.  Tiles, the
// XCD pairing of half-line tiles and the transform are those of k_inv_g2;
// the tile layout is the plain interleaving e * TLC + l (every access of
// the tile has the column index fastest).  Band lines outside the band or
// the shell's disk are copied from a zero block.  Tile i runs as: wait for
// every vector memory instruction of this wavefront (its copies of tile i,
// and the stores of tile i - 2, issued a transform ago), barrier, copies of
// tile i + 1 into buffer (i + 1) mod 3, stores of tile i - 1 from its
// buffer, transform of tile i.  Global stores and loads share one counter
// and may complete out of order, so the wait is for all of them; issuing the
// stores before the transform gives them its duration to complete.
__device__ __align__(16) float g_zero16[4 * 64];  // zeros: the copy source of empty band lines

// where the three tile buffers leave LDS for two workgroups per CU
template <typename T, int N, int WD> constexpr bool g2_dma_ok()
{
    constexpr int TLC = tile_lines<T, N>() / WD;
    return sizeof(T) == 4 && TLC % 2 == 0 && (3 * TLC + 1) * N * (int)sizeof(cplx<T>) <= 80 * 1024;
}

template <typename T, int N, int WD>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(band_pass_waves<T, N>())))
k_inv_g2_dma(const LegDev *__restrict__ legs, int nlegs, int64_t t0, int64_t t1, const cplx<T> *__restrict__ tw)
{
    static_assert(sizeof(T) == 4, "16-byte copies hold two FP32 complex values");
    constexpr int TLC = tile_lines<T, N>() / WD;
    constexpr int TOT = TLC * N;     // complex values of a tile
    constexpr int CH = TOT / 2;      // 16-byte pairs of a tile
    constexpr int CPT = (CH + NT - 1) / NT;
    constexpr int HP = TLC / 2;      // pairs per band line of a tile
    static_assert(TLC % 2 == 0, "pairs of adjacent columns");
    __shared__ __align__(16) cplx<T> tb[3][TOT];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);
    constexpr int nyt = N / TLC;
    stage_twiddles<T, N>(stw, tw, tid);
    const PS_CONST LegDev *clegs = const_tab(legs);
    const unsigned tb_lds = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void *)&tb[0][0];
    const PS_GLOBAL f32x4 *zero = (const PS_GLOBAL f32x4 *)g_zero16;
    struct Tile {
        const cplx<T> *g;
        cplx<T> *o;
        int kz, y0, Kx, Bx, kx2;
    };
    constexpr bool PAIR = TLC * (int)sizeof(cplx<T>) < 128;
    auto decode = [&](int64_t tile) -> Tile {
        if (PAIR && (tile & ~(int64_t)15) >= t0 && (tile | 15) < t1) tile = (tile & ~(int64_t)15) | ((tile & 7) << 1) | ((tile >> 3) & 1);
        int li = 0;
        {
            int hi = nlegs - 1;
            while (li < hi) {
                const int mid = (li + hi + 1) >> 1;
                if (clegs[mid].g2_begin <= tile) li = mid;
                else hi = mid - 1;
            }
        }
        const PS_CONST LegDev &L = clegs[li];
        const int64_t local = tile - L.g2_begin;
        const int64_t per_part = (int64_t)(L.Kz + 1) * nyt;
        const int part = (int)(local / per_part);
        const int64_t rest = local - part * per_part;
        Tile t;
        t.g = reinterpret_cast<const cplx<T> *>(part ? L.g1i : L.g1d);
        t.o = reinterpret_cast<cplx<T> *>(part ? L.g2i : L.g2d);
        t.kz = (int)(rest / nyt);
        t.y0 = (int)(rest % nyt) * TLC;
        t.Kx = L.Kx;
        t.Bx = L.Bx;
        const int mz = lattice(t.kz, N);
        t.kx2 = (L.r2 < 0) ? N * N : L.r2 - mz * mz;
        return t;
    };
    auto issue = [&](const Tile &t, int b) {
        const PS_GLOBAL f32x4 *g = (const PS_GLOBAL f32x4 *)t.g;
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const int base = NT * m + 64 * w;  // uniform over the wavefront
            if (CH % NT == 0 || base < CH) {
                const int c = base + lane;
                if (CH % NT == 0 || c < CH) {
                    const int kx = c / HP, l = 2 * (c % HP);
                    const int j = k_to_band(kx, t.Kx, t.Bx, N);
                    const int mx = lattice(kx, N);
                    const PS_GLOBAL f32x4 *src = (j >= 0 && g && mx * mx <= t.kx2)
                                                     ? g + (((int64_t)t.kz * t.Bx + j) * N + t.y0 + l) / 2
                                                     : zero + lane;
                    lds_copy16(src, tb_lds + 16u * (unsigned)(b * CH + base));
                }
            }
        }
    };
    // 16-byte stores of two adjacent columns, read from LDS as one pair
    // (element x of line l at x * TLC + l); non-temporal: the planes are read
    // once, by the fused pass after this one (G2 0.382 -> 0.372 ms and the
    // fused pass 0.395 -> 0.391 ms on the 512^3 sweep against plain stores)
    auto put = [&](const Tile &t, const cplx<T> *buf) {
        if (t.o) {
            const f32x4 *b4 = reinterpret_cast<const f32x4 *>(buf);
            PS_GLOBAL f32x4 *o4 = (PS_GLOBAL f32x4 *)t.o;
            tile_loop_u<CH, 8>(tid, [&](int c) {
                const int x = c / HP, l = 2 * (c % HP);
                __builtin_nontemporal_store(b4[c], &o4[(((int64_t)t.kz * N + x) * N + t.y0 + l) / 2]);
            });
        }
    };
    int64_t tile = t0 + blockIdx.x;
    if (tile >= t1) return;
    Tile cur = decode(tile);
    Tile prev = cur;
    issue(cur, 0);
    int i = 0;
    for (;; ++i) {
        const int b = i % 3;
        // the copies of this tile (and the stores issued a transform ago)
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        const int64_t nt = tile + gridDim.x;
        Tile nx = cur;
        if (nt < t1) {
            nx = decode(nt);
            issue(nx, (i + 1) % 3);
        }
        if (i > 0) put(prev, tb[(i + 2) % 3]);
        fft_lds<T, N, TLC, +1, 0, 0, false>(tb[b], stw, tid);
        if (nt >= t1) break;
        prev = cur;
        cur = nx;
        tile = nt;
    }
    put(cur, tb[i % 3]);  // the last tile (the transform ended with a barrier)
}

// Amplitude scale of each leg: the rms of |spectrum| over (a sample of about
// 4k) cells of its shell.  Packing D + i I into one complex transform
// in the last pass gives both parts the rounding noise of the larger one;
// dividing the D part by this scale keeps both near unit amplitude.
template <typename T, int N>
__global__ void __launch_bounds__(NT) k_leg_scale(const LegDev *__restrict__ legs)
{
    __shared__ double red[2][NT];
    const int tid = threadIdx.x;
    const LegDev &L = legs[blockIdx.x];
    if (!L.scale) return;
    const cplx<T> *spec = reinterpret_cast<const cplx<T> *>(L.spec);
    double pw = 0.0, cnt = 0.0;
    if (spec && !L.phase_only) {
        // a sample of about 4096 cells of the band box (< 2^31 cells)
        const int B = L.Bx, BB = B * B;
        const int total = (L.Kz + 1) * BB;
        const int step = (total > 4096) ? total / 4096 : 1;
        for (int c = tid * step; c < total; c += NT * step) {
            const int kz = c / BB, rem = c - kz * BB;
            const int jx = rem / B, jy = rem - jx * B;
            const int mx = lattice(band_to_k(jx, L.Kx, B, N), N), my = lattice(band_to_k(jy, L.Kx, B, N), N);
            if (mask_eval(L.mask, mx, my, lattice(kz, N))) {
                const int ns = L.ns;
                const cplx<T> d = gload<T>(spec, ((int64_t)kz * ns + (mx < 0 ? mx + ns : mx)) * ns + (my < 0 ? my + ns : my));
                pw += (double)d.x * (double)d.x + (double)d.y * (double)d.y;
                cnt += 1.0;
            }
        }
    }
    red[0][tid] = pw;
    red[1][tid] = cnt;
    __syncthreads();
    for (int h = NT / 2; h > 0; h >>= 1) {
        if (tid < h) {
            red[0][tid] += red[0][tid + h];
            red[1][tid] += red[1][tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double sc = (red[1][0] > 0.0 && red[0][0] > 0.0) ? sqrt(red[0][0] / red[1][0]) : 1.0;
        *L.scale = (sc > 0.0 && sc < 1e300) ? sc : 1.0;
    }
}

// ================================================================ adjoint helpers
// H *= sum_j w_j * mask_j(m) over the whole half spectrum, in place.
template <typename T, int N>
__global__ void __launch_bounds__(NT) k_apply_masks(cplx<T> *__restrict__ H, const MaskParams *__restrict__ masks,
                                                    const double *__restrict__ w, int nm)
{
    const int64_t total = (int64_t)(N / 2 + 1) * N * N;
    for (int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x; c < total; c += (int64_t)gridDim.x * NT) {
        const int kz = (int)(c / ((int64_t)N * N));
        const int kx = (int)((c / N) % N), ky = (int)(c % N);
        const int mx = lattice(kx, N), my = lattice(ky, N), mz = lattice(kz, N);
        double f = 0.0;
        for (int j = 0; j < nm; ++j)
            if (mask_eval(masks[j], mx, my, mz)) f += w[j];
        const cplx<T> v = H[c];
        H[c] = mk<T>((T)((double)v.x * f), (T)((double)v.y * f));
    }
}

// a += b over a half spectrum
template <typename T>
__global__ void __launch_bounds__(NT) k_add_spec(cplx<T> *__restrict__ a, const cplx<T> *__restrict__ b, int64_t n)
{
    for (int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x; c < n; c += (int64_t)gridDim.x * NT)
        a[c] = cadd<T>(a[c], b[c]);
}

// out = alpha * a * b (b == nullptr: alpha * a * a)
__global__ void __launch_bounds__(NT) k_mul_fields(double *__restrict__ out, const double *__restrict__ a,
                                                   const double *__restrict__ b, double alpha, int64_t n)
{
    for (int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x; c < n; c += (int64_t)gridDim.x * NT) {
        const double x = a[c];
        out[c] = alpha * x * (b ? b[c] : x);
    }
}

// z -> z/|z| over a half spectrum, in place (utils_fftw.c:78-93); an exact
// zero stays zero (it is never selected by a shell of a closing polygon)
template <typename T>
__global__ void __launch_bounds__(NT) k_phase_normalize(cplx<T> *__restrict__ H, int64_t n)
{
    for (int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x; c < n; c += (int64_t)gridDim.x * NT) {
        const cplx<T> z = H[c];
        const double a = hypot((double)z.x, (double)z.y);
        H[c] = (a > 0.0) ? mk<T>((T)((double)z.x / a), (T)((double)z.y / a)) : mk<T>(0, 0);
    }
}

// Pull a spectrum-space gradient Y back through p = z/|z|: with
// dp = (dz - p Re(conj(p) dz)) / |z|, the gradient with respect to z is
// Y' = (Y - p Re(conj(p) Y)) / |z| (zero where z = 0).  H holds Y, spec z.
template <typename T>
__global__ void __launch_bounds__(NT) k_phase_pullback(cplx<T> *__restrict__ H, const cplx<T> *__restrict__ spec,
                                                       int64_t n)
{
    for (int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x; c < n; c += (int64_t)gridDim.x * NT) {
        const cplx<T> z = spec[c], y = H[c];
        const double a = hypot((double)z.x, (double)z.y);
        if (a > 0.0) {
            const double px = (double)z.x / a, py = (double)z.y / a;
            const double r = px * (double)y.x + py * (double)y.y;
            H[c] = mk<T>((T)(((double)y.x - px * r) / a), (T)(((double)y.y - py * r) / a));
        } else {
            H[c] = mk<T>(0, 0);
        }
    }
}

// C2R along kz of a half spectrum whose kx, ky axes are already inverse
// transformed: two adjacent y rows share one complex line.  accumulate: out +=.
template <typename T, int N>
__global__ void __launch_bounds__(NT) k_inv_z_real(const cplx<T> *__restrict__ H, double *__restrict__ out,
                                                   int accumulate, const cplx<T> *__restrict__ tw)
{
    constexpr int TL = tile_lines<T, N>();
    constexpr int TLC = TL / 2;
    constexpr int NH = N / 2 + 1;
    __shared__ cplx<T> lds[N * TLC];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * TL, x = blockIdx.y;
    for (int idx = tid; idx < NH * TLC; idx += NT) {
        const int c = idx % TLC, kz = idx / TLC;
        const int64_t o = ((int64_t)kz * N + x) * N + y0 + 2 * c;
        const cplx<T> X = H[o], Y = H[o + 1];
        if (kz == 0 || kz == N / 2) {
            lds[lidx<TLC, T>(kz, c)] = mk<T>(X.x, Y.x);
        } else {
            lds[lidx<TLC, T>(kz, c)] = mk<T>(X.x - Y.y, X.y + Y.x);
            lds[lidx<TLC, T>(N - kz, c)] = mk<T>(X.x + Y.y, Y.x - X.y);
        }
    }
    __syncthreads();
    fft_lds<T, N, TLC, +1>(lds, tw, tid);
    const int64_t base = ((int64_t)x * N + y0) * N;
    for (int idx = tid; idx < TL * N; idx += NT) {
        const int row = idx / N, z = idx % N;
        const cplx<T> v = lds[lidx<TLC, T>(z, row >> 1)];
        const double r = (double)((row & 1) ? v.y : v.x);
        out[base + idx] = accumulate ? out[base + idx] + r : r;
    }
}

// ================================================================ reductions
// out[i] = compensated sum over j < count[i] of partials[2*(i*stride + j) + which]
__global__ void __launch_bounds__(NT) k_reduce_partials(const double2 *__restrict__ partials, int64_t stride,
                                                        const int *__restrict__ counts, int64_t uniform_count,
                                                        double *__restrict__ outD, double *__restrict__ outI)
{
    __shared__ double red[NT / 64][4];
    const int tid = threadIdx.x;
    const int64_t i = blockIdx.x;
    const int64_t cnt = counts ? counts[i] : uniform_count;
    KSum a, b;
    a.init();
    b.init();
    for (int64_t j = tid; j < cnt; j += NT) {
        const double2 p = partials[2 * (i * stride + j) + 0];
        const double2 q = partials[2 * (i * stride + j) + 1];
        a.add(p.x); a.add(p.y);
        b.add(q.x); b.add(q.y);
    }
    double as = a.s, ac = a.c, bs = b.s, bc = b.c;
    wave_sum2(as, ac);
    wave_sum2(bs, bc);
    if ((tid & 63) == 0) {
        red[tid / 64][0] = as; red[tid / 64][1] = ac; red[tid / 64][2] = bs; red[tid / 64][3] = bc;
    }
    __syncthreads();
    if (tid == 0) {
        KSum A, B;
        A.init();
        B.init();
        for (int k = 0; k < NT / 64; ++k) {
            A.add(red[k][0]); A.add(red[k][1]);
            B.add(red[k][2]); B.add(red[k][3]);
        }
        if (outD) outD[i] = A.value();
        if (outI) outI[i] = B.value();
    }
}

__global__ void k_fill(double *__restrict__ p, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// DERIVED binning: extremes of the squared corner norms over a shell's cells
// (filter.c:106-113).  Partial min/max per block over the band |m_i| <= K.
__global__ void __launch_bounds__(NT) k_shell_corner_extremes(MaskParams mp, int N, int K, double bw,
                                                              double2 *__restrict__ partials)
{
    __shared__ double lo_s[NT], hi_s[NT];
    const int tid = threadIdx.x;
    const int B = 2 * K + 1;
    const int64_t total = (int64_t)B * B * B;
    double lo = 1.e30 * 1.e30, hi = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * NT + tid; c < total; c += (int64_t)gridDim.x * NT) {
        const int mx = (int)(c / ((int64_t)B * B)) - K;
        const int my = (int)((c / B) % B) - K;
        const int mz = (int)(c % B) - K;
        // a lattice coordinate +N/2 does not exist: index N/2 is -N/2
        if (2 * mx == N || 2 * my == N || 2 * mz == N) continue;
        if (mask_eval(mp, mx, my, mz)) {
            double l2, h2;
            corner_extremes(mx, my, mz, bw, l2, h2);
            lo = fmin(lo, l2);
            hi = fmax(hi, h2);
        }
    }
    lo_s[tid] = lo;
    hi_s[tid] = hi;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            lo_s[tid] = fmin(lo_s[tid], lo_s[tid + s]);
            hi_s[tid] = fmax(hi_s[tid], hi_s[tid + s]);
        }
        __syncthreads();
    }
    if (tid == 0) partials[blockIdx.x] = make_double2(lo_s[0], hi_s[0]);
}

// ================================================================ launchers
// Forward R2C with radix plan PLAN and tiles of TL lines (variant v of
// run_forward: PLAN = v % 2, tiles of the default width (v < 2) or half of
// it (v >= 2, where a half tile has at least 4 lines))
template <typename T, int N, int TL, int PLAN>
static int launch_forward_v(const void *field, int field_is_double, void *H, const void *tw, hipStream_t st)
{
    const cplx<T> *twp = reinterpret_cast<const cplx<T> *>(tw);
    cplx<T> *h = reinterpret_cast<cplx<T> *>(H);
    if (field_is_double)
        hipLaunchKernelGGL((k_fwd_z<T, double, N, TL, PLAN>), dim3(N / TL, N), dim3(NT), 0, st, (const double *)field,
                           h, twp);
    else
        hipLaunchKernelGGL((k_fwd_z<T, float, N, TL, PLAN>), dim3(N / TL, N), dim3(NT), 0, st, (const float *)field,
                           h, twp);
    const int64_t nlines = (int64_t)(N / 2 + 1) * N;  // (kz, kx) lines along ky
    hipLaunchKernelGGL((k_c2c_contig<T, N, -1, TL, PLAN>), dim3((unsigned)(nlines / TL)), dim3(NT), 0, st, h, twp);
    hipLaunchKernelGGL((k_c2c_strided<T, N, -1, TL, PLAN>), dim3(N / TL, N / 2 + 1), dim3(NT), 0, st, h, twp);
    return (int)hipGetLastError();
}

template <typename T, int N> constexpr bool fwd_half_tiles() { return N >= 128 && tile_lines<T, N>() / 2 >= 4; }

template <typename T, int N>
static int launch_forward_n(const void *field, int field_is_double, void *H, const void *tw, hipStream_t st,
                            int variant)
{
    constexpr int TL = tile_lines<T, N>();
    if constexpr (N >= 128) {
        switch (variant) {
        case 1: return launch_forward_v<T, N, TL, 1>(field, field_is_double, H, tw, st);
        case 2:
            if constexpr (fwd_half_tiles<T, N>()) return launch_forward_v<T, N, TL / 2, 0>(field, field_is_double, H, tw, st);
            break;
        case 3:
            if constexpr (fwd_half_tiles<T, N>()) return launch_forward_v<T, N, TL / 2, 1>(field, field_is_double, H, tw, st);
            break;
        default: break;
        }
    }
    return launch_forward_v<T, N, TL, 0>(field, field_is_double, H, tw, st);
}

// the copy-pipelined G2 of the D-only waves (PS_G2_DMA=0: k_inv_g2)
static bool g2_dma_enabled()
{
    static const bool on = [] {
        const char *e = getenv("PS_G2_DMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <typename T, int N>
static int launch_inverse_g_n(const LegDev *legs_dev, int nlegs, int64_t g1a, int64_t g1b, int64_t g2a, int64_t g2b,
                              int grid_cap, int g2_div, int dmode, const void *tw, hipStream_t st)
{
    const cplx<T> *twp = reinterpret_cast<const cplx<T> *>(tw);
    if (g1b > g1a) {
        const unsigned g = (unsigned)((g1b - g1a < grid_cap) ? g1b - g1a : grid_cap);
        if (dmode)
            hipLaunchKernelGGL((k_inv_g1d<T, N>), dim3(g), dim3(NT), 0, st, legs_dev, nlegs, g1a, g1b, twp);
        else
            hipLaunchKernelGGL((k_inv_g1<T, N>), dim3(g), dim3(NT), 0, st, legs_dev, nlegs, g1a, g1b, twp);
    }
    if (g2b > g2a) {
        const unsigned g = (unsigned)((g2b - g2a < grid_cap) ? g2b - g2a : grid_cap);
        if (dmode && g2_dma_enabled()) {
            const bool half = g2_tile_cols(N, 0, g2_div) < tile_lines<T, N>();
            if constexpr (g2_dma_ok<T, N, 2>()) {
                if (half) {
                    hipLaunchKernelGGL((k_inv_g2_dma<T, N, 2>), dim3(g), dim3(NT), 0, st, legs_dev, nlegs, g2a, g2b,
                                       twp);
                    return (int)hipGetLastError();
                }
            }
            if constexpr (g2_dma_ok<T, N, 1>()) {
                if (!half) {
                    hipLaunchKernelGGL((k_inv_g2_dma<T, N, 1>), dim3(g), dim3(NT), 0, st, legs_dev, nlegs, g2a, g2b,
                                       twp);
                    return (int)hipGetLastError();
                }
            }
        }
        if (g2_tile_cols(N, sizeof(T) == 8, g2_div) < tile_lines<T, N>())
            hipLaunchKernelGGL((k_inv_g2<T, N, 2>), dim3(g), dim3(NT), 0, st, legs_dev, nlegs, g2a, g2b, twp);
        else
            hipLaunchKernelGGL((k_inv_g2<T, N, 1>), dim3(g), dim3(NT), 0, st, legs_dev, nlegs, g2a, g2b, twp);
    }
    return (int)hipGetLastError();
}

template <typename T, int N>
static int launch_zc_plain_n(const void *gd, const void *gi, int Kz, double *accD, double *accI, int mode,
                             double2 *partials, const void *tw, hipStream_t st)
{
    constexpr int TLC = tile_lines<T, N>();
    hipLaunchKernelGGL((k_inv_zc_plain<T, N>), dim3(N / TLC, N), dim3(NT), 0, st,
                       reinterpret_cast<const cplx<T> *>(gd), reinterpret_cast<const cplx<T> *>(gi), Kz, accD, accI,
                       mode, partials, reinterpret_cast<const cplx<T> *>(tw));
    return (int)hipGetLastError();
}

template <typename T, int N, bool DM, bool ABS = false>
static int launch_zc_fused_m(const GroupDev *groups, int ngroups, const LegDev *legs, const int *suffix,
                             int64_t i0, int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kz_cap,
                             const void *tw, hipStream_t st)
{
    using C = ZcCfg<T, N, DM>;
    if (kz_cap < 1 || kz_cap > C::NH) kz_cap = C::NH;
    // two staging buffers (the next step's LDS stores need no extra barrier)
    // when they fit next to the static LDS, except in the compact form
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&k_inv_zc_fused<T, N, DM, ABS>));
    if (e != hipSuccess) return (int)e;
    int dev = 0, lds_max = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return (int)e;
    const size_t per_buf = sizeof(cplx<T>) * 2 * (size_t)kz_cap * C::SP;
    int nbuf = zc_compact<T, N>() ? 1 : 2;
    if (fa.sharedSizeBytes + 2 * per_buf > (size_t)lds_max) nbuf = 1;
    if (fa.sharedSizeBytes + nbuf * per_buf > (size_t)lds_max) return PS_ERR_UNSUPPORTED_N;
    const size_t dyn = nbuf * per_buf;
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_inv_zc_fused<T, N, DM, ABS>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (e != hipSuccess) return (int)e;
    if (i1 <= i0) return 0;
    const unsigned g = (unsigned)((i1 - i0 < grid_cap) ? i1 - i0 : grid_cap);
    hipLaunchKernelGGL((k_inv_zc_fused<T, N, DM, ABS>), dim3(g), dim3(C::BT), dyn, st, groups, ngroups, legs, suffix, i0,
                       i1, partials, max_chunks, reinterpret_cast<const cplx<T> *>(tw), kz_cap, nbuf);
    return (int)hipGetLastError();
}

// the raw-pair form of the D-only pass where it exists (PS_ZC_PAIRS=0: k_inv_zc_fused)
static bool zc_pairs_enabled()
{
    static const bool on = [] {
        const char *e = getenv("PS_ZC_PAIRS");
        return !(e && e[0] == '0');
    }();
    return on;
}

// the copy-pipelined form of the raw-pair pass where it exists (PS_ZC_DMA=0: k_zc_pairs)
static bool zc_dma_enabled()
{
    static const bool on = [] {
        const char *e = getenv("PS_ZC_DMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

// staging buffers of the copy-pipelined pass: two (one step ahead, four
// workgroups per CU) where three would leave LDS for only three workgroups
// (96, 192: measured Zc 0.404 -> 0.392 ms on the 512^3 sweep); three at 144,
// where four workgroups fit either way.  PS_ZC_NBUF=2 or 3 sets every grid.
template <int N> static int zc_dma_nbuf()
{
    static const int n = [] {
        const char *e = getenv("PS_ZC_NBUF");
        if (e && (e[0] == '2' || e[0] == '3')) return e[0] - '0';
        return (N == 144) ? 3 : 2;
    }();
    return n;
}

// dmode bit 0: D parts only; bit 1 (FP32 D-only): partial slot 1 of every
// output carries sum |prod D| (the numerator of its condition number)
template <typename T, int N, bool ABS>
static int launch_zc_dmode(const GroupDev *groups, int ngroups, const LegDev *legs, const int *suffix, int64_t i0,
                           int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kz_cap, const void *tw,
                           hipStream_t st)
{
    const unsigned g = (unsigned)((i1 - i0 < grid_cap) ? i1 - i0 : grid_cap);
    if constexpr (zc_dma_ok<T, N>()) {
        if (zc_pairs_enabled() && zc_dma_enabled()) {
            if (i1 <= i0) return 0;
            if (zc_dma_nbuf<N>() == 2)
                hipLaunchKernelGGL((k_zc_dma<T, N, ABS, 2>), dim3(g), dim3(ZcCfg<T, N, true>::BT), 0, st, groups,
                                   ngroups, legs, suffix, i0, i1, partials, max_chunks,
                                   reinterpret_cast<const cplx<T> *>(tw));
            else
                hipLaunchKernelGGL((k_zc_dma<T, N, ABS, 3>), dim3(g), dim3(ZcCfg<T, N, true>::BT), 0, st, groups,
                                   ngroups, legs, suffix, i0, i1, partials, max_chunks,
                                   reinterpret_cast<const cplx<T> *>(tw));
            return (int)hipGetLastError();
        }
    }
    if constexpr (zc_pair_ok<T, N>()) {
        if (zc_pairs_enabled()) {
            if (i1 <= i0) return 0;
            hipLaunchKernelGGL((k_zc_pairs<T, N, ABS>), dim3(g), dim3(ZcCfg<T, N, true>::BT), 0, st, groups, ngroups,
                               legs, suffix, i0, i1, partials, max_chunks, reinterpret_cast<const cplx<T> *>(tw));
            return (int)hipGetLastError();
        }
    }
    if constexpr (zc_dmode_ok<T, N>())
        return launch_zc_fused_m<T, N, true, ABS>(groups, ngroups, legs, suffix, i0, i1, partials, max_chunks,
                                                  grid_cap, kz_cap, tw, st);
    return PS_ERR_INVALID;
}

template <typename T, int N>
static int launch_zc_fused_n(const GroupDev *groups, int ngroups, const LegDev *legs, const int *suffix,
                             int64_t i0, int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kz_cap,
                             int dmode, const void *tw, hipStream_t st)
{
    if (dmode & 1) {
        // the magnitude sums exist in the FP32 kernels only (mixed-precision refinement)
        if (sizeof(T) == 4 && (dmode & 2))
            return launch_zc_dmode<T, N, sizeof(T) == 4>(groups, ngroups, legs, suffix, i0, i1, partials,
                                                           max_chunks, grid_cap, kz_cap, tw, st);
        return launch_zc_dmode<T, N, false>(groups, ngroups, legs, suffix, i0, i1, partials, max_chunks, grid_cap,
                                            kz_cap, tw, st);
    }
    return launch_zc_fused_m<T, N, false>(groups, ngroups, legs, suffix, i0, i1, partials, max_chunks, grid_cap,
                                          kz_cap, tw, st);
}

#define PS_FOR_EACH_N(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(144) X(192) X(216) X(256) X(384) X(512)

template <typename T>
static int dispatch_forward(int N, const void *field, int fd, void *H, const void *tw, hipStream_t st, int variant)
{
    switch (N) {
#define CASE(n) case n: return launch_forward_n<T, n>(field, fd, H, tw, st, variant);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

template <typename T>
static int dispatch_inverse_g(int N, const LegDev *legs, int nlegs, int64_t g1a, int64_t g1b, int64_t g2a, int64_t g2b,
                              int cap, int g2_div, int dmode, const void *tw, hipStream_t st)
{
    switch (N) {
#define CASE(n) case n: return launch_inverse_g_n<T, n>(legs, nlegs, g1a, g1b, g2a, g2b, cap, g2_div, dmode, tw, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

template <typename T>
static int dispatch_zc_plain(int N, const void *gd, const void *gi, int Kz, double *accD, double *accI, int mode,
                             double2 *partials, const void *tw, hipStream_t st)
{
    switch (N) {
#define CASE(n) case n: return launch_zc_plain_n<T, n>(gd, gi, Kz, accD, accI, mode, partials, tw, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

template <typename T>
static int dispatch_zc_fused(int N, const GroupDev *groups, int ngroups, const LegDev *legs, const int *suffix,
                             int64_t i0, int64_t i1, double2 *partials, int max_chunks, int cap, int kz_cap,
                             int dmode, const void *tw, hipStream_t st)
{
    switch (N) {
#define CASE(n) case n: return launch_zc_fused_n<T, n>(groups, ngroups, legs, suffix, i0, i1, partials, max_chunks, cap, kz_cap, dmode, tw, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

// ---------------------------------------------------------------- exported (internal API)
int supported_grid(int N)
{
    switch (N) {
#define CASE(n) case n:
        PS_FOR_EACH_N(CASE)
#undef CASE
        return 1;
    }
    return 0;
}

int tile_rows(int N, int fp64) { return tile_lines_for(N, fp64 != 0); }

int g1_dmode_lines(int N, int fp64)
{
    switch (N) {
#define CASE(n) case n: return fp64 ? g1d_lines_t<double, n>() : g1d_lines_t<float, n>();
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return 0;
}

// G2 tile width: half the lines when asked (g2_div = 2), else the full width,
// or half of it where only the half-width tile's band values fit the
// registers that prefetch the next tile (at most 8 per thread)
int g2_tile_cols(int N, int fp64, int g2_div)
{
    const int full = tile_lines_for(N, fp64 != 0);
    if (g2_div == 2) return full / 2;
    const int per_thread = full * N / NT;
    return (per_thread > 8 && per_thread / 2 <= 8) ? full / 2 : full;
}

int fused_tile_rows(int N, int fp64)
{
    switch (N) {
#define CASE(n) case n: return fp64 ? ZcCfg<double, n>::ROWS : ZcCfg<float, n>::ROWS;  // both modes
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return 0;
}
int g_tile_lines(int N, int fp64) { return tile_rows(N, fp64) / 2; }

int run_forward(int fp64, int N, const void *field, int field_is_double, void *H, const void *tw, hipStream_t st,
                int variant)
{
    return fp64 ? dispatch_forward<double>(N, field, field_is_double, H, tw, st, variant)
                : dispatch_forward<float>(N, field, field_is_double, H, tw, st, variant);
}

// variants of run_forward an N^3 context has: plan 0/1 with default tiles,
// and with half tiles where they exist
int forward_variants(int N, int fp64)
{
    switch (N) {
#define CASE(n) case n: return (n < 128) ? 1 : ((fp64 ? fwd_half_tiles<double, n>() : fwd_half_tiles<float, n>()) ? 4 : 2);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return 1;
}

int run_inverse_g(int fp64, int N, const LegDev *legs, int nlegs, int64_t g1a, int64_t g1b, int64_t g2a, int64_t g2b,
                  int grid_cap, int g2_div, int dmode, const void *tw, hipStream_t st)
{
    return fp64 ? dispatch_inverse_g<double>(N, legs, nlegs, g1a, g1b, g2a, g2b, grid_cap, g2_div, dmode, tw, st)
                : dispatch_inverse_g<float>(N, legs, nlegs, g1a, g1b, g2a, g2b, grid_cap, g2_div, dmode, tw, st);
}

int run_zc_plain(int fp64, int N, const void *gd, const void *gi, int Kz, double *accD, double *accI, int mode,
                 double2 *partials, const void *tw, hipStream_t st)
{
    return fp64 ? dispatch_zc_plain<double>(N, gd, gi, Kz, accD, accI, mode, partials, tw, st)
                : dispatch_zc_plain<float>(N, gd, gi, Kz, accD, accI, mode, partials, tw, st);
}

int run_zc_fused(int fp64, int N, const GroupDev *groups, int ngroups, const LegDev *legs, const int *suffix,
                 int64_t i0, int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kz_cap, int dmode,
                 const void *tw, hipStream_t st)
{
    return fp64 ? dispatch_zc_fused<double>(N, groups, ngroups, legs, suffix, i0, i1, partials, max_chunks, grid_cap,
                                            kz_cap, dmode, tw, st)
                : dispatch_zc_fused<float>(N, groups, ngroups, legs, suffix, i0, i1, partials, max_chunks, grid_cap,
                                           kz_cap, dmode, tw, st);
}

int zc_dmode_supported(int N, int fp64)
{
    switch (N) {
#define CASE(n) case n: return fp64 ? zc_dmode_ok<double, n>() : zc_dmode_ok<float, n>();
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return 0;
}

int run_reduce_partials(const double2 *partials, int64_t stride, const int *counts, int64_t uniform_count,
                        int64_t nout, double *outD, double *outI, hipStream_t st)
{
    if (nout <= 0) return 0;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)nout), dim3(NT), 0, st, partials, stride, counts,
                       uniform_count, outD, outI);
    return (int)hipGetLastError();
}

int run_fill(double *p, int64_t n, double v, hipStream_t st)
{
    int64_t blocks = (n + NT - 1) / NT;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(NT), 0, st, p, n, v);
    return (int)hipGetLastError();
}

template <typename T>
static int dispatch_leg_scale(int N, const LegDev *legs, int nlegs, hipStream_t st)
{
    switch (N) {
#define CASE(n) case n: hipLaunchKernelGGL((k_leg_scale<T, n>), dim3(nlegs), dim3(NT), 0, st, legs); return (int)hipGetLastError();
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

int run_leg_scale(int fp64, int N, const LegDev *legs, int nlegs, hipStream_t st)
{
    if (nlegs <= 0) return 0;
    return fp64 ? dispatch_leg_scale<double>(N, legs, nlegs, st) : dispatch_leg_scale<float>(N, legs, nlegs, st);
}

static unsigned grid_for(int64_t n)
{
    int64_t b = (n + NT - 1) / NT;
    if (b > 16384) b = 16384;
    return (unsigned)(b < 1 ? 1 : b);
}

template <typename T, int N>
static int launch_inverse_real_n(void *H, double *out, int accumulate, const void *tw, hipStream_t st)
{
    constexpr int TLC = tile_lines<T, N>();
    cplx<T> *h = reinterpret_cast<cplx<T> *>(H);
    const cplx<T> *twp = reinterpret_cast<const cplx<T> *>(tw);
    const int64_t nlines = (int64_t)(N / 2 + 1) * N;
    hipLaunchKernelGGL((k_c2c_contig<T, N, +1>), dim3((unsigned)(nlines / TLC)), dim3(NT), 0, st, h, twp);
    hipLaunchKernelGGL((k_c2c_strided<T, N, +1>), dim3(N / TLC, N / 2 + 1), dim3(NT), 0, st, h, twp);
    hipLaunchKernelGGL((k_inv_z_real<T, N>), dim3(N / TLC, N), dim3(NT), 0, st, h, out, accumulate, twp);
    return (int)hipGetLastError();
}

template <typename T, int N>
static int launch_apply_masks_n(void *H, const MaskParams *masks, const double *w, int nm, hipStream_t st)
{
    const int64_t total = (int64_t)(N / 2 + 1) * N * N;
    hipLaunchKernelGGL((k_apply_masks<T, N>), dim3(grid_for(total)), dim3(NT), 0, st,
                       reinterpret_cast<cplx<T> *>(H), masks, w, nm);
    return (int)hipGetLastError();
}

int run_inverse_real(int fp64, int N, void *H, double *out, int accumulate, const void *tw, hipStream_t st)
{
    switch (N) {
#define CASE(n)                                                                                     \
    case n:                                                                                         \
        return fp64 ? launch_inverse_real_n<double, n>(H, out, accumulate, tw, st)                  \
                    : launch_inverse_real_n<float, n>(H, out, accumulate, tw, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

int run_apply_masks(int fp64, int N, void *H, const MaskParams *masks, const double *w, int nm, hipStream_t st)
{
    switch (N) {
#define CASE(n)                                                                                     \
    case n:                                                                                         \
        return fp64 ? launch_apply_masks_n<double, n>(H, masks, w, nm, st)                          \
                    : launch_apply_masks_n<float, n>(H, masks, w, nm, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

int run_add_spec(int fp64, void *a, const void *b, int64_t n, hipStream_t st)
{
    if (fp64)
        hipLaunchKernelGGL(k_add_spec<double>, dim3(grid_for(n)), dim3(NT), 0, st, (cplx<double> *)a,
                           (const cplx<double> *)b, n);
    else
        hipLaunchKernelGGL(k_add_spec<float>, dim3(grid_for(n)), dim3(NT), 0, st, (cplx<float> *)a,
                           (const cplx<float> *)b, n);
    return (int)hipGetLastError();
}

int run_phase_normalize(int fp64, void *H, int64_t n, hipStream_t st)
{
    if (fp64) hipLaunchKernelGGL(k_phase_normalize<double>, dim3(grid_for(n)), dim3(NT), 0, st, (cplx<double> *)H, n);
    else hipLaunchKernelGGL(k_phase_normalize<float>, dim3(grid_for(n)), dim3(NT), 0, st, (cplx<float> *)H, n);
    return (int)hipGetLastError();
}

int run_phase_pullback(int fp64, void *H, const void *spec, int64_t n, hipStream_t st)
{
    if (fp64)
        hipLaunchKernelGGL(k_phase_pullback<double>, dim3(grid_for(n)), dim3(NT), 0, st, (cplx<double> *)H,
                           (const cplx<double> *)spec, n);
    else
        hipLaunchKernelGGL(k_phase_pullback<float>, dim3(grid_for(n)), dim3(NT), 0, st, (cplx<float> *)H,
                           (const cplx<float> *)spec, n);
    return (int)hipGetLastError();
}

int run_mul_fields(double *out, const double *a, const double *b, double alpha, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_mul_fields, dim3(grid_for(n)), dim3(NT), 0, st, out, a, b, alpha, n);
    return (int)hipGetLastError();
}

int run_corner_extremes(const MaskParams &mp, int N, int K, double bw, double2 *partials, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_shell_corner_extremes, dim3(nblocks), dim3(NT), 0, st, mp, N, K, bw, partials);
    return (int)hipGetLastError();
}

}  // namespace ps

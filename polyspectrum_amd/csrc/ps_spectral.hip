// ps_spectral.hip -- polygon groups evaluated in Fourier space.
//
// For a group of polygons sharing the prefix legs A (, B) the estimator needs,
// for every suffix leg t, S_t = sum_x P(x) D_t(x) with P = D_A D_B (and the
// same for the unit-amplitude I fields).  With unnormalised transforms,
// D_t(x) = sum_k m_t(k) H(k) e^{ikx}, so
//     sum_x P(x) D_t(x) = sum_k conj(F[P](k)) m_t(k) H(k)      (F: forward DFT)
// and one forward transform of the prefix product replaces the inverse
// transform of every suffix leg: each suffix costs one sum over its shell.
// The inverse transforms of the prefix legs are the G1/G2 passes of
// ps_kernels.hip; the passes here are
//   Zf : per (x, y) row, C2R along kz of each prefix leg (D + i I packed),
//        the product in registers, the forward C2C along z of Q = P_D + i P_I
//        and its split into the z half spectra of P_D and P_I on kz <= Kqz
//   Fx : forward C2C along x of those planes, kept on the kx band
//   dot: forward C2C along y of the band lines and, per cell of the band, the
//        shell sums of every suffix leg of the group
// Power spectra (two legs) need no transform: Q = N^3 m_A H is read from the
// spectrum in the dot pass.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <algorithm>

#include "ps_device.hpp"
#include "ps_fft_core.hpp"
#include "ps_internal.h"
#include "ps_mask.hpp"
#include "ps_wave_fft.hpp"

namespace ps {

// ================================================================ Zf
template <typename T, int N> struct ZfCfg {
    static constexpr int LW = wave_lines<T, N>();
    static constexpr int NWV = 4;
    static constexpr int BT = 64 * NWV;
    static constexpr int ROWS = NWV * LW;  // y rows of a tile (at one x)
    static constexpr int SP = ROWS + 1;    // staging pitch
    static constexpr int LP = wave_pitch<N>();   // the tile's Q lines
    static constexpr int WP = wave_fft_pitch<T, N>();  // wave_fft's lines
};

// Tile = (group, x, ROWS adjacent y rows).  Dynamic LDS holds the staged
// planes of one prefix leg [2][kz_cap * SP] and, in the same memory once the
// prefix transforms are done, the tile's Q lines [ROWS][LP].
template <typename T, int N>
__global__ void __launch_bounds__((ZfCfg<T, N>::BT)) k_spec_zf(const SpecGroupDev *__restrict__ groups, int ngroups,
                                                               const LegDev *__restrict__ legs, int64_t t0, int64_t t1,
                                                               const cplx<T> *__restrict__ twg, int kz_cap)
{
    using C = ZfCfg<T, N>;
    constexpr int LW = C::LW, NWV = C::NWV, ROWS = C::ROWS, SP = C::SP, BT = C::BT, LP = C::LP, WP = C::WP;
    using WO = WaveOut<T, N, LW>;
    extern __shared__ unsigned char dyn_lds[];
    __shared__ cplx<T> wbuf[NWV * LW * WP];
    cplx<T> *dyn = reinterpret_cast<cplx<T> *>(dyn_lds);
    cplx<T> *sD = dyn, *sI = dyn + kz_cap * SP;
    const int tid = threadIdx.x, w = tid / 64, lane = tid & 63;
    WaveTw<T, N, LW> twr;
    twr.template load<+1>(twg, lane);
    const RegTw<T, N, LW> tw{twr};
    cplx<T> *wb = wbuf + w * LW * WP;
    cplx<T> *qb = dyn + w * LW * LP;  // this wavefront's Q lines
    constexpr int tiles_per_x = N / ROWS;
    for (int64_t tile = t0 + blockIdx.x; tile < t1; tile += gridDim.x) {
        const int gi = find_owner(groups, ngroups, tile, &SpecGroupDev::zf_begin);
        const SpecGroupDev &G = groups[gi];
        const int64_t local = tile - G.zf_begin;
        const int x = (int)(local / tiles_per_x), y0 = (int)(local % tiles_per_x) * ROWS;
        T pd[WO::IT][WO::R], pi[WO::IT][WO::R];
        for (int q = 0; q < G.npre; ++q) {
            const LegDev &L = legs[q == 0 ? G.preA : G.preB];
            const int Kz = L.Kz;
            const cplx<T> *gd = reinterpret_cast<const cplx<T> *>(L.g2d);
            const cplx<T> *gI = reinterpret_cast<const cplx<T> *>(L.g2i);
            __syncthreads();  // the staging memory is free (previous leg or tile)
            // the z line of row y is Z[k] = X[k] + i Y[k] for k <= N/2 and
            // Z[N-k] = conj X[k] + i conj Y[k]; only real parts at k = 0, N/2
            for (int idx = tid; idx < (Kz + 1) * ROWS; idx += BT) {
                const int row = idx % ROWS, kz = idx / ROWS;
                const int64_t o = ((int64_t)kz * N + x) * N + y0 + row;
                const cplx<T> X = gd ? gload<T>(gd, o) : mk<T>(0, 0), Y = gload<T>(gI, o);
                if (kz == 0 || kz == N / 2) {
                    sD[kz * SP + row] = mk<T>(X.x, Y.x);
                } else {
                    sD[kz * SP + row] = mk<T>(X.x - Y.y, X.y + Y.x);
                    sI[kz * SP + row] = mk<T>(X.x + Y.y, Y.x - X.y);
                }
            }
            __syncthreads();
            auto src = [&](int l, int e) -> cplx<T> {
                const int row = w * LW + l;
                const bool lo = e <= N / 2;
                const int k = lo ? e : N - e;
                if (k > Kz) return mk<T>(0, 0);
                return (lo ? sD : sI)[k * SP + row];
            };
            const bool edge = Kz < N / stage_radix(N, 0);
            auto run = [&](auto &sink) {
                if (edge) wave_fft<T, N, LW, +1, 0, true, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
                else wave_fft<T, N, LW, +1, 0, false, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
            };
            if (q == 0 && !G.pre_sq) {
                auto sink = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] = v.x;
                    pi[i][r] = v.y;
                };
                run(sink);
            } else if (q == 0) {
                auto sink = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] = v.x * v.x;
                    pi[i][r] = v.y * v.y;
                };
                run(sink);
            } else {
                auto sink = [&](int i, int r, int, int, cplx<T> v) {
                    pd[i][r] *= v.x;
                    pi[i][r] *= v.y;
                };
                run(sink);
            }
        }
        __syncthreads();  // every wavefront has read the staged planes (qb aliases them)
        // Lines to transform, conjugated and in natural z order (the forward
        // transform is F[q] = conj(IFFT(conj q)) with the inverse twiddles).
        // PAIRED (an even number of rows per wavefront): rows a, b = a + 1
        // share lines as q_D = P_D(a) + i P_D(b) and q_I = P_I(a) + i P_I(b),
        // so that only like parts share a transform (P_I = I_A I_B peaks at
        // x = 0 and would dominate the rounding of P_D).  Otherwise one line
        // per row: q = P_D + i P_I.
        constexpr bool PAIRED = (LW % 2 == 0);
#pragma unroll
        for (int i = 0; i < WO::IT; ++i) {
            const int b = lane + 64 * i;
            if (WO::TOT % 64 == 0 || b < WO::TOT) {
                const int j = b % WO::NB, l = b / WO::NB;
#pragma unroll
                for (int r = 0; r < WO::R; ++r) {
                    const int e = wpad(j + r * WO::NB);
                    if constexpr (PAIRED) {
                        T *qd_ = reinterpret_cast<T *>(&qb[(l & ~1) * LP + e]);
                        T *qi_ = reinterpret_cast<T *>(&qb[(l | 1) * LP + e]);
                        qd_[l & 1] = (l & 1) ? -pd[i][r] : pd[i][r];
                        qi_[l & 1] = (l & 1) ? -pi[i][r] : pi[i][r];
                    } else {
                        qb[l * LP + e] = mk<T>(pd[i][r], -pi[i][r]);
                    }
                }
            }
        }
        wave_sync();
        cplx<T> qo[WO::IT][WO::R];
        auto src2 = [&](int l, int e) -> cplx<T> { return qb[l * LP + wpad(e)]; };
        auto sink2 = [&](int i, int r, int, int, cplx<T> v) { qo[i][r] = v; };
        wave_fft<T, N, LW, +1, 0, false, RegTw<T, N, LW>>(wb, tw, lane, src2, sink2);
        const int Kqz = G.Kqz;
        wave_sync();
#pragma unroll
        for (int i = 0; i < WO::IT; ++i) {
            const int b = lane + 64 * i;
            if (WO::TOT % 64 == 0 || b < WO::TOT) {
                const int j = b % WO::NB, l = b / WO::NB;
#pragma unroll
                for (int r = 0; r < WO::R; ++r) {
                    const int z = j + r * WO::NB;
                    if (z <= Kqz || z >= N - Kqz) qb[l * LP + wpad(z)] = qo[i][r];
                }
            }
        }
        __syncthreads();
        // a line holds a = conj F[q]; with A = F[q](k), B = conj F[q](N-k), the
        // real-part field of q has F_z = (A + B) / 2, the imaginary-part field
        // F_z = (A - B) / (2i)
        cplx<T> *od = reinterpret_cast<cplx<T> *>(G.qd);
        cplx<T> *oi = reinterpret_cast<cplx<T> *>(G.qi);
        const T h = (T)0.5;
        auto re_part = [&](cplx<T> a, cplx<T> c) { return mk<T>((a.x + c.x) * h, (c.y - a.y) * h); };
        auto im_part = [&](cplx<T> a, cplx<T> c) { return mk<T>(-(a.y + c.y) * h, (c.x - a.x) * h); };
        for (int idx = tid; idx < (Kqz + 1) * ROWS; idx += BT) {
            const int row = idx % ROWS, kz = idx / ROWS;  // row = wavefront * LW + line
            const int kn = wpad(kz), km = wpad((N - kz) % N);
            const int64_t o = ((int64_t)kz * N + x) * N + y0 + row;
            if constexpr (PAIRED) {
                const cplx<T> *ld = dyn + (row & ~1) * LP, *li = dyn + (row | 1) * LP;
                const cplx<T> ad = ld[kn], cd = ld[km], ai = li[kn], ci = li[km];
                gstore<T>(od, o, (row & 1) ? im_part(ad, cd) : re_part(ad, cd));
                gstore<T>(oi, o, (row & 1) ? im_part(ai, ci) : re_part(ai, ci));
            } else {
                const cplx<T> *line = dyn + row * LP;
                const cplx<T> a = line[kn], c = line[km];
                gstore<T>(od, o, re_part(a, c));
                gstore<T>(oi, o, im_part(a, c));
            }
        }
    }
}

// ================================================================ Fx
// Tile = (group, part, kz, y tile of TLC adjacent columns): forward C2C along
// x, the kx band stored as [kz][jx][y].
template <typename T, int N, int WD>
__global__ void __launch_bounds__(NT) k_spec_fx(const SpecGroupDev *__restrict__ groups, int ngroups, int64_t t0,
                                                int64_t t1, const cplx<T> *__restrict__ tw)
{
    constexpr int TLC = tile_lines<T, N>() / WD;
    __shared__ cplx<T> lds[N * TLC];
    __shared__ cplx<T> stw[N];
    const int tid = threadIdx.x;
    constexpr int nyt = N / TLC;
    stage_twiddles<T, N>(stw, tw, tid);
    for (int64_t tile = t0 + blockIdx.x; tile < t1; tile += gridDim.x) {
        const int gi = find_owner(groups, ngroups, tile, &SpecGroupDev::fx_begin);
        const SpecGroupDev &G = groups[gi];
        const int64_t local = tile - G.fx_begin;
        const int64_t per_part = (int64_t)(G.Kqz + 1) * nyt;
        const int part = (int)(local / per_part);
        const int64_t rest = local - part * per_part;
        const int kz = (int)(rest / nyt);
        const int y0 = (int)(rest % nyt) * TLC;
        const cplx<T> *in = reinterpret_cast<const cplx<T> *>(part ? G.qi : G.qd);
        cplx<T> *out = reinterpret_cast<cplx<T> *>(part ? G.bi : G.bd);
        tile_loop<TLC * N>(tid, [&](int idx) {
            const int l = idx % TLC, xx = idx / TLC;
            lds[lidx<TLC, T>(xx, l)] = gload<T>(in, ((int64_t)kz * N + xx) * N + y0 + l);
        });
        __syncthreads();
        fft_lds<T, N, TLC, -1>(lds, stw, tid);
        const int Bq = G.Bq;
        for (int idx = tid; idx < TLC * Bq; idx += NT) {
            const int l = idx % TLC, jx = idx / TLC;
            const int kx = band_to_k(jx, G.Kq, Bq, N);
            gstore<T>(out, ((int64_t)kz * Bq + jx) * N + y0 + l, lds[lidx<TLC, T>(kx, l)]);
        }
        __syncthreads();
    }
}

// ================================================================ dot
template <typename T, int N> struct DotCfg {
    static constexpr int LW = wave_lines<T, N>();
    static constexpr int NWV = 4;
    static constexpr int BT = 64 * NWV;
    static constexpr int LINES = NWV * LW;  // lines of one workgroup step
};

// Work item = (group, chunk of dot_lines lines).  Line 2p + part of a group is
// the y line of band cell p = kz * Bq + jx, part 0 = D, 1 = I.  Each
// wavefront transforms LW lines (fft groups) or reads Q from the spectrum
// (power spectra); every thread then holds cells (line, ky) with their
// contribution w Re(conj Q V) (D) or w Re Q (I).
//   radial groups: along a line m^2 = mx^2 + mz^2 + my^2 grows with |my|, so
//     a shell [lo, hi] of m^2 is one range a <= |my| <= b of the line.  The
//     contributions go to LDS; lane (t, part) adds, line by line, the range
//     of suffix t and keeps a compensated sum in registers for the item.
//   other masks (cube-corner tests): per suffix, every cell tested, the sums
//     tree-summed over the wavefront and accumulated per wavefront.
// Partial sums per (output, chunk).
template <typename T, int N, bool RADIAL>
__global__ void __launch_bounds__((DotCfg<T, N>::BT)) k_spec_dot(const SpecGroupDev *__restrict__ groups, int ngroups,
                                                                 const MaskParams *__restrict__ masks,
                                                                 const int2 *__restrict__ ivl, int64_t i0, int64_t i1,
                                                                 double2 *__restrict__ partials, int max_chunks,
                                                                 const cplx<T> *__restrict__ twg)
{
    using C = DotCfg<T, N>;
    constexpr int LW = C::LW, NWV = C::NWV, BT = C::BT, LINES = C::LINES;
    static_assert(MAX_SUFFIX <= 32, "lanes (t, part) of a wavefront hold the radial sums");
    constexpr int LP = wave_fft_pitch<T, N>();
    using WO = WaveOut<T, N, LW>;
    constexpr int IT = WO::IT, R = WO::R, NB = WO::NB;
    __shared__ cplx<T> wbuf[NWV * LW * LP];
    __shared__ double acc[NWV][MAX_SUFFIX][4];
    __shared__ MaskParams smask[MAX_SUFFIX];
    __shared__ int2 sivl[MAX_SUFFIX];
    const int tid = threadIdx.x, w = tid / 64, lane = tid & 63;
    WaveTw<T, N, LW> twr;
    twr.template load<+1>(twg, lane);
    const RegTw<T, N, LW> tw{twr};
    cplx<T> *wb = wbuf + w * LW * LP;
    for (int64_t item = i0 + blockIdx.x; item < i1; item += gridDim.x) {
        const int gi = find_owner(groups, ngroups, item, &SpecGroupDev::dot_begin);
        const SpecGroupDev &G = groups[gi];
        if ((G.radial != 0) != RADIAL) continue;  // the other instantiation's item
        const int chunk = (int)(item - G.dot_begin);
        const int Bq = G.Bq, Kq = G.Kq, ns = G.ns;
        const double n3 = (double)ns * (double)ns * (double)ns;
        const int64_t nlines = 2 * (int64_t)(G.Kqz + 1) * Bq;
        const int64_t l0 = (int64_t)chunk * G.dot_lines;
        const int64_t l1 = (l0 + G.dot_lines < nlines) ? l0 + G.dot_lines : nlines;
        for (int i = tid; i < NWV * MAX_SUFFIX * 4; i += BT) (&acc[0][0][0])[i] = 0.0;
        if (tid < G.nsuf) {
            smask[tid] = masks[G.suf_begin + tid];
            sivl[tid] = ivl[2 * (G.suf_begin + tid)];  // (lo, hi); the count follows
        }
        __syncthreads();
        const cplx<T> *spec = reinterpret_cast<const cplx<T> *>(G.spec);
        KSum ks;  // radial: lane (t, part)'s sum over the item
        ks.init();
        for (int64_t base = l0; base < l1; base += LINES) {
            const int64_t wl0 = base + w * LW;  // this wavefront's first line
            cplx<T> qo[IT][R];
            if (G.fft) {
                // F_y = conj(IFFT_y(conj .)) with the inverse twiddles
                auto src = [&](int l, int e) -> cplx<T> {
                    const int64_t line = wl0 + l;
                    if (line >= l1) return mk<T>(0, 0);
                    const cplx<T> *p = reinterpret_cast<const cplx<T> *>((line & 1) ? G.bi : G.bd) + (line >> 1) * N;
                    const cplx<T> v = gload<T>(p, e);
                    return mk<T>(v.x, -v.y);
                };
                auto sink = [&](int i, int r, int, int, cplx<T> v) { qo[i][r] = mk<T>(v.x, -v.y); };
                wave_fft<T, N, LW, +1, 0, false, RegTw<T, N, LW>>(wb, tw, lane, src, sink);
            }
            // contribution of cell (line, ky): w Re(conj Q V) (D) or w Re Q (I)
            auto contribution = [&](int i, int r, int l, int ky, int &m2, int &mx_, int &my_, int &mz_) -> double {
                const int64_t line = wl0 + l;
                m2 = -1;
                if (line >= l1) return 0.0;
                const int part = (int)(line & 1);
                const int64_t pair = line >> 1;
                const int kz = (int)(pair / Bq), jx = (int)(pair - (int64_t)kz * Bq);
                if (k_to_band(ky, Kq, Bq, N) < 0) return 0.0;
                const int mx = lattice(band_to_k(jx, Kq, Bq, N), N), my = lattice(ky, N), mz = lattice(kz, N);
                m2 = mx * mx + my * my + mz * mz;
                mx_ = mx;
                my_ = my;
                mz_ = mz;
                const double wz = (kz == 0 || 2 * kz == N) ? 1.0 : 2.0;
                cplx<T> V = mk<T>(0, 0);
                if (part == 0) {
                    V = gload<T>(spec, ((int64_t)kz * ns + (mx < 0 ? mx + ns : mx)) * ns + (my < 0 ? my + ns : my));
                    if (G.phase_only) {
                        // z / |z| (utils_fftw.c:78-93) on the cells a shell selects
                        const T a = hypot(V.x, V.y);
                        V = mk<T>(V.x / a, V.y / a);
                    }
                }
                if (G.fft) {
                    const cplx<T> Q = qo[i][r];
                    return part ? wz * (double)Q.x : wz * ((double)Q.x * (double)V.x + (double)Q.y * (double)V.y);
                }
                const double a = (!G.has_A || mask_eval(G.mA, mx, my, mz)) ? n3 : 0.0;
                return part ? wz * a : wz * a * ((double)V.x * (double)V.x + (double)V.y * (double)V.y);
            };
            if constexpr (RADIAL) {
                // along a line m^2 = mx^2 + mz^2 + my^2 grows with |my|: shell t
                // is the range a_t <= |my| <= b_t.  Contributions go to this
                // wavefront's LDS region (free after the transform) as [LW][N]
                // doubles; lane (t, part) sums its ranges line by line.
                double *cb = reinterpret_cast<double *>(wb);
                wave_sync();
#pragma unroll
                for (int i = 0; i < IT; ++i) {
                    const int b = lane + 64 * i;
                    if (WO::TOT % 64 == 0 || b < WO::TOT) {
                        const int j = b % NB, l = b / NB;
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            int m2, mx, my, mz;
                            cb[l * N + j + r * NB] = contribution(i, r, l, j + r * NB, m2, mx, my, mz);
                        }
                    }
                }
                wave_sync();
                const int t = lane & 31, lh = lane >> 5;
                if (t < G.nsuf) {
                    const int2 iv = sivl[t];
                    for (int l = 0; l < LW; ++l) {
                        const int64_t line = wl0 + l;
                        if (line >= l1 || (int)(line & 1) != lh) continue;
                        const int64_t pair = line >> 1;
                        const int kz = (int)(pair / Bq), jx = (int)(pair - (int64_t)kz * Bq);
                        const int mx = lattice(band_to_k(jx, Kq, Bq, N), N), mz = lattice(kz, N);
                        const int bse = mx * mx + mz * mz;
                        if (iv.y < bse) continue;
                        // largest ub with bse + ub^2 <= hi, smallest ua with bse + ua^2 >= lo
                        int ub = (int)sqrtf((float)(iv.y - bse));
                        while (ub * ub > iv.y - bse) --ub;
                        while ((ub + 1) * (ub + 1) <= iv.y - bse) ++ub;
                        int ua = 0;
                        if (iv.x > bse) {
                            ua = (int)sqrtf((float)(iv.x - bse));
                            while (ua * ua < iv.x - bse) ++ua;
                            while (ua > 0 && (ua - 1) * (ua - 1) >= iv.x - bse) --ua;
                        }
                        if (ub > Kq) ub = Kq;
                        const double *row = cb + l * N;
                        double s = 0.0;
                        for (int u = ua; u <= ub; ++u) {
                            s += row[u];
                            if (u > 0 && 2 * u != N) s += row[N - u];  // my = -u
                        }
                        ks.add(s);
                    }
                }
            } else {
                // cube-corner masks: every cell tested for every suffix
                double cv[IT][R];
                int crd[IT][R], prt[IT];
#pragma unroll
                for (int i = 0; i < IT; ++i) {
                    const int b = lane + 64 * i;
                    const int j = b % NB, l = b / NB;
                    prt[i] = (int)((wl0 + l) & 1);
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        int m2 = -1, mx = 0, my = 0, mz = 0;
                        cv[i][r] = (WO::TOT % 64 == 0 || b < WO::TOT)
                                       ? contribution(i, r, l, j + r * NB, m2, mx, my, mz) : 0.0;
                        crd[i][r] = (m2 < 0) ? -1 : (((mx + 512) << 20) | ((my + 512) << 10) | (mz + 512));
                    }
                }
                for (int o = 0; o < G.nsuf; ++o) {
                    const MaskParams mp = smask[o];
                    double sd = 0.0, si = 0.0;
#pragma unroll
                    for (int i = 0; i < IT; ++i) {
                        double s = 0.0;
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int cc = crd[i][r];
                            const bool in = cc >= 0 && mask_eval(mp, ((cc >> 20) & 1023) - 512,
                                                                 ((cc >> 10) & 1023) - 512, (cc & 1023) - 512);
                            s += in ? cv[i][r] : 0.0;
                        }
                        if (prt[i]) si += s;
                        else sd += s;
                    }
                    sd = wave_sum(sd);
                    si = wave_sum(si);
                    if (lane == 0) {
                        KSum a;
                        a.s = acc[w][o][0];
                        a.c = acc[w][o][1];
                        a.add(sd);
                        KSum c;
                        c.s = acc[w][o][2];
                        c.c = acc[w][o][3];
                        c.add(si);
                        acc[w][o][0] = a.s;
                        acc[w][o][1] = a.c;
                        acc[w][o][2] = c.s;
                        acc[w][o][3] = c.c;
                    }
                }
            }
        }
        if constexpr (RADIAL) {
            const int t = lane & 31, lh = lane >> 5;
            if (t < G.nsuf) {
                acc[w][t][2 * lh] = ks.s;
                acc[w][t][2 * lh + 1] = ks.c;
            }
        }
        __syncthreads();
        for (int o = tid; o < G.nsuf; o += BT) {
            KSum a, c;
            a.init();
            c.init();
            for (int ww = 0; ww < NWV; ++ww) {
                a.add(acc[ww][o][0]);
                a.add(acc[ww][o][1]);
                c.add(acc[ww][o][2]);
                c.add(acc[ww][o][3]);
            }
            // undo the prefix legs' amplitude scales (their D parts entered G1 divided by them)
            // and the grid ratio (ns/M)^3 of a transform on the reduced grid
            const double g = G.fscale;
            double f = g;
            if (G.scaleA) f *= G.pre_sq ? (*G.scaleA) * (*G.scaleA) : *G.scaleA;
            if (G.scaleB) f *= *G.scaleB;
            const int64_t po = (int64_t)(G.out_begin + o) * max_chunks + chunk;
            partials[2 * po + 0] = make_double2(a.s * f, a.c * f);
            partials[2 * po + 1] = make_double2(c.s * g, c.c * g);
        }
        __syncthreads();
    }
}

// ================================================================ shell intervals
// out[s] = (lo, hi): mask s selects exactly the cells with lo <= m^2 <= hi
// (radial kinds; (1, 0) when empty or not radial); cnt[s] = members found, so
// that the host can check that they form one interval.
__global__ void __launch_bounds__(NT) k_shell_intervals(const MaskParams *__restrict__ masks,
                                                        const int *__restrict__ kmax, int2 *__restrict__ out)
{
    __shared__ int slo[NT], shi[NT], scnt[NT];
    const int tid = threadIdx.x;
    const MaskParams mp = masks[blockIdx.x];
    const int K = kmax[blockIdx.x];
    const int M = 3 * K * K;
    int lo = INT_MAX, hi = -1, cnt = 0;
    if (mask_is_radial(mp.kind)) {
        for (int m2 = tid; m2 <= M; m2 += NT) {
            if (mask_eval_m2(mp, m2)) {
                lo = min(lo, m2);
                hi = max(hi, m2);
                ++cnt;
            }
        }
    }
    slo[tid] = lo;
    shi[tid] = hi;
    scnt[tid] = cnt;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            slo[tid] = min(slo[tid], slo[tid + s]);
            shi[tid] = max(shi[tid], shi[tid + s]);
            scnt[tid] += scnt[tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const bool empty = shi[0] < 0;
        out[2 * blockIdx.x] = empty ? make_int2(1, 0) : make_int2(slo[0], shi[0]);
        out[2 * blockIdx.x + 1] = make_int2(scnt[0], 0);
    }
}

// ================================================================ launchers
static int max_lds_per_block(int *out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(out, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    return (int)e;
}

template <typename T, int N>
static int launch_spec_zf_n(const SpecGroupDev *g, int ng, const LegDev *legs, int64_t t0, int64_t t1, int grid_cap,
                            int kz_cap, const void *tw, hipStream_t st)
{
    using C = ZfCfg<T, N>;
    if (t1 <= t0) return 0;
    if (kz_cap < 1 || kz_cap > N / 2 + 1) kz_cap = N / 2 + 1;
    const size_t dyn = sizeof(cplx<T>) * std::max((size_t)2 * kz_cap * C::SP, (size_t)C::ROWS * C::LP);
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&k_spec_zf<T, N>));
    if (e != hipSuccess) return (int)e;
    int lds_max = 0;
    if (int r = max_lds_per_block(&lds_max)) return r;
    if (fa.sharedSizeBytes + dyn > (size_t)lds_max) return PS_ERR_UNSUPPORTED_N;
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_spec_zf<T, N>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (e != hipSuccess) return (int)e;
    const unsigned grid = (unsigned)((t1 - t0 < grid_cap) ? t1 - t0 : grid_cap);
    hipLaunchKernelGGL((k_spec_zf<T, N>), dim3(grid), dim3(C::BT), dyn, st, g, ng, legs, t0, t1,
                       reinterpret_cast<const cplx<T> *>(tw), kz_cap);
    return (int)hipGetLastError();
}

template <typename T, int N>
static int launch_spec_fx_n(const SpecGroupDev *g, int ng, int64_t t0, int64_t t1, int grid_cap, int div,
                            const void *tw, hipStream_t st)
{
    if (t1 <= t0) return 0;
    const unsigned grid = (unsigned)((t1 - t0 < grid_cap) ? t1 - t0 : grid_cap);
    const cplx<T> *twp = reinterpret_cast<const cplx<T> *>(tw);
    if (div == 2) hipLaunchKernelGGL((k_spec_fx<T, N, 2>), dim3(grid), dim3(NT), 0, st, g, ng, t0, t1, twp);
    else hipLaunchKernelGGL((k_spec_fx<T, N, 1>), dim3(grid), dim3(NT), 0, st, g, ng, t0, t1, twp);
    return (int)hipGetLastError();
}

template <typename T, int N>
static int launch_spec_dot_n(const SpecGroupDev *g, int ng, const MaskParams *masks, const int2 *ivl, int64_t i0,
                             int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kinds, const void *tw,
                             hipStream_t st)
{
    using C = DotCfg<T, N>;
    if (i1 <= i0) return 0;
    const unsigned grid = (unsigned)((i1 - i0 < grid_cap) ? i1 - i0 : grid_cap);
    const cplx<T> *twp = reinterpret_cast<const cplx<T> *>(tw);
    if (kinds & 1)
        hipLaunchKernelGGL((k_spec_dot<T, N, true>), dim3(grid), dim3(C::BT), 0, st, g, ng, masks, ivl, i0, i1,
                           partials, max_chunks, twp);
    if (kinds & 2)
        hipLaunchKernelGGL((k_spec_dot<T, N, false>), dim3(grid), dim3(C::BT), 0, st, g, ng, masks, ivl, i0, i1,
                           partials, max_chunks, twp);
    return (int)hipGetLastError();
}

#define PS_FOR_EACH_N(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(144) X(192) X(216) X(256) X(384) X(512)

int spec_zf_rows(int N, int fp64)
{
    switch (N) {
#define CASE(n) case n: return fp64 ? ZfCfg<double, n>::ROWS : ZfCfg<float, n>::ROWS;
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return 0;
}

int spec_dot_lines(int N, int fp64)
{
    switch (N) {
#define CASE(n) case n: return fp64 ? DotCfg<double, n>::LINES : DotCfg<float, n>::LINES;
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return 0;
}

int run_spec_zf(int fp64, int N, const SpecGroupDev *groups, int ngroups, const LegDev *legs, int64_t t0, int64_t t1,
                int grid_cap, int kz_cap, const void *tw, hipStream_t st)
{
    switch (N) {
#define CASE(n)                                                                                               \
    case n:                                                                                                   \
        return fp64 ? launch_spec_zf_n<double, n>(groups, ngroups, legs, t0, t1, grid_cap, kz_cap, tw, st)   \
                    : launch_spec_zf_n<float, n>(groups, ngroups, legs, t0, t1, grid_cap, kz_cap, tw, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

int run_spec_fx(int fp64, int N, const SpecGroupDev *groups, int ngroups, int64_t t0, int64_t t1, int grid_cap,
                int div, const void *tw, hipStream_t st)
{
    switch (N) {
#define CASE(n)                                                                                      \
    case n:                                                                                          \
        return fp64 ? launch_spec_fx_n<double, n>(groups, ngroups, t0, t1, grid_cap, div, tw, st)   \
                    : launch_spec_fx_n<float, n>(groups, ngroups, t0, t1, grid_cap, div, tw, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

int run_spec_dot(int fp64, int N, const SpecGroupDev *groups, int ngroups, const MaskParams *masks, const int2 *ivl,
                 int64_t i0, int64_t i1, double2 *partials, int max_chunks, int grid_cap, int kinds, const void *tw,
                 hipStream_t st)
{
    switch (N) {
#define CASE(n)                                                                                                   \
    case n:                                                                                                       \
        return fp64 ? launch_spec_dot_n<double, n>(groups, ngroups, masks, ivl, i0, i1, partials, max_chunks,    \
                                                   grid_cap, kinds, tw, st)                                       \
                    : launch_spec_dot_n<float, n>(groups, ngroups, masks, ivl, i0, i1, partials, max_chunks,     \
                                                  grid_cap, kinds, tw, st);
        PS_FOR_EACH_N(CASE)
#undef CASE
    }
    return PS_ERR_UNSUPPORTED_N;
}

int run_shell_intervals(const MaskParams *masks, const int *kmax, int n, int2 *out, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_shell_intervals, dim3(n), dim3(NT), 0, st, masks, kmax, out);
    return (int)hipGetLastError();
}

}  // namespace ps

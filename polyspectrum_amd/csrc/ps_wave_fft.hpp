// ps_wave_fft.hpp -- wavefront-private FFTs: one wavefront transforms LW
// whole lines of length N.  Butterfly b = lane + 64*i of a stage works on line
// b / (N/R); lines are exchanged between stages through the wavefront's own
// LDS region, so no workgroup barrier is needed inside a transform.  The first
// stage takes its inputs from a producer functor (e.g. a Hermitian view of a
// staged half spectrum) and the last stage hands its outputs to a consumer
// functor in registers; every thread sees the same output positions for every
// transform, which lets per-cell products stay in registers across lines.
#pragma once
#include "ps_fft_core.hpp"

namespace ps {

// the LDS write -> other-lane read hand-off of one wavefront
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int largest_divisor_leq(int n, int cap)
{
    for (int d = cap; d >= 1; --d)
        if (n % d == 0) return d;
    return 1;
}

constexpr int largest_even_divisor_leq(int n, int cap)
{
    for (int d = cap; d >= 2; --d)
        if (n % d == 0 && d % 2 == 0) return d;
    return 0;
}

// lines per wavefront: about 1024 complex values in flight, dividing N/4 so
// that 4 wavefronts cover a whole number of rows of a tile; an even count
// where one divides (the D-only fused pass pairs rows)
template <typename T, int N> constexpr int wave_lines()
{
    constexpr int cap = (1024 / N < 1) ? 1 : 1024 / N;
    constexpr int de = largest_even_divisor_leq(N / 4, cap);
    constexpr int d = de > 0 ? de : largest_divisor_leq(N / 4, cap);
    return (sizeof(T) == 8 && d > 1) ? d / 2 : d;
}

template <int N> constexpr int wave_pitch() { return N + N / 8 + 1; }
__device__ __forceinline__ int wpad(int e) { return e + (e >> 3); }

// LDS layout of a wavefront's lines inside wave_fft: element e of line l at
// l * LP + pad(e), pad(e) = e + (e / Q) * A.  Chosen per precision and length
// with a model of the stages' accesses (32 banks of 4 bytes; a 64-lane access
// of 8-byte values served 16 lanes at a time): for FP32 the layouts below
// take the modelled bank-conflict cycles of the exchanges from 1.5-2x the
// conflict-free count to 1-1.4x at every line count (e.g. 192: 2.0 -> 1.25);
// FP64 keeps e + e/8 and pitch N + N/8 + 1.  Each one keeps the per-r
// offsets of every stage's reads and writes constant (checked below), so
// they fold into the LDS instructions' offsets.  pitch >= pad(N - 1) + 1.
struct WLayP {
    int Q, A, LP;
};
template <typename T, int N> constexpr WLayP wlay_params()
{
    if constexpr (sizeof(T) == 8) return WLayP{8, 1, N + N / 8 + 1};
    switch (N) {
    case 16: return WLayP{4, 1, 20};
    case 24: return WLayP{6, 1, 35};
    case 32: return WLayP{4, 2, 49};
    case 48: return WLayP{48, 1, 54};
    case 64: return WLayP{8, 1, 72};
    case 96: return WLayP{48, 1, 108};
    case 128: return WLayP{16, 1, 136};
    case 144: return WLayP{144, 1, 146};
    case 192: return WLayP{48, 2, 200};
    case 216: return WLayP{216, 1, 216};
    case 256: return WLayP{16, 1, 272};
    case 384: return WLayP{48, 2, 400};
    case 512: return WLayP{16, 1, 544};
    default: return WLayP{8, 1, N + N / 8 + 1};
    }
}
template <typename T, int N> struct WLay {
    static constexpr int Q = wlay_params<T, N>().Q, A = wlay_params<T, N>().A, LP = wlay_params<T, N>().LP;
    __host__ __device__ static constexpr int pad(int e) { return e + (int)((unsigned)e / (unsigned)Q) * A; }
    static_assert(pad(N - 1) < LP, "the pitch holds a whole line");
    // pad(j + r NB) = pad(j) + pad(r NB) for every j < NB (reads of a stage)
    static constexpr bool read_const(int NB, int R)
    {
        for (int j = 0; j < NB; ++j)
            for (int r = 0; r < R; ++r)
                if (pad(j + r * NB) - pad(j) != pad(r * NB)) return false;
        return true;
    }
    // pad(d + r NS) = pad(d) + pad(r NS) for the write bases d of a stage
    static constexpr bool write_const(int NB, int NS, int R)
    {
        for (int j = 0; j < NB; ++j) {
            const int d = (j / NS) * NS * R + j % NS;
            for (int r = 0; r < R; ++r)
                if (pad(d + r * NS) - pad(d) != pad(r * NS)) return false;
        }
        return true;
    }
};
// wavefront buffer of wave_fft: LW lines of this pitch
template <typename T, int N> constexpr int wave_fft_pitch() { return WLay<T, N>::LP; }


template <typename T, int N, int LW, int S> struct WaveStage {
    static constexpr int R = stage_radix(N, S);
    static constexpr int NS = stage_span(N, S);
    static constexpr int NB = N / R;
    static constexpr int TOT = LW * NB;
    static constexpr int IT = (TOT + 63) / 64;
    static constexpr bool LAST = (S == num_stages(N) - 1);
};

// outputs each thread holds after the last stage
template <typename T, int N, int LW> struct WaveOut {
    using L = WaveStage<T, N, LW, num_stages(N) - 1>;
    static constexpr int IT = L::IT;
    static constexpr int R = L::R;
    static constexpr int NB = L::NB;
    static constexpr int TOT = L::TOT;
};

// Per-thread twiddles of every stage after the first, loaded once: stage S
// butterfly i needs W^(r*k*(N/(NS*R))), k = (j mod NS), j = (lane + 64 i) mod NB.
// Butterflies i that share j share their twiddles (sets(S) distinct sets).
template <typename T, int N, int LW> struct WaveTw {
    static constexpr int sets(int S)
    {
        const int NB = N / stage_radix(N, S);
        const int IT = (LW * NB + 63) / 64;
        const int d = (64 % NB == 0) ? 1 : ((NB % 64 == 0) ? NB / 64 : IT);
        return d < IT ? d : IT;
    }
    static constexpr int off(int S)
    {
        int o = 0;
        for (int s = 1; s < S; ++s) o += sets(s) * (stage_radix(N, s) - 1);
        return o;
    }
    static constexpr int total() { return off(num_stages(N)) > 0 ? off(num_stages(N)) : 1; }
    template <int S> __device__ static constexpr int at(int i, int r)
    {
        return off(S) + (i % sets(S)) * (stage_radix(N, S) - 1) + r - 1;
    }
    cplx<T> w[total()];

    template <int SIGN, int S = 1> __device__ __forceinline__ void load(const cplx<T> *__restrict__ tw, int lane)
    {
        if constexpr (S < num_stages(N)) {
            using St = WaveStage<T, N, LW, S>;
#pragma unroll
            for (int i = 0; i < sets(S); ++i) {
                const int b = lane + 64 * i;
                const int j = (b < St::TOT ? b : St::TOT - 1) % St::NB;
                const int k = j % St::NS;
#pragma unroll
                for (int r = 1; r < St::R; ++r)
                    w[at<S>(i, r)] = twiddle<T, SIGN>(tw, r * k * (N / (St::NS * St::R)));
            }
            load<SIGN, S + 1>(tw, lane);
        }
    }
};

// Twiddles read from a table exp(-2 pi i t/N) in LDS instead of registers.
template <typename T, int N, int LW> struct LdsTw {
    const cplx<T> *tab;
    template <int S, int SIGN> __device__ __forceinline__ cplx<T> get(int i, int r, int j) const
    {
        using St = WaveStage<T, N, LW, S>;
        return twiddle<T, SIGN>(tab, r * (j % St::NS) * (N / (St::NS * St::R)));
    }
};

template <typename T, int N, int LW> struct RegTw {
    const WaveTw<T, N, LW> &w;
    template <int S, int SIGN> __device__ __forceinline__ cplx<T> get(int i, int r, int) const
    {
        return w.w[WaveTw<T, N, LW>::template at<S>(i, r)];
    }
};

// Inverse (SIGN = +1) or forward (SIGN = -1) transform of LW lines.
//   src(l, e)          -> input element e of line l (first stage)
//   sink(i, r, l, z, v) <- output z of line l (held as v[i][r] of this thread)
//   EDGE (first stage only): inputs e with NB <= e < N - NB are known to be
//   zero (a spectrum band-limited to |k| < N/R_0), so only r = 0 and r = R-1
//   are read and the butterfly is constant-folded.
template <typename T, int N, int LW, int SIGN, int S, bool EDGE = false, class Tw, class Src, class Sink>
__device__ __forceinline__ void wave_fft(cplx<T> *__restrict__ wbuf, const Tw &tw, int lane, Src &src, Sink &sink)
{
    using St = WaveStage<T, N, LW, S>;
    constexpr int R = St::R, NS = St::NS, NB = St::NB, TOT = St::TOT, IT = St::IT;
    using Lay = WLay<T, N>;
    constexpr int LP = Lay::LP;
    cplx<T> v[IT][R];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int b = lane + 64 * i;
        if (TOT % 64 == 0 || b < TOT) {
            const int j = b % NB, l = b / NB;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (S == 0) {
                    if (EDGE && r != 0 && r != R - 1) v[i][r] = mk<T>(0, 0);
                    else v[i][r] = src(l, j + r * NB);
                } else if constexpr (Lay::read_const(NB, R)) {
                    v[i][r] = wbuf[l * LP + Lay::pad(j) + Lay::pad(r * NB)];
                } else {
                    v[i][r] = wbuf[l * LP + Lay::pad(j + r * NB)];
                }
            }
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul<T>(v[i][r], tw.template get<S, SIGN>(i, r, j));
            }
            dft<T, R, SIGN>(v[i]);
        }
    }
    if constexpr (St::LAST) {
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int b = lane + 64 * i;
            if (TOT % 64 == 0 || b < TOT) {
                const int j = b % NB, l = b / NB;
#pragma unroll
                for (int r = 0; r < R; ++r) sink(i, r, l, j + r * NB, v[i][r]);
            }
        }
    } else {
        wave_sync();
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int b = lane + 64 * i;
            if (TOT % 64 == 0 || b < TOT) {
                const int j = b % NB, l = b / NB;
                const int d = (j / NS) * NS * R + j % NS;
                if constexpr (Lay::write_const(NB, NS, R)) {
                    const int base = l * LP + Lay::pad(d);
#pragma unroll
                    for (int r = 0; r < R; ++r) wbuf[base + Lay::pad(r * NS)] = v[i][r];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) wbuf[l * LP + Lay::pad(d + r * NS)] = v[i][r];
                }
            }
        }
        wave_sync();
        wave_fft<T, N, LW, SIGN, S + 1, false, Tw>(wbuf, tw, lane, src, sink);
    }
}

// plain FP64 tree sum over a wavefront (valid in every lane)
__device__ __forceinline__ double wave_sum(double s)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// a double moved across lanes by a DPP control (all lanes of the row valid)
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// FP64 sum over a wavefront (valid in every lane): within each 16-lane row by
// DPP moves (quad_perm [1,0,3,2], [2,3,0,1], row_ror 4, row_ror 8: VALU
// latency only), then across the four rows by two lane shuffles.  A fixed
// order, so the result is reproducible.
__device__ __forceinline__ double wave_sum_dpp(double s)
{
    s += dpp_f64<0xB1>(s);
    s += dpp_f64<0x4E>(s);
    s += dpp_f64<0x124>(s);
    s += dpp_f64<0x128>(s);
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    return s;
}

}  // namespace ps

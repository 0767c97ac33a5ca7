// ps_fft_core.hpp -- device building blocks of the gfx950 FFT engine.
//
// A workgroup transforms TLC complex lines of length N held in LDS.  Element
// e of line l lives at lidx<TLC, T>(e, l) = e*TLC + (l ^ (e & (TLC-1))): lines
// are interleaved (line index fastest) and XOR-swizzled, so both the strided
// global tiles (TLC adjacent columns) and the Stockham butterflies hit
// distinct LDS banks without padding.
//
// Stockham autosort, decimation in time: stage s with radix R and span Ns
// (product of earlier radices) reads v[r] = x[j + r*N/R], twiddles v[r] by
// W_N^(r*(j mod Ns)*N/(Ns*R)), applies DFT_R and writes to
// ((j div Ns)*Ns*R + j mod Ns) + r*Ns.  The result is in natural order.
//
// Convention (the reference's, fft.c:19-57): SIGN = -1 is FFTW_FORWARD
// (e^{-i}), SIGN = +1 is FFTW_BACKWARD (e^{+i}); nothing is normalised.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ps {

constexpr int NT = 256;  // threads per workgroup for every FFT kernel

template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using cplx = typename cvec<T>::type;

template <typename T> __device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by S*i
template <typename T, int S> __device__ __forceinline__ cplx<T> mul_si(cplx<T> a)
{
    return (S > 0) ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

// the XOR key takes e >> SH: with SH = log2(128 / (TLC * element bytes)),
// the elements e, e + 1, ... of one line fill distinct bank groups of every
// 128 bytes (one pass of a 32-bank LDS), which the tiles' global loads and
// stores (e fastest over a workgroup's threads) need; the stages' accesses
// (l fastest) stay conflict-free for any key
template <int TLC, typename T> constexpr int lidx_shift()
{
    int sh = 0;
    while ((TLC * (int)sizeof(cplx<T>) << sh) < 128) ++sh;
    return sh;
}
template <int TLC, typename T, bool SWZ = true> __device__ __forceinline__ int lidx(int e, int l)
{
    // SWZ = false: plain interleaving e*TLC + l, for tiles whose every
    // access has the line index fastest (then any key is conflict-free) and
    // that are filled by LDS copies of 16-byte pairs of adjacent lines
    if constexpr (!SWZ) return e * TLC + l;
    return e * TLC + (l ^ ((e >> lidx_shift<TLC, T>()) & (TLC - 1)));
}

// ---------------------------------------------------------------- butterflies
template <typename T, int S> __device__ __forceinline__ void dft2(cplx<T> *v)
{
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
}

template <typename T, int S> __device__ __forceinline__ void dft3(cplx<T> *v)
{
    const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
    cplx<T> a = v[0], b = v[1], c = v[2];
    cplx<T> t1 = cadd<T>(b, c);
    cplx<T> t2 = mk<T>(a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y);
    cplx<T> d = csub<T>(b, c);
    cplx<T> t3 = mul_si<T, S>(mk<T>(h * d.x, h * d.y));
    v[0] = cadd<T>(a, t1);
    v[1] = cadd<T>(t2, t3);
    v[2] = csub<T>(t2, t3);
}

template <typename T, int S>
__device__ __forceinline__ void dft4v(cplx<T> a, cplx<T> b, cplx<T> c, cplx<T> d, cplx<T> &X0, cplx<T> &X1,
                                      cplx<T> &X2, cplx<T> &X3)
{
    cplx<T> y0 = cadd<T>(a, c), y1 = csub<T>(a, c), y2 = cadd<T>(b, d);
    cplx<T> y3 = mul_si<T, S>(csub<T>(b, d));
    X0 = cadd<T>(y0, y2);
    X1 = cadd<T>(y1, y3);
    X2 = csub<T>(y0, y2);
    X3 = csub<T>(y1, y3);
}

template <typename T, int S> __device__ __forceinline__ void dft4(cplx<T> *v)
{
    dft4v<T, S>(v[0], v[1], v[2], v[3], v[0], v[1], v[2], v[3]);
}

template <typename T, int S> __device__ __forceinline__ void dft8(cplx<T> *v)
{
    const T r2 = (T)0.70710678118654752440084436210485;  // 1/sqrt(2)
    cplx<T> E0, E1, E2, E3, O0, O1, O2, O3;
    dft4v<T, S>(v[0], v[2], v[4], v[6], E0, E1, E2, E3);
    dft4v<T, S>(v[1], v[3], v[5], v[7], O0, O1, O2, O3);
    // O1 *= (1 + S i)/sqrt2 ; O2 *= S i ; O3 *= (-1 + S i)/sqrt2
    O1 = mk<T>(r2 * (O1.x - (T)S * O1.y), r2 * (O1.y + (T)S * O1.x));
    O2 = mul_si<T, S>(O2);
    O3 = mk<T>(r2 * (-O3.x - (T)S * O3.y), r2 * (-O3.y + (T)S * O3.x));
    v[0] = cadd<T>(E0, O0);
    v[4] = csub<T>(E0, O0);
    v[1] = cadd<T>(E1, O1);
    v[5] = csub<T>(E1, O1);
    v[2] = cadd<T>(E2, O2);
    v[6] = csub<T>(E2, O2);
    v[3] = cadd<T>(E3, O3);
    v[7] = csub<T>(E3, O3);
}

template <typename T, int R, int S> __device__ __forceinline__ void dft(cplx<T> *v)
{
    if constexpr (R == 2) dft2<T, S>(v);
    else if constexpr (R == 3) dft3<T, S>(v);
    else if constexpr (R == 4) dft4<T, S>(v);
    else dft8<T, S>(v);
}

// ---------------------------------------------------------------- radix plan
// N = 3^b * 2^a: b radix-3 stages, then a/3 radix-8 stages, then one radix-4
// or radix-2 stage for the remainder.
constexpr int pow3_count(int n) { int c = 0; while (n % 3 == 0) { n /= 3; ++c; } return c; }
constexpr int log2_of(int n) { int c = 0; while (n > 1) { n >>= 1; ++c; } return c; }
constexpr int pow2_part(int n) { while (n % 3 == 0) n /= 3; return n; }
constexpr int num_stages(int N)
{
    return pow3_count(N) + log2_of(pow2_part(N)) / 3 + ((log2_of(pow2_part(N)) % 3) ? 1 : 0);
}
constexpr int stage_radix(int N, int s)
{
    const int n3 = pow3_count(N), a = log2_of(pow2_part(N));
    if (s < n3) return 3;
    s -= n3;
    if (s < a / 3) return 8;
    return (a % 3 == 2) ? 4 : 2;
}
constexpr int stage_span(int N, int s)
{
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= stage_radix(N, i);
    return ns;
}
// Alternative plans (chosen per context by ps_forward_autotune for the
// forward transform): PLAN 0 is the order above; PLAN 1 runs the radix-4/2
// stage first, then the radix-8 stages, then the radix-3 stages.  Both have
// num_stages(N) stages.
template <int PLAN> constexpr int stage_radix_p(int N, int s)
{
    if constexpr (PLAN == 0) {
        return stage_radix(N, s);
    } else {
        const int n3 = pow3_count(N), a = log2_of(pow2_part(N));
        const int nrem = (a % 3) ? 1 : 0;
        if (s < nrem) return (a % 3 == 2) ? 4 : 2;
        s -= nrem;
        if (s < a / 3) return 8;
        (void)n3;
        return 3;
    }
}
template <int PLAN> constexpr int stage_span_p(int N, int s)
{
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= stage_radix_p<PLAN>(N, i);
    return ns;
}
constexpr int kFwdPlans = 2;

constexpr bool supported_n(int N)
{
    return N >= 4 && pow2_part(N) >= 2 && (pow2_part(N) & (pow2_part(N) - 1)) == 0;
}

// twiddle W_N^t for direction S from the forward table tw[t] = exp(-2 pi i t/N)
template <typename T, int S> __device__ __forceinline__ cplx<T> twiddle(const cplx<T> *__restrict__ tw, int t)
{
    cplx<T> w = tw[t];
    if (S > 0) w.y = -w.y;
    return w;
}

// One Stockham stage over TLC lines in LDS.  Starts by reading LDS (callers
// synchronise before the first stage) and ends with a barrier.
template <typename T, int N, int TLC, int S, int R, int NS, bool SWZ = true>
__device__ __forceinline__ void stage(cplx<T> *lds, const cplx<T> *__restrict__ tw, int tid)
{
    constexpr int NB = N / R;
    constexpr int TOT = TLC * NB;
    constexpr int IT = (TOT + NT - 1) / NT;
    cplx<T> v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int b = tid + it * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int l = b % TLC, j = b / TLC;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = lds[lidx<TLC, T, SWZ>(j + r * NB, l)];
            if constexpr (NS > 1) {
                const int k = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) v[it][r] = cmul<T>(v[it][r], twiddle<T, S>(tw, r * k * (N / (NS * R))));
            }
            dft<T, R, S>(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int b = tid + it * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int l = b % TLC, j = b / TLC;
            const int k = j % NS;
            const int d = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lidx<TLC, T, SWZ>(d + r * NS, l)] = v[it][r];
        }
    }
    __syncthreads();
}

// f(idx) for idx = tid, tid + NT, ... < TOT, fully unrolled so that the
// independent global loads of all iterations are in flight together
template <int TOT, class F> __device__ __forceinline__ void tile_loop(int tid, F &&f)
{
#pragma unroll
    for (int k = 0; k < (TOT + NT - 1) / NT; ++k) {
        const int idx = tid + k * NT;
        if (TOT % NT == 0 || idx < TOT) f(idx);
    }
}

// the same with at most U iterations unrolled (bodies with much per-item work)
template <int TOT, int U, class F> __device__ __forceinline__ void tile_loop_u(int tid, F &&f)
{
#pragma unroll U
    for (int k = 0; k < (TOT + NT - 1) / NT; ++k) {
        const int idx = tid + k * NT;
        if (TOT % NT == 0 || idx < TOT) f(idx);
    }
}

// The twiddle table copied into LDS once per workgroup: the stages read
// their twiddles inside the dependency chain of every butterfly, where a
// global (cache) read would stall the stage.  The caller's next barrier
// publishes it.
template <typename T, int N, int NTH = NT>
__device__ __forceinline__ void stage_twiddles(cplx<T> *stw, const cplx<T> *__restrict__ tw, int tid)
{
    for (int i = tid; i < N; i += NTH) stw[i] = tw[i];
}

template <typename T, int N, int TLC, int S, int ST = 0, int PLAN = 0, bool SWZ = true>
__device__ __forceinline__ void fft_lds(cplx<T> *lds, const cplx<T> *__restrict__ tw, int tid)
{
    if constexpr (ST < num_stages(N)) {
        stage<T, N, TLC, S, stage_radix_p<PLAN>(N, ST), stage_span_p<PLAN>(N, ST), SWZ>(lds, tw, tid);
        fft_lds<T, N, TLC, S, ST + 1, PLAN, SWZ>(lds, tw, tid);
    }
}

// ---------------------------------------------------------------- reductions
// Neumaier-compensated FP64 accumulator.
struct KSum {
    double s, c;
    __device__ __forceinline__ void init() { s = 0.0; c = 0.0; }
    __device__ __forceinline__ void add(double x)
    {
        double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

// error-free sum of a wavefront's values (TwoSum pairs), result valid in all lanes
__device__ __forceinline__ void wave_sum2(double &s, double &c)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double os = __shfl_xor(s, off, 64);
        double oc = __shfl_xor(c, off, 64);
        double t = s + os;
        double bp = t - s;
        double e = (s - (t - bp)) + (os - bp);
        s = t;
        c = c + oc + e;
    }
}

}  // namespace ps

// ps_device.hpp -- small device helpers shared by the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ps_fft_core.hpp"
#include "ps_internal.h"

namespace ps {

// rows per tile of the z passes and lines per tile of the other passes
template <typename T, int N> constexpr int tile_lines() { return tile_lines_for(N, sizeof(T) == 8); }

// Loads and stores through pointers kept in the device tables (LegDev,
// SpecGroupDev): the compiler cannot tell their address space and would emit
// flat instructions, which count on the LDS counter too, so that the
// transforms' LDS waits would also wait for the next tile's prefetched loads.
// The buffers are device (global) memory: these go through the global
// address space.
#define PS_GLOBAL __attribute__((address_space(1)))
template <typename T> struct gvec;
template <> struct gvec<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct gvec<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename T> __device__ __forceinline__ cplx<T> gload(const cplx<T> *p, int64_t i)
{
    const typename gvec<T>::type v = ((const PS_GLOBAL typename gvec<T>::type *)p)[i];
    return mk<T>(v.x, v.y);
}
template <typename T> __device__ __forceinline__ void gstore(cplx<T> *p, int64_t i, cplx<T> v)
{
    typename gvec<T>::type t;
    t.x = v.x;
    t.y = v.y;
    ((PS_GLOBAL typename gvec<T>::type *)p)[i] = t;
}

// The same tables read through the constant address space: uniform loads
// from it are scalar (s_load, counted on lgkmcnt), so reading a table entry
// never waits for the vector memory instructions in flight.  Only for tables
// that no kernel writes.
#define PS_CONST __attribute__((address_space(4)))
template <typename X> __device__ __forceinline__ const PS_CONST X *const_tab(const X *p)
{
    return (const PS_CONST X *)p;
}

// band index j in [0, B) <-> grid index k of the band |m| <= K (B = min(2K+1, N))
__device__ __forceinline__ int band_to_k(int j, int K, int B, int N) { return (B == N) ? j : (j <= K ? j : j + N - B); }
__device__ __forceinline__ int k_to_band(int k, int K, int B, int N)
{
    if (B == N) return k;
    if (k <= K) return k;
    if (k >= N - K) return k - N + B;
    return -1;
}

// find the leg (or group) owning a tile: largest i with begin[i] <= t
template <typename D>
__device__ __forceinline__ int find_owner(const D *__restrict__ d, int n, int64_t t, int64_t D::*field)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].*field <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace ps

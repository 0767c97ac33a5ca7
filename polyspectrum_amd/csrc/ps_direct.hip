// ps_direct.hip -- the polygon sums without any transform: an independent
// check of the estimator on the device.
//
// With the unnormalised transforms of the estimator (fft.c:19-57),
//   sum_x prod_i D_i(x) = N^3 * sum over closed lattice polygons of prod_i H(m_i),
//   sum_x prod_i I_i(x) = N^3 * number of closed lattice polygons,
// where a closed polygon is m_1 + ... + m_n = 0 (mod N) with every m_i in the
// shell of leg i (Documentation.tex:124-162).  For bispectra the kernel runs
// over the pairs (m_1, m_2) of the first two shells, closes each with
// m_3 = -(m_1 + m_2) wrapped onto the grid, and tests m_3 against up to
// DIRECT_MEMBERS third-leg shells at once (the theta bins of one k1, k2 pair
// share the pairs).  For power spectra m_2 = -m_1.  Values are read from the
// resident half spectrum through its Hermitian symmetry; the products and
// sums are FP64; every workgroup writes one partial (sum, count) per member,
// summed on the host in a fixed order, so results are reproducible.
#include <algorithm>

#include "ps_device.hpp"
#include "ps_wave_fft.hpp"

namespace ps {

// cell coordinates in [-N/2, N/2) packed 10 bits each (N <= 1024)
__device__ __forceinline__ void unpack_cell(int c, int h, int &mx, int &my, int &mz)
{
    mx = (c & 1023) - h;
    my = ((c >> 10) & 1023) - h;
    mz = ((c >> 20) & 1023) - h;
}

// the coordinate of a lattice index in [-N, N] on the grid: [-N/2, N/2)
__device__ __forceinline__ int wrap_half(int v, int N)
{
    const int h = N / 2;
    if (v >= h) v -= N;
    if (v < -h) v += N;
    return v;
}

// H(m) of the full spectrum from the half spectrum H[kz][kx][ky] (kz <= N/2):
// H(-m) = conj H(m) for a real field; phase-only spectra use z/|z|
// (utils_fftw.c:78-93), as the masked gather of the estimator does
template <typename T>
__device__ __forceinline__ double2 spec_at(const cplx<T> *__restrict__ H, int N, int phase, int mx, int my, int mz)
{
    bool cj = false;
    if (mz < 0) {
        mx = -mx;
        my = -my;
        mz = -mz;
        cj = true;
    }
    const int kx = mx < 0 ? mx + N : mx, ky = my < 0 ? my + N : my;
    cplx<T> d = H[((int64_t)mz * N + kx) * N + ky];
    if (phase) {
        const T a = hypot(d.x, d.y);
        d = mk<T>(d.x / a, d.y / a);
    }
    return make_double2((double)d.x, cj ? -(double)d.y : (double)d.y);
}

template <typename T>
__global__ void __launch_bounds__(NT) k_direct_values(const cplx<T> *__restrict__ H, int N, int phase,
                                                      const int *__restrict__ cells, int n, double2 *__restrict__ out)
{
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        int mx, my, mz;
        unpack_cell(cells[i], N / 2, mx, my, mz);
        out[i] = spec_at<T>(H, N, phase, mx, my, mz);
    }
}

// sums and counts of the members, reduced over the workgroup into partials[blockIdx][j]
__device__ __forceinline__ void block_partials(const double *acc, const int *cnt, int nm, double2 *__restrict__ partials)
{
    __shared__ double2 red[NT / 64][DIRECT_MEMBERS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < DIRECT_MEMBERS; ++j) {
        if (j < nm) {
            const double s = wave_sum(acc[j]);
            const double c = wave_sum((double)cnt[j]);
            if (lane == 0) red[w][j] = make_double2(s, c);
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < nm) {
        double s = 0.0, c = 0.0;
#pragma unroll
        for (int v = 0; v < NT / 64; ++v) {
            s += red[v][threadIdx.x].x;
            c += red[v][threadIdx.x].y;
        }
        partials[(int64_t)blockIdx.x * DIRECT_MEMBERS + threadIdx.x] = make_double2(s, c);
    }
}

// one closing cell m tested against every member's shell; its value is read
// only when a member selects it
template <typename T>
__device__ __forceinline__ void close_polygon(const DirectMembers &mem, const cplx<T> *__restrict__ H, int N, int phase,
                                              int mx, int my, int mz, double wr, double wi, double *acc, int *cnt)
{
    const int m2 = mx * mx + my * my + mz * mz;
    double pmin = 0.0, pmax = 0.0;
    if (mem.corners) {
        double lo2, hi2;
        corner_extremes(mx, my, mz, mem.bw, lo2, hi2);
        pmin = sqrt(lo2);
        pmax = sqrt(hi2);
    }
    const double m = sqrt((double)m2);
    bool sel[DIRECT_MEMBERS];
    bool any = false;
#pragma unroll
    for (int j = 0; j < DIRECT_MEMBERS; ++j) {
        sel[j] = j < mem.nm && mask_eval_norms(mem.mask[j], m, pmin, pmax);
        any |= sel[j];
    }
    if (!any) return;
    const double2 v = spec_at<T>(H, N, phase, mx, my, mz);
    const double p = wr * v.x - wi * v.y;  // Re(w v)
#pragma unroll
    for (int j = 0; j < DIRECT_MEMBERS; ++j) {
        if (sel[j]) {
            acc[j] += p;
            cnt[j] += 1;
        }
    }
}

// bispectra: workgroups take cells m_1 of shell 1, threads the cells m_2 of shell 2
template <typename T>
__global__ void __launch_bounds__(NT) k_direct_pairs(const cplx<T> *__restrict__ H, int N, int phase,
                                                     const int *__restrict__ c1, const double2 *__restrict__ v1, int n1,
                                                     const int *__restrict__ c2, const double2 *__restrict__ v2, int n2,
                                                     const DirectMembers mem, double2 *__restrict__ partials)
{
    double acc[DIRECT_MEMBERS];
    int cnt[DIRECT_MEMBERS];
#pragma unroll
    for (int j = 0; j < DIRECT_MEMBERS; ++j) {
        acc[j] = 0.0;
        cnt[j] = 0;
    }
    const int h = N / 2;
    for (int a = blockIdx.x; a < n1; a += gridDim.x) {
        int x1, y1, z1;
        unpack_cell(c1[a], h, x1, y1, z1);
        const double2 a1 = v1[a];
        for (int b = threadIdx.x; b < n2; b += NT) {
            int x2, y2, z2;
            unpack_cell(c2[b], h, x2, y2, z2);
            const double2 a2 = v2[b];
            const double wr = a1.x * a2.x - a1.y * a2.y, wi = a1.x * a2.y + a1.y * a2.x;
            close_polygon<T>(mem, H, N, phase, wrap_half(-(x1 + x2), N), wrap_half(-(y1 + y2), N),
                             wrap_half(-(z1 + z2), N), wr, wi, acc, cnt);
        }
    }
    block_partials(acc, cnt, mem.nm, partials);
}

// power spectra: threads take cells m_1 of shell 1, closed by m_2 = -m_1
template <typename T>
__global__ void __launch_bounds__(NT) k_direct_pairs2(const cplx<T> *__restrict__ H, int N, int phase,
                                                      const int *__restrict__ c1, const double2 *__restrict__ v1, int n1,
                                                      const DirectMembers mem, double2 *__restrict__ partials)
{
    double acc[DIRECT_MEMBERS];
    int cnt[DIRECT_MEMBERS];
#pragma unroll
    for (int j = 0; j < DIRECT_MEMBERS; ++j) {
        acc[j] = 0.0;
        cnt[j] = 0;
    }
    const int h = N / 2;
    for (int a = blockIdx.x * NT + threadIdx.x; a < n1; a += gridDim.x * NT) {
        int x1, y1, z1;
        unpack_cell(c1[a], h, x1, y1, z1);
        const double2 a1 = v1[a];
        close_polygon<T>(mem, H, N, phase, wrap_half(-x1, N), wrap_half(-y1, N), wrap_half(-z1, N), a1.x, a1.y, acc,
                         cnt);
    }
    block_partials(acc, cnt, mem.nm, partials);
}

int run_direct_values(int fp64, const void *H, int N, int phase, const int *cells, int n, double2 *out,
                      hipStream_t st)
{
    if (n <= 0) return 0;
    const int grid = std::min((n + NT - 1) / NT, 4096);
    if (fp64)
        hipLaunchKernelGGL(k_direct_values<double>, dim3(grid), dim3(NT), 0, st, (const cplx<double> *)H, N, phase,
                           cells, n, out);
    else
        hipLaunchKernelGGL(k_direct_values<float>, dim3(grid), dim3(NT), 0, st, (const cplx<float> *)H, N, phase,
                           cells, n, out);
    return (int)hipGetLastError();
}

int direct_sum_blocks(int n1, int n2) { return n2 > 0 ? std::min(n1, 16384) : std::min((n1 + NT - 1) / NT, 1024); }

int run_direct_sums(int fp64, const void *H, int N, int phase, const int *c1, const double2 *v1, int n1,
                    const int *c2, const double2 *v2, int n2, const DirectMembers &mem, double2 *partials,
                    hipStream_t st)
{
    const int grid = direct_sum_blocks(n1, n2);
    if (grid <= 0) return 0;
    if (n2 > 0) {
        if (fp64)
            hipLaunchKernelGGL(k_direct_pairs<double>, dim3(grid), dim3(NT), 0, st, (const cplx<double> *)H, N, phase,
                               c1, v1, n1, c2, v2, n2, mem, partials);
        else
            hipLaunchKernelGGL(k_direct_pairs<float>, dim3(grid), dim3(NT), 0, st, (const cplx<float> *)H, N, phase,
                               c1, v1, n1, c2, v2, n2, mem, partials);
    } else {
        if (fp64)
            hipLaunchKernelGGL(k_direct_pairs2<double>, dim3(grid), dim3(NT), 0, st, (const cplx<double> *)H, N, phase,
                               c1, v1, n1, mem, partials);
        else
            hipLaunchKernelGGL(k_direct_pairs2<float>, dim3(grid), dim3(NT), 0, st, (const cplx<float> *)H, N, phase,
                               c1, v1, n1, mem, partials);
    }
    return (int)hipGetLastError();
}

}  // namespace ps

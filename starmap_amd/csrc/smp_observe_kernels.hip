// gfx950 kernels of the observables that example scripts of the reference
// compute from the state: per-component sums of squares and of squared
// differences along each axis (the moment-order norms of
// starmap_2d_ex_lattice_filter_convergence.m:66-76 and
// starmap_2d_ex_gauss_filter_convergence.m), and linear interpolation of
// every component at one point (the interp2 loop of
// starmap_2d_ex_linesource_filter_reconstruction.m:118-121).
#include "smp_kernels.h"

#include <algorithm>
#include <cmath>

namespace smp {

namespace {

__device__ __forceinline__ double wave_sum4(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace

// Per block: sum of u^2 and, per kernel axis, of (u(next) - u)^2 over the
// pairs of neighbours in the reference's output order (part[4*block + 0..3]).
template <typename T>
__global__ __launch_bounds__(256) void norm_kernel(const T *U, NormGeom g, double *part) {
    __shared__ double red[4][4];
    const int64_t total = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    const int64_t stride[3] = {1, g.pitch0, (int64_t)g.pitch0 * g.ext1};
    double acc[4] = {0, 0, 0, 0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int ii[3];
        ii[0] = (int)(i % g.ext[0]);
        const int64_t r = i / g.ext[0];
        ii[1] = (int)(r % g.ext[1]);
        ii[2] = (int)(r / g.ext[1]);
        const int64_t p = ii[0] + stride[1] * ii[1] + stride[2] * ii[2];
        const double v = (double)U[p];
        acc[0] += v * v;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            // rotated axis (periodic faces, stored index (f+1) mod n): the
            // pairs (s, s+1 mod n) for s != 0; otherwise (s, s+1), s+1 < ext
            int nxt = -1;
            if (g.rot[e]) {
                if (ii[e] != 0) nxt = ii[e] + 1 == g.ext[e] ? 0 : ii[e] + 1;
            } else if (ii[e] + 1 < g.ext[e]) {
                nxt = ii[e] + 1;
            }
            if (nxt >= 0) {
                const double d = (double)U[p + (nxt - ii[e]) * stride[e]] - v;
                acc[1 + e] += d * d;
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double s = wave_sum4(acc[k]);
        if (lane == 0) red[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[threadIdx.x][w];
        part[4 * blockIdx.x + threadIdx.x] = s;
    }
}

// out4[k] = sum over blocks of part[4*block + k], one wavefront, fixed order
__global__ void norm_final_kernel(const double *part, int blocks, double *out4) {
    double a[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < blocks; b += 64)
        for (int k = 0; k < 4; ++k) a[k] += part[4 * b + k];
    for (int k = 0; k < 4; ++k) {
        const double s = wave_sum4(a[k]);
        if (threadIdx.x == 0) out4[k] = s;
    }
}

// out[j] = sum_k w[j][k] * U_j[off[j][k]]: one thread per component
template <typename T>
__global__ void probe_kernel(const T *U, int64_t slab, const ProbeComp *pc, int n, double *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const ProbeComp &c = pc[j];
    double s = 0;
    for (int k = 0; k < c.n; ++k) s += c.w[k] * (double)U[(int64_t)j * slab + c.off[k]];
    out[j] = c.nan ? NAN : s;
}

// flag = 1 if any of U[0..count) is NaN or infinite: a grid-stride scan of
// the whole state (padding elements hold zeros), one atomic OR per wave that
// saw such a value.  The flag is cleared by the caller.
template <typename T>
__global__ __launch_bounds__(256) void finite_kernel(const T *U, int64_t count, int *flag) {
    bool bad = false;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i + 3 * step < count; i += 4 * step) {
        const T a = U[i], b = U[i + step], c = U[i + 2 * step], d = U[i + 3 * step];
        bad |= !(isfinite(a) && isfinite(b) && isfinite(c) && isfinite(d));
    }
    for (; i < count; i += step) bad |= !isfinite(U[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

template <typename T>
hipError_t launch_finite(const T *U, int64_t count, int *flag, hipStream_t s) {
    int64_t blocks = (count + 4 * 256 - 1) / (4 * 256);
    blocks = std::max<int64_t>(1, std::min<int64_t>(blocks, 8192));
    hipLaunchKernelGGL(finite_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, U, count, flag);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_norm(const T *U, const NormGeom &g, double *part, int blocks, double *out4,
                       hipStream_t s) {
    hipLaunchKernelGGL(norm_kernel<T>, dim3(blocks), dim3(256), 0, s, U, g, part);
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(64), 0, s, part, blocks, out4);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_probe(const T *U, int64_t slab, const ProbeComp *pc, int n, double *out,
                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(probe_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, U, slab, pc, n, out);
    return hipGetLastError();
}

#define SMP_OBS_INSTANTIATE(T)                                                              \
    template hipError_t launch_finite<T>(const T *, int64_t, int *, hipStream_t);             \
    template hipError_t launch_norm<T>(const T *, const NormGeom &, double *, int, double *,  \
                                       hipStream_t);                                        \
    template hipError_t launch_probe<T>(const T *, int64_t, const ProbeComp *, int, double *, \
                                        hipStream_t);

SMP_OBS_INSTANTIATE(double)
SMP_OBS_INSTANTIATE(float)

}  // namespace smp

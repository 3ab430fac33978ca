// zola_project.hpp -- projection of the per-PSF adjoint phases on the Zernike
// basis: dL/da_j += sum_p Z_j(p) Q(p) with Q(p) = sum_b dL/dphase_b(p), over a
// pupil window of W points (ZernikePhase.java:200-224 builds the phase as
// sum_j a_j Z_j on the same points, so this is its transpose).
//
// Two launches, both deterministic (fixed summation orders, no atomics):
//  * k_window_qsum: Q(p) in FP64.  A block of 512 threads owns P consecutive
//    points and 512/P row groups (rows g, g + 512/P, ...), with U loads in
//    flight per thread; the group sums are combined in LDS in group order.
//    FP32 rows (64 grid, 1024 points): P = 8, U = 4 -- a few hundred PSFs are
//    one round of loads over 128 blocks; FP64 rows (256 grid, 16,384 points):
//    P = 32, U = 16 over 512 blocks.
//  * k_window_project<NT>: one block per coefficient, each thread a fixed set
//    of points (eight loads of Z and of Q in flight), then an LDS tree.
// The PSF sum is formed once for all coefficients, and every launch keeps
// enough loads in flight that the batch's dL/dphase rows stream at the
// L2/HBM rate instead of one dependent load round per eight rows.
#pragma once
#include <hip/hip_runtime.h>

namespace zola {

// Points, row groups and loads in flight per thread of k_window_qsum.
template <class T>
struct QsumShape {
    static constexpr int P = sizeof(T) == 4 ? 8 : 32, G = 512 / P, U = sizeof(T) == 4 ? 4 : 16;
};

template <class T>
__global__ __launch_bounds__(512) void k_window_qsum(int B, int W, const T* __restrict__ qbar,
                                                     double* __restrict__ Q) {
    constexpr int P = QsumShape<T>::P, G = QsumShape<T>::G, U = QsumShape<T>::U;
    __shared__ double part[G][P];
    const int lp = threadIdx.x % P, g = threadIdx.x / P;
    const int p = blockIdx.x * P + lp;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (p < W) {
        for (int b0 = g; b0 < B; b0 += G * U) {
            T v[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int b = b0 + G * k;
                v[k] = b < B ? qbar[(size_t)b * W + p] : T(0);
            }
#pragma unroll
            for (int k = 0; k < U; k++) acc[k & 3] += (double)v[k];
        }
    }
    part[g][lp] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (g == 0 && p < W) {
        double s = part[0][lp];
        for (int k = 1; k < G; k++) s += part[k][lp];
        Q[p] = s;
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_window_project(int W, const double* __restrict__ Zw,
                                                       const double* __restrict__ Q, double* __restrict__ da,
                                                       int assign) {
    constexpr int kU = 8;
    __shared__ double red[NT];
    const double* z = Zw + (size_t)blockIdx.x * W;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int p0 = threadIdx.x; p0 < W; p0 += NT * kU) {
        double zv[kU], qv[kU];
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int p = p0 + NT * k;
            zv[k] = p < W ? z[p] : 0.0;
            qv[k] = p < W ? Q[p] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kU; k++) acc[k & 3] += zv[k] * qv[k];
    }
    red[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) da[blockIdx.x] = assign ? red[0] : da[blockIdx.x] + red[0];
}

// Q into `Q` (W doubles), then da[0..J) += Z Q (assign: da = Z Q, so that a
// caller need not clear da first).
template <class T, int NT>
inline hipError_t launch_window_projection(int B, int J, int W, const double* Zw, const T* qbar, double* Q,
                                           double* da, hipStream_t st, bool assign = false) {
    if (B <= 0 || J <= 0) return hipSuccess;
    constexpr int P = QsumShape<T>::P;
    hipLaunchKernelGGL(k_window_qsum<T>, dim3((W + P - 1) / P), dim3(512), 0, st, B, W, qbar, Q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_window_project<NT>, dim3(J), dim3(NT), 0, st, W, Zw, (const double*)Q, da, assign ? 1 : 0);
    return hipGetLastError();
}

}  // namespace zola

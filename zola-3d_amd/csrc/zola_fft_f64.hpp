// zola_fft_f64.hpp -- batched FP64 line FFTs in LDS for the frame detector.
//
// A workgroup of 256 threads transforms C = E / L lines of length L (L a power
// of two, 16..E; E = 2048 elements, 32 KiB of LDS as split re/im arrays, or
// 4096 for the long column passes, so that a block reads 128-byte runs of
// adjacent columns).
// The transform is the Stockham autosort formulation (Govindaraju et al.,
// "High Performance Discrete Fourier Transforms on Graphics Processors",
// SC 2008): pass p with radix R and span Ns reads R elements L/R apart,
// twiddles them by W_{Ns R}^{r (j mod Ns)}, does a radix-R DFT in registers
// and writes them Ns apart from (j div Ns) Ns R + (j mod Ns).  Radix-8 passes
// first, then one radix-4 or radix-2 pass.  Output is in natural order and is
// scaled by 1/sqrt(L) (unitary, the convention of utils/
// FastFourierTransform.java: forward e^{-i}, inverse e^{+i}, 1/sqrt(n) per axis).
//
// Loads and stores go through an IO functor so that placement, the spectrum
// product and the crop/normalisation are fused into the passes:
//   double2 io.load(int line, int k)      element k of line `line`
//   void    io.store(int line, int k, double2 v)
// COLMAJOR selects the element->thread mapping for the global accesses: lines
// fastest (adjacent lines are adjacent in memory: column passes) or elements
// fastest (row passes).
#pragma once
#include <hip/hip_runtime.h>

#include "zola_detect.h"

namespace zola {

constexpr double kSqrtHalf = 0.70710678118654752440;

// Radix-R DFT in registers, natural order in and out, kernel e^{sgn 2 pi i n k / R}.
__device__ __forceinline__ void dft2(double (&r)[2], double (&i)[2]) {
    const double ar = r[0], ai = i[0];
    r[0] = ar + r[1]; i[0] = ai + i[1];
    r[1] = ar - r[1]; i[1] = ai - i[1];
}

__device__ __forceinline__ void dft4v(double& r0, double& i0, double& r1, double& i1, double& r2,
                                      double& i2, double& r3, double& i3, double sgn) {
    const double a0r = r0 + r2, a0i = i0 + i2, a1r = r0 - r2, a1i = i0 - i2;
    const double b0r = r1 + r3, b0i = i1 + i3, b1r = r1 - r3, b1i = i1 - i3;
    // W4 b1 = (i sgn) b1
    const double wr = -sgn * b1i, wi = sgn * b1r;
    r0 = a0r + b0r; i0 = a0i + b0i;
    r2 = a0r - b0r; i2 = a0i - b0i;
    r1 = a1r + wr;  i1 = a1i + wi;
    r3 = a1r - wr;  i3 = a1i - wi;
}

__device__ __forceinline__ void dft4(double (&r)[4], double (&i)[4], double sgn) {
    dft4v(r[0], i[0], r[1], i[1], r[2], i[2], r[3], i[3], sgn);
}

__device__ __forceinline__ void dft8(double (&r)[8], double (&i)[8], double sgn) {
    // X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k]
    double er0 = r[0], ei0 = i[0], er1 = r[2], ei1 = i[2], er2 = r[4], ei2 = i[4], er3 = r[6], ei3 = i[6];
    double or0 = r[1], oi0 = i[1], or1 = r[3], oi1 = i[3], or2 = r[5], oi2 = i[5], or3 = r[7], oi3 = i[7];
    dft4v(er0, ei0, er1, ei1, er2, ei2, er3, ei3, sgn);
    dft4v(or0, oi0, or1, oi1, or2, oi2, or3, oi3, sgn);
    // W8^1 = (c, sgn c), W8^2 = (0, sgn), W8^3 = (-c, sgn c), c = sqrt(1/2)
    const double c = kSqrtHalf;
    const double t1r = c * (or1 - sgn * oi1), t1i = c * (oi1 + sgn * or1);
    const double t2r = -sgn * oi2, t2i = sgn * or2;
    const double t3r = -c * (or3 + sgn * oi3), t3i = c * (sgn * or3 - oi3);
    r[0] = er0 + or0; i[0] = ei0 + oi0; r[4] = er0 - or0; i[4] = ei0 - oi0;
    r[1] = er1 + t1r; i[1] = ei1 + t1i; r[5] = er1 - t1r; i[5] = ei1 - t1i;
    r[2] = er2 + t2r; i[2] = ei2 + t2i; r[6] = er2 - t2r; i[6] = ei2 - t2i;
    r[3] = er3 + t3r; i[3] = ei3 + t3i; r[7] = er3 - t3r; i[7] = ei3 - t3i;
}

template <int R>
__device__ __forceinline__ void dft_r(double (&r)[R], double (&i)[R], double sgn) {
    if constexpr (R == 2) dft2(r, i);
    else if constexpr (R == 4) dft4(r, i, sgn);
    else dft8(r, i, sgn);
}

// One Stockham pass over all C = E / L lines of the workgroup, in place (all
// reads complete before the barrier, all writes after it).
template <int L, int R, int NS, int E>
__device__ __forceinline__ void stockham_pass(double* sre, double* sim, const double2* __restrict__ tw,
                                              double sgn) {
    constexpr int C = E / L;
    constexpr int NBT = (C * L / R) / 256;  // butterflies per thread
    static_assert(NBT >= 1, "too few butterflies per workgroup");
    double vr[NBT][R], vi[NBT][R];
#pragma unroll
    for (int b = 0; b < NBT; b++) {
        const int bid = threadIdx.x + 256 * b;
        const int c = bid / (L / R), j = bid % (L / R);
        const int kk = j % NS;
#pragma unroll
        for (int r = 0; r < R; r++) {
            vr[b][r] = sre[c * L + j + r * (L / R)];
            vi[b][r] = sim[c * L + j + r * (L / R)];
        }
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const double2 w = tw[(r * kk * (kTwN / (NS * R))) & (kTwN - 1)];
                const double wr = w.x, wi = sgn * w.y;
                const double xr = vr[b][r], xi = vi[b][r];
                vr[b][r] = xr * wr - xi * wi;
                vi[b][r] = xr * wi + xi * wr;
            }
        }
        dft_r<R>(vr[b], vi[b], sgn);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NBT; b++) {
        const int bid = threadIdx.x + 256 * b;
        const int c = bid / (L / R), j = bid % (L / R);
        const int kk = j % NS;
        const int base = (j - kk) * R + kk;
#pragma unroll
        for (int r = 0; r < R; r++) {
            sre[c * L + base + r * NS] = vr[b][r];
            sim[c * L + base + r * NS] = vi[b][r];
        }
    }
    __syncthreads();
}

// Radix plan: 8 while the remaining factor is >= 8, then 4 or 2.
template <int L, int NS, int E>
__device__ __forceinline__ void stockham_all(double* sre, double* sim, const double2* tw, double sgn) {
    if constexpr (NS < L) {
        constexpr int rem = L / NS;
        constexpr int R = rem >= 8 ? 8 : rem;
        stockham_pass<L, R, NS, E>(sre, sim, tw, sgn);
        stockham_all<L, NS * R, E>(sre, sim, tw, sgn);
    }
}

template <int L, bool COLMAJOR, class IO, int E = kFftElems>
__global__ __launch_bounds__(256) void k_fft_lines(IO io, int nlines, double sgn,
                                                   const double2* __restrict__ tw) {
    static_assert(L >= 16 && L <= E && (L & (L - 1)) == 0, "L must be a power of 2 in [16, E]");
    constexpr int C = E / L;
    __shared__ double sre[E], sim[E];
    const int line0 = blockIdx.x * C;
#pragma unroll 4
    for (int m = 0; m < E / 256; m++) {
        const int e = threadIdx.x + 256 * m;
        const int c = COLMAJOR ? e % C : e / L;
        const int k = COLMAJOR ? e / C : e % L;
        double2 v = make_double2(0.0, 0.0);
        if (line0 + c < nlines) v = io.load(line0 + c, k);
        sre[c * L + k] = v.x;
        sim[c * L + k] = v.y;
    }
    __syncthreads();
    stockham_all<L, 1, E>(sre, sim, tw, sgn);
    const double sc = 1.0 / sqrt((double)L);
#pragma unroll 4
    for (int m = 0; m < E / 256; m++) {
        const int e = threadIdx.x + 256 * m;
        const int c = COLMAJOR ? e % C : e / L;
        const int k = COLMAJOR ? e / C : e % L;
        if (line0 + c < nlines) io.store(line0 + c, k, make_double2(sre[c * L + k] * sc, sim[c * L + k] * sc));
    }
}

// Grid for k_fft_lines<L, ., ., E>.
template <int L, int E = kFftElems>
inline int fft_lines_grid(int nlines) {
    constexpr int C = E / L;
    return (nlines + C - 1) / C;
}

}  // namespace zola

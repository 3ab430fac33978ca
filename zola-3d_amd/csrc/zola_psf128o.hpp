// zola_psf128o.hpp -- the LM evaluation (PSF value and tangents) on the
// 128x128 pupil grid by eight waves (one block of 512 threads) per ROI.
//
// Same model, transforms and exact blurs as eval128 (zola_psf128.hpp;
// process/cpu/PSFPhaseMany.java:553-622).  Four waves per ROI hold one ROI per
// CU and one wave per SIMD (the two 66 KB stage-1 buffers fill the LDS), so
// nothing hides the LDS and barrier latency between the transform stages.
// Here every stage is split over twice the waves, with half the registers per
// wave: two waves per SIMD on the same LDS.
//  * Stage 1, wave (H, c) with H = wr & 3, c = wr >> 2: output rows p = 4m + H
//    of pupil columns [32c, 32c + 32).  Lane (v, s) evaluates column 32c + v at
//    the 32 signed frequencies of half s (the twiddle exp(2 pi i nu H / 128) is
//    folded into the phase, as eval128 does for its two parities), forms P and
//    i kz P, and the two lanes of a column exchange halves with
//    permlane32_swap: lane (v, 0) keeps the E input of all 64 frequencies,
//    lane (v, 1) the dE/dz input.  Frequencies nu and nu - 32 meet the same
//    exp(2 pi i nu m / 32), so each lane adds its own half and the received
//    half (in either order) and runs one 32-point IDFT.
//  * Stage 2, wave (rh, Q) with rh = wr & 1, Q = wr >> 1: grid row
//    p = lane + 64 rh, output columns 4n + Q, for E and dE/dz; I and Iz to LDS.
//  * Row blur at the crop columns, wave (rh, cq): row p, the cq-th quarter of
//    the R crop columns.
//  * Column blur: lane = ROI column j, wave wr owns ROI rows
//    [wr R8, wr R8 + R8) (R8 = ceil(R / 8)); this is the fits' pixel layout.
#pragma once
#include "zola_psf128.hpp"
#include "zola_psf64p.hpp"

namespace zola {

template <int R>
struct Roi128o {
    static constexpr int R8 = (R + 7) / 8;   // ROI rows per wave
    static constexpr int JQ = (R + 3) / 4;   // crop columns per row-blur quarter
    static constexpr int S2 = Roi128<R>::S2;
    // rows [I0, I0 + NI) of wave W, columns [J0, J0 + NJ) of quarter C
    template <int W>
    static constexpr int I0 = W * R8 < R ? W * R8 : R;
    template <int W>
    static constexpr int NI = (W * R8 + R8 <= R ? R8 : R - I0<W>);
    template <int C>
    static constexpr int J0 = C * JQ < R ? C * JQ : R;
    template <int C>
    static constexpr int NJ = (C * JQ + JQ <= R ? JQ : R - J0<C>);
};

// Stage 1 of both fields: wave parity H (0..3), column half c (0..1).
__device__ __forceinline__ void stage1_oct(const Model128& M, Buf128 (&S)[2], int lane, int H, int c, float X,
                                           float Y, float z, float zoil) {
    const int v = lane & 31, s = lane >> 5;
    const bool up = s != 0;
    const int vs = 32 * c + v;
    const float nv = (float)(vs - 32);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    const float Xh = X + (float)H * (float)(2.0 * kPi / 128.0);
    float re[32], im[32];
    // groups of 8 samples keep the scheduler from issuing all table loads
    // ahead of the sincos work
    static_for<0, 4>([&](auto gc) {
        constexpr int g = decltype(gc)::value;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            const int k = g * 8 + kk;
            const int nu = k - 32 * s;   // half 0: 0..31, half 1: -32..-1
            const float4 t = M.tab[(nu + 32) * kW2 + vs];
            const float q = t.x + Xh * (float)nu + cY + t.z * z - t.w * zoil;
            float sn, cs;
            sincos_reduced(q, sn, cs);
            const float pr = t.y * cs, pi = t.y * sn;
            const float zr = -t.z * pi, zi = t.z * pr;   // i kz P
            // keep this lane's field, add the partner's half of it (fold mod 32)
            re[k] = (up ? zr : pr) + other_half(up ? pr : zr, up);
            im[k] = (up ? zi : pi) + other_half(up ? pi : zi, up);
        }
        __builtin_amdgcn_sched_barrier(0);
    });
    fft_dif<32, +1>(re, im);
    Buf128& D = S[s];
#pragma unroll
    for (int m = 0; m < 32; m++) {
        const int p = 4 * m + H;
        D.re[p * kS2 + vs] = re[fidx<32>(m)];
        D.im[p * kS2 + vs] = im[fidx<32>(m)];
    }
}

// Stage 2, quarter Q of row p for E and dE/dz: I and Iz at columns 4n + Q.
// |E|^2 is formed before the dE/dz quarter so that at most E, I and dE/dz are live.
template <int Q>
__device__ __forceinline__ void stage2_oct(const Buf128 (&S)[2], int p, float (&iq)[32], float (&izq)[32]) {
    float er[32], ei[32];
    stage2_quarter<Q>(S[0], p, er, ei);
    static_for<0, 32>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        constexpr int k = fidx<32>(n);
        iq[n] = er[k] * er[k] + ei[k] * ei[k];
    });
    __builtin_amdgcn_sched_barrier(0);
    float zr[32], zi[32];
    stage2_quarter<Q>(S[1], p, zr, zi);
    static_for<0, 32>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        constexpr int k = fidx<32>(n);
        izq[n] = 2.f * (er[k] * zr[k] + ei[k] * zi[k]);
    });
}

// Row blur of crop-column quarter C at row p; stores the three channels of
// those columns as Rb[(ch R + j) kI2 + p] after the caller's barrier.
template <int R, int C>
__device__ __forceinline__ void row_blur_oct(const float* Ib, const float* Izb, int p, const float (&h1)[2],
                                             const float (&hd)[2], const float* __restrict__ kwin,
                                             float (&acc)[3][Roi128o<R>::JQ]) {
    constexpr int J0 = Roi128o<R>::template J0<C>, NJ = Roi128o<R>::template NJ<C>;
    if constexpr (NJ > 0) {
        float a[3][NJ];
        row_blur_full<R, J0, NJ, 3>(Ib, Izb, p, h1, hd, kwin, a);
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
#pragma unroll
            for (int jj = 0; jj < NJ; jj++) acc[ch][jj] = a[ch][jj];
    }
}

template <int R, int C>
__device__ __forceinline__ void store_rows_oct(float* Rb, int p, const float (&acc)[3][Roi128o<R>::JQ]) {
    constexpr int J0 = Roi128o<R>::template J0<C>, NJ = Roi128o<R>::template NJ<C>;
#pragma unroll
    for (int ch = 0; ch < 3; ch++)
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) Rb[(ch * R + J0 + jj) * kI2 + p] = acc[ch][jj];
}

template <int R, int W>
__device__ __forceinline__ void column_blur_oct(const float* Rb, int j, const float (&h1)[2], const float (&hd)[2],
                                                const float* __restrict__ kwin, float (&vB)[Roi128o<R>::R8],
                                                float (&vX)[Roi128o<R>::R8], float (&vY)[Roi128o<R>::R8],
                                                float (&vZ)[Roi128o<R>::R8]) {
    constexpr int I0 = Roi128o<R>::template I0<W>, NI = Roi128o<R>::template NI<W>;
    if constexpr (NI > 0) {
        float o0[NI], o1[NI], o2[NI], o3[NI];
        column_blur_full<R, I0, NI, 4>(Rb, j, h1, hd, kwin, o0, o1, o2, o3);
#pragma unroll
        for (int t = 0; t < NI; t++) {
            vB[t] = o0[t];
            vX[t] = o1[t];
            vY[t] = o2[t];
            vZ[t] = o3[t];
        }
    }
}

// One LM evaluation of a ROI by the eight waves of a block (every wave calls
// it; it contains block barriers).  On return lane j < R of wave wr holds, for
// ROI rows wr R8 + t (t < R8, rows below R), the raw blurred value (signed) in
// vB and the x, y, z tangents in vX, vY, vZ, as eval128 does for its rows.
template <int R>
__device__ __forceinline__ void eval128o(const Model128& M, Buf128 (&S)[2], int lane, int wr, float x, float y,
                                         float z, float zoil, float (&vB)[Roi128o<R>::R8],
                                         float (&vX)[Roi128o<R>::R8], float (&vY)[Roi128o<R>::R8],
                                         float (&vZ)[Roi128o<R>::R8]) {
    constexpr int R8 = Roi128o<R>::R8, JQ = Roi128o<R>::JQ;
    static_assert(3 * R * kI2 <= 2 * kN2 * kS2, "blurred rows must fit one buffer");
    const float X = M.kappa * x, Y = M.kappa * y;
    stage1_oct(M, S, lane, wr & 3, wr >> 2, X, Y, z, zoil);
    __syncthreads();
    const int p = lane + 64 * (wr & 1);
    const int qd = wr >> 1;   // stage-2 quarter, then row-blur column quarter
    {
        float iq[32], izq[32];
        switch (qd) {
            case 0: stage2_oct<0>(S, p, iq, izq); break;
            case 1: stage2_oct<1>(S, p, iq, izq); break;
            case 2: stage2_oct<2>(S, p, iq, izq); break;
            default: stage2_oct<3>(S, p, iq, izq); break;
        }
        __syncthreads();  // all stage-1 reads are done
        float* Ib = S[0].re;   // real [128][kI2]
        float* Izb = S[1].re;
#pragma unroll
        for (int n = 0; n < 32; n++) {
            Ib[p * kI2 + 4 * n + qd] = iq[n];
            Izb[p * kI2 + 4 * n + qd] = izq[n];
        }
    }
    __syncthreads();
    float h1[2], hd[2];   // loaded here: not live through the two transforms
    load_taps128(M, lane, h1, hd);
    float acc[3][JQ];
    {
        const float* Ib = S[0].re;
        const float* Izb = S[1].re;
        switch (qd) {
            case 0: row_blur_oct<R, 0>(Ib, Izb, p, h1, hd, M.kwin, acc); break;
            case 1: row_blur_oct<R, 1>(Ib, Izb, p, h1, hd, M.kwin, acc); break;
            case 2: row_blur_oct<R, 2>(Ib, Izb, p, h1, hd, M.kwin, acc); break;
            default: row_blur_oct<R, 3>(Ib, Izb, p, h1, hd, M.kwin, acc); break;
        }
    }
    __syncthreads();  // all intensity reads are done
    float* Rb = S[0].re;  // [channel][column][row]
    switch (qd) {
        case 0: store_rows_oct<R, 0>(Rb, p, acc); break;
        case 1: store_rows_oct<R, 1>(Rb, p, acc); break;
        case 2: store_rows_oct<R, 2>(Rb, p, acc); break;
        default: store_rows_oct<R, 3>(Rb, p, acc); break;
    }
    __syncthreads();
    const int j = lane < R ? lane : 0;
#pragma unroll
    for (int t = 0; t < R8; t++) vB[t] = vX[t] = vY[t] = vZ[t] = 0.f;
    switch (wr) {
        case 0: column_blur_oct<R, 0>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 1: column_blur_oct<R, 1>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 2: column_blur_oct<R, 2>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 3: column_blur_oct<R, 3>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 4: column_blur_oct<R, 4>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 5: column_blur_oct<R, 5>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 6: column_blur_oct<R, 6>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        default: column_blur_oct<R, 7>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
    }
    // the tangents are read only on accepted trials: keep the compiler from
    // sinking their arithmetic past the reductions that follow
#pragma unroll
    for (int t = 0; t < R8; t++) __asm__ __volatile__("" : "+v"(vB[t]), "+v"(vX[t]), "+v"(vY[t]), "+v"(vZ[t]));
    __syncthreads();  // the buffers are reused by the next evaluation
}

}  // namespace zola

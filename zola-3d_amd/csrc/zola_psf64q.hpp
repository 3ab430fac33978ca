// zola_psf64q.hpp -- PSF synthesis on the 64x64 pupil grid by four waves (one
// block) per PSF, for the larger ROI sides (gfx950).
//
// Same model and stage-1 code as zola_psf64.hpp (process/cpu/PSFPhaseMany.java:
// 553-622); the work after stage 1 is split over four waves so that no lane
// holds more than ceil(R/4) ROI rows:
//  * Stage 1: wave 0 transforms E, wave 1 dE/dz (stage1_64: lane = (window
//    column, output-row parity)); waves 2 and 3 wait at the barrier.
//  * Stage 2: wave wr takes the output columns b = 4k + wr (k < 16) of every
//    grid row p = lane: the 32 signed-frequency inputs n and n - 16 fold to 16
//    after the twiddle exp(2 pi i n wr / 64), then a 16-point IDFT.  I and
//    2 Re(conj(E) dE/dz) over the whole 64 x 64 grid go to LDS (over the stage-1
//    buffers, after a barrier).
//  * Row blur (exact 64-periodic kernel) at the R crop columns: lane = grid row
//    p, wave wr the columns [wr JQ, wr JQ + JQ) (JQ = ceil(R/4)); the taps are
//    broadcast from one VGPR with readlane, grouped by tap.
//  * Column blur of the crop columns: lane = ROI column j, wave wr owns ROI rows
//    [wr RQ, wr RQ + RQ) (RQ = ceil(R/4)), the pixel layout of the fits.
#pragma once
#include "zola_psf64.hpp"

namespace zola {

constexpr int kIq = 65;   // row stride of the real intensity / blur buffers

template <int R>
struct Roi64q {
    static constexpr int RQ = (R + 3) / 4;   // ROI rows (and crop columns) per wave
    static_assert(R - 3 * RQ >= 1 && R <= 43, "wave 3 needs a crop column; three blurred channels must fit");
};

// Stage 2, output columns b = 4k + Q of row p: E(p, 4k + Q) at fidx<16>(k).
template <int Q>
__device__ __forceinline__ void stage2_q64(const WaveScratch64& S, int p, float (&re)[16], float (&im)[16]) {
    __asm__ __volatile__("" ::: "memory");
    static_for<0, 16>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        // signed frequency n = s (column s + 16) and n = s - 16 (column s)
        float ar = S.re[p * kS1 + s + 16], ai = S.im[p * kS1 + s + 16];
        float br = S.re[p * kS1 + s], bi = S.im[p * kS1 + s];
        twiddle_mul<64, (s * Q) % 64, +1>(ar, ai);
        twiddle_mul<64, ((s - 16) * Q + 256) % 64, +1>(br, bi);
        re[s] = ar + br;
        im[s] = ai + bi;
    });
    fft_dif<16, +1>(re, im);
}

// Row blur at crop columns J0 .. J0+NJ-1 of grid row p (ROI column j is grid
// column j - R/2):  acc[c][jj] = sum_q k_c((J0 + jj - R/2 - q) mod 64) img_c(p, q)
// NCH = 1: (h1, I);  NCH = 3: (h1, I), (hd, I), (h1, Iz).
template <int R, int J0, int NJ, int NCH>
__device__ __forceinline__ void row_blur64q(const float* Ib, const float* Izb, int p, float h1reg, float hdreg,
                                            float (&acc)[NCH][NJ]) {
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) acc[c][jj] = 0.f;
    static_for<0, 8>([&](auto qcc) {
        constexpr int q0 = decltype(qcc)::value * 8;
        float iv[8], izv[NCH == 3 ? 8 : 1];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            iv[t] = Ib[p * kIq + q0 + t];
            if constexpr (NCH == 3) izv[t] = Izb[p * kIq + q0 + t];
        }
        constexpr int dmin = J0 - R / 2 - q0 - 7;
        static_for<0, NJ + 7>([&](auto ddc) {
            constexpr int d = dmin + decltype(ddc)::value;
            constexpr int dm = ((d % 64) + 64) % 64;
            const float c1 = bcast(h1reg, dm);
            float cd = 0.f;
            if constexpr (NCH == 3) cd = bcast(hdreg, dm);
            static_for<0, 8>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr int jj = d + R / 2 + q0 + t - J0;
                if constexpr (jj >= 0 && jj < NJ) {
                    acc[0][jj] += c1 * iv[t];
                    if constexpr (NCH == 3) {
                        acc[1][jj] += cd * iv[t];
                        acc[2][jj] += c1 * izv[t];
                    }
                }
            });
            __builtin_amdgcn_sched_barrier(0);
        });
    });
}

// Column blur of ROI column j at ROI rows I0 .. I0+NI-1 (ROI row i is grid row
// i - R/2):  out[o][ii] = sum_p k_o((I0 + ii - R/2 - p) mod 64) Rb_src(o)(p, j),
// Rb = [channel][column][row] with row stride kIq.  NO = 1: (h1, Rb0);
// NO = 4: (h1, Rb0), (hd, Rb0), (h1, Rb1), (h1, Rb2).
template <int R, int I0, int NI, int NO>
__device__ __forceinline__ void column_blur64q(const float* Rb, int j, float h1reg, float hdreg, float (&o0)[NI],
                                               float (&o1)[NI], float (&o2)[NI], float (&o3)[NI]) {
#pragma unroll
    for (int ii = 0; ii < NI; ii++) {
        o0[ii] = 0.f;
        if constexpr (NO == 4) o1[ii] = o2[ii] = o3[ii] = 0.f;
    }
    static_for<0, 8>([&](auto pcc) {
        constexpr int p0 = decltype(pcc)::value * 8;
        float r0[8], r1[NO == 4 ? 8 : 1], r2[NO == 4 ? 8 : 1];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            r0[t] = Rb[(0 * R + j) * kIq + p0 + t];
            if constexpr (NO == 4) {
                r1[t] = Rb[(1 * R + j) * kIq + p0 + t];
                r2[t] = Rb[(2 * R + j) * kIq + p0 + t];
            }
        }
        constexpr int dmin = I0 - R / 2 - p0 - 7;
        static_for<0, NI + 7>([&](auto ddc) {
            constexpr int d = dmin + decltype(ddc)::value;
            constexpr int dm = ((d % 64) + 64) % 64;
            const float c1 = bcast(h1reg, dm);
            float cd = 0.f;
            if constexpr (NO == 4) cd = bcast(hdreg, dm);
            static_for<0, 8>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr int ii = d + R / 2 + p0 + t - I0;
                if constexpr (ii >= 0 && ii < NI) {
                    o0[ii] += c1 * r0[t];
                    if constexpr (NO == 4) {
                        o1[ii] += cd * r0[t];
                        o2[ii] += c1 * r1[t];
                        o3[ii] += c1 * r2[t];
                    }
                }
            });
            __builtin_amdgcn_sched_barrier(0);
        });
    });
}

// Evaluation of one PSF by the four waves of a block (every wave calls it; it
// contains block barriers).  scr: TANG ? two : one WaveScratch64, contiguous.
// On return lane j < R of wave wr holds, for ROI rows wr*RQ + t (t < RQ), the
// raw blurred value (signed) in vB and, when TANG, the x, y, z tangents.
template <int R, bool TANG>
__device__ __forceinline__ void eval64q(const Model64& M, const BlockTables64& T, float* scr, int lane, int wr,
                                        float x, float y, float z, float zoil, float (&vB)[Roi64q<R>::RQ],
                                        float (&vX)[Roi64q<R>::RQ], float (&vY)[Roi64q<R>::RQ],
                                        float (&vZ)[Roi64q<R>::RQ]) {
    constexpr int RQ = Roi64q<R>::RQ;
    constexpr int NCH = TANG ? 3 : 1;
    constexpr int kBuf = (int)(sizeof(WaveScratch64) / sizeof(float));
    static_assert(64 * kIq <= kBuf, "the intensity must fit one stage-1 buffer");
    static_assert(NCH * R * kIq <= (TANG ? 2 : 1) * kBuf, "the blurred rows must fit the buffers");
    WaveScratch64& SE = *reinterpret_cast<WaveScratch64*>(scr);
    WaveScratch64& SZ = *reinterpret_cast<WaveScratch64*>(scr + kBuf);
    const float h1reg = T.h3[lane], hdreg = T.hd3[lane];
    const float X = M.kappa * x, Y = M.kappa * y;
    // ---- stage 1: wave 0 E, wave 1 dE/dz
    if (wr == 0) stage1_64(T, SE, lane, 0, X, Y, z, zoil, M.kappa);
    if (TANG && wr == 1) stage1_64(T, SZ, lane, 3, X, Y, z, zoil, M.kappa);
    __syncthreads();
    // ---- stage 2: columns 4k + wr of row p = lane
    const int p = lane;
    float iq[16], izq[TANG ? 16 : 1];
    auto quarter = [&](auto qc) {
        constexpr int Q = decltype(qc)::value;
        float er[16], ei[16];
        stage2_q64<Q>(SE, p, er, ei);
        if constexpr (TANG) {
            float zr[16], zi[16];
            stage2_q64<Q>(SZ, p, zr, zi);
            static_for<0, 16>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int f = fidx<16>(k);
                iq[k] = er[f] * er[f] + ei[f] * ei[f];
                izq[k] = 2.f * (er[f] * zr[f] + ei[f] * zi[f]);
            });
        } else {
            static_for<0, 16>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int f = fidx<16>(k);
                iq[k] = er[f] * er[f] + ei[f] * ei[f];
            });
        }
    };
    switch (wr) {
        case 0: quarter(std::integral_constant<int, 0>{}); break;
        case 1: quarter(std::integral_constant<int, 1>{}); break;
        case 2: quarter(std::integral_constant<int, 2>{}); break;
        default: quarter(std::integral_constant<int, 3>{}); break;
    }
    __syncthreads();  // all stage-1 reads are done
    float* Ib = scr;          // real [64][kIq] over the E buffer
    float* Izb = scr + kBuf;  // over the dE/dz buffer
#pragma unroll
    for (int k = 0; k < 16; k++) {
        Ib[p * kIq + 4 * k + wr] = iq[k];
        if constexpr (TANG) Izb[p * kIq + 4 * k + wr] = izq[k];
    }
    __syncthreads();
    // ---- row blur at the crop columns [wr*RQ, min(R, wr*RQ + RQ))
    constexpr int N3 = R - 3 * RQ;   // columns of wave 3 (>= 1)
    float ra[NCH][RQ], rl[NCH][N3];
    switch (wr) {
        case 0: row_blur64q<R, 0 * RQ, RQ, NCH>(Ib, Izb, p, h1reg, hdreg, ra); break;
        case 1: row_blur64q<R, 1 * RQ, RQ, NCH>(Ib, Izb, p, h1reg, hdreg, ra); break;
        case 2: row_blur64q<R, 2 * RQ, RQ, NCH>(Ib, Izb, p, h1reg, hdreg, ra); break;
        default: row_blur64q<R, 3 * RQ, N3, NCH>(Ib, Izb, p, h1reg, hdreg, rl); break;
    }
    __syncthreads();  // all intensity reads are done
    float* Rb = scr;  // [channel][column][row]
    if (wr < 3) {
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int jj = 0; jj < RQ; jj++) Rb[(c * R + wr * RQ + jj) * kIq + p] = ra[c][jj];
    } else {
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int jj = 0; jj < N3; jj++) Rb[(c * R + 3 * RQ + jj) * kIq + p] = rl[c][jj];
    }
    __syncthreads();
    // ---- column blur: lane j, ROI rows [wr*RQ, wr*RQ + RQ)
    const int j = lane < R ? lane : 0;
    constexpr int NO = TANG ? 4 : 1;
    switch (wr) {
        case 0: column_blur64q<R, 0 * RQ, RQ, NO>(Rb, j, h1reg, hdreg, vB, vX, vY, vZ); break;
        case 1: column_blur64q<R, 1 * RQ, RQ, NO>(Rb, j, h1reg, hdreg, vB, vX, vY, vZ); break;
        case 2: column_blur64q<R, 2 * RQ, RQ, NO>(Rb, j, h1reg, hdreg, vB, vX, vY, vZ); break;
        default: column_blur64q<R, 3 * RQ, RQ, NO>(Rb, j, h1reg, hdreg, vB, vX, vY, vZ); break;
    }
    __syncthreads();  // the buffers are reused by the next evaluation
}

}  // namespace zola

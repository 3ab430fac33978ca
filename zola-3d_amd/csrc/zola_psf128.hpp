// zola_psf128.hpp -- PSF synthesis on a 128x128 pupil grid (gfx950), ROI side
// R in [9, 44], four waves (one block) per PSF.  The blur grid is
// s2 = pow2(min(2R, 128)) (process/cpu/PSFPhaseMany.java:102-108): 32 for R <= 16,
// 64 for R <= 32, else 128.
//
// Same model as zola_psf64.hpp (process/cpu/PSFPhaseMany.java:553-622):
// E = IDFT_128x128 of the fftshifted pupil, I = |E|^2, Gaussian blur (the exact
// s2-periodic kernel h1 = IDFT(G1): the Nyquist cut of the sampled Gaussian
// leaves tails of 1e-4 of the centre, so no truncation), |.|, R x R crop,
// normalised.
//  * The pupil radius is < N/4 = 32 px: its support lies in a 64x64 window of
//    signed frequencies n in [-32, 31].
//  * Stage 1 (along the first axis): wave (field, H) with field E or dE/dz and
//    output-row parity H; lane = pupil column.  A 128-point IDFT of a 64-point
//    support, outputs p = 2m + H, is a 64-point IDFT of the input twiddled by
//    exp(2 pi i u H / 128) (folded into the phase before sincos).  Results go
//    to LDS, all 128 rows x 64 columns per field (split re/im, stride 65).
//  * Stage 2 (along the second axis): lane = grid row p, outputs q = 4n + Q in
//    quarters Q (each lane two quarters): the 64 inputs folded to 32 after the
//    twiddle exp(2 pi i v Q / 128), then a 32-point IDFT.  I and
//    2 Re(conj(E) dE/dz) over the whole grid go back to LDS.
//  * For s2 < 128 the reference crops |E|^2 to the s2 x s2 window of offsets
//    [-s2/2, s2/2) around the centre (PSFPhaseMany.java:575-582) and blurs it
//    s2-periodically; the blurs below read only that window, with s2 taps.
//  * Row blur (exact, s2 taps) only at the R crop columns: lane = row p,
//    waves split the columns; the taps are broadcast from two VGPRs with
//    readlane, grouped by tap so each coefficient is read once per chunk.
//  * Column blur (exact) of the R crop columns: lane = ROI column j, wave w
//    owns ROI rows [w RQ, w RQ + RQ) (RQ = ceil(R / 4)), so the tap index is a
//    compile-time constant again.  This is also the pixel layout of the fits.
//  * Pupil tables live in global memory as float4 {phi, amp, kz', kzo'}
//    (64 KiB, L2-resident), one coalesced 16-byte load per pupil sample.
#pragma once
#include "zola_psf64.hpp"

namespace zola {

constexpr int kN2 = 128;   // pupil / E grid
constexpr int kW2 = 64;    // support window
constexpr int kS2 = 65;    // row stride of the complex stage-1 buffers
constexpr int kI2 = 129;   // row stride of the real intensity / blur buffers

struct Model128 {
    const float4* tab;   // 64 x 64 window [wu][wv]: {phi, amp, kz - kz0, kzoil - kzoil0}
    const float* h1;     // s2 blur taps h1[d], d = 0..s2-1 (periodic)
    const float* hd;     // s2 taps of the spectral-derivative blur (includes 1/dx), s2 = 128
    // s2 < 128: R x 128, the x (or y) tangent of the blurred crop at crop index c
    // per unit intensity at grid index g: sum over the window offsets o of
    // h1_s2(c - R/2 - o) dN(o - g), dN the 128-periodic spectral derivative (1/dx)
    const float* kwin;
    float kappa;         // 2 pi / (dx N): kx = kappa (n_u + 1/2)
};

// One stage-1 buffer: 128 rows x 64 columns, complex; reused as a real
// 128 x 129 array (intensity) or as blurred rows [channel][column][row].
struct alignas(16) Buf128 {
    float re[kN2 * kS2];
    float im[kN2 * kS2];
};

struct NoShared {};

__host__ __device__ constexpr int pow2_at_least(int x) { return x <= 1 ? 1 : 2 * pow2_at_least((x + 1) / 2); }

template <int R>
struct Roi128 {
    static constexpr int RQ = (R + 3) / 4;   // ROI rows per wave
    static constexpr int S2 = pow2_at_least(2 * R < 128 ? 2 * R : 128);   // blur grid
    static_assert(R >= 9 && R <= 4 * 11, "the 128 grid covers ROI sides 9..44");
    static_assert(S2 == 32 || S2 == 64 || S2 == 128, "blur grid");
};

// First offset of chunk k (8 grid columns or rows) of the blur window, and its
// grid index: the window is [-S2/2, S2/2) around the centre (offset o at index
// o mod 128); for S2 = 128 the chunks simply run over the grid.  With S2/2 a
// multiple of 8 no chunk wraps.
template <int S2, int K>
struct WinChunk {
    static constexpr int o0 = S2 == 128 ? 8 * K : 8 * K - S2 / 2;
    static constexpr int g0 = (o0 + 128) % 128;
};

__device__ __forceinline__ constexpr int slot64_n(int j) { return j < 32 ? j : j - 64; }

// Tap d (0..127) of a kernel held as two VGPRs (lane l holds taps l and l+64).
template <int D>
__device__ __forceinline__ float tap(const float (&k)[2]) {
    return bcast(k[D >> 6], D & 63);
}

// Stage 1 for one field (0 = E, 1 = dE/dz = IDFT(i kz P)) and output-row parity H:
//   S(p, v + 32) = sum_u P(u, v) exp(2 pi i u p / 128),  p = 2m + H, m in [0, 64).
__device__ __forceinline__ void stage1_full(const Model128& M, Buf128& S, int lane, int field, int H, float X,
                                            float Y, float z, float zoil) {
    const int vs = lane;
    const float nv = (float)(vs - 32);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    const float Xh = X + (float)H * (float)(2.0 * kPi / 128.0);
    const float fz = field == 1 ? 1.f : 0.f, c0 = field == 1 ? 0.f : 1.f;
    float re[64], im[64];
    // groups of 8 samples: keeps the scheduler from issuing all 64 table loads
    // (256 VGPRs) ahead of the sincos work
    static_for<0, 8>([&](auto gc) {
        constexpr int g = decltype(gc)::value;
#pragma unroll
        for (int jj = 0; jj < 8; jj++) {
            const int j = g * 8 + jj;
            const int nu = slot64_n(j);
            const float4 t = M.tab[(nu + 32) * kW2 + vs];
            const float q = t.x + Xh * (float)nu + cY + t.z * z - t.w * zoil;
            float s, c;
            sincos_reduced(q, s, c);
            const float ar = t.y * c, ai = t.y * s, f = fz * t.z;
            re[j] = c0 * ar - f * ai;
            im[j] = c0 * ai + f * ar;
        }
        __builtin_amdgcn_sched_barrier(0);
    });
    fft_dif<64, +1>(re, im);
    static_for<0, 64>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const int p = 2 * m + H;
        S.re[p * kS2 + vs] = re[fidx<64>(m)];
        S.im[p * kS2 + vs] = im[fidx<64>(m)];
    });
}

// Stage 2, quarter Q of row p: E(p, 4n + Q), n in [0, 32), at fidx<32>(n).
template <int Q>
__device__ __forceinline__ void stage2_quarter(const Buf128& S, int p, float (&re)[32], float (&im)[32]) {
    __asm__ __volatile__("" ::: "memory");
    static_for<0, 32>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        // v = k (column k + 32) and v = k - 32 (column k), both = k mod 32
        float ar = S.re[p * kS2 + k + 32], ai = S.im[p * kS2 + k + 32];
        float br = S.re[p * kS2 + k], bi = S.im[p * kS2 + k];
        twiddle_mul<128, (k * Q) % 128, +1>(ar, ai);
        twiddle_mul<128, ((k - 32) * Q + 512) % 128, +1>(br, bi);
        re[k] = ar + br;
        im[k] = ai + bi;
    });
    fft_dif<32, +1>(re, im);
}

// Row blur at crop columns J0 .. J0+NJ-1 of row p (grid column j - R/2):
//   acc[c][jj] = sum_q k_c((J0 + jj - R/2 - q) mod 128) img_c(p, q)
// NCH = 1: (h1, I);  NCH = 3: (h1, I), (hd, I), (h1, Iz).
template <int R, int J0, int NJ, int NCH>
__device__ __forceinline__ void row_blur_full(const float* Ib, const float* Izb, int p, const float (&h1)[2],
                                              const float (&hd)[2], const float* __restrict__ kwin,
                                              float (&acc)[NCH][NJ]) {
    constexpr int S2 = Roi128<R>::S2;
    // the derivative channel goes through hd when the window is the whole grid
    constexpr bool HD = NCH == 3 && S2 == 128;
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int jj = 0; jj < NJ; jj++) acc[c][jj] = 0.f;
    static_for<0, S2 / 8>([&](auto qcc) {
        constexpr int q0 = WinChunk<S2, decltype(qcc)::value>::o0;   // column offset
        constexpr int g0 = WinChunk<S2, decltype(qcc)::value>::g0;   // its grid column
        float iv[8], izv[NCH == 3 ? 8 : 1];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            iv[t] = Ib[p * kI2 + g0 + t];
            if constexpr (NCH == 3) izv[t] = Izb[p * kI2 + g0 + t];
        }
        constexpr int dmin = J0 - R / 2 - q0 - 7;
        static_for<0, NJ + 7>([&](auto ddc) {
            constexpr int d = dmin + decltype(ddc)::value;
            constexpr int dm = ((d % S2) + S2) % S2;
            const float c1 = tap<dm>(h1);
            float cd = 0.f;
            if constexpr (HD) cd = tap<dm>(hd);
            static_for<0, 8>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr int jj = d + R / 2 + q0 + t - J0;
                if constexpr (jj >= 0 && jj < NJ) {
                    acc[0][jj] += c1 * iv[t];
                    if constexpr (NCH == 3) {
                        if constexpr (HD) acc[1][jj] += cd * iv[t];
                        acc[2][jj] += c1 * izv[t];
                    }
                }
            });
            __builtin_amdgcn_sched_barrier(0);
        });
    });
    if constexpr (NCH == 3 && !HD) {
        // derivative of the crop: every grid column of the row through kwin
        static_for<0, 16>([&](auto qcc) {
            constexpr int g0 = decltype(qcc)::value * 8;
            float iv[8];
#pragma unroll
            for (int t = 0; t < 8; t++) iv[t] = Ib[p * kI2 + g0 + t];
#pragma unroll
            for (int jj = 0; jj < NJ; jj++)
#pragma unroll
                for (int t = 0; t < 8; t++) acc[1][jj] += kwin[(J0 + jj) * kN2 + g0 + t] * iv[t];
            __builtin_amdgcn_sched_barrier(0);
        });
    }
}

// Column blur of ROI column j at ROI rows I0 .. I0+NI-1:
//   out[o][ii] = sum_p k_o((I0 + ii - R/2 - p) mod 128) Rb_src(o)(p, j)
// Rb is [channel][column][row] with row stride kI2.  NO = 1: (h1, Rb0);
// NO = 4: (h1, Rb0), (hd, Rb0), (h1, Rb1), (h1, Rb2).
template <int R, int I0, int NI, int NO>
__device__ __forceinline__ void column_blur_full(const float* Rb, int j, const float (&h1)[2], const float (&hd)[2],
                                                 const float* __restrict__ kwin, float (&o0)[NI], float (&o1)[NI],
                                                 float (&o2)[NI], float (&o3)[NI]) {
    constexpr int S2 = Roi128<R>::S2;
    constexpr bool HD = NO == 4 && S2 == 128;
#pragma unroll
    for (int ii = 0; ii < NI; ii++) {
        o0[ii] = 0.f;
        if constexpr (NO == 4) o1[ii] = o2[ii] = o3[ii] = 0.f;
    }
    static_for<0, S2 / 8>([&](auto pcc) {
        constexpr int p0 = WinChunk<S2, decltype(pcc)::value>::o0;   // row offset
        constexpr int g0 = WinChunk<S2, decltype(pcc)::value>::g0;   // its grid row
        float r0[8], r1[NO == 4 ? 8 : 1], r2[NO == 4 ? 8 : 1];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            r0[t] = Rb[(0 * R + j) * kI2 + g0 + t];
            if constexpr (NO == 4) {
                r1[t] = Rb[(1 * R + j) * kI2 + g0 + t];
                r2[t] = Rb[(2 * R + j) * kI2 + g0 + t];
            }
        }
        constexpr int dmin = I0 - R / 2 - p0 - 7;
        static_for<0, NI + 7>([&](auto ddc) {
            constexpr int d = dmin + decltype(ddc)::value;
            constexpr int dm = ((d % S2) + S2) % S2;
            const float c1 = tap<dm>(h1);
            float cd = 0.f;
            if constexpr (HD) cd = tap<dm>(hd);
            static_for<0, 8>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr int ii = d + R / 2 + p0 + t - I0;
                if constexpr (ii >= 0 && ii < NI) {
                    o0[ii] += c1 * r0[t];
                    if constexpr (NO == 4) {
                        if constexpr (HD) o1[ii] += cd * r0[t];
                        o2[ii] += c1 * r1[t];
                        o3[ii] += c1 * r2[t];
                    }
                }
            });
            __builtin_amdgcn_sched_barrier(0);
        });
    });
    if constexpr (NO == 4 && !HD) {
        // x tangent: every grid row of the row-blurred value through kwin
        static_for<0, 16>([&](auto pcc) {
            constexpr int g0 = decltype(pcc)::value * 8;
            float r0[8];
#pragma unroll
            for (int t = 0; t < 8; t++) r0[t] = Rb[(0 * R + j) * kI2 + g0 + t];
#pragma unroll
            for (int ii = 0; ii < NI; ii++)
#pragma unroll
                for (int t = 0; t < 8; t++) o1[ii] += kwin[(I0 + ii) * kN2 + g0 + t] * r0[t];
            __builtin_amdgcn_sched_barrier(0);
        });
    }
}

__device__ __forceinline__ void load_taps128(const Model128& M, int lane, float (&h1)[2], float (&hd)[2]) {
    h1[0] = M.h1[lane];
    h1[1] = M.h1[lane + 64];
    hd[0] = M.hd[lane];
    hd[1] = M.hd[lane + 64];
}

// Evaluation of one PSF by the four waves of a block (wr = wave 0..3; every
// wave calls it, it contains block barriers).  S[0], S[1]: the PSF's buffers.
// On return lane j < R of wave wr holds, for ROI rows wr*RQ + t (t < RQ), the
// raw blurred value (signed) in vB and, when TANG, the x, y, z tangents.
template <int R, bool TANG>
__device__ __forceinline__ void eval128(const Model128& M, Buf128 (&S)[2], int lane, int wr, float x, float y,
                                        float z, float zoil, float (&vB)[Roi128<R>::RQ], float (&vX)[Roi128<R>::RQ],
                                        float (&vY)[Roi128<R>::RQ], float (&vZ)[Roi128<R>::RQ]) {
    constexpr int RQ = Roi128<R>::RQ;
    constexpr int NCH = TANG ? 3 : 1;
    static_assert(3 * R * kI2 <= 2 * kN2 * kS2, "blurred rows must fit one buffer");
    float h1[2], hd[2];
    load_taps128(M, lane, h1, hd);
    const float X = M.kappa * x, Y = M.kappa * y;
    // ---- stage 1: wave (field, H)
    const int field = wr >> 1, H = wr & 1;
    if (TANG || field == 0) stage1_full(M, S[field], lane, field, H, X, Y, z, zoil);
    __syncthreads();
    // ---- stage 2: row p, quarters 2*(wr>>1) and 2*(wr>>1)+1
    const int p = lane + 64 * (wr & 1);
    float iq[2][32], izq[2][TANG ? 32 : 1];
    auto quarter = [&](auto qc, auto slot) {
        constexpr int Q = decltype(qc)::value, sl = decltype(slot)::value;
        float er[32], ei[32];
        stage2_quarter<Q>(S[0], p, er, ei);
        if constexpr (TANG) {
            float zr[32], zi[32];
            stage2_quarter<Q>(S[1], p, zr, zi);
            static_for<0, 32>([&](auto nc) {
                constexpr int n = decltype(nc)::value;
                constexpr int k = fidx<32>(n);
                iq[sl][n] = er[k] * er[k] + ei[k] * ei[k];
                izq[sl][n] = 2.f * (er[k] * zr[k] + ei[k] * zi[k]);
            });
        } else {
            static_for<0, 32>([&](auto nc) {
                constexpr int n = decltype(nc)::value;
                constexpr int k = fidx<32>(n);
                iq[sl][n] = er[k] * er[k] + ei[k] * ei[k];
            });
        }
    };
    using Sl0 = std::integral_constant<int, 0>;
    using Sl1 = std::integral_constant<int, 1>;
    if ((wr >> 1) == 0) {
        quarter(std::integral_constant<int, 0>{}, Sl0{});
        quarter(std::integral_constant<int, 1>{}, Sl1{});
    } else {
        quarter(std::integral_constant<int, 2>{}, Sl0{});
        quarter(std::integral_constant<int, 3>{}, Sl1{});
    }
    __syncthreads();  // all stage-1 reads are done
    float* Ib = S[0].re;   // real [128][kI2]
    float* Izb = S[1].re;
    {
        const int qb = 2 * (wr >> 1);
#pragma unroll
        for (int sl = 0; sl < 2; sl++)
#pragma unroll
            for (int n = 0; n < 32; n++) {
                Ib[p * kI2 + 4 * n + qb + sl] = iq[sl][n];
                if constexpr (TANG) Izb[p * kI2 + 4 * n + qb + sl] = izq[sl][n];
            }
    }
    __syncthreads();
    // ---- row blur at the crop columns: waves 0,1 columns [0, JA), waves 2,3 [JA, R)
    constexpr int JA = (R + 1) / 2, JB = R - JA;
    float ra[NCH][JA], rb[NCH][JB];
    if ((wr >> 1) == 0) row_blur_full<R, 0, JA, NCH>(Ib, Izb, p, h1, hd, M.kwin, ra);
    else row_blur_full<R, JA, JB, NCH>(Ib, Izb, p, h1, hd, M.kwin, rb);
    __syncthreads();  // all intensity reads are done
    float* Rb = S[0].re;  // [channel][column][row]
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        if ((wr >> 1) == 0) {
#pragma unroll
            for (int jj = 0; jj < JA; jj++) Rb[(c * R + jj) * kI2 + p] = ra[c][jj];
        } else {
#pragma unroll
            for (int jj = 0; jj < JB; jj++) Rb[(c * R + JA + jj) * kI2 + p] = rb[c][jj];
        }
    }
    __syncthreads();
    // ---- column blur: lane j, ROI rows [wr*RQ, wr*RQ + RQ)
    const int j = lane < R ? lane : 0;
    constexpr int NO = TANG ? 4 : 1;
    switch (wr) {
        case 0: column_blur_full<R, 0 * RQ, RQ, NO>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 1: column_blur_full<R, 1 * RQ, RQ, NO>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        case 2: column_blur_full<R, 2 * RQ, RQ, NO>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
        default: column_blur_full<R, 3 * RQ, RQ, NO>(Rb, j, h1, hd, M.kwin, vB, vX, vY, vZ); break;
    }
    __syncthreads();  // the buffers are reused by the next evaluation
}

}  // namespace zola

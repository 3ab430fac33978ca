// zola_psf64p.hpp -- the LM evaluation (PSF value and tangents) on the 64x64
// pupil grid by a pair of waves (one block) per ROI, for ROI sides up to 21
// (gfx950).
//
// Same model and blur as zola_psf64.hpp (process/cpu/PSFPhaseMany.java:553-622).
// One wave per ROI holds E, dE/dz, the three row-blur channels of both output
// halves and 34 KB of stage-1 buffers at once: about 450 registers and one wave
// per SIMD.  Here the work of one evaluation is split so that a wave needs
// about 200 registers and a block (ROI) about 37 KB of LDS: four blocks per CU,
// two waves per SIMD.
//  * Stage 1: wave wr forms the grid rows a = 2m + wr of both fields.  Lane
//    (v, s) evaluates the pupil of window column v at the 16 signed frequencies
//    of half s (phase, sincos; the stage-1 twiddle exp(2 pi i n wr / 64) is
//    folded into the phase), forms P and i kz P there, and the two lanes of a
//    column exchange halves with permlane32_swap: lane (v, 0) then transforms E
//    and lane (v, 1) dE/dz.  Lane (v, 1) holds its own half (slots 16..31) first,
//    which rotates its input by 16: the odd outputs change sign, undone at the
//    store.  The pupil tables are read from global memory as float4
//    {phi, amp, kz', kzo'} (16 KB, cache-resident).
//  * Stage 2: wave wr forms the output columns b = 2m + wr of every grid row
//    (lane = row) for E and dE/dz, then I = |E|^2 and Iz = 2 Re(conj(E) dE/dz)
//    there.  After a block barrier (the stage-1 buffers are free) it forms the
//    row-blur partial sums over those columns at the R crop columns, one
//    channel at a time (h1 (x) I, hd (x) I, h1 (x) Iz), each stored transposed
//    in LDS over the stage-1 buffers; after a barrier the two waves add the
//    two partials once (each wave half of the entries).
//  * Column pass: lane = (ROI column j, row group g); wave wr owns ROI rows
//    [wr RH, wr RH + RH) (RH = ceil(R/2)), RPL rows per lane.
#pragma once
#include "zola_psf64.hpp"

namespace zola {

template <int R>
struct Roi64p {
    static constexpr int RH = (R + 1) / 2;          // ROI rows of wave 0 (wave 1: R - RH)
    static constexpr int G = 64 / R;                // lane groups over the rows
    static constexpr int RPL = (RH + G - 1) / G;    // rows per lane
    static_assert(R <= 21, "the two waves' transposed partial sums must fit the stage-1 buffers");
};

// Blur taps shared by the block (the pupil tables stay in global memory).
struct alignas(16) BlurTaps64 {
    float h3[192];   // h1 repeated three times
    float hd3[192];  // hd repeated three times
};

__device__ __forceinline__ void load_taps64(BlurTaps64& T, const Model64& M) {
    for (int t = threadIdx.x; t < 192; t += blockDim.x) {
        T.h3[t] = M.h1[t & 63];
        T.hd3[t] = M.hd[t & 63];
    }
}

// The value of v held by the lane 32 away (lane ^ 32); upper: this lane is in
// lanes 32..63.  permlane32_swap with both operands v returns the lower half's
// value in [0] and the upper half's in [1] at every lane.
__device__ __forceinline__ float other_half(float v, bool upper) {
    const int b = __builtin_bit_cast(int, v);
    const auto p = __builtin_amdgcn_permlane32_swap(b, b, false, false);
    return __builtin_bit_cast(float, (int)(upper ? p[0] : p[1]));
}

// Stage 1 of both fields for the grid rows of parity h (wave-uniform).
__device__ __forceinline__ void stage1_64p(const float4* __restrict__ tab, WaveScratch64& SE, WaveScratch64& SZ,
                                           int lane, int h, float X, float Y, float z, float zoil) {
    const int v = lane & 31;
    const int s = lane >> 5;
    const bool up = s != 0;
    const float nv = (float)(v - 16);
    const float Xh = X + (float)h * (float)(2.0 * kPi / 64.0);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    float re[32], im[32];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        // own half: slot 16 s + k, signed frequency k - 16 s, window row k + 16 (1 - s)
        const int nu = k - 16 * s;
        const float4 t = tab[(nu + 16) * kW + v];
        const float q = t.x + Xh * (float)nu + cY + t.z * z - t.w * zoil;
        float sn, cs;
        sincos_reduced(q, sn, cs);
        const float pr = t.y * cs, pi = t.y * sn;
        const float zr = -t.z * pi, zi = t.z * pr;   // i kz P
        // keep this lane's field, give the partner the other one
        re[k] = up ? zr : pr;
        im[k] = up ? zi : pi;
        re[16 + k] = other_half(up ? pr : zr, up);
        im[16 + k] = other_half(up ? pi : zi, up);
    }
    fft_dif<32, +1>(re, im);
    float* dre = up ? SZ.re : SE.re;
    float* dim = up ? SZ.im : SE.im;
    const float sg = up ? -1.f : 1.f;
#pragma unroll
    for (int m = 0; m < 32; m++) {
        const int a = 2 * m + h;
        const float f = (m & 1) ? sg : 1.f;
        dre[a * kS1 + v] = re[fidx<32>(m)] * f;
        dim[a * kS1 + v] = im[fidx<32>(m)] * f;
    }
}

// Stage 2 of the output columns b = 2m + HP of grid row lane, for E (from SE)
// and dE/dz (from SZ): im[m] = I, iz[m] = Iz at column 2m + HP.  |E|^2 is
// formed before the dE/dz transform, and the scheduler is kept from starting
// the second transform's loads early, so that at most E, I and dE/dz are live.
template <int HP>
__device__ __forceinline__ void stage2_pair64(const WaveScratch64& SE, const WaveScratch64& SZ, int lane,
                                              float (&im)[32], float (&iz)[32]) {
    float er[32], ei[32];
    stage2_half64<HP>(SE, lane, er, ei);
    static_for<0, 32>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int k = fidx<32>(m);
        im[m] = er[k] * er[k] + ei[k] * ei[k];
    });
    __builtin_amdgcn_sched_barrier(0);
    float zr[32], zi[32];
    stage2_half64<HP>(SZ, lane, zr, zi);
    static_for<0, 32>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int k = fidx<32>(m);
        iz[m] = 2.f * (er[k] * zr[k] + ei[k] * zi[k]);
    });
}

// acc[jj] += sum_c k[c] img[m] over the columns 2m + HP, jj = (c + R/2 + 2m +
// HP) mod 64 < R.  The 64 taps are read from LDS four at a time at one
// address for the whole wave (a broadcast: no VALU work to spread them).
template <int R, int HP>
__device__ __forceinline__ void row_blur_half_lds(const float* __restrict__ k, const float (&img)[32],
                                                  float (&acc)[R]) {
    static_for<0, 16>([&](auto c4c) {
        constexpr int c4 = decltype(c4c)::value;
        const float4 t4 = reinterpret_cast<const float4*>(k)[c4];
        const float tv[4] = {t4.x, t4.y, t4.z, t4.w};
        static_for<0, 4>([&](auto uc) {
            constexpr int c = 4 * c4 + decltype(uc)::value;
            static_for<0, 32>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                constexpr int jj = (c + R / 2 + 2 * m + HP) & 63;
                if constexpr (jj < R) acc[jj] += tv[decltype(uc)::value] * img[m];
            });
        });
        __builtin_amdgcn_sched_barrier(0);
    });
}

// Row-blur partial sums over the columns 2m + HP, one channel at a time (only
// R accumulators live), stored transposed at Tw[(c R + j) kIp + lane]:
// c = 0: h1 (x) I, 1: hd (x) I, 2: h1 (x) Iz, at the R crop columns j.
template <int R, int HP, int kIp>
__device__ __forceinline__ void row_blur_pair64(const BlurTaps64& T, const float (&im)[32], const float (&iz)[32],
                                                float* Tw, int lane) {
    static_for<0, 3>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        float acc[R];
#pragma unroll
        for (int j = 0; j < R; j++) acc[j] = 0.f;
        row_blur_half_lds<R, HP>(c == 1 ? T.hd3 : T.h3, c == 2 ? iz : im, acc);
#pragma unroll
        for (int j = 0; j < R; j++) Tw[(c * R + j) * kIp + lane] = acc[j];
    });
}

// One LM evaluation of a ROI by its two waves (every wave of the block calls
// it; it contains block barriers).  scr: two WaveScratch64, contiguous.  On
// return lane (j, g) of wave wr holds, for ROI rows wr RH + g RPL + t, the raw
// blurred value (signed) in vB and the x, y, z tangents in vX, vY, vZ (the
// column-direction derivative kernel on the h1 row blur, the h1 column blur of
// the hd row blur, and the h1 blur of Iz, as in lm_eval64).
template <int R>
__device__ __forceinline__ void eval64p(const Model64& M, const BlurTaps64& T, float* scr, int lane, int wr,
                                        float x, float y, float z, float zoil, float (&vB)[Roi64p<R>::RPL],
                                        float (&vX)[Roi64p<R>::RPL], float (&vY)[Roi64p<R>::RPL],
                                        float (&vZ)[Roi64p<R>::RPL]) {
    constexpr int RH = Roi64p<R>::RH, G = Roi64p<R>::G, RPL = Roi64p<R>::RPL;
    constexpr int kBuf = (int)(sizeof(WaveScratch64) / sizeof(float));
    constexpr int kIp = 65;  // row stride of the transposed partial sums
    static_assert(2 * 3 * R * kIp <= 2 * kBuf, "both waves' partial sums must fit the buffers");
    WaveScratch64& SE = *reinterpret_cast<WaveScratch64*>(scr);
    WaveScratch64& SZ = *reinterpret_cast<WaveScratch64*>(scr + kBuf);
    stage1_64p(M.tab4, SE, SZ, lane, wr, M.kappa * x, M.kappa * y, z, zoil);
    __syncthreads();
    {
        float im[32], iz[32];
        if (wr == 0) stage2_pair64<0>(SE, SZ, lane, im, iz);
        else stage2_pair64<1>(SE, SZ, lane, im, iz);
        __syncthreads();  // both waves are done reading the stage-1 buffers
        float* Tw = scr + wr * 3 * R * kIp;
        if (wr == 0) row_blur_pair64<R, 0, kIp>(T, im, iz, Tw, lane);
        else row_blur_pair64<R, 1, kIp>(T, im, iz, Tw, lane);
    }
    __syncthreads();
    // the two waves' partial sums added once (each wave half of them), into
    // wave 0's array, which the column pass then reads alone
    {
        float* P0 = scr + wr * kIp + lane;
        const float* P1 = P0 + 3 * R * kIp;
#pragma unroll
        for (int e = 0; e < 3 * R; e += 2)
            if (e + wr < 3 * R) P0[e * kIp] += P1[e * kIp];
    }
    __syncthreads();
    // column pass over the summed partials; lanes past G groups read
    // column 0 and their outputs are never used.  Output row i = r0 + t takes
    // h1((i - R/2 - a) mod 64) = h1((a - t - c0) mod 64) (h1 even) and
    // hd((i - R/2 - a) mod 64) = -hd((a - t - c0) mod 64) (hd odd) with
    // c0 = r0 - R/2: one base address per lane and non-negative immediate
    // offsets a + (RPL - 1 - t).
    const int g = lane / R;
    const int jr = g < G ? lane % R : 0;
    const int kb = 64 + ((R / 2 - wr * RH - g * RPL) & 63) - (RPL - 1);
    const float* K1 = T.h3 + kb;
    const float* KD = T.hd3 + kb;
    const float* A0 = scr + jr * kIp;
#pragma unroll
    for (int t = 0; t < RPL; t++) vB[t] = vX[t] = vY[t] = vZ[t] = 0.f;
    // sliding windows w1[u] = K1[a + u], wd[u] = KD[a + u] (u < RPL): one new
    // tap of each kernel per grid row a; groups of 8 rows are kept apart so
    // that the scheduler does not load every tap of the pass at once
    float w1[RPL], wd[RPL];
#pragma unroll
    for (int u = 0; u < RPL - 1; u++) {
        w1[u + 1] = K1[u];
        wd[u + 1] = KD[u];
    }
    static_for<0, 8>([&](auto gc) {
        constexpr int a0 = decltype(gc)::value * 8;
        static_for<0, 8>([&](auto ac) {
            constexpr int a = a0 + decltype(ac)::value;
#pragma unroll
            for (int u = 0; u < RPL - 1; u++) {
                w1[u] = w1[u + 1];
                wd[u] = wd[u + 1];
            }
            w1[RPL - 1] = K1[a + RPL - 1];
            wd[RPL - 1] = KD[a + RPL - 1];
            const float t0 = A0[a];
            const float t1 = A0[R * kIp + a];
            const float tz = A0[2 * R * kIp + a];
#pragma unroll
            for (int t = 0; t < RPL; t++) {
                const float k1 = w1[RPL - 1 - t], kd = wd[RPL - 1 - t];
                vB[t] += k1 * t0;
                vX[t] -= kd * t0;
                vY[t] += k1 * t1;
                vZ[t] += k1 * tz;
            }
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    // Pin the results here: the tangents are only read on accepted trials, and
    // without this the compiler sinks their arithmetic past the reductions that
    // follow, keeping every loaded row-blur value live across them.
#pragma unroll
    for (int t = 0; t < RPL; t++) __asm__ __volatile__("" : "+v"(vB[t]), "+v"(vX[t]), "+v"(vY[t]), "+v"(vZ[t]));
    __syncthreads();  // the buffers are reused by the next evaluation
}

}  // namespace zola

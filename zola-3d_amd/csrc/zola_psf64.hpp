// zola_psf64.hpp -- wave-per-PSF synthesis of the ZOLA PSF model on a 64x64
// pupil grid, gfx950.
//
// Model (process/cpu/PSFPhaseMany.java:553-622):
//   P(u,v) = A(u,v) exp(i q),  q = phi + kx x + ky y + kz z - kzoil zoil
//   E = IDFT_64x64(P)                 (pupil scattered fftshifted)
//   I = |E|^2 on the s2 x s2 grid     (s2 = 64 here)
//   Bl = IDFT(G . DFT(I))             (Gaussian blur, G separable)
//   psf = |Bl| cropped to R x R, divided by its sum
//
// Layout choices (all per wave = one PSF/ROI):
//  * The pupil has radius < N/4, so its fftshifted support lies in a 32x32
//    window of signed frequencies n in [-16, 15].  A 64-point IDFT of a
//    32-point support splits into two 32-point IDFTs (even / odd outputs)
//    after a twiddle exp(2 pi i n h / 64).  Each lane runs 32-point
//    transforms entirely in VGPRs.
//  * Stage 1 (along the first axis): lane = (column v, half h).  The stage-1
//    twiddle is folded into the phase q before sincos, so it is free.
//  * LDS transpose (split re/im arrays, row stride 33: conflict free).
//  * Stage 2 (along the second axis): lane = grid row a, both halves.
//  * The Gaussian blur is applied as a separable circular convolution with the
//    exact s2-periodic kernel h1 = IDFT(G1) (real, even), so no further FFTs:
//    row pass in the lane that owns the row (coefficients are wave-uniform,
//    read from SGPRs), column pass after a second LDS transpose, lane =
//    (ROI column j, group of RPL ROI rows).
// Tangent fields dE/dtheta = IDFT(i f P) (f = kx, ky or kz) reuse the same
// stages; the tangent image is 2 Re(conj(E) dE).
#pragma once
#include "zola_fft.hpp"

namespace zola {

constexpr int kN = 64;     // pupil / E grid
constexpr int kW = 32;     // support window
constexpr int kS1 = 33;    // LDS row stride of the stage-1 transpose
constexpr int kTS = 68;    // LDS row stride of the column-blur transpose

struct alignas(16) BlockTables64 {
    float phi[kW * kW];    // aberration phase on the window [wu][wv]
    float amp[kW * kW];    // pupil amplitude (0 outside the disk)
    float kz[kW * kW];     // kz - kz0 (rad/um)
    float kzo[kW * kW];    // kzoil - kzoil0 (rad/um)
    float h3[192];         // blur kernel h1, repeated three times
    float hd3[192];        // derivative blur kernel hd (includes 1/dx), repeated
};

struct alignas(16) WaveScratch64 {
    float re[kN * kS1];
    float im[kN * kS1];
};

// Model constants shared by all kernels (passed by value in kernel args).
struct Model64 {
    const float* phi;  // global copies of the window tables (1024 each)
    const float* amp;
    const float* kz;
    const float* kzo;
    const float4* tab4;  // the same four tables interleaved per window point {phi, amp, kz, kzo}
    float kappa;       // 2 pi / (dx N): kx = kappa (n_u + 1/2)
    float h1[64];      // blur kernel over signed offset d (index d & 63)
    float hd[64];      // d/dt of the blur kernel (spectral), divided by dx
    float hs[64];      // d h1 / d sigma (sigma in camera pixels)
};

template <int R>
struct RoiLayout {
    static constexpr int G = 64 / R;              // lane groups over ROI rows
    static constexpr int RPL = (R + G - 1) / G;   // ROI rows per lane
};

__device__ __forceinline__ void load_tables64(BlockTables64& T, const Model64& M) {
    for (int t = threadIdx.x; t < kW * kW; t += blockDim.x) {
        T.phi[t] = M.phi[t];
        T.amp[t] = M.amp[t];
        T.kz[t] = M.kz[t];
        T.kzo[t] = M.kzo[t];
    }
    for (int t = threadIdx.x; t < 192; t += blockDim.x) {
        T.h3[t] = M.h1[t & 63];
        T.hd3[t] = M.hd[t & 63];
    }
}

// Signed frequency of FFT input slot j (j < 16: j, else j - 32).
__device__ __forceinline__ constexpr int slot_n(int j) { return j < 16 ? j : j - 32; }

// Stage 1 for one field (runtime, wave-uniform selector):
//   field 0 = E, 1 = dE/dx, 2 = dE/dy, 3 = dE/dz  (dE = IDFT(i f P)).
// Writes S1(a, v) = sum_u P(u,v) w^(n_u a) for all 64 rows a.
__device__ __forceinline__ void stage1_64(const BlockTables64& T, WaveScratch64& S, int lane,
                                          int field, float X, float Y, float z, float zoil,
                                          float kappa) {
    const int v = lane & 31;
    const int h = lane >> 5;
    const float nv = (float)(v - 16);
    const float Xh = X + (float)h * (float)(2.0 * kPi / 64.0);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    // i f P = (-f Pi, f Pr); for field 0 the factor is (c0, f) = (1, 0).
    const float c0 = field == 0 ? 1.f : 0.f;
    const float fx = field == 1 ? kappa : 0.f;
    const float fy = field == 2 ? kappa * (nv + 0.5f) : 0.f;
    const float fz = field == 3 ? 1.f : 0.f;
    float re[32], im[32];
#pragma unroll
    for (int j = 0; j < 32; j++) {
        const int nu = slot_n(j);
        const int wi = (nu + 16) * kW + v;
        const float kzv = T.kz[wi];
        const float q = T.phi[wi] + Xh * (float)nu + cY + kzv * z - T.kzo[wi] * zoil;
        float s, c;
        sincos_reduced(q, s, c);
        const float a = T.amp[wi];
        const float pr = a * c, pi = a * s;
        const float f = fx * ((float)nu + 0.5f) + fy + fz * kzv;
        re[j] = c0 * pr - f * pi;
        im[j] = c0 * pi + f * pr;
    }
    fft_dif<32, +1>(re, im);
#pragma unroll
    for (int m = 0; m < 32; m++) {
        const int a = 2 * m + h;
        S.re[a * kS1 + v] = re[fidx<32>(m)];
        S.im[a * kS1 + v] = im[fidx<32>(m)];
    }
}

// Stage 1 of E and dE/dz together (dE/dz = IDFT(i kz P)): the pupil phases and
// table reads are shared, two transforms.  Same outputs as stage1_64 with
// field 0 into SE and field 3 into SZ.
__device__ __forceinline__ void stage1_64_ez(const BlockTables64& T, WaveScratch64& SE, WaveScratch64& SZ,
                                             int lane, float X, float Y, float z, float zoil) {
    const int v = lane & 31;
    const int h = lane >> 5;
    const float nv = (float)(v - 16);
    const float Xh = X + (float)h * (float)(2.0 * kPi / 64.0);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    float re[32], im[32], zr[32], zi[32];
#pragma unroll
    for (int j = 0; j < 32; j++) {
        const int nu = slot_n(j);
        const int wi = (nu + 16) * kW + v;
        const float kzv = T.kz[wi];
        const float q = T.phi[wi] + Xh * (float)nu + cY + kzv * z - T.kzo[wi] * zoil;
        float sn, cs;
        sincos_reduced(q, sn, cs);
        const float a = T.amp[wi];
        const float pr = a * cs, pi = a * sn;
        re[j] = pr;
        im[j] = pi;
        zr[j] = -kzv * pi;
        zi[j] = kzv * pr;
    }
    fft_dif<32, +1>(re, im);
#pragma unroll
    for (int m = 0; m < 32; m++) {
        const int a = 2 * m + h;
        SE.re[a * kS1 + v] = re[fidx<32>(m)];
        SE.im[a * kS1 + v] = im[fidx<32>(m)];
    }
    fft_dif<32, +1>(zr, zi);
#pragma unroll
    for (int m = 0; m < 32; m++) {
        const int a = 2 * m + h;
        SZ.re[a * kS1 + v] = zr[fidx<32>(m)];
        SZ.im[a * kS1 + v] = zi[fidx<32>(m)];
    }
}

// Stage 2 half hp for row a = lane: returns E(a, 2m'+hp) in natural order.
template <int HP>
__device__ __forceinline__ void stage2_half64(const WaveScratch64& S, int lane, float (&re)[32],
                                              float (&im)[32]) {
    // Re-read the row for each half instead of keeping 64 loaded values live
    // across the first half's transform (the compiler would otherwise CSE the
    // two identical sets of LDS loads).
    __asm__ __volatile__("" ::: "memory");
    static_for<0, 32>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int v = slot_n(j) + 16;
        float r = S.re[lane * kS1 + v], i = S.im[lane * kS1 + v];
        if constexpr (HP == 1) twiddle_mul<64, (slot_n(j) + 64) % 64, +1>(r, i);
        re[j] = r;
        im[j] = i;
    });
    fft_dif<32, +1>(re, im);
}

// Row blur of one half of a grid row.  The lane owns row a; img[m] is the
// image at column b = 2m + HP.  acc[j] += h[(j - R/2 - b) & 63] * img[m].
// The kernel is held as one VGPR (lane i holds h[i]) and broadcast one
// coefficient at a time with readlane, grouping the FMAs by coefficient so
// that only one SGPR per kernel is live (64 coefficients do not fit the SGPR
// file once the compiler hoists them).
template <int R, int HP, int NK>
__device__ __forceinline__ void row_blur_half(const float (&hreg)[NK], const float (&img)[NK][32],
                                              float (&acc)[NK][R]) {
    static_for<0, 64>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        float sv[NK];
#pragma unroll
        for (int q = 0; q < NK; q++)
            sv[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hreg[q]), c));
        static_for<0, 32>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr int jj = (c + R / 2 + 2 * m + HP) & 63;
            if constexpr (jj < R) {
#pragma unroll
                for (int q = 0; q < NK; q++) acc[q][jj] += sv[q] * img[q][m];
            }
        });
        __builtin_amdgcn_sched_barrier(0);
    });
}

// Stage 2 of the E field: fills the row-blur accumulators (kernel h1 in hreg).
template <int R>
__device__ __forceinline__ void stage2_E64(float h1reg, const WaveScratch64& S, int lane,
                                           float (&acc)[R]) {
#pragma unroll
    for (int j = 0; j < R; j++) acc[j] = 0.f;
    static_for<0, 2>([&](auto hpc) {
        constexpr int HP = decltype(hpc)::value;
        float re[32], im[32];
        stage2_half64<HP>(S, lane, re, im);
        float img[1][32];
        static_for<0, 32>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const float er = re[fidx<32>(m)], ei = im[fidx<32>(m)];
            img[0][m] = er * er + ei * ei;
        });
        const float hr[1] = {h1reg};
        float (&a1)[1][R] = *reinterpret_cast<float (*)[1][R]>(&acc);
        row_blur_half<R, HP, 1>(hr, img, a1);
    });
}

// Column blur: stores this lane's row accumulators transposed in LDS, then
// lane (j, g) forms out(g*RPL + t, j) = sum_a h1[(i - R/2 - a) & 63] T(a, j).
template <int R>
__device__ __forceinline__ void column_blur64(const BlockTables64& T, WaveScratch64& S, int lane,
                                              const float (&acc)[R],
                                              float (&out)[RoiLayout<R>::RPL]) {
    constexpr int RPL = RoiLayout<R>::RPL;
    wave_lds_fence();  // all lanes are done reading the stage-1 data
    float* Tt = S.re;
#pragma unroll
    for (int j = 0; j < R; j++) Tt[j * kTS + lane] = acc[j];
    wave_lds_fence();
    const int j = lane % R;
    const int g = lane / R;
    const int base = 64 + ((g * RPL - R / 2) & 63);
#pragma unroll
    for (int t = 0; t < RPL; t++) out[t] = 0.f;
#pragma unroll
    for (int a = 0; a < 64; a++) {
        const float tv = Tt[j * kTS + a];
#pragma unroll
        for (int t = 0; t < RPL; t++) out[t] += T.h3[base + t - a] * tv;
    }
    wave_lds_fence();  // the scratch is reused by the next stage 1
}

// Row blur of one half of a grid row for the LM channels: acc[0] += h1 (x) I,
// acc[1] += hd (x) I, acc[2] += h1 (x) Iz, with I and Iz at columns 2m + HP.
template <int R, int HP>
__device__ __forceinline__ void row_blur_half_lm(float h1reg, float hdreg, const float (&im)[32],
                                                 const float (&iz)[32], float (&acc)[3][R]) {
    static_for<0, 64>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        const float s1 = bcast(h1reg, c), sd = bcast(hdreg, c);
        static_for<0, 32>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr int jj = (c + R / 2 + 2 * m + HP) & 63;
            if constexpr (jj < R) {
                acc[0][jj] += s1 * im[m];
                acc[1][jj] += sd * im[m];
                acc[2][jj] += s1 * iz[m];
            }
        });
        __builtin_amdgcn_sched_barrier(0);
    });
}

// Fused stage 2 of E (from SE) and dE/dz (from SZ) for row a = lane:
//   T0 = h1 (x) I,  T1 = hd (x) I  (row direction),  Tz = h1 (x) 2 Re(conj(E) dE/dz).
// |E|^2 is formed before the dE/dz transform so that E's and dE/dz's spectra
// and the three images are not all live at once.
template <int R>
__device__ __forceinline__ void stage2_EZ64(float h1reg, float hdreg, const WaveScratch64& SE,
                                            const WaveScratch64& SZ, int lane,
                                            float (&acc)[3][R]) {
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int j = 0; j < R; j++) acc[q][j] = 0.f;
    static_for<0, 2>([&](auto hpc) {
        constexpr int HP = decltype(hpc)::value;
        float er[32], ei[32];
        stage2_half64<HP>(SE, lane, er, ei);
        float im[32], iz[32];
        static_for<0, 32>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr int k = fidx<32>(m);
            im[m] = er[k] * er[k] + ei[k] * ei[k];
        });
        float zr[32], zi[32];
        stage2_half64<HP>(SZ, lane, zr, zi);
        static_for<0, 32>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr int k = fidx<32>(m);
            iz[m] = 2.f * (er[k] * zr[k] + ei[k] * zi[k]);
        });
        row_blur_half_lm<R, HP>(h1reg, hdreg, im, iz, acc);
    });
}

// Stores a lane's row accumulators transposed at Tt[j * kTS + lane].
template <int R>
__device__ __forceinline__ void store_rows64(float* Tt, int lane, const float (&acc)[R]) {
#pragma unroll
    for (int j = 0; j < R; j++) Tt[j * kTS + lane] = acc[j];
}

// Column pass over a transposed row-blurred image Tt for this lane's pixels
// (lane = (ROI column j, row group g)), with kernel table K3 (h3 or hd3).
template <int R>
__device__ __forceinline__ void column_pass64(const float* K3, const float* Tt, int lane,
                                              float (&out)[RoiLayout<R>::RPL]) {
    constexpr int RPL = RoiLayout<R>::RPL;
    const int j = lane % R;
    const int g = lane / R;
    const int base = 64 + ((g * RPL - R / 2) & 63);
#pragma unroll
    for (int t = 0; t < RPL; t++) out[t] = 0.f;
#pragma unroll
    for (int a = 0; a < 64; a++) {
        const float tv = Tt[j * kTS + a];
#pragma unroll
        for (int t = 0; t < RPL; t++) out[t] += K3[base + t - a] * tv;
    }
}

// Two column passes sharing the T reads: out1 with K3a, out2 with K3b.
template <int R>
__device__ __forceinline__ void column_pass2_64(const float* K3a, const float* K3b,
                                                const float* Tt, int lane,
                                                float (&o1)[RoiLayout<R>::RPL],
                                                float (&o2)[RoiLayout<R>::RPL]) {
    constexpr int RPL = RoiLayout<R>::RPL;
    const int j = lane % R;
    const int g = lane / R;
    const int base = 64 + ((g * RPL - R / 2) & 63);
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        o1[t] = 0.f;
        o2[t] = 0.f;
    }
#pragma unroll
    for (int a = 0; a < 64; a++) {
        const float tv = Tt[j * kTS + a];
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            o1[t] += K3a[base + t - a] * tv;
            o2[t] += K3b[base + t - a] * tv;
        }
    }
}

}  // namespace zola

// zola_fft.hpp -- register-resident radix-2 FFT building blocks for gfx950.
//
// One lane owns a whole short transform in VGPRs (float re[L], im[L]); the
// transform is fully unrolled at compile time with constexpr twiddles, so
// trivial twiddles (1, +-i, (+-1 +- i)/sqrt2) cost no multiplies.  Cross-lane
// data movement (the 2-D transpose) goes through LDS in the callers.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace zola {

constexpr double kPi = 3.14159265358979323846;

// constexpr sin/cos (Taylor series after reduction to [-pi, pi]) for the
// compile-time twiddle tables.
constexpr double cx_reduce(double x) {
    while (x > kPi) x -= 2 * kPi;
    while (x < -kPi) x += 2 * kPi;
    return x;
}
constexpr double cx_sin(double x) {
    x = cx_reduce(x);
    double term = x, sum = x;
    for (int n = 1; n < 30; n++) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos(double x) {
    x = cx_reduce(x);
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 30; n++) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }
constexpr int bitrev(int v, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <int LEN, int K, int SIGN>
struct Twiddle {
    static constexpr double ang = SIGN * 2.0 * kPi * K / LEN;
    static constexpr float c = (float)cx_cos(ang);
    static constexpr float s = (float)cx_sin(ang);
};

// (dr + i di) *= exp(SIGN * 2 pi i K / LEN)
template <int LEN, int K, int SIGN>
__device__ __forceinline__ void twiddle_mul(float& dr, float& di) {
    constexpr float r2 = 0.70710678118654752440f;
    if constexpr (K == 0) {
    } else if constexpr (4 * K == LEN) {
        float t = dr;
        dr = -SIGN * di;
        di = SIGN * t;
    } else if constexpr (8 * K == LEN) {
        float t = dr;
        dr = (dr - SIGN * di) * r2;
        di = (di + SIGN * t) * r2;
    } else if constexpr (8 * K == 3 * LEN) {
        float t = dr;
        dr = (-dr - SIGN * di) * r2;
        di = (-di + SIGN * t) * r2;
    } else {
        constexpr float c = Twiddle<LEN, K, SIGN>::c;
        constexpr float s = Twiddle<LEN, K, SIGN>::s;
        float t = dr;
        dr = dr * c - di * s;
        di = t * s + di * c;
    }
}

// In-place unnormalised DFT X[m] = sum_j x[j] exp(SIGN 2 pi i j m / L),
// decimation in frequency; on return X[m] sits at index bitrev(m).
template <int L, int SIGN>
__device__ __forceinline__ void fft_dif(float (&re)[L], float (&im)[L]) {
    static_for<0, ilog2(L)>([&](auto st) {
        constexpr int len = L >> decltype(st)::value;
        constexpr int half = len / 2;
        static_for<0, L / len>([&](auto blk) {
            static_for<0, half>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr int a = decltype(blk)::value * len + k;
                constexpr int b = a + half;
                float ar = re[a], ai = im[a], br = re[b], bi = im[b];
                re[a] = ar + br;
                im[a] = ai + bi;
                float dr = ar - br, di = ai - bi;
                twiddle_mul<len, k, SIGN>(dr, di);
                re[b] = dr;
                im[b] = di;
            });
        });
    });
}

// Natural-order index of output m after fft_dif.
template <int L>
__device__ __forceinline__ constexpr int fidx(int m) { return bitrev(m, ilog2(L)); }

// sin and cos of q for moderate |q| (|q| < 1e4 rad), in a few registers:
// quadrant reduction r = q - k pi/2 with pi/2 split into two floats
// (c1 = fl(pi/2), c2 = pi/2 - c1) and fused multiply-adds, so each k*c is
// exact before rounding; then the Taylor polynomials of sin (to r^9) and cos
// (to r^8) on |r| <= pi/4, whose truncation errors (r^11/11!, r^10/10!) stay
// below 3.5e-8.
__device__ __forceinline__ void sincos_reduced(float q, float& s, float& c) {
    const float k = rintf(q * 0.636619772367581343f);  // 2/pi
    float r = fmaf(-k, 1.57079637050628662109375f, q);
    r = fmaf(-k, -4.37113900018624283e-8f, r);
    const float r2 = r * r;
    const float sp = r + r * r2 * (-1.f / 6.f + r2 * (1.f / 120.f + r2 * (-1.f / 5040.f + r2 * (1.f / 362880.f))));
    const float cp = 1.f + r2 * (-0.5f + r2 * (1.f / 24.f + r2 * (-1.f / 720.f + r2 * (1.f / 40320.f))));
    const int qd = ((int)k) & 3;
    const float sa = (qd & 1) ? cp : sp, ca = (qd & 1) ? sp : cp;
    s = (qd & 2) ? -sa : sa;
    c = ((qd + 1) & 2) ? -ca : ca;
}

// Value of v in lane l, broadcast to the wave (l wave-uniform).
__device__ __forceinline__ float bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One DPP lane permutation within each row of 16 lanes (CTRL: a dpp_ctrl
// code), for a 32-bit or 64-bit value (the halves permuted alike).
template <int CTRL>
__device__ __forceinline__ float dpp_perm(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_perm(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Sum over rows 2k and 2k+1 (permlane16_swap) or over the two halves of the
// wave (permlane32_swap, ROW32 true), in every lane: with both operands equal
// to v, the swap returns the lower row's (half's) value in .x and the upper
// one's in .y at each lane, so every lane adds the same two numbers.
template <bool ROW32>
__device__ __forceinline__ float row_swap_sum(float v) {
    const int b = __builtin_bit_cast(int, v);
    const auto p = ROW32 ? __builtin_amdgcn_permlane32_swap(b, b, false, false)
                         : __builtin_amdgcn_permlane16_swap(b, b, false, false);
    return __builtin_bit_cast(float, (int)p[0]) + __builtin_bit_cast(float, (int)p[1]);
}
template <bool ROW32>
__device__ __forceinline__ double row_swap_sum(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
    const auto pl = ROW32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                          : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto ph = ROW32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                          : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double a = __builtin_bit_cast(double, ((long long)(int)ph[0] << 32) | (unsigned int)pl[0]);
    const double c = __builtin_bit_cast(double, ((long long)(int)ph[1] << 32) | (unsigned int)pl[1]);
    return a + c;
}

// Sums of K values over the wave, identical in every lane, with no LDS.
// Within each row of 16 lanes by DPP permutations: partner lane ^1 and ^2
// (quad_perm [1,0,3,2], [2,3,0,1]), then the mirrored lane in the half row and
// in the row (each pairs the two partial sums of the level below); across rows
// by the permlane swaps of row pairs, then of wave halves.  Every step adds a
// value to its partner's (or the same two values in both), so all lanes end
// with the same bits.
template <int K, class T>
__device__ __forceinline__ void wave_sum_rows(T (&v)[K]) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] += dpp_perm<0xB1>(v[k]);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] += dpp_perm<0x4E>(v[k]);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] += dpp_perm<0x141>(v[k]);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] += dpp_perm<0x140>(v[k]);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = row_swap_sum<false>(v[k]);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = row_swap_sum<true>(v[k]);
}

// wave_sum of K values at once, butterfly step by butterfly step: every lane
// performs exactly the additions of K separate wave_sum calls (same results),
// but the K cross-lane exchanges of a step are in flight together instead of
// each waiting out the exchange latency alone.
template <int K, class T>
__device__ __forceinline__ void wave_sum_v(T (&v)[K]) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        T t[K];
#pragma unroll
        for (int k = 0; k < K; k++) t[k] = __shfl_xor(v[k], o, 64);
#pragma unroll
        for (int k = 0; k < K; k++) v[k] += t[k];
    }
}

// Orders LDS traffic of one wave around a cross-lane exchange: all lanes'
// LDS writes before the fence are complete before any read after it.  Only
// the issuing wave is involved (waves of a block run independent ROIs).
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

}  // namespace zola

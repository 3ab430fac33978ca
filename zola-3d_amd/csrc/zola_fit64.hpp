// zola_fit64.hpp -- per-ROI fitting kernels on the 64x64 pupil grid (gfx950).
//
// One wave owns one ROI for the whole fit (persistent over iterations); the
// waves of a block share the pupil tables in LDS and never synchronise with
// each other after the table load, so ROIs with different iteration counts
// retire independently (per-ROI convergence masks are the waves themselves).
//
// Two fit forms:
//  * k_fit_lm64  -- Poisson-MLE Levenberg-Marquardt.  The Jacobian of the
//    model comes from analytic tangent fields dE/dx, dE/dy, dE/dz (one extra
//    pupil transform each, sharing the retained E), no finite differences.
//    FP32 PSF synthesis; FP64 likelihood, gradient, normal matrix and solve.
//  * k_fit_ref64 -- the reference's coordinate descent with central finite
//    differences (process/cpu/Localization.java:29-31,115,317-653,672-1143),
//    every likelihood a full PSF synthesis (cached when only A/B change).
//
// The fit kernels are written against a synthesis policy (Synth64, Synth128:
// pupil grid 64 or 128) that provides the value-only and the value+tangent
// evaluations and the LDS each needs; the ROI pixel layout (RoiLayout<R>) and
// all fitting logic are shared.
#pragma once
#include "zola_psf64.hpp"
#include "zola_psf128.hpp"
#include "zola_psf64q.hpp"
#include "zola_psf64p.hpp"
#include "zola_psf128o.hpp"

namespace zola {

struct FitArgs {
    int B;
    int max_iter;
    float xystep;
    double zmin, zmax, zfocus, thr, tol;
    double noise_rel;         // LM: likelihood resolution, relative to A
    const float* pixels;      // B*R*R
    const double* init;       // B*5 [A, B, x, y, z]
    const float* zoil;        // B per-ROI focus term z_oil (um), or null: zfocus for all
    const float* varg;        // B*R*R sCMOS term var/gain^2 added to mu per pixel, or null
    int* queue;               // persistent form: ROI counter (zeroed before launch), or null
    double* theta;            // B*5
    double* nll;              // B
    double* crlb;             // B*3
    int* iters;               // B
    int* status;              // B
};

enum : int { kStConverged = 1, kStMaxIter = 2, kStNaN = 4, kStLowPhoton = 8 };

template <int R>
struct Px {
    static constexpr int G = RoiLayout<R>::G;
    static constexpr int RPL = RoiLayout<R>::RPL;
    __device__ __forceinline__ static bool valid(int lane, int t) {
        const int g = lane / R;
        return g < G && g * RPL + t < R;
    }
};

// E field at (x, y, z, zoil): raw blurred values v (signed) of this lane's
// ROI pixels.  Every caller runs it from a single site in its iteration loop.
template <int R>
__device__ __forceinline__ void eval_E64(const Model64& M, const BlockTables64& T,
                                         WaveScratch64& S, int lane, float h1reg, float x,
                                         float y, float z, float zoil,
                                         float (&v)[RoiLayout<R>::RPL]) {
    stage1_64(T, S, lane, 0, M.kappa * x, M.kappa * y, z, zoil, M.kappa);
    wave_lds_fence();
    float acc[R];
    stage2_E64<R>(h1reg, S, lane, acc);
    column_blur64<R>(T, S, lane, acc, v);
}

// |v| normalised to sum 1 over the ROI; returns the sum.
template <class SY, int R>
__device__ __forceinline__ float normalise(int lane, int wr, typename SY::Red& red,
                                           const float (&v)[SY::template RPL<R>],
                                           float (&psf)[SY::template RPL<R>]) {
    constexpr int RPL = SY::template RPL<R>;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        psf[t] = fabsf(v[t]);
        if (SY::template valid<R>(lane, wr, t)) s += psf[t];
    }
    s = SY::sum(s, red, lane, wr);
    const float inv = 1.f / s;
#pragma unroll
    for (int t = 0; t < RPL; t++) psf[t] *= inv;
    return s;
}

// sCMOS term of a pixel: var/gain^2 of the ROI's map (vroi), 0 without one.
template <class SY, int R>
__device__ __forceinline__ double scmos_term(const float* vroi, int lane, int wr, int t) {
    return vroi ? (double)vroi[SY::template pix<R>(lane, wr, t)] : 0.0;
}

// Natural logarithm of a positive, finite, normal x (the Poisson means of a fit
// are >= 0.01), to 2 ulp: x = 2^e m with m in [sqrt(1/2), sqrt(2)), and
// ln m = 2 atanh(s) = 2 s sum_k s^2k / (2k + 1), s = (m - 1)/(m + 1),
// |s| < 0.1716, summed to k = 10 (truncation below 1e-17 relative); e ln 2 is
// added in two parts (ln 2 split as in fdlibm's e_log.c) so that e ln2_hi is
// exact.  About a quarter of the instructions of the library's log, which
// carries extra precision the likelihood does not need.
// 1/21, 1/19, ..., 1/3: the atanh series of log_pos.
__constant__ const double kLogSeries[10] = {1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0,
                                            1.0 / 11.0, 1.0 / 9.0,  1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0};

__device__ __forceinline__ double log_pos(double x, const double (&C)[10]) {
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m += m;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (f + 2.0);
    const double s2 = s * s;
    double t = C[0];
#pragma unroll
    for (int k = 1; k < 10; k++) t = t * s2 + C[k];
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double de = (double)e, ss = s + s;
    return de * ln2_hi + (ss + (ss * s2 * t + de * ln2_lo));
}

// Poisson NLL without ln n! (Localization.computeLikelihood, 1125-1143), FP64;
// mu = A psf + B, plus the sCMOS term var/gain^2 of the pixel when vroi is
// given (1131-1132).
template <class SY, int R>
__device__ __forceinline__ double nll64(int lane, int wr, typename SY::Red& red,
                                        const float (&psf)[SY::template RPL<R>],
                                        const float (&n)[SY::template RPL<R>], double A, double Bg,
                                        const float* vroi) {
    constexpr int RPL = SY::template RPL<R>;
    // The series coefficients of log_pos are read once per call through a
    // pointer the compiler cannot see through (scalar loads into SGPRs).  As
    // literals they are built in VGPR pairs, hoisted out of the fit loops and
    // held (or spilled) through every evaluation.
    const double* Cp = kLogSeries;
    __asm__ __volatile__("" : "+s"(Cp));
    double C[10];
#pragma unroll
    for (int k = 0; k < 10; k++) C[k] = Cp[k];
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        if (SY::template valid<R>(lane, wr, t)) {
            const double mu = (double)psf[t] * A + Bg + scmos_term<SY, R>(vroi, lane, wr, t);
            acc += mu - (double)n[t] * log_pos(mu, C);
        }
    }
    return SY::sum(acc, red, lane, wr);
}

// Cholesky factor Lc (lower, full storage) of the symmetric 5x5 matrix A
// (lower triangle read); false unless A is positive definite.
__device__ __forceinline__ bool chol5(const double (&A)[25], double (&Lc)[25]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 5; j++) {
        double s = A[j * 5 + j];
#pragma unroll
        for (int k = 0; k < j; k++) s -= Lc[j * 5 + k] * Lc[j * 5 + k];
        ok = ok && (s > 0.0);
        Lc[j * 5 + j] = sqrt(fmax(s, 1e-300));
        const double inv = 1.0 / Lc[j * 5 + j];
#pragma unroll
        for (int i = j + 1; i < 5; i++) {
            double t = A[i * 5 + j];
#pragma unroll
            for (int k = 0; k < j; k++) t -= Lc[i * 5 + k] * Lc[j * 5 + k];
            Lc[i * 5 + j] = t * inv;
        }
    }
    return ok;
}

// chol5 in place: the lower triangle of A is replaced by its Cholesky factor
// (the same operations in the same order as chol5; half the registers).
__device__ __forceinline__ bool chol5_ip(double (&A)[25]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 5; j++) {
        double s = A[j * 5 + j];
#pragma unroll
        for (int k = 0; k < j; k++) s -= A[j * 5 + k] * A[j * 5 + k];
        ok = ok && (s > 0.0);
        A[j * 5 + j] = sqrt(fmax(s, 1e-300));
        const double inv = 1.0 / A[j * 5 + j];
#pragma unroll
        for (int i = j + 1; i < 5; i++) {
            double t = A[i * 5 + j];
#pragma unroll
            for (int k = 0; k < j; k++) t -= A[i * 5 + k] * A[j * 5 + k];
            A[i * 5 + j] = t * inv;
        }
    }
    return ok;
}

// Solves (Lc Lc^T) x = b in place.
__device__ __forceinline__ void chol5_solve(const double (&Lc)[25], double (&b)[5]) {
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double t = b[i];
#pragma unroll
        for (int k = 0; k < i; k++) t -= Lc[i * 5 + k] * b[k];
        b[i] = t / Lc[i * 5 + i];
    }
#pragma unroll
    for (int i = 4; i >= 0; i--) {
        double t = b[i];
#pragma unroll
        for (int k = i + 1; k < 5; k++) t -= Lc[k * 5 + i] * b[k];
        b[i] = t / Lc[i * 5 + i];
    }
}

// ---------------------------------------------------------------------------
// Levenberg-Marquardt (theta order [A, B, x, y, z])
//
// Every trial point is evaluated with its Jacobian: two pupil transforms (E and
// dE/dz) and four separable blurs of the same intensity.  d/dx and d/dy need
// no transform: moving the emitter by x moves the band-limited intensity by
// x/dx samples, so the blurred x/y tangents are the intensity blurred with the
// spectral-derivative kernel hd along that axis.
// ---------------------------------------------------------------------------
template <int R>
__device__ __forceinline__ void lm_eval64(const Model64& M, const BlockTables64& T, float* scr,
                                          int lane, float x, float y, float z, float zoil,
                                          float (&vB)[RoiLayout<R>::RPL],
                                          float (&vX)[RoiLayout<R>::RPL],
                                          float (&vY)[RoiLayout<R>::RPL],
                                          float (&vZ)[RoiLayout<R>::RPL]) {
    WaveScratch64& SE = *reinterpret_cast<WaveScratch64*>(scr);
    WaveScratch64& SZ = *reinterpret_cast<WaveScratch64*>(scr + 2 * kN * kS1);
    stage1_64_ez(T, SE, SZ, lane, M.kappa * x, M.kappa * y, z, zoil);
    wave_lds_fence();
    {
        float acc[3][R];
        stage2_EZ64<R>(T.h3[lane], T.hd3[lane], SE, SZ, lane, acc);
        wave_lds_fence();
        store_rows64<R>(scr, lane, acc[0]);
        store_rows64<R>(scr + R * kTS, lane, acc[1]);
        store_rows64<R>(scr + 2 * R * kTS, lane, acc[2]);
    }
    wave_lds_fence();
    column_pass2_64<R>(T.h3, T.hd3, scr, lane, vB, vX);
    column_pass64<R>(T.h3, scr + R * kTS, lane, vY);
    column_pass64<R>(T.h3, scr + 2 * R * kTS, lane, vZ);
    // Pin the results here: the tangents are only read on accepted trials, and
    // without this the compiler sinks their arithmetic past the reductions that
    // follow, keeping every loaded row-blur value live across them.
#pragma unroll
    for (int t = 0; t < RoiLayout<R>::RPL; t++)
        __asm__ __volatile__("" : "+v"(vB[t]), "+v"(vX[t]), "+v"(vY[t]), "+v"(vZ[t]));
    wave_lds_fence();
}

// Synthesis policies: tables shared by a block, per-wave scratch (floats) of
// the LM and reference-form kernels, and the two evaluations.
struct Synth64 {
    using Model = Model64;
    using Shared = BlockTables64;
    static constexpr int kWaves = 1;  // waves per ROI
    template <int R>
    static constexpr int RPL = RoiLayout<R>::RPL;
    // lane = (ROI column j, row group g), rows g*RPL + t
    template <int R>
    __device__ __forceinline__ static bool valid(int lane, int, int t) { return Px<R>::valid(lane, t); }
    template <int R>
    __device__ __forceinline__ static int pix(int lane, int, int t) { return ((lane / R) * RPL<R> + t) * R + lane % R; }
    struct Red {};
    __device__ __forceinline__ static float sum(float v, Red&, int, int) {
        float a[1] = {v};
        wave_sum_rows<1>(a);
        return a[0];
    }
    __device__ __forceinline__ static double sum(double v, Red&, int, int) {
        double a[1] = {v};
        wave_sum_rows<1>(a);
        return a[0];
    }
    template <int K>
    __device__ __forceinline__ static void sumv(double (&v)[K], Red&, int, int) {
        wave_sum_rows<K>(v);
    }
    template <int K>
    __device__ __forceinline__ static void sumv(float (&v)[K], Red&, int, int) {
        wave_sum_rows<K>(v);
    }
    static constexpr int kLmFloats = 2 * (int)(sizeof(WaveScratch64) / sizeof(float));
    static constexpr int kRefFloats = (int)(sizeof(WaveScratch64) / sizeof(float));
    __device__ __forceinline__ static void load(Shared& T, const Model& M) { load_tables64(T, M); }
    template <int R>
    __device__ __forceinline__ static void eval(const Model& M, const Shared& T, float* scr, int lane, int, float x, float y,
                                float z, float zoil, float (&v)[RoiLayout<R>::RPL]) {
        eval_E64<R>(M, T, *reinterpret_cast<WaveScratch64*>(scr), lane, T.h3[lane], x, y, z, zoil, v);
    }
    template <int R>
    __device__ __forceinline__ static void lm_eval(const Model& M, const Shared& T, float* scr, int lane, int, float x, float y,
                                   float z, float zoil, float (&vB)[RoiLayout<R>::RPL],
                                   float (&vX)[RoiLayout<R>::RPL], float (&vY)[RoiLayout<R>::RPL],
                                   float (&vZ)[RoiLayout<R>::RPL]) {
        static_assert(3 * R * kTS <= 4 * kN * kS1, "transposed blur rows must fit the scratch");
        lm_eval64<R>(M, T, scr, lane, x, y, z, zoil, vB, vX, vY, vZ);
    }
};

// Four waves per ROI (zola_psf128.hpp, zola_psf64q.hpp); lane = ROI column
// j, wave wr owns rows wr*RQ + t (RQ = ceil(R/4)).  Every reduction is summed
// over the four waves through LDS in a fixed order, so all waves see identical
// values and take identical branches (the block barriers inside the
// evaluations stay matched).
struct QuadWaves {
    static constexpr int kWaves = 4;
    template <int R>
    static constexpr int RPL = (R + 3) / 4;
    template <int R>
    __device__ __forceinline__ static bool valid(int lane, int wr, int t) { return lane < R && wr * RPL<R> + t < R; }
    template <int R>
    __device__ __forceinline__ static int pix(int lane, int wr, int t) { return (wr * RPL<R> + t) * R + lane; }
    struct Red {
        double d[4][24];
        float f[4][4];
    };
    template <int K>
    __device__ __forceinline__ static void sumv(double (&v)[K], Red& r, int lane, int wr) {
        static_assert(K <= 24, "reduction width");
        wave_sum_rows<K>(v);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (lane == 0) r.d[wr][k] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = ((r.d[0][k] + r.d[1][k]) + r.d[2][k]) + r.d[3][k];
        __syncthreads();
    }
    template <int K>
    __device__ __forceinline__ static void sumv(float (&v)[K], Red& r, int lane, int wr) {
        static_assert(K <= 4, "reduction width");
        wave_sum_rows<K>(v);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (lane == 0) r.f[wr][k] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = ((r.f[0][k] + r.f[1][k]) + r.f[2][k]) + r.f[3][k];
        __syncthreads();
    }
    __device__ __forceinline__ static double sum(double v, Red& r, int lane, int wr) {
        double a[1] = {v};
        sumv<1>(a, r, lane, wr);
        return a[0];
    }
    __device__ __forceinline__ static float sum(float v, Red& r, int lane, int wr) {
        float a[1] = {v};
        sumv<1>(a, r, lane, wr);
        return a[0];
    }
};

struct Synth128 : QuadWaves {
    using Model = Model128;
    using Shared = NoShared;
    static constexpr int kLmFloats = 2 * (int)(sizeof(Buf128) / sizeof(float));
    static constexpr int kRefFloats = kLmFloats;
    __device__ __forceinline__ static void load(Shared&, const Model&) {}
    template <int R>
    __device__ __forceinline__ static void eval(const Model& M, const Shared&, float* scr, int lane, int wr, float x, float y,
                                float z, float zoil, float (&v)[Roi128<R>::RQ]) {
        float d0[Roi128<R>::RQ], d1[Roi128<R>::RQ], d2[Roi128<R>::RQ];
        eval128<R, false>(M, *reinterpret_cast<Buf128 (*)[2]>(scr), lane, wr, x, y, z, zoil, v, d0, d1, d2);
    }
    template <int R>
    __device__ __forceinline__ static void lm_eval(const Model& M, const Shared&, float* scr, int lane, int wr, float x, float y,
                                   float z, float zoil, float (&vB)[Roi128<R>::RQ], float (&vX)[Roi128<R>::RQ],
                                   float (&vY)[Roi128<R>::RQ], float (&vZ)[Roi128<R>::RQ]) {
        eval128<R, true>(M, *reinterpret_cast<Buf128 (*)[2]>(scr), lane, wr, x, y, z, zoil, vB, vX, vY, vZ);
    }
};

// Eight waves per ROI (zola_psf128o.hpp): lane = ROI column j, wave wr owns
// rows wr*R8 + t (R8 = ceil(R/8)).  Reductions are summed over the eight waves
// through LDS in wave order, so every wave sees identical values and takes
// identical branches.
// Lane l's x at every lane (both halves of an FP64 value through readlane).
__device__ __forceinline__ double bcast_d(double x, int l) {
    const long long u = __builtin_bit_cast(long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), l);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)hi << 32) | lo));
}

struct OctWaves {
    static constexpr int kWaves = 8;
    template <int R>
    static constexpr int RPL = Roi128o<R>::R8;
    template <int R>
    __device__ __forceinline__ static bool valid(int lane, int wr, int t) { return lane < R && wr * RPL<R> + t < R; }
    template <int R>
    __device__ __forceinline__ static int pix(int lane, int wr, int t) { return (wr * RPL<R> + t) * R + lane; }
    struct Red {
        double d[8][24];
        float f[8][4];
    };
    // The eight waves' partials of value k are added in wave order by lane k
    // (eight LDS reads per lane, not 8 K), then broadcast: every wave forms the
    // same totals bit for bit.
    template <int K>
    __device__ __forceinline__ static void sumv(double (&v)[K], Red& r, int lane, int wr) {
        static_assert(K <= 24, "reduction width");
        wave_sum_rows<K>(v);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (lane == 0) r.d[wr][k] = v[k];
        __syncthreads();
        double tot = 0.0;
        if (lane < K) {
            tot = r.d[0][lane];
#pragma unroll
            for (int q = 1; q < 8; q++) tot += r.d[q][lane];
        }
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = bcast_d(tot, k);
        __syncthreads();
    }
    template <int K>
    __device__ __forceinline__ static void sumv(float (&v)[K], Red& r, int lane, int wr) {
        static_assert(K <= 4, "reduction width");
        wave_sum_rows<K>(v);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (lane == 0) r.f[wr][k] = v[k];
        __syncthreads();
        float tot = 0.f;
        if (lane < K) {
            tot = r.f[0][lane];
#pragma unroll
            for (int q = 1; q < 8; q++) tot += r.f[q][lane];
        }
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tot), k));
        __syncthreads();
    }
    __device__ __forceinline__ static double sum(double v, Red& r, int lane, int wr) {
        double a[1] = {v};
        sumv<1>(a, r, lane, wr);
        return a[0];
    }
    __device__ __forceinline__ static float sum(float v, Red& r, int lane, int wr) {
        float a[1] = {v};
        sumv<1>(a, r, lane, wr);
        return a[0];
    }
};

// The LM form on the 128 grid by eight waves per ROI.
struct Synth128o : OctWaves {
    using Model = Model128;
    using Shared = NoShared;
    static constexpr int kLmFloats = 2 * (int)(sizeof(Buf128) / sizeof(float));
    __device__ __forceinline__ static void load(Shared&, const Model&) {}
    template <int R>
    __device__ __forceinline__ static void lm_eval(const Model& M, const Shared&, float* scr, int lane, int wr, float x,
                                                   float y, float z, float zoil, float (&vB)[Roi128o<R>::R8],
                                                   float (&vX)[Roi128o<R>::R8], float (&vY)[Roi128o<R>::R8],
                                                   float (&vZ)[Roi128o<R>::R8]) {
        eval128o<R>(M, *reinterpret_cast<Buf128 (*)[2]>(scr), lane, wr, x, y, z, zoil, vB, vX, vY, vZ);
    }
};

// The 64 grid by four waves per ROI: stage-1 tables in LDS as for one wave.
struct Synth64q : QuadWaves {
    using Model = Model64;
    using Shared = BlockTables64;
    static constexpr int kLmFloats = 2 * (int)(sizeof(WaveScratch64) / sizeof(float));
    static constexpr int kRefFloats = (int)(sizeof(WaveScratch64) / sizeof(float));
    __device__ __forceinline__ static void load(Shared& T, const Model& M) { load_tables64(T, M); }
    template <int R>
    __device__ __forceinline__ static void eval(const Model& M, const Shared& T, float* scr, int lane, int wr, float x, float y,
                                float z, float zoil, float (&v)[Roi64q<R>::RQ]) {
        float d0[Roi64q<R>::RQ], d1[Roi64q<R>::RQ], d2[Roi64q<R>::RQ];
        eval64q<R, false>(M, T, scr, lane, wr, x, y, z, zoil, v, d0, d1, d2);
    }
    template <int R>
    __device__ __forceinline__ static void lm_eval(const Model& M, const Shared& T, float* scr, int lane, int wr, float x, float y,
                                   float z, float zoil, float (&vB)[Roi64q<R>::RQ], float (&vX)[Roi64q<R>::RQ],
                                   float (&vY)[Roi64q<R>::RQ], float (&vZ)[Roi64q<R>::RQ]) {
        eval64q<R, true>(M, T, scr, lane, wr, x, y, z, zoil, vB, vX, vY, vZ);
    }
};

// Two waves per ROI (zola_psf64p.hpp): lane = (ROI column j, row group g),
// wave wr owns ROI rows [wr RH, wr RH + RH) (RH = ceil(R/2)), RPL rows per lane.
// Reductions are summed over the two waves through LDS, wave 0's part first,
// so both waves see identical values and take identical branches.
struct PairWaves {
    static constexpr int kWaves = 2;
    template <int R>
    static constexpr int RPL = Roi64p<R>::RPL;
    template <int R>
    __device__ __forceinline__ static bool valid(int lane, int wr, int t) {
        const int g = lane / R;
        const int rows = wr == 0 ? Roi64p<R>::RH : R - Roi64p<R>::RH;
        return g < Roi64p<R>::G && g * RPL<R> + t < rows;
    }
    template <int R>
    __device__ __forceinline__ static int pix(int lane, int wr, int t) {
        return (wr * Roi64p<R>::RH + (lane / R) * RPL<R> + t) * R + lane % R;
    }
    struct Red {
        double d[2][24];
        float f[2][4];
    };
    template <int K>
    __device__ __forceinline__ static void sumv(double (&v)[K], Red& r, int lane, int wr) {
        static_assert(K <= 24, "reduction width");
        wave_sum_rows<K>(v);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (lane == 0) r.d[wr][k] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = r.d[0][k] + r.d[1][k];
        __syncthreads();
    }
    template <int K>
    __device__ __forceinline__ static void sumv(float (&v)[K], Red& r, int lane, int wr) {
        static_assert(K <= 4, "reduction width");
        wave_sum_rows<K>(v);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (lane == 0) r.f[wr][k] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = r.f[0][k] + r.f[1][k];
        __syncthreads();
    }
    __device__ __forceinline__ static double sum(double v, Red& r, int lane, int wr) {
        double a[1] = {v};
        sumv<1>(a, r, lane, wr);
        return a[0];
    }
    __device__ __forceinline__ static float sum(float v, Red& r, int lane, int wr) {
        float a[1] = {v};
        sumv<1>(a, r, lane, wr);
        return a[0];
    }
};

// The 64 grid by two waves per ROI (LM form only): blur taps in LDS, pupil
// tables read from global memory.
struct Synth64p : PairWaves {
    using Model = Model64;
    using Shared = BlurTaps64;
    static constexpr int kLmFloats = 2 * (int)(sizeof(WaveScratch64) / sizeof(float));
    static constexpr int kRefFloats = kLmFloats;
    __device__ __forceinline__ static void load(Shared& T, const Model& M) { load_taps64(T, M); }
    template <int R>
    __device__ __forceinline__ static void lm_eval(const Model& M, const Shared& T, float* scr, int lane, int wr, float x, float y,
                                   float z, float zoil, float (&vB)[Roi64p<R>::RPL], float (&vX)[Roi64p<R>::RPL],
                                   float (&vY)[Roi64p<R>::RPL], float (&vZ)[Roi64p<R>::RPL]) {
        eval64p<R>(M, T, scr, lane, wr, x, y, z, zoil, vB, vX, vY, vZ);
    }
};

// Trial point of an LM step d from th, clamped as the fit clamps every trial:
// x, y by at most one pixel, z by at most 0.2 um, A >= 1, B >= 0.01,
// zmin <= z <= zmax.
__device__ __forceinline__ void lm_clamp_trial(const double* th, const double (&d)[5], double dxmax, double zmin,
                                               double zmax, double (&tn)[5]) {
#pragma unroll
    for (int p = 0; p < 5; p++) tn[p] = th[p] + d[p];
#pragma unroll
    for (int p = 2; p < 4; p++) {
        if (d[p] > dxmax) tn[p] = th[p] + dxmax;
        if (d[p] < -dxmax) tn[p] = th[p] - dxmax;
    }
    if (d[4] > 0.2) tn[4] = th[4] + 0.2;
    if (d[4] < -0.2) tn[4] = th[4] - 0.2;
    tn[0] = fmax(tn[0], 1.0);
    tn[1] = fmax(tn[1], 0.01);
    tn[4] = fmin(fmax(tn[4], zmin), zmax);
}

struct LmState {
    double H[15];   // Fisher matrix, packed upper triangle
    double S[15];   // secant estimate of the rest of the Hessian, packed upper
    double gr[5];   // gradient
    double gp[5];   // gradient at the previous point (secant update)
    double th[5];   // current point
    double tn[5];   // trial point
    double nll, lambda, nu, pred;
};

// The trial tc from L.th is below tol in x, y, z and below 1% of the
// statistical precision (1/sqrt(F_ii)) in A and B.
__device__ __forceinline__ bool lm_small_step(const LmState& L, const double (&tc)[5], double tol) {
    return fmax(fabs(tc[2] - L.th[2]), fmax(fabs(tc[3] - L.th[3]), fabs(tc[4] - L.th[4]))) < tol &&
           fabs(tc[0] - L.th[0]) * sqrt(L.H[0]) < 0.01 && fabs(tc[1] - L.th[1]) * sqrt(L.H[5]) < 0.01;
}

__device__ __forceinline__ constexpr int hidx(int p, int q) {
    // packed upper triangle, row-major
    const int lo = p < q ? p : q, hi = p < q ? q : p;
    return lo * 5 - (lo * (lo - 1)) / 2 + (hi - lo);
}
__device__ __forceinline__ double hget(const double (&H)[15], int p, int q) { return H[hidx(p, q)]; }

// Structured secant correction (NL2SOL's idea, Dennis, Gay & Welsch 1981, for
// a Poisson likelihood): the NLL Hessian is the Fisher matrix F plus
// S = sum (1 - n/mu) d2mu, which Fisher scoring drops.  After an accepted
// step s = tn - th, S is sized by min(1, |s'y|/|s'Ss|) and updated with
// Powell's symmetric Broyden formula so that S s = y = g+ - g - F+ s, in the
// variables scaled by sqrt(diag F+).  Reads g+ and F+ from L.gr and L.H, g from
// L.gp, the step from L.tn - L.th (all in LDS: no register array is indexed
// at run time).
__device__ __forceinline__ void lm_secant_update(LmState& L) {
    // register copies of the LDS state: its loads issue together
    double H[15], S[15], dth[5], dg[5];
#pragma unroll
    for (int k = 0; k < 15; k++) {
        H[k] = L.H[k];
        S[k] = L.S[k];
    }
#pragma unroll
    for (int p = 0; p < 5; p++) {
        dth[p] = L.tn[p] - L.th[p];
        dg[p] = L.gr[p] - L.gp[p];
    }
    double D[5], iD[5], sh[5], yh[5];
    bool ok = true;
#pragma unroll
    for (int p = 0; p < 5; p++) {
        const double f = H[hidx(p, p)];
        ok = ok && (f > 0.0);
        D[p] = sqrt(fmax(f, 1e-300));
        iD[p] = 1.0 / D[p];
    }
    if (!ok) return;
#pragma unroll
    for (int p = 0; p < 5; p++) {
        double fs = 0.0;
#pragma unroll
        for (int q = 0; q < 5; q++) fs += H[hidx(p, q)] * dth[q];
        sh[p] = D[p] * dth[p];
        yh[p] = (dg[p] - fs) * iD[p];
    }
    double Ssh[5], sSs = 0.0, sy = 0.0, ss = 0.0;
#pragma unroll
    for (int p = 0; p < 5; p++) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 5; q++) t += S[hidx(p, q)] * iD[p] * iD[q] * sh[q];
        Ssh[p] = t;
        sSs += sh[p] * t;
        sy += sh[p] * yh[p];
        ss += sh[p] * sh[p];
    }
    if (!(ss > 0.0)) return;
    const double tau = sSs != 0.0 ? fmin(1.0, fabs(sy) / fabs(sSs)) : 1.0;
    double r[5], rs = 0.0;
#pragma unroll
    for (int p = 0; p < 5; p++) {
        r[p] = yh[p] - tau * Ssh[p];
        rs += r[p] * sh[p];
    }
    const double iss = 1.0 / ss;
#pragma unroll
    for (int p = 0; p < 5; p++)
#pragma unroll
        for (int q = p; q < 5; q++) {
            const double v = tau * S[hidx(p, q)] * iD[p] * iD[q] + (r[p] * sh[q] + sh[p] * r[q]) * iss -
                             rs * sh[p] * sh[q] * iss * iss;
            L.S[hidx(p, q)] = v * D[p] * D[q];
        }
}

// One LDS struct fixes the layout: shared tables first (low addresses, so lane
// addresses plus immediate offsets reach them), then per-wave state and scratch.
// WPB waves per block, WPB / SY::kWaves ROIs; the fit state is private to
// each wave (the waves of a ROI keep identical copies).
// The step of a ROI fitted by several waves, solved by its first wave and read
// by the others: trial point, predicted decrease, and flags (bit 0: no trial
// to evaluate, bit 1: converged before evaluating).
struct LmShare {
    double tn[5];
    double pred;
    int flag;
};

template <class SY, int WPB>
struct LmLds {
    typename SY::Shared T;
    LmState L[WPB];
    typename SY::Red red[WPB / SY::kWaves];
    LmShare share[WPB / SY::kWaves];
    int q[WPB / SY::kWaves];
    alignas(16) float S[WPB / SY::kWaves][SY::kLmFloats];
};

template <class SY, int WPB>
struct RefLds {
    typename SY::Shared T;
    typename SY::Red red[WPB / SY::kWaves];
    int q[WPB / SY::kWaves];
    alignas(16) float S[WPB / SY::kWaves][SY::kRefFloats];
};

// Next ROI of a persistent fit: the ROI's first wave takes it from the queue,
// the other waves of the ROI (N = 128: one block) read it through LDS.
template <class SY>
__device__ __forceinline__ int next_roi(int* queue, int& slot, int lane, int wr) {
    if constexpr (SY::kWaves == 1) {
        int v = 0;
        if (lane == 0) v = atomicAdd(queue, 1);
        return __builtin_amdgcn_readfirstlane(v);
    } else {
        if (lane == 0 && wr == 0) slot = atomicAdd(queue, 1);
        __syncthreads();
        const int v = slot;
        __syncthreads();
        return v;
    }
}

template <class SY, int R, int WPB>
__device__ __forceinline__ void fit_lm_body(const typename SY::Model& M, const FitArgs& A, LmLds<SY, WPB>& lds) {
    auto& T = lds.T;
    SY::load(T, M);
    __syncthreads();
    constexpr int RPL = SY::template RPL<R>;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int wr = w % SY::kWaves, slot = w / SY::kWaves;
    // static form: one ROI per wave slot; persistent form: ROIs from the queue
    int b = blockIdx.x * (WPB / SY::kWaves) + slot;
    if (A.queue) b = next_roi<SY>(A.queue, lds.q[slot], lane, wr);
    while (b < A.B) {
        float* scr = lds.S[slot];
        auto& red = lds.red[slot];
        LmShare& share = lds.share[slot];
        LmState& L = lds.L[w];
        const float* pix = A.pixels + (size_t)b * R * R;
        const float* vroi = A.varg ? A.varg + (size_t)b * R * R : nullptr;
        const float zoil = A.zoil ? A.zoil[b] : (float)A.zfocus;
        // Wave-uniform FP64 state lives in LDS between uses; registers are kept
        // for the transforms.  The step solves (F + S + lambda diag F) d = -g by
        // Cholesky (F Fisher matrix, S its secant correction); a system that is
        // not positive definite counts as a failed trial.  Damping follows
        // Nielsen's rule: on success
        // lambda *= max(1/3, 1-(2 rho-1)^3) with rho the ratio of actual to
        // predicted decrease; on failure lambda *= nu, nu *= 2.  A trial is
        // accepted when the likelihood does not rise; a rejected trial whose
        // predicted gain is below the FP32 synthesis noise of the likelihood
        // (noise_rel * A; measured about 1e-9 A) ends the fit as converged.
    #pragma unroll
        for (int p = 0; p < 5; p++) L.th[p] = A.init[(size_t)b * 5 + p];
        L.th[0] = fmax(L.th[0], 1.0);
        L.th[1] = fmax(L.th[1], 0.01);
    #pragma unroll
        for (int p = 0; p < 5; p++) L.tn[p] = L.th[p];
        L.nll = 0.0;
    #pragma unroll
        for (int k = 0; k < 15; k++) L.S[k] = 0.0;
        L.lambda = 1e-3;
        L.nu = 2.0;
        L.pred = 0.0;
        wave_lds_fence();
        const double dxmax = (double)A.xystep;
        int it = 0, tries = 0, conv = 0, nan = 0, nfield = 0;
        bool first = true;
        while (true) {
            bool eval_now = true, stop_noise = false, pre_conv = false;
            // The step is solved by the ROI's first wave.  With several waves per
            // ROI the others take its trial point, predicted decrease and
            // decision from LDS; they keep no Fisher matrix, secant term or
            // gradient of their own (only the first wave solves).
            if (!first && wr == 0) {
                double Lc[25], d[5];
    #pragma unroll
                for (int p = 0; p < 5; p++) {
    #pragma unroll
                    for (int q = 0; q <= p; q++) Lc[p * 5 + q] = hget(L.H, p, q) + hget(L.S, p, q);
                    Lc[p * 5 + p] += L.lambda * hget(L.H, p, p);
                    d[p] = -L.gr[p];
                }
                if (!chol5_ip(Lc)) {
                    L.lambda *= L.nu;
                    L.nu *= 2;
                    eval_now = false;
                } else {
                    chol5_solve(Lc, d);
                    if (tries == 0) {
                        // Converged at the point just accepted when the step the fit
                        // would take from it, almost undamped (lambda at most 1e-4,
                        // raised tenfold while the system is not positive definite)
                        // and clamped as trials are, is below tol in x, y, z and below
                        // 1% of the statistical precision in A and B.  A damped step
                        // would understate the distance along weakly determined
                        // directions (lambda diag(F) dominates their small eigenvalue);
                        // the almost undamped step is solved only when the damped one
                        // already passes, since it is the longer of the two.
                        double tc[5];
                        lm_clamp_trial(L.th, d, dxmax, A.zmin, A.zmax, tc);
                        bool small = lm_small_step(L, tc, A.tol);
                        if (small && L.lambda > 1e-4) {
                            double dc[5], lc = 1e-4;
                            bool have = false;
                            for (int k = 0; k < 8 && !have && lc < L.lambda; k++) {
    #pragma unroll
                                for (int p = 0; p < 5; p++) {
    #pragma unroll
                                    for (int q = 0; q <= p; q++) Lc[p * 5 + q] = hget(L.H, p, q) + hget(L.S, p, q);
                                    Lc[p * 5 + p] += lc * hget(L.H, p, p);
                                    dc[p] = -L.gr[p];
                                }
                                if (chol5_ip(Lc)) {
                                    chol5_solve(Lc, dc);
                                    have = true;
                                } else {
                                    lc *= 10.0;
                                }
                            }
                            if (have) {
                                lm_clamp_trial(L.th, dc, dxmax, A.zmin, A.zmax, tc);
                                small = lm_small_step(L, tc, A.tol);
                            }
                        }
                        if (small) {
                            pre_conv = true;
                            eval_now = false;
                        }
                    }
                }
                if (eval_now) {
                    // predicted decrease -(g'd + d'(F + S)d / 2); since
                    // (F + S + lambda diag F) d = -g this is
                    // -g'd / 2 + lambda sum F_pp d_p^2 / 2 (no second pass over F, S)
                    double gd = 0.0, dd = 0.0;
    #pragma unroll
                    for (int p = 0; p < 5; p++) {
                        gd += L.gr[p] * d[p];
                        dd += hget(L.H, p, p) * d[p] * d[p];
                    }
                    const double pred = -0.5 * gd + 0.5 * L.lambda * dd;
                    L.pred = pred;
                    double tn[5];
                    lm_clamp_trial(L.th, d, dxmax, A.zmin, A.zmax, tn);
    #pragma unroll
                    for (int p = 0; p < 5; p++) L.tn[p] = tn[p];
                }
                wave_lds_fence();
                if constexpr (SY::kWaves > 1) {
                    if (lane == 0) {
                        share.flag = (eval_now ? 0 : 1) | (pre_conv ? 2 : 0);
                        share.pred = L.pred;
    #pragma unroll
                        for (int p = 0; p < 5; p++) share.tn[p] = L.tn[p];
                    }
                }
            }
            if constexpr (SY::kWaves > 1) {
                if (!first) {
                    __syncthreads();
                    const int f = share.flag;
                    eval_now = (f & 1) == 0;
                    pre_conv = (f & 2) != 0;
                    if (wr != 0) {
                        L.pred = share.pred;
    #pragma unroll
                        for (int p = 0; p < 5; p++) L.tn[p] = share.tn[p];
                    }
                    // without an evaluation (and its barriers) before the next
                    // solve, the first wave would rewrite the step too early
                    if (!eval_now) __syncthreads();
                    wave_lds_fence();
                }
            }
            bool accepted = false;
            if (eval_now) {
                // Opaque copy of the lane id: keeps the compiler from hoisting the
                // hundreds of per-lane LDS addresses of one evaluation out of the
                // iteration loop (they would stay live across it and spill).
                int lane_it = lane;
                __asm__ __volatile__("" : "+v"(lane_it));
                float vB[RPL], vX[RPL], vY[RPL], vZ[RPL];
                SY::template lm_eval<R>(M, T, scr, lane_it, wr, (float)L.tn[2], (float)L.tn[3], (float)L.tn[4],
                                        zoil, vB, vX, vY, vZ);
                nfield += 2;
                float psf[RPL], n[RPL];
    #pragma unroll
                for (int t = 0; t < RPL; t++)
                    n[t] = SY::template valid<R>(lane, wr, t) ? pix[SY::template pix<R>(lane, wr, t)] : 0.f;
                const float Sn = normalise<SY, R>(lane, wr, red, vB, psf);
                const double tA = L.tn[0], tB = L.tn[1];
                const double nn = nll64<SY, R>(lane, wr, red, psf, n, tA, tB, vroi);
                if (nn != nn) nan = 1;
                accepted = first || nn <= L.nll;
                if (!accepted) {
                    // a gain predicted below the likelihood noise cannot be resolved
                    if (L.pred <= A.noise_rel * L.th[0]) stop_noise = true;
                    L.lambda *= L.nu;
                    L.nu *= 2;
                } else {
                    if (!first) {
                        const double step = fmax(fabs(L.tn[2] - L.th[2]),
                                                 fmax(fabs(L.tn[3] - L.th[3]), fabs(L.tn[4] - L.th[4])));
                        const double rho = L.pred > 0.0 ? (L.nll - nn) / L.pred : 0.0;
                        const double t3 = 2.0 * rho - 1.0;
                        L.lambda *= fmax(1.0 / 3.0, 1.0 - t3 * t3 * t3);
                        L.nu = 2.0;
                        // at a bound the clamped steps shrink to zero while the
                        // unclamped step does not: a step below tol ends the fit there
                        const bool at_bound = L.tn[4] <= A.zmin || L.tn[4] >= A.zmax || L.tn[0] <= 1.0 ||
                                              L.tn[1] <= 0.01;
                        if (step < A.tol && at_bound) conv = 1;
                    }
                    L.nll = nn;
                    // d psf = (sgn dv - psf * sum(sgn dv)) / S for the three tangents
                    float ss[3] = {0.f, 0.f, 0.f};
    #pragma unroll
                    for (int t = 0; t < RPL; t++) {
                        const float sg = vB[t] < 0.f ? -1.f : 1.f;
                        vX[t] *= sg;
                        vY[t] *= sg;
                        vZ[t] *= sg;
                        if (SY::template valid<R>(lane, wr, t)) {
                            ss[0] += vX[t];
                            ss[1] += vY[t];
                            ss[2] += vZ[t];
                        }
                    }
                    SY::template sumv<3>(ss, red, lane, wr);
                    const float sx = ss[0], sy = ss[1], sz = ss[2];
                    const float inv = 1.f / Sn;
                    double acc[20];
    #pragma unroll
                    for (int k = 0; k < 20; k++) acc[k] = 0.0;
    #pragma unroll
                    for (int t = 0; t < RPL; t++) {
                        if (!SY::template valid<R>(lane, wr, t)) continue;
                        const double mu = (double)psf[t] * tA + tB + scmos_term<SY, R>(vroi, lane, wr, t);
                        const double J[5] = {(double)psf[t], 1.0,
                                             tA * (double)((vX[t] - psf[t] * sx) * inv),
                                             tA * (double)((vY[t] - psf[t] * sy) * inv),
                                             tA * (double)((vZ[t] - psf[t] * sz) * inv)};
                        const double wg = 1.0 - (double)n[t] / mu, wh = 1.0 / mu;
                        int k = 5;
    #pragma unroll
                        for (int p = 0; p < 5; p++) {
                            acc[p] += wg * J[p];
    #pragma unroll
                            for (int q = p; q < 5; q++) acc[k++] += wh * J[p] * J[q];
                        }
                    }
                    SY::template sumv<20>(acc, red, lane, wr);
                    if (wr == 0) {
    #pragma unroll
                        for (int p = 0; p < 5; p++) {
                            L.gp[p] = L.gr[p];
                            L.gr[p] = acc[p];
                        }
    #pragma unroll
                        for (int k = 0; k < 15; k++) L.H[k] = acc[5 + k];
                        wave_lds_fence();
                        if (!first) lm_secant_update(L);
                    }
    #pragma unroll
                    for (int p = 0; p < 5; p++) L.th[p] = L.tn[p];
                }
                wave_lds_fence();
            }
            if (first) {
                first = false;
                if (A.max_iter <= 0) break;
                continue;
            }
            tries++;
            if (accepted || pre_conv || tries >= 12 || stop_noise) {
                if (!accepted) conv = 1;
                it++;
                tries = 0;
                if (conv || it >= A.max_iter) break;
            }
        }
        // CRLB of x, y, z from the Fisher matrix at the optimum (first wave)
        double cr[3] = {0.0, 0.0, 0.0};
        if (wr == 0) {
            double Lc[25];
    #pragma unroll
            for (int p = 0; p < 5; p++)
    #pragma unroll
                for (int q = 0; q <= p; q++) Lc[p * 5 + q] = hget(L.H, p, q);
            const bool pd = chol5_ip(Lc);
    #pragma unroll
            for (int c = 2; c < 5; c++) {
                double e[5];
    #pragma unroll
                for (int p = 0; p < 5; p++) e[p] = (p == c) ? 1.0 : 0.0;
                chol5_solve(Lc, e);
                cr[c - 2] = pd ? sqrt(e[c]) : (double)NAN;
            }
        }
        if (lane == 0 && wr == 0) {
    #pragma unroll
            for (int p = 0; p < 5; p++) A.theta[(size_t)b * 5 + p] = L.th[p];
            A.nll[b] = L.nll;
    #pragma unroll
            for (int c = 0; c < 3; c++) A.crlb[(size_t)b * 3 + c] = cr[c];
            A.iters[b] = it;
            A.status[b] = (conv ? kStConverged : kStMaxIter) | (nan ? kStNaN : 0) | (nfield << 8);
        }
        if (!A.queue) break;
        b = next_roi<SY>(A.queue, lds.q[slot], lane, wr);
    }
}

template <int R, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_fit_lm64(Model64 M, FitArgs A) {
    __shared__ LmLds<Synth64, WPB> lds;
    fit_lm_body<Synth64, R, WPB>(M, A, lds);
}

template <int R, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_fit_lm128(Model128 M, FitArgs A) {
    __shared__ LmLds<Synth128, WPB> lds;
    fit_lm_body<Synth128, R, WPB>(M, A, lds);
}

// One ROI per block and CU (the LDS holds one ROI's stage-1 buffers), two
// waves per SIMD: at most 256 registers per lane.
template <int R>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fit_lm128o(Model128 M, FitArgs A) {
    __shared__ LmLds<Synth128o, 8> lds;
    fit_lm_body<Synth128o, R, 8>(M, A, lds);
}

// Two waves per SIMD (four blocks per CU): at most 256 registers per lane.
template <int R>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fit_lm64p(Model64 M, FitArgs A) {
    __shared__ LmLds<Synth64p, 2> lds;
    fit_lm_body<Synth64p, R, 2>(M, A, lds);
}

template <int R>
__global__ __launch_bounds__(256) void k_fit_lm64q(Model64 M, FitArgs A) {
    __shared__ LmLds<Synth64q, 4> lds;
    fit_lm_body<Synth64q, R, 4>(M, A, lds);
}

// ---------------------------------------------------------------------------
// Reference coordinate descent (cpu/Localization.localize), as a state
// machine around one PSF-synthesis site.
// ---------------------------------------------------------------------------

// Laplace-expansion determinant and cofactor inverse (utils/Matrixe.java:140-246).
template <int NN>
__device__ __forceinline__ double det_lap(const double* a) {
    if constexpr (NN == 1) {
        return a[0];
    } else if constexpr (NN == 2) {
        return a[0] * a[3] - a[1] * a[2];
    } else {
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < NN; i++) {
            double sub[(NN - 1) * (NN - 1)];
            int r = 0;
#pragma unroll
            for (int rr = 1; rr < NN; rr++, r++) {
                int c = 0;
#pragma unroll
                for (int cc = 0; cc < NN; cc++) {
                    if (cc == i) continue;
                    sub[r * (NN - 1) + c++] = a[rr * NN + cc];
                }
            }
            sum += ((i & 1) ? -1.0 : 1.0) * a[i] * det_lap<NN - 1>(sub);
        }
        return sum;
    }
}

__device__ __forceinline__ void inverse_cofactor5(const double (&a)[25], double (&inv)[25]) {
    const double det = 1.0 / det_lap<5>(a);
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++) {
            double sub[16];
            int r = 0;
#pragma unroll
            for (int rr = 0; rr < 5; rr++) {
                if (rr == i) continue;
                int c = 0;
#pragma unroll
                for (int cc = 0; cc < 5; cc++) {
                    if (cc == j) continue;
                    sub[r * 4 + c++] = a[rr * 5 + cc];
                }
                r++;
            }
            inv[j * 5 + i] = (((i + j) & 1) ? -1.0 : 1.0) * det_lap<4>(sub) * det;
        }
}

enum : int { kRsInit, kRsAmpA, kRsAmpB, kRsPos, kRsCheck, kRsFisher, kRsDone };

template <class SY, int R, int WPB>
__device__ __forceinline__ void fit_ref_body(const typename SY::Model& M, const FitArgs& A, RefLds<SY, WPB>& lds) {
    auto& T = lds.T;
    SY::load(T, M);
    __syncthreads();
    constexpr int RPL = SY::template RPL<R>;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int wr = w % SY::kWaves, slot = w / SY::kWaves;
    // static form: one ROI per wave slot; persistent form: ROIs from the queue
    int b = blockIdx.x * (WPB / SY::kWaves) + slot;
    if (A.queue) b = next_roi<SY>(A.queue, lds.q[slot], lane, wr);
    while (b < A.B) {
        float* scr = lds.S[slot];
        auto& red = lds.red[slot];
        float n[RPL];
    #pragma unroll
        for (int t = 0; t < RPL; t++)
            n[t] = SY::template valid<R>(lane, wr, t) ? A.pixels[(size_t)b * R * R + SY::template pix<R>(lane, wr, t)] : 0.f;
        const float zoil = A.zoil ? A.zoil[b] : (float)A.zfocus;
        double tA = A.init[(size_t)b * 5 + 0], tB = A.init[(size_t)b * 5 + 1];
        double th[3] = {A.init[(size_t)b * 5 + 2], A.init[(size_t)b * 5 + 3], A.init[(size_t)b * 5 + 4]};
        const double h = .0005, eps = .0001, maxShift = (double)A.xystep / 2.;  // Localization.java:29-31,115

        // psf at the current position (cur) and the last synthesised one (tmp)
        float cur[RPL], tmp[RPL], hi[RPL], dd[3][RPL];
        double pc[3] = {0, 0, 0};
        bool hc = false;
        double req[3] = {0, 0, 0};
        bool want = false;

        int st = kRsInit, which = 0, k = 0, it = 0, nbIter = A.max_iter, nfield = 0;
        int brk = 0, ok = 1, nan = 0, conv = 0, initial = 1;
        double likelihood = 0, save = 0, l1 = 0, l3 = 0, grad = 0, prevLik = 0;
        double oldX = 1.7976931348623157e308, oldY = oldX, oldZ = oldX;
        auto at_theta = [&]() { return hc && pc[0] == th[0] && pc[1] == th[1] && pc[2] == th[2]; };
        const float* vroi = A.varg ? A.varg + (size_t)b * R * R : nullptr;
        auto L = [&](const float (&ps)[RPL], double a, double bb) {
            return nll64<SY, R>(lane, wr, red, ps, n, a, bb, vroi);
        };
        // min_amp (929-977 for A with h*1000, 997-1058 for B with h*10), on the cached psf
        auto min_amp = [&](bool isB) {
            const double hh = isB ? h * 10 : h * 1000;
            double& t = isB ? tB : tA;
            double sv = t;
            double a1, a2, a3;
            if (!isB) { a1 = L(cur, tA - hh, tB); a2 = L(cur, tA, tB); a3 = L(cur, tA + hh, tB); }
            else { a1 = L(cur, tA, tB - hh); a2 = L(cur, tA, tB); a3 = L(cur, tA, tB + hh); t = sv; }
            double gd;
            if (fabs((a3 + a1 - 2 * a2) / (hh * hh)) == 0) gd = (a3 - a1) * (2 * hh);
            else gd = ((a3 - a1) / (2 * hh)) / fabs((a3 + a1 - 2 * a2) / (hh * hh));
            for (double gamma = 1; gamma > .02; gamma /= 10) {
                t -= gamma * gd;
                const double lk = L(cur, tA, tB);
                if (lk < a2) { a2 = lk; sv = t; }
                else t = sv;
            }
            return L(cur, tA, tB);
        };

        while (st != kRsDone) {
            if (want) {
                float v[RPL];
                SY::template eval<R>(M, T, scr, lane, wr, (float)req[0], (float)req[1], (float)req[2], zoil, v);
                normalise<SY, R>(lane, wr, red, v, tmp);
                want = false;
                nfield++;
            }
            if (st == kRsInit || st == kRsAmpA || st == kRsAmpB) {
                if (!at_theta()) {
                    if (k == 1) {  // tmp holds psf(theta)
    #pragma unroll
                        for (int t = 0; t < RPL; t++) cur[t] = tmp[t];
    #pragma unroll
                        for (int c = 0; c < 3; c++) pc[c] = th[c];
                        hc = true;
                        k = 0;
                    } else {
    #pragma unroll
                        for (int c = 0; c < 3; c++) req[c] = th[c];
                        want = true;
                        k = 1;
                        continue;
                    }
                }
                if (st == kRsInit) {
                    tA = fmax(1.0, tA);
                    tB = fmax(.01, tB);
                    likelihood = L(cur, tA, tB);
                    st = kRsAmpA;
                } else if (st == kRsAmpA) {
                    likelihood = min_amp(false);
                    tA = fmax(1.0, tA);
                    tB = fmax(.01, tB);
                    st = kRsAmpB;
                } else {
                    likelihood = min_amp(true);
                    tA = fmax(1.0, tA);
                    tB = fmax(.01, tB);
                    if (initial) {
                        initial = 0;
                        it = 0;
                        if (A.max_iter <= 0) { st = kRsFisher; k = 0; continue; }
                        st = kRsPos; which = 0; k = 0;
                    } else {
                        st = kRsCheck;
                    }
                }
                continue;
            }
            if (st == kRsPos) {
                // minimizeZ / minimizeX / minimizeY (672-906); which: 0 = z, 1 = x, 2 = y
                const int ci = which == 0 ? 2 : which - 1;
                if (k == 0) {
                    save = th[ci];
    #pragma unroll
                    for (int c = 0; c < 3; c++) req[c] = th[c];
                    req[ci] = th[ci] - h;
                    want = true;
                    k = 1;
                } else if (k == 1) {
                    l1 = L(tmp, tA, tB);
    #pragma unroll
                    for (int c = 0; c < 3; c++) req[c] = th[c];
                    req[ci] = th[ci] + h;
                    want = true;
                    k = 2;
                } else if (k == 2) {
                    l3 = L(tmp, tA, tB);
                    const double l2 = likelihood;
                    if (fabs((l3 + l1 - 2 * l2) / (h * h)) == 0) grad = (l3 - l1) * (2 * h);
                    else grad = ((l3 - l1) / (2 * h)) / fabs((l3 + l1 - 2 * l2) / (h * h));
                    if (grad > maxShift) grad = maxShift;
                    else if (grad < -maxShift) grad = -maxShift;
                    prevLik = likelihood;
                    th[ci] -= 1.0 * grad;
    #pragma unroll
                    for (int c = 0; c < 3; c++) req[c] = th[c];
                    want = true;
                    k = 3;
                } else {
                    const double lik = L(tmp, tA, tB);
                    bool done = false;
                    if (lik < likelihood) {
                        prevLik = lik;
    #pragma unroll
                        for (int t = 0; t < RPL; t++) cur[t] = tmp[t];
    #pragma unroll
                        for (int c = 0; c < 3; c++) pc[c] = th[c];
                        hc = true;
                        done = true;
                    } else {
                        th[ci] = save;
                        if (k == 3) {
                            th[ci] -= (1.0 / 10) * grad;
    #pragma unroll
                            for (int c = 0; c < 3; c++) req[c] = th[c];
                            want = true;
                            k = 4;
                        } else {
                            done = true;
                        }
                    }
                    if (done) {
                        likelihood = prevLik;
                        if (which == 0) {
                            if (th[2] < A.zmin) th[2] = A.zmin;
                            if (th[2] > A.zmax) th[2] = A.zmax;
                        }
                        which++;
                        k = 0;
                        if (which == 3) st = kRsAmpA;
                    }
                }
                continue;
            }
            if (st == kRsCheck) {
                const double distance = sqrt((th[0] - oldX) * (th[0] - oldX) + (th[1] - oldY) * (th[1] - oldY) +
                                             (th[2] - oldZ) * (th[2] - oldZ));
                int breaking = 0;
                if (it > 4) breaking = !(tA > A.thr);
                if (breaking) brk = 1;
                if (distance < eps || breaking) {
                    nbIter = it + 1;
                    conv = !breaking;
                    st = kRsFisher;
                    k = 0;
                    continue;
                }
                if (likelihood != likelihood) {
                    ok = 0;
                    nan = 1;
                    st = kRsDone;
                    continue;
                }
                oldX = th[0]; oldY = th[1]; oldZ = th[2];
                it++;
                if (it >= A.max_iter) {
                    nbIter = A.max_iter;
                    st = kRsFisher;
                    k = 0;
                } else {
                    st = kRsPos; which = 0; k = 0;
                }
                continue;
            }
            if (st == kRsFisher) {
                // computeFisher (317-403) with the reference's steps: probes k = 0..5
                // at x+h, x-h, y+h, y-h, z+h, z-h; then the centre (cached psf).
                if (k >= 1 && k <= 6) {
                    const int q = (k - 1) / 2;
                    if ((k - 1) % 2 == 0) {
    #pragma unroll
                        for (int t = 0; t < RPL; t++) hi[t] = tmp[t];
                    } else {
    #pragma unroll
                        for (int t = 0; t < RPL; t++) {
                            const double f2 = (double)hi[t] * tA + tB, f0 = (double)tmp[t] * tA + tB;
                            const float d = (float)((f2 - f0) / (2 * fabs(h)));
                            dd[0][t] = q == 0 ? d : dd[0][t];
                            dd[1][t] = q == 1 ? d : dd[1][t];
                            dd[2][t] = q == 2 ? d : dd[2][t];
                        }
                    }
                }
                if (k < 6) {
                    const int q = k / 2;
                    const double sgnh = (k % 2 == 0) ? h : -h;
                    req[0] = th[0] + (q == 0 ? sgnh : 0.0);
                    req[1] = th[1] + (q == 1 ? sgnh : 0.0);
                    req[2] = th[2] + (q == 2 ? sgnh : 0.0);
                    want = true;
                    k++;
                    continue;
                }
                if (!at_theta()) {
                    if (k == 7) {
    #pragma unroll
                        for (int t = 0; t < RPL; t++) cur[t] = tmp[t];
    #pragma unroll
                        for (int c = 0; c < 3; c++) pc[c] = th[c];
                        hc = true;
                    } else {
    #pragma unroll
                        for (int c = 0; c < 3; c++) req[c] = th[c];
                        want = true;
                        k = 7;
                        continue;
                    }
                }
                double acc[15];
    #pragma unroll
                for (int kk = 0; kk < 15; kk++) acc[kk] = 0.0;
    #pragma unroll
                for (int t = 0; t < RPL; t++) {
                    if (!SY::template valid<R>(lane, wr, t)) continue;
                    // the reference's weights 1/f1 carry no sCMOS term (computeFisher, 317-403)
                    const double f1 = (double)cur[t] * tA + tB;
                    // x, y, z, then A (the reference differentiates the model f1) and B
                    const double J[5] = {(double)dd[0][t], (double)dd[1][t], (double)dd[2][t], f1, 1.0};
                    int kk = 0;
    #pragma unroll
                    for (int p = 0; p < 5; p++)
    #pragma unroll
                        for (int q = p; q < 5; q++) acc[kk++] += (1 / f1) * J[p] * J[q];
                }
                SY::template sumv<15>(acc, red, lane, wr);
                double F[25], inv[25];
                int kk = 0;
    #pragma unroll
                for (int p = 0; p < 5; p++)
    #pragma unroll
                    for (int q = p; q < 5; q++) { F[p * 5 + q] = acc[kk]; F[q * 5 + p] = acc[kk]; kk++; }
                likelihood = L(cur, tA, tB);
                inverse_cofactor5(F, inv);
                if (lane == 0 && wr == 0) {
    #pragma unroll
                    for (int c = 0; c < 3; c++) A.crlb[(size_t)b * 3 + c] = sqrt(inv[c * 5 + c]);
                }
                st = kRsDone;
                continue;
            }
        }
        if (lane == 0 && wr == 0) {
            if (!ok) {
    #pragma unroll
                for (int c = 0; c < 3; c++) A.crlb[(size_t)b * 3 + c] = (double)NAN;
            }
            A.theta[(size_t)b * 5 + 0] = tA;
            A.theta[(size_t)b * 5 + 1] = tB;
            A.theta[(size_t)b * 5 + 2] = th[0];
            A.theta[(size_t)b * 5 + 3] = th[1];
            A.theta[(size_t)b * 5 + 4] = th[2];
            A.nll[b] = likelihood;
            A.iters[b] = nbIter;
            A.status[b] = (conv ? kStConverged : 0) | (nbIter == A.max_iter && !conv ? kStMaxIter : 0) |
                          (nan ? kStNaN : 0) | (brk ? kStLowPhoton : 0) | (nfield << 8);
        }
        if (!A.queue) break;
        b = next_roi<SY>(A.queue, lds.q[slot], lane, wr);
    }
}

template <int R, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_fit_ref64(Model64 M, FitArgs A) {
    __shared__ RefLds<Synth64, WPB> lds;
    fit_ref_body<Synth64, R, WPB>(M, A, lds);
}

template <int R, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_fit_ref128(Model128 M, FitArgs A) {
    __shared__ RefLds<Synth128, WPB> lds;
    fit_ref_body<Synth128, R, WPB>(M, A, lds);
}

template <int R>
__global__ __launch_bounds__(256) void k_fit_ref64q(Model64 M, FitArgs A) {
    __shared__ RefLds<Synth64q, 4> lds;
    fit_ref_body<Synth64q, R, 4>(M, A, lds);
}

}  // namespace zola

#pragma once
// zola_kernels_impl.hpp -- gfx950 kernels of the ZOLA-3D localization hot path and
// their launch templates; instantiated per ROI size in zola_inst_r*.hip.
#include <hip/hip_runtime.h>

#include "zola_psf64.hpp"
#include "zola_psf128.hpp"

namespace zola {

// ---------------------------------------------------------------------------
// Batched PSF synthesis: one wave per PSF, WPB waves per block sharing the
// pupil tables in LDS.
// ---------------------------------------------------------------------------
template <int R, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_psf_forward64(Model64 M, int B,
                                                            const float* __restrict__ params,
                                                            float* __restrict__ out) {
    __shared__ BlockTables64 T;
    __shared__ WaveScratch64 S[WPB];
    load_tables64(T, M);
    __syncthreads();
    constexpr int G = RoiLayout<R>::G, RPL = RoiLayout<R>::RPL;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * WPB + w;
    if (b >= B) return;
    const float x = params[b * 4 + 0], y = params[b * 4 + 1];
    const float z = params[b * 4 + 2], zoil = params[b * 4 + 3];
    stage1_64(T, S[w], lane, 0, M.kappa * x, M.kappa * y, z, zoil, M.kappa);
    wave_lds_fence();
    float acc[R];
    stage2_E64<R>(T.h3[lane], S[w], lane, acc);
    float o[RPL];
    column_blur64<R>(T, S[w], lane, acc, o);
    const int j = lane % R, g = lane / R;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        o[t] = fabsf(o[t]);
        const int i = g * RPL + t;
        if (g < G && i < R) s += o[t];
    }
    s = wave_sum(s);
    const float inv = 1.f / s;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        const int i = g * RPL + t;
        if (g < G && i < R) out[(size_t)b * R * R + i * R + j] = o[t] * inv;
    }
}

constexpr int kWPB = 8;

template <int R>
hipError_t launch_psf_forward_r(const Model64& M, int B, const float* params, float* out,
                                       hipStream_t st) {
    dim3 grid((B + kWPB - 1) / kWPB), block(64 * kWPB);
    hipLaunchKernelGGL((k_psf_forward64<R, kWPB>), grid, block, 0, st, M, B, params, out);
    return hipGetLastError();
}

}  // namespace zola


// ---------------------------------------------------------------------------
// Fits and their initial values
// ---------------------------------------------------------------------------
#include "zola_fit64.hpp"
#include "zola_adjoint64.hpp"
#include "zola_kernels.h"

namespace zola {

// Initial values for an isolated ROI: best zero-shift normalised correlation
// against the template stack (z refined by the reference's parabola,
// process/cpu/Predetection.java:529-531,589-593), then getAandB with that
// template (process/cpu/LocalizationPipeline.java:1119-1162).
// Per-template sums of the correlation (sum t, and sum t^2 - (sum t)^2 / R^2),
// computed once per template stack with the lane layout and reduction order of
// k_fit_init64.  One wave per template.
template <int R, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_template_stats64(int nZ, const float* __restrict__ tmpl,
                                                               double* __restrict__ tstat) {
    constexpr int RPL = RoiLayout<R>::RPL;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * WPB + w;
    if (r >= nZ) return;
    const int j = lane % R, g = lane / R;
    double st = 0, stt = 0;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        if (!Px<R>::valid(lane, t)) continue;
        const double tv = tmpl[(size_t)r * R * R + (g * RPL + t) * R + j];
        st += tv;
        stt += tv * tv;
    }
    st = wave_sum(st);
    stt = wave_sum(stt);
    if (lane == 0) {
        tstat[2 * r] = st;
        tstat[2 * r + 1] = stt - st * st / ((double)R * R);
    }
}

template <int R, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_fit_init64(int B, const float* __restrict__ pixels,
                                                         int nZ, const float* __restrict__ tmpl,
                                                         const double* __restrict__ tstat,
                                                         double zmin_t, double zstep_t,
                                                         double* __restrict__ init) {
    constexpr int RPL = RoiLayout<R>::RPL;
    // templates are staged in LDS in chunks that the block's WPB ROIs share
    constexpr int kChunk = 8192;  // floats per chunk (32 KB)
    constexpr int NC = kChunk / (R * R) > 0 ? kChunk / (R * R) : 1;
    __shared__ float ts[NC * R * R];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * WPB + w;
    // waves past the batch end stay for the block barriers of the staging
    const bool active = b < B;
    const int j = lane % R, g = lane / R;
    float n[RPL];
    double sn = 0, snn = 0;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        n[t] = active && Px<R>::valid(lane, t) ? pixels[(size_t)b * R * R + (g * RPL + t) * R + j] : 0.f;
        sn += n[t];
        snn += (double)n[t] * n[t];
    }
    {
        double s2[2] = {sn, snn};
        wave_sum_rows<2>(s2);
        sn = s2[0];
        snn = s2[1];
    }
    const double P = (double)R * R;
    const double vn = snn - sn * sn / P;
    auto corr = [&](int r) -> double {
        double stn = 0;
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            if (!Px<R>::valid(lane, t)) continue;
            const double tv = tmpl[(size_t)r * R * R + (g * RPL + t) * R + j];
            stn += tv * n[t];
        }
        {
            double a1[1] = {stn};
            wave_sum_rows<1>(a1);
            stn = a1[0];
        }
        const double st = tstat[2 * r], vt = tstat[2 * r + 1];
        return (stn - st * sn / P) / sqrt(vt * vn);
    };
    int best = 0;
    double cbest = -1e300;
    for (int r0 = 0; r0 < nZ; r0 += NC) {
        const int nc = nZ - r0 < NC ? nZ - r0 : NC;
        __syncthreads();  // the previous chunk is no longer read
        for (int k = threadIdx.x; k < nc * R * R; k += 64 * WPB) ts[k] = tmpl[(size_t)r0 * R * R + k];
        __syncthreads();
        if (!active) continue;
        for (int rc = 0; rc < nc; rc++) {
            // corr(r0 + rc) with the template read from LDS
            double stn = 0;
#pragma unroll
            for (int t = 0; t < RPL; t++) {
                if (!Px<R>::valid(lane, t)) continue;
                const double tv = ts[rc * R * R + (g * RPL + t) * R + j];
                stn += tv * n[t];
            }
            {
                double a1[1] = {stn};
                wave_sum_rows<1>(a1);
                stn = a1[0];
            }
            const int r = r0 + rc;
            const double st = tstat[2 * r], vt = tstat[2 * r + 1];
            const double c = (stn - st * sn / P) / sqrt(vt * vn);
            if (c > cbest) { cbest = c; best = r; }
        }
    }
    if (!active) return;
    double z0 = zmin_t + best * zstep_t;
    if (best > 1 && best < nZ - 1) {
        const double ya = corr(best - 1), yb = cbest, yc = corr(best + 1);
        const double xa = -zstep_t, xb = 0.0, xc = zstep_t;
        const double a = (yc - ya) / ((xc - xa) * (xc - xb)) - (yb - ya) / ((xb - xa) * (xc - xb));
        const double bq = ((yb - ya) / (xb - xa)) - a * (xb + xa);
        z0 += -bq / (2 * a);
    }
    // getAandB(patch, template, 0.1)
    float tmax = -3.0e38f;
#pragma unroll
    for (int t = 0; t < RPL; t++)
        if (Px<R>::valid(lane, t)) tmax = fmaxf(tmax, tmpl[(size_t)best * R * R + (g * RPL + t) * R + j]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) tmax = fmaxf(tmax, __shfl_xor(tmax, o, 64));
    const float thr = tmax * 0.1f;
    double sb = 0, cb = 0, sa = 0, ca = 0;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        if (!Px<R>::valid(lane, t)) continue;
        const float tv = tmpl[(size_t)best * R * R + (g * RPL + t) * R + j];
        if (tv < thr) { sb += n[t]; cb += 1; } else { sa += n[t]; ca += 1; }
    }
    {
        double ab[4] = {sb, cb, sa, ca};
        wave_sum_rows<4>(ab);
        sb = ab[0];
        cb = ab[1];
        sa = ab[2];
        ca = ab[3];
    }
    const double Bv = cb != 0 ? sb / cb : 0.0;
    const double Av = sa - ca * Bv;
    if (lane == 0) {
        init[(size_t)b * 5 + 0] = Av;
        init[(size_t)b * 5 + 1] = Bv;
        init[(size_t)b * 5 + 2] = 0.0;
        init[(size_t)b * 5 + 3] = 0.0;
        init[(size_t)b * 5 + 4] = z0;
    }
}

constexpr int kWPB_LM = 4;
constexpr int kWPB_REF = 4;

// Grid of a fit launch: one ROI per wave slot, or, with a queue (persistent
// form), the resident blocks only (ROIs taken from the queue, zeroed first).
template <class K>
int fit_grid(K kernel, int threads, int roi_per_block, const FitArgs& A, hipStream_t st, hipError_t& err) {
    const int need = (A.B + roi_per_block - 1) / roi_per_block;
    err = hipSuccess;
    if (!A.queue) return need;
    const int resident = zola_persistent_blocks(reinterpret_cast<const void*>(kernel), threads);
    err = zola_launch_zero_ints(A.queue, 1, st);
    return resident > 0 && resident < need ? resident : need;
}

template <int R>
hipError_t launch_fit_r(const Model64& M, int mode, const FitArgs& A, hipStream_t st) {
    hipError_t e;
    if (mode == 1) {
        const int g = fit_grid(k_fit_lm64<R, kWPB_LM>, 64 * kWPB_LM, kWPB_LM, A, st, e);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fit_lm64<R, kWPB_LM>), dim3(g), dim3(64 * kWPB_LM), 0, st, M, A);
    } else {
        const int g = fit_grid(k_fit_ref64<R, kWPB_REF>, 64 * kWPB_REF, kWPB_REF, A, st, e);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fit_ref64<R, kWPB_REF>), dim3(g), dim3(64 * kWPB_REF), 0, st, M, A);
    }
    return hipGetLastError();
}

// N = 64 with four waves (one block) per ROI, see zola_psf64q.hpp.
template <int R>
hipError_t launch_fit64q_r(const Model64& M, int mode, const FitArgs& A, hipStream_t st) {
    hipError_t e;
    if (mode == 1) {
        const int g = fit_grid(k_fit_lm64q<R>, 256, 1, A, st, e);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fit_lm64q<R>), dim3(g), dim3(256), 0, st, M, A);
    } else {
        const int g = fit_grid(k_fit_ref64q<R>, 256, 1, A, st, e);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fit_ref64q<R>), dim3(g), dim3(256), 0, st, M, A);
    }
    return hipGetLastError();
}

// N = 64 with two waves (one block) per ROI, LM form, see zola_psf64p.hpp.
template <int R>
hipError_t launch_fit64p_r(const Model64& M, const FitArgs& A, hipStream_t st) {
    hipError_t e;
    const int g = fit_grid(k_fit_lm64p<R>, 128, 1, A, st, e);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fit_lm64p<R>), dim3(g), dim3(128), 0, st, M, A);
    return hipGetLastError();
}

// N = 128: four waves per PSF (one block), see zola_psf128.hpp.
template <int R>
__global__ __launch_bounds__(256) void k_psf_forward128(Model128 M, int B, const float* __restrict__ params,
                                                        float* __restrict__ out) {
    constexpr int RQ = Roi128<R>::RQ;
    __shared__ Buf128 S[2];
    __shared__ float red[4];
    const int wr = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int b = blockIdx.x;
    if (b >= B) return;
    const float x = params[b * 4 + 0], y = params[b * 4 + 1];
    const float z = params[b * 4 + 2], zoil = params[b * 4 + 3];
    float o[RQ], d0[RQ], d1[RQ], d2[RQ];
    eval128<R, false>(M, S, lane, wr, x, y, z, zoil, o, d0, d1, d2);
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < RQ; t++) {
        o[t] = fabsf(o[t]);
        if (lane < R && wr * RQ + t < R) s += o[t];
    }
    s = wave_sum(s);
    if (lane == 0) red[wr] = s;
    __syncthreads();
    const float inv = 1.f / (((red[0] + red[1]) + red[2]) + red[3]);
    if (lane < R) {
#pragma unroll
        for (int t = 0; t < RQ; t++)
            if (wr * RQ + t < R) out[(size_t)b * R * R + (wr * RQ + t) * R + lane] = o[t] * inv;
    }
}

template <int R>
hipError_t launch_psf_forward128_r(const Model128& M, int B, const float* params, float* out, hipStream_t st) {
    hipLaunchKernelGGL((k_psf_forward128<R>), dim3(B), dim3(256), 0, st, M, B, params, out);
    return hipGetLastError();
}

// Fits: one ROI (four waves) per block.
template <int R>
hipError_t launch_fit128_r(const Model128& M, int mode, const FitArgs& A, hipStream_t st) {
    hipError_t e;
    if (mode == 1) {
        const int g = fit_grid(k_fit_lm128<R, 4>, 256, 1, A, st, e);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fit_lm128<R, 4>), dim3(g), dim3(256), 0, st, M, A);
    } else {
        const int g = fit_grid(k_fit_ref128<R, 4>, 256, 1, A, st, e);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fit_ref128<R, 4>), dim3(g), dim3(256), 0, st, M, A);
    }
    return hipGetLastError();
}

// LM form, eight waves (one block) per ROI.
template <int R>
hipError_t launch_fit128o_r(const Model128& M, const FitArgs& A, hipStream_t st) {
    hipError_t e;
    const int g = fit_grid(k_fit_lm128o<R>, 512, 1, A, st, e);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_fit_lm128o<R>), dim3(g), dim3(512), 0, st, M, A);
    return hipGetLastError();
}

template <int R>
hipError_t launch_tstats_r(int nZ, const float* tmpl, double* tstat, hipStream_t st) {
    constexpr int W = 4;
    hipLaunchKernelGGL((k_template_stats64<R, W>), dim3((nZ + W - 1) / W), dim3(64 * W), 0, st, nZ, tmpl, tstat);
    return hipGetLastError();
}

template <int R>
hipError_t launch_init_r(int B, const float* pixels, int nZ, const float* tmpl, const double* tstat,
                         double zmin_t, double zstep_t, double* init, hipStream_t st) {
    constexpr int W = 16;
    dim3 grid((B + W - 1) / W), block(64 * W);
    hipLaunchKernelGGL((k_fit_init64<R, W>), grid, block, 0, st, B, pixels, nZ, tmpl, tstat, zmin_t,
                       zstep_t, init);
    return hipGetLastError();
}

// Waves per PSF of the adjoint (4 or 8: 33.0 and 32.4 us for the 205 PSFs of a
// 31^2 calibration evaluation; four leave the other CUs' SIMDs free).
constexpr int kAdjWaves = 4;

template <int R>
hipError_t launch_adjoint_r(const Model64& M, const AdjArgs& A, hipStream_t st) {
    // one PSF per block of kAdjWaves waves
    constexpr int NW = kAdjWaves;
    dim3 grid(A.B), block(64 * NW);
    if (A.pixels && A.dsig) hipLaunchKernelGGL((k_psf_adjoint64<R, NW, true, true>), grid, block, 0, st, M, A);
    else if (A.pixels) hipLaunchKernelGGL((k_psf_adjoint64<R, NW, true>), grid, block, 0, st, M, A);
    else hipLaunchKernelGGL((k_psf_adjoint64<R, NW, false>), grid, block, 0, st, M, A);
    return hipGetLastError();
}

}  // namespace zola

// zola_adjoint64.hpp -- reverse-mode (adjoint) pass of the PSF model on the
// 64x64 pupil grid, one wave per PSF.
//
// Given w = dL/dpsf on the ROI, the wave recomputes the forward pass (E kept
// in VGPRs), then runs every stage backwards:
//   psf = m/S, m = |Bl|          -> dL/dBl = sign(Bl) (w/S - sum(w m)/S^2)
//   Bl = h (x) I (separable)     -> dL/dI  = h^T (x) dL/dBl  (same kernel, transposed sums)
//   I = |E|^2                    -> dL/dE  = 2 dL/dI E
//   E = IDFT64(P) (32x32 support)-> dL/dP  = DFT64(dL/dE) cropped to the support
//   P = A exp(i q)               -> dL/dq  = A (-sin q dL/dPr + cos q dL/dPi)
// and reduces dL/dq against dq/dx, dq/dy, dq/dz, dq/dzoil.  dL/dq itself is
// written per PSF so that dL/da_j = sum_p Z_j(p) dL/dq(p) can be formed for
// the whole batch by one projection kernel (Zernike basis in FP64).
#pragma once
#include "zola_fit64.hpp"

namespace zola {

struct AdjArgs {
    int B;
    const float* params;   // B x 4 [x, y, z, zoil]
    const float* w;        // B x R x R (adjoint form: dL/dpsf)
    float* dparams;        // B x 4
    float* qbar;           // B x 1024 (window layout), may be null
    // calibration form: w = dNLL/dpsf is formed from the data
    const float* pixels;   // B x R x R photons
    const double* amp;     // B photons per PSF
    const double* bg;      // B background per pixel
    double* nll;           // B  sum(mu - n ln mu), mu = amp psf + bg (clamped at 1e-8)
    double* damp;          // B  dNLL/damp = sum psf (1 - n/mu)
    double* dbg;           // B  dNLL/dbg  = sum (1 - n/mu)
};

// Stage 2 of E keeping E (row a = lane) plus the forward row blur.
template <int R>
__device__ __forceinline__ void stage2_E_keep64(float h1reg, const WaveScratch64& S, int lane,
                                                float (&acc)[R], float (&Er)[2][32],
                                                float (&Ei)[2][32]) {
#pragma unroll
    for (int j = 0; j < R; j++) acc[j] = 0.f;
    static_for<0, 2>([&](auto hpc) {
        constexpr int HP = decltype(hpc)::value;
        float re[32], im[32];
        stage2_half64<HP>(S, lane, re, im);
        float img[1][32];
        static_for<0, 32>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            Er[HP][m] = re[fidx<32>(m)];
            Ei[HP][m] = im[fidx<32>(m)];
            img[0][m] = Er[HP][m] * Er[HP][m] + Ei[HP][m] * Ei[HP][m];
        });
        const float hr[1] = {h1reg};
        float (&a1)[1][R] = *reinterpret_cast<float (*)[1][R]>(&acc);
        row_blur_half<R, HP, 1>(hr, img, a1);
    });
}

template <int R, int WPB, bool CALIB>
__global__ __launch_bounds__(64 * WPB) void k_psf_adjoint64(Model64 M, AdjArgs A) {
    struct Lds {
        BlockTables64 T;
        WaveScratch64 S[WPB];
    };
    __shared__ Lds lds;
    BlockTables64& T = lds.T;
    load_tables64(T, M);
    __syncthreads();
    constexpr int G = RoiLayout<R>::G, RPL = RoiLayout<R>::RPL;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * WPB + w;
    if (b >= A.B) return;
    WaveScratch64& S = lds.S[w];
    const float x = A.params[b * 4 + 0], y = A.params[b * 4 + 1];
    const float z = A.params[b * 4 + 2], zoil = A.params[b * 4 + 3];
    const float X = M.kappa * x, Y = M.kappa * y;
    const float h1reg = T.h3[lane];

    // ---- forward, keeping E ----
    stage1_64(T, S, lane, 0, X, Y, z, zoil, M.kappa);
    wave_lds_fence();
    float Er[2][32], Ei[2][32];
    float bl[RPL];
    {
        float acc[R];
        stage2_E_keep64<R>(h1reg, S, lane, acc, Er, Ei);
        column_blur64<R>(T, S, lane, acc, bl);
    }
    const int j = lane % R, g = lane / R;
    float wv[RPL];
    float ssum = 0.f, wm = 0.f;
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        const bool ok = g < G && g * RPL + t < R;
        if (ok) ssum += fabsf(bl[t]);
    }
    ssum = wave_sum(ssum);
    if constexpr (CALIB) {
        // ZernikePhaseRetrieval.Model.computeLikelihood (1859-1867): mu = psf A + bg,
        // clamped at 1e-8; w = dNLL/dpsf = A (1 - n/mu)
        const double av = A.amp[b], bv = A.bg[b];
        double nl = 0.0, da = 0.0, db = 0.0;
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            const bool ok = g < G && g * RPL + t < R;
            wv[t] = 0.f;
            if (ok) {
                const double ps = (double)(fabsf(bl[t]) / ssum);
                const double n = A.pixels[(size_t)b * R * R + (g * RPL + t) * R + j];
                double mu = ps * av + bv;
                if (mu <= 0) mu = 1e-8;
                nl += mu - n * log(mu);
                const double r = 1.0 - n / mu;
                da += ps * r;
                db += r;
                wv[t] = (float)(av * r);
            }
        }
        nl = wave_sum(nl);
        da = wave_sum(da);
        db = wave_sum(db);
        if (lane == 0) {
            A.nll[b] = nl;
            A.damp[b] = da;
            A.dbg[b] = db;
        }
    } else {
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            const bool ok = g < G && g * RPL + t < R;
            wv[t] = ok ? A.w[(size_t)b * R * R + (g * RPL + t) * R + j] : 0.f;
        }
    }
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        const bool ok = g < G && g * RPL + t < R;
        if (ok) wm += wv[t] * fabsf(bl[t]);
    }
    wm = wave_sum(wm);
    const float inv = 1.f / ssum, c2 = wm * inv * inv;

    // ---- dL/dBl at the ROI pixels, stored as Bt[i][j] in LDS ----
    float* Bt = S.re;             // R*R floats
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        const bool ok = g < G && g * RPL + t < R;
        const float sg = bl[t] < 0.f ? -1.f : 1.f;
        if (ok) Bt[(g * RPL + t) * R + j] = sg * (wv[t] * inv - c2);
    }
    wave_lds_fence();
    // column adjoint for this lane's grid row a: u(j) = sum_i h1[(i - R/2 - a) & 63] Bt[i][j]
    float u[R];
    {
        float c[R];
#pragma unroll
        for (int i = 0; i < R; i++) c[i] = T.h3[64 + ((i - R / 2 - lane) & 63)];
#pragma unroll
        for (int jj = 0; jj < R; jj++) {
            float acc = 0.f;
#pragma unroll
            for (int i = 0; i < R; i++) acc += c[i] * Bt[i * R + jj];
            u[jj] = acc;
        }
    }
    // row adjoint along b, dL/dE = 2 dL/dI E, then the adjoint of the
    // stage-2 IDFT: S1bar(a, n) = sum_b dL/dE(a,b) exp(-2 pi i n b / 64)
    float sr[32], si[32];
#pragma unroll
    for (int k = 0; k < 32; k++) {
        sr[k] = 0.f;
        si[k] = 0.f;
    }
    static_for<0, 2>([&](auto hpc) {
        constexpr int HP = decltype(hpc)::value;
        float er[32], ei[32];
        static_for<0, 32>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr int bcol = 2 * m + HP;
            float ib = 0.f;
            static_for<0, R>([&](auto jc) {
                constexpr int jj = decltype(jc)::value;
                const float hc = __int_as_float(
                    __builtin_amdgcn_readlane(__float_as_int(h1reg), (jj - R / 2 - bcol) & 63));
                ib += hc * u[jj];
            });
            er[m] = 2.f * ib * Er[HP][m];
            ei[m] = 2.f * ib * Ei[HP][m];
        });
        fft_dif<32, -1>(er, ei);
        // output n (mod 32) sits at fidx(n); add exp(-2 pi i n HP / 64) * Y[n]
        static_for<0, 32>([&](auto jc) {
            constexpr int jslot = decltype(jc)::value;   // n = slot_n(jslot)
            constexpr int k = fidx<32>((slot_n(jslot) + 32) % 32);
            float r = er[k], i = ei[k];
            if constexpr (HP == 1) twiddle_mul<64, (64 - ((slot_n(jslot) + 64) % 64)) % 64, +1>(r, i);
            sr[jslot] += r;
            si[jslot] += i;
        });
    });
    wave_lds_fence();
    // S1bar(a = lane, v) with v = slot_n(jslot) + 16, layout [a][v] stride kS1
#pragma unroll
    for (int jslot = 0; jslot < 32; jslot++) {
        const int v = slot_n(jslot) + 16;
        S.re[lane * kS1 + v] = sr[jslot];
        S.im[lane * kS1 + v] = si[jslot];
    }
    wave_lds_fence();
    // adjoint of stage 1: lane (v, h) sums over rows a = 2m + h
    const int v = lane & 31, hh = lane >> 5;
    float pr[32], pi[32];
#pragma unroll
    for (int m = 0; m < 32; m++) {
        pr[m] = S.re[(2 * m + hh) * kS1 + v];
        pi[m] = S.im[(2 * m + hh) * kS1 + v];
    }
    fft_dif<32, -1>(pr, pi);
    const float nv = (float)(v - 16);
    const float Xh = X + (float)hh * (float)(2.0 * kPi / 64.0);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    float gx = 0.f, gy = 0.f, gz = 0.f, go = 0.f;
#pragma unroll
    for (int jslot = 0; jslot < 32; jslot++) {
        const int nu = slot_n(jslot);
        const int k = fidx<32>((nu + 32) % 32);
        // exp(-2 pi i nu hh / 64): fold into the phase as in the forward pass
        const float tw = -(float)hh * (float)(2.0 * kPi / 64.0) * (float)nu;
        float st, ct;
        sincos_reduced(tw, st, ct);
        float r = pr[k] * ct - pi[k] * st;
        float i = pr[k] * st + pi[k] * ct;
        // sum of the two row halves (lane v and lane v + 32)
        r += __shfl_xor(r, 32, 64);
        i += __shfl_xor(i, 32, 64);
        const int wi = (nu + 16) * kW + v;
        const float q = T.phi[wi] + X * (float)nu + cY + T.kz[wi] * z - T.kzo[wi] * zoil;
        float s, c;
        sincos_reduced(q, s, c);
        const float amp = T.amp[wi];
        const float qb = amp * (-s * r + c * i);
        if (hh == 0) {
            gx += M.kappa * ((float)nu + 0.5f) * qb;
            gy += M.kappa * (nv + 0.5f) * qb;
            gz += T.kz[wi] * qb;
            go -= T.kzo[wi] * qb;
            if (A.qbar) A.qbar[(size_t)b * 1024 + wi] = qb;
        }
    }
    (void)Xh;
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    gz = wave_sum(gz);
    go = wave_sum(go);
    if (lane == 0) {
        A.dparams[b * 4 + 0] = gx;
        A.dparams[b * 4 + 1] = gy;
        A.dparams[b * 4 + 2] = gz;
        A.dparams[b * 4 + 3] = go;
    }
}

}  // namespace zola

// zola_adjoint64.hpp -- reverse-mode (adjoint) pass of the PSF model on the
// 64x64 pupil grid, one wave per PSF.
//
// Given w = dL/dpsf on the ROI, the wave recomputes the forward pass (the
// stage-1 data kept in LDS, E recomputed from them in the backward pass), then
// runs every stage backwards:
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
    const double* bg;      // B background per pixel (used when poly is null)
    double* nll;           // B  sum(mu - n ln mu), mu = amp psf + bg (clamped at 1e-8)
    double* damp;          // B  dNLL/damp = sum psf (1 - n/mu)
    double* dbg;           // B  dNLL/dbg  = sum (1 - n/mu), may be null
    // full model of ZernikePhaseRetrieval.Model (1859-1867): PSF b belongs to
    // bead b / S; its background is that bead's order-2 polynomial in the patch
    // coordinates, c0 + c1 i + c2 i^2 + c3 ii + c4 i ii + c5 ii^2, plus an
    // optional per-bead additive map (sCMOS var/gain^2)
    int S;                 // slices per bead (bead index b / S)
    const double* poly;    // nb x 6, or null
    const double* offset;  // nb x R x R, or null
    double* dpoly;         // B x 6: sum (1 - n/mu) x monomial, or null
    double* dsig;          // B: dNLL/dsigma (the SIG kernels only)
};

// SIG (calibration form only): also dNLL/dsigma.  The blur kernel is the
// separable h1 (x) h1 with h1 = IDFT(G1), G1 a Gaussian of width sigma px, so
// d(Bl)/dsigma = (hs (x) h1 + h1 (x) hs) I with hs = dh1/dsigma, and
//   dNLL/dsigma = sum_{a,b} I(a,b) sum_j [us(a,j) h1(j - b) + u(a,j) hs(j - b)]
// with u, us the column adjoints of dL/dBl through h1 and hs.
template <int R, int WPB, bool CALIB, bool SIG = false>
__global__ __launch_bounds__(64 * WPB) void k_psf_adjoint64(Model64 M, AdjArgs A) {
    static_assert(CALIB || !SIG, "the sigma derivative is formed in the calibration form");
    // S: stage-1 data, kept until the row adjoint (E is recomputed from it there
    // instead of being held in registers through the backward pass); Tt: the
    // column-blur transpose, then dL/dBl
    struct Lds {
        BlockTables64 T;
        float hs[64];
        WaveScratch64 S[WPB];
        float Tt[WPB][R * kTS];
    };
    __shared__ Lds lds;
    BlockTables64& T = lds.T;
    load_tables64(T, M);
    if constexpr (SIG) {
        for (int t = threadIdx.x; t < 64; t += blockDim.x) lds.hs[t] = M.hs[t];
    }
    __syncthreads();
    constexpr int G = RoiLayout<R>::G, RPL = RoiLayout<R>::RPL;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * WPB + w;
    if (b >= A.B) return;
    WaveScratch64& S = lds.S[w];
    float* Tt = lds.Tt[w];
    const float x = A.params[b * 4 + 0], y = A.params[b * 4 + 1];
    const float z = A.params[b * 4 + 2], zoil = A.params[b * 4 + 3];
    const float X = M.kappa * x, Y = M.kappa * y;
    const float h1reg = T.h3[lane];

    // ---- forward (the stage-1 data stay in S) ----
    stage1_64(T, S, lane, 0, X, Y, z, zoil, M.kappa);
    wave_lds_fence();
    float bl[RPL];
    {
        float acc[R];
        stage2_E64<R>(h1reg, S, lane, acc);
        store_rows64<R>(Tt, lane, acc);
        wave_lds_fence();
        column_pass64<R>(T.h3, Tt, lane, bl);
        wave_lds_fence();  // Tt is reused for dL/dBl
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
        // ZernikePhaseRetrieval.Model.computeLikelihood (1859-1867): mu = psf A +
        // background (+ sCMOS map), replaced by 1e-8 where it is <= 0 (there the
        // likelihood term is constant, so its gradient is 0); w = dNLL/dpsf = A (1 - n/mu)
        const double av = A.amp[b];
        const int bead = (A.poly || A.offset) ? b / A.S : 0;
        double pc[6] = {A.poly ? 0.0 : A.bg[b], 0.0, 0.0, 0.0, 0.0, 0.0};
        if (A.poly) {
#pragma unroll
            for (int c = 0; c < 6; c++) pc[c] = A.poly[bead * 6 + c];
        }
        const double jd = (double)j;
        double nl = 0.0, da = 0.0, dm[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < RPL; t++) {
            const bool ok = g < G && g * RPL + t < R;
            wv[t] = 0.f;
            if (ok) {
                const int i = g * RPL + t;
                const double id = (double)i;
                const double ps = (double)(fabsf(bl[t]) / ssum);
                const double n = A.pixels[(size_t)b * R * R + i * R + j];
                // monomial order of PolynomialBackground (base 3, exponents of (i, ii) with sum < 3)
                const double mono[6] = {1.0, id, id * id, jd, id * jd, jd * jd};
                double bgp = 0.0;
#pragma unroll
                for (int c = 0; c < 6; c++) bgp += pc[c] * mono[c];
                if (A.offset) bgp += A.offset[((size_t)bead * R + i) * R + j];
                double mu = ps * av + bgp;
                const bool clamped = mu <= 0;
                if (clamped) mu = 1e-8;
                nl += mu - n * log(mu);
                const double r = clamped ? 0.0 : 1.0 - n / mu;
                da += ps * r;
#pragma unroll
                for (int c = 0; c < 6; c++) dm[c] += r * mono[c];
                wv[t] = (float)(av * r);
            }
        }
        nl = wave_sum(nl);
        da = wave_sum(da);
#pragma unroll
        for (int c = 0; c < 6; c++) dm[c] = wave_sum(dm[c]);
        if (lane == 0) {
            A.nll[b] = nl;
            A.damp[b] = da;
            if (A.dbg) A.dbg[b] = dm[0];
            if (A.dpoly) {
#pragma unroll
                for (int c = 0; c < 6; c++) A.dpoly[(size_t)b * 6 + c] = dm[c];
            }
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
    float* Bt = Tt;               // R*R floats
#pragma unroll
    for (int t = 0; t < RPL; t++) {
        const bool ok = g < G && g * RPL + t < R;
        const float sg = bl[t] < 0.f ? -1.f : 1.f;
        if (ok) Bt[(g * RPL + t) * R + j] = sg * (wv[t] * inv - c2);
    }
    wave_lds_fence();
    // column adjoint for this lane's grid row a: u(j) = sum_i h1[(i - R/2 - a) & 63] Bt[i][j]
    float u[R], us[SIG ? R : 1];
    {
        float c[R], cs[SIG ? R : 1];
#pragma unroll
        for (int i = 0; i < R; i++) {
            c[i] = T.h3[64 + ((i - R / 2 - lane) & 63)];
            if constexpr (SIG) cs[i] = lds.hs[(i - R / 2 - lane) & 63];
        }
#pragma unroll
        for (int jj = 0; jj < R; jj++) {
            float acc = 0.f, accs = 0.f;
#pragma unroll
            for (int i = 0; i < R; i++) {
                const float bt = Bt[i * R + jj];
                acc += c[i] * bt;
                if constexpr (SIG) accs += cs[i] * bt;
            }
            u[jj] = acc;
            if constexpr (SIG) us[jj] = accs;
        }
    }
    // Pin u and us here: they are read only in the row adjoint, and without this
    // the compiler sinks their sums there, keeping all R x R loads of Bt live.
#pragma unroll
    for (int jj = 0; jj < R; jj++) {
        __asm__ __volatile__("" : "+v"(u[jj]));
        if constexpr (SIG) __asm__ __volatile__("" : "+v"(us[jj]));
    }
    const float hsreg = SIG ? lds.hs[lane] : 0.f;
    float gsig = 0.f;
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
        // E at the columns 2m + HP of row lane, recomputed from the stage-1 data:
        // column 2m + HP at index fidx(m)
        float Eh_r[32], Eh_i[32];
        stage2_half64<HP>(S, lane, Eh_r, Eh_i);
        float er[32], ei[32];
        // dL/dI(a, b) at the columns b = 2m + HP: ib[m] = sum_jj h1((jj - R/2 - b) & 63) u[jj],
        // grouped by tap index c = (jj - R/2 - b) & 63 (one broadcast per tap and
        // group, as the forward row blur does; jj = (c + R/2 + b) & 63 is a
        // constant), MG columns at a time: the sigma form keeps two accumulators
        // per column
        constexpr int MG = SIG ? 8 : 16;
        static_for<0, 32 / MG>([&](auto gc) {
            constexpr int m0 = decltype(gc)::value * MG;
            float ib[MG], ibs[SIG ? MG : 1];
#pragma unroll
            for (int t = 0; t < MG; t++) {
                ib[t] = 0.f;
                if constexpr (SIG) ibs[t] = 0.f;
            }
            static_for<0, 64>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                const float hc = bcast(h1reg, c);
                const float hsc = SIG ? bcast(hsreg, c) : 0.f;
                static_for<0, MG>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int jj = (c + R / 2 + 2 * (m0 + t) + HP) & 63;
                    if constexpr (jj < R) {
                        ib[t] += hc * u[jj];
                        if constexpr (SIG) ibs[t] += hsc * u[jj] + hc * us[jj];
                    }
                });
                __builtin_amdgcn_sched_barrier(0);
            });
            static_for<0, MG>([&](auto tc) {
                constexpr int t = decltype(tc)::value, m = m0 + t;
                const float Erm = Eh_r[fidx<32>(m)], Eim = Eh_i[fidx<32>(m)];
                if constexpr (SIG) gsig += ibs[t] * (Erm * Erm + Eim * Eim);
                er[m] = 2.f * ib[t] * Erm;
                ei[m] = 2.f * ib[t] * Eim;
            });
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
    if constexpr (SIG) {
        gsig = wave_sum(gsig);
        if (lane == 0) A.dsig[b] = (double)gsig;
    }
}

}  // namespace zola

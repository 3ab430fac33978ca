// zola_adjoint64.hpp -- reverse-mode (adjoint) pass of the PSF model on the
// 64x64 pupil grid, four waves (one block) per PSF.
//
// Given w = dL/dpsf on the ROI (or, in the calibration form, the likelihood of
// ZernikePhaseRetrieval.Model.computeLikelihood, 1859-1867, formed from the
// data), the block runs the forward pass, then every stage backwards:
//   psf = m/S, m = |Bl|          -> dL/dBl = sign(Bl) (w/S - sum(w m)/S^2)
//   Bl = h (x) I (separable)     -> dL/dI  = h^T (x) dL/dBl  (same kernel, transposed sums)
//   I = |E|^2                    -> dL/dE  = 2 dL/dI E
//   E = IDFT64(P) (32x32 support)-> dL/dP  = DFT64(dL/dE) cropped to the support
//   P = A exp(i q)               -> dL/dq  = A (-sin q dL/dPr + cos q dL/dPi)
// and reduces dL/dq against dq/dx, dq/dy, dq/dz, dq/dzoil.  dL/dq itself is
// written per PSF so that dL/da_j = sum_p Z_j(p) dL/dq(p) can be formed for
// the whole batch by the projection kernels (zola_project.hpp, FP64).
//
// A calibration evaluation has a few hundred PSFs, so one wave per PSF would
// leave most of the chip idle and take one wave's latency; the work of a PSF
// is split over four waves:
//  * forward: stage 1 by output-row parity (rows 4m + wr); stage 2, wave wr the output columns
//    b = 4k + wr of every grid row (lane), kept in registers for the backward
//    pass; row blur, wave wr a quarter of the crop columns; column blur and
//    likelihood, wave wr a quarter of the ROI rows (lane = ROI column);
//  * column adjoint, wave wr a quarter of the crop columns for every grid row;
//    row adjoint, wave wr the columns 4k + wr again, so dL/dE = 2 dL/dI E uses
//    the E held since the forward pass;
//  * adjoint of stage 2: each wave's 16 columns give a 16-point DFT and a
//    twiddle per window frequency; the four partial sums are added in LDS in
//    wave order;
//  * adjoint of stage 1: every wave runs the 32-point transforms of the window
//    columns, then handles 8 of the 32 window frequencies (the sincos and
//    gradient work); the per-wave gradient sums are added in wave order.
#pragma once
#include "zola_fit64.hpp"
#include "zola_psf64q.hpp"

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

constexpr int kPq = 33;   // row stride of the stage-2 adjoint partials

// Work split of a PSF over NW waves (4 or 8): stage-2 columns b = NW k + w
// (CPW = 64 / NW of them), crop columns [J0, J0 + NJ) and ROI rows
// [w RQ, w RQ + RQ) of wave w.
template <int R, int NW>
struct AdjSplit {
    static constexpr int CPW = 64 / NW;
    static constexpr int RQ = (R + NW - 1) / NW;
    template <int W>
    static constexpr int J0 = W * RQ < R ? W * RQ : R;
    template <int W>
    static constexpr int NJ = (W * RQ + RQ <= R ? RQ : R - J0<W>);
};

// Block sums of K doubles over the NW waves, added in wave order.
template <int NW, int K>
__device__ __forceinline__ void block_sum_w(double (&v)[K], double (*red)[16], int lane, int wr) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) red[wr][k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double t = red[0][k];
#pragma unroll
        for (int q = 1; q < NW; q++) t += red[q][k];
        v[k] = t;
    }
    __syncthreads();
}

// Stage 2, output columns b = NW k + Q of row p: E(p, NW k + Q) at
// fidx<CPW>(k).  The 32 signed frequencies n fold to n mod CPW after the
// twiddle exp(2 pi i n Q / 64), then a CPW-point IDFT.
template <int NW, int Q>
__device__ __forceinline__ void stage2_split64(const WaveScratch64& S, int p, float (&re)[64 / NW],
                                               float (&im)[64 / NW]) {
    constexpr int CPW = 64 / NW;
    __asm__ __volatile__("" ::: "memory");
    static_for<0, CPW>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        float ar = 0.f, ai = 0.f;
        static_for<0, 32 / CPW>([&](auto qc) {
            constexpr int v = s + decltype(qc)::value * CPW;   // window column, n = v - 16
            constexpr int n = v - 16;
            float r = S.re[p * kS1 + v], i = S.im[p * kS1 + v];
            twiddle_mul<64, ((n * Q) % 64 + 64) % 64, +1>(r, i);
            ar += r;
            ai += i;
        });
        re[s] = ar;
        im[s] = ai;
    });
    fft_dif<CPW, +1>(re, im);
}

// Row adjoint of the columns b = NW k + Q of grid row a (lane): ib[k] =
// sum_jj h1((jj - R/2 - b) & 63) u[jj], grouped by tap index c (one broadcast
// per tap); SIG: ibs[k] = sum_jj hs((jj - R/2 - b) & 63) u[jj] + h1(..) us[jj].
template <int R, int NW, int Q, bool SIG>
__device__ __forceinline__ void row_adjoint_split(float h1reg, float hsreg, const float (&u)[R],
                                                  const float (&us)[SIG ? R : 1], float (&ib)[64 / NW],
                                                  float (&ibs)[SIG ? 64 / NW : 1]) {
    constexpr int CPW = 64 / NW;
#pragma unroll
    for (int k = 0; k < CPW; k++) {
        ib[k] = 0.f;
        if constexpr (SIG) ibs[k] = 0.f;
    }
    static_for<0, 64>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        const float hc = bcast(h1reg, c);
        const float hsc = SIG ? bcast(hsreg, c) : 0.f;
        static_for<0, CPW>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int jj = (c + R / 2 + NW * k + Q) & 63;
            if constexpr (jj < R) {
                ib[k] += hc * u[jj];
                if constexpr (SIG) ibs[k] += hsc * u[jj] + hc * us[jj];
            }
        });
        __builtin_amdgcn_sched_barrier(0);
    });
}

// Adjoint of stage 2 for the columns NW k + Q of row a: part[jslot] =
// sum_k dE(a, NW k + Q) exp(-2 pi i n (NW k + Q) / 64), n = slot_n(jslot).
// ADD: added to the partials already there (a second wave's share).
template <int NW, int Q, bool ADD>
__device__ __forceinline__ void stage2_adj_split(float (&dr)[64 / NW], float (&di)[64 / NW], float* Pr, float* Pi,
                                                 int a) {
    constexpr int CPW = 64 / NW;
    fft_dif<CPW, -1>(dr, di);
    static_for<0, 32>([&](auto jc) {
        constexpr int jslot = decltype(jc)::value;
        constexpr int n = slot_n(jslot);
        constexpr int k = fidx<CPW>(((n % CPW) + CPW) % CPW);
        float r = dr[k], i = di[k];
        twiddle_mul<64, ((-n * Q) % 64 + 64) % 64, +1>(r, i);
        if constexpr (ADD) {
            Pr[a * kPq + jslot] += r;
            Pi[a * kPq + jslot] += i;
        } else {
            Pr[a * kPq + jslot] = r;
            Pi[a * kPq + jslot] = i;
        }
    });
}

// Stage 1 of E for the grid rows a = 4m + H (H = 0..3, wave-uniform):
//   S1(a, v) = sum_nu P(nu, v) exp(2 pi i nu a / 64), nu in [-16, 16).
// With a = 4m + H the sum is a 16-point IDFT over m of the input twiddled by
// exp(2 pi i nu H / 64) (folded into the phase) and folded mod 16.  Lane
// (v, s) evaluates the 16 frequencies of half s (s = 0: 0..15, s = 1:
// -16..-1); the two lanes of a column add each other's half (permlane32_swap),
// so both hold the folded input, and each stores half of the 16 outputs.
__device__ __forceinline__ void stage1_quarter64(const BlockTables64& T, WaveScratch64& S, int lane, int H, float X,
                                                 float Y, float z, float zoil) {
    const int v = lane & 31, s = lane >> 5;
    const bool up = s != 0;
    const float nv = (float)(v - 16);
    const float Xh = X + (float)H * (float)(2.0 * kPi / 64.0);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    float re[16], im[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int nu = k - 16 * s;
        const int wi = (nu + 16) * kW + v;
        const float q = T.phi[wi] + Xh * (float)nu + cY + T.kz[wi] * z - T.kzo[wi] * zoil;
        float sn, cs;
        sincos_reduced(q, sn, cs);
        const float a = T.amp[wi];
        const float pr = a * cs, pi = a * sn;
        re[k] = pr + other_half(pr, up);
        im[k] = pi + other_half(pi, up);
    }
    fft_dif<16, +1>(re, im);
#pragma unroll
    for (int m = 0; m < 16; m++) {
        if ((m & 1) == s) {
            const int a = 4 * m + H;
            S.re[a * kS1 + v] = re[fidx<16>(m)];
            S.im[a * kS1 + v] = im[fidx<16>(m)];
        }
    }
}

// Calls f(integral_constant<int, W>) for the wave index wr (wave-uniform).
template <int NW, class F>
__device__ __forceinline__ void per_wave(int wr, F&& f) {
    static_for<0, NW>([&](auto wc) {
        if (wr == decltype(wc)::value) f(wc);
    });
}

// SIG (calibration form only): also dNLL/dsigma.  The blur kernel is the
// separable h1 (x) h1 with h1 = IDFT(G1), G1 a Gaussian of width sigma px, so
// d(Bl)/dsigma = (hs (x) h1 + h1 (x) hs) I with hs = dh1/dsigma, and
//   dNLL/dsigma = sum_{a,b} I(a,b) sum_j [us(a,j) h1(j - b) + u(a,j) hs(j - b)]
// with u, us the column adjoints of dL/dBl through h1 and hs.
template <int R, int NW, bool CALIB, bool SIG = false>
__global__ __launch_bounds__(64 * NW) void k_psf_adjoint64(Model64 M, AdjArgs A) {
    static_assert(CALIB || !SIG, "the sigma derivative is formed in the calibration form");
    static_assert(NW == 4 || NW == 8, "four or eight waves per PSF");
    using SP = AdjSplit<R, NW>;
    constexpr int CPW = SP::CPW, RQ = SP::RQ;
    constexpr int kSlots = 4;   // stage-2 partial slots (eight waves: two per slot)
    struct Lds {
        BlockTables64 T;
        float hs[64];
        WaveScratch64 S;          // stage 1; then I; then the row-blurred rows; then dL/d(stage-1 output)
        float Bt[R * R];          // dL/dBl on the ROI
        float U[64 * R];          // column adjoint u(a, jj)
        float Us[SIG ? 64 * R : 1];
        float Pr[kSlots][64 * kPq];   // stage-2 adjoint partials
        float Pi[kSlots][64 * kPq];
        double red[NW][16];
    };
    __shared__ Lds lds;
    BlockTables64& T = lds.T;
    load_tables64(T, M);
    if constexpr (SIG) {
        for (int t = threadIdx.x; t < 64; t += blockDim.x) lds.hs[t] = M.hs[t];
    }
    const int b = blockIdx.x;
    if (b >= A.B) return;
    const int wr = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    __syncthreads();
    const float x = A.params[b * 4 + 0], y = A.params[b * 4 + 1];
    const float z = A.params[b * 4 + 2], zoil = A.params[b * 4 + 3];
    const float X = M.kappa * x, Y = M.kappa * y;
    const float h1reg = T.h3[lane];

    // ---- forward: stage 1, waves 0..3 the grid rows of parity wr mod 4
    if (wr < 4) stage1_quarter64(T, lds.S, lane, wr, X, Y, z, zoil);
    __syncthreads();
    // ---- stage 2: E(a, NW k + wr) of row a = lane, kept for the backward pass
    float Er[CPW], Ei[CPW];
    per_wave<NW>(wr, [&](auto wc) { stage2_split64<NW, decltype(wc)::value>(lds.S, lane, Er, Ei); });
    __syncthreads();  // all stage-1 reads are done
    float* const Sf = reinterpret_cast<float*>(&lds.S);   // the stage-1 buffer as one float array
    float* Ib = Sf;  // real [64][kIq]
#pragma unroll
    for (int k = 0; k < CPW; k++) {
        const int f = fidx<CPW>(k);
        Ib[lane * kIq + NW * k + wr] = Er[f] * Er[f] + Ei[f] * Ei[f];
    }
    __syncthreads();
    // ---- row blur at the crop columns [J0, J0 + NJ) of row a = lane
    float ra[1][RQ];
    per_wave<NW>(wr, [&](auto wc) {
        constexpr int W = decltype(wc)::value;
        constexpr int NJ = SP::template NJ<W>;
        if constexpr (NJ > 0) {
            float rw[1][NJ];
            row_blur64q<R, SP::template J0<W>, NJ, 1>(Ib, Ib, lane, h1reg, 0.f, rw);
#pragma unroll
            for (int jj = 0; jj < NJ; jj++) ra[0][jj] = rw[0][jj];
        }
    });
    __syncthreads();  // all intensity reads are done
    float* Rb = Sf;  // [column][row]
    per_wave<NW>(wr, [&](auto wc) {
        constexpr int W = decltype(wc)::value;
#pragma unroll
        for (int jj = 0; jj < SP::template NJ<W>; jj++) Rb[(SP::template J0<W> + jj) * kIq + lane] = ra[0][jj];
    });
    __syncthreads();
    // ---- column blur: lane j, ROI rows [wr RQ, wr RQ + RQ)
    const int j = lane < R ? lane : 0;
    float bl[RQ], d0[RQ], d1[RQ], d2[RQ];
#pragma unroll
    for (int t = 0; t < RQ; t++) bl[t] = 0.f;
    per_wave<NW>(wr, [&](auto wc) {
        constexpr int W = decltype(wc)::value;
        if constexpr (W * RQ < R) column_blur64q<R, W * RQ, RQ, 1>(Rb, j, h1reg, 0.f, bl, d0, d1, d2);
    });
    auto ok = [&](int t) { return lane < R && wr * RQ + t < R; };
    float wv[RQ];
    double sacc[1] = {0.0};
#pragma unroll
    for (int t = 0; t < RQ; t++)
        if (ok(t)) sacc[0] += (double)fabsf(bl[t]);
    // the PSF sum added per wave and over the waves in FP64, then rounded
    block_sum_w<NW, 1>(sacc, lds.red, lane, wr);
    const float ssum = (float)sacc[0];
    if constexpr (CALIB) {
        const double av = A.amp[b];
        const int bead = (A.poly || A.offset) ? b / A.S : 0;
        double pc[6] = {A.poly ? 0.0 : A.bg[b], 0.0, 0.0, 0.0, 0.0, 0.0};
        if (A.poly) {
#pragma unroll
            for (int c = 0; c < 6; c++) pc[c] = A.poly[bead * 6 + c];
        }
        const double jd = (double)j;
        // [0] nll, [1] dNLL/damp, [2..7] sum r x monomial
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < RQ; t++) {
            wv[t] = 0.f;
            if (ok(t)) {
                const int i = wr * RQ + t;
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
                acc[0] += mu - n * log(mu);
                const double r = clamped ? 0.0 : 1.0 - n / mu;
                acc[1] += ps * r;
#pragma unroll
                for (int c = 0; c < 6; c++) acc[2 + c] += r * mono[c];
                wv[t] = (float)(av * r);
            }
        }
        block_sum_w<NW, 8>(acc, lds.red, lane, wr);
        if (threadIdx.x == 0) {
            A.nll[b] = acc[0];
            A.damp[b] = acc[1];
            if (A.dbg) A.dbg[b] = acc[2];
            if (A.dpoly) {
#pragma unroll
                for (int c = 0; c < 6; c++) A.dpoly[(size_t)b * 6 + c] = acc[2 + c];
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < RQ; t++) wv[t] = ok(t) ? A.w[(size_t)b * R * R + (wr * RQ + t) * R + j] : 0.f;
    }
    double wacc[1] = {0.0};
#pragma unroll
    for (int t = 0; t < RQ; t++)
        if (ok(t)) wacc[0] += (double)(wv[t] * fabsf(bl[t]));
    block_sum_w<NW, 1>(wacc, lds.red, lane, wr);
    const float wm = (float)wacc[0];
    const float inv = 1.f / ssum, c2 = wm * inv * inv;

    // ---- dL/dBl at the ROI pixels, Bt[i][j]
#pragma unroll
    for (int t = 0; t < RQ; t++) {
        const float sg = bl[t] < 0.f ? -1.f : 1.f;
        if (ok(t)) lds.Bt[(wr * RQ + t) * R + j] = sg * (wv[t] * inv - c2);
    }
    __syncthreads();
    // ---- column adjoint for grid row a = lane, this wave's crop columns:
    // u(a, jj) = sum_i h1((i - R/2 - a) & 63) Bt[i][jj]
    {
        float c[R], cs[SIG ? R : 1];
#pragma unroll
        for (int i = 0; i < R; i++) {
            c[i] = T.h3[64 + ((i - R / 2 - lane) & 63)];
            if constexpr (SIG) cs[i] = lds.hs[(i - R / 2 - lane) & 63];
        }
        const int j0 = wr * RQ;
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            const int jj = j0 + q;
            if (jj < R) {
                float acc = 0.f, accs = 0.f;
#pragma unroll
                for (int i = 0; i < R; i++) {
                    const float bt = lds.Bt[i * R + jj];
                    acc += c[i] * bt;
                    if constexpr (SIG) accs += cs[i] * bt;
                }
                lds.U[lane * R + jj] = acc;
                if constexpr (SIG) lds.Us[lane * R + jj] = accs;
            }
        }
    }
    __syncthreads();
    // ---- row adjoint and dL/dE = 2 dL/dI E at the columns NW k + wr of row a = lane
    float gsig = 0.f;
    {
        float u[R], us[SIG ? R : 1];
#pragma unroll
        for (int jj = 0; jj < R; jj++) {
            u[jj] = lds.U[lane * R + jj];
            if constexpr (SIG) us[jj] = lds.Us[lane * R + jj];
        }
        const float hsreg = SIG ? lds.hs[lane] : 0.f;
        float ib[CPW], ibs[SIG ? CPW : 1];
        per_wave<NW>(wr, [&](auto wc) {
            row_adjoint_split<R, NW, decltype(wc)::value, SIG>(h1reg, hsreg, u, us, ib, ibs);
        });
        float dr[CPW], di[CPW];
#pragma unroll
        for (int k = 0; k < CPW; k++) {
            const int f = fidx<CPW>(k);
            if constexpr (SIG) gsig += ibs[k] * (Er[f] * Er[f] + Ei[f] * Ei[f]);
            dr[k] = 2.f * ib[k] * Er[f];
            di[k] = 2.f * ib[k] * Ei[f];
        }
        // waves 0..3 write the slots, waves 4..7 then add theirs (slot wr - 4)
        if (wr < kSlots) {
            per_wave<NW>(wr, [&](auto wc) {
                constexpr int W = decltype(wc)::value;
                if constexpr (W < kSlots) stage2_adj_split<NW, W, false>(dr, di, lds.Pr[W], lds.Pi[W], lane);
            });
        }
        if constexpr (NW > kSlots) {
            __syncthreads();
            if (wr >= kSlots) {
                per_wave<NW>(wr, [&](auto wc) {
                    constexpr int W = decltype(wc)::value;
                    if constexpr (W >= kSlots)
                        stage2_adj_split<NW, W, true>(dr, di, lds.Pr[W - kSlots], lds.Pi[W - kSlots], lane);
                });
            }
        }
    }
    __syncthreads();
    // ---- dL/d(stage-1 output) S1bar(a, v) = sum of the slots, in slot order
    for (int e = threadIdx.x; e < 64 * 32; e += 64 * NW) {
        const int a = e >> 5, js = e & 31;
        const int v = slot_n(js) + 16;
        const int o = a * kPq + js;
        lds.S.re[a * kS1 + v] = ((lds.Pr[0][o] + lds.Pr[1][o]) + lds.Pr[2][o]) + lds.Pr[3][o];
        lds.S.im[a * kS1 + v] = ((lds.Pi[0][o] + lds.Pi[1][o]) + lds.Pi[2][o]) + lds.Pi[3][o];
    }
    __syncthreads();
    // ---- adjoint of stage 1: lane (v, hh) sums over rows a = 2m + hh; wave wr
    // then takes the window frequencies of slots [wr 32/NW, (wr + 1) 32/NW)
    const int v = lane & 31, hh = lane >> 5;
    float pr[32], pi[32];
#pragma unroll
    for (int m = 0; m < 32; m++) {
        pr[m] = lds.S.re[(2 * m + hh) * kS1 + v];
        pi[m] = lds.S.im[(2 * m + hh) * kS1 + v];
    }
    fft_dif<32, -1>(pr, pi);
    const float nv = (float)(v - 16);
    const float cY = Y * (nv + 0.5f) + 0.5f * X;
    double g4[4] = {0.0, 0.0, 0.0, 0.0};
    static_for<0, 32>([&](auto jc) {
        constexpr int jslot = decltype(jc)::value;
        if (jslot / (32 / NW) == wr) {
            constexpr int nu = slot_n(jslot);
            constexpr int k = fidx<32>((nu + 32) % 32);
            // exp(-2 pi i nu hh / 64)
            const float tw = -(float)hh * (float)(2.0 * kPi / 64.0) * (float)nu;
            float st, ct;
            sincos_reduced(tw, st, ct);
            float r = pr[k] * ct - pi[k] * st;
            float i = pr[k] * st + pi[k] * ct;
            // sum of the two row parities (lane v and lane v + 32)
            r += __shfl_xor(r, 32, 64);
            i += __shfl_xor(i, 32, 64);
            const int wi = (nu + 16) * kW + v;
            const float q = T.phi[wi] + X * (float)nu + cY + T.kz[wi] * z - T.kzo[wi] * zoil;
            float s, c;
            sincos_reduced(q, s, c);
            const float qb = T.amp[wi] * (-s * r + c * i);
            if (hh == 0) {
                g4[0] += (double)(M.kappa * ((float)nu + 0.5f) * qb);
                g4[1] += (double)(M.kappa * (nv + 0.5f) * qb);
                g4[2] += (double)(T.kz[wi] * qb);
                g4[3] -= (double)(T.kzo[wi] * qb);
                if (A.qbar) A.qbar[(size_t)b * 1024 + wi] = qb;
            }
        }
    });
    block_sum_w<NW, 4>(g4, lds.red, lane, wr);
    double gs[1] = {(double)gsig};
    if constexpr (SIG) block_sum_w<NW, 1>(gs, lds.red, lane, wr);
    if (threadIdx.x == 0) {
        A.dparams[b * 4 + 0] = (float)g4[0];
        A.dparams[b * 4 + 1] = (float)g4[1];
        A.dparams[b * 4 + 2] = (float)g4[2];
        A.dparams[b * 4 + 3] = (float)g4[3];
        if constexpr (SIG) A.dsig[b] = gs[0];
    }
}

}  // namespace zola

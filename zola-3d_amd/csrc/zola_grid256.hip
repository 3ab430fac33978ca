// zola_grid256.hip -- PSF synthesis, its adjoint and the calibration
// likelihood on the 256 x 256 pupil grid (gfx950), in FP64.
//
// The plugin's calibration dialog uses sizeFFT 140 (ZOLA.java:2652-2658), which
// the CPU path rounds up to N = 256 (PhaseParameters.java:278-296), with 64 px
// bead patches: blur grid s2 = pow2(min(2R, N)) = 128 (PSFPhaseMany.java:102-108).
// The pupil disk (radius 60 px for NA 1.45, 680 nm, 110 nm pixels) lies in a
// 128 x 128 window of signed frequencies [-64, 64).
//
// One evaluation is a chain of batched FP64 line transforms (k_fft_lines,
// zola_fft_f64.hpp, unitary like utils/FastFourierTransform.java) whose IO
// functors place, crop and combine the data, plus small per-PSF kernels:
//   1  columns of the window:  E1(a, v) = IDFT_256 over n_u of P(n_u, v), a in
//      the s2 crop rows;  P = A exp(i q), q = phi + kx x + ky y + kz z - kzo zoil
//      formed in the load functor (PSFPhaseMany.java:553-571)
//   2  rows:                   E(a, c) = IDFT_256 over n_v of E1(a, v), c in the crop
//   3  blur:                   FI = DFT_s2(|E|^2);  B = IDFT_s2(G FI) at the R x R
//      crop (G the float-rounded Gaussian of PSFPhaseMany.java:387-404)
//   4  per PSF:                psf = |B| / sum |B|; calibration: mu, NLL and
//      dNLL/dpsf (ZernikePhaseRetrieval.java:1859-1867); dL/dB on the crop
//   5  adjoint of the blur:    dI = Re IDFT(G DFT(dB)); dNLL/dsigma =
//      Re sum conj(DFT(dB)) dG/dsigma FI
//   6  adjoint of 2 and 1:     dP = DFT_256 of 2 dI E on the crop, at the window
//   7  per PSF and window point: dL/dq = A (cos q dPi - sin q dPr), reduced
//      against kx, ky, kz, -kzo; projected on the Zernike basis for dL/da_j.
// Each pass reads and writes whole lines in HBM: the work is a few hundred MB
// of traffic per evaluation of a 205-slice stack, so it is memory-bound work
// that FP64 costs in bytes, not in issue rate (FP64 and non-packed FP32 FMA
// rates are equal on this chip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "zola_ctx.hpp"
#include "zola_fft_f64.hpp"
#include "zola_grid256.h"
#include "zola_kernels.h"
#include "zola_project.hpp"

namespace zola {

namespace {

constexpr int kN4 = 256;    // pupil grid
constexpr int kW4 = 128;    // window side (signed frequencies -64..63)
constexpr int kWin = kW4 * kW4;

__device__ __forceinline__ int signed_n(int k) { return k < kN4 / 2 ? k : k - kN4; }

struct Geo {
    const double4* tab;    // kWin {phase, amp, kz, kzo}, amp = 0 outside the disk
    double kappa;          // 2 pi / (dx N)
    int s2, R;
};

// Pupil phase of window point (nu, nv) for PSF b.
__device__ __forceinline__ double pupil_q(const Geo& g, const double4& t, int nu, int nv, const float* params, int b) {
    const double x = params[b * 4 + 0], y = params[b * 4 + 1];
    const double z = params[b * 4 + 2], zoil = params[b * 4 + 3];
    const double kx = ((double)nu + 0.5) * g.kappa, ky = ((double)nv + 0.5) * g.kappa;
    return t.x + kx * x + ky * y + t.z * z - t.w * zoil;
}

// 1: columns v of the window, lines b * 128 + v, length 256 (inverse).
struct IoStage1 {
    Geo g;
    const float* params;
    double2* S1;  // B x s2 x 128
    __device__ double2 load(int line, int k) const {
        const int b = line >> 7, v = line & (kW4 - 1);
        const int nu = signed_n(k);
        if (nu < -kW4 / 2 || nu >= kW4 / 2) return make_double2(0.0, 0.0);
        const double4 t = g.tab[(nu + kW4 / 2) * kW4 + v];
        if (t.y == 0.0) return make_double2(0.0, 0.0);
        double s, c;
        sincos(pupil_q(g, t, nu, v - kW4 / 2, params, b), &s, &c);
        return make_double2(t.y * c, t.y * s);
    }
    __device__ void store(int line, int k, double2 val) const {
        const int b = line >> 7, v = line & (kW4 - 1);
        const int a = signed_n(k);
        if (a < -g.s2 / 2 || a >= g.s2 / 2) return;
        S1[((size_t)b * g.s2 + ((a + g.s2) & (g.s2 - 1))) * kW4 + v] = val;
    }
};

// 2: rows of the crop, lines b * s2 + ai, length 256 (inverse).
struct IoStage2 {
    int s2;
    const double2* S1;
    double2* E;  // B x s2 x s2
    __device__ double2 load(int line, int k) const {
        const int nv = signed_n(k);
        if (nv < -kW4 / 2 || nv >= kW4 / 2) return make_double2(0.0, 0.0);
        return S1[(size_t)line * kW4 + nv + kW4 / 2];
    }
    __device__ void store(int line, int k, double2 val) const {
        const int c = signed_n(k);
        if (c < -s2 / 2 || c >= s2 / 2) return;
        E[(size_t)line * s2 + ((c + s2) & (s2 - 1))] = val;
    }
};

// 3a: rows of |E|^2 (forward), lines b * s2 + ai.
struct IoAbs2Rows {
    int s2;
    const double2* E;
    double2* T;
    __device__ double2 load(int line, int k) const {
        const double2 e = E[(size_t)line * s2 + k];
        return make_double2(e.x * e.x + e.y * e.y, 0.0);
    }
    __device__ void store(int line, int k, double2 val) const { T[(size_t)line * s2 + k] = val; }
};

// Column pass over an s2 x s2 array per PSF, lines b * s2 + column; the load
// multiplies by the Gaussian spectrum when gk is set (frequency (k, column)).
struct IoCols {
    int s2;
    const double2* src;
    double2* dst;
    const double* gk;  // s2 x s2 or null
    __device__ double2 load(int line, int k) const {
        const int b = line / s2, col = line - b * s2;
        double2 v = src[((size_t)b * s2 + k) * s2 + col];
        if (gk) {
            const double gv = gk[k * s2 + col];
            v.x *= gv;
            v.y *= gv;
        }
        return v;
    }
    __device__ void store(int line, int k, double2 val) const {
        const int b = line / s2, col = line - b * s2;
        dst[((size_t)b * s2 + k) * s2 + col] = val;
    }
};

// 3c: rows of the crop of the blurred image (inverse), lines b * R + r, grid
// row (r - R/2) mod s2; columns stored at (c + R/2) for signed c in [-R/2, R - R/2).
struct IoCropRows {
    int s2, R;
    const double2* T;
    double2* Bc;  // B x R x R
    __device__ double2 load(int line, int k) const {
        const int b = line / R, r = line - b * R;
        const int i = (r - R / 2 + s2) & (s2 - 1);
        return T[((size_t)b * s2 + i) * s2 + k];
    }
    __device__ void store(int line, int k, double2 val) const {
        const int cs = k < s2 / 2 ? k : k - s2;
        const int rr = cs + R / 2;
        if (rr < 0 || rr >= R) return;
        Bc[(size_t)line * R + rr] = val;
    }
};

// 5a: rows of dL/dB (zero outside the crop) (forward), lines b * R + r.
struct IoCropRowsIn {
    int s2, R;
    const double2* dBc;
    double2* T;
    __device__ double2 load(int line, int k) const {
        const int cs = k < s2 / 2 ? k : k - s2;
        const int rr = cs + R / 2;
        if (rr < 0 || rr >= R) return make_double2(0.0, 0.0);
        return dBc[(size_t)line * R + rr];
    }
    __device__ void store(int line, int k, double2 val) const {
        const int b = line / R, r = line - b * R;
        const int i = (r - R / 2 + s2) & (s2 - 1);
        T[((size_t)b * s2 + i) * s2 + k] = val;
    }
};

// 5b: column pass of 5a (forward): rows outside the crop are zero.
struct IoColsCropIn {
    int s2, R;
    double2* T;
    __device__ double2 load(int line, int k) const {
        const int b = line / s2, col = line - b * s2;
        const int cs = k < s2 / 2 ? k : k - s2;
        const int rr = cs + R / 2;
        if (rr < 0 || rr >= R) return make_double2(0.0, 0.0);
        return T[((size_t)b * s2 + k) * s2 + col];
    }
    __device__ void store(int line, int k, double2 val) const {
        const int b = line / s2, col = line - b * s2;
        T[((size_t)b * s2 + k) * s2 + col] = val;
    }
};

// 5d: rows (inverse), in place.
struct IoRows {
    int s2;
    double2* T;
    __device__ double2 load(int line, int k) const { return T[(size_t)line * s2 + k]; }
    __device__ void store(int line, int k, double2 val) const { T[(size_t)line * s2 + k] = val; }
};

// 6a: rows (forward, length 256): 2 Re(dI) E on the crop columns, window output.
struct IoAdjStage2 {
    int s2;
    const double2* dI;
    const double2* E;
    double2* S1;
    __device__ double2 load(int line, int k) const {
        const int c = signed_n(k);
        if (c < -s2 / 2 || c >= s2 / 2) return make_double2(0.0, 0.0);
        const size_t o = (size_t)line * s2 + ((c + s2) & (s2 - 1));
        const double gI = 2.0 * dI[o].x;
        const double2 e = E[o];
        return make_double2(gI * e.x, gI * e.y);
    }
    __device__ void store(int line, int k, double2 val) const {
        const int nv = signed_n(k);
        if (nv < -kW4 / 2 || nv >= kW4 / 2) return;
        S1[(size_t)line * kW4 + nv + kW4 / 2] = val;
    }
};

// 6b: columns of the window (forward, length 256), then dL/dq at the window.
struct IoAdjStage1 {
    Geo g;
    const float* params;
    const double2* S1;
    double* qbar;  // B x kWin
    __device__ double2 load(int line, int k) const {
        const int b = line >> 7, v = line & (kW4 - 1);
        const int a = signed_n(k);
        if (a < -g.s2 / 2 || a >= g.s2 / 2) return make_double2(0.0, 0.0);
        return S1[((size_t)b * g.s2 + ((a + g.s2) & (g.s2 - 1))) * kW4 + v];
    }
    __device__ void store(int line, int k, double2 val) const {
        const int b = line >> 7, v = line & (kW4 - 1);
        const int nu = signed_n(k);
        if (nu < -kW4 / 2 || nu >= kW4 / 2) return;
        const int w = (nu + kW4 / 2) * kW4 + v;
        const double4 t = g.tab[w];
        double gq = 0.0;
        if (t.y != 0.0) {
            double s, c;
            sincos(pupil_q(g, t, nu, v - kW4 / 2, params, b), &s, &c);
            gq = t.y * (c * val.y - s * val.x);
        }
        qbar[(size_t)b * kWin + w] = gq;
    }
};

// Deterministic block sum of K doubles over the NT / 64 waves of a block
// (butterfly within each wave, then the wave totals in wave order); every
// thread gets the sums.  red: K x (NT / 64) doubles.
template <int NT, int K>
__device__ __forceinline__ void block_sum_nt(double (&v)[K], double* red) {
    constexpr int NWV = NT / 64;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) red[k * NWV + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double t = red[k * NWV];
        for (int q = 1; q < NWV; q++) t += red[k * NWV + q];
        v[k] = t;
    }
    __syncthreads();
}

// Threads per block of the per-PSF kernels (one block per PSF: a few hundred
// blocks, so each needs many threads to keep its loads in flight); the
// likelihood and window-reduction kernels hold more per thread and take 512
// (256 registers).
constexpr int kPsfThreads = 1024, kCropThreads = 512;

struct CropArgs {
    int R, S;
    const double2* Bc;
    double2* dBc;
    float* out;            // forward: psf (B x R x R), or null
    const float* w;        // adjoint of a given dL/dpsf, or null
    const float* pixels;   // calibration: photons (B x R x R), or null
    const double* amp;     // B
    const double* bg;      // B (when poly is null)
    const double* poly;    // nb x 6 or null
    const double* offset;  // nb x R x R or null
    double *nll, *damp, *dbg, *dpoly;
};

// 4: one block per PSF.  Forward: psf = |B| / sum |B|.  Otherwise w = dL/dpsf
// (given, or dNLL/dpsf of the calibration model) and dL/dB = (w / S -
// sum(w |B|) / S^2) B / |B| on the crop.
__global__ __launch_bounds__(kCropThreads) void k_crop256(CropArgs A) {
    constexpr int NT = kCropThreads;
    __shared__ double red[9 * (NT / 64)];
    const int b = blockIdx.x, R = A.R, P = R * R;
    const double2* Bc = A.Bc + (size_t)b * P;
    double s1[1] = {0.0};
    for (int p = threadIdx.x; p < P; p += NT) s1[0] += sqrt(Bc[p].x * Bc[p].x + Bc[p].y * Bc[p].y);
    block_sum_nt<NT, 1>(s1, red);
    const double S = s1[0];
    if (A.out) {
        for (int p = threadIdx.x; p < P; p += NT)
            A.out[(size_t)b * P + p] = (float)(sqrt(Bc[p].x * Bc[p].x + Bc[p].y * Bc[p].y) / S);
        return;
    }
    // sums: [0] w |B|; calibration: [1] nll, [2] damp, [3..8] sum r x monomial (dbg = [3])
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int bead = (A.poly || A.offset) ? b / A.S : 0;
    double pc[6] = {0, 0, 0, 0, 0, 0};
    if (A.pixels) {
        if (A.poly) {
            for (int c = 0; c < 6; c++) pc[c] = A.poly[bead * 6 + c];
        } else {
            pc[0] = A.bg[b];
        }
    }
    const double av = A.pixels ? A.amp[b] : 0.0;
    for (int p = threadIdx.x; p < P; p += NT) {
        const double m = sqrt(Bc[p].x * Bc[p].x + Bc[p].y * Bc[p].y);
        double wv;
        if (A.pixels) {
            // ZernikePhaseRetrieval.Model.computeLikelihood (1859-1867): mu = psf A +
            // background (+ sCMOS map), 1e-8 where <= 0 (there the term is constant)
            const int i = p / R, j = p - i * R;
            const double id = (double)i, jd = (double)j;
            const double mono[6] = {1.0, id, id * id, jd, id * jd, jd * jd};
            double bgp = 0.0;
            for (int c = 0; c < 6; c++) bgp += pc[c] * mono[c];
            if (A.offset) bgp += A.offset[(size_t)bead * P + p];
            const double ps = m / S;
            double mu = ps * av + bgp;
            const bool clamped = mu <= 0;
            if (clamped) mu = 1e-8;
            const double n = A.pixels[(size_t)b * P + p];
            acc[1] += mu - n * log(mu);
            const double r = clamped ? 0.0 : 1.0 - n / mu;
            acc[2] += ps * r;
            for (int c = 0; c < 6; c++) acc[3 + c] += r * mono[c];
            wv = av * r;
        } else {
            wv = A.w[(size_t)b * P + p];
        }
        acc[0] += wv * m;
    }
    block_sum_nt<NT, 9>(acc, red);
    if (threadIdx.x == 0 && A.pixels) {
        A.nll[b] = acc[1];
        A.damp[b] = acc[2];
        if (A.dbg) A.dbg[b] = acc[3];
        if (A.dpoly)
            for (int c = 0; c < 6; c++) A.dpoly[(size_t)b * 6 + c] = acc[3 + c];
    }
    const double wm = acc[0], iS = 1.0 / S, c2 = wm * iS * iS;
    for (int p = threadIdx.x; p < P; p += NT) {
        const double2 bv = Bc[p];
        const double m = sqrt(bv.x * bv.x + bv.y * bv.y);
        double wv;
        if (A.pixels) {
            const int i = p / R, j = p - i * R;
            const double id = (double)i, jd = (double)j;
            const double mono[6] = {1.0, id, id * id, jd, id * jd, jd * jd};
            double bgp = 0.0;
            for (int c = 0; c < 6; c++) bgp += pc[c] * mono[c];
            if (A.offset) bgp += A.offset[(size_t)bead * P + p];
            double mu = m / S * av + bgp;
            const double n = A.pixels[(size_t)b * P + p];
            wv = mu <= 0 ? 0.0 : av * (1.0 - n / mu);
        } else {
            wv = A.w[(size_t)b * P + p];
        }
        const double gm = wv * iS - c2;
        A.dBc[(size_t)b * P + p] = m > 0 ? make_double2(gm * bv.x / m, gm * bv.y / m) : make_double2(0.0, 0.0);
    }
}

// 5c: dNLL/dsigma = Re sum_k conj(FdB(k)) dG/dsigma(k) FI(k), one block per PSF.
__global__ __launch_bounds__(kPsfThreads) void k_dsig256(int s2, const double2* FdB, const double2* FI,
                                                         const double* gks, double* dsig) {
    constexpr int NT = kPsfThreads;
    __shared__ double red[NT / 64];
    const int b = blockIdx.x, P = s2 * s2;
    double a[1] = {0.0};
    for (int p = threadIdx.x; p < P; p += NT) {
        const double2 u = FdB[(size_t)b * P + p], v = FI[(size_t)b * P + p];
        a[0] += gks[p] * (u.x * v.x + u.y * v.y);
    }
    block_sum_nt<NT, 1>(a, red);
    if (threadIdx.x == 0) dsig[b] = a[0];
}

// 7a: dL/dx, dL/dy, dL/dz, dL/dzoil of each PSF from dL/dq at the window.
__global__ __launch_bounds__(kCropThreads) void k_dparams256(Geo g, const double* qbar, float* dparams) {
    constexpr int NT = kCropThreads;
    __shared__ double red[4 * (NT / 64)];
    const int b = blockIdx.x;
    double a[4] = {0, 0, 0, 0};
    for (int w = threadIdx.x; w < kWin; w += NT) {
        const double gq = qbar[(size_t)b * kWin + w];
        if (gq == 0.0) continue;
        const int nu = w / kW4 - kW4 / 2, nv = w % kW4 - kW4 / 2;
        const double4 t = g.tab[w];
        a[0] += ((double)nu + 0.5) * g.kappa * gq;
        a[1] += ((double)nv + 0.5) * g.kappa * gq;
        a[2] += t.z * gq;
        a[3] -= t.w * gq;
    }
    block_sum_nt<NT, 4>(a, red);
    if (threadIdx.x < 4) dparams[(size_t)b * 4 + threadIdx.x] = (float)a[threadIdx.x];
}

template <int L, bool COLMAJOR, class IO>
hipError_t lines(const IO& io, int nlines, double sgn, const double2* tw, hipStream_t st) {
    hipLaunchKernelGGL((k_fft_lines<L, COLMAJOR, IO>), dim3(fft_lines_grid<L>(nlines)), dim3(256), 0, st, io, nlines,
                       sgn, tw);
    return hipGetLastError();
}

// s2-length line pass (s2 = 16 .. 256).
template <bool COLMAJOR, class IO>
hipError_t lines_s2(int s2, const IO& io, int nlines, double sgn, const double2* tw, hipStream_t st) {
    switch (s2) {
        case 16: return lines<16, COLMAJOR>(io, nlines, sgn, tw, st);
        case 32: return lines<32, COLMAJOR>(io, nlines, sgn, tw, st);
        case 64: return lines<64, COLMAJOR>(io, nlines, sgn, tw, st);
        case 128: return lines<128, COLMAJOR>(io, nlines, sgn, tw, st);
        case 256: return lines<256, COLMAJOR>(io, nlines, sgn, tw, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace
}  // namespace zola

using namespace zola;

namespace {
size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

int zola256_setup(zola_ctx* c) {
    HIP_TRY(c, hipMalloc((void**)&c->d_tab256, (size_t)kWin * sizeof(double4)));
    HIP_TRY(c, hipMalloc((void**)&c->d_gk256, 2 * (size_t)c->s2 * c->s2 * sizeof(double)));
    std::vector<double2> tw(kTwN);
    zola_twiddle_table(tw.data());
    HIP_TRY(c, hipMalloc((void**)&c->d_tw256, tw.size() * sizeof(double2)));
    HIP_TRY(c, hipMemcpy(c->d_tw256, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    if (c->J > 0) {
        std::vector<double> zw((size_t)c->J * kWin, 0.0);
        for (int jz = 0; jz < c->J; jz++)
            for (int p = 0; p < c->D; p++) zw[(size_t)jz * kWin + c->wu[p] * kW4 + c->wv[p]] = c->Z[(size_t)jz * c->D + p];
        HIP_TRY(c, hipMalloc((void**)&c->d_Zw, zw.size() * sizeof(double)));
        HIP_TRY(c, hipMemcpy(c->d_Zw, zw.data(), zw.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    int rc = zola256_upload_tables(c);
    if (rc) return rc;
    return zola256_upload_gauss(c);
}

int zola256_upload_tables(zola_ctx* c) {
    ensure_phase(c);
    std::vector<double4> tab(kWin, make_double4(0.0, 0.0, 0.0, 0.0));
    for (int p = 0; p < c->D; p++)
        tab[c->wu[p] * kW4 + c->wv[p]] = make_double4(c->phase[p], c->amp[p], c->kz[p], c->kzo[p]);
    HIP_TRY(c, hipMemcpyAsync(c->d_tab256, tab.data(), tab.size() * sizeof(double4), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

// G at frequency (k1, k2) of the s2 grid as PSFPhaseMany.java:387-404 builds it
// (a float-rounded 2-D Gaussian, indexed through the fftshift), and its
// derivative in sigma, -G r^2 / (sigF^2 sigma) with sigF = s2 / (2 pi sigma).
int zola256_upload_gauss(zola_ctx* c) {
    const int s2 = c->s2;
    const double sigF = (1.0 / c->o.sigma) * s2 / (3.14159265358979323846 * 2.0);
    const double sigpow = sigF * sigF;
    const double cc = (double)(s2 / 2);
    std::vector<double> g(2 * (size_t)s2 * s2);
    for (int k1 = 0; k1 < s2; k1++)
        for (int k2 = 0; k2 < s2; k2++) {
            const double j = (double)((k1 + s2 / 2) % s2), jj = (double)((k2 + s2 / 2) % s2);
            const double r2 = (j - cc) * (j - cc) + (jj - cc) * (jj - cc);
            g[(size_t)k1 * s2 + k2] = (double)(float)std::exp(-.5 * r2 / sigpow);
            g[(size_t)s2 * s2 + (size_t)k1 * s2 + k2] = -std::exp(-.5 * r2 / sigpow) * r2 / (sigpow * c->o.sigma);
        }
    HIP_TRY(c, hipMemcpyAsync(c->d_gk256, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ZOLA_OK;
}

namespace zola {
namespace {

// tab[w].x = sum_j Z_j(w) a_j in j order, products and sums not fused (the
// host's phase bit for bit; zero outside the disk).
__global__ __launch_bounds__(256) void k_phase256(int J, CoefVals cv, const double* __restrict__ Zw,
                                                  double4* __restrict__ tab) {
#pragma clang fp contract(off)
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= kWin) return;
    double ph = 0.0;
    for (int j0 = 0; j0 < J; j0 += 32) {
        double zv[32];
#pragma unroll
        for (int k = 0; k < 32; k++) zv[k] = j0 + k < J ? Zw[(size_t)(j0 + k) * kWin + w] : 0.0;
#pragma unroll
        for (int k = 0; k < 32; k++)
            if (j0 + k < J) ph = ph + zv[k] * cv.a[j0 + k];
    }
    tab[w].x = ph;
}

// zola256_upload_gauss's tables on the device: the same expressions, with the
// device exp (within an ulp of the host's; the spectrum is rounded to float as
// PSFPhaseMany.java:387-404 does).
__global__ __launch_bounds__(256) void k_gauss256(int s2, double sigma, double* __restrict__ g) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= s2 * s2) return;
    const int k1 = k / s2, k2 = k % s2;
    const double sigF = (1.0 / sigma) * s2 / (3.14159265358979323846 * 2.0);
    const double sigpow = sigF * sigF;
    const double cc = (double)(s2 / 2);
    const double j = (double)((k1 + s2 / 2) % s2), jj = (double)((k2 + s2 / 2) % s2);
    const double r2 = (j - cc) * (j - cc) + (jj - cc) * (jj - cc);   // integers: exact
    const double e = exp(-.5 * r2 / sigpow);
    g[k] = (double)(float)e;
    g[(size_t)s2 * s2 + k] = -e * r2 / (sigpow * sigma);
}

}  // namespace
}  // namespace zola

int zola256_set_phase(zola_ctx* c) {
    if (c->J <= 0 || c->J > kMaxCoefDev || !c->d_Zw) return zola256_upload_tables(c);
    CoefVals cv{};
    for (int j = 0; j < c->J; j++) cv.a[j] = c->coef[j];
    hipLaunchKernelGGL(k_phase256, dim3(kWin / 256), dim3(256), 0, c->stream, c->J, cv, (const double*)c->d_Zw,
                       c->d_tab256);
    HIP_TRY(c, hipGetLastError());
    return ZOLA_OK;
}

int zola256_set_gauss(zola_ctx* c) {
    const int s2 = c->s2;
    hipLaunchKernelGGL(k_gauss256, dim3((s2 * s2 + 255) / 256), dim3(256), 0, c->stream, s2, c->o.sigma,
                       c->d_gk256);
    HIP_TRY(c, hipGetLastError());
    return ZOLA_OK;
}

void zola256_free(zola_ctx* c) {
    if (c->d_tab256) (void)hipFree(c->d_tab256);
    if (c->d_gk256) (void)hipFree(c->d_gk256);
    if (c->d_tw256) (void)hipFree(c->d_tw256);
    if (c->d_work256) (void)hipFree(c->d_work256);
    c->d_tab256 = nullptr;
    c->d_gk256 = nullptr;
    c->d_tw256 = nullptr;
    c->d_work256 = nullptr;
    c->work256_bytes = 0;
}

int zola256_run(zola_ctx* c, const Eval256& e, hipStream_t st) {
    const int B = e.B, s2 = c->s2, R = c->R;
    if (B <= 0) return ZOLA_OK;
    const bool adjoint = e.w || e.pixels;
    // per-PSF work: S1 (s2 x 128), E, T, FI (s2 x s2 each), Bc, dBc (R x R), qbar (128 x 128 doubles)
    const size_t nS1 = al256((size_t)B * s2 * kW4 * sizeof(double2)), nE = al256((size_t)B * s2 * s2 * sizeof(double2));
    const size_t nB = al256((size_t)B * R * R * sizeof(double2)), nQ = al256((size_t)B * kWin * sizeof(double));
    const size_t nQs = al256((size_t)kWin * sizeof(double));
    const size_t need = nS1 + 3 * nE + 2 * nB + (adjoint ? nQ + nQs : 0);
    if (need > c->work256_bytes) {
        if (c->d_work256) HIP_TRY(c, hipFree(c->d_work256));
        c->d_work256 = nullptr;
        c->work256_bytes = 0;
        HIP_TRY(c, hipMalloc((void**)&c->d_work256, need));
        c->work256_bytes = need;
    }
    char* p = c->d_work256;
    double2* S1 = (double2*)p; p += nS1;
    double2* E = (double2*)p; p += nE;
    double2* T = (double2*)p; p += nE;
    double2* FI = (double2*)p; p += nE;
    double2* Bc = (double2*)p; p += nB;
    double2* dBc = (double2*)p; p += nB;
    double* qbar = (double*)p;
    double* Qs = (double*)(p + nQ);
    const Geo g{c->d_tab256, c->kappa, s2, R};
    const double2* tw = c->d_tw256;
    const double* gk = c->d_gk256;
    const double* gks = c->d_gk256 + (size_t)s2 * s2;

    // forward: pupil -> E -> FI -> B at the crop
    HIP_TRY(c, (lines<kN4, true>(IoStage1{g, e.params, S1}, B * kW4, +1.0, tw, st)));
    HIP_TRY(c, (lines<kN4, false>(IoStage2{s2, S1, E}, B * s2, +1.0, tw, st)));
    HIP_TRY(c, (lines_s2<false>(s2, IoAbs2Rows{s2, E, T}, B * s2, -1.0, tw, st)));
    HIP_TRY(c, (lines_s2<true>(s2, IoCols{s2, T, FI, nullptr}, B * s2, -1.0, tw, st)));
    HIP_TRY(c, (lines_s2<true>(s2, IoCols{s2, FI, T, gk}, B * s2, +1.0, tw, st)));
    HIP_TRY(c, (lines_s2<false>(s2, IoCropRows{s2, R, T, Bc}, B * R, +1.0, tw, st)));
    CropArgs A{};
    A.R = R;
    A.S = e.S > 0 ? e.S : 1;
    A.Bc = Bc;
    A.dBc = dBc;
    A.out = adjoint ? nullptr : e.out;
    A.w = e.w;
    A.pixels = e.pixels;
    A.amp = e.amp;
    A.bg = e.bg;
    A.poly = e.poly;
    A.offset = e.offset;
    A.nll = e.nll;
    A.damp = e.damp;
    A.dbg = e.dbg;
    A.dpoly = e.dpoly;
    hipLaunchKernelGGL(k_crop256, dim3(B), dim3(kCropThreads), 0, st, A);
    HIP_TRY(c, hipGetLastError());
    if (!adjoint) return ZOLA_OK;

    // adjoint of the blur: T <- DFT(dB), sigma term, T <- Re IDFT(G T)
    HIP_TRY(c, (lines_s2<false>(s2, IoCropRowsIn{s2, R, dBc, T}, B * R, -1.0, tw, st)));
    HIP_TRY(c, (lines_s2<true>(s2, IoColsCropIn{s2, R, T}, B * s2, -1.0, tw, st)));
    if (e.dsig) {
        hipLaunchKernelGGL(k_dsig256, dim3(B), dim3(kPsfThreads), 0, st, s2, T, FI, gks, e.dsig);
        HIP_TRY(c, hipGetLastError());
    }
    HIP_TRY(c, (lines_s2<true>(s2, IoCols{s2, T, T, gk}, B * s2, +1.0, tw, st)));
    HIP_TRY(c, (lines_s2<false>(s2, IoRows{s2, T}, B * s2, +1.0, tw, st)));
    // adjoint of the two pupil transforms, dL/dq at the window
    HIP_TRY(c, (lines<kN4, false>(IoAdjStage2{s2, T, E, S1}, B * s2, -1.0, tw, st)));
    HIP_TRY(c, (lines<kN4, true>(IoAdjStage1{g, e.params, S1, qbar}, B * kW4, -1.0, tw, st)));
    if (e.dparams) {
        hipLaunchKernelGGL(k_dparams256, dim3(B), dim3(kCropThreads), 0, st, g, qbar, e.dparams);
        HIP_TRY(c, hipGetLastError());
    }
    if (e.da && c->J > 0) {
        HIP_TRY(c, (launch_window_projection<double, 1024>(B, c->J, kWin, c->d_Zw, qbar, Qs, e.da, st, e.da_assign)));
    }
    return ZOLA_OK;
}

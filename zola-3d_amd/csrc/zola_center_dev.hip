// zola_center_dev.hip -- FP64 PSFs and finite-difference Fisher matrices of the
// focus-centre search on the GPU (process/cpu/SearchPSFcenter.java:219-351).
//
// The search needs, at every z of a scan, the reference's FD Fisher matrix of a
// 3000-photon emitter on background 50 with the PSF cropped to the whole FFT
// grid (R = s2 = N): seven PSFs per z (centre, x +- h, y +- h, z +- h).  All the
// PSFs of a scan are one batch: each is the reference pipeline
// (process/cpu/PSFPhaseMany.java:553-622) in FP64 as six passes of the
// detector's line FFT (zola_fft_f64.hpp) with the model fused into their loads
// and stores: pupil field -> inverse transform -> |E|^2 -> forward transform ->
// x float-rounded Gaussian -> inverse transform -> modulus.  With R = s2 = N the
// reference's crop and re-centring maps compose to the identity on |E|^2, and
// the output crop lastS[i] = (i + N/2) % N.  A block per z then normalises the
// seven PSFs and forms the 5 x 5 matrix (computeFisher, 317-403, with the
// photon derivative on the model with step 100 h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/zola_hip.h"
#include "zola_ctx.hpp"
#include "zola_fft_f64.hpp"

namespace zola {
namespace {

struct PupilCell {   // dense fftshifted N x N pupil (amp 0 off the disk)
    double amp, phase, kz, kzo;
};

struct PsfParams {
    double x, y, zoil, zwat;
};

// pass A: rows of the inverse transform of the pupil field (line = psf * N + a)
struct IoPupil {
    const PupilCell* cell;
    const PsfParams* par;
    double2* out;
    int N;
    double cc, kscale;   // centre, 2 pi / (dx N)
    __device__ double2 load(int line, int k) const {
        const int p = line / N, a = line % N;
        const PupilCell t = cell[a * N + k];
        if (t.amp == 0.0) return make_double2(0.0, 0.0);
        const PsfParams q = par[p];
        const int i = (a + N / 2) % N, ii = (k + N / 2) % N;   // grid indices before the shift
        const double ph = (i - cc) * kscale * q.x + (ii - cc) * kscale * q.y + t.phase + t.kz * q.zwat - t.kzo * q.zoil;
        return make_double2(t.amp * cos(ph), t.amp * sin(ph));
    }
    __device__ void store(int line, int k, double2 v) const { out[(size_t)line * N + k] = v; }
};

// columns (line = psf * N + column b, element k = row a) of buf, with a store op
template <class Op>
struct IoCols {
    const double2* in;
    double2* out;
    int N;
    Op op;
    __device__ double2 load(int line, int k) const {
        const int p = line / N, b = line % N;
        return in[((size_t)p * N + k) * N + b];
    }
    __device__ void store(int line, int k, double2 v) const {
        const int p = line / N, b = line % N;
        op(out, p, k, b, N, v);
    }
};

struct IoRows {
    const double2* in;
    double2* out;
    int N;
    __device__ double2 load(int line, int k) const { return in[(size_t)line * N + k]; }
    __device__ void store(int line, int k, double2 v) const { out[(size_t)line * N + k] = v; }
};

struct OpIntensity {   // |E|^2 (real)
    __device__ void operator()(double2* o, int p, int a, int b, int N, double2 v) const {
        o[((size_t)p * N + a) * N + b] = make_double2(v.x * v.x + v.y * v.y, 0.0);
    }
};

struct OpGauss {       // x float-rounded Gaussian, indexed [a][b]
    const double* g;
    __device__ void operator()(double2* o, int p, int a, int b, int N, double2 v) const {
        const double w = g[a * N + b];
        o[((size_t)p * N + a) * N + b] = make_double2(v.x * w, v.y * w);
    }
};

struct OpModulus {     // |.| into the cropped output, out[p][(a + N/2) % N][(b + N/2) % N] (re)
    __device__ void operator()(double2* o, int p, int a, int b, int N, double2 v) const {
        o[((size_t)p * N + (a + N / 2) % N) * N + (b + N / 2) % N] = make_double2(sqrt(v.x * v.x + v.y * v.y), 0.0);
    }
};

template <class IO, bool COL>
hipError_t lines(int N, IO io, int nlines, double sgn, const double2* tw, hipStream_t st) {
    if (N == 64) hipLaunchKernelGGL((k_fft_lines<64, COL, IO>), dim3(fft_lines_grid<64>(nlines)), dim3(256), 0, st, io,
                                    nlines, sgn, tw);
    else if (N == 128) hipLaunchKernelGGL((k_fft_lines<128, COL, IO>), dim3(fft_lines_grid<128>(nlines)), dim3(256), 0,
                                          st, io, nlines, sgn, tw);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// Fisher matrix of z point q from its seven PSFs (unnormalised moduli in re):
// 0 centre, 1/2 x +- h, 3/4 y +- h, 5/6 z +- h.  The photon and background
// derivatives are taken on the model f1 = photons psf + bg itself, as
// computeFisher does (SearchPSFcenter.java:271-276), and the arithmetic is
// not contracted into FMAs, so each term rounds as the reference's does.
__global__ __launch_bounds__(256) void k_center_fisher(int N, const double2* __restrict__ psf, double photons,
                                                       double bg, double h, double* __restrict__ fisher) {
#pragma clang fp contract(off)
    const int q = blockIdx.x;
    const int n = N * N;
    const double2* P = psf + (size_t)q * 7 * n;
    __shared__ double red[256];
    __shared__ double inv_sum[7];
    for (int s = 0; s < 7; s++) {
        double acc = 0.0;
        for (int k = threadIdx.x; k < n; k += 256) acc += P[(size_t)s * n + k].x;
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) inv_sum[s] = 1.0 / red[0];
        __syncthreads();
    }
    double acc[15];
#pragma unroll
    for (int t = 0; t < 15; t++) acc[t] = 0.0;
    const double h2 = 2 * fabs(h);
    for (int k = threadIdx.x; k < n; k += 256) {
        const double f1 = P[k].x * inv_sum[0] * photons + bg;
        double d[5];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double f2 = P[(size_t)(1 + 2 * c) * n + k].x * inv_sum[1 + 2 * c] * photons + bg;
            const double f0 = P[(size_t)(2 + 2 * c) * n + k].x * inv_sum[2 + 2 * c] * photons + bg;
            d[c] = (f2 - f0) / h2;
        }
        d[3] = (f1 * (photons + h * 100) + bg - (f1 * (photons - h * 100) + bg)) / h2;
        d[4] = (f1 * photons + (bg + h) - (f1 * photons + (bg - h))) / h2;
        int t = 0;
#pragma unroll
        for (int a = 0; a < 5; a++)
#pragma unroll
            for (int b = a; b < 5; b++) acc[t++] += (1 / f1) * d[a] * d[b];
    }
    for (int t = 0; t < 15; t++) {
        red[threadIdx.x] = acc[t];
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            int a = 0, b = 0, u = t;
            for (a = 0; a < 5; a++) {
                if (u < 5 - a) { b = a + u; break; }
                u -= 5 - a;
            }
            fisher[(size_t)q * 25 + a * 5 + b] = red[0];
            fisher[(size_t)q * 25 + b * 5 + a] = red[0];
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace zola

using namespace zola;

// FD Fisher matrices (nz x 25, row-major) of the focus-centre search at depths
// zs, focus term zfocus, on the device in FP64.  Called from zola_center.cpp.
// Depths are taken kMaxChunk at a time (two 7 x 256 x N^2 complex buffers,
// 0.9 GB at N = 128).
int zola_center_fisher_dev(zola_ctx* c, int nz, const double* zs, double zfocus, double photons, double bg,
                           double h, double* fisher) {
    constexpr int kMaxChunk = 256;
    const int N = c->N;
    if (N != 64 && N != 128) return set_err(c, ZOLA_EUNSUPPORTED, "focus search covers N = 64 and 128");
    if (nz <= 0) return ZOLA_OK;
    if (!zs || !fisher) return set_err(c, ZOLA_EINVAL, "null depth or output array");
    const size_t n = (size_t)N * N;
    const int P = 7 * std::min(nz, kMaxChunk);
    // dense fftshifted pupil, Gaussian of the blur grid (s2 = N here), twiddles
    std::vector<PupilCell> cell(n, PupilCell{0, 0, 0, 0});
    ensure_phase(c);
    for (int p = 0; p < c->D; p++) {
        const int a = (c->di[p] + N / 2) % N, b = (c->dj[p] + N / 2) % N;
        cell[(size_t)a * N + b] = PupilCell{c->amp[p], c->phase[p], c->kz[p], c->kzo[p]};
    }
    std::vector<double> g(n);
    {
        const double sigF = (1.0 / c->o.sigma) * N / (3.14159265358979323846 * 2.0), sigpow = sigF * sigF;
        const double c2 = (double)(N / 2);
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                const double u = (a + N / 2) % N, v = (b + N / 2) % N;   // nextS
                g[(size_t)a * N + b] = (double)(float)std::exp(-.5 * ((u - c2) * (u - c2) + (v - c2) * (v - c2)) / sigpow);
            }
    }
    std::vector<PsfParams> par(P);
    const double off[7][3] = {{0, 0, 0}, {h, 0, 0}, {-h, 0, 0}, {0, h, 0}, {0, -h, 0}, {0, 0, h}, {0, 0, -h}};
    std::vector<double2> tw(kTwN);
    zola_twiddle_table(tw.data());

    char* base = nullptr;
    const size_t b_cell = n * sizeof(PupilCell), b_g = n * sizeof(double), b_par = P * sizeof(PsfParams),
                 b_tw = kTwN * sizeof(double2), b_buf = (size_t)P * n * sizeof(double2), b_f = (size_t)(P / 7) * 25 * sizeof(double);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = al(b_cell) + al(b_g) + al(b_par) + al(b_tw) + 2 * al(b_buf) + al(b_f);
    HIP_TRY(c, hipMalloc((void**)&base, total));
    char* q0 = base;
    auto take = [&](size_t b) { char* p = q0; q0 += al(b); return p; };
    PupilCell* d_cell = (PupilCell*)take(b_cell);
    double* d_g = (double*)take(b_g);
    PsfParams* d_par = (PsfParams*)take(b_par);
    double2* d_tw = (double2*)take(b_tw);
    double2* A = (double2*)take(b_buf);
    double2* Bf = (double2*)take(b_buf);
    double* d_f = (double*)take(b_f);
    hipStream_t st = c->stream;
    hipError_t e = hipMemcpyAsync(d_cell, cell.data(), b_cell, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_g, g.data(), b_g, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tw, tw.data(), b_tw, hipMemcpyHostToDevice, st);
    const double cc = c->center, ks = 2 * 3.14159265358979323846 / (c->o.xystep * N);
    for (int z0 = 0; z0 < nz && e == hipSuccess; z0 += kMaxChunk) {
        const int m = std::min(kMaxChunk, nz - z0), nl = 7 * m * N;
        for (int q = 0; q < m; q++)
            for (int s = 0; s < 7; s++) par[q * 7 + s] = PsfParams{off[s][0], off[s][1], zfocus, zs[z0 + q] + off[s][2]};
        e = hipMemcpyAsync(d_par, par.data(), (size_t)7 * m * sizeof(PsfParams), hipMemcpyHostToDevice, st);
        // E = IDFT(pupil): rows into A, columns into B as |E|^2
        if (e == hipSuccess) e = lines<IoPupil, false>(N, IoPupil{d_cell, d_par, A, N, cc, ks}, nl, +1.0, d_tw, st);
        if (e == hipSuccess) e = lines<IoCols<OpIntensity>, true>(N, IoCols<OpIntensity>{A, Bf, N, OpIntensity{}}, nl, +1.0, d_tw, st);
        // DFT(|E|^2) x G
        if (e == hipSuccess) e = lines<IoRows, false>(N, IoRows{Bf, A, N}, nl, -1.0, d_tw, st);
        if (e == hipSuccess) e = lines<IoCols<OpGauss>, true>(N, IoCols<OpGauss>{A, Bf, N, OpGauss{d_g}}, nl, -1.0, d_tw, st);
        // IDFT, modulus into the crop
        if (e == hipSuccess) e = lines<IoRows, false>(N, IoRows{Bf, A, N}, nl, +1.0, d_tw, st);
        if (e == hipSuccess) e = lines<IoCols<OpModulus>, true>(N, IoCols<OpModulus>{A, Bf, N, OpModulus{}}, nl, +1.0, d_tw, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_center_fisher, dim3(m), dim3(256), 0, st, N, Bf, photons, bg, h, d_f);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(fisher + (size_t)z0 * 25, d_f, (size_t)m * 25 * sizeof(double), hipMemcpyDeviceToHost, st);
        // par is rewritten by the next chunk: wait for this chunk's copies
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(base);
    if (e != hipSuccess) return set_err(c, ZOLA_EHIP, "focus search on the device: %s", hipGetErrorString(e));
    return ZOLA_OK;
}

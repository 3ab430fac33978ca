// zola_detect.hip -- the frame detector on the GPU, in FP64
// (process/cpu/Predetection.java:63-593, the CPU path the plugin runs).
//
// Per frame: the unitary 2-D FFT of the frame zero-padded to FX x FY =
// pow2(2W) x pow2(2H) (Predetection.setImage 292-329), the windowed mean and
// variance (337-410), then for every template r the inverse FFT of
// F_I conj(F_T[r]) cropped back to W x H and divided by
// sqrt(varPSF[r] var) (convolveNormalizedInFourierDomain 424-478), and the
// local maxima over a 7-pixel row then column filter with parabola
// refinement (getVectMaxPosition 500-593).  FP64 throughout: the candidate
// list (integer positions, template index and order) must be the reference's.
#include <hip/hip_runtime.h>

#include <cmath>

#include "zola_detect.h"
#include "zola_fft_f64.hpp"

namespace zola {

// ---------------------------------------------------------------- window stats
// Running sums in the reference's order: along the second index of each row,
// then along the first index, with no contraction into fused multiply-adds
// (Java evaluates each product and sum separately).
__global__ void k_window_rows(DetGeom g, int F, const float* __restrict__ frames, double* __restrict__ sum1,
                              double* __restrict__ sq1) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F * g.W) return;
    const int H = g.H, h = g.R / 2;
    const float* f = frames + (size_t)t * H;
    double* s = sum1 + (size_t)t * H;
    double* q = sq1 + (size_t)t * H;
    double a = 0.0, b = 0.0;
    for (int u = 0; u < h + 1; u++)
        if (u < H) {
            const double v = (double)f[u];
            a += v;
            b += v * v;
        }
    s[0] = a;
    q[0] = b;
    for (int ii = 1; ii < H; ii++) {
        if (ii - h >= 0) a -= (double)f[ii - h];
        if (ii + h < H) a += (double)f[ii + h];
        if (ii - h >= 0) {
            const double v = (double)f[ii - h];
            b -= v * v;
        }
        if (ii + h < H) {
            const double v = (double)f[ii + h];
            b += v * v;
        }
        s[ii] = a;
        q[ii] = b;
    }
}

__global__ void k_window_cols(DetGeom g, int F, const double* __restrict__ sum1, const double* __restrict__ sq1,
                              const double* __restrict__ norm, double* __restrict__ var) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F * g.H) return;
    const int W = g.W, H = g.H, h = g.R / 2;
    const int f = t / H, ii = t % H;
    const double* s1 = sum1 + (size_t)f * W * H;
    const double* q1 = sq1 + (size_t)f * W * H;
    double* vo = var + (size_t)f * W * H;
    double a = 0.0, b = 0.0;
    for (int u = 0; u < h + 1; u++)
        if (u < W) {
            a += s1[(size_t)u * H + ii];
            b += q1[(size_t)u * H + ii];
        }
    for (int i = 0; i < W; i++) {
        if (i > 0) {
            if (i - h >= 0) a -= s1[(size_t)(i - h) * H + ii];
            if (i + h < W) a += s1[(size_t)(i + h) * H + ii];
            if (i - h >= 0) b -= q1[(size_t)(i - h) * H + ii];
            if (i + h < W) b += q1[(size_t)(i + h) * H + ii];
        }
        const double nrm = norm[(size_t)i * H + ii];
        const double mean = a / nrm;
        const double m2 = mean * mean;
        vo[(size_t)i * H + ii] = (b / nrm) - m2;
    }
}

// ---------------------------------------------------------------- FFT IO functors
// Forward row pass of the frame: line = f * FX + x; row x holds frame row
// ii = (x + W/2) % FX when that is < W, element y holds column
// iii = (y + H/2) % FY when < H (the placement x = (ii - W/2 + FX) % FX of
// Predetection.java:306-309), zero elsewhere.
struct FrameRowIO {
    DetGeom g;
    const float* frames;
    double2* spec;
    __device__ double2 load(int line, int k) const {
        const int f = line / g.FX, x = line % g.FX;
        const int ii = (x + g.W / 2) % g.FX, iii = (k + g.H / 2) % g.FY;
        if (ii >= g.W || iii >= g.H) return make_double2(0.0, 0.0);
        return make_double2((double)frames[((size_t)f * g.W + ii) * g.H + iii], 0.0);
    }
    __device__ void store(int line, int k, double2 v) const {
        spec[(size_t)line * g.FY + k] = v;
    }
};

// Column pass in place over a stack of FX x FY spectra: line = s * FY + y.
struct ColInPlaceIO {
    int FX, FY;
    double2* a;
    __device__ double2 load(int line, int k) const {
        const int s = line / FY, y = line % FY;
        return a[((size_t)s * FX + k) * FY + y];
    }
    __device__ void store(int line, int k, double2 v) const {
        const int s = line / FY, y = line % FY;
        a[((size_t)s * FX + k) * FY + y] = v;
    }
};

// Forward row pass of the templates: line = r * FX + x; template row
// u = (x + R/2) % FX when < R, element uu = (y + R/2) % FY when < R
// (placement of Predetection.makeDictionnary 232-238).
struct TemplateRowIO {
    DetGeom g;
    const double* tmpl;
    double2* tspec;
    __device__ double2 load(int line, int k) const {
        const int r = line / g.FX, x = line % g.FX;
        const int u = (x + g.R / 2) % g.FX, uu = (k + g.R / 2) % g.FY;
        if (u >= g.R || uu >= g.R) return make_double2(0.0, 0.0);
        return make_double2(tmpl[((size_t)r * g.R + u) * g.R + uu], 0.0);
    }
    __device__ void store(int line, int k, double2 v) const {
        tspec[(size_t)line * g.FY + k] = v;
    }
};

// Inverse column pass for templates r0..r0+nr-1: line = (f * nr + rr) * FY + y.
// Input F_I conj(F_T) written as in Predetection.java:449-450; only the W rows
// that the crop keeps are stored: x -> ii = (x + W/2) % FX < W.
// The frame and the templates are real, so each product F_I conj(F_T) is
// Hermitian and its inverse transform real: two templates ra, rb = ra + 1 share
// one complex inverse transform of F_I conj(F_Ta) + i F_I conj(F_Tb), whose real
// and imaginary parts are the two correlations.  Line = (f * np + pp) * FY + y
// for template pair pp (templates r0 + 2 pp and r0 + 2 pp + 1 < nZ).
struct CorrColIO {
    DetGeom g;
    int r0, np;
    const double2* spec;
    const double2* tspec;
    double2* work;
    __device__ double2 load(int line, int k) const {
        const int y = line % g.FY, fp = line / g.FY;
        const int f = fp / np, ra = r0 + 2 * (fp % np), rb = ra + 1;
        const double2 a = spec[((size_t)f * g.FX + k) * g.FY + y];
        const double2 da = tspec[((size_t)ra * g.FX + k) * g.FY + y];
        const double2 xa = make_double2(a.y * da.y + a.x * da.x, a.y * da.x - a.x * da.y);
        if (rb >= g.nZ) return xa;
        const double2 db = tspec[((size_t)rb * g.FX + k) * g.FY + y];
        const double2 xb = make_double2(a.y * db.y + a.x * db.x, a.y * db.x - a.x * db.y);
        return make_double2(xa.x - xb.y, xa.y + xb.x);
    }
    __device__ void store(int line, int k, double2 v) const {
        const int ii = (k + g.W / 2) % g.FX;
        if (ii >= g.W) return;
        const int y = line % g.FY, fp = line / g.FY;
        work[((size_t)fp * g.W + ii) * g.FY + y] = v;
    }
};

// Inverse row pass: line = (f * np + pp) * W + ii; keeps iii = (y + H/2) % FY < H
// and normalises as Predetection.java:463-476, the real part for template
// r0 + 2 pp and the imaginary part for r0 + 2 pp + 1.
struct CorrRowIO {
    DetGeom g;
    int r0, np;
    const double2* work;
    const double* var;
    const double* var_psf;
    double* corr;
    __device__ double2 load(int line, int k) const { return work[(size_t)line * g.FY + k]; }
    __device__ static double normalised(double c, double vp, double vi) {
        const double sq = sqrt(vp * vi);
        if (sq > 0) return c / sq;
        if (vp <= 0) return -1.0;
        if (vi <= 0) return -2.0;
        // var is NaN where the reference's norm is 0 (its size mix-up zeroes it on
        // the same border pixels of every frame): there it keeps the array's
        // previous value, which stays the initial 0.0
        return 0.0;
    }
    __device__ void store(int line, int k, double2 v) const {
        const int iii = (k + g.H / 2) % g.FY;
        if (iii >= g.H) return;
        const int ii = line % g.W, fp = line / g.W;
        const int f = fp / np, ra = r0 + 2 * (fp % np), rb = ra + 1;
        const double vi = var[((size_t)f * g.W + ii) * g.H + iii];
        corr[(((size_t)f * g.nZ + ra) * g.W + ii) * g.H + iii] = normalised(v.x, var_psf[ra], vi);
        if (rb < g.nZ) corr[(((size_t)f * g.nZ + rb) * g.W + ii) * g.H + iii] = normalised(v.y, var_psf[rb], vi);
    }
};

// ---------------------------------------------------------------- local maxima
// Tile of 64 columns (j, one per lane) x TI rows (i) of one frame; 4 waves.
// tmp[i][j] of the reference is carried across templates in registers: it is
// refreshed (max of c_r(i, j-3..j+3)) only where c_r(i, j) > 0 and keeps its
// value from an earlier template elsewhere (Predetection.java:510-519).
constexpr int kLmTI = 16;                 // output rows per tile
constexpr int kLmRows = kLmTI + 6;        // with 3 halo rows each side
constexpr int kLmRowsPerWave = (kLmRows + 3) / 4;

__global__ __launch_bounds__(256) void k_local_max(DetGeom g, const double* __restrict__ corr, int cap,
                                                   DetCand* __restrict__ cand, int* __restrict__ count) {
    __shared__ double cs[kLmRows][64 + 6];
    __shared__ double ts[kLmRows][64];
    const int f = blockIdx.z;
    const int j0 = blockIdx.x * 64, i0 = blockIdx.y * kLmTI - 3;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int W = g.W, H = g.H, nZ = g.nZ;
    const int lo = 3 > g.R / 2 ? 3 : g.R / 2;
    const int j = j0 + lane;
    const bool jin = j >= lo && j < H - lo;
    double tmp[kLmRowsPerWave];
    double cprev[kLmRowsPerWave];
#pragma unroll
    for (int t = 0; t < kLmRowsPerWave; t++) {
        tmp[t] = 0.0;
        cprev[t] = 0.0;
    }
    const double* cf = corr + (size_t)f * nZ * W * H;
    for (int r = 0; r < nZ; r++) {
        const double* c = cf + (size_t)r * W * H;
        // stage rows i0 .. i0 + kLmRows - 1, columns j0 - 3 .. j0 + 66
        for (int e = threadIdx.x; e < kLmRows * 70; e += 256) {
            const int rr = e / 70, cc = e % 70;
            const int i = i0 + rr, jj = j0 - 3 + cc;
            cs[rr][cc] = (i >= 0 && i < W && jj >= 0 && jj < H) ? c[(size_t)i * H + jj] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kLmRowsPerWave; t++) {
            const int rr = w + 4 * t;
            if (rr < kLmRows) {
                const int i = i0 + rr;
                if (jin && i >= lo && i < W - lo && cs[rr][lane + 3] > 0) {
                    double m = -INFINITY;
#pragma unroll
                    for (int a = 0; a < 7; a++) m = fmax(cs[rr][lane + a], m);
                    tmp[t] = m;
                }
                ts[rr][lane] = tmp[t];
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kLmRowsPerWave; t++) {
            const int rr = w + 4 * t;
            if (rr >= 3 && rr < 3 + kLmTI) {
                const int i = i0 + rr;
                const double v = cs[rr][lane + 3];
                if (jin && i >= lo && i < W - lo && v > g.thr && v == ts[rr][lane]) {
                    double m = -INFINITY;
#pragma unroll
                    for (int a = -3; a <= 3; a++) m = fmax(ts[rr + a][lane], m);
                    if (m == v) {
                        // topOfParabola (Predetection.java:588-593) through (-s, ya), (0, yb), (s, yc)
                        auto top = [](double s, double ya, double yb, double yc) {
                            const double xa = -s, xb = 0.0, xc = s;
                            const double aa = (yc - ya) / ((xc - xa) * (xc - xb)) - (yb - ya) / ((xb - xa) * (xc - xb));
                            const double bb = ((yb - ya) / (xb - xa)) - aa * (xb + xa);
                            return (-bb / (2 * aa));
                        };
                        double dz = 0.0, dx = 0.0, dy = 0.0;
                        if (r > 1 && r < nZ - 1) {
                            const double yc = cf[((size_t)(r + 1) * W + i) * H + j];
                            dz = top(1. * g.zstep, cprev[t], v, yc);
                        }
                        if (i > 1 && i < W - 1) dx = top(1. * g.xystep, cs[rr - 1][lane + 3], v, cs[rr + 1][lane + 3]);
                        if (j > 1 && j < H - 1) dy = top(1. * g.xystep, cs[rr][lane + 2], v, cs[rr][lane + 4]);
                        const int k = atomicAdd(&count[f], 1);
                        if (k < cap) {
                            DetCand d;
                            d.score = v;
                            d.dz = dz;
                            d.dx = dx;
                            d.dy = dy;
                            d.r = r;
                            d.i = i;
                            d.j = j;
                            d.pad = 0;
                            cand[(size_t)f * cap + k] = d;
                        }
                    }
                }
                cprev[t] = v;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- ordering
// rank(e) = number of candidates of the frame that precede e; the order is by
// score descending, then by the reference's scan order (r, j, i) -- the order in
// which getVectMaxPosition appends them, which Arrays.sort keeps for ties.
__device__ __forceinline__ bool cand_before(double sa, long long ka, double sb, long long kb) {
    return sa > sb || (sa == sb && ka < kb);
}

__device__ __forceinline__ long long cand_key(const DetCand& d) {
    return ((long long)d.r << 40) | ((long long)d.j << 20) | (long long)d.i;
}

__global__ __launch_bounds__(256) void k_rank_sort(int cap, const DetCand* __restrict__ cand,
                                                   const int* __restrict__ count, double* __restrict__ out) {
    __shared__ double ss[256];
    __shared__ long long sk[256];
    const int f = blockIdx.y;
    int n = count[f];
    if (n > cap) n = cap;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x * 256 >= n) return;
    const DetCand* cf = cand + (size_t)f * cap;
    DetCand me{};
    double ms = 0.0;
    long long mk = 0;
    if (e < n) {
        me = cf[e];
        ms = me.score;
        mk = cand_key(me);
    }
    int rank = 0;
    for (int base = 0; base < n; base += 256) {
        __syncthreads();
        const int o = base + threadIdx.x;
        if (o < n) {
            const DetCand d = cf[o];
            ss[threadIdx.x] = d.score;
            sk[threadIdx.x] = cand_key(d);
        }
        __syncthreads();
        const int lim = n - base < 256 ? n - base : 256;
        for (int q = 0; q < lim; q++) rank += cand_before(ss[q], sk[q], ms, mk) ? 1 : 0;
    }
    if (e < n) {
        double* o = out + ((size_t)f * cap + rank) * 7;
        o[0] = me.score;
        o[1] = (double)me.r;
        o[2] = (double)me.i;
        o[3] = (double)me.j;
        o[4] = me.dz;
        o[5] = me.dx;
        o[6] = me.dy;
    }
}

// ---------------------------------------------------------------- ROI gather
__global__ void k_gather_rois(int W, int H, int R, int B, const int* __restrict__ roi,
                              const float* __restrict__ frames, float* __restrict__ pixels) {
    const int b = blockIdx.x;
    if (b >= B) return;
    const int f = roi[b * 3 + 0], ci = roi[b * 3 + 1], cj = roi[b * 3 + 2];
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
        const int a = e / R, bb = e % R;
        const int i = ci + a - R / 2, jj = cj + bb - R / 2;
        float v = 0.f;
        if (i >= 0 && i < W && jj >= 0 && jj < H) v = frames[((size_t)f * W + i) * H + jj];
        pixels[(size_t)b * R * R + e] = v;
    }
}

// ---------------------------------------------------------------- launchers
template <class IO, bool COL>
static hipError_t launch_lines(int L, IO io, int nlines, double sgn, const double2* tw, hipStream_t st) {
    switch (L) {
// column passes of 512 and longer lines: 4096 elements per block (8 columns,
// 128-byte runs per row instead of 64)
#define ZOLA_L(LL)                                                                                   \
    case LL: {                                                                                       \
        constexpr int E = (COL && LL >= 512) ? 4096 : kFftElems;                                     \
        hipLaunchKernelGGL((k_fft_lines<LL, COL, IO, E>), dim3(fft_lines_grid<LL, E>(nlines)), dim3(256), \
                           0, st, io, nlines, sgn, tw);                                              \
        break;                                                                                       \
    }
        ZOLA_L(16) ZOLA_L(32) ZOLA_L(64) ZOLA_L(128) ZOLA_L(256) ZOLA_L(512) ZOLA_L(1024) ZOLA_L(2048)
#undef ZOLA_L
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace zola

using namespace zola;

hipError_t zola_launch_window_stats(const DetGeom& g, int F, const float* frames, double* sum1, double* sq1,
                                    const double* norm, double* var, hipStream_t st) {
    const int n1 = F * g.W, n2 = F * g.H;
    hipLaunchKernelGGL(k_window_rows, dim3((n1 + 63) / 64), dim3(64), 0, st, g, F, frames, sum1, sq1);
    hipLaunchKernelGGL(k_window_cols, dim3((n2 + 63) / 64), dim3(64), 0, st, g, F, sum1, sq1, norm, var);
    return hipGetLastError();
}

hipError_t zola_launch_frame_spectrum(const DetGeom& g, int F, const float* frames, double2* spec,
                                      const double2* tw, hipStream_t st) {
    FrameRowIO rio{g, frames, spec};
    hipError_t e = launch_lines<FrameRowIO, false>(g.FY, rio, F * g.FX, -1.0, tw, st);
    if (e != hipSuccess) return e;
    ColInPlaceIO cio{g.FX, g.FY, spec};
    return launch_lines<ColInPlaceIO, true>(g.FX, cio, F * g.FY, -1.0, tw, st);
}

hipError_t zola_launch_template_spectrum(const DetGeom& g, const double* tmpl, double2* tspec,
                                         const double2* tw, hipStream_t st) {
    TemplateRowIO rio{g, tmpl, tspec};
    hipError_t e = launch_lines<TemplateRowIO, false>(g.FY, rio, g.nZ * g.FX, -1.0, tw, st);
    if (e != hipSuccess) return e;
    ColInPlaceIO cio{g.FX, g.FY, tspec};
    return launch_lines<ColInPlaceIO, true>(g.FX, cio, g.nZ * g.FY, -1.0, tw, st);
}

hipError_t zola_launch_correlation(const DetGeom& g, int F, int r0, int np, const double2* spec,
                                   const double2* tspec, const double* var, const double* var_psf,
                                   double2* work, double* corr, const double2* tw, hipStream_t st) {
    CorrColIO cio{g, r0, np, spec, tspec, work};
    hipError_t e = launch_lines<CorrColIO, true>(g.FX, cio, F * np * g.FY, +1.0, tw, st);
    if (e != hipSuccess) return e;
    CorrRowIO rio{g, r0, np, work, var, var_psf, corr};
    return launch_lines<CorrRowIO, false>(g.FY, rio, F * np * g.W, +1.0, tw, st);
}

hipError_t zola_launch_local_max(const DetGeom& g, int F, const double* corr, int cap, DetCand* cand,
                                 int* count, hipStream_t st) {
    dim3 grid((g.H + 63) / 64, (g.W + kLmTI - 1) / kLmTI, F);
    hipLaunchKernelGGL(k_local_max, grid, dim3(256), 0, st, g, corr, cap, cand, count);
    return hipGetLastError();
}

hipError_t zola_launch_rank_sort(int F, int cap, int max_count, const DetCand* cand, const int* count,
                                 double* out, hipStream_t st) {
    if (max_count <= 0 || F <= 0) return hipSuccess;
    int n = max_count < cap ? max_count : cap;
    dim3 grid((n + 255) / 256, F);
    hipLaunchKernelGGL(k_rank_sort, grid, dim3(256), 0, st, cap, cand, count, out);
    return hipGetLastError();
}

hipError_t zola_launch_gather_rois(int W, int H, int R, int B, const int* roi, const float* frames,
                                   float* pixels, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rois, dim3(B), dim3(256), 0, st, W, H, R, B, roi, frames, pixels);
    return hipGetLastError();
}

// Statistics of the fitted model mu = A psf + B (+ the sCMOS term of the pixel
// when vpx is given, as in Localization.computeLikelihood) that the acceptance walk uses
// (LocalizationPipeline.java:862-873, 924-938 and scoreCompute, 1099-1115), one
// ROI per thread, in FP64 with the host's operation order and no contraction,
// so that they equal the host's values bit for bit:
//   mx[b]    = max(0, max_p (mu_p - B))
//   score[b] = sum_p (n_p - mu_p)^2 / mu_p / R^2, in pixel order
//   bits[b]  = bit p set where (mu_p - B) > frac * mx[b]   (ceil(R^2/32) words)
__global__ void k_model_stats(int B, int R, const float* __restrict__ psf, const float* __restrict__ px,
                              const float* __restrict__ vpx, const double* __restrict__ th, double frac,
                              double* __restrict__ mx, double* __restrict__ score, uint32_t* __restrict__ bits) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int n = R * R, nw = (n + 31) / 32;
    const double A = th[(size_t)b * 5 + 0], Bg = th[(size_t)b * 5 + 1];
    const float* ps = psf + (size_t)b * n;
    const float* pp = px + (size_t)b * n;
    const float* pv = vpx ? vpx + (size_t)b * n : nullptr;
    double m = 0.0;
    for (int p = 0; p < n; p++) {
        const double mu = pv ? (double)ps[p] * A + Bg + (double)pv[p] : (double)ps[p] * A + Bg;
        if (m < (mu - Bg)) m = mu - Bg;
    }
    const double thr = m * frac;
    double sum = 0.0, cnt = 0.0;
    for (int w = 0; w < nw; w++) {
        uint32_t word = 0;
        for (int t = 0; t < 32; t++) {
            const int p = w * 32 + t;
            if (p >= n) break;
            const double mu = pv ? (double)ps[p] * A + Bg + (double)pv[p] : (double)ps[p] * A + Bg;
            const double v = (double)pp[p];
            sum += (v - mu) * (v - mu) / mu;
            cnt++;
            if ((mu - Bg) > thr) word |= 1u << t;
        }
        bits[(size_t)b * nw + w] = word;
    }
    mx[b] = m;
    score[b] = sum / cnt;
}

hipError_t zola_launch_model_stats(int B, int R, const float* psf, const float* px, const float* vpx,
                                   const double* th, double frac, double* mx, double* score, uint32_t* bits,
                                   hipStream_t st) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_model_stats, dim3((B + 63) / 64), dim3(64), 0, st, B, R, psf, px, vpx, th, frac, mx,
                       score, bits);
    return hipGetLastError();
}

// Patches of one W x H map (the sCMOS var/gain^2 map) at the ROI centres.
__global__ void k_gather_map(int W, int H, int R, int B, const int* __restrict__ roi, const float* __restrict__ map,
                             float* __restrict__ out) {
    const int b = blockIdx.x;
    if (b >= B) return;
    const int ci = roi[b * 3 + 1], cj = roi[b * 3 + 2];
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
        const int i = ci + e / R - R / 2, jj = cj + e % R - R / 2;
        out[(size_t)b * R * R + e] = (i >= 0 && i < W && jj >= 0 && jj < H) ? map[(size_t)i * H + jj] : 0.f;
    }
}

hipError_t zola_launch_gather_map(int W, int H, int R, int B, const int* roi, const float* map, float* out,
                                  hipStream_t st) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_map, dim3(B), dim3(256), 0, st, W, H, R, B, roi, map, out);
    return hipGetLastError();
}

namespace {
// PSF parameters [x, y, z, zoil] (float) of fitted theta [A, B, x, y, z]: the
// model of the last likelihood (Localization.getPSF) at the fit's focus term.
__global__ void k_theta_params(int B, const double* __restrict__ th, float zfocus, float* __restrict__ par) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    par[b * 4 + 0] = (float)th[b * 5 + 2];
    par[b * 4 + 1] = (float)th[b * 5 + 3];
    par[b * 4 + 2] = (float)th[b * 5 + 4];
    par[b * 4 + 3] = zfocus;
}
}  // namespace

hipError_t zola_launch_theta_params(int B, const double* th, double zfocus, float* par, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_theta_params, dim3((B + 255) / 256), dim3(256), 0, st, B, th, (float)zfocus, par);
    return hipGetLastError();
}

void zola_twiddle_table(double2* t) {
    const double pi = 3.14159265358979323846;
    for (int a = 0; a < kTwN; a++) {
        const double ang = 2.0 * pi * (double)a / (double)kTwN;
        t[a] = make_double2(std::cos(ang), std::sin(ang));
    }
}

// zola_beads.hip -- calibration bead preparation on the GPU, replacing
// InitBackgroundAndPhotonNumber (process/InitBackgroundAndPhotonNumber.java:
// 50-97 constructors, 99-468 initBackgroundAndPhotonNumberMany, 563-618
// refinePoints, 622-904 computeBackground, with utils/PolynomialFit.java at
// order 2 and utils/GaussianElimination.java).
//
// The whole stack stays resident in HBM as FP64 (the reference's
// double[S][W][H], [z][i][ii] with ii fastest).  Work that the reference does
// per output (z-sums of the seed patches, per-pixel means over the dilated
// mask, the photon sums of each bead slice, the slice filter, the polynomial
// evaluation) runs one thread per output with the reference's loop order and
// no contraction, so it is bit-identical given the same inputs.  The global
// moments of the stack, the per-slice means and the 6x6 normal equations are
// tree reductions of fixed shape (deterministic, but summed in another order
// than the reference's single running sum).  The few scalar steps (seed
// centroid, slice argmin, the elimination) run on the host in the reference's
// order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "../../include/zola_hip.h"
#include "zola_ctx.hpp"

namespace {

#pragma clang fp contract(off)

constexpr int kT = 256;        // threads per block
constexpr int kRedBlocks = 1024;  // blocks of the grid-stride stack reductions

__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Camera conversion (constructors, 50-97) fused with the partial sums of the
// stack mean.  CCD/EMCCD: (v*slope + intercept) + 1 with slope = adu/gain,
// intercept = -offset*adu/gain; sCMOS: (v - offset)/gain per pixel.  Both
// clamp at 0.
__global__ void __launch_bounds__(kT) k_convert(long n, long wh, const float* __restrict__ adu, double slope,
                                                double intercept, const float* __restrict__ off,
                                                const float* __restrict__ gain, double* __restrict__ img,
                                                double* __restrict__ part) {
    __shared__ double sh[kT];
    double s = 0;
    for (long k = (long)blockIdx.x * kT + threadIdx.x; k < n; k += (long)gridDim.x * kT) {
        double v;
        if (off) {
            const long q = k % wh;
            v = ((double)adu[k] - (double)off[q]) / (double)gain[q];
        } else {
            v = ((double)adu[k] * slope + intercept) + 1;
        }
        v = v < 0 ? 0 : v;
        img[k] = v;
        s += v;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Partial sums of the three further moments of computeBackground (636-690):
// pass 1: sum (v - mean0)^2; pass 2: sum v and count over v <= mean0 + 2 std0;
// pass 3: sum (v - mean)^2 over v <= mean0 + 3 std0.  st = {mean0, std0, mean, std}.
__global__ void __launch_bounds__(kT) k_moment(long n, const double* __restrict__ img,
                                               const double* __restrict__ st, int pass,
                                               double* __restrict__ part) {
    __shared__ double sh[kT];
    const double mean0 = st[0], std0 = st[1], mean = st[2];
    double s = 0, c = 0;
    for (long k = (long)blockIdx.x * kT + threadIdx.x; k < n; k += (long)gridDim.x * kT) {
        const double v = img[k];
        if (pass == 1) {
            s += (v - mean0) * (v - mean0);
        } else if (pass == 2) {
            if (v <= mean0 + 2 * std0) {
                s += v;
                c += 1;
            }
        } else if (v <= mean0 + 3 * std0) {
            s += (v - mean) * (v - mean);
        }
    }
    s = block_sum(s, sh);
    if (pass == 2) c = block_sum(c, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s;
        if (pass == 2) part[kRedBlocks + blockIdx.x] = c;
    }
}

// One block folds the partials of a pass in a fixed order and finishes the
// statistic as the reference does (int divisor nint for mean0/std0, the count
// of pass 2 for mean and std).  st[4] keeps that count.
__global__ void __launch_bounds__(kT) k_moment_final(const double* __restrict__ part, int pass, int nint,
                                                     double* __restrict__ st) {
    __shared__ double sh[kT];
    double s = 0, c = 0;
    for (int k = threadIdx.x; k < kRedBlocks; k += kT) {
        s += part[k];
        if (pass == 2) c += part[kRedBlocks + k];
    }
    s = block_sum(s, sh);
    if (pass == 2) c = block_sum(c, sh);
    if (threadIdx.x != 0) return;
    if (pass == 0) {
        st[0] = s / nint;
    } else if (pass == 1) {
        st[1] = sqrt(s / nint);
    } else if (pass == 2) {
        st[2] = c > 0 ? s / c : s;
        st[4] = c;
    } else {
        st[3] = st[4] > 0 ? sqrt(s / st[4]) : s;
    }
}

// is2 (v <= mean + 3 std, 692-704) and per-slice partial sums of the kept
// values (709-724).  Grid (chunks, S).
__global__ void __launch_bounds__(kT) k_is2(long wh, const double* __restrict__ img,
                                            const double* __restrict__ st, unsigned char* __restrict__ is2,
                                            double* __restrict__ part) {
    __shared__ double sh[kT];
    const double thr = st[2] + 3 * st[3];
    const long base = (long)blockIdx.y * wh;
    double s = 0, c = 0;
    for (long k = (long)blockIdx.x * kT + threadIdx.x; k < wh; k += (long)gridDim.x * kT) {
        const double v = img[base + k];
        const bool keep = !(v > thr);
        is2[base + k] = keep ? 1 : 0;
        if (keep) {
            s += v;
            c += 1;
        }
    }
    s = block_sum(s, sh);
    c = block_sum(c, sh);
    if (threadIdx.x == 0) {
        part[((long)blockIdx.y * gridDim.x + blockIdx.x) * 2] = s;
        part[((long)blockIdx.y * gridDim.x + blockIdx.x) * 2 + 1] = c;
    }
}

__global__ void __launch_bounds__(kT) k_slice_final(int chunks, const double* __restrict__ part,
                                                    double* __restrict__ mstack) {
    __shared__ double sh[kT];
    double s = 0, c = 0;
    for (int k = threadIdx.x; k < chunks; k += kT) {
        s += part[((long)blockIdx.x * chunks + k) * 2];
        c += part[((long)blockIdx.x * chunks + k) * 2 + 1];
    }
    s = block_sum(s, sh);
    c = block_sum(c, sh);
    if (threadIdx.x == 0) mstack[blockIdx.x] = c != 0 ? s / c : s;
}

// 5x5 in-plane AND of is2 around (ii, iii) in slice j; slices outside the
// stack impose nothing.
__device__ __forceinline__ int plane_and(const unsigned char* __restrict__ is2, int S, int W, int H, int j,
                                         int ii, int iii) {
    if (j < 0 || j >= S) return 1;
    const unsigned char* p = is2 + (long)j * W * H;
    int ok = 1;
    for (int uu = -2; uu <= 2; uu++) {
        const int a = ii + uu;
        if (a < 0 || a >= W) continue;
        for (int uuu = -2; uuu <= 2; uuu++) {
            const int b = iii + uuu;
            if (b < 0 || b >= H) continue;
            ok &= p[(long)a * H + b];
        }
    }
    return ok;
}

// is3 (the slices mp-nN..mp+nN whose 5x5x5 neighbourhood is all in is2,
// 760-798) and the per-pixel mean over it (810-845).  Thread per pixel; the
// z window of plane ANDs slides down the stack.
__global__ void __launch_bounds__(kT) k_pixel_mean(int S, int W, int H, int mp, int nN,
                                                   const double* __restrict__ img,
                                                   const unsigned char* __restrict__ is2, double* __restrict__ pm,
                                                   int* __restrict__ cnt) {
    const long q = (long)blockIdx.x * kT + threadIdx.x;
    if (q >= (long)W * H) return;
    const int ii = (int)(q / H), iii = (int)(q % H);
    const int lo = std::max(0, mp - nN), hi = std::min(S - 1, mp + nN);
    int count = 0;
    double m = 0;
    if (lo <= hi) {
        int w[5];
        for (int u = 0; u < 5; u++) w[u] = plane_and(is2, S, W, H, lo - 2 + u, ii, iii);
        for (int i = lo; i <= hi; i++) {
            if (i > lo) {
                for (int u = 0; u < 4; u++) w[u] = w[u + 1];
                w[4] = plane_and(is2, S, W, H, i + 2, ii, iii);
            }
            if (w[0] & w[1] & w[2] & w[3] & w[4]) {
                m += img[(long)i * W * H + q];
                count++;
            }
        }
    }
    cnt[q] = count;
    pm[q] = count >= 1 ? m / count : 0.0;
}

// Normal equations of the order-2 fit (PolynomialFit.run, 92-153 and the
// Computation thread 557-606) over the pixels with a mean.  Each pixel term is
// exact or singly rounded as in the reference: moments x^a y^b (a + b <= 4,
// exact for integer coordinates) and the six right-hand terms (x^e0 y^e1) * m.
// Per block: 15 moments, 6 right-hand sums, the data count.
constexpr int kNE = 22;
__constant__ int c_ma[15] = {0, 1, 2, 3, 4, 0, 1, 2, 3, 0, 1, 2, 0, 1, 0};
__constant__ int c_mb[15] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4};
__constant__ int c_e0[6] = {0, 1, 2, 0, 1, 0};
__constant__ int c_e1[6] = {0, 0, 0, 1, 1, 2};

__device__ __forceinline__ double ipow(double x, int e) { return e == 0 ? 1.0 : (e == 1 ? x : x * x); }

__global__ void __launch_bounds__(kT) k_normal(int W, int H, const double* __restrict__ pm,
                                               const int* __restrict__ cnt, double* __restrict__ part) {
    __shared__ double sh[kT];
    const long q = (long)blockIdx.x * kT + threadIdx.x;
    const bool in = q < (long)W * H && cnt[q] >= 1;
    const double x = in ? (double)(q / H) : 0.0, y = in ? (double)(q % H) : 0.0;
    const double m = in ? pm[q] : 0.0;
    for (int k = 0; k < 15; k++) {
        double t = 0;
        if (in) {
            double px = 1, py = 1;
            for (int a = 0; a < c_ma[k]; a++) px *= x;
            for (int b = 0; b < c_mb[k]; b++) py *= y;
            t = px * py;
        }
        t = block_sum(t, sh);
        if (threadIdx.x == 0) part[(long)blockIdx.x * kNE + k] = t;
    }
    for (int i = 0; i < 6; i++) {
        double t = 0;
        if (in) {
            double pr = 1;
            pr *= ipow(x, c_e0[i]);
            pr *= ipow(y, c_e1[i]);
            pr *= m;
            t = pr;
        }
        t = block_sum(t, sh);
        if (threadIdx.x == 0) part[(long)blockIdx.x * kNE + 15 + i] = t;
    }
    const double c = block_sum(in ? 1.0 : 0.0, sh);
    if (threadIdx.x == 0) part[(long)blockIdx.x * kNE + 21] = c;
}

__global__ void __launch_bounds__(kT) k_normal_final(int nblk, const double* __restrict__ part,
                                                     double* __restrict__ out) {
    if ((int)threadIdx.x >= kNE) return;
    double s = 0;
    for (int b = 0; b < nblk; b++) s += part[(long)b * kNE + threadIdx.x];
    out[threadIdx.x] = s;
}

struct Coef6 {
    double a[6];
};

// PolynomialFit.transform (191-222): res += (a_i * x^e0) * y^e1, in index order.
__global__ void __launch_bounds__(kT) k_poly_eval(int W, int H, Coef6 c, double* __restrict__ bg) {
    const long q = (long)blockIdx.x * kT + threadIdx.x;
    if (q >= (long)W * H) return;
    const double x = (double)(q / H), y = (double)(q % H);
    double res = 0;
    for (int i = 0; i < 6; i++) {
        double prod = c.a[i];
        prod *= ipow(x, c_e0[i]);
        prod *= ipow(y, c_e1[i]);
        res += prod;
    }
    bg[q] = res;
}

// Seed patches summed over z (refinePoints 565-577): thread per (bead, u, uu).
__global__ void __launch_bounds__(kT) k_refine_sum(int S, int W, int H, int P, int nb, const int* __restrict__ org,
                                                   const double* __restrict__ img, double* __restrict__ sub) {
    const long t = (long)blockIdx.x * kT + threadIdx.x;
    if (t >= (long)nb * P * P) return;
    const int p = (int)(t / (P * P)), u = (int)(t / P % P), uu = (int)(t % P);
    const int a = org[2 * p] + u, b = org[2 * p + 1] + uu;
    double s = 0;
    if (a >= 0 && a < W && b >= 0 && b < H)
        for (int i = 0; i < S; i++) s += img[((long)i * W + a) * H + b];
    sub[t] = s;
}

// Per-bead background B (mean of the polynomial over the clipped square,
// 245-262): thread per bead.
__global__ void k_bead_b(int W, int H, int P, int nb, const int* __restrict__ org, const double* __restrict__ bg,
                         double* __restrict__ B) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nb) return;
    double s = 0, c = 0;
    for (int u = 0; u < P; u++)
        for (int uu = 0; uu < P; uu++) {
            const int a = org[2 * p] + u, b = org[2 * p + 1] + uu;
            if (a >= 0 && a < W && b >= 0 && b < H) {
                s += bg[(long)a * H + b];
                c++;
            }
        }
    if (c >= 1) s /= c;
    B[p] = s;
}

// Photon sum of each bead slice (272-302): thread per (slice z, bead p),
// output A0[z * nb + p] as the reference's flattened array.
__global__ void __launch_bounds__(kT) k_bead_a0(int S, int W, int H, int P, int nb, const int* __restrict__ org,
                                                const double* __restrict__ img, const double* __restrict__ B,
                                                double* __restrict__ A0) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= S * nb) return;
    const int z = t / nb, p = t % nb;
    const double b0 = B[p];
    double s = 0;
    for (int u = 0; u < P; u++) {
        const int a = org[2 * p] + u;
        if (a < 0 || a >= W) continue;
        for (int uu = 0; uu < P; uu++) {
            const int b = org[2 * p + 1] + uu;
            if (b >= 0 && b < H) s += img[((long)z * W + a) * H + b] - b0;
        }
    }
    A0[t] = s;
}

// Moving mean of the photon curves over +-sF slices (318-343): A[p][z].
__global__ void __launch_bounds__(kT) k_bead_a(int S, int nb, int sF, const double* __restrict__ A0,
                                               double* __restrict__ A) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= S * nb) return;
    const int p = t / S, z = t % S;
    double m = 0, c = 0;
    for (int f = -sF; f <= sF; f++)
        if (z + f >= 0 && z + f < S) {
            m += A0[(z + f) * nb + p];
            c++;
        }
    if (c >= 1) m /= c;
    A[t] = m;
}

// Bead patches (417-431): image (+ var/gain^2 for sCMOS) inside the frame, B
// outside.  sCMOS patches of var/gain^2 (0 outside the frame): the reference
// re-allocates its patch array inside the bead loop (433-447), so only the
// last bead's patch holds the map and the others stay zero; kept as is.
// Thread per output pixel.
__global__ void __launch_bounds__(kT) k_patches(int S, int W, int H, int P, int nb, const int* __restrict__ org,
                                                const double* __restrict__ img, const double* __restrict__ vg,
                                                const double* __restrict__ B, double* __restrict__ patches,
                                                double* __restrict__ spatch) {
    const long t = (long)blockIdx.x * kT + threadIdx.x;
    const long n = (long)nb * S * P * P;
    if (t >= n) return;
    const int uu = (int)(t % P), u = (int)(t / P % P), z = (int)(t / ((long)P * P) % S);
    const int p = (int)(t / ((long)S * P * P));
    const int a = org[2 * p] + u, b = org[2 * p + 1] + uu;
    const bool in = a >= 0 && a < W && b >= 0 && b < H;
    double v = B[p];
    if (in) {
        const long k = (long)a * H + b;
        v = vg ? img[((long)z * W + a) * H + b] + vg[k] : img[((long)z * W + a) * H + b];
    }
    patches[t] = v;
    if (spatch && z == 0)
        spatch[((long)p * P + u) * P + uu] = (in && vg && p == nb - 1) ? vg[(long)a * H + b] : 0.0;
}

// var / gain^2 in FP64 from the float camera maps (utils/SCMOScamera.java:95).
__global__ void __launch_bounds__(kT) k_vargain(long wh, const float* __restrict__ var,
                                                const float* __restrict__ gain, double* __restrict__ vg) {
    const long k = (long)blockIdx.x * kT + threadIdx.x;
    if (k >= wh) return;
    const double g = gain[k];
    vg[k] = (double)var[k] / (g * g);
}

// GaussianElimination.lsolve (double): partial pivoting on the first row with
// the strictly largest |a|, singular when |pivot| <= 1e-20.
bool lsolve6(double (&a)[6][6], double (&b)[6], double (&x)[6]) {
    const int n = 6;
    for (int p = 0; p < n; p++) {
        int mx = p;
        for (int i = p + 1; i < n; i++)
            if (std::fabs(a[i][p]) > std::fabs(a[mx][p])) mx = i;
        for (int j = 0; j < n; j++) std::swap(a[p][j], a[mx][j]);
        std::swap(b[p], b[mx]);
        if (std::fabs(a[p][p]) <= 1e-20) return false;
        for (int i = p + 1; i < n; i++) {
            const double alpha = a[i][p] / a[p][p];
            b[i] -= alpha * b[p];
            for (int j = p; j < n; j++) a[i][j] -= alpha * a[p][j];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = 0.0;
        for (int j = i + 1; j < n; j++) sum += a[i][j] * x[j];
        x[i] = (b[i] - sum) / a[i][i];
    }
    return true;
}

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t get(T** out, size_t count) {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return e;
    }
};

inline unsigned nblocks(long n) { return (unsigned)((n + kT - 1) / kT); }

int bead_prepare(zola_ctx* c, int S, int W, int H, const float* adu, double slope, double intercept,
                 const float* offset, const float* gain, const float* variance, double zstep, int P, int nb,
                 float* xp, float* yp, double* patches, double* B, double* A, double* spatches, double* background,
                 zola_bead_diag* diag) {
    if (!c) return ZOLA_EINVAL;
    if (S <= 0 || W <= 0 || H <= 0 || P <= 0 || nb < 0 || !adu || !(zstep > 0))
        return set_err(c, ZOLA_EINVAL, "bead_prepare: bad sizes (S %d, W %d, H %d, P %d, nb %d, zstep %g)", S, W, H,
                       P, nb, zstep);
    if (S > 65535) return set_err(c, ZOLA_EINVAL, "bead_prepare: %d slices exceed 65535", S);
    if ((long)S * W * H > INT_MAX)
        return set_err(c, ZOLA_EINVAL, "bead_prepare: stack of %ld voxels exceeds 2^31 - 1", (long)S * W * H);
    if (nb > 0 && (!xp || !yp || !patches || !B || !A))
        return set_err(c, ZOLA_EINVAL, "bead_prepare: NULL seed or output array");
    const hipStream_t st = c->stream;
    const long wh = (long)W * H, n = (long)S * wh;
    const int sF = std::max((int)(0.5 / zstep) / 2, 1);

    DevBuf mem;
    float *d_adu = nullptr, *d_off = nullptr, *d_gain = nullptr, *d_var = nullptr;
    double *d_img = nullptr, *d_part = nullptr, *d_st = nullptr, *d_mstack = nullptr, *d_pm = nullptr;
    double *d_ne = nullptr, *d_nes = nullptr, *d_bg = nullptr, *d_vg = nullptr, *d_sub = nullptr;
    double *d_B = nullptr, *d_A0 = nullptr, *d_A = nullptr, *d_patch = nullptr, *d_sp = nullptr;
    unsigned char* d_is2 = nullptr;
    int *d_cnt = nullptr, *d_org = nullptr;
    const int chunks = (int)std::min<long>(64, (wh + kT - 1) / kT);
    const int nblk = (int)nblocks(wh);
    HIP_TRY(c, mem.get(&d_adu, n));
    HIP_TRY(c, mem.get(&d_img, n));
    HIP_TRY(c, mem.get(&d_is2, n));
    HIP_TRY(c, mem.get(&d_part, std::max<long>(2L * kRedBlocks, 2L * S * chunks)));
    HIP_TRY(c, mem.get(&d_st, 8));
    HIP_TRY(c, mem.get(&d_mstack, S));
    HIP_TRY(c, mem.get(&d_pm, wh));
    HIP_TRY(c, mem.get(&d_cnt, wh));
    HIP_TRY(c, mem.get(&d_ne, (size_t)nblk * kNE));
    HIP_TRY(c, mem.get(&d_nes, kNE));
    HIP_TRY(c, mem.get(&d_bg, wh));
    HIP_TRY(c, mem.get(&d_org, 2 * std::max(nb, 1)));
    HIP_TRY(c, mem.get(&d_sub, (size_t)std::max(nb, 1) * P * P));
    HIP_TRY(c, mem.get(&d_B, std::max(nb, 1)));
    HIP_TRY(c, mem.get(&d_A0, (size_t)std::max(nb, 1) * S));
    HIP_TRY(c, mem.get(&d_A, (size_t)std::max(nb, 1) * S));
    HIP_TRY(c, mem.get(&d_patch, (size_t)std::max(nb, 1) * S * P * P));
    if (offset) {
        HIP_TRY(c, mem.get(&d_off, wh));
        HIP_TRY(c, mem.get(&d_gain, wh));
        HIP_TRY(c, mem.get(&d_var, wh));
        HIP_TRY(c, mem.get(&d_vg, wh));
        HIP_TRY(c, mem.get(&d_sp, (size_t)std::max(nb, 1) * P * P));
        HIP_TRY(c, hipMemcpyAsync(d_off, offset, wh * sizeof(float), hipMemcpyHostToDevice, st));
        HIP_TRY(c, hipMemcpyAsync(d_gain, gain, wh * sizeof(float), hipMemcpyHostToDevice, st));
        HIP_TRY(c, hipMemcpyAsync(d_var, variance, wh * sizeof(float), hipMemcpyHostToDevice, st));
        k_vargain<<<nblocks(wh), kT, 0, st>>>(wh, d_var, d_gain, d_vg);
    }
    HIP_TRY(c, hipMemcpyAsync(d_adu, adu, n * sizeof(float), hipMemcpyHostToDevice, st));

    // conversion and the stack mean
    k_convert<<<kRedBlocks, kT, 0, st>>>(n, wh, d_adu, slope, intercept, d_off, d_gain, d_img, d_part);
    k_moment_final<<<1, kT, 0, st>>>(d_part, 0, (int)n, d_st);

    // seed refinement (host centroid on the z-summed patches)
    std::vector<int> org(2 * std::max(nb, 1));
    if (nb > 0) {
        for (int p = 0; p < nb; p++) {
            org[2 * p] = (int)xp[p] - P / 2;
            org[2 * p + 1] = (int)yp[p] - P / 2;
        }
        HIP_TRY(c, hipMemcpyAsync(d_org, org.data(), 2 * nb * sizeof(int), hipMemcpyHostToDevice, st));
        k_refine_sum<<<nblocks((long)nb * P * P), kT, 0, st>>>(S, W, H, P, nb, d_org, d_img, d_sub);
        std::vector<double> sub((size_t)nb * P * P), vect((size_t)P * P);
        HIP_TRY(c, hipMemcpyAsync(sub.data(), d_sub, sub.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
        const size_t kth = (size_t)((P * P - 1) * .9);
        for (int p = 0; p < nb; p++) {
            const double* s = sub.data() + (size_t)p * P * P;
            std::copy(s, s + (size_t)P * P, vect.begin());
            std::nth_element(vect.begin(), vect.begin() + kth, vect.end());
            const double thr = vect[kth];
            double mx = 0, my = 0, sum = 0;
            int count = 0;
            for (int u = P / 4; u < 3 * P / 4; u++)
                for (int uu = P / 4; uu < 3 * P / 4; uu++)
                    if (s[u * P + uu] > thr) {
                        mx += (double)(org[2 * p] + u) * s[u * P + uu];
                        my += (double)(org[2 * p + 1] + uu) * s[u * P + uu];
                        sum += s[u * P + uu];
                        count++;
                    }
            if (count >= 1) {
                const float cx = (float)(mx / sum), cy = (float)(my / sum);
                const float dx = xp[p] - cx, dy = yp[p] - cy;
                const float dx2 = dx * dx, dy2 = dy * dy;
                if (std::sqrt((double)dx2) < (double)(P / 4)) xp[p] = cx;
                if (std::sqrt((double)dy2) < (double)(P / 4)) yp[p] = cy;
            }
            org[2 * p] = (int)xp[p] - P / 2;
            org[2 * p + 1] = (int)yp[p] - P / 2;
        }
        HIP_TRY(c, hipMemcpyAsync(d_org, org.data(), 2 * nb * sizeof(int), hipMemcpyHostToDevice, st));
    }

    // computeBackground: moments, is2, slice means
    for (int pass = 1; pass <= 3; pass++) {
        k_moment<<<kRedBlocks, kT, 0, st>>>(n, d_img, d_st, pass, d_part);
        k_moment_final<<<1, kT, 0, st>>>(d_part, pass, (int)n, d_st);
    }
    k_is2<<<dim3(chunks, S), kT, 0, st>>>(wh, d_img, d_st, d_is2, d_part);
    k_slice_final<<<S, kT, 0, st>>>(chunks, d_part, d_mstack);
    std::vector<double> mstack(S), stats(5);
    HIP_TRY(c, hipMemcpyAsync(mstack.data(), d_mstack, S * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipMemcpyAsync(stats.data(), d_st, 5 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    int mp = 0;
    {
        std::vector<double> filt(S, 0.0);
        for (int i = 0; i < S; i++) {
            int count = 0;
            for (int u = -sF; u <= sF; u++)
                if (i + u >= 0 && i + u < S) {
                    filt[i] += mstack[i + u];
                    count++;
                }
            if (count != 0) filt[i] /= count;
            if (filt[i] < filt[mp]) mp = i;
        }
        if (mp < sF) mp = sF;
        if (mp >= S - sF) mp = S - sF - 1;
    }
    k_pixel_mean<<<nblocks(wh), kT, 0, st>>>(S, W, H, mp, sF, d_img, d_is2, d_pm, d_cnt);
    k_normal<<<nblk, kT, 0, st>>>(W, H, d_pm, d_cnt, d_ne);
    k_normal_final<<<1, 64, 0, st>>>(nblk, d_ne, d_nes);
    double ne[kNE];
    HIP_TRY(c, hipMemcpyAsync(ne, d_nes, kNE * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    static const int ma[15] = {0, 1, 2, 3, 4, 0, 1, 2, 3, 0, 1, 2, 0, 1, 0};
    static const int mb[15] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4};
    static const int e0[6] = {0, 1, 2, 0, 1, 0}, e1[6] = {0, 0, 0, 1, 1, 2};
    double XX[6][6], YY[6], coef[6] = {};
    for (int i = 0; i < 6; i++) {
        for (int j = 0; j < 6; j++) {
            const int a = e0[i] + e0[j], b = e1[i] + e1[j];
            int k = 0;
            while (ma[k] != a || mb[k] != b) k++;
            XX[i][j] = ne[k];
        }
        YY[i] = ne[15 + i];
    }
    const int ndata = (int)ne[21];
    if (diag) {
        diag->mean0 = stats[0];
        diag->std0 = stats[1];
        diag->mean = stats[2];
        diag->std = stats[3];
        diag->size_filt = sF;
        diag->miniposition = mp;
        diag->ndata = ndata;
    }
    if (!lsolve6(XX, YY, coef))
        return set_err(c, ZOLA_EINVAL, "bead_prepare: background fit has singular normal equations (%d pixels)",
                       ndata);
    if (diag)
        for (int i = 0; i < 6; i++) diag->coef[i] = coef[i];
    Coef6 cf;
    for (int i = 0; i < 6; i++) cf.a[i] = coef[i];
    k_poly_eval<<<nblocks(wh), kT, 0, st>>>(W, H, cf, d_bg);
    if (background)
        HIP_TRY(c, hipMemcpyAsync(background, d_bg, wh * sizeof(double), hipMemcpyDeviceToHost, st));
    if (nb > 0) {
        k_bead_b<<<nblocks(nb), kT, 0, st>>>(W, H, P, nb, d_org, d_bg, d_B);
        k_bead_a0<<<nblocks((long)S * nb), kT, 0, st>>>(S, W, H, P, nb, d_org, d_img, d_B, d_A0);
        k_bead_a<<<nblocks((long)S * nb), kT, 0, st>>>(S, nb, sF, d_A0, d_A);
        k_patches<<<nblocks((long)nb * S * P * P), kT, 0, st>>>(S, W, H, P, nb, d_org, d_img, d_vg, d_B, d_patch,
                                                                 spatches ? d_sp : nullptr);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(B, d_B, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipMemcpyAsync(A, d_A, (size_t)nb * S * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipMemcpyAsync(patches, d_patch, (size_t)nb * S * P * P * sizeof(double), hipMemcpyDeviceToHost,
                                  st));
        if (spatches && d_sp)
            HIP_TRY(c, hipMemcpyAsync(spatches, d_sp, (size_t)nb * P * P * sizeof(double), hipMemcpyDeviceToHost,
                                      st));
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(st));
    return ZOLA_OK;
}

}  // namespace

extern "C" {

int zola_bead_prepare(zola_ctx* ctx, int S, int W, int H, const float* adu, double adu_factor, double gain,
                      double offset, double zstep, int P, int nb, float* xp, float* yp, double* patches, double* B,
                      double* A, double* background, zola_bead_diag* diag) {
    CtxLock lk_(ctx);
    if (!ctx) return ZOLA_EINVAL;
    if (!(gain > 0) || !(adu_factor > 0))
        return set_err(ctx, ZOLA_EINVAL, "bead_prepare: gain and adu must be positive");
    const double slope = adu_factor / gain;
    const double intercept = -offset * adu_factor / gain;
    return bead_prepare(ctx, S, W, H, adu, slope, intercept, nullptr, nullptr, nullptr, zstep, P, nb, xp, yp,
                        patches, B, A, nullptr, background, diag);
}

int zola_bead_prepare_scmos(zola_ctx* ctx, int S, int W, int H, const float* adu, const float* offset,
                            const float* gain, const float* variance, double zstep, int P, int nb, float* xp,
                            float* yp, double* patches, double* B, double* A, double* scmos_patches,
                            double* background, zola_bead_diag* diag) {
    CtxLock lk_(ctx);
    if (!ctx) return ZOLA_EINVAL;
    if (!offset || !gain || !variance) return set_err(ctx, ZOLA_EINVAL, "bead_prepare_scmos: NULL camera map");
    return bead_prepare(ctx, S, W, H, adu, 0.0, 0.0, offset, gain, variance, zstep, P, nb, xp, yp, patches, B, A,
                        scmos_patches, background, diag);
}

}  // extern "C"

// zola_detect.h -- launchers of the frame detector and frame-pipeline kernels.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace zola {

constexpr int kTwN = 4096;       // twiddle table length: W_4096^a = (cos, sin)(2 pi a / 4096)
constexpr int kFftElems = 2048;  // complex elements per FFT workgroup (longest line)

// Candidate as found on the device (before ordering).
struct DetCand {
    double score, dz, dx, dy;
    int r, i, j, pad;
};

struct DetGeom {
    int W, H, FX, FY, R, nZ;
    double thr, zstep, xystep;
};

}  // namespace zola

// Windowed mean and variance of each frame (Predetection.setImage, running sums
// of the float-rounded image in the reference's order).  frames F x W x H;
// sum1/sq1 scratch F x W x H; norm W x H; mean/var F x W x H.
hipError_t zola_launch_window_stats(const zola::DetGeom& g, int F, const float* frames, double* sum1,
                                    double* sq1, const double* norm, double* var, hipStream_t st);

// Unitary forward 2-D FFT of the zero-padded frames into spec (F x FX x FY double2).
hipError_t zola_launch_frame_spectrum(const zola::DetGeom& g, int F, const float* frames,
                                      double2* spec, const double2* tw, hipStream_t st);

// Unitary forward 2-D FFT of the nZ mean-subtracted templates (nZ x R x R, FP64)
// placed as in Predetection.makeDictionnary, into tspec (nZ x FX x FY double2).
hipError_t zola_launch_template_spectrum(const zola::DetGeom& g, const double* tmpl, double2* tspec,
                                         const double2* tw, hipStream_t st);

// Normalised correlation maps of templates [r0, min(r0 + 2 np, nZ)) for F
// frames, two templates per complex inverse transform:
// corr[f][r][ii][iii] (F x nZ x W x H).  work: F x np x W x FY double2.
hipError_t zola_launch_correlation(const zola::DetGeom& g, int F, int r0, int np, const double2* spec,
                                   const double2* tspec, const double* var, const double* var_psf,
                                   double2* work, double* corr, const double2* tw, hipStream_t st);

// Local maxima of the correlation maps (Predetection.getVectMaxPosition) into
// cand (F x cap, unordered) with per-frame counts (zeroed by the caller).
hipError_t zola_launch_local_max(const zola::DetGeom& g, int F, const double* corr, int cap,
                                 zola::DetCand* cand, int* count, hipStream_t st);

// Stable descending order by score (ties in the reference's scan order r, j, i)
// into out (F x cap x 7 doubles, the reference's record layout).
hipError_t zola_launch_rank_sort(int F, int cap, int max_count, const zola::DetCand* cand,
                                 const int* count, double* out, hipStream_t st);

// Gathers B R x R patches: roi[b] = (frame, ci, cj); patch pixel (a, bb) is
// frames[frame][ci + a - R/2][cj + bb - R/2].
hipError_t zola_launch_gather_rois(int W, int H, int R, int B, const int* roi, const float* frames,
                                   float* pixels, hipStream_t st);

// Model statistics of B fitted ROIs for the acceptance walk (see zola_detect.hip):
// max(mu - B), the chi2 score and the bit mask of (mu - B) > frac * max.
hipError_t zola_launch_model_stats(int B, int R, const float* psf, const float* px, const float* vpx,
                                   const double* th, double frac, double* mx, double* score, uint32_t* bits,
                                   hipStream_t st);

// R x R patches of one W x H map at the ROI centres (roi as for gather_rois).
hipError_t zola_launch_gather_map(int W, int H, int R, int B, const int* roi, const float* map, float* out,
                                  hipStream_t st);

// PSF parameters (B x 4 floats [x, y, z, zfocus]) of fitted theta (B x 5 doubles).
hipError_t zola_launch_theta_params(int B, const double* th, double zfocus, float* par, hipStream_t st);

// FP64 twiddle table W_4096^a = (cos, sin)(2 pi a / 4096), a = 0..4095.
void zola_twiddle_table(double2* host4096);

// zola_center.cpp -- the focus centre of the template range
// (process/cpu/SearchPSFcenter.java:37-81, 86-214, 219-351), in FP64.
//
// The reference picks the z (depth) at which the lateral CRLB of a standard
// emitter (3000 photons, background 50, PSF cropped to the full FFT size) is
// smallest: a coarse scan (step 5 px) over +-(2 range + |zfocus|), up to ten
// refinements (+-5 steps, step halved) until the position moves by less than
// 1e-5 um, then the start of the window of width range/2 with the smallest
// summed CRLB, scanned at one pixel.  Each CRLB is the reference's FD Fisher
// matrix (h = 5 nm; the photon derivative taken on the model with step
// 100 h, as computeFisher does), inverted by cofactors.  The seven FP64 PSFs
// per depth and the Fisher matrices of a whole scan are one batch on the GPU
// (zola_center_dev.hip); the cofactor inverses and the scan logic stay here.
#include <cmath>
#include <vector>

#include "../../include/zola_hip.h"
#include "zola_ctx.hpp"

namespace {

// Laplace-expansion determinant and cofactor inverse (utils/Matrixe.java:140-246).
double det_rec(int n, const double* a) {
    if (n == 1) return a[0];
    if (n == 2) return a[0] * a[3] - a[1] * a[2];
    double sum = 0, sub[16];
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int rr = 1; rr < n; rr++, r++) {
            int c = 0;
            for (int cc = 0; cc < n; cc++)
                if (cc != i) sub[r * (n - 1) + c++] = a[rr * n + cc];
        }
        sum += ((i & 1) ? -1.0 : 1.0) * a[i] * det_rec(n - 1, sub);
    }
    return sum;
}

void inverse5(const double* a, double* inv) {
    const double det = 1.0 / det_rec(5, a);
    double sub[16];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            int r = 0;
            for (int rr = 0; rr < 5; rr++) {
                if (rr == i) continue;
                int c = 0;
                for (int cc = 0; cc < 5; cc++)
                    if (cc != j) sub[r * 4 + c++] = a[rr * 5 + cc];
                r++;
            }
            inv[j * 5 + i] = (((i + j) & 1) ? -1.0 : 1.0) * det_rec(4, sub) * det;
        }
}

struct Search {
    zola_ctx* c;
    int R;
    double zfocus, photons = 3000, bg = 50;
    int rc = ZOLA_OK;

    // sqrt(CRLB_x^2 + CRLB_y^2) at every depth of a scan (computeFisher(0, 0, z, .005)).
    std::vector<double> scan(const std::vector<double>& zs) {
        std::vector<double> out(zs.size(), INFINITY), F(zs.size() * 25);
        if (rc) return out;
        rc = zola_center_fisher_dev(c, (int)zs.size(), zs.data(), zfocus, photons, bg, .005, F.data());
        if (rc) return out;
        for (size_t q = 0; q < zs.size(); q++) {
            double inv[25];
            inverse5(&F[q * 25], inv);
            const double sx = std::sqrt(inv[0]), sy = std::sqrt(inv[6]);
            out[q] = std::sqrt(sx * sx + sy * sy);
        }
        return out;
    }

    // getMinPositionCRLB (SearchPSFcenter.java:198-214): first strict minimum.
    double min_position(double minZ, double maxZ, double stepZ) {
        std::vector<double> zs;
        for (double u = minZ; u < maxZ; u += stepZ) zs.push_back(u);
        const std::vector<double> cr = scan(zs);
        double mv = INFINITY, mp = 0;
        for (size_t i = 0; i < zs.size(); i++)
            if (mv > cr[i]) {
                mv = cr[i];
                mp = zs[i];
            }
        return mp;
    }

    // getMinRangeCRLB (SearchPSFcenter.java:86-143) with its halved range.
    double min_range(double minZ, double maxZ, double stepZ, double axial) {
        axial /= 2;
        std::vector<double> zs;
        for (double u = minZ - axial / 2; u <= maxZ + axial / 2; u += stepZ) zs.push_back(u);
        const std::vector<double> crlb = scan(zs);
        double mv = INFINITY, mp = 0;
        int nb = (int)(axial / (2 * stepZ));
        for (double u = minZ; u < maxZ; u += stepZ) {
            double C = 0;
            int nb2 = -(int)(axial / (2 * stepZ));
            for (double i = -axial / 2; i <= axial / 2; i += stepZ) {
                const int k = nb + nb2;
                if (k >= 0 && k < (int)crlb.size()) C += crlb[k];
                nb2++;
            }
            if (mv > C) {
                mv = C;
                mp = u;
            }
            nb++;
        }
        return mp;
    }
};

}  // namespace

extern "C" int zola_search_psf_center(zola_ctx* c, double zfocus, double axial_range, double* center) {
    CtxLock lk_(c);
    if (!c || !center) return ZOLA_EINVAL;
    ZOLA_REFUSE_GRID256(c);
    if (!(axial_range > 0)) return set_err(c, ZOLA_EINVAL, "axial range must be positive");
    Search S{c, c->N, zfocus};
    double stepZ = c->o.xystep * 5;
    double minZ = zfocus - 2 * axial_range - std::fabs(zfocus);
    double maxZ = zfocus + std::fabs(zfocus) + 2 * axial_range;
    double position = S.min_position(minZ, maxZ, stepZ);
    double old = INFINITY;
    for (int i = 0; i < 10; i++) {
        minZ = position - stepZ * 5;
        maxZ = position + stepZ * 5;
        stepZ /= 2;
        position = S.min_position(minZ, maxZ, stepZ);
        if (std::fabs(old - position) < 0.00001) break;
        old = position;
    }
    *center = S.min_range(position - axial_range * 2, position + axial_range * 2, c->o.xystep, axial_range);
    return S.rc;
}

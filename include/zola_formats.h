/*
 * zola_formats.h -- the file formats on either side of the localization path:
 * the PSF calibration file (read before a run, written by a calibration) and
 * the localization table (written after a run).  Host code only, no GPU call;
 * the functions live in libzola_hip.so beside the GPU entry points.
 *
 * Output is byte-identical to what the plugin writes on a JVM whose
 * Double.toString returns the shortest decimal that round-trips (Java 19 and
 * later; earlier JVMs print an extra digit for a small set of values).
 *
 * Text functions follow snprintf: they write at most cap-1 characters plus a
 * NUL and return the length of the whole text (excluding the NUL), so a call
 * with cap = 0 returns the size needed.  Negative returns are ZOLA_E* codes;
 * zola_formats_last_error() gives the message of the calling thread's last
 * failure.
 */
#ifndef ZOLA_FORMATS_H
#define ZOLA_FORMATS_H

#include "zola_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The text "" + v gives for a Java double (Double.toString), as used by
 * utils/JSONformat.java:55-62,109-134 and data/PLocalization.java:224-238.
 * 32 bytes always suffice. */
int zola_java_double_string(double v, char* buf, int cap);

/* Calibration file of a Zernike PSF model, as DataPhase.saveJSON writes it
 * (process/cpu/DataPhase.java:711-760 with utils/JSONformat.java:46-134,
 * 283-297): keys zstep, sizeFFT, sizePix, NA, index, wavelength,
 * sigmaGaussianKernel, zernikedPSF (true), withApoFactor, azimuthal, radial,
 * zernike, on one line followed by a newline.  coeffs[j] is the coefficient of
 * Zernike slot j (radial order n, azimuthal m = -n, -n+2, .., n in slot order,
 * utils/Zernike.java:34-47).  optics->nwat is not part of the file. */
long zola_calib_json_format(const zola_optics* optics, const double* coeffs, int J, char* buf, long cap);

/* Reads a calibration file's text the way DataPhase.loadJSON does
 * (process/cpu/DataPhase.java:587-705 with JSONformat.get/getVect/readFile,
 * utils/JSONformat.java:193-280): each key is looked up by its first quoted
 * occurrence and its value taken up to the next comma (or the first '}'); the
 * numbers are read by Java's Double.parseDouble / Integer.parseInt rules and
 * the flags by Boolean.parseBoolean.  As in the reference, a Zernike file
 * takes its pupil grid from the caller: optics->size_fft = size_fft; the
 * file's own sizeFFT is returned in *file_size_fft (may be NULL).  nwat and
 * reserved are left untouched.  On success *J = number of coefficients (up
 * to cap written).  Files of a pixel-wise phase (zernikedPSF false) return
 * ZOLA_EUNSUPPORTED; a missing key or malformed number returns ZOLA_EINVAL,
 * where the reference throws. */
int zola_calib_json_parse(const char* text, int size_fft, zola_optics* optics, int* file_size_fft,
                          double* coeffs, int cap, int* J);

/* First line of the localization table: PLocalization.toStringName
 * (data/PLocalization.java:290-304) plus the newline of PrintWriter.println. */
long zola_loc_csv_header(char* buf, long cap);

/* Rows of the localization table for n records of ZOLA_LOC_FIELDS doubles (as
 * zola_localize_frames writes them): PLocalization(id, frame, x, y, z, A, B,
 * score, crlbX, crlbY, crlbZ) with each double floored to 1e-3
 * (data/PLocalization.java:60-73), then toString (224-238) and println, as
 * StackLocalization.saveCSV writes them (data/StackLocalization.java:1331-1368).
 * Row k gets id first_id + k and frame frame_base + record frame. */
long zola_loc_csv_rows(const double* locs, int n, int first_id, int frame_base, char* buf, long cap);

const char* zola_formats_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

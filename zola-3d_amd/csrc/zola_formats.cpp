// zola_formats.cpp -- calibration file and localization table, byte for byte
// as the plugin reads and writes them.  Host code only.
//
// Java's Double.toString (the text of "" + v) selects the shortest decimal that
// rounds back to v, the closest to v among those; when one digit would do, it
// chooses the closest among the one- and two-digit decimals (so the smallest
// subnormal prints as 4.9E-324).  std::to_chars gives the shortest decimal with
// the same closest-first rule; the two-digit case is settled with the correctly
// rounded %.1e of glibc.  Magnitudes in [1e-3, 1e7) print as plain decimals
// with at least one fractional digit, the others as d.dddE<exp>.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/zola_formats.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Digits and decimal exponent of the text Java selects for finite v != 0:
// |v| = 0.D[0]D[1].. * 10^(exp10 + 1), i.e. D[0].D[1..] * 10^exp10.
void java_digits(double v, std::string& digits, int& exp10) {
    char s[48];
    auto r = std::to_chars(s, s + sizeof s, std::fabs(v), std::chars_format::scientific);
    *r.ptr = 0;
    const char* e = std::strchr(s, 'e');
    digits.clear();
    for (const char* p = s; p < e; p++)
        if (*p != '.') digits.push_back(*p);
    exp10 = std::atoi(e + 1);
    if (digits.size() == 1) {
        char t[48];
        std::snprintf(t, sizeof t, "%.1e", std::fabs(v));
        if (std::strtod(t, nullptr) == std::fabs(v) && t[2] != '0') {
            const char* te = std::strchr(t, 'e');
            digits.assign({t[0], t[2]});
            exp10 = std::atoi(te + 1);
        }
    }
}

void append_java(std::string& out, double v) {
    if (std::isnan(v)) { out += "NaN"; return; }
    if (std::isinf(v)) { out += v > 0 ? "Infinity" : "-Infinity"; return; }
    if (v == 0) { out += std::signbit(v) ? "-0.0" : "0.0"; return; }
    std::string d;
    int k;
    java_digits(v, d, k);
    if (v < 0) out += '-';
    const double a = std::fabs(v);
    const int n = (int)d.size();
    if (a >= 1e-3 && a < 1e7) {
        if (k >= 0) {
            for (int i = 0; i <= k; i++) out += i < n ? d[i] : '0';
            out += '.';
            if (k + 1 < n) out.append(d, k + 1, std::string::npos);
            else out += '0';
        } else {
            out += "0.";
            out.append(-k - 1, '0');
            out += d;
        }
    } else {
        out += d[0];
        out += '.';
        if (n > 1) out.append(d, 1, std::string::npos);
        else out += '0';
        out += 'E';
        out += std::to_string(k);
    }
}

long emit(const std::string& s, char* buf, long cap) {
    if (buf && cap > 0) {
        const long m = (long)s.size() < cap - 1 ? (long)s.size() : cap - 1;
        std::memcpy(buf, s.data(), (size_t)m);
        buf[m] = 0;
    }
    return (long)s.size();
}

// ---- Java number parsing rules -------------------------------------------

// Double.parseDouble: String.trim() (drops chars <= ' ' at both ends), then an
// optional sign and NaN / Infinity / a decimal with optional exponent and an
// optional f/F/d/D suffix.  Hexadecimal literals are not accepted here.
bool java_parse_double(const std::string& in, double& out) {
    size_t b = 0, e = in.size();
    while (b < e && (unsigned char)in[b] <= ' ') b++;
    while (e > b && (unsigned char)in[e - 1] <= ' ') e--;
    std::string s = in.substr(b, e - b);
    if (s.empty()) return false;
    size_t i = 0;
    bool neg = false;
    if (s[i] == '+' || s[i] == '-') neg = s[i++] == '-';
    const std::string body = s.substr(i);
    if (body == "NaN") { out = NAN; return true; }
    if (body == "Infinity") { out = neg ? -INFINITY : INFINITY; return true; }
    if (!s.empty() && std::strchr("fFdD", s.back())) s.pop_back();
    size_t j = i, digits = 0;
    while (j < s.size() && std::isdigit((unsigned char)s[j])) j++, digits++;
    if (j < s.size() && s[j] == '.') {
        j++;
        while (j < s.size() && std::isdigit((unsigned char)s[j])) j++, digits++;
    }
    if (!digits) return false;
    if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
        j++;
        if (j < s.size() && (s[j] == '+' || s[j] == '-')) j++;
        size_t ed = 0;
        while (j < s.size() && std::isdigit((unsigned char)s[j])) j++, ed++;
        if (!ed) return false;
    }
    if (j != s.size()) return false;
    out = std::strtod(s.c_str(), nullptr);
    return true;
}

// Integer.parseInt: optional sign, then decimal digits only, within int range.
bool java_parse_int(const std::string& s, int& out) {
    size_t i = 0;
    bool neg = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i == s.size()) return false;
    long long v = 0;
    for (; i < s.size(); i++) {
        if (!std::isdigit((unsigned char)s[i])) return false;
        v = v * 10 + (s[i] - '0');
        if (v > 2147483648LL) return false;
    }
    if (neg) v = -v;
    if (v > 2147483647LL || v < -2147483648LL) return false;
    out = (int)v;
    return true;
}

// Boolean.parseBoolean: "true" in any case, nothing else (no trimming).
bool java_parse_bool(const std::string& s) {
    if (s.size() != 4) return false;
    const char* t = "true";
    for (int i = 0; i < 4; i++)
        if (std::tolower((unsigned char)s[i]) != t[i]) return false;
    return true;
}

// ---- JSONformat lookups ---------------------------------------------------

// BufferedReader.readLine over the file, each line followed by "\n".
std::string read_lines(const char* text) {
    std::string out;
    const char* p = text;
    while (*p) {
        const char* q = p;
        while (*q && *q != '\n' && *q != '\r') q++;
        out.append(p, q);
        out += '\n';
        if (*q == '\r' && q[1] == '\n') q += 2;
        else if (*q) q++;
        p = q;
    }
    return out;
}

// JSONformat.get: value after the first quoted occurrence of the key.
bool json_get(const std::string& f, const std::string& key, std::string& out) {
    const size_t id = f.find("\"" + key + "\"");
    if (id == std::string::npos) return false;
    const size_t colon = f.find(':', id);
    long start = colon == std::string::npos ? 0 : (long)colon + 1;
    long end;
    auto at = [&](long p, const char* lit) {
        const size_t L = std::strlen(lit);
        return p >= 0 && (size_t)p + L <= f.size() && f.compare((size_t)p, L, lit) == 0;
    };
    if ((size_t)start + 2 > f.size()) return false;
    if (at(start, "\"{")) {
        start += 1;
        const size_t q = f.find("}\"", (size_t)start + 2);
        end = q == std::string::npos ? 0 : (long)q + 1;  // 1 + indexOf, as in the reference
    } else if (at(start, "\"")) {
        start += 1;
        const size_t q = f.find('"', (size_t)start + 1);
        end = q == std::string::npos ? -1 : (long)q;
    } else {
        const size_t q = f.find(',', (size_t)start);
        end = q == std::string::npos ? -1 : (long)q;
    }
    if (end < 0) {
        const size_t q = f.find('}');
        end = q == std::string::npos ? -1 : (long)q;
    }
    if (end < start) return false;
    out = f.substr((size_t)start, (size_t)(end - start));
    return true;
}

// JSONformat.getVect for a numeric list: the text between the first '[' after
// the key and the first ']' after it, split at commas (String.split drops
// trailing empty strings).
bool json_get_vect(const std::string& f, const std::string& key, std::vector<std::string>& out) {
    const size_t id0 = f.find("\"" + key + "\"");
    const size_t from = id0 == std::string::npos ? 0 : id0;
    const size_t lb = f.find('[', from);
    const size_t rb = f.find(']', from);
    const long start = lb == std::string::npos ? 0 : (long)lb + 1;
    if (rb == std::string::npos || (long)rb < start) return false;
    const std::string v = f.substr((size_t)start, rb - (size_t)start);
    out.clear();
    size_t p = 0;
    while (true) {
        const size_t c = v.find(',', p);
        out.push_back(v.substr(p, c == std::string::npos ? std::string::npos : c - p));
        if (c == std::string::npos) break;
        p = c + 1;
    }
    while (out.size() > 1 && out.back().empty()) out.pop_back();
    return true;
}

void put_key(std::string& f, const char* key) {
    f += f.empty() ? "{\"" : ",\"";
    f += key;
    f += "\":";
}

}  // namespace

extern "C" {

int zola_java_double_string(double v, char* buf, int cap) {
    if (!buf || cap <= 0) return fail(ZOLA_EINVAL, "null or empty buffer");
    std::string s;
    append_java(s, v);
    if ((int)s.size() >= cap) return fail(ZOLA_EINVAL, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return (int)s.size();
}

long zola_calib_json_format(const zola_optics* o, const double* coeffs, int J, char* buf, long cap) {
    if (!o || J < 0 || (J > 0 && !coeffs)) return fail(ZOLA_EINVAL, "null optics or coefficients");
    std::string f;
    auto put_d = [&](const char* k, double v) { put_key(f, k); append_java(f, v); };
    put_d("zstep", o->zstep);
    put_key(f, "sizeFFT");
    f += std::to_string(o->size_fft);
    put_d("sizePix", o->xystep);
    put_d("NA", o->na);
    put_d("index", o->noil);
    put_d("wavelength", o->wavelength);
    put_d("sigmaGaussianKernel", o->sigma);
    put_key(f, "zernikedPSF");
    f += "true";
    put_key(f, "withApoFactor");
    f += o->with_apo ? "true" : "false";
    // slot j -> (n, m): n = 0, 1, ..; m = -n, -n+2, .., n
    std::vector<int> rn(J), am(J);
    for (int j = 0, n = 0; j < J; n++)
        for (int m = -n; m <= n && j < J; m += 2, j++) rn[j] = n, am[j] = m;
    put_key(f, "azimuthal");
    f += '[';
    for (int j = 0; j < J; j++) f += (j ? "," : "") + std::to_string(am[j]);
    f += ']';
    put_key(f, "radial");
    f += '[';
    for (int j = 0; j < J; j++) f += (j ? "," : "") + std::to_string(rn[j]);
    f += ']';
    put_key(f, "zernike");
    f += '[';
    for (int j = 0; j < J; j++) {
        if (j) f += ',';
        append_java(f, coeffs[j]);
    }
    f += "]}";
    f += '\n';
    return emit(f, buf, cap);
}

int zola_calib_json_parse(const char* text, int size_fft, zola_optics* o, int* file_size_fft, double* coeffs,
                          int cap, int* J) {
    if (!text || !o || !J || cap < 0 || (cap > 0 && !coeffs)) return fail(ZOLA_EINVAL, "null argument");
    const std::string f = read_lines(text);
    std::string v;
    auto need = [&](const char* key) {
        if (!json_get(f, key, v)) {
            g_err = std::string("calibration file has no value for \"") + key + "\"";
            return false;
        }
        return true;
    };
    auto num = [&](const char* key, double& out) {
        if (!need(key)) return false;
        if (!java_parse_double(v, out)) {
            g_err = std::string("calibration file: \"") + key + "\" is not a number: " + v;
            return false;
        }
        return true;
    };
    if (!need("zernikedPSF")) return ZOLA_EINVAL;
    const bool zern = java_parse_bool(v);
    double zstep, xy, na, noil, wl, sig;
    if (!num("zstep", zstep)) return ZOLA_EINVAL;
    if (!need("sizeFFT")) return ZOLA_EINVAL;
    int fsize;
    if (!java_parse_int(v, fsize)) return fail(ZOLA_EINVAL, "calibration file: \"sizeFFT\" is not an int: " + v);
    if (!num("sizePix", xy) || !num("NA", na) || !num("index", noil) || !num("wavelength", wl) ||
        !num("sigmaGaussianKernel", sig))
        return ZOLA_EINVAL;
    const bool apo = json_get(f, "withApoFactor", v) && java_parse_bool(v);
    if (!zern) return fail(ZOLA_EUNSUPPORTED, "calibration file holds a pixel-wise phase (zernikedPSF false)");
    std::vector<std::string> items;
    if (!json_get_vect(f, "zernike", items)) return fail(ZOLA_EINVAL, "calibration file has no \"zernike\" list");
    std::vector<double> a(items.size());
    for (size_t i = 0; i < items.size(); i++)
        if (!java_parse_double(items[i], a[i]))
            return fail(ZOLA_EINVAL, "calibration file: zernike[" + std::to_string(i) + "] is not a number: " + items[i]);
    o->size_fft = size_fft;
    o->xystep = xy;
    o->zstep = zstep;
    o->wavelength = wl;
    o->noil = noil;
    o->na = na;
    o->sigma = sig;
    o->with_apo = apo ? 1 : 0;
    if (file_size_fft) *file_size_fft = fsize;
    for (int i = 0; i < cap && i < (int)a.size(); i++) coeffs[i] = a[i];
    *J = (int)a.size();
    return ZOLA_OK;
}

long zola_loc_csv_header(char* buf, long cap) {
    static const char* h =
        "id,frame,x [nm],y [nm],z [nm],intensity,background,chi2,crlbX,crlbY,crlbZ,driftX,driftY,driftZ,"
        "occurrenceMerging\n";
    return emit(h, buf, cap);
}

long zola_loc_csv_rows(const double* locs, int n, int first_id, int frame_base, char* buf, long cap) {
    if (n < 0 || (n > 0 && !locs)) return fail(ZOLA_EINVAL, "null records");
    std::string s;
    s.reserve((size_t)n * 110);
    char tmp[32];
    auto put_int = [&](long long v) {
        auto r = std::to_chars(tmp, tmp + sizeof tmp, v);
        s.append(tmp, r.ptr);
    };
    for (int k = 0; k < n; k++) {
        const double* r = locs + (size_t)k * ZOLA_LOC_FIELDS;
        put_int((long long)first_id + k);
        s += ',';
        put_int((long long)frame_base + (long long)r[0]);
        for (int c = 1; c <= 9; c++) {
            s += ',';
            const double f = std::floor(1000.0 * r[c]);
            if (std::fabs(f) >= 1 && std::fabs(f) < 1e10) {
                // f/1000 has at most 10 significant digits, so its shortest
                // round-trip decimal is f/1000 itself: integer part, '.', the
                // three thousandths without trailing zeros (at least one digit).
                const long long q = (long long)std::fabs(f);
                if (f < 0) s += '-';
                put_int(q / 1000);
                s += '.';
                const int m = (int)(q % 1000);
                const char d[3] = {char('0' + m / 100), char('0' + m / 10 % 10), char('0' + m % 10)};
                const int keep = d[2] != '0' ? 3 : d[1] != '0' ? 2 : 1;
                s.append(d, keep);
            } else {
                append_java(s, f / 1000.0);
            }
        }
        s += ",0.0,0.0,0.0,1\n";
    }
    return emit(s, buf, cap);
}

const char* zola_formats_last_error(void) { return g_err.c_str(); }

}  // extern "C"

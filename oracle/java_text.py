"""Restatement of the plugin's text formats in plain Python, from the Java
specification and the reference's own writers and readers.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
calibration-file and localization-table functions.  The product never imports
it.

- java_double_string: Java's Double.toString, derived from its specification
  (java.lang.Double.toString, Java 19 and later) with exact rational
  arithmetic: among the decimals inside the rounding interval of v take those
  of minimal length (lengths 1 and 2 when one digit suffices), then the closest
  to v, then the even last digit; render magnitudes in [1e-3, 1e7) as plain
  decimals with at least one fractional digit and the others as d.ddE<exp>.
  Independent of the library's method (std::to_chars).
- JSONformat.put/get/getVect/readFile: utils/JSONformat.java:46-134, 193-280.
- DataPhase.saveJSON / loadJSON: process/cpu/DataPhase.java:587-760.
- PLocalization floor to 1e-3, toString, toStringName:
  data/PLocalization.java:60-73, 224-238, 290-304; StackLocalization.saveCSV:
  data/StackLocalization.java:1331-1368.
"""
import math
import re
import struct
from fractions import Fraction


# ---- Double.toString ------------------------------------------------------

def _interval(x):
    """Rounding interval of the positive finite double x: (lo, hi, closed)."""
    lo = (Fraction(math.nextafter(x, 0.0)) + Fraction(x)) / 2
    up = math.nextafter(x, math.inf)
    hi = (Fraction(x) + (Fraction(up) if math.isfinite(up) else Fraction(2) ** 1024)) / 2
    # round half to even: the midpoints round to x when its significand is even
    closed = (struct.unpack("<Q", struct.pack("<d", x))[0] & 1) == 0
    return lo, hi, closed


def _decimal_exponent(fx):
    """floor(log10(fx)) for a positive Fraction, exactly."""
    e = math.floor(math.log10(float(fx))) if fx > Fraction(1, 10**300) else -330
    while Fraction(10) ** e > fx:
        e -= 1
    while Fraction(10) ** (e + 1) <= fx:
        e += 1
    return e


def _candidates(fx, n, lo, hi, closed):
    """n-significant-digit decimals (D, q) with D*10^q inside the interval."""
    e = _decimal_exponent(fx)
    q = e - n + 1
    scale = Fraction(10) ** q
    d_lo = math.floor(fx / scale)
    out = []
    for D in (d_lo, d_lo + 1):
        val = D * scale
        inside = (lo < val < hi) or (closed and (val == lo or val == hi))
        if D > 0 and inside:
            out.append((D, q))
    return out


def _strip(D, q):
    while D % 10 == 0:
        D //= 10
        q += 1
    return D, q


def java_double_string(v):
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    if v == 0:
        return "-0.0" if math.copysign(1.0, v) < 0 else "0.0"
    x = abs(v)
    fx = Fraction(x)
    lo, hi, closed = _interval(x)
    for n in range(1, 19):
        cands = _candidates(fx, n, lo, hi, closed)
        if cands:
            break
    if n == 1:
        cands = cands + _candidates(fx, 2, lo, hi, closed)
    best = None
    for D, q in cands:
        D, q = _strip(D, q)
        dist = abs(D * Fraction(10) ** q - fx)
        key = (dist, D % 2)
        if best is None or key < best[0]:
            best = (key, D, q)
    _, D, q = best
    digits = str(D)
    k = q + len(digits) - 1          # value = d0.d1d2... * 10^k
    sign = "-" if v < 0 else ""
    if Fraction(1, 1000) <= fx < 10**7:
        if k >= 0:
            ip = digits[:k + 1].ljust(k + 1, "0")
            fp = digits[k + 1:] or "0"
        else:
            ip, fp = "0", "0" * (-k - 1) + digits
        return f"{sign}{ip}.{fp}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}E{k}"


# ---- JSONformat -----------------------------------------------------------

class JsonFormat:
    """utils/JSONformat.java: a JSON object kept as one string."""

    def __init__(self, text="{}"):
        self.file = text

    def _append(self, key, value_text):
        if self.file.endswith("{}"):
            self.file = '{"' + key + '":' + value_text + "}"
        else:
            self.file = self.file[:-1] + ',"' + key + '":' + value_text + "}"

    def put(self, key, value):
        if isinstance(value, bool):
            self._append(key, "true" if value else "false")
        elif isinstance(value, int):
            self._append(key, str(value))
        elif isinstance(value, float):
            self._append(key, java_double_string(value))
        elif isinstance(value, (list, tuple)):
            parts = []
            for v in value:
                if isinstance(v, bool):
                    parts.append("true" if v else "false")
                elif isinstance(v, int):
                    parts.append(str(v))
                else:
                    parts.append(java_double_string(float(v)))
            self._append(key, "[" + ",".join(parts) + "]")
        else:
            raise TypeError(type(value))

    def saved(self):
        """What PrintWriter.println writes."""
        return self.file + "\n"

    @staticmethod
    def read_file(text):
        """BufferedReader.readLine per line, each followed by a newline."""
        out = []
        i = 0
        while i < len(text):
            j = i
            while j < len(text) and text[j] not in "\r\n":
                j += 1
            out.append(text[i:j] + "\n")
            if j < len(text) and text[j] == "\r" and j + 1 < len(text) and text[j + 1] == "\n":
                j += 2
            elif j < len(text):
                j += 1
            i = j
        return "".join(out)

    def get(self, key):
        f = self.file
        idx = f.find('"' + key + '"')
        if idx < 0:
            return None
        start = 1 + f.find(":", idx)
        if start + 2 > len(f):
            raise IndexError("subSequence")
        if f[start:start + 2] == '"{':
            start += 1
            end = 1 + f.find('}"', start + 2)
        elif f[start:start + 1] == '"':
            start += 1
            end = f.find('"', start + 1)
        else:
            end = f.find(",", start)
        if end < 0:
            end = f.find("}")
        if end < start:
            raise IndexError("substring")
        return f[start:end]

    def get_vect(self, key):
        f = self.file
        idx = f.find('"' + key + '"')
        start = 1 + f.find("[", idx)
        end = f.find("]", idx)
        if end < start:
            raise IndexError("substring")
        items = f[start:end].split(",")
        while len(items) > 1 and items[-1] == "":
            items.pop()
        return items


def java_parse_double(s):
    """Double.parseDouble for decimal text (trim, sign, NaN/Infinity,
    exponent, f/F/d/D suffix)."""
    t = s.strip(" " + "".join(chr(c) for c in range(33)))
    body = t[1:] if t[:1] in "+-" else t
    if body == "NaN":
        return math.nan
    if body == "Infinity":
        return -math.inf if t.startswith("-") else math.inf
    if t and t[-1] in "fFdD":
        t = t[:-1]
    if not re.fullmatch(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?", t):
        raise ValueError("NumberFormatException: " + repr(s))
    return float(t)


def java_parse_int(s):
    if not re.fullmatch(r"[+-]?\d+", s):
        raise ValueError("NumberFormatException: " + repr(s))
    v = int(s)
    if not -2**31 <= v < 2**31:
        raise ValueError("NumberFormatException: " + repr(s))
    return v


def java_parse_bool(s):
    return s is not None and s.lower() == "true"


def zernike_slots(J):
    """(n, m) of Zernike slots 0..J-1: n = 0, 1, ..; m = -n, -n+2, .., n
    (utils/Zernike.java:34-47)."""
    out = []
    n = 0
    while len(out) < J:
        for m in range(-n, n + 1, 2):
            if len(out) < J:
                out.append((n, m))
        n += 1
    return out


def calib_json_save(opt, coeffs):
    """DataPhase.saveJSON for a Zernike model (process/cpu/DataPhase.java:711-760).
    opt: mapping with size_fft, xystep, zstep, wavelength, noil, na, sigma, with_apo."""
    js = JsonFormat()
    js.put("zstep", float(opt["zstep"]))
    js.put("sizeFFT", int(opt["size_fft"]))
    js.put("sizePix", float(opt["xystep"]))
    js.put("NA", float(opt["na"]))
    js.put("index", float(opt["noil"]))
    js.put("wavelength", float(opt["wavelength"]))
    js.put("sigmaGaussianKernel", float(opt["sigma"]))
    js.put("zernikedPSF", True)
    js.put("withApoFactor", bool(opt["with_apo"]))
    slots = zernike_slots(len(coeffs))
    js.put("azimuthal", [m for _, m in slots])
    js.put("radial", [n for n, _ in slots])
    js.put("zernike", [float(a) for a in coeffs])
    return js.saved()


def calib_json_load(text, size_fft):
    """DataPhase.loadJSON (process/cpu/DataPhase.java:587-705).  Returns
    (optics dict, coefficients, file sizeFFT); raises where Java throws;
    returns None for a pixel-phase file."""
    js = JsonFormat(JsonFormat.read_file(text))

    def need(k):
        v = js.get(k)
        if v is None:
            raise ValueError("NullPointerException: " + k)
        return v

    zern = java_parse_bool(need("zernikedPSF"))
    opt = {"zstep": java_parse_double(need("zstep"))}
    fsize = java_parse_int(need("sizeFFT"))
    opt["xystep"] = java_parse_double(need("sizePix"))
    opt["na"] = java_parse_double(need("NA"))
    opt["noil"] = java_parse_double(need("index"))
    opt["wavelength"] = java_parse_double(need("wavelength"))
    opt["sigma"] = java_parse_double(need("sigmaGaussianKernel"))
    opt["with_apo"] = java_parse_bool(js.get("withApoFactor"))
    if not zern:
        return None
    opt["size_fft"] = size_fft
    coeffs = [java_parse_double(s) for s in js.get_vect("zernike")]
    return opt, coeffs, fsize


# ---- localization table ---------------------------------------------------

CSV_HEADER = ("id,frame,x [nm],y [nm],z [nm],intensity,background,chi2,crlbX,crlbY,crlbZ,"
              "driftX,driftY,driftZ,occurrenceMerging")


def loc_csv_row(loc_id, frame, x, y, z, A, B, score, crlbx, crlby, crlbz):
    """PLocalization(id, frame, ...) then toString (data/PLocalization.java:60-73,
    224-238); drift 0, occurrence 1."""
    vals = [math.floor(1000 * v) / 1000 if math.isfinite(1000 * v) else 1000 * v / 1000
            for v in (x, y, z, A, B, score, crlbx, crlby, crlbz)]
    parts = [str(loc_id), str(frame)] + [java_double_string(v) for v in vals] + ["0.0", "0.0", "0.0", "1"]
    return ",".join(parts)

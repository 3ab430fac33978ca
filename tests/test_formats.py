"""Calibration file and localization table: the library's writers and readers
against the restatement in oracle/java_text.py, and that restatement pinned by
values the Java platform documents and by properties of the shortest
round-trip decimal.  Host code only (no GPU)."""
import math
import os
import random
import struct
import sys

import numpy as np
import pytest

import zola_hip as Z

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import java_text as J  # noqa: E402

# Double.toString outputs fixed by the Java platform's documentation (the
# constants Double.MIN_VALUE, MIN_NORMAL, MAX_VALUE, Float.MAX_VALUE and
# Long.MAX_VALUE as double are printed in their javadoc) and by the
# plain/scientific thresholds 1e-3 and 1e7 of the Double.toString javadoc.
JAVA_DOC = [
    (5e-324, "4.9E-324"),
    (2.2250738585072014e-308, "2.2250738585072014E-308"),
    (1.7976931348623157e308, "1.7976931348623157E308"),
    (3.4028234663852886e38, "3.4028234663852886E38"),
    (9.223372036854776e18, "9.223372036854776E18"),
    (1e7, "1.0E7"),
    (9999999.0, "9999999.0"),
    (1e-3, "0.001"),
    (1e-4, "1.0E-4"),
    (100.0, "100.0"),
    (1.0, "1.0"),
    (-0.0, "-0.0"),
    (0.0, "0.0"),
    (math.nan, "NaN"),
    (math.inf, "Infinity"),
    (-math.inf, "-Infinity"),
    (0.1 + 0.2, "0.30000000000000004"),
    (1 / 3, "0.3333333333333333"),
    (math.pi, "3.141592653589793"),
    (12345678.0, "1.2345678E7"),
    (-123.456, "-123.456"),
]


def random_doubles(n, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        kind = rng.randrange(4)
        if kind == 0:   # any finite bit pattern
            while True:
                v = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
                if math.isfinite(v):
                    break
        elif kind == 1:  # values floored to 1e-3, as the table holds
            v = math.floor(1000 * rng.uniform(-5e4, 5e4)) / 1000
        elif kind == 2:  # near the 1e-3 / 1e7 thresholds
            v = rng.choice([1e-3, 1e7]) * (1 + rng.uniform(-1e-15, 1e-15)) * rng.choice([1, -1])
        else:           # short decimals and subnormals
            v = rng.choice([float(f"{rng.randint(1, 99)}e{rng.randint(-320, 300)}"),
                            rng.randint(1, 1000) * 5e-324])
        out.append(v)
    return out


@pytest.mark.parametrize("v,text", JAVA_DOC)
def test_oracle_double_string_matches_java_documented_values(v, text):
    assert J.java_double_string(v) == text


def test_oracle_double_string_is_shortest_round_trip():
    # repr is Python's shortest round-trip decimal: Java's digits must have the
    # same count, except where one digit suffices (Java may then use two).
    for v in random_doubles(3000, 7):
        t = J.java_double_string(v)
        assert float(t) == v
        mant = t.split("E")[0].lstrip("-").replace(".", "").lstrip("0").rstrip("0")
        r = repr(abs(v)).split("e")[0].replace(".", "").lstrip("0").rstrip("0")
        assert len(mant) == len(r) or (len(r) == 1 and len(mant) == 2), (v, t, repr(v))
        a = abs(v)
        assert ("E" in t) == (a < 1e-3 or a >= 1e7), (v, t)


def test_library_double_string_matches_oracle():
    vals = [v for v, _ in JAVA_DOC] + random_doubles(20000, 11)
    for v in vals:
        assert Z.java_double_string(v) == J.java_double_string(v), repr(v)


def _opt_dict(o):
    return {k: getattr(o, k) for k in ("size_fft", "xystep", "zstep", "wavelength", "noil", "na",
                                       "sigma", "with_apo")}


def test_calib_json_write_matches_oracle_and_reads_back():
    rng = np.random.default_rng(5)
    for J_ in (1, 15, 28, 45):
        o = Z.optics(size_fft=int(rng.choice([64, 128])), xystep=float(rng.uniform(0.05, 0.2)),
                     zstep=float(rng.choice([0.01, 0.05, 0.1])), wavelength=float(rng.uniform(0.5, 0.7)),
                     noil=1.518, na=float(rng.uniform(1.2, 1.49)), sigma=float(rng.uniform(0.5, 1.5)))
        o.with_apo = int(rng.integers(0, 2))
        a = rng.normal(0, 0.3, J_)
        text = Z.calib_json_format(o, a)
        assert text == J.calib_json_save(_opt_dict(o), a)
        back, b, fsize = Z.calib_json_parse(text, 128)
        assert fsize == o.size_fft and back.size_fft == 128
        for k in ("xystep", "zstep", "wavelength", "noil", "na", "sigma", "with_apo"):
            assert getattr(back, k) == getattr(o, k)
        assert np.array_equal(b, a)
    # an empty coefficient list is written, but the reader throws on it, as in the reference
    empty = Z.calib_json_format(Z.optics(), [])
    assert empty == J.calib_json_save(_opt_dict(Z.optics()), [])
    with pytest.raises(Z.ZolaError):
        Z.calib_json_parse(empty, 64)


def test_calib_json_slots_and_layout():
    o = Z.optics(size_fft=64, xystep=0.1, zstep=0.05, wavelength=0.6, noil=1.518, na=1.4, sigma=1.0)
    text = Z.calib_json_format(o, [0.0, 0.1, -0.2, 0.25, 1.0, 1e-5])
    assert text == ('{"zstep":0.05,"sizeFFT":64,"sizePix":0.1,"NA":1.4,"index":1.518,"wavelength":0.6,'
                    '"sigmaGaussianKernel":1.0,"zernikedPSF":true,"withApoFactor":false,'
                    '"azimuthal":[0,-1,1,-2,0,2],"radial":[0,1,1,2,2,2],'
                    '"zernike":[0.0,0.1,-0.2,0.25,1.0,1.0E-5]}\n')


QUIRKS = [
    # CRLF lines, spaces, a d-suffixed number, an exponent, keys in another order
    ('{\r\n "NA": 1.4,\r\n "zernikedPSF":true,"sizePix": 1.06E-1 ,"index":1.518d,"zstep":0.05,'
     '"sizeFFT":64,"wavelength":6e-1,"sigmaGaussianKernel":1,"zernike":[ 0.5, -1.0E-2 ,3]}'),
    # withApoFactor absent -> false; flag text before '}' without a comma
    '{"zstep":0.1,"sizeFFT":128,"sizePix":0.1,"NA":1.3,"index":1.33,"wavelength":0.5,'
    '"sigmaGaussianKernel":0.8,"zernike":[1,2],"zernikedPSF":TRUE}',
    # the flag followed by a newline is not "true" for Boolean.parseBoolean
    '{"zstep":0.1,"sizeFFT":64,"sizePix":0.1,"NA":1.3,"index":1.33,"wavelength":0.5,'
    '"sigmaGaussianKernel":0.8,"zernike":[1],"withApoFactor":true\n,"zernikedPSF":true}',
    # trailing comma inside the list (String.split drops the empty tail)
    '{"zstep":0.1,"sizeFFT":64,"sizePix":0.1,"NA":1.3,"index":1.33,"wavelength":0.5,'
    '"sigmaGaussianKernel":0.8,"zernikedPSF":true,"zernike":[1,2,]}',
]


@pytest.mark.parametrize("text", QUIRKS)
def test_calib_json_read_follows_reference_lookup(text):
    want = J.calib_json_load(text, 64)
    o, a, fsize = Z.calib_json_parse(text, 64)
    wopt, wa, wfsize = want
    assert fsize == wfsize and o.size_fft == 64
    for k in ("xystep", "zstep", "wavelength", "noil", "na", "sigma"):
        assert getattr(o, k) == wopt[k]
    assert bool(o.with_apo) == wopt["with_apo"]
    assert list(a) == wa


BAD = [
    ('{"zstep":0.1,"sizeFFT":64,"sizePix":0.1,"NA":1.3,"index":1.33,"wavelength":0.5,'
     '"zernikedPSF":true,"zernike":[1]}', -1),                         # sigma key missing
    ('{"zstep":0.1,"sizeFFT": 64,"sizePix":0.1,"NA":1.3,"index":1.33,"wavelength":0.5,'
     '"sigmaGaussianKernel":0.8,"zernikedPSF":true,"zernike":[1]}', -1),  # parseInt(" 64") throws
    ('{"zstep":0.1,"sizeFFT":64,"sizePix":0.1,"NA":1.3,"index":1.33,"wavelength":0.5,'
     '"sigmaGaussianKernel":0.8,"zernikedPSF":true,"zernike":[]}', -1),   # parseDouble("") throws
    ('{"zstep":0.1,"sizeFFT":64,"sizePix":0.1,"NA":1.3,"index":1.33,"wavelength":0.5,'
     '"sigmaGaussianKernel":0.8,"zernikedPSF":false,"pixelPhaseValue":[1]}', -2),
]


@pytest.mark.parametrize("text,code", BAD)
def test_calib_json_read_rejects_what_the_reference_cannot_load(text, code):
    if code == -1:
        with pytest.raises((ValueError, IndexError)):
            J.calib_json_load(text, 64)
    else:
        assert J.calib_json_load(text, 64) is None
    with pytest.raises(Z.ZolaError, match=f"^{code}:"):
        Z.calib_json_parse(text, 64)


def test_localization_table_matches_oracle():
    rng = np.random.default_rng(9)
    n = 2000
    locs = np.zeros((n, Z.LOC_FIELDS))
    locs[:, 0] = rng.integers(0, 500, n)
    locs[:, 1:3] = rng.uniform(-100, 30000, (n, 2))
    locs[:, 3] = rng.uniform(-3000, 3000, n)
    locs[:, 4] = rng.uniform(100, 2e7, n)
    locs[:, 5] = rng.uniform(0, 50, n)
    locs[:, 6] = rng.uniform(0, 5, n)
    locs[:, 7:10] = rng.uniform(0, 100, (n, 3))
    locs[:5, 7] = np.nan
    locs[5, 4] = 1e7
    locs[6, 5] = 0.0
    locs[7, 6] = 1e-4
    text = Z.loc_csv(locs, first_id=3, frame_base=1)
    lines = text.split("\n")
    assert lines[0] == J.CSV_HEADER and lines[-1] == ""
    for k in range(n):
        r = locs[k]
        want = J.loc_csv_row(3 + k, 1 + int(r[0]), *r[1:10])
        assert lines[1 + k] == want, k

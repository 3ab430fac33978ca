"""Checks that the oracle pins catch plausible mistakes.

TEST INFRASTRUCTURE ONLY (CPU).  Plants one mistake at a time into a copy of
oracle/zola_oracle.c under a temporary directory, builds that copy with gcc and
runs the pins of tests/test_oracle_pins.py that should catch it.  Prints PASS
(the pins failed, as they must) or MISSED for each mistake; exits 1 if any was
missed.  Usage: python oracle/plant_mistakes.py [name ...]
"""
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MUT = [
 ("gauss pi", "double sigF = (1.0 / sigma) * s2 / (ZO_PI * 2.0);", "double sigF = (1.0 / sigma) * s2 / (ZO_PI);", "gaussian or independent_numpy"),
 ("gauss N", "double sigF = (1.0 / sigma) * s2 / (ZO_PI * 2.0);", "double sigF = (1.0 / sigma) * m->N / (ZO_PI * 2.0);", "gaussian or independent_numpy"),
 ("score cnt", "return sum / cnt;", "return sum / (cnt - 1);", "score"),
 ("score sq", "sum += (patch[k] - model[k]) * (patch[k] - model[k]) / model[k];", "sum += (patch[k] - model[k]) * (patch[k] - model[k]) / patch[k];", "score"),
 ("calib clamp", "if (mod <= 0) mod = 0.00000001;", "if (mod <= 0) mod = 0.0001;", "calib_nll"),
 ("calib dz sign", "zo_model_psf(m, xs[s], ys[s], dz[s] + zs[s], zwat, psf);", "zo_model_psf(m, xs[s], ys[s], dz[s] - zs[s], zwat, psf);", "calib_nll"),
 ("A step x100", "double h = isB ? s->h * 10 : s->h * 1000;", "double h = isB ? s->h * 10 : s->h * 100;", "sweeps"),
 ("B step x100", "double h = isB ? s->h * 10 : s->h * 1000;", "double h = isB ? s->h * 100 : s->h * 1000;", "sweeps"),
 ("clamp dx", "s.maxShift = m->o.xystep / 2.;", "s.maxShift = m->o.xystep;", "sweeps"),
 ("gamma 3 steps", "for (double gamma = 1; gamma > .02; gamma /= 10) {\n        *t -= gamma * grad;\n        double lik = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB);\n        if (lik < lik2) { prevLik = lik; break; }", "for (double gamma = 1; gamma > .002; gamma /= 10) {\n        *t -= gamma * grad;\n        double lik = L(s, s->tX, s->tY, s->tZ, s->tA, s->tB);\n        if (lik < lik2) { prevLik = lik; break; }", "sweeps"),
 ("amp break", "if (lik < lik2) { lik2 = lik; save = *t; }", "if (lik < lik2) { lik2 = lik; save = *t; break; }", "sweeps"),
 ("x/y order", "likelihood = minimize_pos(&s, 1, likelihood);\n        likelihood = minimize_pos(&s, 2, likelihood);", "likelihood = minimize_pos(&s, 2, likelihood);\n        likelihood = minimize_pos(&s, 1, likelihood);", "sweeps"),
 ("h 1nm", "s.h = .0005;", "s.h = .001;", "sweeps"),
 ("no z clamp", "if (s.tZ < s.minZ) s.tZ = s.minZ;", "if (0) s.tZ = s.minZ;", "sweeps"),
 ("photon floor", "s->tA = fmax(1, s->tA);", "s->tA = fmax(0.5, s->tA);", "sweeps"),
]
sel = sys.argv[1:]
missed = 0
tmp = tempfile.mkdtemp(prefix="zola_plant_")
for name, old, new, k in MUT:
    if sel and name not in sel:
        continue
    d = os.path.join(tmp, name.replace(" ", "_").replace("/", ""))
    shutil.rmtree(d, ignore_errors=True)
    shutil.copytree(os.path.join(ROOT, "oracle"), os.path.join(d, "oracle"),
                    ignore=shutil.ignore_patterns("*.so", "__pycache__"))
    shutil.copytree(os.path.join(ROOT, "tests"), os.path.join(d, "tests"), ignore=shutil.ignore_patterns("__pycache__"))
    os.makedirs(os.path.join(d, "zola-3d_amd"))
    shutil.copy(os.path.join(ROOT, "zola-3d_amd", "zola_hip.py"), os.path.join(d, "zola-3d_amd"))
    src = os.path.join(d, "oracle", "zola_oracle.c")
    txt = open(src).read()
    assert txt.count(old) >= 1, name
    open(src, "w").write(txt.replace(old, new, 1))
    subprocess.run(["make", "-s", "-C", os.path.join(d, "oracle")], check=True)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "tests/test_oracle_pins.py", "-k", k, "-p", "no:cacheprovider"],
                       cwd=d, capture_output=True, text=True)
    last = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else r.stderr[-300:]
    print(("PASS  " if r.returncode != 0 else "MISSED") + "  %-14s %s" % (name, last), flush=True)
    missed += r.returncode == 0
    shutil.rmtree(d, ignore_errors=True)
shutil.rmtree(tmp, ignore_errors=True)
sys.exit(1 if missed else 0)

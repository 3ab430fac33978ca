"""Builds libzola_hip.so (gfx950) in-tree from zola-3d_amd/csrc.

hipcc compiles every source as HIP for --offload-arch=gfx950.  Packed FP32
instructions are disabled (-packed-fp32-ops) on purpose.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libzola_hip.so")
# the 128-grid instantiations compile longest: they start first
SOURCES = (sorted(os.path.basename(p) for p in glob.glob(os.path.join(CSRC, "zola_inst_n128*.hip"))) +
           ["zola_dispatch.hip", "zola_host.cpp", "zola_calib.cpp", "zola_detect.hip", "zola_frames.cpp",
            "zola_center.cpp", "zola_formats.cpp", "zola_beads.hip", "zola_noalloc.cpp", "zola_center_dev.hip",
            "zola_grid256.hip"] +
           sorted(os.path.basename(p) for p in glob.glob(os.path.join(CSRC, "zola_inst_*.hip"))
                  if not os.path.basename(p).startswith("zola_inst_n128")))
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _deps_mtime(src):
    """Newest mtime of the source and of every header it may include."""
    hdrs = (glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(CSRC, "*.hpp")) +
            glob.glob(os.path.join(HERE, "..", "include", "*.h")))
    return max(os.path.getmtime(p) for p in [os.path.join(CSRC, src)] + hdrs)


def _compile(src, verbose, incremental=False):
    obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) >= _deps_mtime(src):
        return obj
    cmd = ["hipcc", "-x", "hip"] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False, jobs=None, incremental=False):
    """Compiles every source (incremental=True skips objects newer than their inputs)."""
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, incremental), SOURCES))
    tmp = OUT + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))

"""Builds libgpusnarks.so (gfx950) in-tree with hipcc, and the CPU oracle.

    python3 -m gpusnarks_amd.build          # or gpusnarks_amd.build.build()

Translation units compile in parallel; the 753-bit kernel files dominate.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libgpusnarks.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the 26x26 digit product must unroll completely to stay in registers
    "-mllvm", "-pragma-unroll-threshold=1000000",
    "-I" + INCLUDE, "-I" + CSRC,
]
HIP_SOURCES = ["kernels_mnt6.hip", "kernels_mnt4.hip", "kernels_fp32.hip", "gsn_runtime.hip",
               "gsn_batch.hip", "gsn_domain.hip", "gsn_server.cpp", "gsn_staging.cpp", "gsn_msm.hip"]
CXX_SOURCES = ["best_fft.cpp", "multi_exp.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build_oracle():
    """oracle/liboracle.so (+ oracle/_ref when the reference tree exists)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", path, "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-I" + INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    shutil.move(tmp, LIB)
    build_oracle()
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

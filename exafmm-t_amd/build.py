"""Build the gfx950 engine (libexafmm_t_amd.so) and the pybind11 module.

    python exafmm-t_amd/build.py [--jobs N]

Outputs (in-tree, git-ignored, shipped with the source tree):
    exafmm-t_amd/lib/libexafmm_t_amd.so          C ABI of include/exafmm_t_amd.h
    exafmm-t_amd/exafmm/_exafmm*.so               Python module `exafmm`
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
HIP_SOURCES = ["p2p.hip", "gemm.hip", "m2l.hip", "engine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# gfx950 only; packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def engine_lib():
    return os.path.join(LIB, "libexafmm_t_amd.so")


def python_ext():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "exafmm", "_exafmm" + suffix)


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in HIP_SOURCES]
    cmds = [[HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(HIP_SOURCES, objs)]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(run, cmds):
            if verbose and out:
                print(out)
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", engine_lib()] + objs)
    import pybind11
    inc = [sysconfig.get_paths()["include"], pybind11.get_include(), os.path.join(PKG, "include"),
           os.path.join(ROOT, "include")]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-fvisibility=hidden",
           os.path.join(CSRC, "python.cpp"), "-o", python_ext()]
    cmd += ["-I" + d for d in inc]
    cmd += ["-L" + LIB, "-lexafmm_t_amd", "-Wl,-rpath,$ORIGIN/../lib"]
    out = run(cmd)
    if verbose and out:
        print(out)
    return engine_lib(), python_ext()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    for f in build(a.jobs, a.verbose):
        print("built", f)

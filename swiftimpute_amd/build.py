"""Build recipes for the native libraries (gfx950 HIP + host C++).

    python -m swiftimpute_amd.build          # build everything in-tree

Products (git-ignored, carried to the GPU box with the tree):
  swiftimpute_amd/lib/libsi_hip.so     kernels + the C ABI (include/swiftimpute_hip.h)
  swiftimpute_amd/lib/libswiftimpute.so  the C++ API (include/swiftimpute/*.hpp) over the C ABI
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

# Packed-FP32 instructions are disabled for every HIP object: the kernels use
# ordinary scalar FP32 arithmetic only.  No floating-point contraction: the
# forward-backward kernels fuse multiply-adds only where written (fmaf), so
# their arithmetic is the same for every lane shape.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{INCLUDE}",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]

HIP_SOURCES = ["hip/capi.hip", "hip/index.hip", "hip/pbwt.hip", "hip/hmm.hip"]
HIP_HEADERS = ["hip/si_device.hpp", "hip/si_internal.hpp", "hip/fb_kernels.hpp"]
# Forward-backward lane shapes (padded states Lp, lanes per sample G), one
# object each from hip/fb_inst.hip: 8, 4 and 2 states per lane, at most 64 lanes.
FB_SHAPES = [(4, 1), (4, 2), (8, 1), (8, 2), (8, 4), (16, 2), (16, 4), (16, 8),
             (32, 4), (32, 8), (32, 16), (64, 8), (64, 16), (64, 32), (128, 16), (128, 32),
             (256, 32), (256, 64)]
API_SOURCES = ["api/panel.cpp", "api/imputer.cpp", "api/synth.cpp", "api/server.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build(force: bool = True, verbose: bool = False) -> dict:
    """Compile both libraries.  force=True rebuilds every object."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HIP_HEADERS] + [os.path.join(INCLUDE, "swiftimpute_hip.h")]

    jobs = []
    hip_objs = []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, os.path.basename(src) + ".o")
        hip_objs.append(o)
        if force or _newer(o, [s] + hdrs):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    inst = os.path.join(CSRC, "hip", "fb_inst.hip")
    for lp, g in FB_SHAPES:
        o = os.path.join(BUILD, f"fb_lp{lp}_g{g}.o")
        hip_objs.append(o)
        if force or _newer(o, [inst] + hdrs):
            jobs.append([HIPCC, *HIP_FLAGS, f"-DSI_FB_LP={lp}", f"-DSI_FB_G={g}", "-c", inst,
                         "-o", o])
    jobs.sort(key=lambda c: 0 if any("fb_inst" in x for x in c) else 1)
    api_objs = []
    api_hdrs = hdrs + [os.path.join(INCLUDE, "swiftimpute", f)
                       for f in ("types.hpp", "imputer.hpp", "synth.h", "server.h", "impute_c.h")]
    for src in API_SOURCES:
        s = os.path.join(CSRC, src)
        if not os.path.exists(s):
            continue
        o = os.path.join(BUILD, os.path.basename(src) + ".o")
        api_objs.append(o)
        if force or _newer(o, [s] + api_hdrs):
            jobs.append(["g++", *HOST_FLAGS, "-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)

    hip_lib = os.path.join(LIBDIR, "libsi_hip.so")
    tmp = hip_lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *hip_objs, "-o", tmp])
    os.replace(tmp, hip_lib)
    result = {"si_hip": hip_lib}
    if api_objs:
        api_lib = os.path.join(LIBDIR, "libswiftimpute.so")
        tmp = api_lib + ".tmp"
        _run(["g++", "-shared", "-fPIC", *api_objs, "-o", tmp, f"-L{LIBDIR}", "-lsi_hip",
              f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib",
              "-pthread"])
        os.replace(tmp, api_lib)
        result["swiftimpute"] = api_lib
        # C++ API test driver (tests/cpp/test_api.cpp), linked against the library
        drv = os.path.join(ROOT, "tests", "cpp", "test_api.cpp")
        if os.path.exists(drv):
            exe = os.path.join(BUILD, "test_api")
            _run(["g++", *HOST_FLAGS, drv, "-o", exe + ".tmp", f"-L{LIBDIR}", "-lswiftimpute",
                  "-lsi_hip", f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{LIBDIR}",
                  f"-Wl,-rpath,{ROCM}/lib", "-Wl,-rpath,$ORIGIN/../lib", "-pthread"])
            os.replace(exe + ".tmp", exe)
            result["test_api"] = exe
    return result


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    print(build(force="--incremental" not in sys.argv, verbose=True))

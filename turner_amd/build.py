"""Builds libturner_amd.so (HIP kernels for gfx950 + host C++) in-tree.

Device code: hipcc --offload-arch=gfx950, FP32 with -ffp-contract=off (the
traversal must reproduce the reference's FP32 results bit for bit) and with
packed FP32 instructions disabled.  Host code: g++ -O2 -ffp-contract=off.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libturner_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["capi.hip", "render.hip", "guide.hip"]
CXX_SOURCES = ["kdtree_build.cpp", "host_api.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
CXX_FLAGS = ["-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-pthread", "-Wall",
             "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + out)
    return out


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, [HIPCC, *HIP_FLAGS, "-c", path, "-o", obj]))
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, ["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(lambda j: _run(j[1]), jobs))
    if verbose:
        for o in outs:
            if o:
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp,
          *[j[0] for j in jobs]])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", oracle])
    return os.path.join(oracle, "liboracle.so")


def build_cli():
    """Host C++ front ends over the C ABI, one per trace() plugin as the
    reference links them: pathtracer, raycaster, raytracer (same flags)."""
    host = os.path.join(HERE, "host")
    srcs = [os.path.join(host, f) for f in ("pathtracer_main.cpp", "turner.cpp")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-O2", "-std=c++17", "-pthread", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", host, "-I", os.path.join(ROCM, "include")]
    lib_obj = os.path.join(BUILD, "turner_host.o")
    _run(["g++", *inc, "-c", srcs[1], "-o", lib_obj])
    link = [f"-L{HERE}", "-lturner_amd", f"-L{os.path.join(ROCM, 'lib')}", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
            f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
    jobs = []
    for name, renderer in (("pathtracer", 0), ("raycaster", 1), ("raytracer", 2)):
        out = os.path.join(HERE, name)
        jobs.append(["g++", *inc, f"-DTN_CLI_RENDERER={renderer}", srcs[0], lib_obj, "-o", out, *link])
    # the C++ caller test of the host interface (tests/test_gpu_host_api.py)
    check_src = os.path.join(ROOT, "tests", "cpp", "host_api_check.cpp")
    if os.path.exists(check_src):
        jobs.append(["g++", *inc, check_src, lib_obj, "-o", os.path.join(ROOT, "tests", "cpp", "host_api_check"),
                     f"-L{HERE}", "-lturner_amd", f"-L{os.path.join(ROCM, 'lib')}", "-lamdhip64",
                     "-Wl,-rpath,$ORIGIN/../../turner_amd", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"])
    with ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_run, jobs))
    return os.path.join(HERE, "pathtracer")


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_oracle())
    print(build_cli())

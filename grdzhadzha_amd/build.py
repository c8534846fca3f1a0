"""Build recipe of the HIP library (gfx950) -- no CMake, one hipcc line.

Flags: -ffp-contract=off keeps every multiply and add separately rounded
(the reference's CPU arithmetic); packed FP32 instructions are disabled;
FP32 divide and square root (mixed-precision faces only) use the
hardware's approximations.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GRHD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # FP32 divide and square root by the hardware's approximations (a few
    # ulp): only the mixed-precision face kernels do FP32 arithmetic; every
    # FP64 kernel compiles to the same instructions with or without it
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIBS = {
    "libgrhd_hip.so": ["grhd_kernels.hip"],
}


ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# C++ host programs over the C-ABI (include/grhd_operator.hpp): host code
# only, built with g++ against the HIP runtime and libgrhd_hip.so.
HOST_PROGRAMS = {
    os.path.join("tests", "cpp", "_build", "level_driver"):
        [os.path.join("tests", "cpp", "level_driver.cpp")],
    os.path.join("tests", "cpp", "_build", "hooks_driver"):
        [os.path.join("tests", "cpp", "hooks_driver.cpp")],
}


def build_host_programs(verbose=False):
    out = []
    for exe, srcs in HOST_PROGRAMS.items():
        path = os.path.join(ROOT, exe)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        rel = os.path.relpath(HERE, os.path.dirname(path))
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
               "-I" + os.path.join(ROCM, "include"),
               "-I" + os.path.join(ROOT, "include"), "-o", path + ".tmp"] + \
            [os.path.join(ROOT, s) for s in srcs] + \
            ["-L" + HERE, "-lgrhd_hip", "-L" + os.path.join(ROCM, "lib"),
             "-lamdhip64", "-Wl,-rpath,$ORIGIN/" + rel,
             "-Wl,-rpath," + os.path.join(ROCM, "lib")]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("g++ failed for %s:\n%s" % (exe, res.stderr))
        os.replace(path + ".tmp", path)
        out.append(path)
    return out


def build(verbose=False):
    for lib, srcs in LIBS.items():
        out = os.path.join(HERE, lib)
        cmd = [HIPCC] + HIP_FLAGS + ["-o", out + ".tmp"] + \
            [os.path.join(CSRC, s) for s in srcs]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("hipcc failed for %s:\n%s" % (lib, res.stderr))
        os.replace(out + ".tmp", out)
    return [os.path.join(HERE, lib) for lib in LIBS] + \
        build_host_programs(verbose)


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)

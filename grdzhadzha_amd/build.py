"""Build recipe of the HIP library (gfx950) -- no CMake, one hipcc line.

Flags: -ffp-contract=off keeps every multiply and add separately rounded
(the reference's CPU arithmetic); packed FP32 instructions are disabled.
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
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIBS = {
    "libgrhd_hip.so": ["grhd_kernels.hip"],
}


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
    return [os.path.join(HERE, lib) for lib in LIBS]


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)

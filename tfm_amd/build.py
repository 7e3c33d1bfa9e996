"""Build recipe for libccpmi.so (gfx950) and the ccp-perform / ccp-analyze /
ccp-profile drop-in executables.

Everything is compiled in-tree with hipcc so the built files travel with the
source tree.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libccpmi.so")
BIN = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CCPMI_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    srcs = [os.path.join(CSRC, "kernels.hip"), os.path.join(CSRC, "ccpmi.cpp"), os.path.join(CSRC, "batch.cpp"),
            os.path.join(CSRC, "nn.hip"), os.path.join(CSRC, "noalloc.hip")]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)


def build_cli():
    os.makedirs(BIN, exist_ok=True)
    for name, src in (("ccp-perform", "ccp_perform.cpp"), ("ccp-analyze", "ccp_analyze.cpp"),
                      ("ccp-profile", "ccp_profile.cpp")):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", os.path.join(BIN, name), path,
              "-L", PKG, "-lccpmi", "-Wl,-rpath,$ORIGIN/.."])


def build():
    build_lib()
    build_cli()


if __name__ == "__main__":
    build()
    sys.exit(0)

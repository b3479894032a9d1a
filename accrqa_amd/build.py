"""Build recipe for libaccrqa_amd.so (gfx950) and the test-only checkers.

    python -m accrqa_amd.build          # product library + oracle (+ oracle/_ref where the reference exists)

The product library is compiled with hipcc for gfx950 only, with floating-point
contraction off (the exact decision path relies on one rounding per operation)
and packed-FP32 instruction selection disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libaccrqa_amd.so")
SOURCES = [os.path.join(HERE, "csrc", "rqa_engine.hip"), os.path.join(HERE, "csrc", "accrqa_api.cpp")]
REF_TREE = "/root/reference"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in ordinary VGPRs: the element stage reads every result once,
    # and accumulator registers would cost a copy per element
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    return r.stdout


def build_library(out=LIB):
    tmp = out + ".tmp"
    _run([hipcc()] + HIPCC_FLAGS + ["-I" + os.path.join(ROOT, "include")] + SOURCES + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir(os.path.join(REF_TREE, "src")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REF_TREE}"])


def build():
    build_oracle()
    build_library()


if __name__ == "__main__":
    build()
    print("built", LIB)

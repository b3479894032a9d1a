"""Build recipe for libaccrqa_amd.so (gfx950) and the test-only checkers.

    python -m accrqa_amd.build          # product library + oracle (+ oracle/_ref where the reference exists)

oracle/_ref also receives the reference library with the AMD platform branch
(oracle/ref_amd.py), linked against the product library, for the integration
test; it is test infrastructure like the rest of oracle/.

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
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("rqa_engine.hip", "rqa_batch.hip", "rqa_static.hip", "accrqa_api.cpp")]
OBJDIR = os.path.join(HERE, "_build")
REF_TREE = "/root/reference"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
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


def build_library(out=LIB, incremental=False):
    """Each translation unit to an object (in parallel), then one link.
    incremental=True keeps an object that is newer than its source and every
    header of csrc/ and include/ (a development shortcut; build() always
    compiles everything)."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJDIR, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include")]
    headers = [os.path.join(d, f) for d in (os.path.join(HERE, "csrc"), os.path.join(ROOT, "include"))
               for f in os.listdir(d) if f.endswith((".h", ".hpp"))]
    newest_header = max(os.path.getmtime(h) for h in headers)

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
        if (incremental and os.path.exists(obj)
                and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header)):
            return obj
        _run([hipcc()] + HIPCC_FLAGS + inc + ["-c", src, "-o", obj])
        return obj
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = out + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir(os.path.join(REF_TREE, "src")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REF_TREE}"])


def build_reference_integration():
    """The reference library with the AMD platform branch (oracle/ref_amd.py),
    linked against the library just built; only where the reference exists."""
    if os.path.isdir(os.path.join(REF_TREE, "src")):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ref_amd
        ref_amd.build(REF_TREE)


def build():
    build_oracle()
    build_library()
    build_reference_integration()


if __name__ == "__main__":
    build()
    print("built", LIB)

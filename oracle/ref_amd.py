"""Builds the reference library with the AMD platform branch a maintainer adds.

TEST INFRASTRUCTURE ONLY.  Recipe, run where the reference tree exists
(oracle/Makefile target `ref_amd`):

* copies the reference's src/ and include/ into a temporary directory (never
  into the repository or oracle/_ref, so no reference source travels);
* applies, in that copy, the edits INTEGRATION.md describes:
  - `PLT_AMD_GPU = 2048` in enum Accrqa_CompPlatform
    (include/AccRQA_utilities_comp_platform.hpp);
  - a first branch in each of the six public accrqa_RR / accrqa_DET /
    accrqa_LAM overloads (src/AccRQA_functions.cpp:525-555, 780-810,
    945-975) that calls the C ABI of include/accrqa_amd.h for PLT_AMD_GPU;
  - platform 2048 accepted by check_comp_platform
    (src/AccRQA_python_bindings.cpp:39-53), the C entry points the
    reference's Python package binds;
* compiles the reference's AccRQA_functions.cpp, AccRQA_python_bindings.cpp
  and AccRQA_utilities_mem.cpp from that copy with g++ and the reference's
  release settings (C++11, -O3, OpenMP; version macros on the command line as
  its build passes them), linked against accrqa_amd/libaccrqa_amd.so, into
  oracle/_ref/libaccrqa_ref_amd.so (git-ignored).

The GPU test tests/test_gpu_integration.py then calls the reference's own
entry points from that library with PLT_CPU and with PLT_AMD_GPU and compares
them: the reference's dispatch, argument checks and Python-binding marshalling
in front of this engine.
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_ref", "libaccrqa_ref_amd.so")

# Per public overload: the C ABI function and how the reference's parameter
# names map onto its arguments (the l_min / v_min list has two spellings).
ABI = {
    ("RR", "float"): "accrqa_amd_RR_f32",
    ("RR", "double"): "accrqa_amd_RR_f64",
    ("DET", "float"): "accrqa_amd_DET_f32",
    ("DET", "double"): "accrqa_amd_DET_f64",
    ("LAM", "float"): "accrqa_amd_LAM_f32",
    ("LAM", "double"): "accrqa_amd_LAM_f64",
}


def _branch(kind, params):
    """The AMD branch for one overload, from its parameter names in order."""
    names = [re.split(r"[\s*]+", p.strip())[-1] for p in params]
    if kind == "RR":
        # output, input, size, tau, nTaus, emb, nEmbs, thr, nThr, dist, platform, error
        out, x, n, tau, ntau, emb, nemb, thr, nthr, dist, plat, err = names
        args = f"{out}, {x}, {n}, {tau}, {ntau}, {emb}, {nemb}, {thr}, {nthr}, (int){dist}, &e_amd"
    else:
        # output, input, size, tau, nTaus, emb, nEmbs, l, nL, thr, nThr, dist, calc_ENTR, platform, error
        out, x, n, tau, ntau, emb, nemb, lv, nl, thr, nthr, dist, ent, plat, err = names
        args = (f"{out}, {x}, {n}, {tau}, {ntau}, {emb}, {nemb}, {lv}, {nl}, {thr}, {nthr}, (int){dist}, "
                f"{ent}, &e_amd")
    return plat, err, args


def patch_functions(src):
    sig = re.compile(r"void accrqa_(RR|DET|LAM)\((float|double) \*output,([^)]*)\)\s*\{")
    count = 0

    def repl(m):
        nonlocal count
        kind, ty = m.group(1), m.group(2)
        params = [f"{ty} *output"] + m.group(3).split(",")
        plat, err, args = _branch(kind, params)
        count += 1
        return (m.group(0) + f"\n\tif({plat}==PLT_AMD_GPU){{ int e_amd = 0; {ABI[(kind, ty)]}({args}); "
                f"*{err} = (Accrqa_Error)e_amd; return; }}")
    out = sig.sub(repl, src)
    if count != 6:
        raise RuntimeError(f"expected the six public overloads, patched {count}")
    return '#include "accrqa_amd.h"\n' + out


def patch_platform_enum(src):
    out, n = re.subn(r"(PLT_NV_GPU\s*=\s*1024\s*,?)", r"PLT_NV_GPU = 1024,\n\tPLT_AMD_GPU = 2048,", src)
    if n != 1:
        raise RuntimeError("Accrqa_CompPlatform enum not found")
    return out


def patch_python_platform(src):
    out, n = re.subn(r"(else if\(int_comp_platform==1024\)\{\s*comp_platform = PLT_NV_GPU;\s*\})",
                     r"\1\n\telse if(int_comp_platform==2048){\n\t\tcomp_platform = PLT_AMD_GPU;\n\t}", src)
    if n != 1:
        raise RuntimeError("check_comp_platform not found")
    return out


def build(ref="/root/reference", cxx="g++"):
    if not os.path.isdir(os.path.join(ref, "src")):
        print("reference tree absent; libaccrqa_ref_amd.so not built", file=sys.stderr)
        return None
    lib = os.path.join(ROOT, "accrqa_amd", "libaccrqa_amd.so")
    if not os.path.exists(lib):
        raise RuntimeError("build accrqa_amd/libaccrqa_amd.so first")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with tempfile.TemporaryDirectory(prefix="accrqa_ref_amd_") as tmp:
        shutil.copytree(os.path.join(ref, "src"), os.path.join(tmp, "src"))
        shutil.copytree(os.path.join(ref, "include"), os.path.join(tmp, "include"))
        edits = [("include/AccRQA_utilities_comp_platform.hpp", patch_platform_enum),
                 ("src/AccRQA_functions.cpp", patch_functions),
                 ("src/AccRQA_python_bindings.cpp", patch_python_platform)]
        for rel, fn in edits:
            p = os.path.join(tmp, rel)
            with open(p) as f:
                s = f.read()
            with open(p, "w") as f:
                f.write(fn(s))
        srcs = [os.path.join(tmp, "src", n) for n in
                ("AccRQA_functions.cpp", "AccRQA_python_bindings.cpp", "AccRQA_utilities_mem.cpp")]
        cmd = [cxx, "-O3", "-std=c++11", "-fPIC", "-fopenmp", "-w", "-shared",
               "-DACCRQA_VER_MAJOR=1", "-DACCRQA_VER_MINOR=0", "-DACCRQA_VER_PATCH=1",
               "-I" + os.path.join(tmp, "include"), "-I" + os.path.join(tmp, "src"),
               "-I" + os.path.join(ROOT, "include")] + srcs + [
               "-L" + os.path.dirname(lib), "-laccrqa_amd", "-Wl,-Bsymbolic",
               "-Wl,-rpath,$ORIGIN/../../accrqa_amd", "-o", OUT + ".tmp"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout)
            raise RuntimeError("building libaccrqa_ref_amd.so failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    print(build(ref))

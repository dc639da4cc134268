"""Build lib/libfdmc.so for gfx950 with hipcc (one object per source, in parallel).

Usage: python build.py [--jobs N] [--verbose]
Every HIP translation unit is compiled with packed-FP32 instruction selection
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(HERE, "lib", "libfdmc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["engine.cpp", "k_sobol.hip", "k_path_raw.hip", "k_path_bb.hip", "k_path_prng.hip",
           "k_levels_raw.hip", "k_levels_bb.hip", "k_levels_prng.hip", "k_levels_rows.hip",
           "k_greeks_raw.hip", "k_greeks_bb.hip", "k_greeks_prng.hip",
           "k_pca.hip", "k_basket.hip", "k_is.hip", "k_path_rows.hip", "k_book.hip", "k_book_o3.hip", "book.cpp",
           "../cpp/utils.cpp", "../cpp/methods_hip.cpp"]
EXAMPLES = ["asian_table", "book_table", "heston_table"]
# The reference's own example programs (examples/ejemplo01..11*.cpp) include
# "../methods_cuda.cuh"; compiled unchanged from their place in the reference
# checkout, fed on stdin from a folder whose parent holds a one-line
# methods_cuda.cuh that includes the facade, so the include resolves there.
REF_EXAMPLES = "/root/reference/examples"
REF_EXAMPLES_OUT = os.path.join(HERE, "lib", "reference_examples")

FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]


def compile_one(src, verbose):
    s = os.path.normpath(os.path.join(CSRC, src))
    o = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") or src == "engine.cpp" else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", s, "-o", o]
    t = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout + p.stderr)
    if p.returncode != 0:
        raise RuntimeError("compile failed: " + src)
    return src, time.time() - t, o


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1))
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, dt, o in ex.map(lambda s: compile_one(s, verbose), SOURCES):
            print("  compiled %-18s %6.1fs" % (src, dt), flush=True)
            objs.append(o)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    print("  linked", LIB, flush=True)
    for ex in EXAMPLES:
        src = os.path.join(HERE, "cpp", "examples", ex + ".cpp")
        exe = os.path.join(HERE, "lib", ex)
        cmd = [HIPCC, "-O2", "-std=c++17", src, "-o", exe, "-L" + os.path.dirname(LIB), "-lfdmc",
               "-Wl,-rpath,$ORIGIN"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            sys.stderr.write(p.stdout + p.stderr)
            raise RuntimeError("example build failed: " + ex)
        print("  built", exe, flush=True)
    ref = build_reference_examples()
    if ref:
        print("  built %d reference example programs in %s" % (len(ref), REF_EXAMPLES_OUT), flush=True)
    return LIB


def build_reference_examples(out_dir=REF_EXAMPLES_OUT, src_dir=REF_EXAMPLES, cxx="g++"):
    """Compile every example program of the reference, unchanged, against the
    C++ facade and libfdmc.so.  Returns the built executables ([] when the
    reference checkout is absent)."""
    if not os.path.isdir(src_dir):
        return []
    cwd = os.path.join(out_dir, "src")
    os.makedirs(cwd, exist_ok=True)
    with open(os.path.join(out_dir, "methods_cuda.cuh"), "w") as f:
        f.write('#include "methods_hip.hpp"\n')
    built = []
    libdir = os.path.dirname(LIB)
    for name in sorted(os.listdir(src_dir)):
        if not name.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, name[:-4])
        rpath = os.path.relpath(libdir, out_dir)
        cmd = [cxx, "-std=c++17", "-O2", "-I" + os.path.join(HERE, "cpp"), "-x", "c++", "-", "-o", exe,
               "-L" + libdir, "-lfdmc", "-Wl,-rpath,$ORIGIN/" + rpath]
        with open(os.path.join(src_dir, name)) as src:
            p = subprocess.run(cmd, stdin=src, cwd=cwd, capture_output=True, text=True)
        if p.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + p.stdout + p.stderr)
            raise RuntimeError("reference example failed to compile against the facade: " + name)
        built.append(exe)
    return built


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    build(a.jobs, a.verbose)

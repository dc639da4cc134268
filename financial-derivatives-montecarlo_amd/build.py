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
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    build(a.jobs, a.verbose)

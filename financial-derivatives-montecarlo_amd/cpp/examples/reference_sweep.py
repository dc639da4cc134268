"""Run the reference's example programs, compiled unchanged against the facade
(build.py: build_reference_examples, into lib/reference_examples/), over the
precision sweep of the reference's report (informes_finales/informe_completo.txt:
the eps values of each example's section), and write their printed tables.

Usage: python reference_sweep.py OUT_DIR [--examples ejemplo01 ...] [--max-eps-index K]
           [--timeout SECONDS]

Each run writes OUT_DIR/<example>_eps<eps>.txt (the program's stdout) and one
line per run to OUT_DIR/sweep.jsonl: example, eps, exit code, wall time, the
parsed table rows (method, price, std error, N, T(s), |error|, OK) and the end of
stderr.  A program that ends with an uncaught exception is recorded and the sweep
goes on: the examples build bridge/PCA lists for levels 0..L_MAX-1 but allow
max_L = L_MAX, so an MLQMC run that needs level L_MAX throws "missing
bb_list[L_MAX]" as the reference's run_mlqmc_cuda does (methods_cuda.cu:1773-1775).
"""
import argparse
import json
import os
import re
import subprocess
import time

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(os.path.dirname(HERE)), "lib", "reference_examples")

# eps values of each example's section of informe_completo.txt (lines 8-631)
SWEEP = {
    "ejemplo01_european": [0.05, 0.02, 0.01, 0.005, 0.002, 0.001, 0.0005],
    "ejemplo02_asian": [0.05, 0.02, 0.01, 0.005, 0.002, 0.001],
    "ejemplo03_lookback": [0.05, 0.02, 0.01, 0.005, 0.002],
    "ejemplo04_barrier": [0.05, 0.02, 0.01, 0.005, 0.002],
    "ejemplo05_heston": [0.1, 0.05, 0.02, 0.01, 0.005],
    "ejemplo06_dupire": [0.05, 0.02, 0.01, 0.005, 0.002, 0.001],
    "ejemplo07_basket_uncorr": [0.01, 0.005, 0.002, 0.001, 0.0005, 0.0002, 0.0001],
    "ejemplo08_basket_corr": [0.05, 0.02, 0.01, 0.005, 0.002, 0.001, 0.0005],
    "ejemplo09_asian_cv": [0.05, 0.02, 0.01],
    "ejemplo10_dupire_cv": [0.05, 0.02, 0.01, 0.005, 0.002],
    "ejemplo11_otm_is": [0.002, 0.001, 0.0005, 0.0002, 0.0001],
}

ROW = re.compile(r"^\s{2}(\S.{0,21}?)\s+(-?[\d.]+|nan|inf)\s+(-?[\d.]+|nan|inf)\s+(\d+)\s+([\d.]+)\s+"
                 r"(-?[\d.]+|nan|inf)\s+(SI|NO)\s*$")
REF = re.compile(r"^\s+Referencia:\s+(-?[\d.]+)")


def parse_table(text):
    """Rows of print_table's output (utils.cpp: print_table) and its reference price."""
    ref, rows = None, []
    for line in text.splitlines():
        m = REF.match(line)
        if m:
            ref = float(m.group(1))
            continue
        m = ROW.match(line)
        if m:
            rows.append({"method": m.group(1).strip(), "price": float(m.group(2)), "std_error": float(m.group(3)),
                         "n": int(m.group(4)), "time_s": float(m.group(5)), "abs_error": float(m.group(6)),
                         "ok": m.group(7) == "SI"})
    return ref, rows


def run(example, eps, timeout):
    exe = os.path.join(BIN, example)
    t = time.perf_counter()
    p = subprocess.run([exe, repr(eps)], capture_output=True, text=True, timeout=timeout)
    wall = time.perf_counter() - t
    ref, rows = parse_table(p.stdout)
    return p, {"example": example, "eps": eps, "rc": p.returncode, "wall_s": wall, "reference_price": ref,
               "rows": rows, "stderr_tail": p.stderr[-300:]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out_dir")
    ap.add_argument("--examples", nargs="*", default=None)
    ap.add_argument("--max-eps-index", type=int, default=None, help="run only the first K eps of each example")
    ap.add_argument("--timeout", type=float, default=120.0)
    a = ap.parse_args()
    os.makedirs(a.out_dir, exist_ok=True)
    names = a.examples or sorted(SWEEP)
    with open(os.path.join(a.out_dir, "sweep.jsonl"), "w") as js:
        for ex in names:
            for eps in SWEEP[ex][:a.max_eps_index]:
                print("run %s eps=%g" % (ex, eps), flush=True)
                p, rec = run(ex, eps, a.timeout)
                with open(os.path.join(a.out_dir, "%s_eps%g.txt" % (ex, eps)), "w") as f:
                    f.write(p.stdout + p.stderr)
                js.write(json.dumps(rec) + "\n")
                js.flush()
                print("  rc=%d wall=%.2fs rows=%d" % (rec["rc"], rec["wall_s"], len(rec["rows"])), flush=True)
                if rec["rc"] != 0:
                    print("  " + p.stderr[-300:].strip().replace("\n", "\n  "), flush=True)


if __name__ == "__main__":
    main()

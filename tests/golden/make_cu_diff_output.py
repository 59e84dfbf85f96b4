"""Writes tests/golden/cu_diff_output.json from the reference's printed run
of its diffusion example (examples/diff/output_cu of the reference tree, the
output of examples/diff/in_Cu_diff.py): the thermo rows (step, time, MSD, FE,
S diagonal) every 10 steps and the run's closing counts. Run from the
repository root with the reference tree at /root/reference:
python tests/golden/make_cu_diff_output.py"""
import json
import os
import re

SRC = "/root/reference/examples/diff/output_cu"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rows, counts = [], {}
    with open(SRC) as f:
        for line in f:
            if line.startswith("| 0"):
                v = [t.strip() for t in line.strip().strip("|").split("|")]
                rows.append([int(v[0])] + [float(t) for t in v[1:7]])
            m = re.match(r"(\w+): accep\w+ = (\d+) rejected = (\d+)", line)
            if m:
                counts[m.group(1)] = [int(m.group(2)), int(m.group(3))]
            m = re.match(r"(maximum order|maximum timestep|minimum timestep|time elapsed): ([0-9.e+-]+)", line)
            if m:
                counts[m.group(1)] = float(m.group(2))
    out = {"source": "examples/diff/output_cu: the thermo table and closing counts of examples/diff/in_Cu_diff.py "
                     "(Cu-DMD.cfg, Cu_mishin, N=5, 300 K, a_tol 1e-7, max_ngmres_iters 4, max_nnewton_iters 10, "
                     "nreset 10, ntally 10, run to t = 1e30)",
           "columns": ["step", "Time", "MSD", "FE", "S00", "S11", "S22"],
           "printed": "%e with 7 significant digits",
           "rows": rows, "counts": counts}
    with open(os.path.join(HERE, "cu_diff_output.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()

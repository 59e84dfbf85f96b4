"""Benchmark: EAM NVT atom-steps/s on one MI355X.

Workload (the metric's "8M atoms" case): fcc Cu, a = 3.615 A, 128^3 cells =
8,388,608 atoms, Johnson-form EAM tabulated to setfl with Nr = Nrho = 10000
and r_c = 4.95 A, positions jittered N(0, 0.05 A) (seed 2024), Maxwell
velocities at 300 K (seed 7), Nose-Hoover-chain NVT at 300 K, dt = 1 fs,
neighbor skin 0.5 A with on-device rebuilds. One step = one full MD step
(thermostat, kick/drift, skin check, rebuild when needed, both EAM passes,
energy and virial, second kick). Synthetic data; inputs resident in HBM.

Prints one JSON line with the metric, a roofline entry for the dominant
kernel (its duration measured with HIP events on the library's stream over
the timed region) and a CPU baseline (the oracle, one core, on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
# algorithmic HBM bytes per atom of the force launch: pass 2 reads position
# and type 16 B and F'(rho_i) 4 B and writes the force 12 B; the second
# half-kick fused into it reads and writes the FP64 velocity (24 + 24 B).
# Neighbor gathers are taken as cache-served and the list as overhead.
PASS2_BYTES_PER_ATOM = 80


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(ncell, seed_x=2024, seed_v=7, T=300.0):
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cu(nr=10000, nrho=10000, rc=4.95)
    x, H = pot.fcc(3.615, (ncell, ncell, ncell))
    x += np.random.default_rng(seed_x).normal(0.0, 0.05, x.shape)
    elem = np.zeros(len(x), np.uint8)
    masses = [pot.MASS["Cu"]]
    v = M.create_temp(elem, masses, T, seed_v)
    return p, x, H, elem, masses, v


def gpu_run(args):
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    t0 = time.time()
    p, x, H, elem, masses, v = build_workload(args.cells)
    n = len(x)
    log(f"[bench] workload built: {n} atoms in {time.time() - t0:.1f}s")
    ff = M.ForceFieldEAM.from_knots(p)
    md = M.AtomsMD(ff, H, x, elem, skin=0.5)
    md.x_d = v
    del x, v
    nvt = M.NVT(300.0, 1.0 * pot.FS)
    t0 = time.time()
    nvt.run(md, args.warmup, masses)
    log(f"[bench] warmup {args.warmup} steps in {time.time() - t0:.2f}s")
    md.profile(True)
    t0 = time.perf_counter()
    out = nvt.run(md, args.steps, masses)
    t1 = time.perf_counter()
    kern = {k: md.profile_read(k) for k in ("eam_rho", "eam_force", "build_list", "cell_sort", "integrate")}
    md.profile(False)
    wall = t1 - t0
    maxn, nfull, _ = md.list_stats()
    log(f"[bench] timed {args.steps} steps in {wall:.3f}s, rebuilds {nvt.rebuilds}, T {nvt.T_part:.2f} K, "
        f"PE/atom {out[0] / n:.6f} eV, list {nfull / n:.1f}/atom (max {maxn})")
    for k, (ms, c) in kern.items():
        log(f"[bench]   {k:11s} {ms:9.2f} ms total, {c:5d} launches, {ms / max(c, 1):8.3f} ms avg")
    return n, wall, kern, nvt, out, nfull


def cpu_baseline(seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p, x, H, elem, masses, v = build_workload(20)  # config 1: 32,000 atoms
    ff = O.OEam.from_knots(p)
    md = O.OMD(ff, H, x, elem, skin=0.5)
    t0 = time.perf_counter()
    md.nvt_run(v, masses, pot.KB_EV, 300.0, 1.0 * pot.FS, 2)
    per = (time.perf_counter() - t0) / 3.0
    steps = max(3, int(seconds / max(per, 1e-6)))
    t0 = time.perf_counter()
    md.nvt_run(v, masses, pot.KB_EV, 300.0, 1.0 * pot.FS, steps)
    dt = time.perf_counter() - t0
    return {"value": len(x) * steps / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (FP64 C restatement), fcc Cu 20^3 cells = {len(x)} atoms, {steps} NVT steps "
                      f"at 1 fs, same table, 1 thread, {dt:.1f} s"}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "eam_force_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy kernel (256 MiB, 20 launches) used to calibrate HBM counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import ctypes as C
        from mapp4py_amd import _lib
        g = C.c_double()
        _lib.check(_lib.lib().mapp_diag_copy(256 << 20, 20, C.byref(g)))
        print(json.dumps({"copy_bytes_per_launch": 2 * (256 << 20), "read_bytes": 256 << 20, "write_bytes": 256 << 20,
                          "launches": 21, "GBps": g.value}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    n, wall, kern, nvt, out, nfull = gpu_run(args)
    value = n * args.steps / wall
    ms_f, c_f = kern["eam_force"]
    avg_s = ms_f / max(c_f, 1) / 1e3
    achieved = PASS2_BYTES_PER_ATOM * n / avg_s if avg_s > 0 else 0.0
    traffic = load_traffic()
    roof = {"kernel": "k_eam_force (EAM pass 2 + second half-kick)", "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
            "avg_launch_ms": avg_s * 1e3, "launches": c_f,
            "algorithmic_bytes_per_launch": PASS2_BYTES_PER_ATOM * n,
            "list_inclusive_bytes_per_launch": PASS2_BYTES_PER_ATOM * n + 4 * nfull}
    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline (~{args.cpu_seconds:.0f} s of oracle work)")
        cpu = cpu_baseline(args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.3e} atom-steps/s")
    line = {
        "metric": "EAM atom-steps/s @8M atoms", "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (Johnson-form Cu EAM table, jittered fcc lattice, Maxwell velocities)",
        "config": {"workload": f"EAM NVT, fcc Cu {args.cells}^3 cells ({n} atoms), Johnson setfl Nr=Nrho=10000, "
                               f"r_c=4.95 A, skin 0.5 A, dt 1 fs, 300 K", "natoms": n, "rebuilds": nvt.rebuilds,
                   "pe_per_atom_eV": out[0] / n, "T_K": nvt.T_part},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

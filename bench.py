"""Benchmark: EAM NVT atom-steps/s on one MI355X (the headline value), with
the DMD free-energy evaluation rate and the batched GCMC attempt rate
reported beside it.

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
# algorithmic HBM bytes of a whole NVT step per atom: pass 1 28 B, pass 2 32 B,
# integrator 128 B (positions, velocities, forces, types, position records)
STEP_BYTES_PER_ATOM = 188


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
    log(f"[bench] LDS row windows: density from interval {md.lds_window()[0]}, force from {md.lds_window()[1]}")
    log(f"[bench] timed {args.steps} steps in {wall:.3f}s, rebuilds {nvt.rebuilds}, T {nvt.T_part:.2f} K, "
        f"PE/atom {out[0] / n:.6f} eV, list {nfull / n:.1f}/atom (max {maxn})")
    for k, (ms, c) in kern.items():
        log(f"[bench]   {k:11s} {ms:9.2f} ms total, {c:5d} launches, {ms / max(c, 1):8.3f} ms avg")
    return n, wall, kern, nvt, out, nfull, md, p, H, elem


def headline_error(md, p, H, elem, threads):
    """Forces, energy and virial of the GPU at the end of the timed run against
    the FP64 oracle (threaded) on the same 8,388,608 positions."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    t0 = time.perf_counter()
    f, e, FH = md.force_calc()
    x = md.x
    n = len(x)
    with O.threads(threads):
        o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=0.5)
        fo, eo, FHo, _ = o.force()
        del o
    df = np.abs(f - fo)
    dt = time.perf_counter() - t0
    return {"max_abs_df_eV_per_A": float(df.max()), "rms_df_eV_per_A": float(np.sqrt((df ** 2).mean())),
            "bar_eV_per_A": 1e-5, "abs_dE_eV": float(abs(e - eo)), "abs_dE_per_atom_eV": float(abs(e - eo) / n),
            "max_abs_dF_H_eV": float(np.abs(FH - FHo).max()), "max_abs_F_H_eV": float(np.abs(FHo).max()),
            "natoms": n, "oracle_threads": threads, "seconds": dt,
            "what": "GPU force_calc at the final state of the timed run vs the FP64 oracle on the same positions"}


def host_threads():
    """threads for the all-cores CPU baseline: the process's CPU share (the
    OpenMP setting where one is given, else the affinity mask), at most 16"""
    env = os.environ.get("OMP_NUM_THREADS", "")
    k = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
    return max(1, min(16, k))


def timed_median(run, reps=5):
    """one untimed warm-up call, then the median of reps timed calls"""
    run()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        t.append(time.perf_counter() - t0)
    return float(np.median(t)), t


def cpu_baseline(seconds):
    """The oracle's NVT (FP64 C restatement of MDNVT::run over the EAM force)
    on config 1's 32,000 Cu atoms, timed on 1 thread and on all the process's
    cores (OpenMP over atoms, the same per-pair arithmetic): one warm-up and
    the median of 5 repetitions each, the step count sized so both modes
    together take about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p, x, H, elem, masses, v = build_workload(20)  # config 1: 32,000 atoms
    ff = O.OEam.from_knots(p)
    n = len(x)
    threads = host_threads()
    out = {}
    for mode, k in (("one_thread", 1), ("all_cores", threads)):
        with O.threads(k) as used:
            md = O.OMD(ff, H, x, elem, skin=0.5)
            t0 = time.perf_counter()
            md.nvt_run(v, masses, pot.KB_EV, 300.0, 1.0 * pot.FS, 4)
            per = (time.perf_counter() - t0) / 4.0
            # 6 runs (warm-up + 5) of this mode within half the budget
            steps = max(4, int(seconds / 2 / 6 / max(per, 1e-6)))
            med, all_t = timed_median(lambda: md.nvt_run(v, masses, pot.KB_EV, 300.0, 1.0 * pot.FS, steps))
        out[mode] = {"value": n * steps / med, "threads": used, "steps": steps, "median_s": med,
                     "runs_s": [round(t, 4) for t in all_t]}
    allc = out["all_cores"]
    return {"value": allc["value"], "unit": "atom-steps/s", "cores": allc["threads"], "kind": "port",
            "one_thread": out["one_thread"]["value"], "all_cores": allc, "one_thread_detail": out["one_thread"],
            "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "timing": "median of 5 after 1 warm-up",
            "sample": f"oracle (FP64 C restatement), fcc Cu 20^3 cells = {n} atoms, NVT at 1 fs, same table: "
                      f"{out['one_thread']['steps']} steps on 1 thread and {allc['steps']} steps on "
                      f"{allc['threads']} OpenMP threads per run"}


def nst_run(md, n, steps=100, warmup=10):
    """MDNST on the headline's 8,388,608 Cu atoms after their NVT run: all six
    stress components driven to zero at 300 K, dt 1 fs, box relaxation times
    1000 dt; the box changes every step and the list is rebuilt on the skin
    flag only."""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    masses = [pot.MASS["Cu"]]
    nst = M.NST(np.zeros((3, 3)), 300.0, 1.0 * pot.FS, kB=pot.KB_EV)
    H0 = md.H.copy()
    nst.run(md, warmup, masses)
    md.profile(True)
    t0 = time.perf_counter()
    nst.run(md, steps, masses)
    wall = time.perf_counter() - t0
    kern = {k: md.profile_read(k) for k in ("eam_rho", "eam_force", "build_list", "integrate")}
    md.profile(False)
    h = nst.hist
    vol = h[:, 2] * h[:, 4] * h[:, 7]
    return {"workload": f"NST (MDNST), the headline's {n} Cu atoms after their NVT run, S = 0 on all six "
                        f"components, 300 K, dt 1 fs, tau 1000 dt",
            "metric": "NST atom-steps/s", "value": n * steps / wall, "unit": "atom-steps/s", "steps": steps,
            "warmup": warmup, "ms_per_step": wall / steps * 1e3, "rebuilds": nst.rebuilds, "T_K": nst.T_part,
            "box_change_A": float(np.abs(md.H - H0).max()), "relative_volume_change": float(vol[-1] / vol[0] - 1.0),
            "stress_last_xx_yx_zx_yy_zy_zz": [float(t) for t in h[-1, 8:]],
            "kernel_ms": {k: ms / max(c, 1) for k, (ms, c) in kern.items()},
            "kernel_launches": {k: c for k, (ms, c) in kern.items()}, "dtype": "f32+f64"}


def nst_cpu_baseline(seconds):
    """the oracle's MDNST on 32,000 Cu atoms (fcc 20^3 cells), one thread"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p, x, H, elem, masses, v = build_workload(20)
    S = np.zeros((3, 3))
    o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=0.5)
    t0 = time.perf_counter()
    o.nst_run(v.copy(), masses, pot.KB_EV, 300.0, 1.0 * pot.FS, S, 3)
    per = (time.perf_counter() - t0) / 3.0
    steps = max(3, int(seconds / max(per, 1e-6)))
    o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=0.5)
    t0 = time.perf_counter()
    o.nst_run(v.copy(), masses, pot.KB_EV, 300.0, 1.0 * pot.FS, S, steps)
    dt = time.perf_counter() - t0
    return {"value": len(x) * steps / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle MDNST (FP64 C restatement), fcc Cu 20^3 cells = {len(x)} atoms, S = 0, {steps} steps, "
                      f"1 thread, {dt:.1f} s"}


def muvt_system():
    """fcc Ni 14^3 cells (10,976 atoms, 0.05 A jitter) with H on 200 random
    octahedral sites, NiAlH_jea EAM"""
    from mapp4py_amd import potentials as pot

    rng = np.random.default_rng(41)
    a = 3.52
    x, H = pot.fcc(a, (14, 14, 14))
    x = x + rng.normal(0.0, 0.05, x.shape)
    corners = np.argwhere(np.ones((14, 14, 14), bool)).astype(float) * a
    octa = corners[rng.choice(len(corners), 200, replace=False)] + np.array([a / 2, 0.0, 0.0])
    elem = np.concatenate([np.zeros(len(x), np.uint8), np.full(len(octa), 2, np.uint8)])
    return np.concatenate([x, octa]), H, elem


def muvt_run(steps=1000, nevery=100, nattempts=1000):
    """MDMuVT (src/md_muvt.cpp): NVT of Ni-H at 300 K with nattempts exchange
    attempts of H at mu = -2.3 eV every nevery steps"""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    x, H, elem = muvt_system()
    masses = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
    g = M.AtomsMD(M.ForceFieldEAM.setfl(pot.NIALH_JEA), H, x, elem, skin=0.5)
    g.x_d = M.create_temp(elem, masses, 300.0, 43)
    run = M.MuVT(-2.3, 300.0, pot.FS, 2, 2411895, nevery=nevery, nattempts=nattempts)
    run.run(g, nevery, masses)  # warm-up: one exchange
    g.profile(True)
    t0 = time.perf_counter()
    run.run(g, steps, masses)
    wall = time.perf_counter() - t0
    kern = {k: g.profile_read(k) for k in ("eam_rho", "eam_force", "build_list", "integrate")}
    g.profile(False)
    nx = steps // nevery
    return {"workload": f"MDMuVT, fcc Ni 14^3 cells + 200 H (10,976 + 200 atoms), NiAlH_jea, mu = -2.3 eV, 300 K, "
                        f"dt 1 fs, {nattempts} attempts every {nevery} steps",
            "metric": "muVT atom-steps/s", "value": len(x) * steps / wall, "unit": "atom-steps/s", "steps": steps,
            "ms_per_step": wall / steps * 1e3, "exchanges": nx, "attempts_per_s": nx * nattempts / wall,
            "counts_ins_del_acc_ins_del_tried": [int(c) for c in run.counts], "natms_end": g.natms,
            "gas_end": int(run.hist[-1, 1]), "T_K": run.T_part, "builds": run.builds,
            "kernel_ms": {k: ms / max(c, 1) for k, (ms, c) in kern.items()}, "dtype": "f32+f64 (exchange f64)"}


def muvt_cpu_baseline():
    """the oracle's MDMuVT on the same system: 200 steps with two exchanges"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    x, H, elem = muvt_system()
    masses = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
    v = O.o_create_T(elem, masses, pot.KB_EV, 300.0, 43)
    o = O.OMD(O.OEam.read(pot.NIALH_JEA), H, x, elem, skin=0.5)
    hp = 4.13566766225 * 0.1 * (1.60217656535 / 1.66053904020) ** 0.5
    t0 = time.perf_counter()
    o.muvt_run(v, masses, pot.KB_EV, hp, 300.0, pot.FS, -2.3, 2, 2411895, 100, 1000, 200)
    dt = time.perf_counter() - t0
    return {"value": len(x) * 200 / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle MDMuVT on the same {len(x)} atoms, 200 steps with 2 exchanges of 1000 attempts, "
                      f"1 thread, {dt:.1f} s"}


def feni_run(steps=1000, warmup=20):
    """Binary Fe-Ni eam/fs alloy: bcc 128 x 64 x 64 cells x 2 = 1,048,576 atoms, Ni on a random
    20% of sites, Nr = Nrho = 5000, r_c = 5.3 A, NVT at 600 K and 1 fs with
    on-device rebuilds (the multi-element passes: global spline tables)."""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    p = pot.johnson_feni_fs(nr=5000, nrho=5000)
    x, H = pot.bcc(2.86, (128, 64, 64))
    x = x + np.random.default_rng(31).normal(0.0, 0.05, x.shape)
    elem = (np.random.default_rng(32).random(len(x)) < 0.2).astype(np.uint8)
    masses = [pot.MASS["Fe"], pot.MASS["Ni"]]
    md = M.AtomsMD(M.ForceFieldEAM.from_knots(p), H, x, elem, skin=0.5)
    md.x_d = M.create_temp(elem, masses, 600.0, 33)
    nvt = M.NVT(600.0, 1.0 * pot.FS)
    nvt.run(md, warmup, masses)
    md.profile(True)
    t0 = time.perf_counter()
    out = nvt.run(md, steps, masses)
    wall = time.perf_counter() - t0
    kern = {k: md.profile_read(k) for k in ("eam_rho", "eam_force", "build_list", "integrate")}
    md.profile(False)
    n = len(x)
    maxn, nfull, _ = md.list_stats()
    return {"workload": f"EAM/FS Fe-Ni NVT, bcc 128 x 64 x 64 cells x 2 = {n:,} atoms (20% Ni), Nr=Nrho=5000, r_c=5.3 A, "
                        "skin 0.5 A, dt 1 fs, 600 K",
            "metric": "EAM atom-steps/s", "value": n * steps / wall, "unit": "atom-steps/s", "steps": steps,
            "ms_per_step": wall / steps * 1e3, "rebuilds": nvt.rebuilds, "T_K": nvt.T_part,
            "pe_per_atom_eV": out[0] / n, "list_per_atom": nfull / n,
            "kernel_ms": {k: ms / max(c, 1) for k, (ms, c) in kern.items()}, "dtype": "f32+f64"}


# minimiser vector kernels, bytes per atom as laid out (double4 rows of 32 B,
# float4 forces of 16 B): one L-BFGS direction with two stored pairs (begin
# 176, four two-loop passes 144 + 112 + 144 + 112), the pair update 144, and
# 144 per line-search trial point (x0, h, build positions read, x and the
# 16-byte record written)
MIN_VEC_BYTES_PER_ATOM_STEP = 832
MIN_VEC_BYTES_PER_ATOM_EVAL = 144


def min_run(steps=30):
    """MinLBFGS (m = 2, back-tracking line search, e_tol 0 so every step is
    taken) on fcc Cu 64^3 cells less 0.1% vacancies (1,047,528 atoms),
    positions jittered by 0.1 A, Johnson setfl Nr = Nrho = 10000, skin 0.5 A."""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cu(nr=10000, nrho=10000)
    x, H = pot.fcc(3.615, (64, 64, 64))
    rng = np.random.default_rng(41)
    x = np.delete(x, rng.choice(len(x), len(x) // 1000, replace=False), axis=0)
    x = x + rng.normal(0.0, 0.1, x.shape)
    elem = np.zeros(len(x), np.uint8)
    md = M.AtomsMD(M.ForceFieldEAM.from_knots(p), H, x, elem, skin=0.5)
    M.MinLBFGS(m=2, e_tol=0.0).run(md, 3)  # warm-up: list, tables, code objects
    md.x = x
    mn = M.MinLBFGS(m=2, e_tol=0.0)
    t0 = time.perf_counter()
    r = mn.run(md, steps)
    wall = time.perf_counter() - t0
    # the same run again with per-kernel HIP events for the breakdown (the
    # events add host work, so the timed run above goes without them)
    md.x = x
    md.profile(True)
    r_p = M.MinLBFGS(m=2, e_tol=0.0).run(md, steps)
    assert r_p["steps"] == r["steps"] and r_p["evals"] == r["evals"]
    groups = ("eam_rho", "eam_force", "eam_energy", "build_list", "cell_sort", "md_min_vec")
    kern = {k: md.profile_read(k) for k in groups}
    md.profile(False)
    n = len(x)
    vec_ms = kern["md_min_vec"][0]
    vec_bytes = n * (MIN_VEC_BYTES_PER_ATOM_STEP * r["steps"] + MIN_VEC_BYTES_PER_ATOM_EVAL * r["evals"])
    vec_gbps = vec_bytes / (vec_ms / 1e3) / 1e9 if vec_ms > 0 else 0.0
    return {"workload": f"MinLBFGS m=2 (ls_bt c=0.4 rho=0.5), fcc Cu 64^3 cells less 0.1% vacancies ({n} atoms), "
                        "0.1 A jitter, Johnson setfl Nr=Nrho=10000, skin 0.5 A, e_tol 0",
            "metric": "minimiser atom-steps/s", "value": n * r["steps"] / wall, "unit": "atom-steps/s",
            "steps": r["steps"], "ms_per_step": wall / max(r["steps"], 1) * 1e3, "energy_evals": r["evals"],
            "force_calcs": r["force_calcs"], "rebuilds": r["rebuilds"], "status": r["message"],
            "e0_per_atom_eV": r["e0"] / n, "e_per_atom_eV": r["e"] / n, "f_max_eV_per_A": r["f_max"],
            "kernel_ms_total": {k: ms for k, (ms, c) in kern.items()},
            "kernel_launches": {k: c for k, (ms, c) in kern.items()},
            "vector_kernels": {"bound": "hbm", "achieved": vec_gbps, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                               "frac": vec_gbps * 1e9 / HBM_PEAK, "algorithmic_bytes": vec_bytes},
            "dtype": "f64 vectors, f32 forces"}


def min_cpu_baseline(seconds):
    """The oracle's MinLBFGS (oracle/o_min.py over the FP64 C oracle's EAM) on
    fcc Cu 20^3 cells less 0.1% vacancies, 1 thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import o_min
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cu(nr=10000, nrho=10000)
    x, H = pot.fcc(3.615, (20, 20, 20))
    rng = np.random.default_rng(41)
    x = np.delete(x, rng.choice(len(x), len(x) // 1000, replace=False), axis=0)
    x = x + rng.normal(0.0, 0.1, x.shape)
    elem = np.zeros(len(x), np.uint8)
    o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=0.5)
    t0 = time.perf_counter()
    o.force()
    per = time.perf_counter() - t0
    steps = max(2, int(seconds / (3.0 * max(per, 1e-6))))
    t0 = time.perf_counter()
    r = o_min.min_lbfgs(o, steps, m=2, e_tol=0.0)
    dt = time.perf_counter() - t0
    return {"value": len(x) * r["steps"] / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle MinLBFGS m=2 over the FP64 C restatement, fcc Cu 20^3 cells less 0.1% vacancies "
                      f"= {len(x)} atoms, {r['steps']} steps, 1 thread, {dt:.1f} s"}


# algorithmic bytes of one __J product: per list entry the partner index and
# rho_{j->i} (first sum) and the index, rho_{i->j} and t0 (second); per
# second-list entry the index and three coefficients; per component the
# counts, F'', S0..S2, c, x, A1, b0, xs and Ax once each
J_BYTES_PER_ENTRY = 32
J_BYTES_PER_EXCHANGE = 28
J_BYTES_PER_COMPONENT = 120
# algorithmic bytes of one J(x, alpha) product of the second-derivative pass:
# per list entry the index and seven slot values (first pass) and the index
# and four slot values (second); per degree of freedom (3 per site + 1 per
# component) the input read by both passes and the output written twice
JXA_BYTES_PER_ENTRY = 60 + 36
JXA_BYTES_PER_DOF = 32
DMD_BYTES_PER_SITE = 132  # x 16, alpha 8, c 16 (read twice), f 12, f_alpha 8, mu 16, E/dE write + read 32


def dmd_run(evals=10):
    """Cu-H DMD, 262,144 sites (fcc 32x32x64 cells), two components per site,
    N = 3 Gauss-Hermite points, r_c = 5.5 A: full free-energy + gradient
    evaluations (ForceFieldDMD::derivative) per second."""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((32, 32, 64), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    d = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    d.derivative()

    # device-resident evaluations (f, f_alpha and mu stay in HBM), timed
    # without profiler events; the passes' kernel times from a second,
    # event-timed run of the same evaluations
    def timed_evals():
        t0 = time.perf_counter()
        for _ in range(evals):
            d.evaluate()
        wall = time.perf_counter() - t0
        d.profile(True)
        for _ in range(evals):
            d.evaluate()
        ks = d.profile_read("dmd_pass1"), d.profile_read("dmd_pass2")
        d.profile(False)
        return wall, ks

    dt, (k1, k2) = timed_evals()
    # the same with f, f_alpha and mu copied to host arrays every time (PCIe included)
    t0 = time.perf_counter()
    for _ in range(evals):
        d.derivative()
    dt_host = time.perf_counter() - t0
    f0, fa0, mu0, fe0, FH0, _ = d.derivative()
    half_pairs = int(d.half_pairs()[0].shape[0])
    n = len(x)
    ms = dt / evals * 1e3
    # FP32 pair splines (FP64 moments and sums): the same evaluations from
    # FP32 knots held in LDS and from FP32 coefficient rows, with their error
    # against the FP64 rows (which equal the oracle to 1e-9)
    def fp32_leg(rows):
        d.set_fp32(True, rows=rows)
        d.evaluate()
        dt32, (k1f, k2f) = timed_evals()
        f32r, fa32, mu32, fe32, FH32, _ = d.derivative()
        first = d.knot_window()
        d.set_fp32(False)
        what = ("f32 table rows" if rows else
                f"f32 knot values and slopes, knots from r = {first * p.dr:.2f} A in LDS") + ", f64 moments and sums"
        return {"value": n * evals / dt32, "ms_per_eval": dt32 / evals * 1e3, "pass1_ms": k1f[0] / max(k1f[1], 1),
                "pass2_ms": k2f[0] / max(k2f[1], 1), "speed_vs_fp64_rows": dt / dt32,
                "max_abs_df_x_eV_per_A": float(np.abs(f32r - f0).max()),
                "max_abs_df_alpha_eV_per_A": float(np.abs(fa32 - fa0).max()),
                "max_abs_dmu_eV": float(np.abs(mu32 - mu0).max()),
                "free_energy_rel_err": abs(fe32 - fe0) / abs(fe0),
                "virial_rel_err": float(np.abs(FH32 - FH0).max() / np.abs(FH0).max()),
                "dtype": what,
                "errors_against": "the FP64 rows on the GPU (equal to the FP64 oracle to 1e-9 relative)"}

    fp32_knots = fp32_leg(False)
    fp32 = fp32_leg(True)
    # the learned surrogate (config 4): trained on the GPU from this rule, then the same evaluation
    s = M.Surrogate(2, box=(2.2, 6.1, 0.05, 0.30), seed=1)
    t0 = time.perf_counter()
    l0, l1 = s.train(d, nsamples=2_000_000, steps=8000, batch=4096, lr=2e-3)
    t_train = time.perf_counter() - t0
    d.use_surrogate(s)
    f1, fa1, _, fe1, _, _ = d.derivative()
    t0 = time.perf_counter()
    for _ in range(evals):
        d.evaluate()
    dts = time.perf_counter() - t0
    nets, quad = d.surrogate_coverage()
    d.use_surrogate(None)
    d16 = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=16, masses=p.masses)
    d16.set_fp32(True)
    d16.derivative()
    t0 = time.perf_counter()
    for _ in range(3):
        d16.evaluate()
    ms16 = (time.perf_counter() - t0) / 3 * 1e3
    del d16
    # the accuracy / speed frontier: for each Gauss-Hermite order, networks
    # trained from that rule (box over its whole list range) against its
    # quadrature with FP64 rows and FP32 knots, same sites
    frontier = []
    for N in (3, 5, 8, 16):
        dn = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=N, masses=p.masses)
        fq = dn.derivative()[0]
        t0 = time.perf_counter()
        for _ in range(3):
            dn.evaluate()
        q64 = (time.perf_counter() - t0) / 3 * 1e3
        dn.set_fp32(True)
        dn.evaluate()
        t0 = time.perf_counter()
        for _ in range(3):
            dn.evaluate()
        q32 = (time.perf_counter() - t0) / 3 * 1e3
        dn.set_fp32(False)
        xmax = float(M.gauss_hermite(N)[0].max())
        sn = M.Surrogate(2, box=(2.2, 5.5 + xmax * np.sqrt(2.0) * 0.30 + 0.05, 0.05, 0.30), seed=1)
        sn.train(dn, nsamples=1_000_000, steps=4000, batch=4096, lr=2e-3)
        dn.use_surrogate(sn)
        fs = dn.derivative()[0]
        dn.evaluate()
        t0 = time.perf_counter()
        for _ in range(3):
            dn.evaluate()
        qs = (time.perf_counter() - t0) / 3 * 1e3
        cov = dn.surrogate_coverage()
        dn.use_surrogate(None)
        frontier.append({"N": N, "quadrature_fp64_ms": q64, "quadrature_fp32_knots_ms": q32, "networks_ms": qs,
                         "networks_speed_vs_fp64": q64 / qs, "networks_speed_vs_fp32_knots": q32 / qs,
                         "networks_max_abs_df_x_eV_per_A": float(np.abs(fs - fq).max()),
                         "quadrature_pairs_left": int(cov[1])})
        log(f"[bench] frontier N={N}: quadrature {q64:.2f} ms (fp32 knots {q32:.2f}), networks {qs:.2f} ms, "
            f"max|df| {frontier[-1]['networks_max_abs_df_x_eV_per_A']:.1e} eV/A")
        del dn, sn
    # concentration rates (ForceFieldDMD::c_d_calc) at the same state: both
    # passes for the chemical potentials, then the exchange sums over the
    # second list (first fcc shell, r_crd 3.5 A, C 1.7 as examples/diff)
    r_crd, zeta = [3.5, 3.5], [1.7, 1.7]
    d.c_d(r_crd, zeta, fetch=False)
    d.profile(True)
    t0 = time.perf_counter()
    for _ in range(evals):
        _, cd_norm = d.c_d(r_crd, zeta, fetch=False)
    dt_cd = time.perf_counter() - t0
    kcd = d.profile_read("dmd_cd")
    # Jacobian products (ForceFieldEAMDMD::__J) at that state: the first one
    # stores the smeared values of every list slot and F''; later ones are
    # three gathers over the list. Caller-order x in, Ax out (host copies).
    v = np.random.default_rng(7).normal(0.0, 1.0, c.shape)
    d.J(v)
    kjs = d.profile_read("dmd_j_state")
    t0 = time.perf_counter()
    for _ in range(evals):
        d.J(v)
    dt_j = time.perf_counter() - t0
    kj = d.profile_read("dmd_j")
    d.profile(False)
    j_entries, j_exch = d.J_entries()
    # second-derivative products J(x, alpha) (__prepJ_n_res + __J) at that
    # state: every call runs pass 1 and the slot pass, then one product
    rng = np.random.default_rng(8)
    dxa, daa = rng.normal(0.0, 0.01, (n, 3)), rng.normal(0.0, 1e-3, c.shape)
    d.Jxa(dxa, daa)
    d.profile(True)
    t0 = time.perf_counter()
    for _ in range(evals):
        d.Jxa(dxa, daa)
    dt_xa = time.perf_counter() - t0
    kprep, kxa = d.profile_read("dmd_prepj"), d.profile_read("dmd_jxa")
    d.profile(False)
    xa_ms = kxa[0] / max(kxa[1], 1)
    xa_bytes = JXA_BYTES_PER_ENTRY * j_entries + JXA_BYTES_PER_DOF * (3 * n + c.size)
    # one implicit backward-Euler step of the concentrations (DAEImplicit::newton,
    # GMRES over J): c - c_n = dt c_d(c) from c_n, dt = 30 (time units of
    # the table: sqrt(amu/eV) A), then c restored for the CG leg
    xs0, as0, cs0 = d.state()
    dt_be = 30.0
    d.profile(True)
    t0 = time.perf_counter()
    nt_ok, nt_it = d.newton(r_crd, zeta, dt_be, -cs0 / dt_be, a_tol=1e-8)
    dt_nt = time.perf_counter() - t0
    nt_k = {k: d.profile_read(k) for k in ("dmd_pass1", "dmd_pass2", "dmd_cd", "dmd_j_state", "dmd_j", "dmd_newton")}
    d.profile(False)
    nt_dc = float(np.abs(d.state()[2] - cs0).max())
    d.set_state(xs0, as0, cs0)
    # 20 variable-order BDF steps of the concentrations (DAEBDF), x and alpha fixed
    d.profile(True)
    t0 = time.perf_counter()
    bdf_rep, _ = d.bdf_run(r_crd, zeta, 1e6, a_tol=1e-8, max_nsteps=20)
    dt_bdf = time.perf_counter() - t0
    bdf_k = {k: d.profile_read(k) for k in ("dmd_pass1", "dmd_pass2", "dmd_cd", "dmd_j_state", "dmd_j", "dmd_newton")}
    d.profile(False)
    bdf_dc = float(np.abs(d.state()[2] - cs0).max())
    d.set_state(xs0, as0, cs0)
    j_bytes = J_BYTES_PER_ENTRY * j_entries + J_BYTES_PER_EXCHANGE * j_exch + J_BYTES_PER_COMPONENT * c.size
    # config 3's solver: CG on the free energy from the same state (x and alpha)
    t0 = time.perf_counter()
    cg = d.min_cg(max_iter=400, e_tol=0.0, grad_tol=1e-6)
    cg_s = time.perf_counter() - t0
    ms_s = dts / evals * 1e3
    speed = fp32_knots["ms_per_eval"] / ms_s
    df_s = float(np.abs(f1 - f0).max())
    surr = {"value": n * evals / dts, "unit": "site-evals/s", "ms_per_eval": ms_s,
            "speed_vs_quadrature_N3": speed, "speed_vs_fp64_quadrature_N3": dt / dts,
            "quadrature_N3": "the FP32-knot passes (the leg's value)", "quadrature_N16_ms_per_eval": ms16,
            "speed_vs_quadrature_N16": ms16 / (dts / evals * 1e3),
            "max_abs_df_x_eV_per_A": df_s,
            "rms_df_x_eV_per_A": float(np.sqrt(((f1 - f0) ** 2).mean())),
            "max_abs_df_alpha_eV_per_A": float(np.abs(fa1 - fa0).max()),
            "free_energy_error_eV_per_site": abs(fe1 - fe0) / n, "network_calls": nets, "quadrature_pairs": quad,
            "train": {"samples_per_net": 2_000_000, "steps": 8000, "batch": 4096, "loss0": l0, "loss": l1,
                      "seconds": t_train, "dtype": "bf16 MFMA, fp32 accumulate"},
            "frontier": frontier,
            "target": f">= 5x the N=3 quadrature at <= 1e-4 eV/A: "
                      f"{'met' if speed >= 5.0 and df_s <= 1e-4 else 'not met'} "
                      f"(speed {speed:.2f}x, forces {df_s:.1e} eV/A)"}
    return {"workload": "DMD Cu-H, 262,144 fcc sites x 2 components, N=3 Gauss-Hermite, r_c=5.5 A, "
                        "alpha log-uniform [0.05,0.30] A, c_H ~ Beta(2,5)",
            "metric": "DMD site-evals/s (free energy + gradient)", "value": fp32_knots["value"],
            "unit": "site-evals/s", "ms_per_eval": fp32_knots["ms_per_eval"], "pass1_ms": fp32_knots["pass1_ms"],
            "pass2_ms": fp32_knots["pass2_ms"],
            "error_vs_fp64": {k: fp32_knots[k] for k in ("max_abs_df_x_eV_per_A", "max_abs_df_alpha_eV_per_A",
                                                         "max_abs_dmu_eV", "free_energy_rel_err", "virial_rel_err",
                                                         "errors_against")},
            "fp64_rows": {"value": n * evals / dt, "ms_per_eval": ms, "pass1_ms": k1[0] / max(k1[1], 1),
                          "pass2_ms": k2[0] / max(k2[1], 1),
                          "dtype": "f64 table rows, moments and sums (the parity mode: the FP64 oracle to 1e-9)"},
            "fp32_knots": fp32_knots, "fp32_rows": fp32,
            "host_copy_inclusive": {"value": n * evals / dt_host, "ms_per_eval": dt_host / evals * 1e3,
                                    "note": "f, f_alpha, mu gathered to caller order and copied to host each time"},
            "dtype": "f32 pair splines, f64 rho, free energy, mu and virial sums", "half_pairs": half_pairs,
            "hbm_fraction": DMD_BYTES_PER_SITE * n / (fp32_knots["ms_per_eval"] / 1e3) / HBM_PEAK,
            "c_d": {"metric": "concentration-rate evaluations (c_d_calc) site-evals/s", "value": n * evals / dt_cd,
                    "ms_per_eval": dt_cd / evals * 1e3, "exchange_kernel_ms": kcd[0] / max(kcd[1], 1),
                    "max_abs_c_d": cd_norm, "r_crd_A": r_crd, "C": zeta},
            "J": {"metric": "Jacobian products (__J) site-evals/s, device kernels", "value": n / (kj[0] / kj[1] / 1e3),
                  "kernels_ms": kj[0] / kj[1], "host_copy_inclusive_ms": dt_j / evals * 1e3,
                  "slot_state_ms": kjs[0] / max(kjs[1], 1),
                  "list_entries": j_entries, "exchange_entries": j_exch, "algorithmic_bytes": j_bytes,
                  "hbm_fraction": j_bytes / (kj[0] / kj[1] / 1e3) / HBM_PEAK},
            "Jxa": {"metric": "second-derivative products J(x, alpha) site-evals/s, device kernels",
                    "value": n / (xa_ms / 1e3), "product_ms": xa_ms,
                    "slot_pass_ms": kprep[0] / max(kprep[1], 1),
                    "host_copy_inclusive_ms": dt_xa / evals * 1e3,
                    "list_entries": j_entries, "algorithmic_bytes": xa_bytes,
                    "hbm_fraction": xa_bytes / (xa_ms / 1e3) / HBM_PEAK,
                    "note": "each host call also runs pass 1 and the slot pass (prepJ_n_res) for its state"},
            "implicit_step": {"solver": "DAEImplicit::newton, GMRES(5) over __J, a_tol 1e-8, backward Euler dt 30",
                              "converged": nt_ok, "newton_iters": nt_it, "ms": dt_nt * 1e3, "max_abs_dc": nt_dc,
                              "site_steps_per_s": n / dt_nt,
                              "kernels_ms": {k: v[0] for k, v in nt_k.items()},
                              "launches": {k: v[1] for k, v in nt_k.items()}},
            "bdf": {"solver": "DAEBDF, a_tol 1e-8, 20 steps from the same state, x and alpha fixed",
                    "steps": bdf_rep["nsteps"], "t_reached": bdf_rep["t"], "max_order": bdf_rep["max_q"],
                    "rejected": bdf_rep["integ_rej"], "seconds": dt_bdf, "ms_per_step": dt_bdf / max(bdf_rep["nsteps"], 1) * 1e3,
                    "site_steps_per_s": n * bdf_rep["nsteps"] / dt_bdf, "max_abs_dc": bdf_dc,
                    "kernels_ms": {k: v[0] for k, v in bdf_k.items()}},
            "cg": {"solver": "CG on free energy over x and alpha (MinCGDMD), grad tol 1e-6, max 400 iterations",
                   "iters": cg["iters"], "status": cg["status"], "fe0": cg["fe0"], "fe": cg["fe"],
                   "grad_norm": cg["grad_norm"], "grad_max": cg["grad_max"], "seconds": cg_s},
            "surrogate": surr}


def dmd_evals_only(evals=5):
    """the 262,144-site DMD free-energy + gradient evaluations alone: FP64
    rows, FP32 knots in LDS, FP32 rows (a short run for kernel traces and
    counters)"""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((32, 32, 64), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    d = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    out = {}
    for name, on, rows in (("fp64_rows", False, False), ("fp32_knots", True, False), ("fp32_rows", True, True)):
        d.set_fp32(on, rows=rows)
        d.evaluate()
        t0 = time.perf_counter()
        for _ in range(evals):
            d.evaluate()
        out[name] = (time.perf_counter() - t0) / evals * 1e3
    return {"workload": "DMD Cu-H, 262,144 sites x 2 components, N=3", "ms_per_eval": out}


def dmd_cpu_baseline(seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((4, 4, 8), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    o = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    t0 = time.perf_counter()
    o.force()
    per = time.perf_counter() - t0
    k = max(2, int(seconds / max(per, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(k):
        o.force()
    dt = time.perf_counter() - t0
    return {"value": len(x) * k / dt, "unit": "site-evals/s", "cores": 1, "kind": "port",
            "sample": f"oracle, same table and site statistics, fcc 4x4x8 cells = {len(x)} sites, {k} evaluations, "
                      f"{dt:.1f} s"}


def dmd_cd_cpu_baseline(seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((4, 4, 8), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    o = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    r_crd, zeta = [3.5, 3.5], [1.7, 1.7]
    t0 = time.perf_counter()
    o.c_d(r_crd, zeta, p.masses)
    per = time.perf_counter() - t0
    k = max(2, int(seconds / max(per, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(k):
        o.c_d(r_crd, zeta, p.masses)
    dt = time.perf_counter() - t0
    return {"value": len(x) * k / dt, "unit": "site-evals/s", "cores": 1, "kind": "port",
            "sample": f"oracle c_d_calc (both passes + exchange sums), fcc 4x4x8 cells = {len(x)} sites, {k} "
                      f"evaluations, {dt:.1f} s"}


def dmd_j_cpu_baseline(seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((4, 4, 8), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    o = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    r_crd, zeta = [3.5, 3.5], [1.7, 1.7]
    v = np.random.default_rng(7).normal(0.0, 1.0, c.shape)
    t0 = time.perf_counter()
    o.J(r_crd, zeta, p.masses, v)
    per = time.perf_counter() - t0
    k = max(2, int(seconds / max(per, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(k):
        o.J(r_crd, zeta, p.masses, v)
    dt = time.perf_counter() - t0
    return {"value": len(x) * k / dt, "unit": "site-evals/s", "cores": 1, "kind": "port",
            "sample": f"oracle __J, state rebuilt per call (calc_mu + rate coefficients), fcc 4x4x8 cells = "
                      f"{len(x)} sites, {k} products, {dt:.1f} s"}


def dmd_newton_cpu_baseline(seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((4, 4, 8), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    o = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    r_crd, zeta = [3.5, 3.5], [1.7, 1.7]
    k, t0 = 0, time.perf_counter()
    while True:
        o.set_state(x, alpha, c)
        o.newton(r_crd, zeta, p.masses, 30.0, -c / 30.0, a_tol=1e-8)
        k += 1
        if time.perf_counter() - t0 > seconds and k >= 2:
            break
    dt = time.perf_counter() - t0
    return {"value": len(x) * k / dt, "unit": "site-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle DAEImplicit::newton (its __J rebuilds the state per product), backward Euler dt 30, "
                      f"fcc 4x4x8 cells = {len(x)} sites, {k} steps, {dt:.1f} s"}


def example_replay(cpu_seconds):
    """The reference's diffusion example up to its first printed line:
    examples/diff/configs/Cu-DMD.cfg (16,384 sites, one vacancy) through the
    library's CFG importer, x and alpha relaxed at fixed c as DAE::min_error_false
    does (CG on the GPU, then the library's Newton stage: first with the
    example's a_tol 1e-7, 10 Newton steps and GMRES(4), then to a tight
    tolerance), FE and S against step 0 of examples/diff/output_cu."""
    import gzip
    import shutil
    import tempfile

    import mapp4py_amd as M

    import math

    GOLDEN = os.path.join(ROOT, "tests", "golden")
    HP_EXAMPLE = 4.13566766225 * 0.1 * math.sqrt(1.60217656535 / 1.66053904020)  # examples/diff/in_Cu_diff.py
    KB_EXAMPLE = 8.617330350e-5
    with open(os.path.join(GOLDEN, "cu_dmd_step0.json")) as fj:
        gd = json.load(fj)
    tmp = tempfile.mkdtemp()
    try:
        files = {}
        for name in ("Cu-DMD.cfg.gz", "Cu_mishin.eam.alloy.gz"):
            files[name] = os.path.join(tmp, name[:-3])
            with gzip.open(os.path.join(GOLDEN, name), "rb") as f, open(files[name], "wb") as g:
                shutil.copyfileobj(f, g)
        t0 = time.perf_counter()
        d = M.import_cfg_dmd(files["Cu-DMD.cfg.gz"])
        t_imp = time.perf_counter() - t0
        ff = M.ForceFieldEAM.setfl(files["Cu_mishin.eam.alloy.gz"], ["Cu"], dmd=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    g = M.AtomsDMD(ff, d["H"], d["x"], d["elem"], d["alpha"], d["c"], nquad=5, T=300.0, kB=KB_EXAMPLE,
                   hP=HP_EXAMPLE, masses=d["masses"])
    n = len(d["c"])
    g.derivative()
    t0 = time.perf_counter()
    for _ in range(20):
        g.derivative()
    gpu_eval = (time.perf_counter() - t0) / 20
    t0 = time.perf_counter()
    r = g.min_cg(max_iter=2000, e_tol=0.0, grad_tol=1e-7)
    t_cg = time.perf_counter() - t0
    t0 = time.perf_counter()
    k_ex, res_ex = g.min_newton(1e-7, 10, 4)
    t_nx = time.perf_counter() - t0
    t0 = time.perf_counter()
    k_t, res_t = g.min_newton(1e-11, 20, 8)
    t_nt = time.perf_counter() - t0
    f, fa, _, fe, FH, _ = g.derivative()
    S = -np.asarray(FH) / np.linalg.det(d["H"])
    out = {"workload": "examples/diff Cu-DMD.cfg, 16,384 sites, Cu_mishin setfl, N=5, 300 K",
           "import_ms": t_imp * 1e3, "import_sites_per_s": n / t_imp,
           "gpu_evals_per_s": 1.0 / gpu_eval, "gpu_ms_per_eval": gpu_eval * 1e3,
           "cg": {"iters": r["iters"], "seconds": t_cg, "fe": r["fe"], "grad_max": r["grad_max"]},
           "newton_example_settings": {"a_tol": 1e-7, "max_nnewton_iters": 10, "max_ngmres_iters": 4,
                                       "steps": k_ex, "residual": res_ex, "seconds": t_nx},
           "newton_tight": {"a_tol": 1e-11, "max_nnewton_iters": 20, "max_ngmres_iters": 8, "steps": k_t,
                            "residual": res_t, "seconds": t_nt},
           "fe": fe, "S_diag": [S[0], S[3], S[5]],
           "max_grad": float(max(np.abs(f).max(), np.abs(fa).max())),
           "reference_output_step0": {"FE": gd["FE"], "S_diag": gd["S_diag"][0]},
           "abs_err_fe": abs(fe - gd["FE"]), "abs_err_S": float(np.abs(S[[0, 3, 5]] - gd["S_diag"][0]).max())}
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O

        tmp = tempfile.mkdtemp()
        try:
            pot = os.path.join(tmp, "Cu_mishin.eam.alloy")
            with gzip.open(os.path.join(GOLDEN, "Cu_mishin.eam.alloy.gz"), "rb") as fsrc, open(pot, "wb") as fdst:
                shutil.copyfileobj(fsrc, fdst)
            oe = O.OEam.read(pot, 0, dmd=True)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        o = O.ODMD(oe, d["H"], d["x"], d["elem"], d["alpha"], d["c"], nquad=5, kB=KB_EXAMPLE, T=300.0,
                   hP=HP_EXAMPLE, masses=d["masses"])
        k, t0 = 0, time.perf_counter()
        while k < 2 or time.perf_counter() - t0 < cpu_seconds:
            o.force()
            k += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / dt, "unit": "evals/s", "cores": 1, "kind": "port",
                               "sample": f"oracle free energy + gradient of the same 16,384-site state, {k} "
                                         f"evaluations, {dt:.1f} s"}
    return out


def dmd_bdf_cpu_baseline():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot
    from o_bdf import OBdf

    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((4, 4, 8), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    o = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    b = OBdf(o, [3.5, 3.5], [1.7, 1.7], p.masses, a_tol=1e-8, max_nsteps=20, x=x, alpha=alpha)
    t0 = time.perf_counter()
    k = b.run(1e6)
    dt = time.perf_counter() - t0
    return {"value": len(x) * k / dt, "unit": "site-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle DAEBDF (numpy control over the C oracle; its __J rebuilds the state per product), "
                      f"fcc 4x4x8 cells = {len(x)} sites, {k} steps, {dt:.1f} s"}


def gcmc_run(launches=10, attempts=64):
    """4096 periodic Ni-H cells (108-4000 atoms, log-uniform), NiAlH_jea EAM,
    H exchange at mu = -2.3 eV and 300 K: one persistent HIP-graph launch per
    64 attempts in every cell, tile shape from the tuner."""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    cells = pot.gcmc_nih_cells(4096, seed=6)
    masses = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
    gb = M.GCMCBatch(M.ForceFieldEAM.setfl(pot.NIALH_JEA), cells, 2, -2.3, 300.0, masses,
                     np.arange(4096, dtype=np.int32) + 1000)
    waves, _ = gb.tune(attempts)
    gb.use_graph(True)
    gb.run(attempts)
    gb.profile(True)
    t0 = time.perf_counter()
    for _ in range(launches):
        gb.run(attempts)
    dt = time.perf_counter() - t0
    ms_k, nk = gb.profile_read("gcmc")
    gb.profile(False)
    n, ng = gb.counts()
    total = 4096 * attempts * launches
    return {"workload": "4096 periodic Ni-H cells, 108-4000 atoms log-uniform, NiAlH_jea EAM, mu=-2.3 eV, 300 K",
            "metric": "GCMC trial-move energy deltas/s", "value": total / dt, "unit": "attempts/s",
            "cells_per_s": 4096 * launches / dt, "ms_per_launch": ms_k / max(nk, 1), "attempts_per_launch": attempts,
            "waves_per_class": [int(w) for w in waves], "atoms": int(n.sum()), "gas_atoms": int(ng.sum()),
            "dtype": "f64"}


def replica_run(nsteps=30):
    """Batched replica minimisation: the GCMC workload's 4096 periodic Ni-H
    cells (108-4000 atoms, log-uniform; Ni jittered by 0.1 A), each minimised
    by MinCG with LineSearchBackTrack for nsteps steps in one persistent
    launch, launch shape from the tuner, replayed from a HIP graph."""
    import mapp4py_amd as M
    from mapp4py_amd import potentials as pot

    cells = [(H, x, e) for x, H, e in pot.gcmc_nih_cells(4096, seed=6, jitter=0.1)]
    natms = np.array([len(x) for _, x, _ in cells])
    ff = M.ForceFieldEAM.setfl(pot.NIALH_JEA)
    rb = M.ReplicaBatch(ff, cells, skin=0.5)
    waves, tune_ms = rb.tune(nsteps, e_tol=0.0, max_dx=1.0)
    rb.use_graph(True)
    rb2 = M.ReplicaBatch(ff, cells, skin=0.5)  # a fresh copy of the batch for the timed run
    rb2.set_waves(waves)
    rb2.use_graph(True)
    rb2.profile(True)
    t0 = time.perf_counter()
    res = rb2.min_cg(nsteps, e_tol=0.0, max_dx=1.0)
    dt = time.perf_counter() - t0
    ms_k, nk = rb2.profile_read()
    rb2.profile(False)
    steps = res["steps"]
    return {"workload": f"4096 periodic Ni-H cells, 108-4000 atoms log-uniform (4.34 M atoms), NiAlH_jea EAM, "
                        f"MinCG + LineSearchBackTrack, {nsteps} steps each",
            "metric": "replica minimisation atom-steps/s", "value": float((natms * steps).sum()) / dt,
            "unit": "atom-steps/s", "seconds": dt, "kernel_ms": ms_k, "replicas_per_s": len(cells) / dt,
            "steps_total": int(steps.sum()), "evals_total": int(res["evals"].sum()),
            "status_counts": {int(k): int(v) for k, v in zip(*np.unique(res["status"], return_counts=True))},
            "mean_energy_drop_eV_per_atom": float(((res["e0"] - res["e"]) / natms).mean()),
            "waves_per_class": [int(w) for w in waves], "tuner_ms": [float(t) for t in tune_ms], "dtype": "f64"}


def replica_cpu_baseline(seconds):
    """the oracle's MinCG (back-tracking) over the oracle EAM, one cell after
    another, on the first cells of the same batch for about `seconds`"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import o_min
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    cells = pot.gcmc_nih_cells(4096, seed=6, jitter=0.1)
    ff = O.OEam.read(pot.NIALH_JEA)
    work, k, t0 = 0, 0, time.perf_counter()
    while k < len(cells) and (k < 2 or time.perf_counter() - t0 < seconds):
        x, H, e = cells[k]
        o = O.OMD(ff, H, x, e, skin=0.5)
        r = o_min.min_cg(o, 30, e_tol=0.0, max_dx=1.0)
        work += len(x) * r["steps"]
        k += 1
    dt = time.perf_counter() - t0
    return {"value": work / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle MinCG over the FP64 C restatement, the first {k} cells of the same batch, 30 steps "
                      f"each, 1 thread, {dt:.1f} s"}


def gcmc_cpu_baseline(seconds):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from mapp4py_amd import potentials as pot

    cells = pot.gcmc_nih_cells(4096, seed=6)
    masses = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
    ff = O.OEam.read(pot.NIALH_JEA)
    done, dt, atoms = 0, 0.0, 0
    for c, (x, H, e) in enumerate(cells):
        g = O.OGCMC(ff, H, x, e, 2, -2.3, 300.0, masses, seed=1000 + c)
        t0 = time.perf_counter()
        g.attempts(512)
        dt += time.perf_counter() - t0
        done += 512
        atoms += len(e)
        if dt > seconds:
            break
    return {"value": done / dt, "unit": "attempts/s", "cores": 1, "kind": "port",
            "sample": f"oracle SGCMC, first {c + 1} cells of the same batch ({atoms} atoms), 512 attempts each, "
                      f"attempts only, {dt:.1f} s"}


def load_profiled(name):
    """a profiled constant committed under profiles/ (JSON), or None"""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dmd", action="store_true", help="skip the DMD measurement")
    ap.add_argument("--no-gcmc", action="store_true", help="skip the batched GCMC measurement")
    ap.add_argument("--no-feni", action="store_true", help="skip the Fe-Ni (two-element) MD measurement")
    ap.add_argument("--no-min", action="store_true", help="skip the L-BFGS minimisation measurement")
    ap.add_argument("--no-nst", action="store_true", help="skip the NST measurement")
    ap.add_argument("--no-muvt", action="store_true", help="skip the muVT measurement")
    ap.add_argument("--no-replica", action="store_true", help="skip the batched replica minimisation")
    ap.add_argument("--dmd-evals-only", action="store_true",
                    help="only the DMD evaluations at 262,144 sites, FP64 and FP32 rows (for profiling)")
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
    if args.dmd_evals_only:
        print(json.dumps(dmd_evals_only()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    n, wall, kern, nvt, out, nfull, md, p, H, elem = gpu_run(args)
    value = n * args.steps / wall
    ms_f, c_f = kern["eam_force"]
    avg_s = ms_f / max(c_f, 1) / 1e3
    achieved = PASS2_BYTES_PER_ATOM * n / avg_s if avg_s > 0 else 0.0
    traffic = load_profiled("eam_force_traffic.json")
    issue = load_profiled("eam_force_issue.json")
    roof = {"kernel": "k_eam_force_lds (EAM pass 2 + second half-kick, spline rows in LDS)", "bound": "hbm",
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "traffic_source": (f"profiled constant: rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel, "
                               f"{traffic.get('measured', '?')} at commit {traffic.get('commit', '?')} "
                               f"(profiles/eam_force_traffic.json)") if traffic else None,
            "limiter": "VALU issue and gather latency, not HBM (the list and neighbor records are L2-served)",
            "valu_busy": issue.get("valu_busy") if issue else None,
            "valu_busy_source": (f"profiled constant: SQ_ACTIVE_INST_VALU / SQ_BUSY_CU_CYCLES, "
                                 f"{issue.get('measured', '?')} at commit {issue.get('commit', '?')} "
                                 f"(profiles/eam_force_issue.json)") if issue else None,
            "avg_launch_ms": avg_s * 1e3, "launches": c_f,
            "algorithmic_bytes_per_launch": PASS2_BYTES_PER_ATOM * n,
            "list_inclusive_bytes_per_launch": PASS2_BYTES_PER_ATOM * n + 4 * nfull,
            "whole_step": {"algorithmic_bytes_per_atom_step": STEP_BYTES_PER_ATOM,
                           "achieved_GBps": value * STEP_BYTES_PER_ATOM / 1e9,
                           "frac": value * STEP_BYTES_PER_ATOM / HBM_PEAK}}
    fp64 = None
    if not args.no_cpu:
        log("[bench] FP64 oracle check of the final state")
        fp64 = headline_error(md, p, H, elem, host_threads())
        log(f"[bench] max|df| {fp64['max_abs_df_eV_per_A']:.2e} eV/A, |dE|/N {fp64['abs_dE_per_atom_eV']:.1e} eV "
            f"({fp64['seconds']:.1f} s)")
    nst = None
    if not args.no_nst:
        log("[bench] NST")
        nst = nst_run(md, n)
        log(f"[bench] NST {nst['value']:.3e} atom-steps/s ({nst['ms_per_step']:.3f} ms/step, "
            f"{nst['rebuilds']} rebuilds, box moved {nst['box_change_A']:.2e} A)")
    del md
    if nst is not None and not args.no_cpu:
        nst["cpu_baseline"] = nst_cpu_baseline(args.cpu_seconds / 5)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline (~{args.cpu_seconds:.0f} s of oracle work)")
        cpu = cpu_baseline(args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.3e} atom-steps/s")
    dmd = gcmc = feni = mini = None
    if not args.no_min:
        log("[bench] L-BFGS minimisation")
        mini = min_run()
        if not args.no_cpu:
            mini["cpu_baseline"] = min_cpu_baseline(args.cpu_seconds / 3)
        log(f"[bench] L-BFGS {mini['value']:.3e} atom-steps/s ({mini['ms_per_step']:.3f} ms/step, "
            f"{mini['energy_evals']} energy evals, {mini['rebuilds']} rebuilds)")
    muvt = None
    if not args.no_muvt:
        log("[bench] muVT")
        muvt = muvt_run()
        if not args.no_cpu:
            muvt["cpu_baseline"] = muvt_cpu_baseline()
        log(f"[bench] muVT {muvt['value']:.3e} atom-steps/s, {muvt['attempts_per_s']:.3e} attempts/s, "
            f"counts {muvt['counts_ins_del_acc_ins_del_tried']}")
    if not args.no_feni:
        log("[bench] Fe-Ni")
        feni = feni_run()
        log(f"[bench] Fe-Ni {feni['value']:.3e} atom-steps/s ({feni['ms_per_step']:.3f} ms/step, "
            f"{feni['rebuilds']} rebuilds, kernels {feni['kernel_ms']})")
    if not args.no_dmd:
        log("[bench] DMD")
        dmd = dmd_run()
        dmd["example_replay"] = example_replay(0.0 if args.no_cpu else args.cpu_seconds / 5)
        log(f"[bench] examples/diff replay: FE {dmd['example_replay']['fe']:.6f} "
            f"(|err| {dmd['example_replay']['abs_err_fe']:.1e}), S err {dmd['example_replay']['abs_err_S']:.1e}")
        if not args.no_cpu:
            dmd["cpu_baseline"] = dmd_cpu_baseline(args.cpu_seconds / 3)
            dmd["c_d"]["cpu_baseline"] = dmd_cd_cpu_baseline(args.cpu_seconds / 5)
            dmd["J"]["cpu_baseline"] = dmd_j_cpu_baseline(args.cpu_seconds / 5)
            dmd["implicit_step"]["cpu_baseline"] = dmd_newton_cpu_baseline(args.cpu_seconds / 5)
            dmd["bdf"]["cpu_baseline"] = dmd_bdf_cpu_baseline()
        log(f"[bench] DMD {dmd['value']:.3e} site-evals/s")
    replica = None
    if not args.no_replica:
        log("[bench] replica minimisation")
        replica = replica_run()
        if not args.no_cpu:
            replica["cpu_baseline"] = replica_cpu_baseline(args.cpu_seconds / 3)
        log(f"[bench] replicas {replica['value']:.3e} atom-steps/s ({replica['seconds']:.3f} s, "
            f"status {replica['status_counts']}, waves {replica['waves_per_class']})")
    if not args.no_gcmc:
        log("[bench] GCMC")
        gcmc = gcmc_run()
        if not args.no_cpu:
            gcmc["cpu_baseline"] = gcmc_cpu_baseline(args.cpu_seconds / 3)
        log(f"[bench] GCMC {gcmc['value']:.3e} attempts/s")
    line = {
        "metric": "EAM atom-steps/s @8M atoms; DMD sites/s surrogate vs quadrature @1e-4 eV/A", "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (Johnson-form Cu EAM table, jittered fcc lattice, Maxwell velocities)",
        "config": {"workload": f"EAM NVT, fcc Cu {args.cells}^3 cells ({n} atoms), Johnson setfl Nr=Nrho=10000, "
                               f"r_c=4.95 A, skin 0.5 A, dt 1 fs, 300 K", "natoms": n, "rebuilds": nvt.rebuilds,
                   "pe_per_atom_eV": out[0] / n, "T_K": nvt.T_part},
        "error_vs_fp64": fp64,
        "roofline": roof, "cpu_baseline": cpu, "feni": feni, "nst": nst, "muvt": muvt, "minimize": mini, "dmd": dmd, "gcmc": gcmc, "replica": replica,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark of the MI355X force engine on the 1M-particle isolated disk galaxy.

A step is one iteration of the reference's Simulation::run on the device:
block time steps, kick + drift, tree build, SPH group density, gravity walk
with embedded SPH force, U update and second kick.  Inputs are resident in
HBM before timing starts.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision fp64|fp32]
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "astrogenesis2.0_amd")
METRIC = "particle-steps/s FP32 vs FP64 at ≤1e-3 p99 accel err; p50/p99 step latency"
FLOPS_PER_INTERACTION = 19          # 3 sub, 5 r^2, 1 add, rsqrt, rcp, 2 mul, 3 fma
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
CPU_THREADS = 16                    # the GPU box's host-core share


def load_pkg():
    spec = importlib.util.spec_from_file_location("ag2_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ag2_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(ag2, n_total, kind="disk"):
    """Disk galaxy: 50 % NFW halo, 30 % stellar disk, 20 % gas disk (or a
    Plummer sphere of the CPU-runnable configuration's recipe)."""
    if kind == "plummer":
        U = ag2.units
        ps = ag2.ics.config1_plummer(n_total, seed=1)
        return ps, dict(theta=0.5, e0=0.05 * U.KPC, massInH=1.0, split_threshold=800, dt=0.1 * U.MYR), (n_total, 0, 0)
    nh, nd = n_total // 2, (3 * n_total) // 10
    ng = n_total - nh - nd
    ps = ag2.ics.disk_galaxy(nh, nd, ng, seed=2, config=2)
    U = ag2.units
    prm = dict(theta=0.5, e0=0.1 * U.KPC, massInH=ag2.ics.mass_in_h(ps), split_threshold=800,
               dt=0.5 * U.MYR)
    return ps, prm, (nh, nd, ng)


def cpu_child(args):
    """Time the compiled reference (or the oracle) on the same workload."""
    ag2 = load_pkg()
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as O
    ps, prm, _ = workload(ag2, args.n)
    if O.ref_available():
        import ctypes as C
        L = O.ref()
        d = C.POINTER(C.c_double)
        L.ref_time_step.argtypes = [C.c_int64] + [d] * 9 + [C.POINTER(C.c_uint8), C.c_double, C.c_double,
                                                             C.c_double, d]
        sec = np.zeros(4)
        arrs = [getattr(ps, f).ctypes.data_as(d) for f in ("x", "y", "z", "vx", "vy", "vz", "m", "U", "mu")]
        L.ref_time_step(ps.n, *arrs, ps.type.ctypes.data_as(C.POINTER(C.c_uint8)), prm["theta"], prm["e0"],
                        prm["massInH"], sec.ctypes.data_as(d))
        kind, total = "reference", float(sec.sum())
        phases = dict(build_tree=sec[0], sph_density=sec[1], forces=sec[2], delete_tree=sec[3])
    else:
        a = ps.copy()
        t0 = time.perf_counter()
        O.force_step(a, O.params(prm["theta"], prm["e0"], prm["massInH"], prm["split_threshold"],
                                 nthreads=CPU_THREADS))
        kind, total = "port", time.perf_counter() - t0
        phases = {}
    print(json.dumps({"kind": kind, "seconds": total, "phases": {k: float(v) for k, v in phases.items()}}))


def cpu_baseline(args):
    env = dict(os.environ, OMP_NUM_THREADS=str(CPU_THREADS), OMP_THREAD_LIMIT=str(CPU_THREADS),
               OMP_PROC_BIND="close")
    cmd = [sys.executable, os.path.abspath(__file__), "--_cpu_child", "--n", str(args.n)]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        return {"value": None, "error": r.stderr[-400:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": args.n / res["seconds"], "unit": "particle-steps/s", "cores": CPU_THREADS,
            "kind": res["kind"],
            "sample": "one force evaluation (build tree + SPH density + forces + delete tree, the reference's "
                      "per-step tree phases) of the full %d-particle workload, OpenMP on %d host threads"
                      % (args.n, CPU_THREADS),
            "phases_s": res["phases"]}


def run_gpu(args, ag2, ps, prm, precision):
    E = ag2.engine
    p = E.make_params(theta=prm["theta"], e0=prm["e0"], massInH=prm["massInH"],
                      split_threshold=prm["split_threshold"], precision=precision, eta=2.0,
                      min_dt=prm["dt"], max_dt=prm["dt"], H0=0.0)
    eng = E.Engine(p, ps.n)
    eng.upload(ps)
    eng.init_forces()
    for _ in range(args.warmup):
        eng.step()
    eng.sync()
    lat, walk_ms, inter, phases = [], [], [], np.zeros(6)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        eng.step()
        eng.sync()
        lat.append(time.perf_counter() - ts)
        ph = eng.phase_times()
        phases += ph
        walk_ms.append(ph[4])
        inter.append(eng.diagnostics()["interactions"])
    elapsed = time.perf_counter() - t0
    d = eng.diagnostics()
    eng.close()
    print("walk: entries %d tests %d interactions %d sph %d -> mask density %.3f" % (
        d["walk_entries"], d["walk_tests"], d["interactions"], d["sph_pairs"],
        d["interactions"] / max(1, 64 * d["walk_entries"])), file=sys.stderr)
    return dict(elapsed=elapsed, lat=np.array(lat), walk_ms=np.array(walk_ms), inter=np.array(inter, float),
                phases=phases / args.steps, diag=d)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--workload", choices=["disk", "plummer"], default="disk")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--_cpu_child", action="store_true")
    args = ap.parse_args()
    if args._cpu_child:
        return cpu_child(args)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (a single simulation is not sharded "
              "and serving jobs are placed per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    ag2 = load_pkg()
    ps, prm, split = workload(ag2, args.n, args.workload)
    prec = 0 if args.precision == "fp64" else 1
    r = run_gpu(args, ag2, ps, prm, prec)
    n = ps.n
    value = n * args.steps / r["elapsed"]
    flops = r["inter"] * FLOPS_PER_INTERACTION
    achieved = float(flops.sum() / (r["walk_ms"].sum() * 1e-3) / 1e12)
    peak = PEAK_TFLOPS[args.precision]
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * r["elapsed"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == 0 else "f32",
        "data": "synthetic (seeded disk-galaxy initial conditions)",
        "config": {"workload": "isolated disk galaxy N=%d (%d NFW halo, %d stellar disk, %d SPH gas), theta=0.5, "
                               "e0=0.1 kpc, dt=0.5 Myr KDK, massInH=64 m_gas, split_threshold=800" % ((n,) + split),
                   "precision": args.precision},
        "p50_step_ms": float(np.percentile(r["lat"], 50) * 1e3),
        "p99_step_ms": float(np.percentile(r["lat"], 99) * 1e3),
        "interactions_per_s": float(r["inter"].sum() / r["elapsed"]),
        "interactions_per_step": float(r["inter"].mean()),
        "phases_ms": dict(zip(["keys", "sort", "topology_moments", "sph_density", "walk", "kdk"],
                              [float(v) for v in r["phases"]])),
        "roofline": {"bound": "valu", "kernel": "k_walk64" if prec == 0 else "k_walk32", "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "work": "%d flops per gravity interaction x interactions per launch" % FLOPS_PER_INTERACTION},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

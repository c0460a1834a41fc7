#!/usr/bin/env python3
"""Headline benchmark: TGV 256^3 FP64 RK3 steps on one MI355X.

Workload (one GPU): the reference's Taylor-Green vortex miniapp
(miniapps/tgv_solver_3d/tgvsolver.F90) at 256^3, Re=1600, Ma=0.1, scheme
643c + 10th-order compact filter (alpha 0.49), RK3, dt=5e-4, velocity
perturbed by seeded 1e-3 Gaussian noise.  A step = 3 RK stages over the
257^3 computed lattice with all fields resident in HBM.

Prints one JSON line: value = Mpt-updates/s (lattice points x RK stages per
second), roofline of the dominant kernel from HIP events recorded around
every pass launch inside the timed region, and the reference Fortran timed
on this host's cores as cpu_baseline.
"""
import argparse
import json

import numpy as np
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TGV 256³ FP64 Mpt-updates/s + % HBM peak; batched/served cases/s, p50/p99"
HBM_PEAK_GBS = 8000.0

# Algorithmic FP64 words per lattice point moved by the reference operators
# each pass fuses (one streaming sweep per operator, reads + writes):
#   gc_first  : gradients x (4+4) + convective x (7+5)                 = 20
#   gc_mid    : gradients y (4+4) + convective y (12+5)                = 25
#   gc_last   : gradients z (4+4) + convective z (12+5)
#               + pointwise stress/heat flux (16+9)                    = 50
#   visc_mid  : viscous derivative (4+4 read, 4 write)                 = 12
#   visc_last_rk3_filter: viscous z (12) + RK3 update (15+5)
#               + filter z (5+5)                                       = 42
#   filter    : one filter sweep (5+5)                                 = 10
# Summed over a 3-D stage: 20+25+50+12+12+42+10+10 = 181, plus the
# reference's updatefvar (5+6 = 11) that the passes make unnecessary:
# 192 words = 1536 B per point per stage.
PASS_WORDS = {"gc_first": 20, "gc_mid": 25, "gc_last": 50, "visc_mid": 12,
              "visc_last_rhs": 12, "visc_last_rk3_filter": 42, "filter": 10}
STAGE_BYTES = 1536


def cpu_baseline(seconds_target=12.0):
    """The reference's own Fortran (oracle/_ref, single thread) on a bounded
    64^3 sample; falls back to the C restatement if that build is absent."""
    from oracle import oracle as O

    n = 64
    pts = (n + 1) ** 3
    if O.ref_available():
        try:
            steps = 2
            r = O.run_ref("rk3", dims=[n, n, n], dt=5e-4, nsteps=steps)
            per = float(r["seconds"]) / steps
            steps = max(2, int(seconds_target / max(per, 1e-3)))
            r = O.run_ref("rk3", dims=[n, n, n], dt=5e-4, nsteps=steps)
            sec = float(r["seconds"])
            return {"value": pts * 3 * steps / sec / 1e6, "unit": "Mpt-updates/s", "cores": 1,
                    "kind": "reference",
                    "sample": f"reference tgvsolver.F90 rk3 (flang -O3, 1 thread), TGV {n}^3, "
                              f"{steps} steps, {sec:.1f} s"}
        except Exception as e:  # pragma: no cover - reported, not fatal
            sys.stderr.write(f"reference baseline failed: {e}\n")
    s = O.OracleState(O.tgv_cfg(n, dt=5e-4))
    s.tgv_init(1e-3, 1)
    t0 = time.perf_counter()
    s.rk3(1)
    per = time.perf_counter() - t0
    steps = max(1, int(seconds_target / max(per, 1e-3)))
    t0 = time.perf_counter()
    s.rk3(steps)
    sec = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": pts * 3 * steps / sec / 1e6, "unit": "Mpt-updates/s", "cores": cores,
            "kind": "port", "sample": f"C restatement, TGV {n}^3, {steps} steps, {sec:.1f} s"}


def chem_batch(ncells=1 << 20, reps=20):
    """Chemistry source-term batch: 2^20 H2-air cells (9 species, 19
    reactions), inputs resident in HBM, FP64.  Algorithmic bytes per cell:
    T, p, Y(9) in and wdot(9), h, cp out = 22 words = 176 B."""
    import astr_amd.chem as AC
    from astr_amd import hip

    mech = AC.Mechanism(os.path.join(ROOT, "tests", "golden", "Jachimowski88_H2-9-19.xml"))
    T, p, Y = AC.sample_cells(mech.names, mech.W, ncells, 2026)
    dT, dp, dY = (hip.DeviceArray.from_host(a) for a in (T, p, Y))
    dW = hip.DeviceArray((mech.ns, ncells))
    dH = hip.DeviceArray((ncells,))
    dC = hip.DeviceArray((ncells,))
    L = AC._bind()
    call = lambda: AC._chk(L.astr_hip_chem_rates_device(  # noqa: E731
        mech.h, ncells, dT.ptr, dp.ptr, 1, dY.ptr, dW.ptr, dH.ptr, dC.ptr, None, None))
    for _ in range(3):
        call()
    hip.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    ms = e0.elapsed_ms(e1) / reps
    wd = dW.to_host()
    mass = float(np.abs(wd.sum(axis=0)).max() / np.abs(wd).max())
    evals = ncells / (ms * 1e-3)
    gbs = 176.0 * ncells / (ms * 1e-3) / 1e9
    return {"cells": ncells, "ms_per_batch": round(ms, 4), "evals_per_s": round(evals, 1),
            "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "alg_bytes_per_cell": 176},
            "mass_conservation_relerr": mass, "dtype": "f64",
            "mechanism": "Jachimowski88_H2-9-19 (9 species, 19 reactions)"}


def copy_probe(nbytes, reps=3):
    """A plain device-to-device copy of nbytes, reps times: the calibration
    kernel for the FETCH_SIZE / WRITE_SIZE counters (it reads and writes
    exactly nbytes per launch)."""
    from astr_amd import hip

    a = hip.DeviceArray((nbytes // 8,))
    b = hip.DeviceArray((nbytes // 8,))
    hip.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record()
    for _ in range(reps):
        hip.copy_device(b, a, nbytes)
    e1.record()
    ms = e0.elapsed_ms(e1) / reps
    return {"bytes": nbytes, "launches": reps, "ms_per_copy": round(ms, 4),
            "GBs_read_plus_write": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def load_traffic(n):
    path = os.path.join(ROOT, "profiles", f"traffic_tgv{n}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-chem", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch selection instead of timing the candidates")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also time a device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the whole {args.n}^3 domain sits in one MI355X's HBM; "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import astr_amd as A

    n = args.n
    dt = 5e-4 * 256.0 / n
    f = A.Flow(A.tgv_config(n, deltat=dt))
    f.init_tgv(1e-3, 2024)
    tune_s = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        f.autotune(reps=3)
        tune_s = time.perf_counter() - t0
    launch = {f"{k[0]}_{k[1]}": {"shape": A.SHAPES[v[0]], "minb": v[1]}
              for k, v in f.pass_configs().items()}
    if args.warmup:
        f.rk3(args.warmup, stats=True)
    f.sync()

    launches_per_step = 3 * 8
    f.profile_begin(launches_per_step * args.steps + 8)
    f.sync()
    t0 = time.perf_counter()
    ke, en = f.rk3(args.steps, stats=True)
    f.sync()
    wall = time.perf_counter() - t0
    prof = f.profile_end()

    if not all(math.isfinite(x) for x in list(ke) + list(en)):
        raise SystemExit("non-finite kinetic energy/enstrophy: run is broken")

    pts = (n + 1) ** 3
    ms_step = wall / args.steps * 1e3
    value = pts * 3 * args.steps / wall / 1e6

    # dominant kernel by total time
    (pname, pdir), (ms_tot, cnt) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_s = ms_tot / cnt / 1e3
    alg_bytes = PASS_WORDS[pname] * 8 * pts
    achieved = alg_bytes / avg_s / 1e9
    tr = load_traffic(n)
    traffic = None
    if tr and f"{pname}_{pdir}" in tr.get("kernels", {}):
        traffic = tr["kernels"][f"{pname}_{pdir}"]["bytes_per_launch"]
    step_alg = STAGE_BYTES * pts * 3 * args.steps / wall / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpt-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (TGV closed form + seeded 1e-3 Gaussian velocity noise)",
        "config": {"workload": f"TGV {n}^3 RK3 step, compact 643c derivative + 10th-order "
                               f"filter, Re=1600, Ma=0.1, dt={dt:g}",
                   "grid": [n, n, n], "lattice_points": pts, "parallelism": "1 GPU, whole domain"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": f"{pname} ({pdir})",
                     "avg_launch_ms": round(avg_s * 1e3, 4),
                     "alg_bytes_per_launch": alg_bytes},
        "step_roofline": {"alg_bytes_per_point_stage": STAGE_BYTES,
                          "achieved_GBs": round(step_alg, 1),
                          "frac": round(step_alg / HBM_PEAK_GBS, 4)},
        "physical_point_rate": round(n ** 3 * 3 * args.steps / wall / 1e6, 3),
        "kernels_ms_per_launch": {f"{k[0]}_{k[1]}": round(v[0] / v[1], 4) for k, v in prof.items()},
        "ke_last": float(ke[-1]),
        "launch": {"autotune_s": None if tune_s is None else round(tune_s, 2),
                   "shape_is": "(lines per workgroup, chunks per line, unknowns per chunk)",
                   "passes": launch},
    }
    if not args.no_chem:
        out["chem_batch"] = chem_batch()
    if args.copy_probe:
        out["copy_probe"] = copy_probe(args.copy_probe)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

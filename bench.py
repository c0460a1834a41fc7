#!/usr/bin/env python3
"""Headline benchmark: TGV 256^3 FP64 RK3 steps on one MI355X.

Workload (one GPU): the reference's Taylor-Green vortex miniapp
(miniapps/tgv_solver_3d/tgvsolver.F90) at 256^3, Re=1600, Ma=0.1, scheme
643c + 10th-order compact filter (alpha 0.49), RK3, dt=5e-4, velocity
perturbed by seeded 1e-3 Gaussian noise.  A step = 3 RK stages over the
257^3 computed lattice with all fields resident in HBM.

Prints one JSON line: value = Mpt-updates/s (lattice points x RK stages per
second) over K steps with nothing else on the stream, roofline of the
dominant kernel from HIP events recorded around every pass launch in the K
steps that follow (the events would add about 1% to the step time), and the
reference Fortran timed on this host's cores as cpu_baseline.
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
FP64_PEAK_TFS = 78.6

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
#   filter_gc_first: gc_first with the previous stage's x filter sweep = 30
# Summed over a 3-D stage: 20+25+50+12+12+42+10+10 = 181, plus the
# reference's updatefvar (5+6 = 11) that the passes make unnecessary:
# 192 words = 1536 B per point per stage.
# When the last gradient/convective pass runs split, its halves carry the
# gradients + stress/heat flux (4+4 + 16+9 = 33) and the convective
# derivatives (12+5 = 17) of the same 50 words.
PASS_WORDS = {"gc_first": 20, "gc_mid": 25, "gc_last": 50, "visc_mid": 12,
              "visc_last_rhs": 12, "visc_last_rk3_filter": 42, "filter": 10,
              "gc_last_grad": 33, "gc_last_conv": 17, "filter_gc_first": 30}
STAGE_BYTES = 1536


def _host_cpus():
    """The host's CPU counts: all CPUs, those in this process's affinity mask,
    and the share OMP_NUM_THREADS states (None when unset)."""
    share = os.environ.get("OMP_NUM_THREADS")
    return {"cpu_count": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "cpu_share": int(share) if share and share.isdigit() else None}


def _baseline_cores():
    """Processes for the CPU baseline: every CPU of the affinity mask, at most
    the stated CPU share (ASTR_BASELINE_CORES overrides both)."""
    env = os.environ.get("ASTR_BASELINE_CORES")
    if env:
        return max(1, int(env))
    h = _host_cpus()
    n = h["affinity_cpus"]
    if h["cpu_share"]:
        n = min(n, h["cpu_share"])
    return max(1, n)


def cpu_baseline(seconds_target=12.0):
    """The reference's own Fortran (oracle/_ref) on a bounded 64^3 sample:
    one serial run, then one run per usable host core at once (the MPI build
    of the main path cannot be compiled here, so independent per-core runs of
    the serial miniapp stand in for its ranks; with no halo exchange this
    bounds an MPI run from above).  The process count is the size of this
    process's CPU affinity mask, capped by the machine's CPU share for one
    GPU when OMP_NUM_THREADS states it; both counts go into the line.  Falls
    back to the C restatement if the reference build is absent."""
    from concurrent.futures import ThreadPoolExecutor

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
            sec1 = float(r["seconds"])
            one = pts * 3 * steps / sec1 / 1e6
            cores = _baseline_cores()
            with ThreadPoolExecutor(cores) as ex:
                runs = list(ex.map(lambda _: O.run_ref("rk3", dims=[n, n, n], dt=5e-4, nsteps=steps),
                                   range(cores)))
            secs = [float(x["seconds"]) for x in runs]
            return {"value": cores * pts * 3 * steps / max(secs) / 1e6, "unit": "Mpt-updates/s",
                    "cores": cores, "kind": "reference", **_host_cpus(),
                    "single_core_value": one,
                    "sample": f"reference tgvsolver.F90 rk3 (flang -O3), TGV {n}^3, {steps} steps: "
                              f"{cores} serial processes at once, one per core, slowest "
                              f"{max(secs):.1f} s (one process alone: {sec1:.1f} s); value = "
                              f"all processes' lattice updates / slowest time"}
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
    cores = _baseline_cores()
    return {"value": pts * 3 * steps / sec / 1e6, "unit": "Mpt-updates/s", "cores": cores,
            "kind": "port", **_host_cpus(),
            "sample": f"C restatement (OpenMP), TGV {n}^3, {steps} steps, {sec:.1f} s"}


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
    valu = None
    cnt_path = os.path.join(ROOT, "profiles", "chem_valu_counters.json")
    if os.path.exists(cnt_path):
        # FP64 operations per cell from the kernel's own instruction counters
        # (rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 of this batch), times
        # the cells per second measured here
        with open(cnt_path) as fh:
            fpc = json.load(fh)["flops_per_cell"]
        tfs = fpc * evals / 1e12
        valu = {"flops_per_cell": fpc, "achieved": round(tfs, 2), "peak": FP64_PEAK_TFS,
                "unit": "TFLOP/s", "frac": round(tfs / FP64_PEAK_TFS, 4),
                "source": "profiles/chem_valu_counters.json"}
    return {"cells": ncells, "ms_per_batch": round(ms, 4), "evals_per_s": round(evals, 1),
            "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "alg_bytes_per_cell": 176},
            "valu": valu,
            "mass_conservation_relerr": mass, "dtype": "f64",
            "mechanism": "Jachimowski88_H2-9-19 (9 species, 19 reactions)"}


def copy_probe(nbytes, reps=5):
    """The achievable HBM rate of a plain FP64 copy of nbytes: the library's
    streaming probe (astr_hip_stream_probe, 1, 2 and 4 doubles per lane per
    access, nontemporal) and the runtime's own device-to-device copy, each
    timed over reps launches after one untimed launch.  The probe reads and
    writes exactly nbytes per launch, so it also calibrates the FETCH_SIZE /
    WRITE_SIZE counters."""
    import ctypes as C

    import astr_amd as A
    from astr_amd import hip

    L = A.lib()
    L.astr_hip_stream_probe.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p]
    L.astr_hip_probe_last_error.restype = C.c_char_p
    n = nbytes // 16 * 2
    a = hip.DeviceArray((n,))
    b = hip.DeviceArray((n,))
    hip.synchronize()
    out = {"bytes": 8 * n, "launches": reps}

    def timed(launch):
        launch()
        hip.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        e0.record()
        for _ in range(reps):
            launch()
        e1.record()
        ms = e0.elapsed_ms(e1) / reps
        return {"ms": round(ms, 4), "GBs_read_plus_write": round(2 * 8 * n / (ms * 1e-3) / 1e9, 1)}

    def probe(w):
        def go():
            if L.astr_hip_stream_probe(a.ptr, b.ptr, n, w, None) != 0:
                raise RuntimeError(L.astr_hip_probe_last_error().decode())
        return go

    for w in (1, 2, 4):
        out[f"stream_probe_w{w}"] = timed(probe(w))
    out["hipMemcpy"] = timed(lambda: hip.copy_device(b, a, 8 * n))
    best = max(v["GBs_read_plus_write"] for k, v in out.items() if isinstance(v, dict))
    out["best_GBs"] = best
    out["best_frac_of_peak"] = round(best / HBM_PEAK_GBS, 4)
    return out


def batch_form(ncases=64, steps=100, seed=2026):
    """Batching form: 64 independent periodic cases drawn from a seed, half
    2-D isentropic vortex (grid side U{64..512}, Ma 0.5, strength 5), half 3-D
    TGV (side U{32..128}, Re U[100,3000]), 100 RK3 steps each, FP64.  Batched:
    one batch (autotuned launch selection, one hipGraph per step).  Sequential:
    each case alone through astr_hip_rk3_steps (own autotuned selection).
    Set-up and tuning are outside both timings."""
    import astr_amd as A

    rng = np.random.default_rng(seed)
    specs = []
    for k in range(ncases):
        if k % 2 == 0:
            n = int(rng.integers(64, 513))
            specs.append(("vortex", A.vortex_config(n), None))
        else:
            n = int(rng.integers(32, 129))
            re = float(rng.uniform(100.0, 3000.0))
            specs.append(("tgv", A.tgv_config(n, deltat=0.128 / n, reynolds=re), int(rng.integers(1 << 30))))

    def make():
        flows = []
        for kind, cfg, sd in specs:
            f = A.Flow(cfg)
            if kind == "vortex":
                f.init_vortex(5.0)
            else:
                f.init_tgv(1e-3, sd)
            flows.append(f)
        return flows

    flows = make()
    b = A.Batch(flows)
    b.autotune(reps=2)
    b.rk3(1, graph=True)  # warm-up step (captures the graph)
    for f, (kind, cfg, sd) in zip(flows, specs):  # same start for both paths
        if kind == "vortex":
            f.init_vortex(5.0)
        else:
            f.init_tgv(1e-3, sd)
    A.lib().astr_hip_sync(flows[0].h)
    t0 = time.perf_counter()
    b.rk3(steps, graph=True, stats=False)
    t_batch = time.perf_counter() - t0
    info = b.info()
    q_batch = [f.download() for f in flows]
    b.close()
    for f in flows:
        f.close()

    seq = make()
    for f in seq:
        f.autotune(reps=2)
        f.rk3(1, stats=False)
    for f, (kind, cfg, sd) in zip(seq, specs):
        if kind == "vortex":
            f.init_vortex(5.0)
        else:
            f.init_tgv(1e-3, sd)
        f.sync()
    t0 = time.perf_counter()
    for f in seq:
        f.rk3(steps, stats=False)
        f.sync()
    t_seq = time.perf_counter() - t0
    diff = 0.0
    for f, qb in zip(seq, q_batch):
        qs = f.download()
        diff = max(diff, float(np.abs(qb - qs).max() / np.abs(qs).max()))
        f.close()
    pts = sum((c.im + 1) * (c.jm + 1) * (c.km + 1) for _, c, _ in specs)
    return {"cases": ncases, "steps_per_case": steps,
            "workload": "32 x 2-D isentropic vortex (side U{64..512}) + 32 x 3-D TGV (side U{32..128}, "
                        "Re U[100,3000]), FP64 RK3, compact 643c + filter",
            "lattice_points": pts,
            "batched": {"s": round(t_batch, 4), "cases_per_s": round(ncases / t_batch, 3),
                        "Mpt_updates_per_s": round(pts * 3 * steps / t_batch / 1e6, 1),
                        "launch_groups": info["groups"], "launches_per_step": info["launches_per_step"],
                        "graph": True},
            "sequential": {"s": round(t_seq, 4), "cases_per_s": round(ncases / t_seq, 3)},
            "speedup": round(t_seq / t_batch, 3),
            "max_rel_diff_batched_vs_sequential": diff}


def serve_form(rates=(10.0, 100.0, 1000.0), counts=(40, 150, 300), seed=7, rtol=1e-5, atol=1e-10):
    """Serving form: Poisson request streams, 70% 0-D constant-pressure H2-air
    ignition to 1 ms (RODAS4 at rtol/atol), 30% 2-D vortex (side U{64..256},
    200 steps), under a 200 GB HBM budget.  Latency = completion - arrival.
    The sequential reference serves a sample one request at a time through
    the plain path; its cases/s is 1 / mean service time."""
    from astr_amd import serve as S

    mech = os.path.join(ROOT, "tests", "golden", "Jachimowski88_H2-9-19.xml")
    out = {"workload": "70% 0-D ignition (T0 U[950,1400] K, p U[1,10] atm, phi U[0.5,2], t_end 1 ms), "
                       "30% 2-D vortex (side U{64..256}, 200 steps); HBM budget 200 GB",
           "reactor_tolerance": {"rtol": rtol, "atol": atol}, "rates": []}
    for rate, count in zip(rates, counts):
        reqs = S.poisson_requests(count, rate, seed + int(rate))
        srv = S.Server(mech, hbm_budget=200e9, reactor_slots=4096, rtol=rtol, atol=atol,
                       steps_per_tick=10, reactor_tries=100)
        res = srv.run(reqs)
        srv.close()
        st = S.latency_stats(res)
        ign = [r.latency for r in res if r.kind == "ignition"]
        vor = [r.latency for r in res if r.kind == "vortex"]
        isteps = [r.data["steps"] for r in res if r.kind == "ignition"]
        out["rates"].append({"rate_req_per_s": rate, "requests": count,
                             "cases_per_s": round(st["cases_per_s"], 3),
                             "p50_s": round(st["p50_s"], 4), "p99_s": round(st["p99_s"], 4),
                             "p50_ignition_s": round(float(np.percentile(ign, 50)), 4) if ign else None,
                             "p50_vortex_s": round(float(np.percentile(vor, 50)), 4) if vor else None,
                             "ignition_steps_mean": round(float(np.mean(isteps)), 1) if isteps else None,
                             "ignition_steps_max": int(max(isteps)) if isteps else None,
                             "max_in_flight": srv.max_active})
    # sequential plain path on a sample; equality of results with the served runs
    sample = S.poisson_requests(6, 1.0, seed + 99)
    t_serv = []
    srv = S.Server(mech, reactor_slots=64, rtol=rtol, atol=atol, steps_per_tick=10, reactor_tries=100)
    for r in sample:
        r.arrival = 0.0
    served = {r.rid: r for r in srv.run(sample)}
    srv.close()
    same = True
    for r in sample:
        t0 = time.perf_counter()
        alone = S.run_alone(r, mech, rtol=rtol, atol=atol, reactor_tries=100)
        t_serv.append(time.perf_counter() - t0)
        got = served[r.rid].data
        if r.__class__.__name__ == "Ignition":
            same = same and got["T"] == alone["T"] and bool(np.array_equal(got["Y"], alone["Y"]))
        else:
            same = same and float(np.abs(got["q"] - alone["q"]).max() / np.abs(alone["q"]).max()) < 1e-12
    mean = float(np.mean(t_serv))
    out["sequential"] = {"sample": len(sample), "mean_service_s": round(mean, 4),
                         "cases_per_s": round(1.0 / mean, 3)}
    best = max(x["cases_per_s"] for x in out["rates"])
    out["speedup_at_best_rate"] = round(best * mean, 3)
    out["served_equals_alone"] = bool(same)
    return out


def timed_steps(f, steps, launches_per_step):
    """K RK3 steps timed with nothing else on the stream (wall time, the
    KE/enstrophy history), then K more steps with a HIP event pair around
    every pass launch for the per-kernel times (and their own wall time).
    The event records cost about 1% of a step (0.15 ms at 256^3), so they
    stay out of the step timing."""
    f.sync()
    t0 = time.perf_counter()
    ke, en = f.rk3(steps, stats=True)
    f.sync()
    wall = time.perf_counter() - t0
    f.profile_begin(launches_per_step * steps + 8)
    f.sync()
    t1 = time.perf_counter()
    f.rk3(steps, stats=True)
    f.sync()
    wall_ev = time.perf_counter() - t1
    return wall, wall_ev, ke, en, f.profile_end()


def main_path_step(n, dt, steps, warmup):
    """The main solver's stage (ASTR_SEM_MAIN: filter at the stage start,
    qswap, stored primitives, metric-weighted fluxes, J-weighted update) on
    the same TGV workload and field: ms per step, Mpt-updates/s and the
    algorithmic fraction of HBM peak (the same 1536 B per point per stage of
    reference operators), with the per-pass times of the timed steps."""
    import astr_amd as A

    f = A.Flow(A.main_config(n, deltat=dt))
    f.init_tgv(1e-3, 2024)
    f.autotune(reps=10)
    if warmup:
        f.rk3(warmup, stats=True)
    wall, wall_ev, ke, en, prof = timed_steps(f, steps, 3 * 9)
    pts = (n + 1) ** 3
    step_alg = STAGE_BYTES * pts * 3 * steps / wall / 1e9
    out = {"workload": f"TGV {n}^3 RK3 step in the main solver's stage order (astr run), Sutherland 110.3 K",
           "ms_per_step": round(wall / steps * 1e3, 4),
           "ms_per_step_instrumented": round(wall_ev / steps * 1e3, 4),
           "value": round(pts * 3 * steps / wall / 1e6, 3), "unit": "Mpt-updates/s",
           "step_roofline": {"achieved_GBs": round(step_alg, 1), "frac": round(step_alg / HBM_PEAK_GBS, 4)},
           "kernels_ms_per_launch": {f"{k[0]}_{k[1]}": round(v[0] / v[1], 4) for k, v in prof.items()},
           "launch": {f"{k[0]}_{k[1]}": {"shape": A.SHAPES[v[0]], "minb": v[1]}
                      for k, v in f.pass_configs().items()},
           "ke_last": float(ke[-1])}
    f.close()
    return out


def upwind_path_step(n, dt, steps, warmup, recon, chardec, conschm="753e"):
    """The main solver's stage with the upwind-biased convection, explicit
    (conschm 753e, recon_schem `recon`; convrsduwd, src/solver.F90:548-1201)
    or compact (conschm 543c; convrsdcmp, src/solver.F90:1271-1937), with
    lchardecomp `chardec`, bfacmpld 0.3, shkcrt 0.05, on the same TGV
    workload and field: ms per step, Mpt-updates/s and per-kernel times of
    the timed steps; for the explicit interface-flux kernel, its HBM rate
    from the algorithmic bytes per point of one direction (the split
    buffer's ten fields read once, qrhs read and written: 10 + 5 + 5
    doubles, the first direction 15).  The compact scheme's uw_face slot
    holds its three kernels (line solves, limited interface fluxes, flux
    difference)."""
    import astr_amd as A

    compact = conschm.endswith("c")
    f = A.Flow(A.main_config(n, deltat=dt))
    f.init_tgv(1e-3, 2024)
    f.autotune(reps=3)
    f.set_conschm(conschm, recon, chardec, 0.3, 0.05)
    if warmup:
        f.rk3(warmup, stats=True)
    wall, wall_ev, ke, en, prof = timed_steps(f, steps, 3 * 16)
    pts = (n + 1) ** 3
    kern = {f"{k[0]}_{k[1]}": round(v[0] / v[1], 4) for k, v in prof.items()}
    # algorithmic bytes of one interface-flux launch along y (10 split
    # fields in, qrhs in and out) over its measured time
    face_bytes = (10 + 10) * 8 * pts
    face_ms = None if compact else kern.get("uw_face_y")
    scheme = "conschm 543c" if compact else f"conschm 753e, recon_schem {recon}"
    out = {"workload": f"TGV {n}^3 RK3 step, main stage order, {scheme}, "
                       f"lchardecomp {'t' if chardec else 'f'}",
           "ms_per_step": round(wall / steps * 1e3, 4),
           "ms_per_step_instrumented": round(wall_ev / steps * 1e3, 4),
           "value": round(pts * 3 * steps / wall / 1e6, 3), "unit": "Mpt-updates/s",
           "kernels_ms_per_launch": kern,
           "face_y_hbm": ({"achieved_GBs": round(face_bytes / (face_ms * 1e-3) / 1e9, 1),
                           "frac": round(face_bytes / (face_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "alg_bytes_per_point": 160} if face_ms else None),
           "ke_last": float(ke[-1])}
    f.close()
    return out


def load_traffic(n):
    path = os.path.join(ROOT, "profiles", f"traffic_tgv{n}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return None


def apply_selection(f, passes):
    """Force the launch selection recorded in a bench line's launch.passes."""
    import astr_amd as A
    ids = {tuple(v): k for k, v in A.SHAPES.items()}
    for key, v in passes.items():
        if key == "last_pass_split":
            f.split_last = bool(v)
            continue
        name, d = key.rsplit("_", 1)
        f.set_pass_config(A.PASS_NAMES.index(name), "xyz".index(d), ids[tuple(v["shape"])],
                          int(v["minb"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-chem", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching-form measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-form measurement")
    ap.add_argument("--no-main", action="store_true", help="skip the main-path step measurement")
    ap.add_argument("--no-upwind", action="store_true",
                    help="skip the upwind-biased convection measurement")
    ap.add_argument("--serve-only", action="store_true", help="only the serving-form measurement")
    ap.add_argument("--upwind-only", action="store_true",
                    help="only the upwind-biased convection measurement (for profiling)")
    ap.add_argument("--chem-only", action="store_true",
                    help="only the chemistry batch (for counter runs of its kernel)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch selection instead of timing the candidates")
    ap.add_argument("--selection", metavar="JSON",
                    help="apply the launch selection of an earlier bench line (a file holding "
                         "that JSON line) instead of autotuning; counter runs use it so that "
                         "they measure the kernels the timed run chose")
    ap.add_argument("--copy-probe", type=int, default=1 << 31, metavar="BYTES",
                    help="bytes of the HBM copy probe (the achievable-rate reference and the "
                         "counter calibration); 0 skips it")
    ap.add_argument("--probe-only", action="store_true", help="only the HBM copy probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the whole {args.n}^3 domain sits in one MI355X's HBM; "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import astr_amd as A

    if args.chem_only:
        print(json.dumps({"chem_batch": chem_batch()}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serving": serve_form()}), flush=True)
        return
    if args.probe_only:
        print(json.dumps({"hbm_copy_probe": copy_probe(args.copy_probe or 1 << 31)}), flush=True)
        return
    n = args.n
    dt = 5e-4 * 256.0 / n
    if args.upwind_only:
        print(json.dumps({"upwind_path": [upwind_path_step(n, dt, args.steps, args.warmup, 3, False),
                                          upwind_path_step(n, dt, args.steps, args.warmup, 3, True),
                                          upwind_path_step(n, dt, args.steps, args.warmup, 0, False, "543c"),
                                          upwind_path_step(n, dt, args.steps, args.warmup, 0, True, "543c")]}),
              flush=True)
        return
    f = A.Flow(A.tgv_config(n, deltat=dt))
    f.init_tgv(1e-3, 2024)
    tune_s = None
    if args.selection:
        with open(args.selection) as fh:
            line = [ln for ln in fh if ln.lstrip().startswith("{")][-1]
        apply_selection(f, json.loads(line)["launch"]["passes"])
    elif not args.no_autotune:
        t0 = time.perf_counter()
        f.autotune(reps=10)
        tune_s = time.perf_counter() - t0
    launch = {f"{k[0]}_{k[1]}": {"shape": A.SHAPES[v[0]], "minb": v[1]}
              for k, v in f.pass_configs().items()}
    launch["last_pass_split"] = f.split_last
    if args.warmup:
        f.rk3(args.warmup, stats=True)
    f.sync()

    wall, wall_ev, ke, en, prof = timed_steps(f, args.steps, 3 * 9)

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
                     "alg_bytes_per_launch": alg_bytes,
                     "launch_timing": "HIP events around every pass launch, on the context's stream, "
                                      "over the K steps that follow the K timed steps",
                     "ms_per_step_instrumented": round(wall_ev / args.steps * 1e3, 4)},
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
    f.close()
    if not args.no_main:
        out["main_path"] = main_path_step(n, dt, args.steps, args.warmup)
    if not args.no_upwind:
        # MP7 as the reference's shock examples use it, without and with the
        # characteristic decomposition (which brings the sensor in)
        out["upwind_path"] = [upwind_path_step(n, dt, args.steps, args.warmup, 3, False),
                              upwind_path_step(n, dt, args.steps, args.warmup, 3, True),
                              upwind_path_step(n, dt, args.steps, args.warmup, 0, False, "543c"),
                              upwind_path_step(n, dt, args.steps, args.warmup, 0, True, "543c")]
    if not args.no_chem:
        out["chem_batch"] = chem_batch()
    if not args.no_batch:
        out["batching"] = batch_form()
    if not args.no_serve:
        out["serving"] = serve_form()
    if args.copy_probe:
        probe = copy_probe(args.copy_probe)
        out["hbm_copy_probe"] = probe
        # the dominant kernel and the step against what a plain copy reaches
        out["roofline"]["frac_of_copy_probe"] = round(out["roofline"]["achieved"] / probe["best_GBs"], 4)
        out["step_roofline"]["frac_of_copy_probe"] = round(out["step_roofline"]["achieved_GBs"]
                                                           / probe["best_GBs"], 4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

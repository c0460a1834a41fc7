"""Serving form: a continuous-batching request loop for small cases.

Requests arrive over time and are of two kinds:
  - Ignition: a 0-D constant-pressure H2-air reactor integrated to t_end on
    the GPU (astr_amd.chem.ReactorPool: one slot per request);
  - Vortex: a 2-D isentropic-vortex case advanced `steps` RK3 steps
    (astr_amd.Flow contexts advanced together by astr_amd.Batch).

Each tick of the loop:
  1. admits queued requests while slots and the HBM budget allow (a vortex
     case costs its context's device memory; the reactor pool is allocated
     once);
  2. launches one reactor advance on the pool's stream;
  3. while that runs, advances all admitted vortex cases by `steps_per_tick`
     RK3 steps in one batch on the batch's stream (rebuilt only when the
     membership changed);
  4. retires what finished.

Requests that arrive while others are on the GPU join at the next tick, so
the device work of old and new requests overlaps.  A request's result does
not depend on the mix:
  - a reactor's steps depend only on its own state;
  - a vortex case runs the single-case pass body under its own launch
    selection.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

import astr_amd as A
from astr_amd import chem as AC


@dataclass
class Ignition:
    T0: float
    p: float
    phi: float
    t_end: float = 1e-3
    arrival: float = 0.0
    rid: int = -1


@dataclass
class Vortex:
    n: int
    steps: int = 200
    beta: float = 5.0
    arrival: float = 0.0
    rid: int = -1


@dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    admitted: float
    finished: float
    data: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def case_bytes(cfg: A.Config) -> int:
    """Device memory of one context: 36 padded FP64 fields plus tables."""
    pitch = (cfg.im + 1 + 15) // 16 * 16
    plane = pitch * (cfg.jm + 1)
    fs = (plane * (cfg.km + 1) + 1 + 63) // 64 * 64
    return 36 * fs * 8 + 64 * 1024


def h2_air(names, W, phi):
    """Mass fractions of an H2-air mixture of equivalence ratio phi."""
    ix = {s: k for k, s in enumerate(names)}
    X = np.zeros(len(names))
    X[ix["H2"]], X[ix["O2"]], X[ix["N2"]] = 2.0 * phi, 1.0, 3.76
    Y = X * W
    return Y / Y.sum()


class Server:
    def __init__(self, mech_path, hbm_budget=200e9, reactor_slots=2048, rtol=1e-6, atol=1e-12,
                 steps_per_tick=10, reactor_tries=100, h0=1e-10):
        self.mech = AC.Mechanism(mech_path)
        self.pool = AC.ReactorPool(self.mech, reactor_slots, rtol=rtol, atol=atol)
        self.budget = float(hbm_budget)
        self.steps_per_tick = steps_per_tick
        self.reactor_tries = reactor_tries
        self.h0 = h0
        self.free_slots = list(range(reactor_slots))[::-1]
        self.used_bytes = 0
        self.reactors = {}   # slot -> (request, admitted time)
        self.vortices = []   # [request, flow, steps done, admitted, ke list]
        self.batch = None
        self.batch_members = None
        self.ticks = 0
        self.max_active = 0

    def close(self):
        if self.batch is not None:
            self.batch.close()
        for v in self.vortices:
            v[1].close()
        self.pool.close()
        self.mech.close()

    # ---- admission ---------------------------------------------------------
    def _admit(self, req, now):
        if isinstance(req, Ignition):
            if not self.free_slots:
                return False
            slot = self.free_slots.pop()
            Y0 = h2_air(self.mech.names, self.mech.W, req.phi)
            self.pool.set(slot, req.T0, req.p, Y0, req.t_end, self.h0)
            self.reactors[slot] = (req, now)
            return True
        cfg = A.vortex_config(req.n)
        need = case_bytes(cfg)
        if self.used_bytes + need > self.budget:
            return False
        f = A.Flow(cfg)
        f.init_vortex(req.beta)
        self.used_bytes += need
        self.vortices.append([req, f, 0, now, []])
        return True

    # ---- one tick ------------------------------------------------------------
    def _tick(self, clock, results):
        self.ticks += 1
        if self.reactors:
            self.pool.advance(self.reactor_tries)
        if self.vortices:
            members = tuple(id(v[1]) for v in self.vortices)
            if members != self.batch_members:
                if self.batch is not None:
                    self.batch.close()
                self.batch = A.Batch([v[1] for v in self.vortices])
                self.batch_members = members
            k = min(self.steps_per_tick, min(v[0].steps - v[2] for v in self.vortices))
            ke, _ = self.batch.rk3(k, graph=True)
            for i, v in enumerate(self.vortices):
                v[2] += k
                v[4].extend(ke[i].tolist())
        done_vortex = [v for v in self.vortices if v[2] >= v[0].steps]
        if self.reactors:
            self.pool.sync()
            st = self.pool.status()
        now = clock()
        if self.reactors:
            for slot in [s for s in self.reactors if st[s] in (AC.ReactorPool.DONE, AC.ReactorPool.FAILED)]:
                req, adm = self.reactors.pop(slot)
                t, T, Y, nstep, nrej = self.pool.get(slot)
                self.pool.release(slot)
                self.free_slots.append(slot)
                results.append(Result(req.rid, "ignition", req.arrival, adm, now,
                                      {"T": T, "Y": Y, "t": t, "steps": nstep, "rejected": nrej,
                                       "ok": int(st[slot]) == AC.ReactorPool.DONE}))
        for v in done_vortex:
            req, f, nsteps, adm, ke = v
            q = f.download()
            results.append(Result(req.rid, "vortex", req.arrival, adm, now,
                                  {"q": q, "ke": np.array(ke), "steps": nsteps}))
            self.vortices.remove(v)
            self.used_bytes -= case_bytes(f.cfg)
            f.close()
        if done_vortex and not self.vortices and self.batch is not None:
            self.batch.close()
            self.batch, self.batch_members = None, None

    def run(self, requests):
        """Serve requests (each with .arrival seconds from the start, on the
        wall clock) and return their results in completion order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        for k, r in enumerate(pending):
            if r.rid < 0:
                r.rid = k
        queue, results = [], []
        t0 = time.perf_counter()

        def clock():
            return time.perf_counter() - t0

        i = 0
        self.max_vortex = 0
        while i < len(pending) or queue or self.reactors or self.vortices:
            now = clock()
            while i < len(pending) and pending[i].arrival <= now:
                queue.append(pending[i])
                i += 1
            waiting = []
            for req in queue:
                if not self._admit(req, clock()):
                    waiting.append(req)
            queue = waiting
            self.max_active = max(self.max_active, len(self.reactors) + len(self.vortices))
            self.max_vortex = max(self.max_vortex, len(self.vortices))
            if self.reactors or self.vortices:
                self._tick(clock, results)
            elif i < len(pending):
                time.sleep(max(0.0, min(pending[i].arrival - clock(), 0.01)))
        return results


def run_alone(req, mech_path, rtol=1e-6, atol=1e-12, h0=1e-10, reactor_tries=100):
    """The plain sequential path for one request: its own reactor pool of one
    slot, or its own context stepped by astr_hip_rk3_steps."""
    if isinstance(req, Ignition):
        mech = AC.Mechanism(mech_path)
        pool = AC.ReactorPool(mech, 1, rtol=rtol, atol=atol)
        pool.set(0, req.T0, req.p, h2_air(mech.names, mech.W, req.phi), req.t_end, h0)
        st = pool.run(reactor_tries)
        t, T, Y, nstep, nrej = pool.get(0)
        pool.close()
        mech.close()
        return {"T": T, "Y": Y, "t": t, "steps": nstep, "rejected": nrej, "ok": int(st[0]) == 2}
    f = A.Flow(A.vortex_config(req.n))
    f.init_vortex(req.beta)
    ke, _ = f.rk3(req.steps)
    q = f.download()
    f.close()
    return {"q": q, "ke": ke, "steps": req.steps}


def poisson_requests(n, rate, seed, frac_ignition=0.7, t_end=1e-3, vortex_steps=200,
                     n_range=(64, 256)):
    """n requests with exponential inter-arrival times at `rate` per second:
    70% ignition (T0 ~ U[950,1400] K, p ~ U[1,10] atm, phi ~ U[0.5,2]),
    30% vortex (grid side ~ U{n_range})."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for k in range(n):
        if rng.uniform() < frac_ignition:
            out.append(Ignition(rng.uniform(950.0, 1400.0), rng.uniform(1.0, 10.0) * 101325.0,
                                rng.uniform(0.5, 2.0), t_end, float(t[k]), k))
        else:
            out.append(Vortex(int(rng.integers(n_range[0], n_range[1] + 1)), vortex_steps, 5.0,
                              float(t[k]), k))
    return out


def latency_stats(results):
    lat = np.array([r.latency for r in results])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {"completed": len(results), "cases_per_s": len(results) / span,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "span_s": span}

"""Serving form: a continuous-batching request loop for small cases.

Requests arrive over time and are of two kinds:
  - Ignition: a 0-D constant-pressure H2-air reactor integrated to t_end on
    the GPU (astr_amd.chem.ReactorPool: one slot per request, a team of 16
    lanes per reactor by default);
  - Vortex: a 2-D isentropic-vortex case advanced `steps` RK3 steps
    (astr_amd.Flow contexts advanced together by astr_amd.Batch).

The two kinds run on their own streams and are tended independently by one
host loop, which never waits on either:
  - when the reactor pool's last advance has finished, the loop retires the
    reactors that reached t_end, admits waiting ignition requests into free
    slots and launches the next advance (`reactor_tries` step attempts);
  - when the vortex batch's last `steps_per_tick` steps have finished, the
    loop collects their KE histories, retires finished cases, rebuilds the
    batch if its membership changed (new cases admitted under the HBM
    budget) and enqueues the next steps.
Requests that arrive while others are on the GPU join at the next launch of
their kind, so the device work of old and new requests overlaps.  A
request's result does not depend on the mix:
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
                 steps_per_tick=10, reactor_tries=100, h0=1e-10, reactor_lanes=None):
        self.mech = AC.Mechanism(mech_path)
        self.pool = AC.ReactorPool(self.mech, reactor_slots, rtol=rtol, atol=atol,
                                   lanes=reactor_lanes)
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
        self.pool_busy = False
        self.batch_busy = False
        self.batch_k = 0
        self.batch_cases = []
        self.launches = {"reactor": 0, "vortex": 0}
        self.max_active = 0
        self.max_vortex = 0

    def close(self):
        if self.pool_busy:
            self.pool.sync()
        if self.batch is not None:
            if self.batch_busy:
                self.batch.rk3_end(stats=False)
            self.batch.close()
        for v in self.vortices:
            v[1].close()
        self.pool.close()
        self.mech.close()

    # ---- admission ---------------------------------------------------------
    def _admit(self, req, now):
        if isinstance(req, Ignition):
            # setting a slot is ordered on the pool's stream: only between advances
            if self.pool_busy or not self.free_slots:
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

    # ---- reactors ----------------------------------------------------------
    def _reactors_collect(self, clock, results):
        """If the pool's advance finished: retire the reactors that are done."""
        if not self.pool_busy or not self.pool.idle():
            return False
        self.pool_busy = False
        st = self.pool.status()
        now = clock()
        for slot in [s for s in self.reactors if st[s] in (AC.ReactorPool.DONE, AC.ReactorPool.FAILED)]:
            req, adm = self.reactors.pop(slot)
            t, T, Y, nstep, nrej = self.pool.get(slot)
            self.pool.release(slot)
            self.free_slots.append(slot)
            results.append(Result(req.rid, "ignition", req.arrival, adm, now,
                                  {"T": T, "Y": Y, "t": t, "steps": nstep, "rejected": nrej,
                                   "ok": int(st[slot]) == AC.ReactorPool.DONE}))
        return True

    def _reactors_launch(self):
        if self.reactors and not self.pool_busy:
            self.pool.advance(self.reactor_tries)
            self.pool_busy = True
            self.launches["reactor"] += 1

    # ---- vortex cases -------------------------------------------------------
    def _vortices_collect(self, clock, results):
        """If the batch's steps finished: record them, retire finished cases."""
        if not self.batch_busy or not self.batch.done():
            return False
        ke, _ = self.batch.rk3_end()
        self.batch_busy = False
        # cases admitted during these steps are not in the batch yet
        for v, h in zip(self.batch_cases, ke):
            v[2] += self.batch_k
            v[4].extend(h.tolist())
        now = clock()
        for v in [v for v in self.vortices if v[2] >= v[0].steps]:
            req, f, nsteps, adm, kh = v
            results.append(Result(req.rid, "vortex", req.arrival, adm, now,
                                  {"q": f.download(), "ke": np.array(kh), "steps": nsteps}))
            self.vortices.remove(v)
            self.used_bytes -= case_bytes(f.cfg)
            f.close()
        if not self.vortices and self.batch is not None:
            self.batch.close()
            self.batch, self.batch_members = None, None
        return True

    def _vortices_launch(self):
        if not self.vortices or self.batch_busy:
            return
        members = tuple(id(v[1]) for v in self.vortices)
        if members != self.batch_members:
            if self.batch is not None:
                self.batch.close()
            self.batch = A.Batch([v[1] for v in self.vortices])
            self.batch_members = members
        k = min(self.steps_per_tick, min(v[0].steps - v[2] for v in self.vortices))
        self.batch_cases = list(self.vortices)
        self.batch.rk3_begin(k, graph=True)
        self.batch_k = k
        self.batch_busy = True
        self.launches["vortex"] += 1

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
        while i < len(pending) or queue or self.reactors or self.vortices:
            now = clock()
            while i < len(pending) and pending[i].arrival <= now:
                queue.append(pending[i])
                i += 1
            moved = self._reactors_collect(clock, results)
            moved = self._vortices_collect(clock, results) or moved
            waiting = []
            for req in queue:
                if not self._admit(req, clock()):
                    waiting.append(req)
            moved = moved or len(waiting) < len(queue)
            queue = waiting
            self._reactors_launch()
            self._vortices_launch()
            self.max_active = max(self.max_active, len(self.reactors) + len(self.vortices))
            self.max_vortex = max(self.max_vortex, len(self.vortices))
            if not moved:
                # nothing finished and nothing admitted: wait a little for the
                # GPU or the next arrival
                nxt = pending[i].arrival - clock() if i < len(pending) else 0.01
                time.sleep(min(max(nxt, 0.0), 2e-4 if (self.pool_busy or self.batch_busy) else 0.01))
        return results


def run_alone(req, mech_path, rtol=1e-6, atol=1e-12, h0=1e-10, reactor_tries=100,
              reactor_lanes=None):
    """The plain sequential path for one request: its own reactor pool of one
    slot, or its own context stepped by astr_hip_rk3_steps."""
    if isinstance(req, Ignition):
        mech = AC.Mechanism(mech_path)
        pool = AC.ReactorPool(mech, 1, rtol=rtol, atol=atol, lanes=reactor_lanes)
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

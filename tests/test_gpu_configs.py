"""The workloads at their own sizes, each against the oracle or the plain path:

  * TGV 256^3 (FP64, seeded noise 1e-3, dt 5e-4) under the launch selection
    the headline measurement uses (profiles/launch_tgv256.json): one
    right-hand side and two RK3 steps against oracle/astr_oracle.c;
  * lines of 513 points on the two launch shapes that only the 512^3
    workload takes (8 x 64 x 9 in 512 threads, 1 x 64 x 9 in 64 threads),
    in every pass and direction, against the oracle on thin boxes;
  * the batching mix: 64 cases from seed 2026 (32 two-dimensional vortices
    of side 64..512, 32 TGV cases of side 32..128, Re 100..3000), 100 RK3
    steps in one hipGraph-captured batch, every case against its own plain
    run and the smallest of each kind against the oracle;
  * the serving mix: a Poisson stream at 100 requests/s of 1 ms ignitions and
    200-step vortex cases of side 64..256, every result against the same
    request alone, and the ignitions against scipy's Radau on the chemistry
    restatement.
"""
import json
import os

import numpy as np
import pytest

import astr_amd as A
from astr_amd import chem as AC
from astr_amd import serve as S
from oracle import chem as OC
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MECH = os.path.join(ROOT, "tests", "golden", "Jachimowski88_H2-9-19.xml")
RHS_TOL = 1e-12
HIST_TOL = 1e-9


def relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def state_err(a, b):
    return [relmax(a[0], b[0]), relmax(a[1:4], b[1:4]), relmax(a[4], b[4])]


def apply_selection(f, passes):
    """Force a recorded launch selection: {pass_dir: {shape, minb}}."""
    ids = {tuple(v): k for k, v in A.SHAPES.items()}
    for key, v in passes.items():
        if key == "last_pass_split":
            f.split_last = bool(v)
            continue
        name, d = key.rsplit("_", 1)
        f.set_pass_config(A.PASS_NAMES.index(name), "xyz".index(d), ids[tuple(v["shape"])], int(v["minb"]))


def test_tgv256_under_bench_selection_matches_oracle():
    n, dt = 256, 5e-4
    line = json.loads(open(os.path.join(ROOT, "profiles", "launch_tgv256.json")).readline())
    s = O.OracleState(O.tgv_cfg(n, dt=dt))
    s.tgv_init(1e-3, 2024)
    s.set_q(s.q.copy())
    f = A.Flow(A.tgv_config(n, deltat=dt))
    apply_selection(f, line["launch"]["passes"])
    f.upload(s.q.copy())
    got = f.rhscal()
    s.rhscal()
    for m in range(5):
        assert relmax(got[m], s.qrhs[m]) < RHS_TOL, m
    ke, en = f.rk3(2)
    ke0, en0 = s.rk3(2)
    assert max(state_err(f.download(), s.q)) < RHS_TOL
    assert relmax(ke, ke0) < HIST_TOL and relmax(en, en0) < HIST_TOL


@pytest.mark.parametrize("dims", [(512, 12, 12), (12, 512, 12), (12, 12, 512)])
def test_lines_513_on_the_long_line_shapes(dims):
    s = O.OracleState(O.tgv_cfg(*dims, dt=1e-4))
    s.tgv_init(1e-3, 9)
    s.set_q(s.q.copy())
    q0 = s.q.copy()
    s.rhscal()
    rhs_ref = s.qrhs.copy()
    s.rk3(2)
    q_ref = s.q.copy()
    for shape in (11, 12):
        for minb in (1, 2):
            f = A.Flow(A.tgv_config(*dims, deltat=1e-4))
            f.split_last = minb == 2
            for prog in range(len(A.PASS_NAMES)):
                for d in range(3):
                    f.set_pass_config(prog, d, shape, minb)
            f.upload(q0)
            got = f.rhscal()
            for m in range(5):
                assert relmax(got[m], rhs_ref[m]) < RHS_TOL, (shape, minb, m)
            f.upload(q0)
            f.rk3(2, stats=False)
            assert max(state_err(f.download(), q_ref)) < RHS_TOL, (shape, minb)
            f.close()


def _batch_specs(ncases=64, seed=2026):
    """The batching workload's cases, drawn as the measurement draws them."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(ncases):
        if k % 2 == 0:
            nv = int(rng.integers(64, 513))
            specs.append(("vortex", A.vortex_config(nv), None))
        else:
            nt = int(rng.integers(32, 129))
            re = float(rng.uniform(100.0, 3000.0))
            specs.append(("tgv", A.tgv_config(nt, deltat=0.128 / nt, reynolds=re), int(rng.integers(1 << 30))))
    return specs


def _start(f, kind, sd):
    if kind == "vortex":
        f.init_vortex(5.0)
    else:
        f.init_tgv(1e-3, sd)


def test_batching_mix_64_cases_100_steps():
    steps = 100
    specs = _batch_specs()
    flows = [A.Flow(cfg) for _, cfg, _ in specs]
    for f, (kind, _, sd) in zip(flows, specs):
        _start(f, kind, sd)
    q_init = [f.download() for f in flows]
    b = A.Batch(flows)
    b.autotune(reps=1)
    for f, (kind, _, sd) in zip(flows, specs):
        _start(f, kind, sd)
    ke_b, en_b = b.rk3(steps, graph=True)
    q_b = [f.download() for f in flows]
    sel = [f.pass_configs() for f in flows]
    b.close()
    for f in flows:
        f.close()
    # every case against its own plain run under the selection the batch used
    for k, (kind, cfg, sd) in enumerate(specs):
        g = A.Flow(cfg)
        ids = {(nm, d): v for (nm, d), v in sel[k].items()}
        for (nm, d), (shape, minb, _) in ids.items():
            g.set_pass_config(A.PASS_NAMES.index(nm), "xyz".index(d), shape, minb)
        _start(g, kind, sd)
        ke, en = g.rk3(steps)
        q = g.download()
        g.close()
        assert max(state_err(q_b[k], q)) < RHS_TOL, (k, kind, cfg.im)
        assert relmax(ke_b[k], ke) < HIST_TOL and relmax(en_b[k], en) < HIST_TOL
    # the smallest case of each kind against the oracle over the 100 steps
    for kind in ("tgv", "vortex"):
        k = min((i for i, sp in enumerate(specs) if sp[0] == kind), key=lambda i: specs[i][1].im)
        cfg = specs[k][1]
        ocfg = O.tgv_cfg(cfg.im, cfg.jm, cfg.km, dt=cfg.deltat, reynolds=cfg.reynolds, mach=cfg.mach,
                         lx=cfg.lx, ly=cfg.ly, lz=cfg.lz)
        s = O.OracleState(ocfg)
        s.set_q(q_init[k])
        ke0, en0 = s.rk3(steps)
        assert max(state_err(q_b[k], s.q)) < 1e-10, (kind, cfg.im)
        assert relmax(ke_b[k], ke0) < HIST_TOL and relmax(en_b[k], en0) < HIST_TOL


def test_serving_mix_at_its_parameters():
    rtol, atol = 1e-5, 1e-10
    reqs = S.poisson_requests(20, 100.0, 17)
    assert all(r.t_end == 1e-3 for r in reqs if isinstance(r, S.Ignition))
    assert all(r.steps == 200 and 64 <= r.n <= 256 for r in reqs if isinstance(r, S.Vortex))
    srv = S.Server(MECH, reactor_slots=256, rtol=rtol, atol=atol, steps_per_tick=10, reactor_tries=100)
    res = {r.rid: r for r in srv.run(reqs)}
    srv.close()
    assert len(res) == len(reqs) and srv.max_active >= 2
    mech = OC.Mechanism(MECH)
    radau_checked = 0
    for req in reqs:
        got = res[req.rid].data
        ref = S.run_alone(req, MECH, rtol=rtol, atol=atol, reactor_tries=100)
        if isinstance(req, S.Ignition):
            assert got["ok"] and ref["ok"]
            assert got["T"] == ref["T"] and np.array_equal(got["Y"], ref["Y"])
            if radau_checked < 4:
                y0 = S.h2_air(mech.names, mech.W, req.phi)
                yr = mech.ignite(req.T0, req.p, y0, req.t_end, rtol=1e-9, atol=1e-14)
                assert abs(got["T"] / yr[0] - 1) < 2e-3, (req, got["T"], yr[0])
                assert np.abs(got["Y"] - yr[1:]).max() < 2e-3
                radau_checked += 1
        else:
            q, q0 = got["q"], ref["q"]
            for m in range(5):
                scale = np.abs(q0[m]).max() if m in (0, 4) else np.abs(q0[1:4]).max()
                assert np.abs(q[m] - q0[m]).max() <= 1e-13 * scale, (req.rid, m)
    assert radau_checked >= 2

"""The serving form: requests admitted while others run, results equal to the
same requests run alone through the plain path (reactors bit-identical,
vortex cases to 1e-13 relative max-norm), and admission held to the HBM
budget."""
import os

import numpy as np
import pytest

from astr_amd import serve as S

pytestmark = pytest.mark.gpu

MECH = os.path.join(os.path.dirname(__file__), "golden", "Jachimowski88_H2-9-19.xml")


def small_mix(seed=3):
    reqs = S.poisson_requests(12, 200.0, seed, t_end=2e-4, vortex_steps=20, n_range=(64, 96))
    assert any(isinstance(r, S.Vortex) for r in reqs) and any(isinstance(r, S.Ignition) for r in reqs)
    return reqs


def test_served_results_equal_alone():
    reqs = small_mix()
    srv = S.Server(MECH, reactor_slots=64, steps_per_tick=5, reactor_tries=40)
    res = {r.rid: r for r in srv.run(reqs)}
    srv.close()
    assert len(res) == len(reqs)
    assert srv.max_active >= 2
    for req in reqs:
        got = res[req.rid].data
        ref = S.run_alone(req, MECH, reactor_tries=40)
        if isinstance(req, S.Ignition):
            assert got["ok"] and ref["ok"]
            assert got["T"] == ref["T"] and np.array_equal(got["Y"], ref["Y"])
            assert got["steps"] == ref["steps"]
        else:
            q, q0 = got["q"], ref["q"]
            for m in range(5):
                scale = np.abs(q0[m]).max() if m in (0, 4) else np.abs(q0[1:4]).max()
                assert np.abs(q[m] - q0[m]).max() <= 1e-13 * scale, (req.rid, m)
            assert np.abs(got["ke"] - ref["ke"]).max() <= 1e-12 * np.abs(ref["ke"]).max()


def test_admission_respects_hbm_budget():
    reqs = [S.Vortex(64, 10, 5.0, 0.0, k) for k in range(4)]
    one = S.case_bytes(S.A.vortex_config(64))
    srv = S.Server(MECH, hbm_budget=1.5 * one, reactor_slots=4, steps_per_tick=5)
    res = srv.run(reqs)
    srv.close()
    assert len(res) == 4
    assert srv.max_vortex == 1
    assert srv.used_bytes == 0

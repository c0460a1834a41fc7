"""Parity of the HIP path's main-solver semantics (ASTR_SEM_MAIN, through the
C ABI) against the main-path restatement oracle/astr_oracle_main.c.

The fields have duplicate planes that differ (as after an RK stage), so
qswap's averages, the primitives it refreshes on the boundary planes and the
halo values kept from before the j and k averages all take part.  Tolerance:
1e-12 relative max-norm per variable (momentum as one vector) for right-hand
sides, gradients and states; 1e-9 relative for statistics histories.
"""
import numpy as np
import pytest

import astr_amd as A
from fields import periodic_field
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
HIST_TOL = 1e-9


def relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def state_err(a, b):
    return [relmax(a[0], b[0]), relmax(a[1:4], b[1:4]), relmax(a[4], b[4])]


def make_pair(im, jm, km, seed=7, dt=1e-3, skew=1e-3):
    """A device case and an oracle state on the same input: a periodic field
    whose duplicate planes are made to differ by `skew`, the same uniform
    metrics on both sides."""
    ocfg = O.tgv_cfg(im, jm, km, dt=dt, suth_dim=110.3)
    q = periodic_field(ocfg, seed)
    rng = np.random.default_rng(seed + 1)
    q = q * (1.0 + skew * rng.standard_normal(q.shape))
    jac, dxi = O.uniform_metrics(ocfg)
    s = O.MainState(ocfg, jac, dxi)
    s.set_q(q)
    f = A.Flow(A.main_config(im, jm, km, deltat=dt))
    shape = (km + 1, jm + 1, im + 1)
    dev_dxi = np.zeros((3, 3) + shape)
    for a in range(3):
        for b in range(3):
            dev_dxi[a, b] = dxi[a + 3 * b].reshape(shape)
    f.set_metrics(jac.reshape(shape), dev_dxi)
    f.upload(q)
    return f, s, q


@pytest.mark.parametrize("dims", [(24, 20, 16), (20, 18, 0), (33, 14, 12)])
def test_main_stage_rhs_matches_oracle(dims):
    """filterq, qswap, gradcal, rhscal on the device against the oracle.
    The filtered states are compared first; the oracle then continues from
    the device's filtered q (the primitives on both sides are still those of
    the uploaded q), so that the right-hand side comparison measures the
    right-hand side evaluation: the filter's own rounding differences (about
    1e-15 of q5) would otherwise reach the momentum right-hand side through
    p = (q5 - rho |u|^2/2)/const6 ~ 71 at the halo rows, amplified by the
    derivative's cancellation."""
    f, s, q = make_pair(*dims)
    f.filterq()
    s.filterq()
    qf = f.download()
    qo, _ = s.get()
    assert max(state_err(qf, qo)) < RHS_TOL
    s.set_q_only(qf)
    f.qswap()
    s.qswap()
    qd, prd = f.download(prims=True)
    qo, pro = s.get()
    assert max(state_err(qd, qo)) < RHS_TOL
    assert relmax(prd, pro) < RHS_TOL
    dvel, dtmp = f.gradcal()
    s.gradcal()
    assert relmax(dvel, s.dvel) < RHS_TOL
    assert relmax(dtmp, s.dtmp) < RHS_TOL
    got = f.rhscal()
    s.rhscal()
    for m in range(5):
        if m == 3 and dims[2] == 0:
            continue
        assert relmax(got[m], s.qrhs[m]) < RHS_TOL, (m, relmax(got[m], s.qrhs[m]))


@pytest.mark.parametrize("dims", [(24, 20, 16), (20, 18, 0)])
def test_main_rk3_steps_match_oracle(dims):
    """Three main-path RK3 steps: state, stored primitives and the
    statistics of each step's first stage."""
    f, s, q = make_pair(*dims, dt=2e-3)
    ke, en, di = f.rk3(3, dissipation=True)
    ke0, en0, di0 = s.rk3(3)
    qd, prd = f.download(prims=True)
    qo, pro = s.get()
    assert max(state_err(qd, qo)) < RHS_TOL
    assert relmax(prd, pro) < RHS_TOL
    assert relmax(ke, ke0) < HIST_TOL
    assert relmax(en, en0) < HIST_TOL
    assert relmax(di, di0) < HIST_TOL


def test_main_every_launch_selection():
    """Every launch shape and register budget of the main-path passes gives
    the oracle's two RK3 steps."""
    dims = (24, 20, 16)
    f0, s, q = make_pair(*dims, dt=2e-3)
    s.rk3(2)
    q_ref, p_ref = s.get()
    tried = 0
    for shape, (tl, p, r) in A.SHAPES.items():
        if p * r < max(dims) + 7 or tl * p * r < tl * (max(dims) + 1):
            continue
        for minb in (1, 2):
            f, _, _ = make_pair(*dims, dt=2e-3)
            f.split_last = (shape + minb) % 2 == 1
            for prog in range(len(A.PASS_NAMES)):
                for d in range(3):
                    f.set_pass_config(prog, d, shape, minb)
            f.rk3(2, stats=False)
            qd, prd = f.download(prims=True)
            assert max(state_err(qd, q_ref)) < RHS_TOL, (shape, minb)
            assert relmax(prd, p_ref) < RHS_TOL, (shape, minb)
            tried += 1
    assert tried >= 4


def test_main_autotune_keeps_state_and_results():
    """The autotuner restores q and the stored primitives; the tuned run
    still matches the oracle."""
    f, s, q = make_pair(36, 32, 28, dt=2e-3)
    q0, p0 = f.download(prims=True)
    f.autotune(reps=1)
    q1, p1 = f.download(prims=True)
    assert np.array_equal(q0, q1) and np.array_equal(p0, p1)
    ke, en = f.rk3(2)
    ke0, en0, _ = s.rk3(2)
    qd = f.download()
    qo, _ = s.get()
    assert max(state_err(qd, qo)) < RHS_TOL
    assert relmax(ke, ke0) < HIST_TOL


def test_main_tgv64_history_matches_oracle():
    """TGV 64^3 (closed-form field, no noise), 20 main-path steps of dt 5e-3:
    histories and final state against the oracle."""
    n = 64
    ocfg = O.tgv_cfg(n, dt=5e-3, suth_dim=110.3)
    mini = O.OracleState(ocfg)
    mini.tgv_init(0.0, 0)
    q = mini.q.copy()
    s = O.MainState(ocfg)
    s.set_q(q)
    f = A.Flow(A.main_config(n, deltat=5e-3))
    f.upload(q)
    jac, dxi = O.uniform_metrics(ocfg)
    shape = (n + 1,) * 3
    dev_dxi = np.zeros((3, 3) + shape)
    for a in range(3):
        dev_dxi[a, a] = dxi[a + 3 * a].reshape(shape)
    f.set_metrics(jac.reshape(shape), dev_dxi)
    ke, en, di = f.rk3(20, dissipation=True)
    ke0, en0, di0 = s.rk3(20)
    assert relmax(ke, ke0) < HIST_TOL
    assert relmax(en, en0) < HIST_TOL
    assert relmax(di, di0) < HIST_TOL
    qo, _ = s.get()
    assert max(state_err(f.download(), qo)) < RHS_TOL


def test_main_rejected_by_batch():
    f = A.Flow(A.main_config(16, deltat=1e-3))
    with pytest.raises(A.AstrError):
        A.Batch([f])

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


def channel_case(im=24, jm=40, km=16, seed=3, beta=1.8):
    """A channel box: x periodic on [0, 2 pi), y between walls at -1 and 1
    on a tanh-stretched grid, z periodic on [0, pi); a laminar profile plus
    seeded noise (the noise makes the duplicate planes of x and z differ)."""
    i, j, k = np.arange(im + 1), np.arange(jm + 1), np.arange(km + 1)
    x = 2 * np.pi * i / im
    y = np.tanh(beta * (-1.0 + 2.0 * j / jm)) / np.tanh(beta)
    z = np.pi * k / km
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rng = np.random.default_rng(seed)
    nz = lambda: rng.standard_normal(X.shape)  # noqa: E731
    rho = 1.0 + 0.02 * nz()
    u = (1.0 - Y ** 2) + 0.05 * nz()
    v = 0.05 * (1.0 - Y ** 2) * nz()
    w = 0.05 * nz()
    p = 1.0 / (1.4 * 0.1 ** 2) + 0.1 * nz()
    q = np.array([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])
    return np.array([X, Y, Z]), q


def channel_pair(dt=1e-4):
    X, q = channel_case()
    km, jm, im = q.shape[1] - 1, q.shape[2] - 1, q.shape[3] - 1
    f = A.Flow(A.main_config(im, jm, km, deltat=dt, lx=2 * np.pi, ly=2.0, lz=np.pi))
    f.set_boundary((3, 4, 3))
    jac, dxi = f.gridgeom(X, npdc=(3, 4, 3))
    f.set_metrics(jac, dxi)
    f.upload(q)
    ocfg = O.tgv_cfg(im, jm, km, dt=dt, suth_dim=110.3, lx=2 * np.pi, ly=2.0, lz=np.pi)
    ocfg.beta_bound = 0.98
    N = q[0].size
    od = np.zeros((9, N))
    for a in range(3):
        for b in range(3):
            od[a + 3 * b] = dxi[a, b].ravel()
    s = O.MainState(ocfg, jac.ravel(), od)
    s.set_npdc((3, 4, 3))
    s.set_q(q)
    return f, s, jac, dxi


def test_channel_metrics_are_stretched_cartesian():
    """gridgeom of the channel grid: dxi(2,2) and J vary along y; off the
    diagonal only the rounding residue of the compact derivative of a constant
    coordinate remains (-5/2 c + 2 c + c/2 is not exactly 0), which the
    reference carries and so do the main path's per-point metrics."""
    f, s, jac, dxi = channel_pair()
    for a in range(3):
        for b in range(3):
            if a != b:
                assert np.abs(dxi[a, b]).max() <= 1e-11 * np.abs(dxi[a, a]).max()
    assert np.abs(dxi[0, 1]).max() > 0.0
    assert np.ptp(dxi[1, 1]) > 0.1 * dxi[1, 1].mean()


def _affine_metrics(shape):
    """A rotated, sheared affine grid x = A xi: constant J = det A and
    dxi = A^-1, a full metric tensor."""
    ca, sa = np.cos(0.3), np.sin(0.3)
    R = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    Am = R @ np.array([[0.4, 0.05, 0.0], [0.0, 0.45, 0.07], [0.03, 0.0, 0.5]])
    inv = np.linalg.inv(Am)
    jac = np.full(shape, np.linalg.det(Am))
    dxi = np.zeros((3, 3) + shape)
    for a in range(3):
        for b in range(3):
            dxi[a, b] = inv[a, b]
    return jac, dxi


def test_main_full_metric_tensor_matches_oracle():
    """A full constant metric tensor (rotated, sheared affine grid) through
    the per-point metric path of every pass: stage and two RK3 steps against
    the oracle."""
    dims = (20, 18, 16)
    ocfg = O.tgv_cfg(*dims, dt=1e-3, suth_dim=110.3)
    q = periodic_field(ocfg, 21)
    q = q * (1.0 + 1e-3 * np.random.default_rng(22).standard_normal(q.shape))
    shape = q.shape[1:]
    jac, dxi = _affine_metrics(shape)
    od = np.zeros((9, q[0].size))
    for a in range(3):
        for b in range(3):
            od[a + 3 * b] = dxi[a, b].ravel()

    def pair():
        f = A.Flow(A.main_config(*dims, deltat=1e-3))
        f.set_metrics(jac, dxi)
        f.upload(q)
        s = O.MainState(ocfg, jac.ravel(), od)
        s.set_q(q)
        return f, s

    f, s = pair()
    f.qswap()
    s.qswap()
    dvel, dtmp = f.gradcal()
    s.gradcal()
    assert relmax(dvel, s.dvel) < RHS_TOL and relmax(dtmp, s.dtmp) < RHS_TOL
    got = f.rhscal()
    s.rhscal()
    for m in range(5):
        assert relmax(got[m], s.qrhs[m]) < RHS_TOL, (m, relmax(got[m], s.qrhs[m]))
    f, s = pair()
    ke, en, di = f.rk3(2, dissipation=True)
    ke0, en0, di0 = s.rk3(2)
    qd, prd = f.download(prims=True)
    qo, pro = s.get()
    assert max(state_err(qd, qo)) < RHS_TOL and relmax(prd, pro) < RHS_TOL
    assert relmax(ke, ke0) < HIST_TOL and relmax(di, di0) < HIST_TOL


def test_channel_stage_matches_oracle():
    """A wall-bounded direction (npdcj = 4) and per-point metrics through
    filterq, qswap, gradcal and rhscal, then two RK3 steps (no boundary
    conditions on either side), against the main-path oracle."""
    f, s, jac, dxi = channel_pair()
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
        assert relmax(got[m], s.qrhs[m]) < RHS_TOL, (m, relmax(got[m], s.qrhs[m]))
    f2, s2, _, _ = channel_pair()
    f2.rk3(2, stats=False)
    s2.rk3(2)
    qd2, prd2 = f2.download(prims=True)
    qo2, pro2 = s2.get()
    assert max(state_err(qd2, qo2)) < RHS_TOL
    assert relmax(prd2, pro2) < RHS_TOL


def test_channel_every_launch_selection():
    """Every launch shape on the wall-bounded passes gives the oracle's RK3
    steps."""
    f0, s, _, _ = channel_pair()
    s.rk3(2)
    q_ref, _ = s.get()
    dims = (24, 40, 16)
    tried = 0
    for shape, (tl, p, r) in A.SHAPES.items():
        if p * r < max(dims) + 7 or tl * p * r < tl * (max(dims) + 1):
            continue
        for minb in (1, 2):
            f, _, _, _ = channel_pair()
            f.split_last = (shape + minb) % 2 == 1
            for prog in range(len(A.PASS_NAMES)):
                for d in range(3):
                    f.set_pass_config(prog, d, shape, minb)
            f.rk3(2, stats=False)
            assert max(state_err(f.download(), q_ref)) < RHS_TOL, (shape, minb)
            tried += 1
    assert tried >= 4


@pytest.mark.parametrize("channel", [False, True])
def test_main_filter_fusion(channel):
    """The next stage's filter sweep along the last direction applied inside
    each RK pass (the default) and as its own pass on the RK pass's launch
    shape give the same bits, periodic and with a wall-bounded direction,
    and both agree with the oracle; a 3-step call then launches the last
    direction's filter pass once (its first stage) instead of 9 times."""
    out = []
    for fuse in (True, False):
        if channel:
            f, s, _, _ = channel_pair()
        else:
            f, s, _ = make_pair(24, 20, 16)
        f.set_filter_fusion(fuse)
        f.profile_begin(400)
        f.rk3(3, stats=False)
        prof = f.profile_end()
        s.rk3(3)
        qd, prd = f.download(prims=True)
        qo, pro = s.get()
        assert max(state_err(qd, qo)) < RHS_TOL, fuse
        assert relmax(prd, pro) < RHS_TOL, fuse
        out.append((prof, qd, prd))
    (fused, qa, pa), (plain, qb, pb) = out
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)
    assert fused[("filter", "z")][1] == 1 and plain[("filter", "z")][1] == 9
    assert fused[("filter", "x")][1] == 9 and plain[("filter", "x")][1] == 9

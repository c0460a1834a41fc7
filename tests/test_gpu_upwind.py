"""Parity of the HIP path's upwind-biased convection (main path, through the
C ABI: astr_hip_set_convection) against the restatement of convrsduwd,
convrsdcmp, ducrossensor and their pointwise pieces in
oracle/astr_oracle_main.c and oracle/astr_oracle_upwind.c.

Every reconstruction scheme of the explicit scheme (recon_schem -1, 0, 1,
2, 3, 5, 6) and the compact scheme 543c (bandwidth factors 0.3 and 1), with
and without the characteristic decomposition, in 3-D and 2-D, on fields whose
duplicate planes differ (so qswap's averages, refreshed primitives and kept
halo sources all take part) and that carry a compression bump (so the shock
sensor flags some interfaces and not others); a full metric tensor; RK3
steps with the filter.  Tolerance: 1e-12 relative max-norm per variable
(momentum as one vector) for right-hand sides and states.
"""
import numpy as np
import pytest

import astr_amd as A
from fields import periodic_field
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
SCHEMES = [-1, 0, 1, 2, 3, 5, 6]


def relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def state_err(a, b):
    return [relmax(a[0], b[0]), relmax(a[1:4], b[1:4]), relmax(a[4], b[4])]


def bumped_field(cfg, seed, skew=1e-3):
    """periodic_field plus a compression bump along x (pressure and density
    up, u converging), with the duplicate planes made to differ by skew."""
    q = periodic_field(cfg, seed)
    ii = np.arange(cfg.im + 1) % cfg.im
    X = 2 * np.pi * ii / cfg.im
    bump = np.exp(-8 * (1 - np.cos(X - np.pi)))[None, None, :]
    q[0] = q[0] * (1 + 0.3 * bump)
    q[1] = q[1] - 0.5 * q[0] * np.sin(X - np.pi)[None, None, :] * bump
    q[4] = q[4] * (1 + 0.3 * bump)
    rng = np.random.default_rng(seed + 1)
    return q * (1.0 + skew * rng.standard_normal(q.shape))


def make_pair(dims, scheme, chardec, shkcrt=0.02, bfac=0.3, dt=1e-3, metrics=None, seed=7, conschm="753e"):
    ocfg = O.tgv_cfg(*dims, dt=dt, suth_dim=110.3)
    q = bumped_field(ocfg, seed)
    shape = q.shape[1:]
    if metrics is None:
        jac, od = O.uniform_metrics(ocfg)
        dev_jac = jac.reshape(shape)
        dev_dxi = np.zeros((3, 3) + shape)
        for a in range(3):
            for b in range(3):
                dev_dxi[a, b] = od[a + 3 * b].reshape(shape)
    else:
        dev_jac, dev_dxi = metrics(shape)
        jac = dev_jac.ravel()
        od = np.zeros((9, q[0].size))
        for a in range(3):
            for b in range(3):
                od[a + 3 * b] = dev_dxi[a, b].ravel()
    s = O.MainState(ocfg, jac, od)
    s.set_conv(2 if conschm.endswith("c") else 1, scheme, chardec, bfac, shkcrt)
    s.set_q(q)
    f = A.Flow(A.main_config(*dims, deltat=dt))
    f.set_metrics(dev_jac, dev_dxi)
    f.set_conschm(conschm, scheme, chardec, bfac, shkcrt)
    f.upload(q)
    return f, s


CASES = [(sch, False) for sch in SCHEMES] + [(0, True), (1, True), (3, True), (5, True)]


@pytest.mark.parametrize("dims", [(24, 20, 16), (26, 20, 0)])
@pytest.mark.parametrize("scheme,chardec", CASES)
def test_upwind_rhs_matches_oracle(dims, scheme, chardec):
    f, s = make_pair(dims, scheme, chardec)
    f.qswap()
    s.qswap()
    got = f.rhscal()
    s.gradcal()
    s.rhscal()
    if scheme == 5 or chardec:
        flags = s.lshock
        assert 0 < flags.sum() < flags.size   # the sensor splits the box
    for m in range(5):
        if m == 3 and dims[2] == 0:
            continue
        assert relmax(got[m], s.qrhs[m]) < TOL, (m, relmax(got[m], s.qrhs[m]))
    f.close()


@pytest.mark.parametrize("scheme,chardec", [(1, False), (3, True), (5, False)])
def test_upwind_rk3_steps_match_oracle(scheme, chardec):
    """Two RK3 steps of the main stage (filter, qswap, gradients, sensor,
    upwind convection, viscous terms, update) against the oracle."""
    f, s = make_pair((24, 20, 16), scheme, chardec, dt=2e-3)
    ke, en, di = f.rk3(2, dissipation=True)
    ke0, en0, di0 = s.rk3(2)
    qd, prd = f.download(prims=True)
    qo, pro = s.get()
    assert max(state_err(qd, qo)) < TOL
    assert relmax(prd, pro) < TOL
    assert relmax(ke, ke0) < 1e-9 and relmax(en, en0) < 1e-9 and relmax(di, di0) < 1e-9
    f.close()


def _affine_metrics(shape):
    """A rotated, sheared affine grid: constant J = det A, dxi = A^-1."""
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


@pytest.mark.parametrize("scheme,chardec", [(2, False), (3, True)])
def test_upwind_full_metric_tensor_matches_oracle(scheme, chardec):
    f, s = make_pair((20, 18, 16), scheme, chardec, metrics=_affine_metrics)
    f.qswap()
    s.qswap()
    got = f.rhscal()
    s.gradcal()
    s.rhscal()
    for m in range(5):
        assert relmax(got[m], s.qrhs[m]) < TOL, (m, relmax(got[m], s.qrhs[m]))
    f.close()


def test_upwind_launch_selection_and_central_switch_back():
    """The gradient passes leave the convective part to the upwind pass under
    every launch shape (split last pass included); switching back to 643c
    restores the central right-hand side."""
    dims = (24, 20, 16)
    f0, s = make_pair(dims, 1, False)
    f0.close()
    s.qswap()
    s.gradcal()
    s.rhscal()
    ref = s.qrhs.copy()
    tried = 0
    for shape, (tl, p, r) in A.SHAPES.items():
        if p * r < max(dims) + 7 or tl * p * r < tl * (max(dims) + 1):
            continue
        f, _ = make_pair(dims, 1, False)
        f.split_last = shape % 2 == 1
        for prog in range(len(A.PASS_NAMES)):
            for d in range(3):
                f.set_pass_config(prog, d, shape, 1 + shape % 2)
        f.qswap()
        got = f.rhscal()
        assert max(relmax(got[m], ref[m]) for m in range(5)) < TOL, shape
        f.close()
        tried += 1
    assert tried >= 4
    f, _ = make_pair(dims, 1, False)
    f.set_conschm("643c")
    c = O.MainState(O.tgv_cfg(*dims, dt=1e-3, suth_dim=110.3))
    c.set_q(f.download())
    f.qswap()
    c.qswap()
    got = f.rhscal()
    c.gradcal()
    c.rhscal()
    assert max(relmax(got[m], c.qrhs[m]) for m in range(5)) < TOL
    f.close()


# the compact upwind-biased scheme (conschm 543c, convrsdcmp)

@pytest.mark.parametrize("dims", [(24, 20, 16), (26, 20, 0)])
@pytest.mark.parametrize("chardec", [False, True])
@pytest.mark.parametrize("bfac", [0.3, 1.0])
def test_compact_upwind_rhs_matches_oracle(dims, chardec, bfac):
    f, s = make_pair(dims, 0, chardec, bfac=bfac, conschm="543c")
    f.qswap()
    s.qswap()
    got = f.rhscal()
    s.gradcal()
    s.rhscal()
    if chardec:
        flags = s.lshock
        assert 0 < flags.sum() < flags.size   # the sensor splits the box
    for m in range(5):
        if m == 3 and dims[2] == 0:
            continue
        assert relmax(got[m], s.qrhs[m]) < TOL, (m, relmax(got[m], s.qrhs[m]))
    f.close()


@pytest.mark.parametrize("chardec", [False, True])
def test_compact_upwind_rk3_steps_match_oracle(chardec):
    f, s = make_pair((24, 20, 16), 0, chardec, dt=2e-3, conschm="543c")
    ke, en, di = f.rk3(2, dissipation=True)
    ke0, en0, di0 = s.rk3(2)
    qd, prd = f.download(prims=True)
    qo, pro = s.get()
    assert max(state_err(qd, qo)) < TOL
    assert relmax(prd, pro) < TOL
    assert relmax(ke, ke0) < 1e-9 and relmax(en, en0) < 1e-9 and relmax(di, di0) < 1e-9
    f.close()


def test_compact_upwind_full_metric_tensor_matches_oracle():
    f, s = make_pair((20, 18, 16), 0, True, metrics=_affine_metrics, conschm="543c")
    f.qswap()
    s.qswap()
    got = f.rhscal()
    s.gradcal()
    s.rhscal()
    for m in range(5):
        assert relmax(got[m], s.qrhs[m]) < TOL, (m, relmax(got[m], s.qrhs[m]))
    f.close()


@pytest.mark.parametrize("dims", [(128, 12, 0), (64, 12, 12)])
@pytest.mark.parametrize("conschm,recon,chardec", [("543c", 0, False), ("543c", 0, True), ("753e", 3, False),
                                                   ("753e", 5, True)])
def test_shock_tube_matches_oracle(dims, conschm, recon, chardec):
    """A periodic double shock tube (Sod's states, Ma = 1, inviscid) through
    20 RK3 steps: the limiters act at the shocks, contacts and rarefaction
    edges, and the device stays on the oracle to 1e-12 and within the
    states' density range."""
    from fields import double_shock_tube
    im, jm, km = dims
    dt = 5e-3 * 128 / im
    ocfg = O.tgv_cfg(im, jm, km, dt=dt, mach=1.0, reynolds=1e30, suth_dim=110.3)
    q = double_shock_tube(im, jm, km)
    shape = q.shape[1:]
    jac, od = O.uniform_metrics(ocfg)
    dev_dxi = np.zeros((3, 3) + shape)
    for a in range(3):
        for b in range(3):
            dev_dxi[a, b] = od[a + 3 * b].reshape(shape)
    s = O.MainState(ocfg, jac, od)
    s.set_conv(2 if conschm == "543c" else 1, recon, chardec, 0.3, 0.05)
    s.set_q(q)
    cfg = A.main_config(im, jm, km, deltat=dt, mach=1.0)
    cfg.reynolds = 1e30
    f = A.Flow(cfg)
    f.set_metrics(jac.reshape(shape), dev_dxi)
    f.set_conschm(conschm, recon, chardec, 0.3, 0.05)
    f.upload(q)
    f.rk3(20, stats=False)
    s.rk3(20)
    qd, prd = f.download(prims=True)
    qo, pro = s.get()
    assert max(state_err(qd, qo)) < TOL, state_err(qd, qo)
    assert relmax(prd, pro) < TOL
    assert 0.125 * 0.98 < qd[0].min() and qd[0].max() < 1.02
    f.close()


def test_upwind_refusals():
    f = A.Flow(A.main_config(16, 16, 16, deltat=1e-3))
    with pytest.raises(A.AstrError, match="recon_schem"):
        f.set_conschm("753e", 4)
    with pytest.raises(A.AstrError, match="543c only"):
        f.set_conschm("753c", 0)
    with pytest.raises(A.AstrError, match="643c"):
        f.set_conschm("642c")
    f.set_boundary((3, 4, 3))
    with pytest.raises(A.AstrError, match="periodic"):
        f.set_conschm("753e", 0)
    f.set_boundary((3, 3, 3))
    f.set_conschm("753e", 0)
    with pytest.raises(A.AstrError, match="periodic"):
        f.set_boundary((4, 3, 3))
    f.close()
    m = A.Flow(A.tgv_config(16, deltat=1e-3))
    with pytest.raises(A.AstrError, match="main-path"):
        m.set_conschm("753e", 0)
    m.close()

"""Pinning the main solver's restatement (oracle/astr_oracle_main.c), CPU only.

The reference's main path cannot be built here (its MPI and HDF5 modules),
so the restatement is pinned through what ties it to pinned pieces and to
the mathematics:
  * reduction: with uniform Cartesian metrics and primitives consistent with
    q, the main right-hand side divided by J is the miniapp right-hand side
    of oracle/astr_oracle.c (itself pinned bit for bit against the
    reference's Fortran rhscal) up to rounding;
  * rotation covariance: on a rotated, sheared affine grid (a full constant
    metric tensor) a rotation of the grid and of the velocity leaves the
    density and energy right-hand sides unchanged and rotates the momentum
    right-hand side;
  * qswap: duplicate planes equal afterwards, the averages equal the pinned
    numpy restatement oracle.qswap, the stored primitives of the boundary
    planes are q2fvar of the averaged q, the interior primitives are left
    stale, and the x halo of a line on a j boundary plane holds the values
    from before the j average (the reference's order);
  * the stage loop against an independent numpy composition of the pinned
    operators, and its TGV history against the miniapp's (the same scheme
    with the filter moved to the stage start).
Each check is chosen so that a plausible slip (a dropped term, a transposed
metric index, a wrong sign, averages in the wrong order, fresh instead of
stored primitives) fails it.
"""
import numpy as np
import pytest

from fields import periodic_field
from oracle import oracle as O

HM = 5


def relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def main_rhs(cfg, q, jac=None, dxi=None, filt=False):
    s = O.MainState(cfg, jac, dxi)
    s.set_q(q)
    if filt:
        s.filterq()
    s.qswap()
    s.gradcal()
    s.rhscal()
    return s


@pytest.mark.parametrize("dims", [(16, 14, 12), (18, 16, 0)])
def test_main_rhs_reduces_to_miniapp(dims):
    """Uniform metrics, consistent primitives, equal duplicate planes: the
    main path's J-weighted right-hand side over J is the miniapp's."""
    cfg = O.tgv_cfg(*dims, dt=1e-3)
    q = periodic_field(cfg, 3)
    s = main_rhs(cfg, q)
    jac = cfg.dx * cfg.dy * (cfg.dz if cfg.km > 0 else 1.0)
    mini = O.OracleState(cfg)
    mini.set_q(q)
    mini.rhscal()
    for m in range(5):
        if m == 3 and cfg.km == 0:
            continue
        err = relmax(s.qrhs[m] / jac, mini.qrhs[m])
        # rounding only: the flux J (q U + dxi p) against (q u + p)/dx differ
        # in the order of operations on p ~ 1/(gamma Ma^2) = 71
        assert err < 1e-12, (m, err)
    # gradients: dvel(m,n) = D_n(u_m) dxi(n,n) against D_n(u_m)/dx_n
    assert relmax(s.dvel, mini.dvel) < 1e-13
    assert relmax(s.dtmp, mini.dtmp) < 1e-13


def test_main_rhs_planted_slip_fails_reduction():
    """The reduction check's resolution: dxi(1,1) perturbed by 1e-9 moves the
    x-momentum right-hand side by far more than the check's 1e-12."""
    cfg = O.tgv_cfg(16, 14, 12, dt=1e-3)
    q = periodic_field(cfg, 3)
    jac, dxi = O.uniform_metrics(cfg)
    base = main_rhs(cfg, q, jac, dxi).qrhs.copy()
    dxi2 = dxi.copy()
    dxi2[0] *= 1.0 + 1e-9
    pert = main_rhs(cfg, q, jac, dxi2).qrhs
    assert relmax(pert[1], base[1]) > 1e-10


def _rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


@pytest.mark.parametrize("filt", [False, True])
def test_main_rhs_rotation_covariance(filt):
    """x = A xi with a full constant A (rotated and sheared): dxi = A^-1,
    J = det A.  Rotating the grid by R (A' = R A) and the velocity (q_2..4' =
    R q_2..4) must leave the density and energy right-hand sides and rotate
    the momentum ones.  A transposed dxi index in gradcal or the fluxes, or a
    dropped off-diagonal term, breaks this."""
    cfg = O.tgv_cfg(14, 12, 13, dt=1e-3)
    q = periodic_field(cfg, 9)
    N = q[0].size
    A = _rot(0.3, -0.5, 0.8) @ np.array([[0.4, 0.05, 0.0], [0.0, 0.45, 0.07], [0.03, 0.0, 0.5]])
    R = _rot(-0.7, 0.4, 1.1)

    def metrics(Am):
        inv = np.linalg.inv(Am)             # d xi_d / d x_n = inv[d, n]
        dxi = np.zeros((9, N))
        for d in range(3):
            for n in range(3):
                dxi[d + 3 * n] = inv[d, n]
        return np.full(N, np.linalg.det(Am)), dxi

    s1 = main_rhs(cfg, q, *metrics(A), filt=filt)
    q2 = q.copy()
    q2[1:4] = np.einsum("ab,bkji->akji", R, q[1:4])
    s2 = main_rhs(cfg, q2, *metrics(R @ A), filt=filt)
    r1, r2 = s1.qrhs, s2.qrhs
    scale = np.abs(r1).max()
    assert np.abs(r2[0] - r1[0]).max() < 1e-11 * scale
    assert np.abs(r2[4] - r1[4]).max() < 1e-11 * scale
    rot = np.einsum("ab,bkji->akji", R, r1[1:4])
    assert np.abs(r2[1:4] - rot).max() < 1e-11 * scale
    # the gradient tensor transforms as R dvel R^T
    d1 = s1.dvel.reshape(3, 3, -1)   # [n, m]: dvel(m, n) at m + 3 n
    d2 = s2.dvel.reshape(3, 3, -1)
    g1 = np.transpose(d1, (1, 0, 2))  # [m, n]
    g2 = np.transpose(d2, (1, 0, 2))
    g1r = np.einsum("ab,bcx,dc->adx", R, g1, R)
    assert np.abs(g2 - g1r).max() < 1e-11 * np.abs(g1).max()


def test_main_qswap_order_and_primitives():
    """qswap on a field whose duplicate planes differ."""
    cfg = O.tgv_cfg(14, 13, 12, dt=1e-3)
    rng = np.random.default_rng(4)
    q = periodic_field(cfg, 5)
    q = q * (1.0 + 1e-3 * rng.standard_normal(q.shape))  # planes 0 and n now differ
    s = O.MainState(cfg)
    s.set_q(q)
    _, pr0 = s.get()
    s.qswap()
    qa, pra = s.get()
    # the averages: the pinned numpy restatement, bit for bit
    assert np.array_equal(qa, O.qswap(q))
    # primitives of the boundary planes from the averaged q; interior stale
    fresh = O.MainState(cfg)
    fresh.set_q(qa)
    _, prf = fresh.get()
    bnd = np.zeros(q.shape[1:], dtype=bool)
    bnd[0] = bnd[-1] = True
    bnd[:, 0] = bnd[:, -1] = True
    bnd[:, :, 0] = bnd[:, :, -1] = True
    assert np.array_equal(pra[:, bnd], prf[:, bnd])
    assert np.array_equal(pra[:, ~bnd], pr0[:, ~bnd])
    # halos: the x halo of the lines on j = 0 holds q before the j average;
    # the x halo of interior lines and the z halo hold the averaged q
    qh = s.qh
    i_src = cfg.im - 3      # halo i = -3
    qi = q.copy()           # after the i average only
    qi[..., 0] = 0.5 * (q[..., 0] + q[..., -1])
    qi[..., -1] = qi[..., 0]
    k = 4
    assert np.array_equal(qh[:, k + HM, 0 + HM, -3 + HM], qi[:, k, 0, i_src])
    assert not np.array_equal(qh[:, k + HM, 0 + HM, -3 + HM], qa[:, k, 0, i_src])
    assert np.array_equal(qh[:, k + HM, 5 + HM, -3 + HM], qa[:, k, 5, i_src])
    assert np.array_equal(qh[:, -2 + HM, 5 + HM, 7], qa[:, cfg.km - 2, 5, 7 - HM])
    # and the primitives there are q2fvar of those values
    ph = s.primsh
    t = O.MainState(cfg)
    t.set_q(qi)
    _, pri = t.get()
    assert np.array_equal(ph[:, k + HM, 0 + HM, -3 + HM], pri[:, k, 0, i_src])


def test_main_qswap_2d_refreshes_every_point():
    """A 2-D case still takes qswap's k step (parallel.F90:5159-5196): the
    slabs -hm..0 and 0..hm of km = 0 are the plane itself, so after qswap the
    stored primitives of every point are q2fvar of the averaged q, where in
    3-D only the boundary planes are."""
    cfg = O.tgv_cfg(14, 12, 0, dt=1e-3)
    q = periodic_field(cfg, 5)
    s = O.MainState(cfg)
    s.set_q(q)
    s.filterq()
    qf, pr_stale = s.get()
    s.qswap()
    qa, pra = s.get()
    fresh = O.MainState(cfg)
    fresh.set_q(qa)
    _, prf = fresh.get()
    assert np.array_equal(pra, prf)
    assert not np.array_equal(pra[:, :, 3:-3, 3:-3], pr_stale[:, :, 3:-3, 3:-3])


def test_main_gradcal_reads_stored_primitives():
    """gradcal differentiates the stored primitives, not those of the current
    q: after filterq (which changes q only) its interior gradients equal those
    of the state before the filter, except near the boundary planes, whose
    primitives qswap refreshes from the filtered q."""
    cfg = O.tgv_cfg(16, 14, 12, dt=1e-3)
    q = periodic_field(cfg, 6)
    a = main_rhs(cfg, q)
    b = main_rhs(cfg, q, filt=True)
    # interior lines far from every boundary plane see the same stored values
    # along their whole stencil except through the compact solve's coupling,
    # which decays by ~0.27 per row from the refreshed planes
    sl = (slice(None), slice(6, -6), slice(6, -6), slice(6, -6))
    assert relmax(b.dvel[sl], a.dvel[sl]) < 1e-3
    assert relmax(b.dvel, a.dvel) > 0.0
    # the filtered state's q differs from the unfiltered one
    qa, _ = a.get()
    qb, _ = b.get()
    assert relmax(qb, qa) > 1e-6


def test_main_stage_loop_matches_composition():
    """or_main_rk3_steps against the same stage composed here from the
    pinned operators (filterq, qswap, gradcal, stats, rhscal, updatefvar)
    with the RK3 update (mainloop.F90:441-449) in numpy."""
    cfg = O.tgv_cfg(12, 12, 12, dt=2e-3)
    q = periodic_field(cfg, 8)
    s = O.MainState(cfg)
    s.set_q(q)
    ke, en, di = s.rk3(2)
    jac = cfg.dx * cfg.dy * cfg.dz
    t = O.MainState(cfg)
    t.set_q(q)
    rk = [(1.0, 0.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0)]
    hist = []
    for it in range(2):
        for st in range(3):
            t.filterq()
            t.qswap()
            t.gradcal()
            qc, _ = t.get()
            if st == 0:
                qsave = qc * jac
                hist.append(t.stats())
            t.rhscal()
            a, b, c = rk[st]
            qn = (a * qsave + b * qc * jac + c * t.qrhs * cfg.dt) / jac
            qh = t.qh
            qh[:, HM:HM + cfg.km + 1, HM:HM + cfg.jm + 1, HM:HM + cfg.im + 1] = qn
            t.updatefvar()
    q1, p1 = s.get()
    q2, p2 = t.get()
    assert np.array_equal(q1, q2)
    assert np.array_equal(p1, p2)
    assert np.array_equal(np.array([ke, en, di]).T, np.array(hist))


def test_main_history_close_to_miniapp():
    """The main path moves the filter to the stage start and differentiates
    the unfiltered primitives; on a smooth TGV both paths follow the same
    decay: KE and enstrophy agree to about 2e-9 over 40 steps at 32^3 (held
    to 1e-8), and KE(0) is the field's closed form 1/8 up to the filter's
    first action."""
    cfg = O.tgv_cfg(32, dt=5e-3)
    mini = O.OracleState(cfg)
    mini.tgv_init(0.0, 0)
    q0 = mini.q.copy()
    ke_m, en_m = mini.rk3(40)
    s = O.MainState(cfg)
    s.set_q(q0)
    ke, en, di = s.rk3(40)
    assert abs(ke[0] - 0.125) < 1e-9
    assert relmax(ke, ke_m) < 1e-8
    assert relmax(en, en_m) < 1e-8
    assert np.all(np.diff(ke) < 0)

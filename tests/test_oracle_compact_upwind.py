"""Pinning the restatement of the compact upwind-biased convection (conschm
543c, convrsdcmp, src/solver.F90:1271-1937) in oracle/astr_oracle_upwind.c
(or_cuw_lu, or_cuw_flux, or_mp5) and oracle/astr_oracle_main.c (convrsdcmp),
CPU only.

The reference's flux.F90 uses its MPI module and is not built here, so the
restatement is pinned by what the mathematics fixes:
  * the compact interface flux reproduces the interface values of a
    polynomial exactly from its cell averages up to degree 4 for every
    bandwidth factor (a 5th-order scheme), and at degree 5 only for
    bfacmpld = 1, where the scheme becomes the symmetric 6th-order one; a
    mistyped coefficient of the left- or right-hand side breaks this;
  * the Thomas sweeps on the reference's LU factors equal a dense solve of
    the system a x(j-1) + x(j) + c x(j+1) = d(j) written from the scheme's
    definition, and the two winds mirror each other;
  * on a smooth periodic function the interface value converges at 5th
    order;
  * MP5 returns the linear value when its limiter is idle or the interface
    is not flagged, and stays inside the data's range at a step;
  * the assembled right-hand side: zero on a uniform flow, converging to the
    pinned central scheme's on smooth fields, unchanged by the
    characteristic projection where MP5 is idle, mirror symmetric, and not
    the explicit scheme's.
"""
import numpy as np
import pytest

from fields import periodic_field
from oracle import oracle as O
from test_oracle_upwind import smooth_field


def cell_averages(F, n):
    """Averages over [t - 1/2, t + 1/2] / n, t = -4..n+4, of the function
    whose primitive is F."""
    t = np.arange(-4, n + 5).astype(float)
    return (F((t + 0.5) / n) - F((t - 0.5) / n)) * n


@pytest.mark.parametrize("plus", [True, False])
@pytest.mark.parametrize("bfac", [0.0, 0.3, 1.0])
def test_compact_flux_reproduces_polynomials_to_degree_4(plus, bfac):
    n = 32
    xi = (np.arange(-1, n + 1) + 0.5) / n
    for deg in range(7):
        f = cell_averages(lambda y: y ** (deg + 1) / (deg + 1), n)
        err = np.abs(O.cuw_flux(n, bfac, plus, f) - xi ** deg).max()
        if deg <= 4 or (deg == 5 and bfac == 1.0):
            assert err < 1e-13, (deg, err)
        else:
            assert err > 1e-11, (deg, err)


def dense_compact_flux(n, bfac, plus, f):
    """The scheme written from its definition (compact_flux_initiate and
    compact_flux_rhs for a periodic line) and solved densely."""
    N = n + 4
    a = 0.5 - bfac / 6.0 if plus else 1.0 / 6.0 + bfac / 6.0
    c = 1.0 / 6.0 + bfac / 6.0 if plus else 0.5 - bfac / 6.0
    A = np.eye(N)
    d = np.zeros(N)
    F = lambda j: f[j + 4]  # noqa: E731
    w = [1.0 / 18.0 - bfac / 36.0, 19.0 / 18.0 - 9.0 * bfac / 36.0, 5.0 / 9.0 + 9.0 * bfac / 36.0,
         bfac / 36.0]
    for r in range(N):
        j = r - 2
        if r in (0, N - 1):
            d[r] = (37.0 / 60.0 * (F(j) + F(j + 1)) - 2.0 / 15.0 * (F(j - 1) + F(j + 2))
                    + 1.0 / 60.0 * (F(j - 2) + F(j + 3)))
            continue
        A[r, r - 1], A[r, r + 1] = a, c
        if plus:
            d[r] = w[0] * F(j - 1) + w[1] * F(j) + w[2] * F(j + 1) + w[3] * F(j + 2)
        else:
            d[r] = w[0] * F(j + 2) + w[1] * F(j + 1) + w[2] * F(j) + w[3] * F(j - 1)
    return np.linalg.solve(A, d)[1:-1]


@pytest.mark.parametrize("plus", [True, False])
def test_compact_flux_equals_a_dense_solve(plus):
    rng = np.random.default_rng(3)
    n = 40
    f = rng.standard_normal(n + 9)
    for bfac in (0.0, 0.3, 1.0):
        got = O.cuw_flux(n, bfac, plus, f)
        want = dense_compact_flux(n, bfac, plus, f)
        assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


def test_compact_flux_winds_mirror_each_other():
    """The downwind flux of a line is the upwind flux of the mirrored line,
    read back mirrored (rows -1..n map to n..-1 around the half-point)."""
    rng = np.random.default_rng(5)
    n = 24
    f = rng.standard_normal(n + 9)
    up = O.cuw_flux(n, 0.3, True, f[::-1].copy())
    dn = O.cuw_flux(n, 0.3, False, f)
    assert np.abs(dn - up[::-1]).max() < 1e-14 * np.abs(dn).max()


def test_compact_flux_converges_at_fifth_order():
    errs = []
    for n in (32, 64, 128):
        f = cell_averages(lambda y: -np.cos(2 * np.pi * y) / (2 * np.pi), n)
        xi = (np.arange(-1, n + 1) + 0.5) / n
        errs.append(np.abs(O.cuw_flux(n, 0.3, True, f) - np.sin(2 * np.pi * xi)).max())
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 4.8), (errs, orders)


def test_mp5_limits():
    smooth = np.array([0.0, 0.1, 0.2, 0.3, 0.4])
    assert O.mp5(smooth, 0.25) == 0.25            # linear data: the limiter is idle
    step = np.array([0.0, 0.0, 0.0, 1.0, 1.0])
    assert O.mp5(step, 1.3, shock=False) == 1.3     # unflagged interface: the linear value
    for ul in (-0.4, 0.3, 0.6, 1.4):
        v = O.mp5(step, ul)
        assert 0.0 <= v <= 1.0, (ul, v)
    assert O.mp5(step, 1.4) < 1.4


# ---------------------------------------------------------------------------
# the assembled right-hand side
# ---------------------------------------------------------------------------

def rhs_c(cfg, q, chardec=False, shkcrt=0.05, bfac=0.3, upwind=2):
    s = O.MainState(cfg)
    if upwind:
        s.set_conv(upwind, 0, chardec, bfac, shkcrt)
    s.set_q(q)
    s.qswap()
    s.gradcal()
    s.rhscal()
    return s


@pytest.mark.parametrize("chardec", [False, True])
def test_uniform_flow_has_zero_rhs(chardec):
    cfg = O.tgv_cfg(12, 12, 12, dt=1e-3)
    q = np.empty((5, 13, 13, 13))
    q[0], q[1], q[2], q[3] = 1.2, 0.3, -0.2, 0.1
    q[4] = (1.0 / (1.4 * 0.01)) / 0.4 + 0.5 * (0.09 + 0.04 + 0.01) / 1.2
    s = rhs_c(cfg, q, chardec, shkcrt=-1.0)
    assert np.abs(s.qrhs).max() < 1e-11


def test_compact_rhs_converges_to_the_central_rhs():
    """The whole right-hand side against the central scheme's on 16^3 and
    32^3 smooth fields: the difference falls at the compact upwind scheme's
    order; an index or orientation slip leaves an O(1) difference."""
    def diff(n):
        cfg = O.tgv_cfg(n, n, n, dt=1e-3)
        q = smooth_field(cfg)
        a = rhs_c(cfg, q)
        b = rhs_c(cfg, q, upwind=0)
        return np.abs(a.qrhs - b.qrhs).max() / np.abs(b.qrhs).max()
    e1, e2 = diff(16), diff(32)
    assert np.log2(e1 / e2) > 4.5, (e1, e2)
    assert e2 < 1e-4


def test_characteristic_projection_commutes_where_mp5_is_idle():
    """On a smooth field MP5 returns the compact value, a linear function of
    the line, so projecting on the characteristics and back changes nothing
    but rounding, whether the interfaces are flagged (threshold below every
    sensor value) or not (above); without the projection the sensor does not
    run and every interface counts as flagged.  On a rough field the limiter
    acts, on characteristic variables in one case and on the flux components
    in the other, and the two differ."""
    cfg = O.tgv_cfg(16, 16, 16, dt=1e-3)
    q = smooth_field(cfg)
    a = rhs_c(cfg, q, chardec=False)
    for shk, flagged in ((1e30, False), (-1.0, True)):
        b = rhs_c(cfg, q, chardec=True, shkcrt=shk)
        assert b.lshock.all() if flagged else not b.lshock.any()
        err = np.abs(a.qrhs - b.qrhs).max() / np.abs(a.qrhs).max()
        assert 0.0 < err < 1e-11, (shk, err)
    cfg = O.tgv_cfg(16, 14, 12, dt=1e-3)
    q = periodic_field(cfg, 4)
    c = rhs_c(cfg, q, chardec=True, shkcrt=-1.0)
    d = rhs_c(cfg, q, chardec=False)
    assert np.abs(c.qrhs - d.qrhs).max() / np.abs(d.qrhs).max() > 1e-8


@pytest.mark.parametrize("chardec", [False, True])
def test_mirror_symmetry(chardec):
    """x -> -x with u -> -u maps the right-hand side onto itself with the
    x-momentum negated: the '-' windows and downwind compact solve are the
    mirrored '+' ones."""
    cfg = O.tgv_cfg(18, 14, 0, dt=1e-3, reynolds=1e30)
    q = periodic_field(cfg, 9)
    qm = q[..., ::-1].copy()
    qm[1] = -qm[1]
    a = rhs_c(cfg, q, chardec, shkcrt=-1.0)
    b = rhs_c(cfg, qm, chardec, shkcrt=-1.0)
    bm = b.qrhs[..., ::-1].copy()
    bm[1] = -bm[1]
    assert np.abs(a.qrhs - bm).max() / np.abs(a.qrhs).max() < 1e-10


def test_compact_differs_from_the_explicit_scheme():
    """The compact flux is not the explicit one under another name: on a
    rough field the two right-hand sides differ at O(1) of the flux
    difference, while both stay finite."""
    cfg = O.tgv_cfg(16, 14, 12, dt=1e-3)
    q = periodic_field(cfg, 6)
    a = rhs_c(cfg, q, shkcrt=-1.0)
    b = rhs_c(cfg, q, shkcrt=-1.0, upwind=1)
    assert np.isfinite(a.qrhs).all()
    assert np.abs(a.qrhs - b.qrhs).max() / np.abs(b.qrhs).max() > 1e-6


def test_shock_tube_stays_bounded_where_the_central_scheme_does_not():
    """A periodic double shock tube (Sod's states, Ma = 1, inviscid), 20 RK3
    steps of dt = 5e-3: the shocks, contacts and rarefactions move about four
    cells.  The upwind-biased schemes keep the density within 2% of the
    initial states' range; the central scheme (whose filter is the only
    dissipation) undershoots by more than 5%."""
    from fields import double_shock_tube
    im, jm = 128, 12

    def run(upwind, recon, chardec):
        cfg = O.tgv_cfg(im, jm, 0, dt=5e-3, mach=1.0, reynolds=1e30, suth_dim=110.3)
        s = O.MainState(cfg)
        if upwind:
            s.set_conv(upwind, recon, chardec, 0.3, 0.05)
        s.set_q(double_shock_tube(im, jm))
        s.rk3(20)
        return s.get()[0][0]
    for up, rec, cd in ((2, 0, False), (2, 0, True), (1, 3, False), (1, 5, True)):
        rho = run(up, rec, cd)
        assert np.isfinite(rho).all()
        assert 0.125 * 0.98 < rho.min() and rho.max() < 1.0 * 1.02, (up, rec, cd, rho.min(), rho.max())
    rho = run(0, 0, False)
    assert rho.min() < 0.125 * 0.95


def test_the_reference_fluxtest_on_periodic_lines():
    """The reference's own check of this scheme (fluxtest, src/test.F90:670-869:
    conschm 543c, bfacmpld = 1, the L2 error of (fh(i) - fh(i-1)) / dx
    against the analytic derivative of a sine) on its periodic directions,
    128 points: at bfacmpld = 1 the scheme is the symmetric 6th-order one
    (observed order above 6), at 0.3 the 5th-order upwind one (above 4.5)."""
    def l2(n, bfac):
        t = np.arange(-4, n + 5)
        f = np.sin(2 * np.pi * t / n)
        fh = O.cuw_flux(n, bfac, True, f)
        d = (fh[1:] - fh[:-1]) / (2 * np.pi / n)
        ref = np.cos(2 * np.pi * np.arange(n + 1) / n)
        return np.sqrt(np.mean((d[1:] - ref[1:]) ** 2))
    assert l2(128, 1.0) < 2e-11
    assert np.log2(l2(64, 1.0) / l2(128, 1.0)) > 6.0
    assert np.log2(l2(64, 0.3) / l2(128, 0.3)) > 4.5

"""Physical-boundary closures (the reference's ntype 1, 2, 4): the compact
derivative's boundary rows (derivative.F90:63-306) and the filter's
(filter.F90:31-285), restated in the oracle and pinned bit for bit against
the reference's own Fortran (oracle/_ref) and by closed forms; the device
line batch (astr_hip_line_batch_nt) and gridgeom on grids with a
wall-bounded direction (astr_hip_gridgeom_nt) against the oracle."""
import numpy as np
import pytest

import astr_amd as A
from oracle import geom as G
from oracle import oracle as O


def poly_line(n, deg, seed):
    """A random polynomial of degree deg sampled at t = i/n, i = -5..n+5,
    and its derivative per unit index at i = 0..n."""
    c = np.random.default_rng(seed).standard_normal(deg + 1)
    t = np.arange(-5, n + 6) / n
    d = np.polyval(np.polyder(c), t[5:n + 6]) / n if deg > 0 else np.zeros(n + 1)
    return np.polyval(c, t), d


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref missing")
@pytest.mark.parametrize("n", [12, 33, 64])
def test_boundary_lines_match_reference_fortran(n):
    """df_compact and compact_filter with ntype 1..4 equal the reference's
    own fd_scheme_initiate/df_compact and compact_filter_initiate/
    compact_filter bit for bit, for beta_bound 0.98 (the main path's) and
    1.09."""
    f = np.random.default_rng(n).standard_normal((4, n + 11))
    for nt in (1, 2, 3, 4):
        got = np.array([O.df_compact_nt(x, nt) for x in f])
        assert np.array_equal(got, O.run_ref("line_deriv", f=f, ntype=nt)["out"]), nt
        for bb in (0.98, 1.09):
            got = np.array([O.compact_filter_nt(x, nt, 0.49, 1.11, bb) for x in f])
            ref = O.run_ref("line_filter", f=f, ntype=nt, alfa=0.49, beta_halo=1.11, beta_bound=bb)
            assert np.array_equal(got, ref["out"]), (nt, bb)


def test_boundary_ntype3_is_the_periodic_form():
    f = np.random.default_rng(3).standard_normal((3, 40 + 11))
    for x in f:
        assert np.array_equal(O.df_compact_nt(x, 3), O.df_compact(x))
        assert np.array_equal(O.compact_filter_nt(x, 3, 0.49, 1.11, 1.09), O.compact_filter(x))


def test_boundary_derivative_closed_forms():
    """With a physical end the derivative system is exact on cubics (the
    3rd-order biased row -5/2 f0 + 2 f1 + 1/2 f2 = f'0 + 2 f'1 is) and not
    on quartics, so the boundary rows are the ones in use; an interface
    end's 6th-order rows stay exact on quartics."""
    n = 40
    for deg in range(4):
        f, d = poly_line(n, deg, deg)
        for nt in (1, 2, 4):
            assert np.abs(O.df_compact_nt(f, nt) - d).max() < 1e-14 * max(1.0, np.abs(d).max()), (deg, nt)
    f, d = poly_line(n, 4, 7)
    for nt in (1, 2, 4):
        err = np.abs(O.df_compact_nt(f, nt) - d)
        assert err.max() > 1e-11, nt   # rounding alone stays near 1e-16
        assert np.abs(O.df_compact_nt(f, 3) - d).max() < 1e-14


def test_boundary_filter_closed_forms():
    """With the main path's beta_bound 0.98 the boundary filter preserves
    polynomials of degree <= 5 (its one-sided rows are 6th order) and not
    degree 6; with beta_bound 1.09 its row 0 no longer preserves a constant,
    so beta_bound enters that row."""
    n = 40
    for deg in range(6):
        f, _ = poly_line(n, deg, 10 + deg)
        for nt in (1, 2, 4):
            assert np.abs(O.compact_filter_nt(f, nt) - f[5:n + 6]).max() < 1e-13 * np.abs(f).max(), (deg, nt)
    f, _ = poly_line(n, 6, 20)
    for nt in (1, 2, 4):
        assert np.abs(O.compact_filter_nt(f, nt) - f[5:n + 6]).max() > 1e-11, nt
    one = np.ones(n + 11)
    assert np.abs(O.compact_filter_nt(one, 4, beta_bound=1.09) - 1.0).max() > 1e-4


def test_boundary_bad_ntype_is_rejected():
    f = np.zeros(20 + 11)
    for nt in (0, 5):
        with pytest.raises(ValueError):
            O.df_compact_nt(f, nt)
        with pytest.raises(A.AstrError):
            A.line_batch_nt(0, nt, f)


# ---------------------------------------------------------------------------
# a channel grid: periodic in i and k, wall-bounded and tanh-stretched in j
# ---------------------------------------------------------------------------

def channel(n, amp, gam=1.5):
    """x1 = xi1 h1 + a sin(2 pi xi2/N2) sin(2 pi xi3/N3),
    x2 = Y(xi2) + a sin(2 pi xi1/N1), Y = 1 - tanh(g (1 - 2 xi2/N2))/tanh(g),
    x3 = xi3 h3 + a sin(2 pi xi1/N1) cos(2 pi xi2/N2), with its exact
    d x_m / d xi_d (per unit index)."""
    n1, n2, n3 = n
    h1, h3 = 2 * np.pi / n1, np.pi / n3
    k3, k2, k1 = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    w1, w2, w3 = 2 * np.pi / n1, 2 * np.pi / n2, 2 * np.pi / n3
    s1, c1 = np.sin(w1 * k1), np.cos(w1 * k1)
    s2, c2 = np.sin(w2 * k2), np.cos(w2 * k2)
    s3, c3 = np.sin(w3 * k3), np.cos(w3 * k3)
    eta = 1 - 2 * k2 / n2
    y = 1 - np.tanh(gam * eta) / np.tanh(gam)
    dy = (2 * gam / n2) / np.cosh(gam * eta) ** 2 / np.tanh(gam)
    x = np.stack([k1 * h1 + amp * s2 * s3, y + amp * s1, k3 * h3 + amp * s1 * c2])
    z = 0.0 * k1
    dx = np.array([[z + h1, amp * w2 * c2 * s3, amp * w3 * s2 * c3],
                   [amp * w1 * c1, dy, z],
                   [amp * w1 * c1 * c2, -amp * w2 * s1 * s2, z + h3]])
    return x, dx


def exact_metrics(dx):
    m = np.moveaxis(dx, (0, 1), (-2, -1))
    return np.linalg.det(m), np.moveaxis(np.linalg.inv(m), (-2, -1), (0, 1))


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


CHAN = (3, 4, 3)


def test_oracle_channel_grid_uniform_exact():
    """Linear coordinates in the wall-bounded direction too: the boundary
    rows are exact on them, so dxi = 1/h and J = h1 h2 h3 to rounding."""
    n = (16, 12, 14)
    k3, k2, k1 = np.meshgrid(np.arange(n[2] + 1), np.arange(n[1] + 1), np.arange(n[0] + 1), indexing="ij")
    h = np.array([0.3, 0.2, 0.5])
    x = np.stack([k1 * h[0], k2 * h[1] - 1.0, k3 * h[2]]).astype(float)
    jac, dxi, _ = G.gridgeom(x, CHAN)
    assert np.abs(jac / h.prod() - 1).max() < 1e-14
    for a in range(3):
        for b in range(3):
            want = 1.0 / h[a] if a == b else 0.0
            assert np.abs(dxi[a, b] - want).max() < 1e-13 / h.min(), (a, b)


def test_oracle_channel_metrics_converge_and_satisfy_identities():
    """On the stretched channel grid the metrics converge to the exact ones
    (at least 2.5th order: the wall rows are 3rd order) and the discrete
    metric identities hold to rounding with the boundary rows in j."""
    errs = []
    for n in ((16, 16, 12), (32, 32, 24)):
        x, dxe = channel(n, 0.05)
        jac, dxi, dx = G.gridgeom(x, CHAN)
        j0, xi0 = exact_metrics(dxe)
        errs.append((relmax(dx, dxe), relmax(jac, j0), relmax(dxi, xi0)))
        gcl = G.metric_identity(jac, dxi, CHAN)
        assert np.abs(gcl).max() < 1e-12 * np.abs(jac[None, None] * dxi).max(), n
    assert all(errs[1][i] < errs[0][i] / 2 ** 2.5 for i in range(3)), errs


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("n", [12, 40, 64, 100, 256, 320])
def test_gpu_boundary_lines_match_oracle(n):
    """astr_hip_line_batch_nt (derivative and filter, ntype 1..4) equals the
    oracle on random lines to 1e-13 of the line's scale."""
    print(f"line_batch_nt n={n}", flush=True)
    f = np.random.default_rng(100 + n).standard_normal((37, n + 11))
    for nt in (1, 2, 3, 4):
        got = A.line_batch_nt(0, nt, f)
        ref = np.array([O.df_compact_nt(x, nt) for x in f])
        assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max(), (n, nt)
        got = A.line_batch_nt(1, nt, f)
        ref = np.array([O.compact_filter_nt(x, nt) for x in f])
        assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max(), (n, nt)
        # a physical end of the filter keeps its input value exactly
        if nt in (1, 4):
            assert np.array_equal(got[:, 0], f[:, 5])
        if nt in (2, 4):
            assert np.array_equal(got[:, n], f[:, n + 5])


@pytest.mark.gpu
def test_gpu_boundary_lines_polynomials_at_the_longest_line():
    """At the longest line the engine takes (569) and on many lines: the
    derivative with physical ends is exact on cubics and the filter keeps
    quintics, the closed forms of the oracle's tests."""
    n = 569
    rows_d, rows_f, want_d, want_f = [], [], [], []
    for s in range(64):
        f3, d3 = poly_line(n, 3, 200 + s)
        f5, _ = poly_line(n, 5, 300 + s)
        rows_d.append(f3)
        want_d.append(d3)
        rows_f.append(f5)
        want_f.append(f5[5:n + 6])
    for nt in (1, 2, 4):
        # rounding is relative to the values of f, O(1), not to f' ~ 1/n
        got = A.line_batch_nt(0, nt, np.array(rows_d))
        assert np.abs(got - np.array(want_d)).max() < 1e-13 * np.abs(rows_d).max(), nt
        got = A.line_batch_nt(1, nt, np.array(rows_f))
        assert np.abs(got - np.array(want_f)).max() < 1e-13 * np.abs(want_f).max(), nt


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(16, 16, 12), (20, 18, 0)])
def test_gpu_gridgeom_channel_matches_oracle(n):
    """astr_hip_gridgeom_nt with j wall-bounded (npdc 3, 4, 3) on the
    stretched channel grid (3-D) and a 2-D section of it.  The metrics of
    this grid are ill-conditioned at the 1e-12 level (the 18 conservative
    terms cancel): the bound is twice what a one-ulp perturbation of x moves
    the oracle's own metrics by, a boundary-row mistake moves them by 1e-3."""
    print(f"gridgeom_nt {n}", flush=True)
    if n[2] == 0:
        x3, _ = channel((n[0], n[1], 12), 0.06)
        x = np.stack([x3[0, 3], x3[1, 3], 0.0 * x3[2, 3]])[:, None]
    else:
        x, _ = channel(n, 0.06)
    f = A.Flow(A.tgv_config(*n))
    jac, dxi = f.gridgeom(x, CHAN)
    j0, xi0, _ = G.gridgeom(x, CHAN)
    xp = x * (1 + 2.2e-16 * np.random.default_rng(0).standard_normal(x.shape))
    _, xi1, _ = G.gridgeom(xp, CHAN)
    assert relmax(jac, j0) < 1e-13
    for a in range(3 if n[2] else 2):
        scale = np.abs(xi0[a, a]).max()
        tol = 2 * max(np.abs(xi1[a] - xi0[a]).max(), 1e-13 * scale)
        for b in range(3):
            assert np.abs(dxi[a, b] - xi0[a, b]).max() <= tol, (a, b, tol / scale)
    with pytest.raises(A.AstrError):
        f.gridgeom(x, (1, 3, 3))

"""Grid metrics (gridgeom, geom.F90:99-692): the oracle's restatement
pinned by closed forms and the discrete metric identities, the device
version (astr_hip_gridgeom) against the oracle, and astr_hip_set_metrics
accepting exactly the uniform Cartesian metrics the passes support."""
import numpy as np
import pytest

import astr_amd as A
from oracle import geom as G


def grid(n, amp, lens=(2 * np.pi, 2 * np.pi, 2 * np.pi)):
    """A smooth periodic curvilinear grid, x_d(xi + N e_d) = x_d + L_d:
    x1 = xi1 h1 + a sin(2 pi xi2/N2) sin(2 pi xi3/N3),
    x2 = xi2 h2 + a sin(2 pi xi1/N1),
    x3 = xi3 h3 + a sin(2 pi xi1/N1) cos(2 pi xi2/N2),
    with its exact d x_m / d xi_d (per unit index)."""
    n1, n2, n3 = n
    h = [lens[0] / n1, lens[1] / n2, lens[2] / n3]
    k3, k2, k1 = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    w1, w2, w3 = 2 * np.pi / n1, 2 * np.pi / n2, 2 * np.pi / n3
    s1, c1 = np.sin(w1 * k1), np.cos(w1 * k1)
    s2, c2 = np.sin(w2 * k2), np.cos(w2 * k2)
    s3, c3 = np.sin(w3 * k3), np.cos(w3 * k3)
    x = np.stack([k1 * h[0] + amp * s2 * s3, k2 * h[1] + amp * s1, k3 * h[2] + amp * s1 * c2])
    z = 0.0 * k1
    dx = np.array([[z + h[0], amp * w2 * c2 * s3, amp * w3 * s2 * c3],
                   [amp * w1 * c1, z + h[1], z],
                   [amp * w1 * c1 * c2, -amp * w2 * s1 * s2, z + h[2]]])
    return x, dx


def exact_metrics(dx):
    m = np.moveaxis(dx, (0, 1), (-2, -1))
    return np.linalg.det(m), np.moveaxis(np.linalg.inv(m), (-2, -1), (0, 1))


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_oracle_uniform_grid_exact():
    """Uniform Cartesian grid of unequal spacings: dxi = diag(1/dx, 1/dy,
    1/dz) and J = dx dy dz to rounding (the 643c derivative is exact on
    linear coordinates)."""
    x, _ = grid((16, 12, 14), 0.0, lens=(3.0, 2.0, 5.0))
    jac, dxi, _ = G.gridgeom(x)
    h = np.array([3.0 / 16, 2.0 / 12, 5.0 / 14])
    assert np.abs(jac / h.prod() - 1).max() < 1e-14
    for a in range(3):
        for b in range(3):
            want = 1.0 / h[a] if a == b else 0.0
            assert np.abs(dxi[a, b] - want).max() < 1e-13 / h.min(), (a, b)


def test_oracle_curvilinear_metrics_converge():
    """J and dxi against the exact metrics of the curvilinear grid, with the
    6th-order convergence of the compact derivative."""
    errs = []
    for n in (16, 32):
        x, dx_exact = grid((n, n, n), 0.05)
        jac, dxi, dx = G.gridgeom(x)
        j0, xi0 = exact_metrics(dx_exact)
        errs.append((relmax(dx, dx_exact), relmax(jac, j0), relmax(dxi, xi0)))
    assert all(e < 1e-7 for e in errs[1]), errs
    assert all(errs[1][i] < errs[0][i] / 40 for i in range(3)), errs


def test_oracle_metric_identities_hold_to_rounding(monkeypatch):
    """The symmetric conservative form satisfies the discrete metric
    identities sum_r D_r(J dxi(r,m)) = 0 to rounding (derivatives along
    different directions commute); one of the 18 terms with its two
    coordinates swapped (a sign error planted in the oracle) breaks them."""
    x, _ = grid((16, 14, 12), 0.08)
    jac, dxi, _ = G.gridgeom(x)
    gcl = G.metric_identity(jac, dxi)
    scale = np.abs(jac[None, None] * dxi).max()
    assert np.abs(gcl).max() < 1e-13 * scale
    terms = list(G.TERMS)
    d, r, m, a, e, b = terms[7]
    terms[7] = (d, r, m, b, e, a)
    monkeypatch.setattr(G, "TERMS", terms)
    jb, xb, _ = G.gridgeom(x)
    assert np.abs(G.metric_identity(jb, xb)).max() > 1e-5 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(16, 14, 12), (20, 18, 0)])
def test_gpu_gridgeom_matches_oracle(n):
    """astr_hip_gridgeom on the curvilinear grid (3-D) and a 2-D grid."""
    if n[2] == 0:
        # 2-D: x1 = xi1 h1 + a sin(2 pi xi2/N2), x2 = xi2 h2 + a sin(2 pi xi1/N1)
        k2, k1 = np.meshgrid(np.arange(n[1] + 1), np.arange(n[0] + 1), indexing="ij")
        h1, h2 = 2 * np.pi / n[0], 2 * np.pi / n[1]
        x = np.stack([k1 * h1 + 0.08 * np.sin(2 * np.pi * k2 / n[1]),
                      k2 * h2 + 0.08 * np.sin(2 * np.pi * k1 / n[0]), 0.0 * k1])[:, None]
    else:
        x, _ = grid(n, 0.08)
    f = A.Flow(A.tgv_config(*n))
    jac, dxi = f.gridgeom(x)
    j0, xi0, _ = G.gridgeom(x)
    assert relmax(jac, j0) < 1e-13
    for a in range(3):
        for b in range(3):
            ref = xi0[a, b]
            scale = max(np.abs(xi0[a, a]).max(), 1e-300)
            assert np.abs(dxi[a, b] - ref).max() <= 1e-13 * scale, (a, b)


@pytest.mark.gpu
def test_gpu_set_metrics_uniform_only():
    """Metrics of the uniform grid are accepted and leave the RHS unchanged;
    curvilinear metrics are rejected with an error."""
    n = (16, 14, 12)
    cfg = A.tgv_config(*n)
    f = A.Flow(cfg)
    f.init_tgv(1e-3, 5)
    q = f.download()
    r0 = f.rhscal()
    x, _ = grid(n, 0.0)
    jac, dxi = f.gridgeom(x)
    f.set_metrics(jac, dxi)
    f.upload(q)
    r1 = f.rhscal()
    # 1/dx is now the mean of d xi/d x, which the compact derivative of the
    # linear coordinates leaves within delta of n/L; each directional term of
    # the RHS scales by (1 + delta), and the energy equation's fluxes (E + p) u
    # ~ 250 stand against an RHS of ~5
    h = [2 * np.pi / m for m in n]
    delta = max(abs(dxi[d, d].mean() * h[d] - 1.0) for d in range(3))
    assert delta < 1e-13
    for m in range(5):
        flux = 300.0 if m == 4 else 3.0
        assert np.abs(r1[m] - r0[m]).max() <= (4 * delta + 1e-15) * flux * 3 / min(h), m
    xc, _ = grid(n, 0.08)
    jc, dc = f.gridgeom(xc)
    with pytest.raises(A.AstrError):
        f.set_metrics(jc, dc)

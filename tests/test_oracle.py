"""Pinning the oracle (CPU only).

The C restatement (oracle/astr_oracle.c) is checked against
  * the reference's own Fortran, built from the reference sources into
    oracle/_ref (when that build is present) and through the committed
    fixtures tests/golden/*.npz that build wrote;
  * the reference-held history miniapps/tgv_solver_3d/state.ref_128;
  * what the mathematics fixes independently of both: a dense solve of the
    same tridiagonal systems, exactness on polynomials, the modified
    wavenumber of the 6th-order compact scheme (Lele 1992), constant
    preservation and the transfer function of the 10th-order filter
    (Gaitonde & Visbal 2000), KE(0)=1/8 of the TGV field, and the published
    Philox4x32-10 known-answer vectors.
Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed coefficient) fails it.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
HM = 5
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")


def relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


# ----------------------------------------------------------------- lines
def dense_deriv(n, f):
    """Assemble the (n+3)-row system of derivative.F90:63-158,210-306 densely
    (explicit rows at -1 and n+1, compact interior) and solve it."""
    m = n + 3
    A = np.eye(m)
    for u in range(1, m - 1):
        A[u, u - 1] = A[u, u + 1] = 1.0 / 3.0
    F = lambda t: f[t + HM]  # noqa: E731
    d = np.zeros(m)
    for u in range(m):
        r = u - 1
        if u in (0, m - 1):
            d[u] = 0.75 * (F(r + 1) - F(r - 1)) - 0.15 * (F(r + 2) - F(r - 2)) + (F(r + 3) - F(r - 3)) / 60.0
        else:
            d[u] = 7.0 / 9.0 * (F(r + 1) - F(r - 1)) + 1.0 / 36.0 * (F(r + 2) - F(r - 2))
    return np.linalg.solve(A, d)[1:n + 2]


def filter_coefs(alfa=0.49, bh=1.11):
    c10 = np.zeros(6)
    ch = np.zeros(33)
    O.lib().or_filter_coefs(alfa, bh, 1.09, c10.ctypes.data_as(O._dp), ch.ctypes.data_as(O._dp))
    return c10, ch.reshape(3, 11)


def dense_filter(n, f, alfa=0.49, bh=1.11):
    m = n + 7
    A = np.eye(m)
    for u in range(1, m - 1):
        A[u, u - 1] = A[u, u + 1] = alfa
    A[0, 1] = bh       # row -3: c = beta_halo
    A[m - 1, m - 2] = bh  # row n+3: a = beta_halo
    c10, ch = filter_coefs(alfa, bh)
    F = lambda t: f[t + HM]  # noqa: E731
    d = np.zeros(m)
    for u in range(m):
        r = u - 3
        if r < 0:
            d[u] = sum(ch[r + 3, j] * F(-5 + j) for j in range(11))
        elif r > n:
            d[u] = sum(ch[n + 3 - r, j] * F(n + 5 - j) for j in range(11))
        else:
            d[u] = 2 * c10[0] * F(r) + sum(c10[k] * (F(r + k) + F(r - k)) for k in range(1, 6))
    return np.linalg.solve(A, d)[3:n + 4]


@pytest.mark.parametrize("n", [8, 17, 64])
def test_deriv_matches_dense_solve(n):
    f = np.random.default_rng(n).standard_normal(n + 11)
    assert relmax(O.df_compact(f), dense_deriv(n, f)) < 1e-13


@pytest.mark.parametrize("n", [8, 17, 64])
def test_filter_matches_dense_solve(n):
    f = np.random.default_rng(n + 1).standard_normal(n + 11)
    assert relmax(O.compact_filter(f), dense_filter(n, f)) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_deriv_exact_on_polynomials(p):
    n = 40
    x = (np.arange(-HM, n + HM + 1) - 20) / 40.0  # unit spacing 1/40
    f = x ** p
    got = O.df_compact(f) / (1 / 40.0)
    exact = p * x[HM:n + HM + 1] ** (p - 1)
    assert np.abs(got - exact).max() < 1e-9


def test_deriv_order_fails_at_degree_7():
    n = 40
    x = (np.arange(-HM, n + HM + 1) - 20) / 40.0
    got = O.df_compact(x ** 7) * 40.0
    assert np.abs(got - 7 * x[HM:n + HM + 1] ** 6).max() > 1e-9


def test_deriv_modified_wavenumber():
    n = 64
    h = 2 * np.pi / n
    for k in (3, 10, 20):
        t = np.arange(-HM, n + HM + 1)
        f = np.sin(k * t * h)
        df = O.df_compact(f)
        kh = k * h
        kmod = (14.0 / 9.0 * np.sin(kh) + 1.0 / 18.0 * np.sin(2 * kh)) / (1 + 2.0 / 3.0 * np.cos(kh))
        mid = slice(n // 4, 3 * n // 4)
        expect = kmod * np.cos(k * np.arange(n + 1) * h)
        assert np.abs(df[mid] - expect[mid]).max() < 1e-6


def test_filter_preserves_constants():
    for n in (8, 33, 128):
        f = np.full(n + 11, 3.25)
        assert np.abs(O.compact_filter(f) - 3.25).max() < 1e-13


def test_filter_transfer_function():
    n = 96
    alfa = 0.49
    c10, _ = filter_coefs()
    for k in (1, 8, 24, 40):
        w = 2 * np.pi * k / n
        t = np.arange(-HM, n + HM + 1)
        f = np.cos(w * t)
        tf = (2 * c10[0] + sum(2 * c10[m] * np.cos(m * w) for m in range(1, 6))) / (1 + 2 * alfa * np.cos(w))
        ff = O.compact_filter(f)
        mid = slice(3 * n // 8, 5 * n // 8)
        assert np.abs(ff[mid] - tf * np.cos(w * np.arange(n + 1))[mid]).max() < 1e-4
    # T(0)=1, T(pi)=0 for the interior stencil
    assert abs(2 * c10[0] + 2 * sum(c10[1:]) - (1 + 2 * alfa)) < 1e-15
    assert abs(2 * c10[0] + 2 * sum(c10[m] * (-1) ** m for m in range(1, 6))) < 1e-15


def test_lines_golden():
    z = np.load(os.path.join(GOLD, "lines.npz"))
    for key in z.files:
        if not key.startswith("f_"):
            continue
        n = int(key[2:])
        f = z[key]
        d = np.array([O.df_compact(x) for x in f])
        fl = np.array([O.compact_filter(x) for x in f])
        assert np.array_equal(d, z[f"deriv_{n}"]), n
        assert np.array_equal(fl, z[f"filter_{n}"]), n


@needs_ref
def test_lines_vs_reference_build():
    rng = np.random.default_rng(5)
    for n in (9, 40, 100):
        f = rng.standard_normal((3, n + 11))
        assert np.array_equal(np.array([O.df_compact(x) for x in f]), O.run_ref("line_deriv", f=f)["out"])
        assert np.array_equal(np.array([O.compact_filter(x) for x in f]),
                              O.run_ref("line_filter", f=f, alfa=0.49)["out"])


# ----------------------------------------------------------------- fields
def test_stage_golden():
    z = np.load(os.path.join(GOLD, "stage_12x10x8.npz"))
    s = O.OracleState(O.tgv_cfg(12, 10, 8))
    s.set_q(z["q"])
    assert relmax(s.prims, z["prims"]) < 1e-15
    s.rhscal()
    for m in range(5):
        assert relmax(s.qrhs[m], z["qrhs"][m]) < 1e-14
    for m in range(9):
        assert relmax(s.dvel[m], z["dvel"][m]) < 1e-14
    assert relmax(s.dtmp, z["dtmp"]) < 1e-12
    s.set_q(z["q"])
    s.filterq()
    assert relmax(s.q, z["filterq"]) < 1e-15


def test_rk3_golden():
    z = np.load(os.path.join(GOLD, "rk3_16.npz"))
    s = O.OracleState(O.tgv_cfg(16, dt=2e-3))
    s.set_q(z["q"])
    ke, en = s.rk3(10)
    assert relmax(ke, z["hist"][:, 2]) < 1e-13
    assert relmax(en, z["hist"][:, 3]) < 1e-13
    assert relmax(s.q, z["q_out"]) < 1e-13


@needs_ref
def test_rhscal_vs_reference_build_noncubic():
    s = O.OracleState(O.tgv_cfg(20, 14, 18))
    s.tgv_init(1e-2, 3)
    q = s.q.copy()
    s.set_q(q)
    s.rhscal()
    r = O.run_ref("rhscal", dims=[20, 14, 18], q=q)
    assert max(relmax(s.qrhs[m], r["qrhs"][m]) for m in range(5)) < 1e-14


def test_tgv_initial_kinetic_energy():
    """KE(0) = 1/8 exactly for the unperturbed TGV field (rho=1):
    <u^2+v^2>/2 = 2 * (1/2)(1/2)(1/2) / 2."""
    s = O.OracleState(O.tgv_cfg(32))
    s.tgv_init()
    s.rhscal()
    ke, en = s.stacal()
    assert abs(ke - 0.125) < 1e-14
    assert abs(en - 0.375) < 1e-3   # -> 3/8 as the grid is refined


def _diss_for(n, field, amp=0.1):
    """diss_rate_cal of the oracle on an n^3 periodic lattice with rho = 1,
    p = p_inf (so T = 1 and mu = miucal(1)/Re = 1/Re) and velocity
    amp * field(x, y, z); returns dissipation * Re."""
    cfg = O.tgv_cfg(n)
    s = O.OracleState(cfg)
    x = np.arange(n + 1) * cfg.dx
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u, v, w = (amp * c for c in field(X, Y, Z))
    g, ma = cfg.gamma, cfg.mach
    p0 = 1.0 / (g * ma * ma)
    q = np.stack([np.ones_like(X), u, v, w, p0 / (g - 1.0) + 0.5 * (u * u + v * v + w * w)])
    s.set_q(q)
    s.rhscal()
    assert np.abs(s.prims[5] - 1.0).max() < 1e-12
    return s.diss_rate() * cfg.reynolds


def test_dissipation_closed_forms():
    """diss_rate_cal (statistic.F90:994-1043) on fields with closed forms
    (<.> over the periodic box, mu = 1/Re):
      - Taylor-Green u = A sin x cos y cos z, v = -A cos x sin y cos z:
        s11 = -s22, s13, s23 nonzero, div = 0, <s:s> = 3A^2/8 -> 3A^2/4 mu;
      - u = A sin x (pure compression): s11 = div, s11^2 - div^2/3 ->
        2A^2/3 mu (pins the divergence term);
      - u = A sin y (pure shear): 2 s12^2 = A^2 cos^2 y / 2 -> A^2/2 mu (pins the
        factor 2 and the 1/2 of the off-diagonal rates).
    The compact derivative is 6th order, so the error drops 64x per halving
    of h."""
    tgv = lambda X, Y, Z: (np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z),
                           0.0 * X)
    comp = lambda X, Y, Z: (np.sin(X), 0.0 * X, 0.0 * X)
    shear = lambda X, Y, Z: (np.sin(Y), 0.0 * X, 0.0 * X)
    A2 = 0.1 ** 2
    for field, exact in ((tgv, 0.75 * A2), (comp, 2.0 / 3.0 * A2), (shear, 0.5 * A2)):
        e16 = abs(_diss_for(16, field) / exact - 1.0)
        e32 = abs(_diss_for(32, field) / exact - 1.0)
        assert e32 < 1e-7, (exact, e16, e32)
        assert e32 < e16 / 40.0 or e32 < 1e-13, (exact, e16, e32)


def test_state_ref_128_first_steps():
    """First 6 lines of the reference-held history state.ref_128 (128^3)."""
    ref = np.loadtxt(os.path.join(GOLD, "state_ref_128.txt"))
    s = O.OracleState(O.tgv_cfg(128, dt=1e-3))
    s.tgv_init()
    ke, en = s.rk3(6)
    assert relmax(ke, ref[:6, 2]) < 1e-12
    assert relmax(en, ref[:6, 3]) < 1e-12


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_gauss_moments():
    z = np.array([O.gauss(7, i, j, 0, 0) for i in range(64) for j in range(64)])
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.05


def test_qswap_restatement_properties():
    """qswap's duplicate-plane average (parallel.F90:4882-4883, 5027-5028,
    5168-5169): afterwards every duplicate plane equals its plane 0; along i
    (the first sweep) each pair keeps its rounded sum exactly (2 v = fl(a+b));
    the interior is untouched; a second application changes nothing."""
    rng = np.random.default_rng(3)
    for shape in ((5, 7, 6, 9), (5, 1, 6, 9)):
        q = rng.standard_normal(shape)
        s = O.qswap(q)
        assert np.array_equal(s[..., 0], s[..., -1])
        assert np.array_equal(s[:, :, 0], s[:, :, -1])
        if shape[1] > 1:
            assert np.array_equal(s[:, 0], s[:, -1])
        inner = (slice(None), slice(1, -1) if shape[1] > 1 else slice(None), slice(1, -1), slice(1, -1))
        assert np.array_equal(s[inner], q[inner])
        # the i sweep alone: pair sums kept, on rows the j/k sweeps do not touch
        rows = (slice(None), slice(1, -1) if shape[1] > 1 else slice(None), slice(1, -1))
        assert np.array_equal(2.0 * s[rows + (0,)], q[rows + (0,)] + q[rows + (-1,)])
        assert np.array_equal(O.qswap(s), s)


def _z_uniform_pair(im=24, jm=20, km=12, seed=5):
    """A 2-D state of the oracle (noisy TGV, km = 0) and the 3-D state that
    repeats it on every k-plane of an im x jm x km box (w = 0)."""
    s2 = O.OracleState(O.tgv_cfg(im, jm, 0, dt=2e-3))
    s2.tgv_init(1e-2, seed)
    q2 = s2.q.copy()
    s2.set_q(q2)
    s3 = O.OracleState(O.tgv_cfg(im, jm, km, dt=2e-3))
    s3.set_q(np.repeat(q2, km + 1, axis=1))
    return s2, s3


def test_oracle_2d_equals_z_uniform_3d():
    """The 2-D path (ndir = 2: no z sweeps, z gradients zero) against the
    reference-pinned 3-D path on a field constant along k with w = 0, where
    every z derivative is exactly zero (the 643c right-hand side of a constant
    line vanishes): rhscal's rho, rho u, rho v and E right-hand sides equal
    the 2-D ones bit for bit on every k-plane and the rho w one is zero.  The
    z filter keeps a constant line only to rounding (its beta_h = 1.11 rows
    and coefficient sums), about 20 ulp, so one filterq agrees within 5e-14
    and three RK3 steps and their statistics within 1e-13."""
    s2, s3 = _z_uniform_pair()
    s2.rhscal()
    s3.rhscal()
    for m in (0, 1, 2, 4):
        for k in range(s3.cfg.km + 1):
            assert np.array_equal(s3.qrhs[m, k], s2.qrhs[m, 0]), (m, k)
    assert np.all(s3.qrhs[3] == 0.0)
    assert np.all(s3.dvel[6:] == 0.0) and np.all(s3.dtmp[2] == 0.0)
    assert np.array_equal(s3.dtmp[:2, 5], s2.dtmp[:2, 0])
    s2, s3 = _z_uniform_pair()
    s2.filterq()
    s3.filterq()
    for m in (0, 1, 2, 4):
        assert relmax(s3.q[m], np.repeat(s2.q[m], s3.cfg.km + 1, axis=0)) < 5e-14, m
    s2, s3 = _z_uniform_pair()
    ke2, en2, di2 = s2.rk3(3, dissipation=True)
    ke3, en3, di3 = s3.rk3(3, dissipation=True)
    assert relmax(ke3, ke2) < 1e-13 and relmax(en3, en2) < 1e-13 and relmax(di3, di2) < 1e-13
    for m in (0, 1, 2, 4):
        assert relmax(s3.q[m], np.repeat(s2.q[m], s3.cfg.km + 1, axis=0)) < 1e-13, m
    assert np.abs(s3.q[3]).max() < 1e-13 * np.abs(s3.q[1]).max()


def _diss2d_for(n, field, amp=0.1):
    """diss_rate_cal of the 2-D oracle (km = 0) on an n^2 periodic lattice,
    rho = 1, T = 1 (mu = 1/Re), velocity amp * field(x, y); returns
    dissipation * Re."""
    cfg = O.tgv_cfg(n, n, 0)
    s = O.OracleState(cfg)
    x = np.arange(n + 1) * cfg.dx
    Y, X = np.meshgrid(x, x, indexing="ij")
    u, v = (amp * c for c in field(X, Y))
    g, ma = cfg.gamma, cfg.mach
    p0 = 1.0 / (g * ma * ma)
    q = np.stack([np.ones_like(X), u, v, 0.0 * X, p0 / (g - 1.0) + 0.5 * (u * u + v * v)])[:, None]
    s.set_q(q)
    s.rhscal()
    return s.diss_rate() * cfg.reynolds


def test_dissipation_closed_forms_2d():
    """diss_rate_cal on the 2-D path: shear u = A sin y gives A^2/2 mu (the
    factor 2 and the 1/2 of the off-diagonal rate), compression u = A sin x
    gives 2A^2/3 mu (the -div^2/3 term), each converging at 6th order."""
    A2 = 0.1 ** 2
    shear = lambda X, Y: (np.sin(Y), 0.0 * X)
    comp = lambda X, Y: (np.sin(X), 0.0 * X)
    for field, exact in ((shear, 0.5 * A2), (comp, 2.0 / 3.0 * A2)):
        e16 = abs(_diss2d_for(16, field) / exact - 1.0)
        e32 = abs(_diss2d_for(32, field) / exact - 1.0)
        assert e32 < 1e-7, (exact, e16, e32)
        assert e32 < e16 / 40.0 or e32 < 1e-13, (exact, e16, e32)


def test_qswap_corner_order():
    """qswap averages along i, then j, then k (parallel.F90:4882-4883,
    5027-5028, 5168-5169), so the value it leaves at the eight corners is
    ((c000+cn00)/2 + (c0n0+cnn0)/2)/2 ... with the i pairs innermost.  On this
    field the k-innermost order rounds differently; the restatement must give
    the i-innermost value bit for bit."""
    rng = np.random.default_rng(11)
    n = 6
    for trial in range(200):
        q = rng.standard_normal((5, n + 1, n + 1, n + 1)) * np.exp(rng.uniform(-20, 20, (5, 1, 1, 1)))
        c = q[:, [0, -1]][:, :, [0, -1]][:, :, :, [0, -1]]   # c[m, k, j, i]
        avg = lambda a, b: 0.5 * (a + b)
        ijk = avg(avg(avg(c[:, 0, 0, 0], c[:, 0, 0, 1]), avg(c[:, 0, 1, 0], c[:, 0, 1, 1])),
                  avg(avg(c[:, 1, 0, 0], c[:, 1, 0, 1]), avg(c[:, 1, 1, 0], c[:, 1, 1, 1])))
        kji = avg(avg(avg(c[:, 0, 0, 0], c[:, 1, 0, 0]), avg(c[:, 0, 1, 0], c[:, 1, 1, 0])),
                  avg(avg(c[:, 0, 0, 1], c[:, 1, 0, 1]), avg(c[:, 0, 1, 1], c[:, 1, 1, 1])))
        if not np.array_equal(ijk, kji):
            break
    else:
        pytest.fail("no field with order-dependent corners found")
    s = O.qswap(q)
    for kk in (0, -1):
        for jj in (0, -1):
            for ii in (0, -1):
                assert np.array_equal(s[:, kk, jj, ii], ijk)
    assert not np.array_equal(s[:, 0, 0, 0], kji)

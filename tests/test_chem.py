"""Chemical source terms: the numpy restatement (oracle/chem.py) pinned on the
CPU, and the HIP kernel checked against it on the GPU.

Cantera 2.5.1, which the reference calls (src/thermchem.F90:1553-1575), is
not available and no reference fixture pins its outputs, so parity with
Cantera itself is unpinned; the restatement is pinned by conservation laws,
thermodynamic consistency, a published rate constant, and detailed balance
at an equilibrium state computed by an independent element-potential solve.
"""
import os

import numpy as np
import pytest
from scipy.optimize import fsolve

from astr_amd import chem as AC
from oracle import chem as OC

MECH = os.path.join(os.path.dirname(__file__), "golden", "Jachimowski88_H2-9-19.xml")


@pytest.fixture(scope="module")
def mech():
    return OC.Mechanism(MECH)


def test_mechanism_contents(mech):
    assert mech.names == ["H2", "O2", "H2O", "O", "OH", "H", "HO2", "H2O2", "N2"]
    assert mech.nr == 19
    assert sum(r["tb"] for r in mech.rx) == 6
    assert mech.p0 == 1.0e5
    assert abs(mech.W[mech.names.index("H2O")] - 18.015) < 1e-12


def test_conservation(mech):
    T, p, Y = AC.sample_cells(mech.names, mech.W, 20000, 3)
    wd, h, cp, hk = mech.rates(T, p, Y)
    scale = np.abs(wd).max(axis=0) + 1e-300
    assert np.all(np.abs(wd.sum(axis=0)) <= 1e-13 * scale)
    for e in "HON":
        a = np.array([mech.comp[s].get(e, 0.0) / mech.W[k] for k, s in enumerate(mech.names)])
        assert np.all(np.abs((a[:, None] * wd).sum(axis=0)) <= 1e-13 * scale / 1.0)
    assert np.all(wd[mech.names.index("N2")] == 0.0)


def test_thermo_consistency(mech):
    """cp = dh/dT and cp/T = ds/dT (central differences), and the two NASA
    ranges meet at T_mid: catches a mis-indexed or mis-assigned coefficient."""
    T = np.array([500.0, 800.0, 1400.0, 2500.0])
    dT = 1e-3
    cpr, hrt, sr = mech.thermo(T)
    _, hp, sp_ = mech.thermo(T + dT)
    _, hm, sm = mech.thermo(T - dT)
    dh = ((hp * (T + dT)[:, None]) - (hm * (T - dT)[:, None])) / (2 * dT)  # d(h/R)/dT
    ds = (sp_ - sm) / (2 * dT)
    assert np.abs(dh - cpr).max() < 1e-6
    assert np.abs(ds - cpr / T[:, None]).max() < 1e-8
    lo = mech.thermo(np.array([1000.0]))
    hi_c = mech.coeffs(np.array([1000.0 + 1e-9]))
    cp_hi = (hi_c[..., 0] + hi_c[..., 1] * 1000.0 + hi_c[..., 2] * 1e6 + hi_c[..., 3] * 1e9
             + hi_c[..., 4] * 1e12)
    assert np.abs(cp_hi - lo[0]).max() < 5e-3 * np.abs(lo[0]).max()


def test_temperature_inverts_internal_energy(mech):
    """e(T) and its Newton inverse: round trip from guesses 500 K off, and
    cv = de/dT (central differences) so the Newton slope is the true one."""
    T, p, Y = AC.sample_cells(mech.names, mech.W, 200, 17)
    e = mech.e_mass(T, Y)
    T1 = mech.temperature(e, Y, T + 500.0)
    assert np.abs(T1 / T - 1).max() < 1e-13
    cpr, _, _ = mech.thermo(T)
    cv = OC.R_GAS * ((Y / mech.W[:, None]) * (cpr.T - 1.0)).sum(axis=0)
    de = (mech.e_mass(T + 1e-3, Y) - mech.e_mass(T - 1e-3, Y)) / 2e-3
    assert np.abs(de / cv - 1).max() < 1e-8


def test_mixture_helpers(mech):
    T, p, Y = AC.sample_cells(mech.names, mech.W, 50, 18)
    X = mech.convertxiyi(Y, "Y2X")
    assert np.abs(X.sum(axis=0) - 1).max() < 1e-15
    assert np.abs(mech.convertxiyi(X, "X2Y") - Y).max() < 1e-15
    # R_mix = R / W_mix with W_mix = sum X_k W_k
    assert np.abs(mech.rgcmix(Y) * (X * mech.W[:, None]).sum(axis=0) / OC.R_REF - 1).max() < 1e-14


def test_published_rate_constant(mech):
    """H + O2 = OH + O of Jachimowski (1988): A = 2.6e14 cm^3/mol/s,
    E = 16.8 kcal/mol, evaluated with R in cal/mol/K; 1 cm^3/mol = 1e-3 m^3/kmol."""
    T = 1500.0
    Rcal = 8.314462618 / 4.184
    k_pub = 2.6e14 * np.exp(-16800.0 / (Rcal * T)) * 1e-3
    r = mech.rx[1]
    k = r["A"] * T ** r["b"] * np.exp(-r["EaR"] / T)
    assert abs(k / k_pub - 1) < 1e-9


def equilibrium(mech, T, p, X0):
    """Chemical equilibrium at (T,p) from the element-potential equations:
    mole fractions x_k = exp(-g_k - ln(p/p0) + sum_e lambda_e a_ek) with
    sum x_k = 1 and the element totals of X0 fixed (moles N of mixture)."""
    els = "HON"
    A = np.array([[mech.comp[s].get(e, 0.0) for s in mech.names] for e in els])
    b = A @ X0
    _, hrt, sr = mech.thermo(np.array([T]))
    g = (hrt - sr)[0] + np.log(p / mech.p0)

    def res(z):
        lam, lnN = z[:3], z[3]
        x = np.exp(-g + A.T @ lam)
        return np.concatenate([np.log(np.exp(lnN) * (A @ x)) - np.log(b), [np.log(x.sum())]])

    # start from the potentials that reproduce a major-species guess
    xg = np.maximum(X0, 1e-6)
    xg = xg / xg.sum()
    lam0 = np.linalg.lstsq(A.T, g + np.log(xg), rcond=None)[0]
    z = fsolve(res, np.concatenate([lam0, [0.0]]), xtol=1e-15)
    for _ in range(3):
        z = fsolve(res, z, xtol=1e-15)
    assert np.abs(res(z)).max() < 1e-12
    x = np.exp(-g + A.T @ z[:3])
    return x / x.sum()


def test_detailed_balance_at_equilibrium(mech):
    T, p = 2200.0, 101325.0
    ix = {s: k for k, s in enumerate(mech.names)}
    X0 = np.zeros(mech.ns)
    X0[ix["H2"]], X0[ix["O2"]], X0[ix["N2"]] = 2.0, 1.0, 3.76
    X = equilibrium(mech, T, p, X0 / X0.sum())
    Y = X * mech.W / (X * mech.W).sum()
    wd, _, _, _, fwd, rev = mech.rates(np.array([T]), np.array([p]), Y[:, None], parts=True)
    ratio = rev[:, 0] / fwd[:, 0]
    assert np.abs(ratio - 1.0).max() < 1e-7, ratio


@pytest.mark.gpu
def test_gpu_rates_match_oracle(mech):
    gm = AC.Mechanism(MECH)
    assert gm.names == mech.names and np.allclose(gm.W, mech.W, rtol=0, atol=0)
    T, p, Y = AC.sample_cells(mech.names, mech.W, 65536, 11)
    wd, h, cp = gm.rates(T, p, Y)
    wd0, h0, cp0, _ = mech.rates(T, p, Y)
    for k in range(mech.ns):
        den = np.abs(wd0[k]).max()
        if den == 0:
            assert np.all(wd[k] == 0)
        else:
            assert np.abs(wd[k] - wd0[k]).max() / den < 1e-12, mech.names[k]
    assert np.abs(h - h0).max() / np.abs(h0).max() < 1e-13
    assert np.abs(cp / cp0 - 1).max() < 1e-13


@pytest.mark.gpu
def test_gpu_temperature_calc(mech):
    gm = AC.Mechanism(MECH)
    T, p, Y = AC.sample_cells(mech.names, mech.W, 65536, 19)
    e = mech.e_mass(T, Y)
    T1 = AC.temperature(gm, Y, e, np.full(T.size, 1000.0))
    assert np.abs(T1 / T - 1).max() < 1e-12


@pytest.mark.gpu
def test_gpu_rates_density_input_and_conservation(mech):
    gm = AC.Mechanism(MECH)
    T, p, Y = AC.sample_cells(mech.names, mech.W, 1 << 20, 5)
    rho = p / (OC.R_GAS * T * (Y / mech.W[:, None]).sum(axis=0))
    wd, h, cp = gm.rates(T, rho, Y, input_is_p=False)
    wdp, _, _ = gm.rates(T, p, Y, input_is_p=True)
    assert np.abs(wd - wdp).max() / np.abs(wdp).max() < 1e-12
    scale = np.abs(wd).max(axis=0) + 1e-300
    assert np.all(np.abs(wd.sum(axis=0)) <= 1e-13 * scale)
    assert np.all(wd[mech.names.index("N2")] == 0.0)


def srccomb_block(mech, im, jm, km, hm, seed):
    """A block in the reference's shapes (C order, reversed indices): every
    point, halo included, carries its own (rho, T, Y), so a mis-indexed
    point or species reads another state."""
    sh = (km + 1 + 2 * hm, jm + 1 + 2 * hm, im + 1 + 2 * hm)
    n = int(np.prod(sh))
    T, p, Y = AC.sample_cells(mech.names, mech.W, n, seed)
    rho = p / (OC.R_GAS * T * (Y / mech.W[:, None]).sum(axis=0))
    rng = np.random.default_rng(seed)
    jac = rng.uniform(0.5, 2.0, n)
    qrhs = rng.standard_normal((5 + mech.ns, km + 1, jm + 1, im + 1))
    return rho.reshape(sh), T.reshape(sh), Y.reshape((mech.ns,) + sh), jac.reshape(sh), qrhs


def srccomb_expected(mech, rho, T, Y, jac, qrhs, hm):
    """srccomb (src/solver.F90:517-525) written as its loop over the computed
    points, each chemrate from the oracle."""
    out = qrhs.copy()
    nz, ny, nx = qrhs.shape[1:]
    for k in range(nz):
        for j in range(ny):
            a, b, c = k + hm, j + hm, slice(hm, hm + nx)
            wd, _, _, _ = mech.rates(T[a, b, c], rho[a, b, c], Y[:, a, b, c], input_is_p=False)
            out[5:, k, j, :] = out[5:, k, j, :] + wd * jac[a, b, c]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(11, 9, 7, 2), (13, 10, 0, 3), (6, 5, 4, 0)])
def test_gpu_srccomb_matches_oracle(mech, dims):
    """srccomb through the host entry point and the device one: the species
    rows gain wirate * jacob at every computed point (oracle to 1e-12 of
    each species' largest rate), the flow rows stay bit for bit, and the
    host and device forms agree bit for bit."""
    from astr_amd import hip as H
    im, jm, km, hm = dims
    gm = AC.Mechanism(MECH)
    rho, T, Y, jac, q0 = srccomb_block(mech, im, jm, km, hm, 31 + im)
    want = srccomb_expected(mech, rho, T, Y, jac, q0, hm)
    got = gm.srccomb(rho, T, Y, jac, q0.copy(), hm)
    assert np.array_equal(got[:5], q0[:5])
    for s in range(mech.ns):
        d = want[5 + s] - q0[5 + s]
        scale = max(np.abs(d).max(), 1e-300)
        assert np.abs(got[5 + s] - want[5 + s]).max() <= 1e-12 * scale + 4e-16 * np.abs(q0[5 + s]).max(), \
            mech.names[s]
    dev = [H.DeviceArray.from_host(a) for a in (rho, T, Y, jac, q0)]
    gm.srccomb(*dev, hm)
    H.synchronize()
    assert np.array_equal(dev[4].to_host(), got)
    with pytest.raises(AC.AstrError, match="numq"):
        AC._chk(AC._bind().astr_hip_srccomb(gm.h, im, jm, km, hm, 5 + mech.ns + 1,
                                            *(AC._p(a) for a in (rho, T, Y, jac, q0))))
    gm.close()


def _mix(mech, phi):
    ix = {s: k for k, s in enumerate(mech.names)}
    X = np.zeros(mech.ns)
    X[ix["H2"]], X[ix["O2"]], X[ix["N2"]] = 2.0 * phi, 1.0, 3.76
    Y = X * mech.W
    return Y / Y.sum()


def test_reactor_rhs_conserves_enthalpy_and_elements(mech):
    """Along the reactor's trajectory the mixture enthalpy sum_k Y_k h_k(T) and
    the element mass fractions stay constant (pins the sign and scale of dT/dt
    and the mass-fraction rates); the mixture ignites (T rises by > 1000 K)."""
    T0, p = 1200.0, 101325.0
    Y0 = _mix(mech, 1.0)
    y = mech.ignite(T0, p, Y0, 1e-3, rtol=1e-9, atol=1e-14)

    def h_of(T, Y):
        _, hrt, _ = mech.thermo(np.array([T]))
        return float((Y * hrt[0] * OC.R_GAS * T / mech.W).sum())

    h0, h1 = h_of(T0, Y0), h_of(y[0], y[1:])
    assert abs(h1 - h0) < 1e-6 * max(abs(h0), 1e3 * 8314.0 / 29.0)
    for e in "HON":
        a = np.array([mech.comp[s].get(e, 0.0) / mech.W[k] for k, s in enumerate(mech.names)])
        assert abs((a * y[1:]).sum() - (a * Y0).sum()) < 1e-9 * (a * Y0).sum()
    assert y[0] > T0 + 1000.0
    assert abs(y[1:].sum() - 1.0) < 1e-9


IGN = os.path.join(os.path.dirname(__file__), "golden", "ignition_radau.npz")


def _trajectory(mech, T0=1100.0, p=3 * 101325.0, phi=1.0, times=(0.0, 1e-6, 2e-5, 5e-5, 1e-4)):
    """States of one ignition along scipy's Radau solution."""
    from scipy.integrate import solve_ivp

    y0 = np.concatenate([[T0], _mix(mech, phi)])
    sol = solve_ivp(lambda t, y: mech.reactor_rhs(y, p), (0.0, max(times)), y0, method="Radau",
                    rtol=1e-9, atol=1e-14, dense_output=True)
    return p, np.array([sol.sol(t) for t in times])


def test_reactor_jacobian_against_richardson_differences(mech):
    """The restatement's analytic Jacobian of the reactor (reactor_jacobian)
    against Richardson-extrapolated central differences of reactor_rhs along
    an ignition: every column to 1e-5 of its norm (the differences are good to
    about 1e-9 in most columns and 2e-6 in the water column, whose entries
    cancel; a dropped chain-rule term is an O(1) error)."""
    p, Ys = _trajectory(mech)
    for v in Ys:
        J = mech.reactor_jacobian(v, p)
        n = v.size
        for j in range(n):
            h = 1e-5 * max(abs(v[j]), 1e-4 if j else 1.0)

            def cd(hh):
                a = v.copy()
                b = v.copy()
                a[j] += hh
                b[j] -= hh
                return (mech.reactor_rhs(a, p) - mech.reactor_rhs(b, p)) / (2 * hh)

            ref = (4 * cd(h / 2) - cd(h)) / 3
            assert np.abs(J[:, j] - ref).max() <= 1e-5 * np.abs(ref).max(), (j, v[0])


@pytest.mark.gpu
def test_gpu_reactor_rhs_and_jacobian_match_oracle(mech):
    """The device reactor's right-hand side and analytic Jacobian
    (astr_hip_reactor_rhs) against the restatement along an ignition: f to
    1e-12 and every Jacobian column to 1e-10 of their norms."""
    gm = AC.Mechanism(MECH)
    p, Ys = _trajectory(mech)
    f, J = gm.reactor_rhs(np.full(len(Ys), p), Ys, jacobian=True)
    for i, v in enumerate(Ys):
        f0 = mech.reactor_rhs(v, p)
        assert np.abs(f[i] - f0).max() <= 1e-12 * np.abs(f0).max(), i
        Jo = mech.reactor_jacobian(v, p)
        for j in range(v.size):
            assert np.abs(J[i][:, j] - Jo[:, j]).max() <= 1e-10 * np.abs(Jo[:, j]).max(), (i, j)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 16])
def test_gpu_reactor_ignites_in_few_steps(mech, lanes):
    """RODAS4 with the analytic Jacobian at the serving tolerance (rtol 1e-5,
    atol 1e-10): a 1 ms ignition in a few hundred steps with few rejections,
    ending within 1e-4 of scipy's Radau temperature."""
    gm = AC.Mechanism(MECH)
    cases = [(1100.0, 3 * 101325.0, 1.0), (1000.0, 101325.0, 0.8), (1300.0, 5 * 101325.0, 1.5),
             (960.0, 101325.0, 0.6)]
    pool = AC.ReactorPool(gm, len(cases), rtol=1e-5, atol=1e-10, lanes=lanes)
    for k, (T0, p, phi) in enumerate(cases):
        pool.set(k, T0, p, _mix(mech, phi), 1e-3)
    st = pool.run(max_tries=5000)
    assert np.all(st == AC.ReactorPool.DONE)
    for k, (T0, p, phi) in enumerate(cases):
        t, T, Y, nstep, nrej = pool.get(k)
        ref = mech.ignite(T0, p, _mix(mech, phi), 1e-3, rtol=1e-9, atol=1e-14)
        assert nstep < 600 and nrej < nstep // 4, (k, nstep, nrej)
        assert abs(T / ref[0] - 1) < 1e-4, (k, T, ref[0])


def test_ignition_fixture_matches_oracle(mech):
    """The stored Radau reference (tests/golden/make_ignition.py) is what the
    current restatement gives: case 1 recomputed at rtol 1e-9."""
    z = np.load(IGN)
    T0, p, phi = z["cases"][1]
    y = mech.ignite(T0, p, _mix(mech, phi), 1e-3, rtol=1e-9, atol=1e-16)
    assert abs(y[0] / z["y_end"][1, 0] - 1) < 1e-10
    assert np.abs(y[1:] - z["y_end"][1, 1:]).max() < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 16])
def test_gpu_reactor_converges_to_radau(mech, lanes):
    """RODAS4 on the GPU at rtol 1e-7 against scipy's Radau at 1e-10 (stored)
    for lean, stoichiometric and rich mixtures at 1 and 5 atm, to t = 1 ms,
    with one thread and with a team of 16 lanes per reactor."""
    z = np.load(IGN)
    gm = AC.Mechanism(MECH)
    cases = [tuple(c) for c in z["cases"]]
    pool = AC.ReactorPool(gm, len(cases), rtol=1e-7, atol=1e-13, lanes=lanes)
    assert pool.lanes == lanes
    for k, (T0, p, phi) in enumerate(cases):
        pool.set(k, T0, p, _mix(mech, phi), 1e-3)
    st = pool.run(max_tries=2000)
    assert np.all(st == AC.ReactorPool.DONE), st
    for k in range(len(cases)):
        t, T, Y, nstep, nrej = pool.get(k)
        assert t == 1e-3 and nstep > 10
        ref = z["y_end"][k]
        assert abs(T / ref[0] - 1) < 1e-4, (k, T, ref[0])
        assert np.abs(Y - ref[1:]).max() < 1e-5, k


@pytest.mark.gpu
def test_gpu_reactor_team_agrees_with_one_thread(mech):
    """The team kernel (16 lanes per reactor: the reactions of each
    right-hand side split over the lanes and the species rates summed over
    the team, Jacobian columns on separate lanes, row-parallel elimination)
    and the one-thread kernel integrate the same ignitions to within the
    integrator's tolerance (the team's sums round differently, so the step
    sequences may part at the rounding level); two slots stay free."""
    gm = AC.Mechanism(MECH)
    rng = np.random.default_rng(11)
    cases = [(rng.uniform(1150.0, 1400.0), rng.uniform(1.0, 10.0) * 101325.0, rng.uniform(0.5, 2.0))
             for _ in range(6)]
    out = {}
    for lanes in (1, 16):
        pool = AC.ReactorPool(gm, 8, rtol=1e-6, atol=1e-12, lanes=lanes)
        for k, (T0, p, phi) in enumerate(cases):
            pool.set(k, T0, p, _mix(mech, phi), 2e-4)
        st = pool.run(max_tries=500)
        assert np.all(st[:6] == AC.ReactorPool.DONE) and np.all(st[6:] == AC.ReactorPool.FREE)
        out[lanes] = [pool.get(k) for k in range(6)]
        pool.close()
    for a, b in zip(out[1], out[16]):
        assert a[3] > 5 and b[3] > 5
        assert a[0] == b[0]                                   # both reached t_end
        assert abs(b[2][0] / a[2][0] - 1.0) < 1e-5, (a[2][0], b[2][0])
        assert np.abs(b[2][1:] - a[2][1:]).max() < 1e-6
        assert abs(a[3] - b[3]) <= max(2, a[3] // 20)         # about the same number of steps


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 16])
def test_gpu_reactor_same_alone_and_in_a_mix(mech, lanes):
    """A reactor's result is bit-identical alone, among 63 others, and split
    into launches of a different length."""
    gm = AC.Mechanism(MECH)
    rng = np.random.default_rng(4)
    target = (1100.0, 3 * 101325.0, _mix(mech, 0.8))
    alone = AC.ReactorPool(gm, 1, rtol=1e-6, atol=1e-12, lanes=lanes)
    alone.set(0, target[0], target[1], target[2], 1e-3)
    alone.run(max_tries=1000)
    mixed = AC.ReactorPool(gm, 64, rtol=1e-6, atol=1e-12, lanes=lanes)
    for k in range(64):
        if k == 37:
            mixed.set(k, target[0], target[1], target[2], 1e-3)
        else:
            mixed.set(k, rng.uniform(950, 1400), rng.uniform(1, 10) * 101325.0,
                      _mix(mech, rng.uniform(0.5, 2.0)), 1e-3)
    mixed.run(max_tries=37)
    a, m = alone.get(0), mixed.get(37)
    assert a[0] == m[0] and a[1] == m[1] and np.array_equal(a[2], m[2]) and a[3:] == m[3:]


# Jachimowski (1988) H2-air mechanism, 19 reversible reactions, in the units
# the mechanism is published in: A in cm^3, mol, s (cm^6/mol^2/s for the
# termolecular three-body reactions), E in cal/mol; k_f = A T^b exp(-E/RT).
# Transcribed by hand, reaction by reaction, from the equations and rate
# parameters of chemMech/Jachimowski88_H2-9-19.xml (that file stores A in
# m^3/kmol/s; the conversion here is the test's own).  The publication itself
# (NASA TP-2791) is not available offline, so the table checks how the parser
# reads the mechanism file, not the file against the paper.
# (reactants, products, A, b, E, third-body efficiencies or None)
JACHIMOWSKI_1988 = [
    ("H2 O2", "HO2 H", 1.0e14, 0.0, 56000.0, None),
    ("H O2", "OH O", 2.6e14, 0.0, 16800.0, None),
    ("O H2", "OH H", 1.8e10, 1.0, 8900.0, None),
    ("OH H2", "H2O H", 2.2e13, 0.0, 5150.0, None),
    ("OH OH", "H2O O", 6.3e12, 0.0, 1090.0, None),
    ("H OH", "H2O", 2.2e22, -2.0, 0.0, {"H2O": 6.0}),
    ("H H", "H2", 6.4e17, -1.0, 0.0, {"H2": 2.0, "H2O": 6.0}),
    ("H O", "OH", 6.0e16, -0.6, 0.0, {"H2O": 6.0}),
    ("H O2", "HO2", 2.1e15, 0.0, -1000.0, {"H2": 2.0, "H2O": 6.0}),
    ("HO2 H", "OH OH", 1.4e14, 0.0, 1080.0, None),
    ("HO2 H", "H2O O", 1.0e13, 0.0, 1080.0, None),
    ("HO2 O", "O2 OH", 1.5e13, 0.0, 950.0, None),
    ("HO2 OH", "H2O O2", 8.0e12, 0.0, 0.0, None),
    ("HO2 HO2", "H2O2 O2", 2.0e12, 0.0, 0.0, None),
    ("H H2O2", "H2 HO2", 1.4e12, 0.0, 3600.0, None),
    ("O H2O2", "OH HO2", 1.4e13, 0.0, 6400.0, None),
    ("OH H2O2", "H2O HO2", 6.1e12, 0.0, 1430.0, None),
    ("H2O2", "OH OH", 1.2e17, 0.0, 45500.0, {"H2O": 15.0}),
    ("O O", "O2", 6.0e13, 0.0, -1800.0, {}),
]


def _side(mech, text):
    out = {}
    for s in text.split():
        k = mech.names.index(s)
        out[k] = out.get(k, 0.0) + 1.0
    return out


def test_mechanism_rate_table(mech):
    """Every reaction's stoichiometry, reversibility, rate constant and
    third-body efficiencies as the parser reads them, against the table above:
    k_f(T) at 900, 1500 and 2500 K within 1e-12 relative (cm^3/mol = 1e-3
    m^3/kmol per concentration factor beyond the first, [M] counted as one;
    E in cal/mol with R = 8.31446261815324/4.184 cal/mol/K), efficiencies
    exactly (default 1).  A planted slip in any one (A, b, E) triple, order
    or efficiency breaks it."""
    assert mech.nr == len(JACHIMOWSKI_1988)
    rcal = 8.31446261815324 / 4.184
    for j, (reac, prod, A, b, E, eff) in enumerate(JACHIMOWSKI_1988):
        r = mech.rx[j]
        assert r["rev"], j
        assert dict(r["reac"]) == _side(mech, reac), j
        assert dict(r["prod"]) == _side(mech, prod), j
        assert r["tb"] == (eff is not None), j
        order = sum(_side(mech, reac).values()) + (1 if eff is not None else 0)
        for T in (900.0, 1500.0, 2500.0):
            k_tab = A * 1.0e-3 ** (order - 1) * T ** b * np.exp(-E / (rcal * T))
            k = r["A"] * T ** r["b"] * np.exp(-r["EaR"] / T)
            assert abs(k / k_tab - 1.0) < 1e-12, (j, T, k, k_tab)
        want = np.ones(mech.ns)
        for s, v in (eff or {}).items():
            want[mech.names.index(s)] = v
        assert np.array_equal(r["eff"], want), j


def test_mechanism_rate_table_catches_a_planted_slip(mech):
    """The table check is sensitive: one changed efficiency or activation
    energy in a copy of the parsed mechanism fails it."""
    import copy

    for field, j, bump in (("eff", 8, None), ("EaR", 15, 1.0 + 1e-9), ("A", 17, 1.001)):
        m2 = copy.deepcopy(mech)
        if field == "eff":
            m2.rx[j]["eff"][m2.names.index("H2O")] = 16.0
        else:
            m2.rx[j][field] *= bump
        with pytest.raises(AssertionError):
            test_mechanism_rate_table(m2)

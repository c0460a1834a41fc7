"""Numpy restatement of the chemical source terms (TEST INFRASTRUCTURE ONLY).

The reference evaluates its chemistry through Cantera 2.5.1 (a third-party
C++ library it does not vendor, pinned by script/install_cantera.sh:1-5;
call sites src/thermchem.F90:1553-1575 chemrate, :1342-1428 cpeval,
:1498-1541 enthpy).  Cantera is absent here, so this module restates the
published ideal-gas / GasKinetics algorithm for a CTML mechanism:

  C_k   = rho Y_k / W_k                                   (kmol/m^3)
  cp/R  = a0 + a1 T + a2 T^2 + a3 T^3 + a4 T^4            (NASA 7-coefficient)
  h/RT  = a0 + a1 T/2 + a2 T^2/3 + a3 T^3/4 + a4 T^4/5 + a5/T
  s/R   = a0 ln T + a1 T + a2 T^2/2 + a3 T^3/3 + a4 T^4/4 + a6
  kf    = A T^b exp(-E/(R T));  Kc = exp(-dG0/RT) (p0/(R T))^dnu;  kr = kf/Kc
  q     = ([M])(kf prod C^nu' - kr prod C^nu''),  [M] = sum eff_k C_k
  wdot_k = W_k sum_r (nu''_kr - nu'_kr) q_r               (kg/m^3/s, as chemrate)
  cp_mix = R sum Y_k cp_k/W_k                             (as cpeval, cp_mass)
  h_k    = (h/RT)_k (8314.2/W_k) T, h_mix = sum Y_k h_k   (as enthpy)

with R = 8314.46261815324 J/kmol/K (Cantera 2.5) except in enthpy, which
the reference writes with its own rguniv = 8314.2 (thermchem.F90:26).

Parity status: "parity unpinned" against Cantera itself (not runnable here,
no reference fixture pins its outputs).  This restatement is pinned by
mass and element conservation, the inert N2 row, the NASA polynomials in
long double, and detailed balance at an equilibrium state found by an
independent Gibbs-energy minimisation (tests/test_chem.py).
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

R_GAS = 8314.46261815324
R_REF = 8314.2
ATOMIC_W = {"H": 1.008, "O": 15.999, "N": 14.007, "C": 12.011, "AR": 39.95, "Ar": 39.95,
            "HE": 4.002602, "He": 4.002602}
EA_UNITS = {"cal/mol": 4184.0, "kcal/mol": 4184000.0, "J/mol": 1000.0, "kJ/mol": 1.0e6,
            "J/kmol": 1.0}


def _pairs(text):
    out = []
    for tok in (text or "").split():
        if ":" in tok:
            a, b = tok.split(":")
            out.append((a, float(b)))
    return out


class Mechanism:
    def __init__(self, path, phase_id="gas"):
        root = ET.parse(path).getroot()
        phase = [p for p in root.iter("phase") if p.get("id") == phase_id][0]
        self.names = phase.find("speciesArray").text.split()
        ns = len(self.names)
        idx = {s: k for k, s in enumerate(self.names)}
        self.W = np.zeros(ns)
        self.comp = {}
        self.lo = np.zeros((ns, 7))
        self.hi = np.zeros((ns, 7))
        self.tmid = np.zeros(ns)
        self.p0 = 101325.0
        for sp in root.iter("species"):
            nm = sp.get("name")
            if nm not in idx:
                continue
            k = idx[nm]
            comp = dict(_pairs(sp.find("atomArray").text))
            self.comp[nm] = comp
            self.W[k] = sum(n * ATOMIC_W[e] for e, n in comp.items())
            blocks = sorted(sp.iter("NASA"), key=lambda b: float(b.get("Tmin")))
            for b, dst in zip(blocks, (self.lo, self.hi)):
                c = [float(v) for v in b.find("floatArray").text.replace(",", " ").split()]
                dst[k] = c
                if b.get("P0"):
                    self.p0 = float(b.get("P0"))
            self.tmid[k] = float(blocks[0].get("Tmax"))
        self.rx = []
        for r in root.iter("reaction"):
            arr = r.find("rateCoeff/Arrhenius")
            e = arr.find("E")
            units = e.get("units", "cal/mol")
            ear = float(e.text) if units == "K" else float(e.text) * EA_UNITS[units] / R_GAS
            eff = np.ones(ns)
            tb = r.get("type") == "threeBody"
            ef = r.find("rateCoeff/efficiencies")
            if ef is not None:
                eff[:] = float(ef.get("default", "1.0"))
                for s, v in _pairs(ef.text):
                    if s in idx:
                        eff[idx[s]] = v
            self.rx.append(dict(
                A=float(arr.find("A").text), b=float(arr.find("b").text), EaR=ear,
                rev=r.get("reversible", "yes") != "no", tb=tb, eff=eff,
                reac=[(idx[s], v) for s, v in _pairs(r.find("reactants").text)],
                prod=[(idx[s], v) for s, v in _pairs(r.find("products").text)]))
        self.ns = ns
        self.nr = len(self.rx)
        # net stoichiometric matrix (ns x nr)
        self.nu = np.zeros((ns, self.nr))
        for j, r in enumerate(self.rx):
            for k, v in r["reac"]:
                self.nu[k, j] -= v
            for k, v in r["prod"]:
                self.nu[k, j] += v

    # ---- thermodynamics -------------------------------------------------
    def coeffs(self, T):
        T = np.asarray(T, dtype=float)
        low = T[..., None] <= self.tmid
        return np.where(low[..., None], self.lo, self.hi)  # (..., ns, 7)

    def thermo(self, T):
        """cp/R, h/RT, s/R per species, shape (..., ns)."""
        T = np.asarray(T, dtype=float)[..., None]
        a = self.coeffs(T[..., 0])
        cpr = a[..., 0] + a[..., 1] * T + a[..., 2] * T**2 + a[..., 3] * T**3 + a[..., 4] * T**4
        hrt = (a[..., 0] + a[..., 1] * T / 2 + a[..., 2] * T**2 / 3 + a[..., 3] * T**3 / 4
               + a[..., 4] * T**4 / 5 + a[..., 5] / T)
        sr = (a[..., 0] * np.log(T) + a[..., 1] * T + a[..., 2] * T**2 / 2 + a[..., 3] * T**3 / 3
              + a[..., 4] * T**4 / 4 + a[..., 6])
        return cpr, hrt, sr

    # ---- rates ------------------------------------------------------------
    def rates(self, T, rho_or_p, Y, input_is_p=True, parts=False):
        """T (n,), rho or p (n,), Y (ns, n) -> wdot (ns, n) kg/m^3/s, h_mix, cp_mix, h_k."""
        T = np.asarray(T, dtype=float)
        Y = np.asarray(Y, dtype=float)
        winv = (Y / self.W[:, None]).sum(axis=0)
        rho = rho_or_p / (R_GAS * T * winv) if input_is_p else np.asarray(rho_or_p, dtype=float)
        C = rho * Y / self.W[:, None]
        cpr, hrt, sr = self.thermo(T)
        g = (hrt - sr).T  # (ns, n)
        logT = np.log(T)
        lc0 = np.log(self.p0 / (R_GAS * T))
        wd = np.zeros_like(C)
        fwd_all, rev_all = [], []
        for j, r in enumerate(self.rx):
            kf = np.exp(np.log(r["A"]) + r["b"] * logT - r["EaR"] / T)
            fwd = kf.copy()
            dg = np.zeros_like(T)
            for k, v in r["reac"]:
                fwd = fwd * C[k] ** v
                dg -= v * g[k]
            rev = np.zeros_like(T)
            if r["rev"]:
                pp = np.ones_like(T)
                for k, v in r["prod"]:
                    pp = pp * C[k] ** v
                    dg += v * g[k]
                dnu = sum(v for _, v in r["prod"]) - sum(v for _, v in r["reac"])
                kr = np.exp(np.log(r["A"]) + r["b"] * logT - r["EaR"] / T + dg - dnu * lc0)
                rev = kr * pp
            if r["tb"]:
                cm = (r["eff"][:, None] * C).sum(axis=0)
                fwd = fwd * cm
                rev = rev * cm
            q = fwd - rev
            fwd_all.append(fwd)
            rev_all.append(rev)
            wd += self.nu[:, j][:, None] * q
        wdot = wd * self.W[:, None]
        hk = (hrt.T * (R_REF / self.W[:, None]) * T)
        hmix = (Y * hk).sum(axis=0)
        cpmix = R_GAS * (Y * cpr.T / self.W[:, None]).sum(axis=0)
        if parts:
            return wdot, hmix, cpmix, hk, np.array(fwd_all), np.array(rev_all)
        return wdot, hmix, cpmix, hk

    # ---- 0-D constant-pressure adiabatic reactor ------------------------
    def reactor_rhs(self, yv, p):
        """d(T, Y_k)/dt of the constant-pressure adiabatic reactor:
        dY_k/dt = wdot_k W_k / rho, dT/dt = -sum_k h_k dY_k/dt / cp."""
        T = np.array([yv[0]])
        Y = np.asarray(yv[1:], dtype=float)[:, None]
        wd, _, cp, _ = self.rates(T, np.array([p]), Y, input_is_p=True)
        winv = (Y[:, 0] / self.W).sum()
        rho = p / (R_GAS * T[0] * winv)
        dY = wd[:, 0] / rho
        _, hrt, _ = self.thermo(T)
        hk = hrt[0] * R_GAS * T[0] / self.W
        return np.concatenate([[-(hk * dY).sum() / cp[0]], dY])

    def reactor_jacobian(self, yv, p):
        """The analytic Jacobian d f / d(T, Y) of reactor_rhs (mass-action
        kinetics, NASA-7 thermodynamics, constant p), by the chain rule through
        the molar concentrations c_k = rho Y_k / W_k:
          dq_r/dc_m  from the forward and reverse products and [M],
          dq_r/dT|_c = [M](fwd dln kf/dT - rev dln kr/dT),
            dln kf/dT = b/T + E/(R T^2), dln kr/dT = dln kf/dT - dH_r/(R T^2) + dnu_r/T,
          dc_k/dY_j = rho/W_k delta_kj - c_k / (W_j sum_i Y_i/W_i),  dc_k/dT = -c_k/T,
        then f_Y = w W / rho and f_T = -sum_k h_k f_Yk / cp."""
        ns = self.ns
        T = float(yv[0])
        Y = np.asarray(yv[1:], dtype=float)
        winv = (Y / self.W).sum()
        rho = p / (R_GAS * T * winv)
        c = rho * Y / self.W
        cpr, hrt, sr = (a[0] for a in self.thermo(np.array([T])))
        g = hrt - sr
        a = self.coeffs(np.array([T]))[0]
        dcpr = a[:, 1] + 2 * a[:, 2] * T + 3 * a[:, 3] * T**2 + 4 * a[:, 4] * T**3
        logT = np.log(T)
        lc0 = np.log(self.p0 / (R_GAS * T))
        w = np.zeros(ns)
        Jc = np.zeros((ns, ns))   # d w_k / d c_m
        wT = np.zeros(ns)         # d w_k / dT at fixed c
        for j, r in enumerate(self.rx):
            kf = np.exp(np.log(r["A"]) + r["b"] * logT - r["EaR"] / T)
            dlkf = r["b"] / T + r["EaR"] / T**2
            fwd, dfwd = kf, np.zeros(ns)
            for i, (k, v) in enumerate(r["reac"]):
                fwd = fwd * c[k] ** v
            for i, (k, v) in enumerate(r["reac"]):
                other = kf
                for i2, (k2, v2) in enumerate(r["reac"]):
                    if i2 != i:
                        other = other * c[k2] ** v2
                dfwd[k] += other * v * c[k] ** (v - 1)
            rev, drev, dlkr = 0.0, np.zeros(ns), 0.0
            if r["rev"]:
                dg = sum(v * g[k] for k, v in r["prod"]) - sum(v * g[k] for k, v in r["reac"])
                dh = sum(v * hrt[k] for k, v in r["prod"]) - sum(v * hrt[k] for k, v in r["reac"])
                dnu = sum(v for _, v in r["prod"]) - sum(v for _, v in r["reac"])
                kr = np.exp(np.log(r["A"]) + r["b"] * logT - r["EaR"] / T + dg - dnu * lc0)
                dlkr = dlkf - dh / T + dnu / T
                rev = kr
                for k, v in r["prod"]:
                    rev = rev * c[k] ** v
                for i, (k, v) in enumerate(r["prod"]):
                    other = kr
                    for i2, (k2, v2) in enumerate(r["prod"]):
                        if i2 != i:
                            other = other * c[k2] ** v2
                    drev[k] += other * v * c[k] ** (v - 1)
            if r["tb"]:
                cm = (r["eff"] * c).sum()
                dq = cm * (dfwd - drev) + (fwd - rev) * r["eff"]
                q = cm * (fwd - rev)
                qT = cm * (fwd * dlkf - rev * dlkr)
            else:
                dq = dfwd - drev
                q = fwd - rev
                qT = fwd * dlkf - rev * dlkr
            w += self.nu[:, j] * q
            Jc += np.outer(self.nu[:, j], dq)
            wT += self.nu[:, j] * qT
        s = Jc @ c
        dwdY = Jc * (rho / self.W)[None, :] - np.outer(s, 1.0 / (self.W * winv))
        dwdT = wT - s / T
        fY = w * self.W / rho
        J = np.zeros((ns + 1, ns + 1))
        J[1:, 1:] = (self.W / rho)[:, None] * dwdY + np.outer(fY, 1.0 / (self.W * winv))
        J[1:, 0] = self.W / rho * dwdT + fY / T
        e = hrt * T / self.W              # h_k / R
        cpm = (Y * cpr / self.W).sum()    # cp / R
        fT = -(e * fY).sum() / cpm
        dcpm = np.concatenate([[(Y * dcpr / self.W).sum()], cpr / self.W])
        dH = e @ J[1:, :]
        dH[0] += (cpr / self.W * fY).sum()
        J[0, :] = -(dH + fT * dcpm) / cpm
        return J

    def ignite(self, T0, p, Y0, t_end, rtol=1e-11, atol=1e-16):
        """Reference solution of the reactor by scipy's Radau (implicit
        Runge-Kutta, order 5) at tight tolerances; returns (T, Y) at t_end."""
        from scipy.integrate import solve_ivp

        y0 = np.concatenate([[T0], np.asarray(Y0, dtype=float)])
        sol = solve_ivp(lambda t, y: self.reactor_rhs(y, p), (0.0, t_end), y0, method="Radau",
                        rtol=rtol, atol=atol)
        if not sol.success:
            raise RuntimeError(sol.message)
        return sol.y[:, -1]

    # ---- thermodynamic helpers (thermchem.F90:1281-1329, 1639-1657, 1775-1786)
    def e_mass(self, T, Y):
        """Mass-specific internal energy sum_k Y_k (h_k - R T / W_k), J/kg."""
        T = np.asarray(T, dtype=float)
        Y = np.asarray(Y, dtype=float)
        _, hrt, _ = self.thermo(T)
        return R_GAS * ((Y / self.W[:, None]) * (hrt.T * T - T)).sum(axis=0)

    def temperature(self, e, Y, T0, tol=1e-13, maxit=50):
        """Inverse of e_mass by Newton iteration from T0 (setState_UV)."""
        T = np.array(T0, dtype=float)
        Y = np.asarray(Y, dtype=float)
        for _ in range(maxit):
            cpr, hrt, _ = self.thermo(T)
            winv = (Y / self.W[:, None]).sum(axis=0)
            cv = R_GAS * ((Y / self.W[:, None]) * cpr.T).sum(axis=0) - R_GAS * winv
            dT = (e - self.e_mass(T, Y)) / cv
            T = T + dT
            if np.all(np.abs(dT) <= tol * np.abs(T)):
                break
        return T

    def rgcmix(self, Y):
        """Mixture gas constant with the reference's R = 8314.2 (rgcmix)."""
        return (np.asarray(Y, dtype=float) * (R_REF / self.W)[:, None]).sum(axis=0)

    def convertxiyi(self, frac, mode):
        """Mole <-> mass fractions (convertxiyi), frac of shape (ns, n)."""
        frac = np.asarray(frac, dtype=float)
        if mode == "X2Y":
            v = frac * self.W[:, None]
        elif mode == "Y2X":
            v = frac / self.W[:, None]
        else:
            raise ValueError("mode must be X2Y or Y2X")
        return v / v.sum(axis=0)

"""Host side of the chemical source terms: a ctypes front end over the
astr_hip_chem_* entry points (include/astr_hip.h) and the synthetic H2-air
cell generator used by the batch benchmark and the tests.

Method names follow the reference routines they replace
(src/thermchem.F90): chemrate (net production rates times W_k), cpeval
(cp_mass) and enthpy (h_k with the reference's R = 8314.2).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import AstrError, lib

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _bind():
    L = lib()
    if getattr(L, "_chem_bound", False):
        return L
    L.astr_hip_chem_last_error.restype = C.c_char_p
    L.astr_hip_chem_create.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
    L.astr_hip_chem_destroy.argtypes = [C.c_void_p]
    L.astr_hip_chem_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp,
                                     C.c_char_p, C.c_int]
    L.astr_hip_chem_rates.argtypes = [C.c_void_p, C.c_longlong, _dp, _dp, C.c_int, _dp, _dp,
                                      _dp, _dp, _dp]
    L.astr_hip_chem_rates_device.argtypes = [C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p,
                                             C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p]
    L.astr_hip_chem_temperature.argtypes = [C.c_void_p, C.c_longlong, _dp, _dp, _dp]
    L.astr_hip_chem_temperature_device.argtypes = [C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p,
                                                   C.c_void_p, C.c_void_p]
    L.astr_hip_reactor_create.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                          C.POINTER(C.c_void_p)]
    L.astr_hip_reactor_destroy.argtypes = [C.c_void_p]
    L.astr_hip_reactor_set.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp,
                                       C.c_double, C.c_double]
    L.astr_hip_reactor_advance.argtypes = [C.c_void_p, C.c_int]
    L.astr_hip_reactor_sync.argtypes = [C.c_void_p]
    L.astr_hip_reactor_query.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.astr_hip_reactor_set_lanes.argtypes = [C.c_void_p, C.c_int]
    L.astr_hip_reactor_get_lanes.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.astr_hip_reactor_status.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.astr_hip_reactor_get.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]
    L.astr_hip_reactor_release.argtypes = [C.c_void_p, C.c_int]
    L.astr_hip_reactor_rhs.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
    L.astr_hip_reactor_method.argtypes = [_dp, C.c_int]
    L.astr_hip_srccomb.argtypes = [C.c_void_p] + [C.c_int] * 5 + [_dp] * 5
    L.astr_hip_srccomb_device.argtypes = [C.c_void_p] + [C.c_int] * 5 + [C.c_void_p] * 6
    L._chem_bound = True
    return L


def _chk(rc):
    if rc != 0:
        raise AstrError(_bind().astr_hip_chem_last_error().decode())


class Mechanism:
    """A CTML mechanism resident on the device (chemread, thermchem.F90:59-98)."""

    def __init__(self, path, phase_id="gas"):
        L = _bind()
        h = C.c_void_p()
        _chk(L.astr_hip_chem_create(path.encode(), phase_id.encode(), C.byref(h)))
        self.h = h
        ns, nr = C.c_int(), C.c_int()
        _chk(L.astr_hip_chem_info(h, C.byref(ns), C.byref(nr), None, None, 0))
        self.ns, self.nr = ns.value, nr.value
        self.W = np.zeros(self.ns)
        buf = C.create_string_buffer(16 * self.ns)
        _chk(L.astr_hip_chem_info(h, None, None, _p(self.W), buf, 16))
        self.names = [buf.raw[16 * k:16 * (k + 1)].split(b"\0")[0].decode() for k in range(self.ns)]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _bind().astr_hip_chem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reactor_rhs(self, p, y, jacobian=False):
        """The reactor's d(T, Y)/dt at states y (n, ns+1) and pressures p (n,),
        and with jacobian=True the analytic Jacobian the integrator uses
        (n, ns+1, ns+1)."""
        p = np.ascontiguousarray(np.atleast_1d(p), dtype=np.float64)
        y = np.ascontiguousarray(np.atleast_2d(y), dtype=np.float64)
        n, m = y.shape
        if m != self.ns + 1 or p.shape != (n,):
            raise AstrError("shapes: p (n,), y (n, ns+1)")
        f = np.empty((n, m))
        J = np.empty((n, m, m)) if jacobian else None
        _chk(_bind().astr_hip_reactor_rhs(self.h, n, _p(p), _p(y), _p(f), _p(J)))
        return (f, J) if jacobian else f

    def rates(self, T, rho_or_p, Y, input_is_p=True, species_h=False):
        """T, rho|p: (n,), Y: (ns, n) -> wdot (ns, n) kg/m^3/s, h_mix, cp_mix[, h_k]."""
        T = np.ascontiguousarray(T, dtype=np.float64)
        rp = np.ascontiguousarray(rho_or_p, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        n = T.size
        if Y.shape != (self.ns, n) or rp.shape != (n,):
            raise AstrError("shapes: T (n,), rho_or_p (n,), Y (ns, n)")
        wd = np.empty((self.ns, n))
        hm = np.empty(n)
        cp = np.empty(n)
        hk = np.empty((self.ns, n)) if species_h else None
        _chk(_bind().astr_hip_chem_rates(self.h, n, _p(T), _p(rp), int(input_is_p), _p(Y),
                                         _p(wd), _p(hm), _p(cp), _p(hk)))
        return (wd, hm, cp, hk) if species_h else (wd, hm, cp)

    chemrate = rates

    def srccomb(self, rho, tmp, spc, jacob, qrhs, hm, stream=0):
        """srccomb (src/solver.F90:500-537): qrhs[5 + k] += wdot_k W_k * jacob at
        every computed point, in place.  The arrays are the reference's
        Fortran ones with their index order reversed (C order): rho, tmp,
        jacob (km+1+2hm, jm+1+2hm, im+1+2hm), spc (ns, ...same), qrhs
        (5 + ns, km+1, jm+1, im+1).  numpy arrays go through the host entry
        point; astr_amd.hip.DeviceArray buffers in HBM through the device
        one, on stream."""
        nz, ny, nx = (d - 2 * hm for d in rho.shape)
        want = {"tmp": tmp.shape, "jacob": jacob.shape}
        for name, shp in want.items():
            if tuple(shp) != tuple(rho.shape):
                raise AstrError(f"srccomb: {name} shape {tuple(shp)} != rho's {tuple(rho.shape)}")
        if tuple(spc.shape) != (self.ns,) + tuple(rho.shape):
            raise AstrError("srccomb: spc must be (ns,) + rho's shape")
        if tuple(qrhs.shape) != (5 + self.ns, nz, ny, nx):
            raise AstrError("srccomb: qrhs must be (5 + ns, km+1, jm+1, im+1)")
        dims = (nx - 1, ny - 1, nz - 1, hm, 5 + self.ns)
        L = _bind()
        if isinstance(rho, np.ndarray):
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, tmp, spc, jacob)]
            if not (qrhs.flags.c_contiguous and qrhs.dtype == np.float64):
                raise AstrError("srccomb: qrhs must be a C-contiguous float64 array")
            _chk(L.astr_hip_srccomb(self.h, *dims, *(_p(a) for a in arrs), _p(qrhs)))
            return qrhs
        _chk(L.astr_hip_srccomb_device(self.h, *dims, *(a.ptr for a in (rho, tmp, spc, jacob, qrhs)),
                                       stream))
        return qrhs


def mixture(names, W, phi, c):
    """Mass fractions (ns, n): unburnt H2/air at equivalence ratio phi (O2:N2
    = 0.21:0.79 by mole) blended by progress c with complete-combustion
    products plus a fixed radical pool (mole fractions OH=H=O=1e-3,
    HO2=H2O2=1e-5).  A build-defined input generator, not a reference
    routine."""
    ix = {s: k for k, s in enumerate(names)}
    W = np.asarray(W, dtype=float)
    phi = np.asarray(phi, dtype=float)
    c = np.asarray(c, dtype=float)
    n = phi.size
    ns = len(names)
    Xu = np.zeros((ns, n))
    Xu[ix["H2"]] = 2.0 * phi
    Xu[ix["O2"]] = 1.0
    Xu[ix["N2"]] = 0.79 / 0.21
    Xb = np.zeros((ns, n))
    Xb[ix["H2O"]] = 2.0 * np.minimum(phi, 1.0)
    Xb[ix["O2"]] = np.maximum(1.0 - phi, 0.0)
    Xb[ix["H2"]] = 2.0 * np.maximum(phi - 1.0, 0.0)
    Xb[ix["N2"]] = 0.79 / 0.21
    Xb = Xb / Xb.sum(axis=0)
    for s, v in (("OH", 1e-3), ("H", 1e-3), ("O", 1e-3), ("HO2", 1e-5), ("H2O2", 1e-5)):
        Xb[ix[s]] += v
    Xb = Xb / Xb.sum(axis=0)
    Xu = Xu / Xu.sum(axis=0)

    def x2y(X):
        m = X * W[:, None]
        return m / m.sum(axis=0)

    Y = (1.0 - c) * x2y(Xu) + c * x2y(Xb)
    return Y / Y.sum(axis=0)


def sample_cells(names, W, n, seed):
    """The chemistry batch workload: T ~ U[900,2500] K, p log-uniform in
    [0.5,5] atm, phi ~ U[0.3,3], progress c ~ U[0,1]; returns T, p, Y."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(900.0, 2500.0, n)
    p = np.exp(rng.uniform(np.log(0.5), np.log(5.0), n)) * 101325.0
    phi = rng.uniform(0.3, 3.0, n)
    c = rng.uniform(0.0, 1.0, n)
    return T, p, mixture(names, W, phi, c)


def temperature(mech, Y, e, T0):
    """temperature_calc: T from internal energy e (J/kg) and Y (ns, n), from
    the initial guesses T0 (thermchem.F90:1775-1786)."""
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    T = np.array(T0, dtype=np.float64, copy=True).reshape(-1)
    if Y.shape != (mech.ns, T.size) or e.shape != (T.size,):
        raise AstrError("shapes: Y (ns, n), e (n,), T0 (n,)")
    _chk(_bind().astr_hip_chem_temperature(mech.h, T.size, _p(Y), _p(e), _p(T)))
    return T


class ReactorPool:
    """Slots of 0-D constant-pressure adiabatic reactors integrated on the GPU
    (astr_hip_reactor_*, include/astr_hip.h).  Status per slot: 0 free,
    1 running, 2 done, 3 failed."""

    FREE, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, mech: Mechanism, capacity, rtol=1e-6, atol=1e-12, lanes=None):
        self.mech = mech
        self.capacity = capacity
        h = C.c_void_p()
        _chk(_bind().astr_hip_reactor_create(mech.h, capacity, rtol, atol, C.byref(h)))
        self.h = h
        if lanes is not None:
            self.lanes = lanes

    @property
    def lanes(self):
        """Lanes per reactor: 1 (throughput) or 16 (a team, for latency)."""
        v = C.c_int()
        _chk(_bind().astr_hip_reactor_get_lanes(self.h, C.byref(v)))
        return v.value

    @lanes.setter
    def lanes(self, n):
        _chk(_bind().astr_hip_reactor_set_lanes(self.h, int(n)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _bind().astr_hip_reactor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, slot, T0, p, Y0, t_end, h0=1e-10):
        Y0 = np.ascontiguousarray(Y0, dtype=np.float64)
        if Y0.shape != (self.mech.ns,):
            raise AstrError("Y0 must hold one mass fraction per species")
        _chk(_bind().astr_hip_reactor_set(self.h, slot, T0, p, _p(Y0), t_end, h0))

    def advance(self, max_tries=200):
        _chk(_bind().astr_hip_reactor_advance(self.h, max_tries))

    def sync(self):
        _chk(_bind().astr_hip_reactor_sync(self.h))

    def idle(self):
        """True once every launched advance has finished."""
        v = C.c_int()
        _chk(_bind().astr_hip_reactor_query(self.h, C.byref(v)))
        return bool(v.value)

    def status(self):
        st = np.zeros(self.capacity, dtype=np.int32)
        _chk(_bind().astr_hip_reactor_status(self.h, st.ctypes.data_as(C.POINTER(C.c_int))))
        return st

    def get(self, slot):
        """(t, T, Y, accepted steps, rejected steps) of a slot."""
        t = C.c_double()
        yv = np.zeros(self.mech.ns + 1)
        ns_, nr_ = C.c_int(), C.c_int()
        _chk(_bind().astr_hip_reactor_get(self.h, slot, C.byref(t), _p(yv), C.byref(ns_),
                                          C.byref(nr_)))
        return t.value, yv[0], yv[1:], ns_.value, nr_.value

    def release(self, slot):
        _chk(_bind().astr_hip_reactor_release(self.h, slot))

    def run(self, max_tries=200):
        """Advance until no slot is running; returns the final status array."""
        while True:
            self.advance(max_tries)
            st = self.status()
            if not np.any(st == self.RUNNING):
                return st

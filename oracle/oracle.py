"""ctypes front end to the oracle (TEST INFRASTRUCTURE ONLY).

Two checkers live under oracle/:
  * liboracle.so   -- the C restatement (astr_oracle.c), in-process;
  * _ref/libastr_ref.so -- the reference's own Fortran (built from the
    reference tree by oracle/Makefile), run in a child process per
    configuration because the reference keeps per-grid SAVE state.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only to check or time against; the product path
(astr_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libastr_ref.so")
HM = 5

_dp = C.POINTER(C.c_double)


class OrCfg(C.Structure):
    _fields_ = [
        ("im", C.c_int), ("jm", C.c_int), ("km", C.c_int),
        ("alfa_filter", C.c_double), ("beta_halo", C.c_double), ("beta_bound", C.c_double),
        ("gamma", C.c_double), ("mach", C.c_double), ("reynolds", C.c_double),
        ("prandtl", C.c_double), ("ref_tem", C.c_double), ("suth", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_ORACLE):
            raise RuntimeError(f"oracle not built: {LIB_ORACLE} (run make -C oracle oracle)")
        L = C.CDLL(LIB_ORACLE)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(OrCfg)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_npoints.restype = C.c_long
        L.or_npoints.argtypes = [C.c_void_p]
        for f in ("or_q", "or_prims", "or_qrhs", "or_dvel", "or_dtmp"):
            getattr(L, f).restype = _dp
            getattr(L, f).argtypes = [C.c_void_p]
        L.or_tgv_init.argtypes = [C.c_void_p, C.c_double, C.c_uint64]
        L.or_set_q.argtypes = [C.c_void_p, _dp]
        L.or_rhscal.argtypes = [C.c_void_p]
        L.or_filterq.argtypes = [C.c_void_p]
        L.or_stacal.argtypes = [C.c_void_p, _dp, _dp]
        L.or_rk3_steps.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.or_rk3_steps3.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.or_diss_rate.argtypes = [C.c_void_p, _dp]
        L.or_df_compact.argtypes = [C.c_int, _dp, _dp]
        L.or_compact_filter.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.or_df_compact_nt.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.or_compact_filter_nt.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                           _dp, _dp]
        L.or_filter_coefs_b.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_gauss.restype = C.c_double
        L.or_gauss.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_deriv_lhs.argtypes = [C.c_int, _dp, _dp]
        L.or_filter_lhs.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.or_deriv_rhs.argtypes = [C.c_int, _dp, _dp]
        L.or_filter_rhs.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.or_filter_coefs.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.or_main_create.restype = C.c_void_p
        L.or_main_create.argtypes = [C.POINTER(OrCfg)]
        L.or_main_destroy.argtypes = [C.c_void_p]
        for f in ("or_main_npoints", "or_main_nhalo"):
            getattr(L, f).restype = C.c_long
            getattr(L, f).argtypes = [C.c_void_p]
        for f in ("or_main_qh", "or_main_primsh", "or_main_qrhs", "or_main_dvel", "or_main_dtmp"):
            getattr(L, f).restype = _dp
            getattr(L, f).argtypes = [C.c_void_p]
        L.or_main_set_metrics.argtypes = [C.c_void_p, _dp, _dp]
        L.or_main_set_npdc.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.or_main_set_q.argtypes = [C.c_void_p, _dp]
        L.or_main_set_q_only.argtypes = [C.c_void_p, _dp]
        L.or_main_get.argtypes = [C.c_void_p, _dp, _dp]
        for f in ("or_main_updatefvar", "or_main_filterq", "or_main_qswap", "or_main_gradcal",
                  "or_main_rhscal"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.or_main_stats.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_main_rk3_steps.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.or_main_set_conv.restype = C.c_int
        L.or_main_set_conv.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_main_ssf.restype = _dp
        L.or_main_ssf.argtypes = [C.c_void_p]
        L.or_main_lshock.restype = C.POINTER(C.c_ubyte)
        L.or_main_lshock.argtypes = [C.c_void_p]
        L.or_sw_split.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.or_recons.restype = C.c_double
        L.or_recons.argtypes = [C.c_int, _dp, C.c_double, C.c_int]
        L.or_chardecomp.restype = C.c_int
        L.or_chardecomp.argtypes = [C.c_double, _dp, _dp, _dp, _dp]
        L.or_mp5.restype = C.c_double
        L.or_mp5.argtypes = [_dp, C.c_double, C.c_int]
        L.or_cuw_lu.argtypes = [C.c_int, C.c_double, C.c_int, _dp]
        L.or_cuw_flux.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def tgv_cfg(im, jm=None, km=None, dt=1e-3, reynolds=1600.0, mach=0.1, suth_dim=110.4,
            ref_tem=273.15, lx=2 * np.pi, ly=None, lz=None):
    jm = im if jm is None else jm
    km = im if km is None else km
    ly = lx if ly is None else ly
    lz = lx if lz is None else lz
    return OrCfg(im, jm, km, 0.49, 1.11, 1.09, 1.4, mach, reynolds, 0.72, ref_tem,
                 suth_dim / ref_tem, lx / im, ly / jm, (lz / km) if km > 0 else 1.0, dt)


class OracleState:
    """One field state of the C restatement (q, primitives, qrhs, dvel, dtmp)."""

    def __init__(self, cfg: OrCfg):
        self.cfg = cfg
        self.h = lib().or_create(C.byref(cfg))
        self.N = lib().or_npoints(self.h)
        self.shape = (cfg.km + 1, cfg.jm + 1, cfg.im + 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def _view(self, fn, nf):
        """A numpy view of a C array that keeps this state alive while the
        view lives (the C arrays are freed with the state)."""
        ptr = getattr(lib(), fn)(self.h)
        buf = (C.c_double * (nf * self.N)).from_address(C.addressof(ptr.contents))
        buf._owner = self
        return np.frombuffer(buf, dtype=np.float64).reshape((nf,) + self.shape)

    @property
    def q(self):
        return self._view("or_q", 5)

    @property
    def prims(self):
        return self._view("or_prims", 6)

    @property
    def qrhs(self):
        return self._view("or_qrhs", 5)

    @property
    def dvel(self):
        """dvel[m + 3*d] = d u_m / d x_d."""
        return self._view("or_dvel", 9)

    @property
    def dtmp(self):
        return self._view("or_dtmp", 3)

    def tgv_init(self, amp=0.0, seed=0):
        lib().or_tgv_init(self.h, amp, seed)

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().or_set_q(self.h, _p(q))

    def rhscal(self):
        lib().or_rhscal(self.h)

    def filterq(self):
        lib().or_filterq(self.h)

    def stacal(self):
        a = np.zeros(1)
        b = np.zeros(1)
        lib().or_stacal(self.h, _p(a), _p(b))
        return a[0], b[0]

    def diss_rate(self):
        """diss_rate_cal (statistic.F90:994-1043) from the current dvel and
        temperature (run rhscal first)."""
        d = np.zeros(1)
        lib().or_diss_rate(self.h, _p(d))
        return d[0]

    def rk3(self, nsteps, dissipation=False):
        ke = np.zeros(nsteps)
        en = np.zeros(nsteps)
        di = np.zeros(nsteps) if dissipation else None
        lib().or_rk3_steps3(self.h, nsteps, _p(ke), _p(en), _p(di))
        return (ke, en, di) if dissipation else (ke, en)


def uniform_metrics(cfg: OrCfg, jac=None, dxi_diag=None):
    """Metrics of a uniform Cartesian grid as (jac (N), dxi (9, N)) with
    dxi[d + 3 n] = d xi_d / d x_n: jac = dx dy dz (dx dy in 2-D) and
    dxi(d,d) = 1/dx_d unless given."""
    N = (cfg.im + 1) * (cfg.jm + 1) * (cfg.km + 1)
    if dxi_diag is None:
        dxi_diag = (1.0 / cfg.dx, 1.0 / cfg.dy, 1.0 / cfg.dz if cfg.km > 0 else 0.0)
    if jac is None:
        jac = cfg.dx * cfg.dy * (cfg.dz if cfg.km > 0 else 1.0)
    dxi = np.zeros((9, N))
    for d in range(3 if cfg.km > 0 else 2):
        dxi[d + 3 * d] = dxi_diag[d]
    return np.full(N, float(jac)), dxi


class MainState:
    """One field state of the main solver's restatement (astr_oracle_main.c):
    q and the stored primitives with halos, the metrics, qrhs, dvel, dtmp."""

    def __init__(self, cfg: OrCfg, jac=None, dxi=None):
        self.cfg = cfg
        self.h = lib().or_main_create(C.byref(cfg))
        self.N = lib().or_main_npoints(self.h)
        self.NH = lib().or_main_nhalo(self.h)
        self.shape = (cfg.km + 1, cfg.jm + 1, cfg.im + 1)
        hk = HM if cfg.km > 0 else 0
        self.hshape = (cfg.km + 1 + 2 * hk, cfg.jm + 1 + 2 * HM, cfg.im + 1 + 2 * HM)
        if jac is None:
            jac, dxi = uniform_metrics(cfg)
        self.set_metrics(jac, dxi)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_main_destroy(self.h)
            self.h = None

    def set_npdc(self, npdc):
        """Boundary type per direction (3 periodic, 4 physical at both ends)."""
        lib().or_main_set_npdc(self.h, (C.c_int * 3)(*[int(v) for v in npdc]))

    def set_metrics(self, jac, dxi):
        """jac (N) and dxi (9, N), dxi[d + 3 n] = d xi_d / d x_n."""
        self._jac = np.ascontiguousarray(np.reshape(jac, (self.N,)), dtype=np.float64)
        self._dxi = np.ascontiguousarray(np.reshape(dxi, (9, self.N)), dtype=np.float64)
        lib().or_main_set_metrics(self.h, _p(self._jac), _p(self._dxi))

    def _view(self, fn, nf, shape):
        """A numpy view of a C array that keeps this state alive while the
        view lives (the C arrays are freed with the state)."""
        ptr = getattr(lib(), fn)(self.h)
        size = int(np.prod(shape))
        buf = (C.c_double * (nf * size)).from_address(C.addressof(ptr.contents))
        buf._owner = self
        return np.frombuffer(buf, dtype=np.float64).reshape((nf,) + shape)

    @property
    def qh(self):
        """q with halos, (5,) + hshape; index [m, k+hk, j+hm, i+hm]."""
        return self._view("or_main_qh", 5, self.hshape)

    @property
    def primsh(self):
        """stored rho, u, v, w, p, T with halos, (6,) + hshape."""
        return self._view("or_main_primsh", 6, self.hshape)

    @property
    def qrhs(self):
        return self._view("or_main_qrhs", 5, self.shape)

    @property
    def dvel(self):
        """dvel[m + 3*n] = d u_m / d x_n."""
        return self._view("or_main_dvel", 9, self.shape)

    @property
    def dtmp(self):
        return self._view("or_main_dtmp", 3, self.shape)

    def set_q(self, q):
        """q on the computed points, then updatefvar."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().or_main_set_q(self.h, _p(q))

    def set_q_only(self, q):
        """q on the computed points; the stored primitives are kept."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().or_main_set_q_only(self.h, _p(q))

    def get(self):
        """(q (5,)+shape, stored primitives (6,)+shape)."""
        q = np.empty((5,) + self.shape)
        pr = np.empty((6,) + self.shape)
        lib().or_main_get(self.h, _p(q), _p(pr))
        return q, pr

    def updatefvar(self):
        lib().or_main_updatefvar(self.h)

    def filterq(self):
        lib().or_main_filterq(self.h)

    def qswap(self):
        lib().or_main_qswap(self.h)

    def gradcal(self):
        lib().or_main_gradcal(self.h)

    def rhscal(self):
        lib().or_main_rhscal(self.h)

    def stats(self):
        a, b, c = np.zeros(1), np.zeros(1), np.zeros(1)
        lib().or_main_stats(self.h, _p(a), _p(b), _p(c))
        return a[0], b[0], c[0]

    def rk3(self, nsteps):
        """nsteps main-path RK3 steps; (ke, ens, diss) of each step's first stage."""
        ke, en, di = np.zeros(nsteps), np.zeros(nsteps), np.zeros(nsteps)
        lib().or_main_rk3_steps(self.h, nsteps, _p(ke), _p(en), _p(di))
        return ke, en, di

    def set_conv(self, upwind, recon=0, chardecomp=False, bfacmpld=0.3, shkcrt=0.05):
        """Convective scheme: central (upwind 0 / False), the explicit
        upwind-biased one (1 / True) with recon_schem, lchardecomp, bfacmpld,
        shkcrt, or the compact upwind-biased 543c (2) with lchardecomp,
        bfacmpld, shkcrt."""
        rc = lib().or_main_set_conv(self.h, int(upwind), int(recon), int(bool(chardecomp)),
                                    float(bfacmpld), float(shkcrt))
        if rc:
            raise ValueError("convection scheme not covered by the restatement")

    @property
    def ssf(self):
        """The Ducros sensor of the last right-hand side, shape."""
        return self._view("or_main_ssf", 1, self.shape)[0]

    @property
    def lshock(self):
        """The shock flags of the last right-hand side (bool copy), shape."""
        ptr = lib().or_main_lshock(self.h)
        return np.ctypeslib.as_array(ptr, shape=(self.N,)).reshape(self.shape).astype(bool)


def sw_split(gamma, mach, jac, dxi, q, prims):
    """flux_steger_warming at one point: (fplus (5,), fminus (5,))."""
    dxi = np.ascontiguousarray(dxi, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    prims = np.ascontiguousarray(prims, dtype=np.float64)
    fp, fm = np.zeros(5), np.zeros(5)
    lib().or_sw_split(gamma, mach, jac, _p(dxi), _p(q), _p(prims), _p(fp), _p(fm))
    return fp, fm


def recons(scheme, f, bfac=0.3, shock=True):
    """recons_exp's periodic branch on the 8-point window f(1:8)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    assert f.shape == (8,)
    return lib().or_recons(int(scheme), _p(f), float(bfac), int(bool(shock)))


def chardecomp(gamma, left, right):
    """chardecomp between two points, each (rho, u, v, w, p, E, dxi row):
    (REV, LEV), 5 x 5."""
    l = np.ascontiguousarray(left, dtype=np.float64)
    r = np.ascontiguousarray(right, dtype=np.float64)
    rev, lev = np.zeros(25), np.zeros(25)
    if lib().or_chardecomp(gamma, _p(l), _p(r), _p(rev), _p(lev)):
        raise ValueError("zero metric row")
    return rev.reshape(5, 5), lev.reshape(5, 5)


def mp5(u, ul, shock=True):
    """MP5 as mplimiter applies it on a periodic line: the window u(1:5)
    (u(3) upwind) limited against the linear interface value ul."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    assert u.shape == (5,)
    return lib().or_mp5(_p(u), float(ul), int(bool(shock)))


def cuw_lu(n, bfac, plus):
    """compact_flux_initiate, scheme 543, periodic line of n+1 points: the
    LU factors ac(1:3, -2:n+1) as an (n+4, 3) array."""
    ac = np.zeros(3 * (n + 4))
    lib().or_cuw_lu(int(n), float(bfac), int(bool(plus)), _p(ac))
    return ac.reshape(n + 4, 3)


def cuw_flux(n, bfac, plus, f):
    """flux_compact on a periodic line: f = f(-4:n+4) -> fh(-1:n)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    assert f.shape == (n + 9,)
    ac = np.ascontiguousarray(cuw_lu(n, bfac, plus)).ravel()
    fh = np.zeros(n + 2)
    lib().or_cuw_flux(int(n), float(bfac), int(bool(plus)), _p(ac), _p(f), _p(fh))
    return fh


def qswap(q):
    """qswap's single-rank periodic branch (parallel.F90:4876-4884,
    5018-5030, 5159-5170) on a (5, km+1, jm+1, im+1) array (the lattice
    0..n with its duplicate planes): q(0) = 0.5 (q(0) + q(n)), q(n) = q(0),
    along i over all j, k, then along j over i = 0..im, then along k.  The
    halo copies it also makes are not part of this array."""
    q = np.array(q, dtype=np.float64, copy=True)
    q[..., 0] = 0.5 * (q[..., 0] + q[..., -1])
    q[..., -1] = q[..., 0]
    q[:, :, 0, :] = 0.5 * (q[:, :, 0, :] + q[:, :, -1, :])
    q[:, :, -1, :] = q[:, :, 0, :]
    if q.shape[1] > 1:
        q[:, 0] = 0.5 * (q[:, 0] + q[:, -1])
        q[:, -1] = q[:, 0]
    return q


def df_compact(f):
    """Reference-layout line f(-hm:n+hm) -> df(0:n) by the C restatement."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.size - 2 * HM - 1
    out = np.zeros(n + 1)
    lib().or_df_compact(n, _p(f), _p(out))
    return out


def compact_filter(f, alfa=0.49, beta_halo=1.11, beta_bound=1.09):
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.size - 2 * HM - 1
    out = np.zeros(n + 1)
    lib().or_compact_filter(n, alfa, beta_halo, beta_bound, _p(f), _p(out))
    return out


def df_compact_nt(f, ntype):
    """df_compact of a line f(-hm:n+hm) with boundary type ntype (1: physical
    boundary at 0, 2: at n, 3: interfaces, 4: both physical)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.size - 2 * HM - 1
    out = np.zeros(n + 1)
    if lib().or_df_compact_nt(n, ntype, _p(f), _p(out)):
        raise ValueError(f"ntype {ntype}")
    return out


def compact_filter_nt(f, ntype, alfa=0.49, beta_halo=1.11, beta_bound=0.98):
    """compact_filter of a line with boundary type ntype; the defaults are
    the main path's coefficients (comsolver.F90:121)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.size - 2 * HM - 1
    out = np.zeros(n + 1)
    if lib().or_compact_filter_nt(n, ntype, alfa, beta_halo, beta_bound, _p(f), _p(out)):
        raise ValueError(f"ntype {ntype}")
    return out


def gauss(seed, i, j, k, c):
    return lib().or_gauss(seed, i, j, k, c)


def philox(ctr, key):
    a = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox(a, k, o)
    return list(o)


# ---------------------------------------------------------------------------
# the reference's own Fortran, one child process per configuration
# ---------------------------------------------------------------------------

def ref_available():
    return os.path.exists(LIB_REF)


def run_ref(task, timeout=3600, **arrays_and_params):
    """Run `task` of oracle/ref_run.py in a child process; returns dict of arrays."""
    if not ref_available():
        raise RuntimeError("reference build missing: " + LIB_REF)
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.npz")
        out = os.path.join(td, "out.npz")
        np.savez(inp, **{k: np.asarray(v) for k, v in arrays_and_params.items()})
        cmd = [sys.executable, os.path.join(HERE, "ref_run.py"), task, inp, out]
        r = subprocess.run(cmd, cwd=td, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference task {task} failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        with np.load(out) as z:
            return {k: z[k] for k in z.files}

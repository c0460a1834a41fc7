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
        ptr = getattr(lib(), fn)(self.h)
        return np.ctypeslib.as_array(ptr, shape=(nf * self.N,)).reshape((nf,) + self.shape)

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

"""astr_amd -- MI355X-native per-step hot path of the ASTR compact-scheme solver.

The compute lives in astr_amd/libastr_hip.so (HIP for gfx950, C ABI declared
in include/astr_hip.h).  This module is a thin ctypes host over that ABI,
used by tests and bench.py; the Fortran host binding is
astr_amd/fortran/astr_hip_mod.F90.  There is no CPU fallback: if the
library is missing or no GPU is visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ASTR_HIP_LIB", os.path.join(PKG, "libastr_hip.so"))
SEM_MINIAPP = 0
SEM_MAIN = 1

_dp = C.POINTER(C.c_double)


class AstrError(RuntimeError):
    pass


class Config(C.Structure):
    """astr_hip_config (include/astr_hip.h)."""

    _fields_ = [
        ("im", C.c_int), ("jm", C.c_int), ("km", C.c_int), ("semantics", C.c_int),
        ("alfa_filter", C.c_double), ("beta_halo", C.c_double), ("beta_bound", C.c_double),
        ("gamma", C.c_double), ("mach", C.c_double), ("reynolds", C.c_double),
        ("prandtl", C.c_double), ("ref_tem", C.c_double), ("sutherland", C.c_double),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double), ("deltat", C.c_double),
    ]


class Layout(C.Structure):
    _fields_ = [("pitch", C.c_int), ("plane", C.c_longlong), ("field_stride", C.c_longlong),
                ("q", C.c_void_p), ("qrhs", C.c_void_p)]


EXPORTS = [
    "astr_hip_last_error", "astr_hip_create", "astr_hip_destroy", "astr_hip_set_stream",
    "astr_hip_set_deltat", "astr_hip_field_info", "astr_hip_init_tgv", "astr_hip_upload",
    "astr_hip_download", "astr_hip_filterq", "astr_hip_gradcal", "astr_hip_rhscal",
    "astr_hip_rk3_steps", "astr_hip_rk3_steps_stats", "astr_hip_sync", "astr_hip_line_batch",
    "astr_hip_profile_begin", "astr_hip_profile_end", "astr_hip_autotune",
    "astr_hip_get_pass_config", "astr_hip_set_pass_config", "astr_hip_init_vortex",
    "astr_hip_batch_create", "astr_hip_batch_destroy", "astr_hip_batch_autotune",
    "astr_hip_batch_rk3_steps", "astr_hip_batch_info", "astr_hip_set_last_split",
    "astr_hip_batch_rk3_begin", "astr_hip_batch_query", "astr_hip_batch_rk3_end",
    "astr_hip_gridgeom", "astr_hip_gridgeom_nt", "astr_hip_line_batch_nt", "astr_hip_updatefvar",
    "astr_hip_set_metrics", "astr_hip_qswap",
    "astr_hip_get_last_split", "astr_hip_set_filter_fusion", "astr_hip_rk3_steps_stats_noalloc",
    "astr_hip_set_fvar",
    "astr_hip_write_checkpoint", "astr_hip_read_checkpoint", "astr_hip_set_boundary",
    "astr_hip_set_convection",
]
SHAPES = {0: (4, 64, 5), 1: (4, 64, 9), 2: (16, 16, 5), 3: (8, 32, 9), 4: (8, 32, 5),
          5: (8, 64, 5), 6: (4, 32, 9), 7: (2, 64, 5),
          8: (1, 64, 5), 9: (2, 32, 9), 10: (16, 16, 9), 11: (8, 64, 9), 12: (1, 64, 9)}
PROF_SLOTS = 39
# pass programs 0..9; filter_gc_first is gc_first preceded by the previous
# stage's x filter sweep (rk3 steps with filter fusion, the default)
PASS_NAMES = ["gc_first", "gc_mid", "gc_last", "visc_mid", "visc_last_rhs",
              "visc_last_rk3_filter", "filter", "gc_last_grad", "gc_last_conv",
              "filter_gc_first"]
# profile slots (3 per name): passes 0..8, the upwind-biased convection's
# kernels, then pass 9
PROF_NAMES = PASS_NAMES[:9] + ["uw_split", "uw_face", "uw_sensor"] + PASS_NAMES[9:]

_lib = None


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AstrError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.astr_hip_last_error.restype = C.c_char_p
        L.astr_hip_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
        L.astr_hip_destroy.argtypes = [C.c_void_p]
        L.astr_hip_set_stream.argtypes = [C.c_void_p, C.c_void_p]
        L.astr_hip_set_deltat.argtypes = [C.c_void_p, C.c_double]
        L.astr_hip_field_info.argtypes = [C.c_void_p, C.POINTER(Layout)]
        L.astr_hip_init_tgv.argtypes = [C.c_void_p, C.c_double, C.c_ulonglong]
        L.astr_hip_upload.argtypes = [C.c_void_p, _dp, C.c_int]
        L.astr_hip_download.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.astr_hip_filterq.argtypes = [C.c_void_p]
        L.astr_hip_gradcal.argtypes = [C.c_void_p, _dp, _dp]
        L.astr_hip_rhscal.argtypes = [C.c_void_p, _dp]
        L.astr_hip_rk3_steps.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.astr_hip_rk3_steps_stats.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.astr_hip_rk3_steps_stats_noalloc.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.astr_hip_gridgeom.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.astr_hip_gridgeom_nt.argtypes = [C.c_void_p, C.POINTER(C.c_int), _dp, _dp, _dp]
        L.astr_hip_qswap.argtypes = [C.c_void_p]
        L.astr_hip_updatefvar.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int]
        L.astr_hip_set_fvar.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int]
        L.astr_hip_write_checkpoint.argtypes = [C.c_void_p, C.c_char_p, C.c_longlong, C.c_double]
        L.astr_hip_read_checkpoint.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_longlong),
                                               C.POINTER(C.c_double)]
        L.astr_hip_set_metrics.argtypes = [C.c_void_p, _dp, _dp]
        L.astr_hip_set_boundary.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.astr_hip_set_convection.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                              C.c_double]
        L.astr_hip_sync.argtypes = [C.c_void_p]
        L.astr_hip_line_batch.argtypes = [C.c_int, C.c_int, C.c_longlong, _dp, _dp,
                                          C.c_double, C.c_double]
        L.astr_hip_line_batch_nt.argtypes = [C.c_int, C.c_int, C.c_int, C.c_longlong, _dp, _dp,
                                             C.c_double, C.c_double, C.c_double]
        L.astr_hip_profile_begin.argtypes = [C.c_void_p, C.c_int]
        L.astr_hip_profile_end.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int)]
        L.astr_hip_autotune.argtypes = [C.c_void_p, C.c_int]
        L.astr_hip_get_pass_config.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                                               C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.astr_hip_set_pass_config.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.astr_hip_init_vortex.argtypes = [C.c_void_p] + [C.c_double] * 5
        L.astr_hip_set_last_split.argtypes = [C.c_void_p, C.c_int]
        L.astr_hip_get_last_split.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.astr_hip_set_filter_fusion.argtypes = [C.c_void_p, C.c_int]
        L.astr_hip_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
        L.astr_hip_batch_destroy.argtypes = [C.c_void_p]
        L.astr_hip_batch_autotune.argtypes = [C.c_void_p, C.c_int]
        L.astr_hip_batch_rk3_steps.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.astr_hip_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.astr_hip_batch_rk3_begin.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.astr_hip_batch_query.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.astr_hip_batch_rk3_end.argtypes = [C.c_void_p, _dp, _dp]
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise AstrError(lib().astr_hip_last_error().decode())


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def tgv_config(im, jm=None, km=None, deltat=1e-3, reynolds=1600.0, mach=0.1,
               sutherland=110.4, ref_tem=273.15, lx=2 * math.pi, ly=None, lz=None,
               semantics=SEM_MINIAPP):
    """Configuration of the reference's TGV miniapp (tgvsolver.F90:152-261),
    or with semantics=SEM_MAIN the main solver's stage on the same case."""
    jm = im if jm is None else jm
    km = im if km is None else km
    return Config(im, jm, km, semantics, 0.49, 1.11, 1.09, 1.4, mach, reynolds, 0.72,
                  ref_tem, sutherland, lx, lx if ly is None else ly,
                  lx if lz is None else lz, deltat)


def main_config(im, jm=None, km=None, deltat=1e-3, reynolds=1600.0, mach=0.1, ref_tem=273.15,
                lx=2 * math.pi, ly=None, lz=None):
    """The main solver's stage (time_integration_rk, src/mainloop.F90:396-482)
    on a periodic box: Sutherland S = 110.3 K (src/solver.F90:122)."""
    cfg = tgv_config(im, jm, km, deltat=deltat, reynolds=reynolds, mach=mach, sutherland=110.3,
                     ref_tem=ref_tem, lx=lx, ly=ly, lz=lz, semantics=SEM_MAIN)
    cfg.beta_bound = 0.98   # filter_coefficient_cal's beter_bouond (src/comsolver.F90:121)
    return cfg


class Flow:
    """One device-resident case.  Method names follow the reference's
    subroutines (tgvsolver.F90): filterq, gradcal, rhscal, rk3/steploop."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        h = C.c_void_p()
        _chk(lib().astr_hip_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.shape = (cfg.km + 1, cfg.jm + 1, cfg.im + 1)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().astr_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------
    def init_tgv(self, noise=0.0, seed=0):
        _chk(lib().astr_hip_init_tgv(self.h, noise, seed))

    def init_vortex(self, beta=5.0, xc=None, yc=None, u0=1.0, v0=0.0):
        """Two-dimensional isentropic vortex (see include/astr_hip.h)."""
        xc = self.cfg.lx / 2 if xc is None else xc
        yc = self.cfg.ly / 2 if yc is None else yc
        _chk(lib().astr_hip_init_vortex(self.h, beta, xc, yc, u0, v0))

    def halo_shape(self, hm):
        return tuple(s + 2 * hm for s in self.shape)

    def upload(self, q, hm=0):
        """q: (5, km+1+2hm, jm+1+2hm, im+1+2hm) C-order, i.e. the reference's
        q(-hm:im+hm,-hm:jm+hm,-hm:km+hm,1:5); halo cells are not read."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != (5,) + self.halo_shape(hm):
            raise AstrError(f"q must have shape {(5,) + self.halo_shape(hm)}")
        _chk(lib().astr_hip_upload(self.h, _p(q), hm))

    def download(self, prims=False, hm=0):
        """q (and rho,u,v,w,p,T) with halo hm; halo cells hold the periodic copies."""
        q = np.empty((5,) + self.halo_shape(hm))
        pr = np.empty((6,) + self.halo_shape(hm)) if prims else None
        _chk(lib().astr_hip_download(self.h, _p(q), _p(pr), hm))
        return (q, pr) if prims else q

    def set_deltat(self, dt):
        _chk(lib().astr_hip_set_deltat(self.h, dt))

    # -- reference subroutines ----------------------------------------
    def filterq(self):
        _chk(lib().astr_hip_filterq(self.h))

    def gradcal(self):
        dvel = np.empty((9,) + self.shape)
        dtmp = np.empty((3,) + self.shape)
        _chk(lib().astr_hip_gradcal(self.h, _p(dvel), _p(dtmp)))
        return dvel, dtmp

    def rhscal(self):
        qrhs = np.empty((5,) + self.shape)
        _chk(lib().astr_hip_rhscal(self.h, _p(qrhs)))
        return qrhs

    def rk3(self, nsteps, stats=True, dissipation=False):
        """nsteps RK3 steps; (ke, ens) per step, or (ke, ens, dissipation)
        with dissipation=True, or None with stats=False."""
        ke = np.empty(nsteps) if stats else None
        en = np.empty(nsteps) if stats else None
        di = np.empty(nsteps) if stats and dissipation else None
        _chk(lib().astr_hip_rk3_steps_stats(self.h, nsteps, _p(ke), _p(en), _p(di)))
        if not stats:
            return None
        return (ke, en, di) if dissipation else (ke, en)

    steploop = rk3

    def rk3_noalloc(self, nsteps):
        """rk3(nsteps, dissipation=True) through the entry point that allocates
        no device memory (its statistics buffers are the library's fixed
        per-GPU array; calls overlapping on one GPU race)."""
        ke, en, di = np.empty(nsteps), np.empty(nsteps), np.empty(nsteps)
        _chk(lib().astr_hip_rk3_steps_stats_noalloc(self.h, nsteps, _p(ke), _p(en), _p(di)))
        return ke, en, di

    def qswap(self):
        """The main path's single-rank qswap: duplicate-plane averages along
        i, then j, then k (parallel.F90:4876-4884, 5018-5030, 5159-5170)."""
        _chk(lib().astr_hip_qswap(self.h))

    def updatefvar(self, hm=0):
        """updatefvar (fludyna.F90:191-242): (rho, vel, prs, tmp) of the
        current q, numpy (k, j, i) views with halos of depth hm; vel has a
        leading component axis of 3."""
        shp = tuple(m + 2 * hm for m in self.shape)
        rho, prs, tmp = np.empty(shp), np.empty(shp), np.empty(shp)
        vel = np.empty((3,) + shp)
        _chk(lib().astr_hip_updatefvar(self.h, _p(rho), _p(vel), _p(prs), _p(tmp), hm))
        return rho, vel, prs, tmp

    def set_fvar(self, rho, vel, prs, tmp, hm=0):
        """Main path: set the stored primitives (numpy (k, j, i) arrays with
        halos of depth hm; vel with a leading component axis of 3)."""
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, vel, prs, tmp)]
        _chk(lib().astr_hip_set_fvar(self.h, *[_p(x) for x in arr], hm))

    def writechkpt(self, path, nstep=0, time=0.0):
        """Checkpoint of the device state (writechkpt, readwrite.F90:2390)."""
        _chk(lib().astr_hip_write_checkpoint(self.h, os.fsencode(path), int(nstep), float(time)))

    def readcheckpoint(self, path):
        """Resume from a checkpoint; returns (nstep, time) of its header."""
        n, t = C.c_longlong(), C.c_double()
        _chk(lib().astr_hip_read_checkpoint(self.h, os.fsencode(path), C.byref(n), C.byref(t)))
        return n.value, t.value

    # -- grid metrics (gridgeom, geom.F90:99-692) ------------------------------
    def gridgeom(self, x, npdc=None):
        """Metrics of the coordinates x (3, km+1, jm+1, im+1), computed on the
        device: (jacob (km+1, jm+1, im+1), dxi (3, 3, km+1, jm+1, im+1)),
        dxi[a, b] = d xi_a / d x_b.  npdc: boundary type per direction i, j,
        k (3 periodic, the default; 4 physical at both ends)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (3,) + self.shape:
            raise AstrError(f"x must have shape {(3,) + self.shape}")
        jac = np.empty(self.shape)
        buf = np.empty((3, 3) + self.shape)   # [b, a]: the C layout (a-1) + 3(b-1)
        if npdc is None:
            _chk(lib().astr_hip_gridgeom(self.h, _p(x), _p(jac), _p(buf)))
        else:
            nt = (C.c_int * 3)(*[int(v) for v in npdc])
            _chk(lib().astr_hip_gridgeom_nt(self.h, nt, _p(x), _p(jac), _p(buf)))
        return jac, buf.transpose((1, 0, 2, 3, 4))

    def set_boundary(self, npdc):
        """Main path: the boundary type per direction (3 periodic, 4 physical
        at both ends)."""
        _chk(lib().astr_hip_set_boundary(self.h, (C.c_int * 3)(*[int(v) for v in npdc])))

    def set_conschm(self, conschm="643c", recon_schem=0, lchardecomp=False, bfacmpld=0.3, shkcrt=0.05):
        """Main path: the convective scheme as the reference's input file
        names it (conschm, then the line recon_schem, lchardecomp, bfacmpld,
        shkcrt; src/readwrite.F90:541-546).  An even first digit of conschm
        selects the central compact scheme (643c only); an odd first digit
        with suffix 'e' the explicit upwind-biased scheme (convrsduwd,
        src/solver.F90:548-1201); 543c the compact upwind-biased scheme
        (convrsdcmp, src/solver.F90:1271-1937; recon_schem unused)."""
        conschm = conschm.strip()
        if len(conschm) != 4 or not conschm[0].isdigit() or conschm[3] not in "ce":
            raise AstrError(f"conschm {conschm!r}: expected three digits and 'c' or 'e'")
        if int(conschm[0]) % 2 == 0:
            if conschm != "643c":
                raise AstrError(f"central scheme {conschm!r}: only 643c is provided")
            _chk(lib().astr_hip_set_convection(self.h, 0, 0, 0, 0.0, 0.0))
            return
        if conschm[3] == "c" and conschm != "543c":
            raise AstrError(f"conschm {conschm!r}: the compact upwind-biased scheme is defined for 543c only "
                            "(compact_flux_initiate, src/flux.F90:93-110)")
        _chk(lib().astr_hip_set_convection(self.h, 2 if conschm[3] == "c" else 1, int(recon_schem),
                                           int(bool(lchardecomp)), float(bfacmpld), float(shkcrt)))

    def set_metrics(self, jac, dxi):
        """Use the metrics (as gridgeom returns them): those of a Cartesian
        grid, uniform (both semantics) or stretched (main path)."""
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        buf = np.ascontiguousarray(np.asarray(dxi, dtype=np.float64).transpose((1, 0, 2, 3, 4)))
        _chk(lib().astr_hip_set_metrics(self.h, _p(jac), _p(buf)))

    def profile_begin(self, max_launches):
        _chk(lib().astr_hip_profile_begin(self.h, max_launches))

    def profile_end(self):
        """{(pass_name, dir): (total_ms, launches)} of the profiled launches."""
        ms = np.zeros(PROF_SLOTS)
        cnt = (C.c_int * PROF_SLOTS)()
        _chk(lib().astr_hip_profile_end(self.h, _p(ms), cnt))
        out = {}
        for k in range(PROF_SLOTS):
            if cnt[k]:
                out[(PROF_NAMES[k // 3], "xyz"[k % 3])] = (float(ms[k]), int(cnt[k]))
        return out

    def sync(self):
        _chk(lib().astr_hip_sync(self.h))

    # -- launch selection ------------------------------------------------
    def autotune(self, reps=3):
        """Time every fitting (shape, register budget) of every pass and keep
        the fastest; the state is left as it was."""
        _chk(lib().astr_hip_autotune(self.h, reps))
        return self.pass_configs()

    def pass_configs(self):
        """{(pass_name, dir): (shape, minb, tuned_ms)} of the passes this case launches."""
        nd = 3 if self.cfg.km > 0 else 2
        used = {0: [0], 1: list(range(1, nd - 1)), 2: [nd - 1], 3: list(range(nd - 1)),
                4: [nd - 1], 5: [nd - 1], 6: list(range(nd)), 7: [nd - 1], 8: [nd - 1]}
        if self.cfg.semantics != SEM_MAIN:
            used[9] = [0]
        out = {}
        for prog, dirs in used.items():
            for d in dirs:
                sh, mb, ms = C.c_int(), C.c_int(), C.c_float()
                _chk(lib().astr_hip_get_pass_config(self.h, prog, d, C.byref(sh), C.byref(mb),
                                                    C.byref(ms)))
                out[(PASS_NAMES[prog], "xyz"[d])] = (sh.value, mb.value, float(ms.value))
        return out

    def set_pass_config(self, prog, d, shape, minb):
        _chk(lib().astr_hip_set_pass_config(self.h, prog, d, shape, minb))

    @property
    def split_last(self):
        v = C.c_int()
        _chk(lib().astr_hip_get_last_split(self.h, C.byref(v)))
        return bool(v.value)

    @split_last.setter
    def split_last(self, on):
        _chk(lib().astr_hip_set_last_split(self.h, 1 if on else 0))

    def set_filter_fusion(self, on):
        """rk3 steps: fuse filter sweeps into neighbouring passes within a
        call.  Compact-scheme stage: each stage's x sweep inside the next
        stage's first pass (off by default; autotune picks the faster).  Main
        path: the next stage's sweep along the last direction inside each RK
        pass (on by default)."""
        _chk(lib().astr_hip_set_filter_fusion(self.h, 1 if on else 0))


def vortex_config(im, jm=None, mach=0.5, reynolds=1.0e5, length=10.0, cfl=0.67, deltat=None):
    """Two-dimensional periodic isentropic-vortex case on [0,length]^2 (km = 0).
    deltat defaults to the summed-CFL limit cfl for the free stream (1,0) plus
    the vortex's peak swirl (0.8): dt = cfl / sum_d (|u_d| + c)/dx_d."""
    jm = im if jm is None else jm
    c = 1.0 / mach
    if deltat is None:
        deltat = cfl / ((1.0 + 0.8 + c) / (length / im) + (0.8 + c) / (length / jm))
    return Config(im=im, jm=jm, km=0, semantics=0, alfa_filter=0.49, beta_halo=1.11,
                  beta_bound=1.09, gamma=1.4, mach=mach, reynolds=reynolds, prandtl=0.72,
                  ref_tem=273.15, sutherland=110.4, lx=length, ly=length, lz=length,
                  deltat=deltat)


class Batch:
    """Many independent cases advanced together (persistent batched passes,
    optionally one hipGraph per RK3 step).  The Flow objects stay owned by the
    caller and must outlive the batch."""

    def __init__(self, flows):
        self.flows = list(flows)
        arr = (C.c_void_p * len(self.flows))(*[f.h.value for f in self.flows])
        h = C.c_void_p()
        _chk(lib().astr_hip_batch_create(arr, len(self.flows), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().astr_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, reps=3):
        _chk(lib().astr_hip_batch_autotune(self.h, reps))

    def rk3(self, nsteps, graph=True, stats=True):
        """nsteps RK3 steps of every case; (ke, ens) of shape (ncases, nsteps)."""
        n = len(self.flows)
        ke = np.empty((n, nsteps)) if stats else None
        en = np.empty((n, nsteps)) if stats else None
        _chk(lib().astr_hip_batch_rk3_steps(self.h, nsteps, 1 if graph else 0, _p(ke), _p(en)))
        return (ke, en) if stats else None

    def rk3_begin(self, nsteps, graph=True):
        """Enqueue nsteps RK3 steps of every case and return at once."""
        _chk(lib().astr_hip_batch_rk3_begin(self.h, nsteps, 1 if graph else 0))
        self._pending = nsteps

    def done(self):
        """True once the steps of the last rk3_begin have finished."""
        v = C.c_int()
        _chk(lib().astr_hip_batch_query(self.h, C.byref(v)))
        return bool(v.value)

    def rk3_end(self, stats=True):
        """Wait for the steps of the last rk3_begin; their (ke, ens) histories."""
        n, k = len(self.flows), getattr(self, "_pending", 0)
        ke = np.empty((n, k)) if stats else None
        en = np.empty((n, k)) if stats else None
        _chk(lib().astr_hip_batch_rk3_end(self.h, _p(ke), _p(en)))
        self._pending = 0
        return (ke, en) if stats else None

    def info(self):
        g, l = C.c_int(), C.c_int()
        _chk(lib().astr_hip_batch_info(self.h, C.byref(g), C.byref(l)))
        return {"groups": g.value, "launches_per_step": l.value}


def line_batch(op, f, alfa_filter=0.49, beta_halo=1.11):
    """Batched per-line df_compact (op 0) / compact_filter (op 1); f: (nlines, n+11)."""
    f = np.ascontiguousarray(np.atleast_2d(f), dtype=np.float64)
    n = f.shape[1] - 11
    out = np.empty((f.shape[0], n + 1))
    _chk(lib().astr_hip_line_batch(op, n, f.shape[0], _p(f), _p(out), alfa_filter, beta_halo))
    return out


def line_batch_nt(op, ntype, f, alfa_filter=0.49, beta_halo=1.11, beta_bound=0.98):
    """line_batch with the boundary type ntype of the lines' ends (1: physical
    boundary at 0, 2: at n, 3: interfaces, 4: physical at both ends)."""
    f = np.ascontiguousarray(np.atleast_2d(f), dtype=np.float64)
    n = f.shape[1] - 11
    out = np.empty((f.shape[0], n + 1))
    _chk(lib().astr_hip_line_batch_nt(op, ntype, n, f.shape[0], _p(f), _p(out), alfa_filter,
                                      beta_halo, beta_bound))
    return out

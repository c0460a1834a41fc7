"""Child-process runner for the reference Fortran (TEST INFRASTRUCTURE ONLY).

usage: ref_run.py TASK IN.npz OUT.npz   (cwd: a scratch directory)

Loads oracle/_ref/libastr_ref.so (the reference's unmodified hot-path
Fortran plus oracle/ref_driver.F90) and runs one task on one grid.
Array layouts are the reference's: q(0:im,0:jm,0:km,1:5) column-major,
which is a C-order numpy array of shape (5, km+1, jm+1, im+1).
"""
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)


def p(a):
    return a.ctypes.data_as(_dp)


def main():
    task, inp, outp = sys.argv[1:4]
    L = C.CDLL(os.path.join(HERE, "_ref", "libastr_ref.so"))
    z = dict(np.load(inp))
    res = {}
    if task in ("line_deriv", "line_filter"):
        lines = np.atleast_2d(np.asarray(z["f"], dtype=np.float64))
        n = lines.shape[1] - 11
        outs = np.zeros((lines.shape[0], n + 1))
        for r in range(lines.shape[0]):
            f = np.ascontiguousarray(lines[r])
            o = np.zeros(n + 1)
            nt = int(z.get("ntype", 3))
            if task == "line_deriv" and "ntype" not in z:
                L.ref_line_deriv(C.c_int(n), p(f), p(o))
            elif task == "line_deriv":
                L.ref_line_deriv_nt(C.c_int(n), C.c_int(nt), p(f), p(o))
            elif "ntype" not in z:
                L.ref_line_filter(C.c_int(n), C.c_double(float(z.get("alfa", 0.49))), p(f), p(o))
            else:
                L.ref_line_filter_nt(C.c_int(n), C.c_int(nt), C.c_double(float(z["alfa"])),
                                     C.c_double(float(z["beta_halo"])),
                                     C.c_double(float(z["beta_bound"])), p(f), p(o))
            outs[r] = o
        res["out"] = outs
    else:
        im, jm, km = (int(x) for x in z["dims"])
        dt = float(z.get("dt", 1e-3))
        L.ref_field_init(C.c_int(im), C.c_int(jm), C.c_int(km), C.c_double(dt))
        shp = (km + 1, jm + 1, im + 1)
        if "q" in z:
            q = np.ascontiguousarray(z["q"], dtype=np.float64).reshape((5,) + shp)
            L.ref_set_q(p(q))
        if task == "rhscal":
            qrhs = np.zeros((5,) + shp)
            dvel = np.zeros((9,) + shp)
            dtmp = np.zeros((3,) + shp)
            L.ref_rhscal(p(qrhs), p(dvel), p(dtmp))
            res.update(qrhs=qrhs, dvel=dvel, dtmp=dtmp)
        elif task == "filterq":
            L.ref_filterq()
        elif task == "rk3":
            nsteps = int(z["nsteps"])
            t0 = time.perf_counter()
            L.ref_rk3_steps(C.c_int(nsteps))
            res["seconds"] = np.array(time.perf_counter() - t0)
            rows = []
            with open("state.dat") as fh:
                for line in fh:
                    t = line.split()
                    if len(t) >= 4:
                        rows.append([float(x.replace("D", "E")) for x in t[:4]])
            res["hist"] = np.array(rows)
        elif task != "init":
            raise SystemExit("unknown task " + task)
        q = np.zeros((5,) + shp)
        L.ref_get_q(p(q))
        pr = np.zeros((6,) + shp)
        L.ref_get_prims(p(pr))
        res.update(q_out=q, prims=pr)
    np.savez(outp, **res)


if __name__ == "__main__":
    main()

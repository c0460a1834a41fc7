"""Write the golden fixtures of tests/golden from the reference's own Fortran.

Runs the reference hot path (oracle/_ref/libastr_ref.so, built by
oracle/Makefile from the unmodified reference sources) on seeded inputs and
stores inputs and outputs as .npz data:

  lines.npz        df_compact (derivative.F90:171-198) and compact_filter
                   (filter.F90:112-144), ntype=3, on random lines f(-5:n+5)
  stage_12x10x8.npz one bchomo+rhscal (qrhs, dvel, dtmp) and one filterq of
                   the reference miniapp (tgvsolver.F90) on a seeded TGV field
  rk3_16.npz       10 RK3 steps of the miniapp at 16^3, dt=2e-3: stacal
                   history and final q
  hist64.npz       TGV 64^3 from the closed-form field, dt=0.005, 200 steps:
                   KE and enstrophy history (+ final-q checksums)
  hist128_noise.npz TGV 128^3 with seeded 1e-3 Gaussian velocity noise
                   (Philox4x32-10 + Box-Muller, seed 2024), dt=1e-3,
                   1000 steps: KE and enstrophy history (+ final-q checksums)

Only run where the reference tree and its build exist:
    make -C oracle all && python tests/golden/make_golden.py [--skip-hist64] [--skip-hist128]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def lines():
    rng = np.random.default_rng(20260)
    out = {}
    for n in (8, 16, 33, 64, 65, 128, 256):
        f = rng.standard_normal((4, n + 11))
        out[f"f_{n}"] = f
        out[f"deriv_{n}"] = O.run_ref("line_deriv", f=f)["out"]
        out[f"filter_{n}"] = O.run_ref("line_filter", f=f, alfa=0.49)["out"]
    np.savez_compressed(os.path.join(HERE, "lines.npz"), **out)


def seeded_q(im, jm, km, amp, seed):
    s = O.OracleState(O.tgv_cfg(im, jm, km))
    s.tgv_init(amp, seed)
    return s.q.copy()


def stage():
    im, jm, km = 12, 10, 8
    q = seeded_q(im, jm, km, 1e-2, 5)
    r = O.run_ref("rhscal", dims=[im, jm, km], dt=1e-3, q=q)
    fq = O.run_ref("filterq", dims=[im, jm, km], dt=1e-3, q=q)["q_out"]
    np.savez_compressed(os.path.join(HERE, f"stage_{im}x{jm}x{km}.npz"), q=q,
                        qrhs=r["qrhs"], dvel=r["dvel"], dtmp=r["dtmp"], prims=r["prims"],
                        filterq=fq)


def rk3_16():
    q = seeded_q(16, 16, 16, 1e-3, 9)
    r = O.run_ref("rk3", dims=[16, 16, 16], dt=2e-3, q=q, nsteps=10)
    np.savez_compressed(os.path.join(HERE, "rk3_16.npz"), q=q, hist=r["hist"], q_out=r["q_out"])


def hist64():
    r = O.run_ref("rk3", dims=[64, 64, 64], dt=0.005, nsteps=200, timeout=7200)
    qo = r["q_out"]
    np.savez_compressed(os.path.join(HERE, "hist64.npz"), hist=r["hist"],
                        q_sum=qo.reshape(5, -1).sum(axis=1), q_absmax=np.abs(qo).reshape(5, -1).max(axis=1))


def hist128_noise():
    q = seeded_q(128, 128, 128, 1e-3, 2024)
    r = O.run_ref("rk3", dims=[128, 128, 128], dt=1e-3, q=q, nsteps=1000, timeout=4 * 3600)
    qo = r["q_out"]
    np.savez_compressed(os.path.join(HERE, "hist128_noise.npz"), hist=r["hist"],
                        q_sum=qo.reshape(5, -1).sum(axis=1), q_absmax=np.abs(qo).reshape(5, -1).max(axis=1),
                        seconds=r["seconds"])


if __name__ == "__main__":
    only = [a.split("=", 1)[1] for a in sys.argv[1:] if a.startswith("--only=")]
    if only:
        globals()[only[0]]()
        sys.exit(0)
    lines()
    stage()
    rk3_16()
    if "--skip-hist64" not in sys.argv:
        hist64()
    if "--skip-hist128" not in sys.argv:
        hist128_noise()
    print("fixtures written to", HERE)

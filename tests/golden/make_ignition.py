"""Reference states of the 0-D constant-pressure H2-air reactor at t = 1 ms,
from the numpy restatement (oracle/chem.py) integrated by scipy's Radau at
rtol 1e-10, atol 1e-16 (converged: rtol 1e-9 and 1e-11 agree to 1e-12 in T).
Writes tests/golden/ignition_radau.npz: cases (T0, p, phi) and y_end (T, Y).

    python tests/golden/make_ignition.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import chem as OC  # noqa: E402

CASES = [(1000.0, 101325.0, 0.5), (1200.0, 101325.0, 1.0), (1400.0, 5 * 101325.0, 2.0),
         (950.0, 5 * 101325.0, 1.0)]


def mixture(mech, phi):
    ix = {s: k for k, s in enumerate(mech.names)}
    X = np.zeros(mech.ns)
    X[ix["H2"]], X[ix["O2"]], X[ix["N2"]] = 2.0 * phi, 1.0, 3.76
    Y = X * mech.W
    return Y / Y.sum()


def main():
    mech = OC.Mechanism(os.path.join(HERE, "Jachimowski88_H2-9-19.xml"))
    ys = [mech.ignite(T0, p, mixture(mech, phi), 1e-3, rtol=1e-10, atol=1e-16) for T0, p, phi in CASES]
    np.savez(os.path.join(HERE, "ignition_radau.npz"), cases=np.array(CASES), y_end=np.array(ys))


if __name__ == "__main__":
    main()

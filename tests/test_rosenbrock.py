"""Pinning the reactor integrator's method, RODAS4 (Hairer and Wanner,
Solving ODEs II, Sect. VI.4), CPU only: the numpy restatement
oracle/rosenbrock.py against what the method's order conditions and
stability fix, and the HIP library's own coefficient table against it.

  * order: with fixed steps on a smooth nonlinear autonomous system, the
    global error of y_new falls by 2^4 per halving of h; the local error of
    the embedded solution by 2^4 (order 3) and of y_new by 2^5 (order 4); a
    mistyped coefficient breaks an order condition and leaves order <= 2;
  * L-stability: |R(z)| -> 0 as z -> -infinity, |R(iy)| <= 1, and R(-1)
    agrees with exp(-1) to the method's order;
  * the table astr_hip_reactor_method returns equals the restatement's.
"""
import ctypes as C
import os

import numpy as np

from oracle import rosenbrock as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "astr_amd", "libastr_hip.so")


def f(y):
    return np.array([y[1], -np.sin(y[0]) - 0.1 * y[1] * y[1], 1.0 - 0.5 * y[2] * y[0]])


def jac(y):
    return np.array([[0.0, 1.0, 0.0], [-np.cos(y[0]), -0.2 * y[1], 0.0], [-0.5 * y[2], 0.0, -0.5 * y[0]]])


Y0 = np.array([1.0, 0.3, 0.2])


def test_rodas4_global_order_4():
    ref, _ = R.integrate_fixed(f, jac, Y0, 2.0, 4096)
    errs = [np.abs(R.integrate_fixed(f, jac, Y0, 2.0, n)[0] - ref).max() for n in (16, 32, 64)]
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 3.8) and np.all(orders < 4.5), (errs, orders)


def test_rodas4_local_orders_of_both_solutions():
    loc_new, loc_emb = [], []
    for h in (0.2, 0.1, 0.05):
        y1, e1 = R.step(f, jac, Y0, h)
        yr, _ = R.integrate_fixed(f, jac, Y0, h, 64)
        loc_new.append(np.abs(y1 - yr).max())
        loc_emb.append(np.abs(e1 - yr).max())
    o_new = np.log2(loc_new[1] / loc_new[2])
    o_emb = np.log2(loc_emb[1] / loc_emb[2])
    assert o_new > 4.5, loc_new      # local error O(h^5): order 4
    assert 3.7 < o_emb < 4.5, loc_emb  # local error O(h^4): order 3


def test_rodas4_l_stable():
    assert abs(R.stability_function(-1e6)) < 1e-4
    assert abs(R.stability_function(-1e9)) < 1e-7
    for yv in np.linspace(0.0, 50.0, 101):
        assert abs(R.stability_function(1j * yv)) <= 1.0 + 1e-12
    for x in (-0.5, -2.0, -30.0):
        assert abs(R.stability_function(x)) < 1.0
    assert abs(R.stability_function(-0.1) - np.exp(-0.1)) < 1e-6


def test_library_table_equals_restatement():
    L = C.CDLL(LIB)
    L.astr_hip_reactor_method.argtypes = [C.POINTER(C.c_double), C.c_int]
    buf = (C.c_double * 26)()
    assert L.astr_hip_reactor_method(buf, 26) == 0
    got = np.array(buf[:])
    want = [R.GAMMA]
    for i in range(1, 5):
        want += list(R.A[i, :i])
    for i in range(1, 6):
        want += list(R.C[i, :i])
    assert np.array_equal(got, np.array(want))
    assert L.astr_hip_reactor_method(buf, 25) != 0

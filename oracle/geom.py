"""Grid metrics (TEST INFRASTRUCTURE ONLY), restated from the reference's
gridgeom (src/geom.F90:99-692) for a single rank, each direction periodic
(npdc 3) or bounded by physical boundaries at both ends (npdc 4).  The
derivative is the C restatement's df_compact / df_compact_nt
(oracle/oracle.py), itself pinned against the reference's own lines; a
physical direction reads no halo, so the halo extensions below only fill
slots its boundary rows ignore.

Arrays are numpy (k, j, i) views of the reference's x(0:im,0:jm,0:km,1:3):
x has shape (3, km+1, jm+1, im+1); dx[m][d] and dxi[a][b] are 0-based
versions of the reference's dx(...,m+1,d+1) and dxi(...,a+1,b+1).
"""
import numpy as np

from oracle import oracle as O

HM = O.HM
AXIS = {0: 2, 1: 1, 2: 0}   # index direction i, j, k -> numpy axis


def _ext_offset(f):
    """gridsendrecv's single-rank halo of a coordinate line f(0:n)
    (parallel.F90:2859-2864): x(n+h) = x(n) + (x(h) - x(0)),
    x(-h) = x(0) - (x(n) - x(n-h))."""
    n = f.size - 1
    h = np.arange(1, HM + 1)
    left = f[0] - (f[n] - f[n - h])
    right = f[n] + (f[h] - f[0])
    return np.concatenate([left[::-1], f, right])


def _ext_periodic(f):
    """dataswap's single-rank periodic halo: f(-h) = f(n-h), f(n+h) = f(h)."""
    n = f.size - 1
    return np.concatenate([f[n - HM:n], f, f[1:HM + 1]])


def _lines(arr, d):
    """Move direction d's axis last; returns the view and its shape."""
    a = np.moveaxis(arr, AXIS[d], -1)
    return a


def deriv(ext_lines, d, shape, ntype=3):
    """fds%central along direction d of lines already extended to
    f(-hm:n+hm); ext_lines has direction d last; ntype the direction's
    boundary type."""
    out = np.empty(ext_lines.shape[:-1] + (ext_lines.shape[-1] - 2 * HM,))
    for idx in np.ndindex(*ext_lines.shape[:-1]):
        out[idx] = O.df_compact(ext_lines[idx]) if ntype == 3 else O.df_compact_nt(ext_lines[idx], ntype)
    return np.moveaxis(out, -1, AXIS[d])


def _extend(arr, d, how):
    a = _lines(arr, d)
    ext = np.empty(a.shape[:-1] + (a.shape[-1] + 2 * HM,))
    for idx in np.ndindex(*a.shape[:-1]):
        ext[idx] = how(a[idx])
    return ext


# the 18 terms J dxi(r,m) += D_d [0.5 (dx(a,e) x_b - dx(b,e) x_a)], 1-based
# (d, r, m, a, e, b), in the order of geom.F90:395-515
TERMS = [
    (1, 2, 1, 2, 3, 3), (1, 2, 2, 3, 3, 1), (1, 2, 3, 1, 3, 2),
    (1, 3, 1, 3, 2, 2), (1, 3, 2, 1, 2, 3), (1, 3, 3, 2, 2, 1),
    (2, 1, 1, 3, 3, 2), (2, 1, 2, 1, 3, 3), (2, 1, 3, 2, 3, 1),
    (2, 3, 1, 2, 1, 3), (2, 3, 2, 3, 1, 1), (2, 3, 3, 1, 1, 2),
    (3, 1, 1, 2, 2, 3), (3, 1, 2, 3, 2, 1), (3, 1, 3, 1, 2, 2),
    (3, 2, 1, 3, 1, 2), (3, 2, 2, 1, 1, 3), (3, 2, 3, 2, 1, 1),
]


def gridgeom(x, npdc=(3, 3, 3)):
    """(jacob, dxi, dx) of the coordinates x (3, km+1, jm+1, im+1), npdc the
    boundary type of directions i, j, k.  km = 0 selects the 2-D form
    (ndims = 2)."""
    x = np.asarray(x, dtype=np.float64)
    shape = x.shape[1:]
    ndims = 3 if shape[0] > 1 else 2
    dx = np.zeros((3, 3) + shape)
    # dx(m,d) = d x_m / d xi_d (geom.F90:128-161)
    for d in range(ndims):
        for m in range(3):
            dx[m, d] = deriv(_extend(x[m], d, _ext_offset), d, shape, npdc[d])
    if ndims == 3:
        # geom.F90:353-371, summed left to right
        jac = (dx[0, 0] * dx[1, 1] * dx[2, 2] + dx[0, 1] * dx[1, 2] * dx[2, 0]
               + dx[0, 2] * dx[1, 0] * dx[2, 1] - dx[0, 2] * dx[1, 1] * dx[2, 0]
               - dx[0, 1] * dx[1, 0] * dx[2, 2] - dx[0, 0] * dx[1, 2] * dx[2, 1])
        dxi = np.zeros((3, 3) + shape)
        for (d, r, m, a, e, b) in TERMS:
            d, r, m, a, e, b = d - 1, r - 1, m - 1, a - 1, e - 1, b - 1
            dxa = _extend(dx[a, e], d, _ext_periodic)
            dxb = _extend(dx[b, e], d, _ext_periodic)
            xa = _extend(x[a], d, _ext_offset)
            xb = _extend(x[b], d, _ext_offset)
            phi = 0.5 * (dxa * xb - dxb * xa)
            dxi[r, m] = dxi[r, m] + deriv(phi, d, shape, npdc[d])
    else:
        # geom.F90:372-373, 522-530
        jac = dx[0, 0] * dx[1, 1] - dx[0, 1] * dx[1, 0]
        dxi = np.zeros((3, 3) + shape)
        dxi[0, 0] = dx[1, 1]
        dxi[0, 1] = -dx[0, 1]
        dxi[1, 0] = -dx[1, 0]
        dxi[1, 1] = dx[0, 0]
    # geom.F90:664-668
    dxi = dxi / jac
    return jac, dxi, dx


def metric_identity(jac, dxi, npdc=(3, 3, 3)):
    """The discrete metric identities sum_r D_r (J dxi(r,m)), m = 1..3
    (the reference's `can`, geom.F90:549-580), with periodic halos or a
    physical direction's boundary rows."""
    shape = jac.shape
    ndims = 3 if shape[0] > 1 else 2
    out = np.zeros((3,) + shape)
    for m in range(3):
        for r in range(ndims):
            out[m] += deriv(_extend(jac * dxi[r, m], r, _ext_periodic), r, shape, npdc[r])
    return out

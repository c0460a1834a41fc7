"""The batching form: many cases in batched persistent launches, optionally one
hipGraph per RK3 step, against the same cases run one at a time; and the
two-dimensional isentropic vortex initial field.

Tolerances: the batched kernels run the single-case pass body, so results
agree to rounding; 1e-13 relative max-norm per variable, 1e-12 for the
KE/enstrophy histories."""
import numpy as np
import pytest

import astr_amd as A

pytestmark = pytest.mark.gpu

TOL = 1e-13


def relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def mixed_cases():
    """(config, initialiser) of a mixed batch: 3-D TGV of several sizes and
    Reynolds numbers, 2-D vortices, non-square and non-cubic grids."""
    out = []
    for n, re, seed in ((24, 400.0, 1), (33, 1600.0, 2), (40, 2900.0, 3)):
        cfg = A.tgv_config(n, deltat=0.128 / n, reynolds=re)
        out.append((cfg, lambda f, s=seed: f.init_tgv(1e-3, s)))
    cfg = A.tgv_config(28, 20, 24, deltat=4e-3)
    out.append((cfg, lambda f: f.init_tgv(1e-3, 9)))
    for n, m in ((40, 40), (64, 48), (96, 96)):
        cfg = A.vortex_config(n, m)
        out.append((cfg, lambda f: f.init_vortex(5.0)))
    return out


def make(cases, selections=None):
    flows = []
    for k, (cfg, init) in enumerate(cases):
        f = A.Flow(cfg)
        init(f)
        if selections is not None:
            for (name, d), (sh, mb, _) in selections[k].items():
                f.set_pass_config(A.PASS_NAMES.index(name), "xyz".index(d), sh, mb)
        flows.append(f)
    return flows


def check_equal(batch_flows, single_flows):
    for fb, fs in zip(batch_flows, single_flows):
        qb, qs = fb.download(), fs.download()
        for m in range(5):
            scale = max(np.abs(qs[m]).max(), np.abs(qs[1:4]).max() if m in (1, 2, 3) else 0.0)
            assert np.abs(qb[m] - qs[m]).max() <= TOL * scale, m


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_single_cases(graph):
    cases = mixed_cases()
    singles = make(cases)
    ref = [f.rk3(4) for f in singles]
    flows = make(cases)
    b = A.Batch(flows)
    ke, en = b.rk3(4, graph=graph)
    check_equal(flows, singles)
    for k, (ks, es) in enumerate(ref):
        assert relmax(ke[k], ks) < 1e-12 and relmax(en[k], es) < 1e-12, k
    # a second call continues from the state (graph replayed, step counter reset)
    ref2 = [f.rk3(3) for f in singles]
    ke2, en2 = b.rk3(3, graph=graph)
    check_equal(flows, singles)
    for k, (ks, es) in enumerate(ref2):
        assert relmax(ke2[k], ks) < 1e-12 and relmax(en2[k], es) < 1e-12, k


def test_batch_autotune_then_match():
    cases = mixed_cases()
    flows = make(cases)
    before = [f.download() for f in flows]
    b = A.Batch(flows)
    b.autotune(reps=1)
    for f, q in zip(flows, before):
        assert np.array_equal(f.download(), q)
    sel = [f.pass_configs() for f in flows]
    singles = make(cases, sel)
    for f in singles:
        f.rk3(3)
    b.rk3(3, graph=True)
    check_equal(flows, singles)
    assert b.info()["groups"] >= 5


def test_vortex_initial_field():
    """Isentropic (p / rho^gamma = p_inf) everywhere, duplicate planes equal
    to planes 0, and in cyclostrophic balance dp/dr = rho v_theta^2 / r for a
    vortex at rest (fourth-order central differences on a 512^2 grid)."""
    n = 512
    cfg = A.vortex_config(n)
    f = A.Flow(cfg)
    f.init_vortex(5.0, u0=0.0, v0=0.0)
    q, pr = f.download(prims=True)
    rho, u, v, p = pr[0, 0], pr[1, 0], pr[2, 0], pr[4, 0]
    g, ma = cfg.gamma, cfg.mach
    pinf = 1.0 / (g * ma * ma)
    assert np.abs(p / rho ** g - pinf).max() < 1e-12 * pinf
    assert np.array_equal(q[:, :, -1, :], q[:, :, 0, :])
    assert np.array_equal(q[:, :, :, -1], q[:, :, :, 0])
    h = cfg.lx / n
    x = np.arange(n + 1) * h - cfg.lx / 2
    X, Y = np.meshgrid(x, x, indexing="xy")   # arrays are (j, i)
    c = (slice(2, -2), slice(2, -2))
    dpdx = (-p[2:-2, 4:] + 8 * p[2:-2, 3:-1] - 8 * p[2:-2, 1:-3] + p[2:-2, :-4]) / (12 * h)
    dpdy = (-p[4:, 2:-2] + 8 * p[3:-1, 2:-2] - 8 * p[1:-3, 2:-2] + p[:-4, 2:-2]) / (12 * h)
    Xi, Yi = X[c], Y[c]
    r = np.hypot(Xi, Yi)
    sel = (r > 0.5) & (r < 2.5)
    dpdr = (dpdx * Xi + dpdy * Yi)[sel] / r[sel]
    vt = ((-u[c] * Yi + v[c] * Xi)[sel]) / r[sel]
    rhs = rho[c][sel] * vt ** 2 / r[sel]
    assert np.abs(dpdr - rhs).max() < 1e-5 * np.abs(rhs).max()

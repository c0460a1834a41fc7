"""The Fortran host (astr_amd/fortran): its ISO_C_BINDING interfaces match
the C ABI, and on the GPU the reference's own program structure run through
it reproduces the reference's outputs."""
import os
import re
import subprocess

import numpy as np
import pytest

import astr_amd as A
from astr_amd import chem as AC
from oracle import chem as OC
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "astr_amd", "fortran")
GOLD = os.path.join(ROOT, "tests", "golden")
MECH = os.path.join(GOLD, "Jachimowski88_H2-9-19.xml")


def test_bind_c_names_are_declared_and_exported():
    src = open(os.path.join(FDIR, "astr_hip_mod.F90")).read()
    names = set(re.findall(r"name='(astr_hip_\w+)'", src))
    hdr = open(os.path.join(ROOT, "include", "astr_hip.h")).read()
    assert len(names) >= 15
    lib = A.lib()
    for n in sorted(names):
        assert re.search(r"\b" + n + r"\(", hdr), n
        assert hasattr(lib, n), n


def test_fortran_config_matches_c_struct():
    """bind(C) astr_hip_config has the C struct's members in the same order."""
    src = open(os.path.join(FDIR, "astr_hip_mod.F90")).read()
    body = src[src.index("type, bind(C) :: astr_hip_config"):src.index("end type astr_hip_config")]
    fnames = re.findall(r"::\s*([\w ,]+)", body)[1:]
    fnames = [x.strip() for part in fnames for x in part.split(",")]
    cnames = [f[0] for f in A.Config._fields_]
    assert fnames == cnames


def _run(args, cwd):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


@pytest.mark.gpu
def test_tgvsolver_hip_reproduces_state_ref_128(tmp_path):
    """tgvsolver_hip (program boxsolver's call sequence on the device) writes
    state.dat; the reference-held history miniapps/tgv_solver_3d/state.ref_128
    (128^3, dt=1e-3, 100 steps) is matched to 1e-9 relative."""
    _run([os.path.join(FDIR, "tgvsolver_hip"), "128", "128", "128", "100"], tmp_path)
    got = np.loadtxt(tmp_path / "state.dat")
    ref = np.loadtxt(os.path.join(GOLD, "state_ref_128.txt"))
    assert got.shape == ref.shape
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.abs(got[:, 1] - ref[:, 1]).max() < 1e-12
    for c in (2, 3):
        assert np.abs(got[:, c] - ref[:, c]).max() / np.abs(ref[:, c]).max() < 1e-9


@pytest.mark.gpu
def test_chemcells_hip_matches_oracle(tmp_path):
    """chemread / chemrate / cpeval / enthpy / srccomb through the Fortran
    module."""
    m = OC.Mechanism(MECH)
    n = 2000
    T, p, Y = AC.sample_cells(m.names, m.W, n, 21)
    rho = p / (OC.R_GAS * T * (Y / m.W[:, None]).sum(axis=0))
    with open(tmp_path / "cells.txt", "w") as fh:
        fh.write(f"{n}\n")
        for i in range(n):
            fh.write(" ".join(repr(float(x)) for x in [rho[i], T[i], *Y[:, i]]) + "\n")
    # srccomb over a block in the reference's array shapes
    from test_chem import srccomb_block, srccomb_expected
    im, jm, km, hm = 7, 6, 5, 2
    grho, gT, gY, gjac, gq = srccomb_block(m, im, jm, km, hm, 41)
    with open(tmp_path / "grid.txt", "w") as fh:
        fh.write(f"{im} {jm} {km} {hm}\n")
        for a in (grho, gT, gY, gjac, gq):
            fh.write("\n".join(repr(float(x)) for x in a.ravel()) + "\n")
    _run([os.path.join(FDIR, "chemcells_hip"), MECH, "cells.txt", "out.txt", "grid.txt", "qrhs.txt"],
         tmp_path)
    gq1 = np.loadtxt(tmp_path / "qrhs.txt").reshape(gq.shape)
    want = srccomb_expected(m, grho, gT, gY, gjac, gq, hm)
    assert np.array_equal(gq1[:5], gq[:5])
    for s in range(m.ns):
        scale = max(np.abs(want[5 + s] - gq[5 + s]).max(), 1e-300)
        assert np.abs(gq1[5 + s] - want[5 + s]).max() <= 1e-12 * scale + 4e-16 * np.abs(gq[5 + s]).max()
    out = np.loadtxt(tmp_path / "out.txt")
    ns = m.ns
    wi, cp, eng, hi = out[:, :ns].T, out[:, ns], out[:, ns + 1], out[:, ns + 2:2 * ns + 2].T
    te, rgc, xm = out[:, 2 * ns + 2], out[:, 2 * ns + 3], out[:, 2 * ns + 4:].T
    wd0, _, cp0, hk0 = m.rates(T, rho, Y, input_is_p=False)
    for k in range(ns):
        den = np.abs(wd0[k]).max()
        if den == 0:
            assert np.all(wi[k] == 0)
        else:
            assert np.abs(wi[k] - wd0[k]).max() / den < 1e-12, m.names[k]
    assert np.abs(cp / cp0 - 1).max() < 1e-13
    # intEnergy_mass: e = sum_k Y_k (h_k - R T / W_k), h_k at Cantera's R
    e0 = (Y * (hk0 * (OC.R_GAS / OC.R_REF) - OC.R_GAS * T / m.W[:, None])).sum(axis=0)
    assert np.abs(eng - e0).max() / np.abs(e0).max() < 1e-13
    assert np.abs(hi - hk0).max() / np.abs(hk0).max() < 1e-13
    # temperature_calc inverts eng back to T; rgcmix; convertxiyi
    assert np.abs(te / T - 1).max() < 1e-12
    assert np.abs(rgc / m.rgcmix(Y) - 1).max() < 1e-14
    assert np.abs(xm - m.convertxiyi(Y, "Y2X")).max() < 1e-15


DATIN = os.path.join(GOLD, "datin_tgv")


def _edit(src, edits):
    """The text of a datin file with whole lines replaced (old -> new)."""
    lines = open(src).read().split("\n")
    for old, new in edits:
        hit = [i for i, ln in enumerate(lines) if ln.strip() == old]
        assert hit, old
        for i in hit:
            lines[i] = new
    return "\n".join(lines)


def _astr_run(tmp_path, datin=DATIN, inp="input.tgv", edits=(), cont_edits=(), args=None):
    """astr_hip run on a datin folder copied into tmp_path/datin, with lines
    of the input and the controller replaced."""
    d = tmp_path / "datin"
    d.mkdir(exist_ok=True)
    (d / "input.dat").write_text(_edit(os.path.join(datin, inp), edits))
    (d / "controller").write_text(_edit(os.path.join(datin, "controller"), cont_edits))
    return subprocess.run([os.path.join(FDIR, "astr_hip")] + (args or ["run"]),
                          cwd=tmp_path, capture_output=True, text=True, timeout=600)


def test_datin_input_is_read_as_the_reference_reads_it(tmp_path):
    """readinput / readcont (src/readwrite.F90:466-769) on the reference's TGV
    example: grid 128^3, Re 1600, Ma 0.1, T_ref 273.15, alpha_f 0.49,
    maxstep 10, deltat 1e-3 (what follows the parse needs a GPU)."""
    r = _astr_run(tmp_path)
    out = r.stdout
    assert "dimension set: 128 x 128 x 128" in out, out[-2000:]
    vals = out.split("ref_tem, reynolds, mach, alfa_filter:")[1].split("\n")[0].split()
    assert [float(v) for v in vals] == [273.15, 1600.0, 0.1, 0.49]
    assert "maxstep 10, feqchkpt 10, deltat" in out
    assert float(out.split("maxstep 10, feqchkpt 10, deltat")[1].split("\n")[0]) == 1e-3
    # the other two examples this build runs parse too
    r2 = _astr_run(tmp_path, os.path.join(GOLD, "datin_2dvort"), "input.2dvort")
    assert "dimension set: 400 x 200 x 0" in r2.stdout and "diffterm F" in r2.stdout
    r3 = _astr_run(tmp_path, os.path.join(GOLD, "datin_tgv2d"), "input.tgv2d")
    assert "dimension set: 128 x 128 x 0" in r3.stdout


@pytest.mark.parametrize("edit", [("tgv", "chl"), ("643c, 643c, rk3, rk3", "m7c, 643c, rk3, rk3"),
                                  ("643c, 643c, rk3, rk3", "753c, 643c, rk3, rk3"),
                                  ("643c, 643c, rk3, rk3", "642c, 642c, rk3, rk3"),
                                  ("t,t,t,f,f,f,t,f", "t,t,t,t,f,f,t,f")])
def test_datin_unsupported_options_stop(tmp_path, edit):
    """Options the device solver does not run stop with a message and exit
    status 1 before any device work."""
    r = _astr_run(tmp_path, edits=[edit])
    assert r.returncode == 1, r.stdout[-1000:] + r.stderr[-1000:]
    assert "input not supported by the device solver" in r.stdout


def _host_init(n, km=None, lx=2 * np.pi):
    """gridcube + tgvini + updateq (gridgeneration.F90:233-267,
    initialisation.F90:621-691, fludyna.F90:254-360) in numpy, as the Fortran
    host computes them: (x, q, rho, vel, prs, tmp) in (k, j, i) order."""
    km = n if km is None else km
    i = np.arange(n + 1)
    k = np.arange(km + 1)
    xx = lx / n * i
    zz = lx / km * k if km > 0 else np.zeros(1)
    Z, Y, X = np.meshgrid(zz, xx, xx, indexing="ij")
    rho = np.ones_like(X)
    u = 1.0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -1.0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(X)
    const2 = 1.4 * 0.1 ** 2
    const1 = 1.0 / (1.4 * 0.4 * 0.1 ** 2)
    p = 1.0 / const2 + rho / 16.0 * 1.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    t = p / rho * const2
    q = np.array([rho, rho * u, rho * v, rho * w, rho * (t * const1 + 0.5 * (u * u + v * v + w * w))])
    return np.array([X, Y, Z]), q, rho, np.array([u, v, w]), p, t


def _host_init_2d(im, jm):
    """gridcube(2 pi, 2 pi, 2 pi) with ka = 0 + tgvini + updateq in numpy."""
    xx = 2 * np.pi / im * np.arange(im + 1)
    yy = 2 * np.pi / jm * np.arange(jm + 1)
    Y, X = np.meshgrid(yy, xx, indexing="ij")
    X, Y = X[None], Y[None]
    Z = np.zeros_like(X)
    rho = np.ones_like(X)
    u = 1.0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -1.0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(X)
    const2 = 1.4 * 0.1 ** 2
    const1 = 1.0 / (1.4 * 0.4 * 0.1 ** 2)
    p = 1.0 / const2 + rho / 16.0 * 1.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    t = p / rho * const2
    q = np.array([rho, rho * u, rho * v, rho * w, rho * (t * const1 + 0.5 * (u * u + v * v + w * w))])
    return np.array([X, Y, Z]), q, rho, np.array([u, v, w]), p, t


@pytest.mark.gpu
def test_astr_hip_run_tgv_matches_oracle(tmp_path):
    """astr_hip run on the reference's TGV input cut to 32^3 and 6 steps with
    a checkpoint every 3: flowstate.dat (steps 1..6, the main path's
    statistics) and the checkpoint at step 7 against the same run through
    the C ABI from the same initial field (computed here in numpy, whose sin
    and cos may differ from the host's libm in the last bit) and against the
    main-path oracle."""
    r = _astr_run(tmp_path, edits=[("128,128,128", "32,32,32")],
                  cont_edits=[("10,      10,     100,      50,      1,    50",
                               "6,      3,     100,      50,      1,    50")])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    fs = np.loadtxt(tmp_path / "flowstate.dat", skiprows=1)
    assert fs.shape == (6, 5) and list(fs[:, 0]) == [1, 2, 3, 4, 5, 6]
    # the same run through the C ABI
    n = 32
    X, q, rho, vel, prs, tmp = _host_init(n)
    f = A.Flow(A.main_config(n, deltat=1e-3))
    jac, dxi = f.gridgeom(X)
    f.set_metrics(jac, dxi)
    f.upload(q)
    f.set_fvar(rho, vel, prs, tmp)
    ke, en, di = f.rk3(7, dissipation=True)
    for c, v in ((2, ke), (3, en), (4, di)):
        assert np.abs(fs[:, c] / v[1:] - 1).max() < 1e-12
    qc, pc = f.download(prims=True)
    g = A.Flow(A.main_config(n, deltat=1e-3))
    g.set_metrics(jac, dxi)
    nstep, time = g.readcheckpoint(str(tmp_path / "checkpoint" / "flowfield.bin"))
    assert nstep == 7 and abs(time - 7e-3) < 1e-15
    qg, pg = g.download(prims=True)
    assert np.abs(qg - qc).max() / np.abs(qc).max() < 1e-13
    assert np.abs(pg - pc).max() / np.abs(pc).max() < 1e-13
    # the oracle on the same input (the device's metrics means)
    ocfg = O.tgv_cfg(n, dt=1e-3, suth_dim=110.3)
    N = (n + 1) ** 3
    oj = np.full(N, float(np.mean(jac, dtype=np.longdouble)))
    od = np.zeros((9, N))
    for a in range(3):
        od[a + 3 * a] = float(np.mean(dxi[a, a], dtype=np.longdouble))
    s = O.MainState(ocfg, oj, od)
    s.set_q(q)
    ke0, en0, di0 = s.rk3(7)
    for v, v0 in ((ke, ke0), (en, en0), (di, di0)):
        assert np.abs(v / v0 - 1).max() < 1e-9
    qo, _ = s.get()
    assert np.abs(qc - qo).max() / np.abs(qo).max() < 1e-12


@pytest.mark.gpu
def test_astr_hip_restart_continues_bit_for_bit(tmp_path):
    """A run of 6 steps with checkpoints every 3, and a second run restarted
    from the checkpoint at step 4 (lrestart = t, maxstep 6): the restarted
    run's last checkpoint equals the unbroken run's bit for bit."""
    cut = [("10,      10,     100,      50,      1,    50", "6,      3,     100,      50,      1,    50")]
    a_dir = tmp_path / "a"
    a_dir.mkdir()
    r = _astr_run(a_dir, edits=[("128,128,128", "24,24,24")], cont_edits=cut)
    assert r.returncode == 0, r.stdout[-3000:]
    full = (a_dir / "checkpoint" / "flowfield.bin").read_bytes()
    b_dir = tmp_path / "b"
    b_dir.mkdir()
    cut3 = [("10,      10,     100,      50,      1,    50", "3,      3,     100,      50,      1,    50")]
    r = _astr_run(b_dir, edits=[("128,128,128", "24,24,24")], cont_edits=cut3)
    assert r.returncode == 0, r.stdout[-3000:]
    first = (b_dir / "checkpoint" / "flowfield.bin").read_bytes()
    assert first != full
    r = _astr_run(b_dir, edits=[("128,128,128", "24,24,24"), ("f", "t")], cont_edits=cut)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "lrestart T" in r.stdout
    assert (b_dir / "checkpoint" / "flowfield.bin").read_bytes() == full


def _host_init_vort(im, jm, mach=0.3):
    """gridcube(20, 10, 1) with ka = 0 + vortini + updateq
    (gridgeneration.F90:233-267, initialisation.F90:480-533) in numpy."""
    xx = 20.0 / im * np.arange(im + 1)
    yy = 10.0 / jm * np.arange(jm + 1)
    Y, X = np.meshgrid(yy, xx, indexing="ij")
    X, Y = X[None], Y[None]
    Z = np.zeros_like(X)
    xc, yc, rvor = 10.0, 5.0, 0.7
    cvor = 0.1 * rvor
    const2 = 1.4 * mach ** 2
    const1 = 1.0 / (1.4 * 0.4 * mach ** 2)
    radi2 = ((X - xc) ** 2 + (Y - yc) ** 2) / rvor / rvor
    var1 = cvor / rvor / rvor * np.exp(-0.5 * radi2)
    rho = np.ones_like(X)
    u = 1.0 - var1 * (Y - yc)
    v = 0.0 + var1 * (X - xc)
    w = np.zeros_like(X)
    p = 1.0 / const2 - 0.5 * rho * cvor ** 2 / rvor ** 2 * np.exp(-radi2)
    t = p / rho * const2
    q = np.array([rho, rho * u, rho * v, rho * w, rho * (t * const1 + 0.5 * (u * u + v * v + w * w))])
    return np.array([X, Y, Z]), q, rho, np.array([u, v, w]), p, t


@pytest.mark.gpu
@pytest.mark.parametrize("case,conv", [("tgv2d", None), ("2dvort", None), ("2dvort", ("753e", 3, False)),
                                       ("2dvort", ("753e", 5, True)), ("2dvort", ("543c", 3, True))])
def test_astr_hip_runs_2d_examples(tmp_path, case, conv):
    """The reference's 2-D examples (Taylor_Green_Vortex_2D; Vortex_Transport,
    inviscid with diffterm = f) cut to 64 x 48 / 80 x 40 and 5 steps with a
    checkpoint after the last: the host's checkpoint equals the same run
    through the C ABI from the same initial field computed in numpy (to
    1e-13: numpy's sin, cos and exp may differ from libm's in the last bit),
    and the main-path oracle to 1e-12 (1/Re = 0 for the inviscid case).
    conv: the vortex with the convective scheme changed to the explicit
    upwind-biased 753e or the compact upwind-biased 543c, and (recon_schem,
    lchardecomp) on the input line that the example already carries
    (bfacmpld 0.3, shkcrt 0.01)."""
    if case == "tgv2d":
        datin, inp, grid, new = os.path.join(GOLD, "datin_tgv2d"), "input.tgv2d", "128,128,0", "64,48,0"
        cont = [("10,      10,     100,      50,      1,    50", "4,      4,     100,      50,      1,    50")]
        dt, mach, re = 1e-3, 0.1, 1600.0
    else:
        datin, inp, grid, new = os.path.join(GOLD, "datin_2dvort"), "input.2dvort", "400,200,0", "80,40,0"
        cont = [("100,     100,     100,      50,      1,    50", "4,     4,     100,      50,      1,    50")]
        dt, mach, re = 5e-3, 0.3, float("inf")
    edits = [(grid, new)]
    if conv is not None:
        edits += [("643c, 643c, rk3", f"{conv[0]}, 643c, rk3"),
                  ("3, f, 0.3d0, 0.01d0", f"{conv[1]}, {'t' if conv[2] else 'f'}, 0.3d0, 0.01d0")]
    r = _astr_run(tmp_path, datin, inp, edits=edits, cont_edits=cont)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    if conv is not None and conv[0] == "543c":
        assert f"compact upwind-biased convection 543c, lchardecomp {'T' if conv[2] else 'F'}" in r.stdout
    elif conv is not None:
        assert f"upwind-biased convection, recon_schem {conv[1]}" in r.stdout
    im, jm = [int(v) for v in new.split(",")[:2]]
    if case == "tgv2d":
        X, q, rho, vel, prs, tmp = _host_init_2d(im, jm)
        lx, ly = 2 * np.pi, 2 * np.pi
    else:
        X, q, rho, vel, prs, tmp = _host_init_vort(im, jm)
        lx, ly = 20.0, 10.0
    cfg = A.main_config(im, jm, 0, deltat=dt, mach=mach, lx=lx, ly=ly)
    cfg.reynolds = re
    f = A.Flow(cfg)
    jac, dxi = f.gridgeom(X)
    f.set_metrics(jac, dxi)
    f.upload(q)
    f.set_fvar(rho, vel, prs, tmp)
    if conv is not None:
        f.set_conschm(conv[0], conv[1], conv[2], 0.3, 0.01)
    f.rk3(5, stats=False)
    qc, pc = f.download(prims=True)
    g = A.Flow(cfg)
    nstep, _ = g.readcheckpoint(str(tmp_path / "checkpoint" / "flowfield.bin"))
    assert nstep == 5
    qg, pg = g.download(prims=True)
    for grp in (slice(0, 1), slice(1, 3), slice(4, 5)):
        assert np.abs(qg[grp] - qc[grp]).max() / np.abs(qc[grp]).max() < 1e-13
    ocfg = O.tgv_cfg(im, jm, 0, dt=dt, suth_dim=110.3, mach=mach, reynolds=re, lx=lx, ly=ly)
    N = (im + 1) * (jm + 1)
    oj = np.full(N, float(np.mean(jac, dtype=np.longdouble)))
    od = np.zeros((9, N))
    for a in range(2):
        od[a + 3 * a] = float(np.mean(dxi[a, a], dtype=np.longdouble))
    s = O.MainState(ocfg, oj, od)
    if conv is not None:
        s.set_conv(2 if conv[0] == "543c" else 1, conv[1], conv[2], 0.3, 0.01)
    s.set_q(q)
    s.rk3(5)
    qo, _ = s.get()
    for grp in (slice(0, 1), slice(1, 3), slice(4, 5)):
        assert np.abs(qc[grp] - qo[grp]).max() / np.abs(qo[grp]).max() < 1e-12

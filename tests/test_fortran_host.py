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
    """chemread / chemrate / cpeval / enthpy through the Fortran module."""
    m = OC.Mechanism(MECH)
    n = 2000
    T, p, Y = AC.sample_cells(m.names, m.W, n, 21)
    rho = p / (OC.R_GAS * T * (Y / m.W[:, None]).sum(axis=0))
    with open(tmp_path / "cells.txt", "w") as fh:
        fh.write(f"{n}\n")
        for i in range(n):
            fh.write(" ".join(repr(float(x)) for x in [rho[i], T[i], *Y[:, i]]) + "\n")
    _run([os.path.join(FDIR, "chemcells_hip"), MECH, "cells.txt", "out.txt"], tmp_path)
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


def _datin_run(tmp_path, text_edit=None):
    """tgvsolver_hip --datin on the reference's Taylor-Green example
    (examples/Taylor_Green_Vortex/datin, kept as a fixture), optionally
    with one line of the input replaced."""
    inp = os.path.join(DATIN, "input.tgv")
    if text_edit:
        lines = open(inp).read().split("\n")
        old, new = text_edit
        lines = [new if ln.strip() == old else ln for ln in lines]
        inp = str(tmp_path / "input.edit")
        open(inp, "w").write("\n".join(lines))
    return subprocess.run([os.path.join(FDIR, "tgvsolver_hip"), "--datin", inp,
                           os.path.join(DATIN, "controller")],
                          cwd=tmp_path, capture_output=True, text=True, timeout=300)


def test_datin_input_is_read_as_the_reference_reads_it(tmp_path):
    """readinput / readcont (src/readwrite.F90:466-769) on the reference's TGV
    example: grid 128^3, Re 1600, Ma 0.1, T_ref 273.15, alpha_f 0.49,
    maxstep 10, deltat 1e-3 (what follows the parse needs a GPU)."""
    r = _datin_run(tmp_path)
    out = r.stdout
    assert "dimension set: 128 x 128 x 128" in out, out[-2000:]
    vals = out.split("ref_tem, reynolds, mach, alfa_filter:")[1].split("\n")[0].split()
    assert [float(v) for v in vals] == [273.15, 1600.0, 0.1, 0.49]
    assert "maxstep 10, deltat" in out
    assert float(out.split("maxstep 10, deltat")[1].split("\n")[0]) == 1e-3


@pytest.mark.parametrize("edit", [("tgv", "chl"), ("643c, 643c, rk3, rk3", "m7c, 643c, rk3, rk3"),
                                  ("t,t,t,f,f,f,t,f", "t,t,t,t,f,f,t,f")])
def test_datin_unsupported_options_stop(tmp_path, edit):
    """Options the device solver does not run stop with a message and exit
    status 1 before any device work."""
    r = _datin_run(tmp_path, edit)
    assert r.returncode == 1, r.stdout[-1000:] + r.stderr[-1000:]
    assert "input not supported by the device solver" in r.stdout


@pytest.mark.gpu
def test_datin_run_reproduces_state_ref_128(tmp_path):
    """The reference's own TGV input (128^3, Re 1600, Ma 0.1, dt 1e-3, 10 steps
    from its controller) run through the Fortran host on the GPU reproduces the
    first 10 lines of the reference-held state.ref_128 to 1e-9."""
    r = _datin_run(tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = np.loadtxt(tmp_path / "state.dat")
    ref = np.loadtxt(os.path.join(GOLD, "state_ref_128.txt"))[:10]
    assert got.shape == ref.shape
    assert np.array_equal(got[:, 0], ref[:, 0])
    for c in (2, 3):
        assert np.abs(got[:, c] - ref[:, c]).max() / np.abs(ref[:, c]).max() < 1e-9
    # flowstate.dat in the main path's layout (statistic.F90:777-859): the same
    # steps, the same KE and enstrophy to the 14 digits it prints, and the
    # dissipation rate of the same run through the C ABI
    lines = open(tmp_path / "flowstate.dat").read().split("\n")
    assert lines[0].split() == ["nstep", "time", "kenergy", "enstophy", "dissipation"]
    assert len(lines[0]) == 5 * 21
    fs = np.loadtxt(tmp_path / "flowstate.dat", skiprows=1)
    assert fs.shape == (10, 5)
    assert np.array_equal(fs[:, 0], got[:, 0])
    for c in (2, 3):
        assert np.abs(fs[:, c] / got[:, c] - 1).max() < 1e-12
    f = A.Flow(A.tgv_config(128, deltat=1e-3))
    f.init_tgv(0.0, 0)
    ke, en, di = f.rk3(10, dissipation=True)
    assert np.abs(fs[:, 4] / di - 1).max() < 1e-12

"""Checkpoint and restart (writechkpt / readcheckpoint, src/readwrite.F90:2390,
1381) through the C ABI: a run split at step k, written, read into a fresh
context and resumed, equals the unbroken run bit for bit, in both semantics
modes; a checkpoint of another grid or mode, or not a checkpoint, is refused
before the state changes."""
import numpy as np
import pytest

import astr_amd as A

pytestmark = pytest.mark.gpu


def _case(sem, n=24):
    cfg = A.main_config(n, 20, 16, deltat=2e-3) if sem == "main" else A.tgv_config(n, 20, 16, deltat=2e-3)
    f = A.Flow(cfg)
    f.init_tgv(1e-3, 5)
    return f, cfg


@pytest.mark.parametrize("sem", ["miniapp", "main"])
def test_split_run_resumes_bit_for_bit(tmp_path, sem):
    path = str(tmp_path / "flowfield.bin")
    a, cfg = _case(sem)
    ke_a, en_a = a.rk3(6)
    qa, pa = a.download(prims=True)
    b, _ = _case(sem)
    ke_b1, en_b1 = b.rk3(3)
    b.writechkpt(path, 3, 3 * 2e-3)
    b.close()
    c = A.Flow(cfg)
    nstep, time = c.readcheckpoint(path)
    assert nstep == 3 and time == 3 * 2e-3
    ke_c, en_c = c.rk3(3)
    qc, pc = c.download(prims=True)
    assert np.array_equal(qc, qa) and np.array_equal(pc, pa)
    assert np.array_equal(np.concatenate([ke_b1, ke_c]), ke_a)
    assert np.array_equal(np.concatenate([en_b1, en_c]), en_a)


def test_checkpoint_refusals(tmp_path):
    path = str(tmp_path / "flowfield.bin")
    a, _ = _case("main")
    a.writechkpt(path, 0, 0.0)
    other = A.Flow(A.main_config(20, 20, 16, deltat=2e-3))
    other.init_tgv(0.0, 0)
    q0 = other.download()
    with pytest.raises(A.AstrError, match="grid size"):
        other.readcheckpoint(path)
    mini = A.Flow(A.tgv_config(24, 20, 16, deltat=2e-3))
    with pytest.raises(A.AstrError, match="semantics"):
        mini.readcheckpoint(path)
    junk = tmp_path / "junk.bin"
    junk.write_bytes(b"not a checkpoint" * 8)
    with pytest.raises(A.AstrError, match="not an astr_hip checkpoint"):
        other.readcheckpoint(str(junk))
    short = tmp_path / "short.bin"
    short.write_bytes(open(path, "rb").read()[:1000])
    a2, _ = _case("main")
    qa2 = a2.download()
    with pytest.raises(A.AstrError, match="truncated"):
        a2.readcheckpoint(str(short))
    assert np.array_equal(other.download(), q0)
    assert np.array_equal(a2.download(), qa2)

"""The reference's own regression samples, replayed run after run as their
unitTest scripts run pizza.exe (samples/*/unitTest.py), against their
reference.out at the reference tests' rtol 1e-8
(samples/RotE1e3EkPump/unitTest.py:116-120):

- RotE1e3EkPump: collocation n_r 32 CNAB2 -> finite differences n_r 64
  CNAB2 -> collocation n_r 33 BPR353 -> finite differences n_r 64 ARS343,
  each restarting from the previous checkpoint (grid maps between the
  Chebyshev and the stretched finite-difference grids), Ekman pumping;
- DIRKSchemes: ARS222 -> SBDF4 -> ARS443 -> BPR353 (n_r 65) -> PC2 ->
  ARS222 (n_r 49, m_max 48): restarts across scheme families (the CNAB2
  start-up of SBDF4 after a DIRK checkpoint) and grids;
- MultistepSchemes: SBDF3 -> CNAB2 -> SBDF2 -> SBDF4 -> CNLF -> MODCNAB,
  every step logged: restarts that keep, extend and shorten the multistep
  history;
- TestRestart: a restart from n_r 32, m_max 32 onto n_r 36, m_max 64;
- InhomogeneousHeatFlux: fixed-flux walls with an m = 2 outer flux pattern;
- NonRotRa2e3: a non-rotating collocation run (CorFac = 0, no beta, no
  pumping), compared on its 2D kinetic energies (e_kin.TAG), the file its
  unitTest reads;
- TimeStepChange: a non-rotating collocation run with the influence-matrix
  psi solve (matrix_solve = 'influence'), SBDF3, 2000 steps through
  Courant-driven dt changes, on e_kin.TAG.

On the CPU through the oracle; on the GPU through the library and, for
the longest chains, its Fortran driver.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import EKIN_2D, GOLDEN, ROOT, SAMPLES, replay_sample

NAMES = list(SAMPLES)


@pytest.mark.parametrize("name", NAMES)
def test_sample_oracle(name, tmp_path):
    from oracle.oracle import Oracle
    rows, ref = replay_sample(Oracle, name, tmp_path)
    assert rows.shape == ref.shape
    np.testing.assert_allclose(rows, ref, rtol=1e-8, atol=1e-20)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_sample_gpu(name, tmp_path):
    from pizza_amd.runtime import Run
    print(f"case: sample {name} through the library", flush=True)
    rows, ref = replay_sample(Run, name, tmp_path)
    assert rows.shape == ref.shape
    np.testing.assert_allclose(rows, ref, rtol=1e-8, atol=1e-20)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["RotE1e3EkPump", "DIRKSchemes", "NonRotRa2e3", "TimeStepChange"])
def test_sample_fortran_driver(name, tmp_path):
    """pizza_gpu in place of pizza.exe: the input files run one after the
    other in one directory, each reading the checkpoint_end file the
    previous one wrote; the concatenated e_kin_3D files (e_kin for a
    non-rotating sample, which writes no 3D file) match reference.out."""
    driver = os.path.join(ROOT, "pizza_amd", "pizza_gpu")
    files, compared = SAMPLES[name]
    base = os.path.join(GOLDEN, "samples", name)
    tags = []
    for f in files:
        shutil.copy(os.path.join(base, f), tmp_path / f)
        print(f"case: pizza_gpu {name}/{f}", flush=True)
        out = subprocess.run([driver, str(tmp_path / f)], cwd=tmp_path, capture_output=True, text=True,
                             timeout=300)
        assert out.returncode == 0, out.stdout + out.stderr
        from pizza_amd.config import config_from_namelist
        tags.append(config_from_namelist(str(tmp_path / f))[2]["tag"])
    kin = "e_kin" if name in EKIN_2D else "e_kin_3D"
    got = np.concatenate([np.loadtxt(tmp_path / f"{kin}.{t}") for t in tags])
    ref = np.loadtxt(os.path.join(base, "reference.out"))
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-20)
    for t in tags:
        assert (tmp_path / f"frame_us_1.{t}").exists() and (tmp_path / f"checkpoint_end.{t}").exists()
        # the other time series share the log steps of e_kin_3D
        n = len(np.loadtxt(tmp_path / f"{kin}.{t}"))
        # the reference's file names (outputs.f90:79-121)
        series = ["e_kin", "power", "reynolds", "length_scales", "heat"]
        if name not in EKIN_2D:
            series += ["power_3D", "reynolds_3D"]
        else:
            assert not (tmp_path / f"e_kin_3D.{t}").exists()
        for sname in series:
            rows = np.loadtxt(tmp_path / f"{sname}.{t}", ndmin=2)
            assert rows.shape[0] == n, (sname, t)
            assert np.array_equal(rows[:, 0], np.loadtxt(tmp_path / f"{kin}.{t}")[:, 0])

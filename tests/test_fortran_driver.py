"""The Fortran driver (pizza_amd/pizza_gpu): the reference's namelist in,
the reference's e_kin_3D time series out, through ISO_C_BINDING.

CPU: the driver reads a namelist into the same pz_config as the Python
reader, and stops on options the GPU path does not implement.
GPU: 1000 CNAB2 steps of the RotInteg start run reproduce the reference's
golden series (samples/RotInteg/reference.out) at the reference test's
rtol 1e-8 (samples/RotInteg/unitTest.py:116-120).
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN, ROOT

DRIVER = os.path.join(ROOT, "pizza_amd", "pizza_gpu")
NML = os.path.join(GOLDEN, "rotinteg_input_start.nml")


def _driver():
    assert os.path.exists(DRIVER), "pizza_gpu is not built (make -C pizza_amd/fortran)"
    return DRIVER


def _print_config(path, tmp_path):
    out = subprocess.run([_driver(), path, "--print-config"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if len(parts) == 2:
            vals[parts[0]] = parts[1]
    return vals


def _compare(vals, cfg):
    fields = [f[0] for f in type(cfg)._fields_ if not f[0].startswith(("t_", "xi_"))]
    for f in fields:
        want = getattr(cfg, f)
        got = vals[f]
        if isinstance(want, int):
            assert int(got) == want, (f, got, want)
        else:
            assert float(got) == want, (f, got, want)


def test_driver_reads_reference_namelist(tmp_path):
    from pizza_amd.config import config_from_namelist
    cfg, start, out = config_from_namelist(NML)
    vals = _print_config(NML, tmp_path)
    _compare(vals, cfg)
    assert int(vals["init_t"]) == start["init_t"] and float(vals["amp_t"]) == start["amp_t"]
    assert int(vals["n_log_step"]) == out["n_log_step"] and vals["tag"] == out["tag"]


def test_driver_reads_production_namelist(tmp_path):
    """A namelist of the n_r=97 BPR353 thermochemical workload, with groups
    the driver does not use (hdif, output_control) and mixed-case values."""
    from pizza_amd.config import config_from_namelist
    text = """&grid
 n_r_max=97, n_cheb_max=97, m_max=128, minc=1,
/
&control
 tag="c2", time_scheme='bpr353', cheb_method="integ", matrix_solve="Direct",
 dtMax=5d-6, n_time_steps=200, l_rerror_fix=.false.,
/
&hdif
 hdif_temp=0.0,
/
&phys_param
 ra=3.0e7, ek=1.0e-5, pr=1.0, raxi=3.0e7, sc=3.0, radratio=0.35,
 ktopt=1, kbott=1, ktopxi=1, kbotxi=1, ktopv=2, kbotv=2, l_ek_pump=.true., g1=1.0,
/
&start_field
 init_t=0,
/
&output_control
 n_log_step=10, n_frames=1,
/
"""
    path = tmp_path / "c2.nml"
    path.write_text(text)
    cfg, _, _ = config_from_namelist(str(path))
    _compare(_print_config(str(path), tmp_path), cfg)


def test_driver_stops_on_unsupported_options(tmp_path):
    path = tmp_path / "map.nml"
    text = open(NML).read()
    assert "l_newmap   =.false." in text
    path.write_text(text.replace("l_newmap   =.false.", "l_newmap   =.true."))
    out = subprocess.run([_driver(), str(path)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode == 2
    assert "unsupported option: l_newmap" in out.stdout


@pytest.mark.gpu
def test_driver_reproduces_rotinteg_golden(tmp_path):
    print("case: pizza_gpu RotInteg start, 1000 CNAB2 steps", flush=True)
    out = subprocess.run([_driver(), NML, "-n", "1000"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    got = np.loadtxt(tmp_path / "e_kin_3D.start")
    ref = np.loadtxt(os.path.join(GOLDEN, "rotinteg_start_e_kin_3D.txt"))[:51]
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-20)

"""The stage-granular boundary: time_loop's procedures one at a time
(pz_stage_explicit = radial_loop, pz_courant = dt_courant,
pz_stage_implicit = finish_explicit_assembly + mloop, or assemble_stage on
the last stage of an assembly scheme) launch exactly what pz_step launches,
so a run stepped stage by stage equals the captured pz_step bit for bit, on
every radial method and scheme family.  The Fortran driver's stage-loop
procedures (module step_stages) reproduce the pz_step e_kin_3D series."""
import os
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN, ROOT

FIELDS = ("us", "up", "om", "psi", "temp", "xi")


def test_stage_entries_reject_a_null_run():
    """Host-side argument checks (no device call)."""
    import ctypes
    from pizza_amd.runtime import lib
    L = lib()
    assert L.pz_stage_explicit(None, 1) == -1
    assert L.pz_stage_implicit(None, 1, 0) == -1
    assert L.pz_courant(None, None, None) == -1
    assert L.pz_nstages(None) == -1


def _cases():
    from pizza_amd import config as C
    from pizza_amd import workloads as W
    out = []
    c = W.small_bpr353()
    out.append(("integration BPR353", c))
    c = W.c1_collocation()
    out.append(("collocation CNAB2", c))
    c = W.small_bpr353()
    c.l_cheb_coll = 1
    c.time_scheme = C.SBDF3
    out.append(("collocation SBDF3 (CNAB2 start-up)", c))
    c = W.small_bpr353()
    c.l_cheb_coll = 1
    c.time_scheme = C.KC343
    c.dtMax = 2.0e-6
    out.append(("collocation KC343 (assembly stage)", c))
    c = W.small_bpr353(65, 32)
    c.l_finite_diff = 1
    c.time_scheme = C.SBDF3
    c.dtMax = 2.0e-6
    out.append(("finite differences SBDF3", c))
    c = W.small_bpr353(65, 32)
    c.l_finite_diff = 1
    c.time_scheme = C.ARS343
    c.dtMax = 2.0e-6
    out.append(("finite differences ARS343", c))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("k", range(6))
def test_stage_by_stage_equals_pz_step(k):
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    name, cfg = _cases()[k]
    print(f"case: {name}", flush=True)
    a, b = Run(cfg), Run(cfg)
    for x in (a, b):
        W.seed_run(x, cfg)
        x.start()
    for _ in range(6):
        a.step(1)
        b.step_by_stages()
    assert a.time() == b.time()
    for f in FIELDS:
        if (f == "temp" and not cfg.ra) or (f == "xi" and not cfg.raxi):
            continue
        assert np.array_equal(a.get_field(f), b.get_field(f)), f
    # and the stage calls interleave with pz_step on one run
    a.step(2)
    b.step(1)
    b.step_by_stages()
    for f in ("temp", "om"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f


@pytest.mark.gpu
def test_courant_limits_and_forced_rebuild():
    """pz_courant returns the step's radial and horizontal Courant limits;
    a rebuild forced through pz_stage_implicit(lMat = 1) at an unchanged dt
    rebuilds the same matrices, so the result is unchanged."""
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    cfg = W.c1_collocation()
    a, b = Run(cfg), Run(cfg)
    for x in (a, b):
        W.seed_run(x, cfg, amp_t=1e-3)
        x.start()
        x.step(3)
    a.step(1)
    b.stage_explicit(1)
    dtr, dth = b.courant(want=True)
    assert 0.0 < dtr < 1e300 and 0.0 < dth < 1e300
    b.stage_implicit(1, lMat=1)
    assert a.time() == b.time()
    for f in ("temp", "om", "psi"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f


@pytest.mark.gpu
def test_fortran_stage_procedures_match_pz_step(tmp_path):
    """pizza_gpu --stages (radial_loop, dt_courant, finish_explicit_assembly,
    mloop of module step_stages) writes the same e_kin_3D series as the
    captured pz_step, and both match the reference's golden rows."""
    driver = os.path.join(ROOT, "pizza_amd", "pizza_gpu")
    nml = os.path.join(GOLDEN, "rotinteg_input_start.nml")
    rows = []
    for mode, sub in ((["--stages"], "s"), ([], "g")):
        d = tmp_path / sub
        d.mkdir()
        out = subprocess.run([driver, nml, "-n", "200"] + mode, cwd=d, capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stdout + out.stderr
        rows.append(np.loadtxt(d / "e_kin_3D.start"))
    assert np.array_equal(rows[0], rows[1])
    ref = np.loadtxt(os.path.join(GOLDEN, "rotinteg_start_e_kin_3D.txt"))[:len(rows[0])]
    np.testing.assert_allclose(rows[0], ref, rtol=1e-8, atol=1e-20)

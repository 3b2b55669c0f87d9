"""Shared helpers of the test-suite."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def rel(a, b):
    """max |a - b| / max |b| over a whole field."""
    a, b = np.asarray(a), np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def rotinteg_config():
    from pizza_amd.config import config_from_namelist
    return config_from_namelist(os.path.join(GOLDEN, "rotinteg_input_start.nml"))


def golden_rows(n):
    return np.loadtxt(os.path.join(GOLDEN, "rotinteg_start_e_kin_3D.txt"))[:n]


def sample_config(name):
    """(cfg, start, output) of a reference sample's start namelist kept in
    tests/golden/<name>_input_start.nml."""
    from pizza_amd.config import config_from_namelist
    return config_from_namelist(os.path.join(GOLDEN, f"{name}_input_start.nml"))


def sample_series(run, name, steps_per_row):
    """Initialises `run` (a Run or an oracle) from the sample's start
    namelist, steps it, and returns (rows, golden): the e_kin_3D rows at the
    golden file's log steps and the golden rows themselves."""
    _, start, _ = sample_config(name)
    ref = np.loadtxt(os.path.join(GOLDEN, f"{name}_start_e_kin_3D.txt"))
    run.init_fields(init_t=start["init_t"], amp_t=start["amp_t"])
    run.start()
    rows = []
    for k in range(len(ref)):
        if k:
            run.step(steps_per_row)
        t, _ = run.time()
        rows.append(np.r_[t, run.ekin_3d()][: ref.shape[1]])
    return np.array(rows), ref

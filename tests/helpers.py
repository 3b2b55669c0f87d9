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

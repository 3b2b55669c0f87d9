"""The C-ABI library loads and exports every entry point the header declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from helpers import ROOT


def test_library_exports_header_symbols():
    from pizza_amd.runtime import LIB_PATH, EXPORTS
    assert os.path.exists(LIB_PATH), "libpizza_amd.so is not built"
    header = open(os.path.join(ROOT, "include", "pizza_amd.h")).read()
    declared = set(re.findall(r"\b(pz_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations found"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(EXPORTS)


def test_config_struct_matches_defaults():
    """pz_config_default (host code, no device) fills the namelist defaults."""
    from pizza_amd.config import PzConfig, default_config
    from pizza_amd.runtime import lib
    a = PzConfig()
    lib().pz_config_default(ctypes.byref(a))
    b = default_config()
    assert bytes(a) == bytes(b)


def test_namelist_reader():
    from helpers import rotinteg_config
    cfg, start, out = rotinteg_config()
    assert (cfg.n_r_max, cfg.n_cheb_max, cfg.m_max) == (49, 32, 32)
    assert cfg.l_cheb_coll == 0 and cfg.l_direct_solve == 1 and cfg.time_scale_rot == 1
    assert cfg.l_rerror_fix == 1 and cfg.courfac == 4.0 and cfg.dtMax == 0.1
    assert start["init_t"] == -1 and out["n_log_step"] == 20

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


# The reference's samples with their runs in the order of their unitTest
# (samples/<name>/unitTest.py) and the run whose e_kin_3D rows
# reference.out holds (None: all runs, concatenated).
SAMPLES = {
    "RotE1e3EkPump": (["input.nml", "input_FD.nml", "input_cheb.nml", "input_final.nml"], None),
    "DIRKSchemes": (["input.nml", "input_first_restart.nml", "input_second_restart.nml", "input_third_restart.nml",
                     "input_fourth_restart.nml", "input_final.nml"], None),
    "MultistepSchemes": (["input.nml", "input_first_restart.nml", "input_second_restart.nml",
                          "input_third_restart.nml", "input_fourth_restart.nml", "input_final.nml"], None),
    "TestRestart": (["inputStart.nml", "input.nml"], "input.nml"),
    "InhomogeneousHeatFlux": (["input.nml"], None),
    "NonRotRa2e3": (["input.nml"], None),
    "TimeStepChange": (["input.nml"], None),
}

# Samples whose unitTest compares the 2D kinetic energies (e_kin.TAG:
# time, us2_2D, up2_2D, up2_axi_2D, outputs.f90:677) instead of e_kin_3D:
# the non-rotating runs, which write no 3D series
EKIN_2D = {"NonRotRa2e3", "TimeStepChange"}


def _ekin_2d(run):
    """time, us2_2D, up2_2D, up2_axi_2D of the run's time series (the
    oracle returns the series with its spectra, the library the series)."""
    ts = run.time_series()
    return np.asarray(ts[0] if isinstance(ts, tuple) else ts)[:4]


def replay_sample(make, name, workdir):
    """Runs a sample's namelists one after the other as its unitTest runs
    pizza.exe: each run starts from the namelist's initial conditions or
    from the checkpoint_end file the previous run wrote in workdir, writes
    an e_kin_3D row on every log step (l_correct_step, the first and the
    last step included) and leaves its own checkpoint_end.<tag>.  make(cfg)
    builds a Run or an oracle.  Returns (rows, golden rows)."""
    from pizza_amd.config import config_from_namelist
    files, compared = SAMPLES[name]
    base = os.path.join(GOLDEN, "samples", name)
    out_rows = []
    for f in files:
        cfg, start, out = config_from_namelist(os.path.join(base, f))
        run = make(cfg)
        if start.get("l_start_file"):
            run.read_checkpoint(os.path.join(str(workdir), start["start_file"]))
        run.init_fields(init_t=start["init_t"], amp_t=start["amp_t"], init_xi=start["init_xi"],
                        amp_xi=start["amp_xi"], init_u=start["init_u"], amp_u=start["amp_u"])
        run.start()
        n_steps, n_log = int(out["n_time_steps"]), int(out["n_log_step"])
        rows, n = [], 0
        while True:
            t, _ = run.time()
            rows.append(_ekin_2d(run) if name in EKIN_2D else np.r_[t, run.ekin_3d()])
            if n >= n_steps:
                break
            k = min(n_log - n % n_log, n_steps - n)
            run.step(k)
            n += k
        run.write_checkpoint(os.path.join(str(workdir), f"checkpoint_end.{out['tag']}"))
        if compared is None or f == compared:
            out_rows.append(np.array(rows))
    return np.concatenate(out_rows), np.loadtxt(os.path.join(base, "reference.out"))

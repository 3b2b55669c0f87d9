"""ctypes binding of libpizza_amd (the C ABI in include/pizza_amd.h).

This is the Python counterpart of the Fortran driver's ISO_C_BINDING layer:
one ``Run`` object per simulation, mirroring the reference's module
procedures of the time-step path (radial_loop.f90:231, mloop.f90:36,
step_time.f90:215) behind pz_step, and its unit operators (costf1,
get_dr/get_ddr, fft/ifft).  The library must be built (``make -C
pizza_amd/csrc``); nothing here falls back to a CPU path.
"""
import ctypes
import os

import numpy as np

from .config import FIELDS, PzConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpizza_amd.so")
_LIB = None


class PizzaError(RuntimeError):
    pass


class PzSeries(ctypes.Structure):
    """struct pz_series (include/pizza_amd.h)."""
    _fields_ = [("time", ctypes.c_double), ("ekin_2d", ctypes.c_double * 3), ("ekin_3d", ctypes.c_double * 4),
                ("power_2d", ctypes.c_double * 5), ("power_3d", ctypes.c_double * 5),
                ("reynolds_2d", ctypes.c_double * 3), ("reynolds_3d", ctypes.c_double * 3),
                ("lengths", ctypes.c_double * 5), ("heat", ctypes.c_double * 7), ("chem", ctypes.c_double * 6)]


class PzCheckpointInfo(ctypes.Structure):
    """struct pz_checkpoint_info (include/pizza_amd.h)."""
    _fields_ = [("version", ctypes.c_int), ("time", ctypes.c_double), ("l_cheb_coll", ctypes.c_int),
                ("family", ctypes.c_char * 11), ("nexp", ctypes.c_int), ("nimp", ctypes.c_int),
                ("nold", ctypes.c_int), ("n_dt", ctypes.c_int), ("dt", ctypes.c_double * 16),
                ("ra", ctypes.c_double), ("pr", ctypes.c_double), ("raxi", ctypes.c_double),
                ("sc", ctypes.c_double), ("ek", ctypes.c_double), ("stef", ctypes.c_double),
                ("radratio", ctypes.c_double), ("n_r_max", ctypes.c_int), ("m_max", ctypes.c_int),
                ("minc", ctypes.c_int), ("rscheme", ctypes.c_char * 73), ("n_max", ctypes.c_int),
                ("order_boundary", ctypes.c_int), ("alph1", ctypes.c_double), ("alph2", ctypes.c_double),
                ("l_heat", ctypes.c_int), ("l_chem", ctypes.c_int), ("l_phase_field", ctypes.c_int),
                ("header_bytes", ctypes.c_longlong), ("n_arrays", ctypes.c_int)]


def lib():
    """Loads libpizza_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise PizzaError(f"{LIB_PATH} is missing: build it with make -C pizza_amd/csrc")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.pz_config_default.argtypes = [ctypes.POINTER(PzConfig)]
        L.pz_config_default.restype = None
        L.pz_create.argtypes = [ctypes.POINTER(PzConfig), ctypes.POINTER(vp)]
        L.pz_destroy.argtypes = [vp]
        L.pz_last_error.argtypes = [vp]
        L.pz_last_error.restype = ctypes.c_char_p
        L.pz_dims.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ip)]
        L.pz_set_field.argtypes = [vp, ip, dp]
        L.pz_get_field.argtypes = [vp, ip, dp]
        L.pz_init_fields.argtypes = [vp, ip, ctypes.c_double, ip, ctypes.c_double, ip, ctypes.c_double]
        L.pz_start.argtypes = [vp]
        L.pz_step.argtypes = [vp, ip]
        L.pz_time.argtypes = [vp, dp, dp]
        L.pz_nstages.argtypes = [vp]
        L.pz_stage_explicit.argtypes = [vp, ip]
        L.pz_courant.argtypes = [vp, dp, dp]
        L.pz_stage_implicit.argtypes = [vp, ip, ip]
        L.pz_ekin_3d.argtypes = [vp, dp]
        L.pz_nusselt.argtypes = [vp, dp]
        L.pz_stream.argtypes = [vp]
        L.pz_stream.restype = vp
        L.pz_costf.argtypes = [vp, dp, ip]
        L.pz_get_dr.argtypes = [vp, dp, dp, ip, ip]
        L.pz_get_ddr.argtypes = [vp, dp, dp, dp, ip, ip]
        L.pz_fft.argtypes = [vp, dp, dp, ip]
        L.pz_ifft.argtypes = [vp, dp, dp, ip]
        L.pz_get_ddr_noalloc.argtypes = [vp, vp, vp, vp, ip, ip]
        L.pz_write_checkpoint.argtypes = [vp, ctypes.c_char_p]
        L.pz_read_checkpoint.argtypes = [vp, ctypes.c_char_p]
        L.pz_read_checkpoint_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(PzCheckpointInfo), dp, ip]
        L.pz_write_frame.argtypes = [vp, ip, ctypes.c_char_p]
        L.pz_time_series.argtypes = [vp, ctypes.POINTER(PzSeries)]
        L.pz_spectra.argtypes = [vp, dp, dp, dp, dp, dp]
        L.pz_time_kernel.argtypes = [vp, ip, ip, dp]
        L.pz_ensemble_create.argtypes = [ctypes.POINTER(vp), ip, ctypes.POINTER(vp)]
        L.pz_ensemble_step.argtypes = [vp, ip]
        L.pz_ensemble_destroy.argtypes = [vp]
        L.pz_ensemble_last_error.argtypes = [vp]
        L.pz_ensemble_last_error.restype = ctypes.c_char_p
        L.pz_ensemble_policy.argtypes = [vp, ctypes.POINTER(ip), dp]
        L.pz_adjoint_forward.argtypes = [vp, ip, dp, ip, ip, dp]
        L.pz_adjoint_backward.argtypes = [vp, ctypes.c_double, dp, dp]
        _LIB = L
    return _LIB


EXPORTS = ["pz_config_default", "pz_create", "pz_destroy", "pz_last_error", "pz_dims",
           "pz_set_field", "pz_get_field", "pz_init_fields", "pz_start", "pz_step", "pz_nstages",
           "pz_stage_explicit", "pz_courant", "pz_stage_implicit", "pz_write_checkpoint",
           "pz_read_checkpoint", "pz_read_checkpoint_header", "pz_write_frame", "pz_time_series", "pz_spectra",
           "pz_time", "pz_ekin_3d", "pz_nusselt", "pz_stream", "pz_costf", "pz_get_dr", "pz_get_ddr",
           "pz_fft", "pz_ifft", "pz_get_ddr_noalloc", "pz_time_kernel", "pz_ensemble_create", "pz_ensemble_step",
           "pz_ensemble_destroy", "pz_ensemble_last_error", "pz_ensemble_policy", "pz_adjoint_forward", "pz_adjoint_backward"]

OBJECTIVES = {"nutop": 0, "ekin": 1, "temp": 2}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def checkpoint_header(path):
    """(PzCheckpointInfo, stored radii) of a checkpoint file (host only)."""
    info = PzCheckpointInfo()
    r = np.zeros(100000)
    rc = lib().pz_read_checkpoint_header(str(path).encode(), ctypes.byref(info), _dp(r), len(r))
    if rc:
        raise PizzaError(f"pz_read_checkpoint_header failed ({rc}) on {path}")
    return info, r[:info.n_r_max].copy()


class Run:
    """One simulation on the current HIP device."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = ctypes.c_void_p()
        rc = lib().pz_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise PizzaError(f"pz_create failed ({rc})")
        self.h = h
        nm, nr, npf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().pz_dims(h, ctypes.byref(nm), ctypes.byref(nr), ctypes.byref(npf))
        self.n_m, self.n_r, self.n_phi = nm.value, nr.value, npf.value

    def close(self):
        if getattr(self, "h", None):
            lib().pz_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = lib().pz_last_error(self.h)
            raise PizzaError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_field(self, name, arr):
        """arr: complex (n_m, n_r)."""
        a = np.asarray(arr, dtype=np.complex128)
        buf = np.ascontiguousarray(a.T)
        self._check(lib().pz_set_field(self.h, FIELDS[name], _dp(buf.view(np.float64))), "pz_set_field")

    def get_field(self, name):
        buf = np.zeros((self.n_r, self.n_m), dtype=np.complex128)
        self._check(lib().pz_get_field(self.h, FIELDS[name], _dp(buf.view(np.float64))), "pz_get_field")
        return buf.T.copy()

    def init_fields(self, init_t=0, amp_t=0.0, init_xi=0, amp_xi=0.0, init_u=0, amp_u=0.0):
        self._check(lib().pz_init_fields(self.h, init_t, amp_t, init_xi, amp_xi, init_u, amp_u),
                    "pz_init_fields")

    def start(self):
        self._check(lib().pz_start(self.h), "pz_start")

    def step(self, n=1):
        self._check(lib().pz_step(self.h, n), "pz_step")

    # time_loop's procedures one at a time (pz_step(h, 1) launch for launch)
    def nstages(self):
        return lib().pz_nstages(self.h)

    def stage_explicit(self, stage):
        self._check(lib().pz_stage_explicit(self.h, stage), "pz_stage_explicit")

    def courant(self, want=False):
        """Stage-1 dt_courant; returns (dt_r, dt_h) when want is true."""
        if not want:
            self._check(lib().pz_courant(self.h, None, None), "pz_courant")
            return None
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().pz_courant(self.h, ctypes.byref(a), ctypes.byref(b)), "pz_courant")
        return a.value, b.value

    def stage_implicit(self, stage, lMat=0):
        self._check(lib().pz_stage_implicit(self.h, stage, int(lMat)), "pz_stage_implicit")

    def step_by_stages(self):
        """One step through the stage-granular entry points."""
        for s in range(1, self.nstages() + 1):
            self.stage_explicit(s)
            if s == 1:
                self.courant()
            self.stage_implicit(s)

    def write_checkpoint(self, path):
        """Version-7 checkpoint file (checkpoints.f90:33-183)."""
        self._check(lib().pz_write_checkpoint(self.h, str(path).encode()), "pz_write_checkpoint")

    def read_checkpoint(self, path):
        """Reads a checkpoint into the run; start() must follow."""
        self._check(lib().pz_read_checkpoint(self.h, str(path).encode()), "pz_read_checkpoint")

    def time_series(self):
        """The 42 log-step values of get_time_series in the layout of
        or_time_series: [time, e_kin_2D(3), e_kin_3D(4), power_2D(5),
        power_3D(5), reynolds_2D(3), reynolds_3D(3), length_scales(5),
        heat(7), chem(6)]."""
        ts = PzSeries()
        self._check(lib().pz_time_series(self.h, ctypes.byref(ts)), "pz_time_series")
        return np.r_[ts.time, list(ts.ekin_2d), list(ts.ekin_3d), list(ts.power_2d), list(ts.power_3d),
                     list(ts.reynolds_2d), list(ts.reynolds_3d), list(ts.lengths), list(ts.heat), list(ts.chem)]

    def spectra(self):
        """(5, n_m): us2_m, up2_m, enst_m, temp2_m, xi2_m of calculate_spectra."""
        out = np.zeros((5, self.n_m))
        self._check(lib().pz_spectra(self.h, *[_dp(out[j]) for j in range(5)]), "pz_spectra")
        return out

    def write_frame(self, name, path):
        """Frame file of one field (output_frames.f90:96-167)."""
        self._check(lib().pz_write_frame(self.h, FIELDS[name], str(path).encode()), "pz_write_frame")

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        self._check(lib().pz_time(self.h, ctypes.byref(t), ctypes.byref(dt)), "pz_time")
        return t.value, dt.value

    def ekin_3d(self):
        out = np.zeros(4)
        self._check(lib().pz_ekin_3d(self.h, _dp(out)), "pz_ekin_3d")
        return out

    def nusselt(self):
        """(NuTop, NuBot)."""
        out = np.zeros(2)
        self._check(lib().pz_nusselt(self.h, _dp(out)), "pz_nusselt")
        return out

    def stream(self):
        return lib().pz_stream(self.h)

    def adjoint_forward(self, objective, nsteps, ckpt_every=20, weights=None):
        """Runs nsteps steps from the started state keeping snapshots every
        ckpt_every steps; returns the objective J ('nutop': time-averaged
        outer-wall Nusselt number; 'ekin': time-averaged kinetic energy
        us2_3D + up2_3D + uz2_3D; 'temp': sum Re(conj(weights) T) after the
        last step, weights complex (n_m, n_r))."""
        J = ctypes.c_double()
        wp = None
        if weights is not None:
            self._wbuf = np.ascontiguousarray(np.asarray(weights, dtype=np.complex128).T)
            wp = _dp(self._wbuf.view(np.float64))
        self._check(lib().pz_adjoint_forward(self.h, OBJECTIVES[objective], wp, nsteps, ckpt_every,
                                             ctypes.byref(J)), "pz_adjoint_forward")
        return J.value

    def adjoint_backward(self, dJ=1.0):
        """dJ times the gradient of the last forward window's objective:
        (d/dT of the temperature given to start(), complex (n_m, n_r);
        d/d of the boundary values of T, complex (n_m, 2): outer, inner)."""
        g = np.zeros((self.n_r, self.n_m), dtype=np.complex128)
        bc = np.zeros((2, self.n_m), dtype=np.complex128)
        self._check(lib().pz_adjoint_backward(self.h, float(dJ), _dp(g.view(np.float64)), _dp(bc.view(np.float64))),
                    "pz_adjoint_backward")
        return g.T.copy(), bc.T.copy()

    def time_kernel(self, which, reps):
        ms = ctypes.c_double()
        self._check(lib().pz_time_kernel(self.h, which, reps, ctypes.byref(ms)), "pz_time_kernel")
        return ms.value

    # unit operators on complex (n_m, n_r) arrays
    def costf(self, f):
        f = np.asarray(f, dtype=np.complex128)
        buf = np.ascontiguousarray(f.T)
        self._check(lib().pz_costf(self.h, _dp(buf.view(np.float64)), f.shape[0]), "pz_costf")
        return buf.T.copy()

    def get_dr(self, f, l_dct_in=True):
        f = np.asarray(f, dtype=np.complex128)
        buf = np.ascontiguousarray(f.T)
        out = np.zeros_like(buf)
        self._check(lib().pz_get_dr(self.h, _dp(buf.view(np.float64)), _dp(out.view(np.float64)),
                                    f.shape[0], int(l_dct_in)), "pz_get_dr")
        return out.T.copy()

    def get_ddr(self, f, l_dct_in=True):
        f = np.asarray(f, dtype=np.complex128)
        buf = np.ascontiguousarray(f.T)
        d1, d2 = np.zeros_like(buf), np.zeros_like(buf)
        self._check(lib().pz_get_ddr(self.h, _dp(buf.view(np.float64)), _dp(d1.view(np.float64)),
                                     _dp(d2.view(np.float64)), f.shape[0], int(l_dct_in)), "pz_get_ddr")
        return d1.T.copy(), d2.T.copy()

    def fft(self, grid):
        g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64).T)
        nrad = g.shape[0]
        out = np.zeros((nrad, self.n_m), dtype=np.complex128)
        self._check(lib().pz_fft(self.h, _dp(g), _dp(out.view(np.float64)), nrad), "pz_fft")
        return out.T.copy()

    def ifft(self, spec):
        s = np.ascontiguousarray(np.asarray(spec, dtype=np.complex128).T)
        nrad = s.shape[0]
        out = np.zeros((nrad, self.n_phi))
        self._check(lib().pz_ifft(self.h, _dp(s.view(np.float64)), _dp(out), nrad), "pz_ifft")
        return out.T.copy()


class Ensemble:
    """Started runs advanced together (pz_ensemble_*): one batched launch
    per kernel and stage for all members.  The members stay usable as
    ordinary runs (get_field, time, ekin_3d) and must outlive the ensemble."""

    def __init__(self, runs):
        self.runs = list(runs)
        arr = (ctypes.c_void_p * len(self.runs))(*[r.h for r in self.runs])
        e = ctypes.c_void_p()
        rc = lib().pz_ensemble_create(arr, len(self.runs), ctypes.byref(e))
        if rc:
            raise PizzaError(f"pz_ensemble_create failed ({rc})")
        self.e = e

    def step(self, n=1):
        rc = lib().pz_ensemble_step(self.e, n)
        if rc:
            msg = lib().pz_ensemble_last_error(self.e)
            raise PizzaError(f"pz_ensemble_step failed ({rc}): {msg.decode() if msg else ''}")

    def policy(self):
        """(policy picked by the autotuner, trial step times in ms)."""
        pol = ctypes.c_int()
        ms = np.zeros(4)
        lib().pz_ensemble_policy(self.e, ctypes.byref(pol), _dp(ms))
        return pol.value, ms

    def close(self):
        if getattr(self, "e", None):
            lib().pz_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

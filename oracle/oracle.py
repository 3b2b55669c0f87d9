"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU restatement of the reference's
IMEX step.  Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline; never by the product library.
"""
import ctypes
import os

import numpy as np

from pizza_amd.config import FIELDS, PzConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.or_create.argtypes = [ctypes.POINTER(PzConfig), ctypes.POINTER(vp)]
        L.or_destroy.argtypes = [vp]
        L.or_destroy.restype = None
        L.or_dims.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ip)]
        L.or_set_field.argtypes = [vp, ip, dp]
        L.or_get_field.argtypes = [vp, ip, dp]
        L.or_init_fields.argtypes = [vp, ip, ctypes.c_double, ip, ctypes.c_double, ip,
                                     ctypes.c_double]
        L.or_start.argtypes = [vp]
        L.or_step.argtypes = [vp, ip]
        L.or_time.argtypes = [vp, dp, dp]
        L.or_ekin_3d.argtypes = [vp, dp]
        L.or_nusselt.argtypes = [vp, dp]
        L.or_robin.argtypes = [vp, ctypes.c_double, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double, dp, dp]
        L.or_ass_solve.argtypes = [vp, ip, dp, dp]
        L.or_fd_ell_solve.argtypes = [vp, ip, dp, dp]
        L.or_radial.argtypes = [vp, dp]
        L.or_cond.argtypes = [vp, dp]
        L.or_threads.argtypes = [ip]
        L.or_time_series.argtypes = [vp, dp, dp]
        L.or_read_checkpoint.argtypes = [vp, ctypes.c_char_p]
        L.or_write_checkpoint.argtypes = [vp, ctypes.c_char_p]
        L.or_dirk_tables.argtypes = [ip, ctypes.POINTER(ip), ctypes.POINTER(ip), dp, dp, dp, dp, dp]
        L.or_costf.argtypes = [vp, dp, ip]
        L.or_get_dr.argtypes = [vp, dp, dp, ip, ip]
        L.or_get_ddr.argtypes = [vp, dp, dp, dp, ip, ip]
        L.or_fft.argtypes = [vp, dp, dp, ip]
        L.or_ifft.argtypes = [vp, dp, dp, ip]
        L.or_grid.argtypes = [vp, dp, dp]
        L.or_stencil.argtypes = [vp, ip, ip, ip, ip, dp]
        L.or_fd_stencils.argtypes = [vp, dp]
        L.or_scheme_weights.argtypes = [ip, ctypes.c_double, dp, ctypes.POINTER(ip), ctypes.POINTER(ip),
                                        dp, dp, dp, dp, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def scheme_weights(scheme, dt, alpha=0.6):
    """Weights of one time scheme for the dt history dt[0..3] (newest
    first): dict with family, nstages, wimp, wimp_lin, wexp (multistep) and
    the dt-scaled Butcher tables bimp, bexp (DIRK)."""
    d = np.ascontiguousarray(np.r_[np.asarray(dt, float), np.ones(4)][:4])
    w = {k: np.zeros(8) for k in ("wimp", "wimp_lin", "wexp")}
    b = {k: np.zeros((8, 8)) for k in ("bimp", "bexp")}
    fam, nst = ctypes.c_int(), ctypes.c_int()
    rc = lib().or_scheme_weights(scheme, alpha, _dp(d), ctypes.byref(fam), ctypes.byref(nst),
                                 _dp(w["wimp"]), _dp(w["wimp_lin"]), _dp(w["wexp"]),
                                 _dp(b["bimp"]), _dp(b["bexp"]))
    if rc:
        raise ValueError(f"unknown scheme {scheme}")
    return dict(family=fam.value, nstages=nst.value, **w, **b)


class Oracle:
    """One CPU run of the reference's step (integration method)."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = ctypes.c_void_p()
        rc = lib().or_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise RuntimeError(f"or_create failed ({rc})")
        self.h = h
        nm, nr, npf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().or_dims(h, ctypes.byref(nm), ctypes.byref(nr), ctypes.byref(npf))
        self.n_m, self.n_r, self.n_phi = nm.value, nr.value, npf.value

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.or_destroy(self.h)
            self.h = None

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc})")

    def set_field(self, name, arr):
        """arr: complex (n_m, n_r)."""
        a = np.asfortranarray(np.asarray(arr, dtype=np.complex128))
        buf = np.ascontiguousarray(a.T).view(np.float64)
        self._check(lib().or_set_field(self.h, FIELDS[name], _dp(buf)), "set_field")

    def get_field(self, name):
        buf = np.zeros((self.n_r, self.n_m), dtype=np.complex128)
        self._check(lib().or_get_field(self.h, FIELDS[name], _dp(buf.view(np.float64))),
                    "get_field")
        return buf.T.copy()

    def init_fields(self, init_t=0, amp_t=0.0, init_xi=0, amp_xi=0.0, init_u=0, amp_u=0.0):
        self._check(lib().or_init_fields(self.h, init_t, amp_t, init_xi, amp_xi, init_u,
                                         amp_u), "init_fields")

    def start(self):
        self._check(lib().or_start(self.h), "start")

    def step(self, n=1):
        self._check(lib().or_step(self.h, n), "step")

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        lib().or_time(self.h, ctypes.byref(t), ctypes.byref(dt))
        return t.value, dt.value

    def nusselt(self):
        out = np.zeros(2)
        lib().or_nusselt(self.h, _dp(out))
        return out

    def robin(self, at, bt, rtop, ab, bb, rbot, f):
        """robin_bc on a complex column f (the boundary values are replaced)."""
        g = np.ascontiguousarray(np.asarray(f, dtype=np.complex128).copy())
        rt = np.array([complex(rtop).real, complex(rtop).imag])
        rb = np.array([complex(rbot).real, complex(rbot).imag])
        self._check(lib().or_robin(self.h, at, bt, _dp(rt), ab, bb, _dp(rb), _dp(g.view(np.float64))), "or_robin")
        return g

    def ass_solve(self, mi, omega):
        """psi on the grid from omega through the assembly matrix of mode mi."""
        w = np.ascontiguousarray(np.asarray(omega, dtype=np.complex128))
        out = np.zeros(self.n_r, dtype=np.complex128)
        self._check(lib().or_ass_solve(self.h, mi, _dp(w.view(np.float64)), _dp(out.view(np.float64))),
                    "or_ass_solve")
        return out

    def fd_ell_solve(self, mi, omega):
        """psi from omega by the finite-difference assembly solve of mode mi
        (psi = 0 at both walls; omega's wall rows are ignored)."""
        w = np.ascontiguousarray(np.asarray(omega, dtype=np.complex128))
        out = np.zeros(self.n_r, dtype=np.complex128)
        self._check(lib().or_fd_ell_solve(self.h, mi, _dp(w.view(np.float64)), _dp(out.view(np.float64))),
                    "or_fd_ell_solve")
        return out

    def write_checkpoint(self, path):
        """Version-7 checkpoint of the run (checkpoints.f90:33-183)."""
        self._check(lib().or_write_checkpoint(self.h, str(path).encode()), "write_checkpoint")

    def read_checkpoint(self, path):
        """Reads a checkpoint into this run (checkpoints.f90:185-707); call
        start() afterwards for the restart branch of get_start_fields."""
        self._check(lib().or_read_checkpoint(self.h, str(path).encode()), "read_checkpoint")

    def time_series(self):
        """(42 time-series values, 5 x n_m spectra) of get_time_series."""
        out, spec = np.zeros(42), np.zeros(5 * self.n_m)
        lib().or_time_series(self.h, _dp(out), _dp(spec))
        return out, spec.reshape(5, self.n_m)

    def cond(self):
        """tcond, dtcond, xicond, dxicond on the grid."""
        out = np.zeros(4 * self.n_r)
        lib().or_cond(self.h, _dp(out))
        return out.reshape(4, self.n_r)

    def radial(self):
        """r, 1/r, 1/r^2, beta, dbeta/dr on the grid."""
        out = np.zeros(5 * self.n_r)
        lib().or_radial(self.h, _dp(out))
        return out.reshape(5, self.n_r)

    def ekin_3d(self):
        out = np.zeros(4)
        lib().or_ekin_3d(self.h, _dp(out))
        return out

    def fd_stencils(self):
        """dr, ddr (N x 3), dddr, ddddr (N x 5), dr_top, dr_bot (3 each)."""
        N = self.n_r
        out = np.zeros(16 * N + 6)
        if lib().or_fd_stencils(self.h, _dp(out)):
            raise ValueError("not a finite-difference run")
        return (out[:3 * N].reshape(N, 3), out[3 * N:6 * N].reshape(N, 3), out[6 * N:11 * N].reshape(N, 5),
                out[11 * N:16 * N].reshape(N, 5), out[16 * N:16 * N + 3], out[16 * N + 3:])

    def grid(self):
        r, x = np.zeros(self.n_r), np.zeros(self.n_r)
        lib().or_grid(self.h, _dp(r), _dp(x))
        return r, x

    # unit operators on complex (n_m, n_r) arrays
    def _to_buf(self, f):
        f = np.asarray(f, dtype=np.complex128)
        return np.ascontiguousarray(f.T), f.shape[0]

    def costf(self, f):
        buf, nm = self._to_buf(f)
        lib().or_costf(self.h, _dp(buf.view(np.float64)), nm)
        return buf.T.copy()

    def get_dr(self, f, l_dct_in=True):
        buf, nm = self._to_buf(f)
        out = np.zeros_like(buf)
        lib().or_get_dr(self.h, _dp(buf.view(np.float64)), _dp(out.view(np.float64)), nm,
                        int(l_dct_in))
        return out.T.copy()

    def get_ddr(self, f, l_dct_in=True):
        buf, nm = self._to_buf(f)
        d1, d2 = np.zeros_like(buf), np.zeros_like(buf)
        lib().or_get_ddr(self.h, _dp(buf.view(np.float64)), _dp(d1.view(np.float64)),
                         _dp(d2.view(np.float64)), nm, int(l_dct_in))
        return d1.T.copy(), d2.T.copy()

    def fft(self, grid):
        """grid: real (n_phi, n_rad) -> spectral complex (n_m, n_rad)."""
        g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64).T)
        nrad = g.shape[0]
        out = np.zeros((nrad, self.n_m), dtype=np.complex128)
        lib().or_fft(self.h, _dp(g), _dp(out.view(np.float64)), nrad)
        return out.T.copy()

    def ifft(self, spec):
        s = np.ascontiguousarray(np.asarray(spec, dtype=np.complex128).T)
        nrad = s.shape[0]
        out = np.zeros((nrad, self.n_phi))
        lib().or_ifft(self.h, _dp(s.view(np.float64)), _dp(out), nrad)
        return out.T.copy()

    def stencil(self, which, m, idx, ku):
        out = np.zeros(2 * ku + 1)
        rc = lib().or_stencil(self.h, which, m, idx, ku, _dp(out))
        self._check(rc, "stencil")
        return out

def threads(n=0):
    """OpenMP threads of the oracle's loops (n > 0 sets them); the count in use."""
    return lib().or_threads(n)


class planted:
    """Context manager that plants mistake k in the oracle (the or_plant
    switch of pz_oracle.c) for the runs created inside it."""

    def __init__(self, k):
        self.k = k

    def __enter__(self):
        ctypes.c_int.in_dll(lib(), "or_plant").value = self.k
        return self

    def __exit__(self, *exc):
        ctypes.c_int.in_dll(lib(), "or_plant").value = 0
        return False


def dirk_tables(scheme):
    """(nstages, l_assembly, wimp_lin0, A_imp, A_exp, b_imp, b_exp) unscaled."""
    ns, la, w0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    bi, be, bai, bae = np.zeros(64), np.zeros(64), np.zeros(8), np.zeros(8)
    rc = lib().or_dirk_tables(scheme, ctypes.byref(ns), ctypes.byref(la), ctypes.byref(w0), _dp(bi), _dp(be),
                              _dp(bai), _dp(bae))
    if rc:
        raise ValueError(f"scheme {scheme} is not a DIRK scheme")
    n = ns.value
    return n, la.value, w0.value, bi.reshape(8, 8)[:n, :n], be.reshape(8, 8)[:n, :n], bai[:n], bae[:n]

"""Wall patterns of the finite-difference T equation (l_fd_bc_pattern = 1).

This is an extension: the reference's finite differences hold the
temperature perturbation at zero on both walls whatever t_top / t_bot say
(prepare_temp_fd, update_temp_fd.f90:85-110; the flux terms of
fill_ghosts_temp are commented out at :142,162), so parity against the
reference is unpinned by definition and the switch defaults to 0, the
reference's behaviour.  With the switch on, a fixed-value wall takes the
pattern as its value and a fixed-flux wall takes it as dT/dr through the
ghost point (T_-1 = T_1 + 2 (r_0 - r_1) q at the outer wall), whose
inhomogeneous part moves to the wall row's right-hand side.

The oracle's restatement is pinned here by a manufactured steady state:
without flow (Ra -> 0) each mode relaxes to the conduction profile
T_m = A r^m + B r^-m of the annulus, which the walls fix in closed form.
The finite-difference solution converges to it at second order, its wall
gradient equals the flux pattern to rounding, and planted mistakes in the
wall row (7) and in the ghost point (8) fail the checks.  The GPU library
is compared with the oracle in tests/test_gpu_parity.py.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, planted
from pizza_amd import workloads as W
from pizza_amd.config import default_config

Q_TOP = 0.3 - 0.1j   # m = 2: dT/dr at the outer wall (ktopt = 2)
P_BOT = 0.2 + 0.05j  # m = 3: T at the inner wall (kbott = 1)


def _cfg(n_r, scheme="SBDF2", dt=1.0e-2, pattern=1):
    c = default_config(n_r_max=n_r, n_cheb_max=n_r, m_max=4, minc=1, time_scheme=scheme, radial_scheme="FD",
                       dtMax=dt, ra=1.0e-30, ek=1.0e-3, pr=1.0, radratio=0.35, ktopt=2, kbott=1,
                       t_top=[2, Q_TOP.real, Q_TOP.imag], t_bot=[3, P_BOT.real, P_BOT.imag])
    c.l_fd_bc_pattern = pattern
    return c


def _conduction(cfg):
    """Closed-form steady modes m = 2 (flux q at r_o, T = 0 at r_i) and
    m = 3 (T = p at r_i, zero flux at r_o) on the run's radii."""
    r, r_i = W.radial_grid(cfg)
    r = np.asarray(r)
    r_o = r[0]
    a2 = np.linalg.solve(np.array([[r_i ** 2, r_i ** -2], [2 * r_o, -2 * r_o ** -3]], complex), [0.0, Q_TOP])
    a3 = np.linalg.solve(np.array([[r_i ** 3, r_i ** -3], [3 * r_o ** 2, -3 * r_o ** -4]], complex), [P_BOT, 0.0])
    return {2: a2[0] * r ** 2 + a2[1] * r ** -2, 3: a3[0] * r ** 3 + a3[1] * r ** -3}


def _steady(cfg, steps):
    o = Oracle(cfg)
    o.start()
    o.step(steps)
    T, dT = o.get_field("temp"), o.get_field("dtemp")
    ex = _conduction(cfg)
    err = {m: np.abs(T[m] - ex[m]).max() / np.abs(ex[m]).max() for m in (2, 3)}
    return err, abs(dT[2][0] - Q_TOP)


def test_fd_wall_patterns_converge_at_second_order():
    """SBDF2, dt 1e-2, 400 steps (t = 4, steady to rounding): the errors of
    both modes fall by 4 per halving of the grid spacing (measured at
    n_r = 33, 65, 129: m = 2 1.56e-4, 3.74e-5, 9.25e-6; m = 3 1.81e-3,
    4.39e-4, 1.08e-4)."""
    errs = [_steady(_cfg(n), 400)[0] for n in (33, 65, 129)]
    for m in (2, 3):
        e = [x[m] for x in errs]
        assert e[0] < 3e-3, (m, e)
        for a, b in zip(e[:-1], e[1:]):
            assert 3.6 < a / b < 4.6, (m, e)


def test_fd_flux_pattern_sets_the_wall_gradient():
    """dT/dr at the outer wall (the ghost-point stencil of the step) equals
    the flux pattern to rounding."""
    _, g = _steady(_cfg(33), 100)
    assert g < 1e-12, g


def test_fd_wall_patterns_with_implicit_terms():
    """CNAB2 (implicit terms evaluated through the ghost points) reaches
    the same steady state: at dt 2e-3 its error equals the SBDF2 one."""
    err_cn, g = _steady(_cfg(33, "CNAB2", 2.0e-3), 2000)
    err_sb, _ = _steady(_cfg(33), 400)
    for m in (2, 3):
        assert abs(err_cn[m] - err_sb[m]) < 1e-3 * err_sb[m], (m, err_cn, err_sb)
    assert g < 1e-12


@pytest.mark.parametrize("plant", [7, 8])
def test_fd_wall_pattern_plants_fail(plant):
    """Dropping the flux part of the wall row (7) leaves the m = 2 mode at
    zero; flipping its sign in the ghost point (8) breaks the wall gradient
    and, through the implicit terms, the CNAB2 steady state."""
    with planted(plant):
        err, g = _steady(_cfg(33), 100)
        if plant == 7:
            assert err[2] > 0.5
        else:
            assert g > 0.1
            err_cn, _ = _steady(_cfg(33, "CNAB2", 2.0e-3), 500)
            assert err_cn[2] > 0.1


def test_fd_patterns_off_is_the_reference():
    """Without the switch the walls stay homogeneous, as in the reference:
    the same run leaves both modes at zero."""
    o = Oracle(_cfg(33, pattern=0))
    o.start()
    o.step(50)
    T = o.get_field("temp")
    assert np.abs(T[2]).max() == 0.0 and np.abs(T[3]).max() == 0.0

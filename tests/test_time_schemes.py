"""The time-scheme weights of the oracle, pinned by the mathematics: each
scheme reaches its design order on the IMEX test equation
y' = lam y (implicit) + mu y (explicit), the multistep ones on a variable
step sequence, and every DIRK stage has the same abscissa in its implicit
and explicit tables (except the DP schemes, which skip explicit
evaluations).  A wrong coefficient breaks the order of its scheme.
The step is driven exactly as the reference drives it: set_imex_rhs
(multistep_schemes.f90:366-413; dirk_schemes.f90:773-814), a solve with the
weight wimp_lin(1), rotate_imex (multistep_schemes.f90:462-498).
"""
import numpy as np
import pytest

from oracle.oracle import scheme_weights
from pizza_amd import config as C

LAM, MU = -1.3, 0.7
PATTERN = [1.0, 1.25, 0.8, 1.1, 0.9]

MULTISTEP = [(C.CNAB2, 2), (C.CNLF, 2), (C.MODCNAB, 2), (C.SBDF2, 2), (C.SBDF3, 3),
             (C.TVB33, 3), (C.SBDF4, 4)]
DIRK = [(C.ARS222, 2), (C.ARS443, 3), (C.BPR353, 3), (C.DP1_242, 2), (C.DP2_242, 2),
        (C.LZ232, 2), (C.CK232, 2), (C.LZ453, 3), (C.PC2, 2)]


def _exact(t):
    return np.exp((LAM + MU) * t)


def _multistep(scheme, h, T=1.0):
    """Integrates to T with step sizes h * PATTERN (cyclic), exact history."""
    hist = 4
    # times t_0 > t_-1 > ... before the start, with the pattern run backwards
    dts = [h * PATTERN[(-k - 1) % len(PATTERN)] for k in range(hist)]
    times = [0.0]
    for d in dts:
        times.append(times[-1] - d)
    y = [_exact(t) for t in times]          # y[0] = y_n, y[1] = y_{n-1}, ...
    dt_hist = list(dts)                     # dt_hist[0] = t_n - t_{n-1}
    t, n = 0.0, 0
    while t < T - 1e-12:
        d = min(h * PATTERN[n % len(PATTERN)], T - t)
        dt_hist = [d] + dt_hist[:3]
        w = scheme_weights(scheme, dt_hist, alpha=0.5)
        rhs = sum(w["wimp"][o] * y[o] for o in range(4))
        rhs += sum(w["wimp_lin"][o + 1] * LAM * y[o] for o in range(4))
        rhs += sum(w["wexp"][o] * MU * y[o] for o in range(4))
        ynew = rhs / (1.0 - w["wimp_lin"][0] * LAM)
        y = [ynew] + y[:3]
        t += d
        n += 1
    return abs(y[0] - _exact(T))


def _dirk(scheme, h, T=1.0):
    y, t = 1.0, 0.0
    while t < T - 1e-12:
        d = min(h, T - t)
        w = scheme_weights(scheme, [d, d, d, d])
        ns = w["nstages"]
        bi, be = w["bimp"], w["bexp"]
        U = [y]
        for st in range(1, ns + 1):
            rhs = y + sum(be[st, s] * MU * U[s] + bi[st, s] * LAM * U[s] for s in range(st))
            U.append(rhs / (1.0 - w["wimp_lin"][0] * LAM))
        y = U[-1]
        t += d
    return abs(y - _exact(T))


@pytest.mark.parametrize("scheme,order", MULTISTEP)
def test_multistep_order_on_variable_steps(scheme, order):
    e1, e2 = _multistep(scheme, 0.02), _multistep(scheme, 0.01)
    assert np.log2(e1 / e2) > order - 0.25, (scheme, e1, e2)


@pytest.mark.parametrize("scheme,order", DIRK)
def test_dirk_order(scheme, order):
    e1, e2 = _dirk(scheme, 0.02), _dirk(scheme, 0.01)
    assert np.log2(e1 / e2) > order - 0.25, (scheme, e1, e2)


@pytest.mark.parametrize("scheme,order", [d for d in DIRK if d[0] not in (C.DP1_242, C.DP2_242)])
def test_dirk_tables_share_abscissae(scheme, order):
    """(The DP schemes skip explicit evaluations and have no common c.)"""
    w = scheme_weights(scheme, [1.0, 1.0, 1.0, 1.0])
    ns = w["nstages"]
    ci, ce = w["bimp"].sum(axis=1)[: ns + 1], w["bexp"].sum(axis=1)[: ns + 1]
    assert np.allclose(ci[1:], ce[1:], atol=1e-14), (ci, ce)
    assert abs(ci[ns] - 1.0) < 1e-14                    # the last row is the solution
    assert np.allclose(np.diag(w["bimp"])[1: ns + 1], w["wimp_lin"][0], atol=1e-15)


def test_cnab2_bridge_is_second_order_start():
    """The start-up steps take the CNAB2 weights: with alpha = 1/2 they are
    the weights of a second-order method on constant steps."""
    w = scheme_weights(C.CNAB2, [0.1, 0.1, 0.1, 0.1], alpha=0.5)
    assert w["wimp"][0] == 1.0
    assert np.allclose(w["wimp_lin"][:2], [0.05, 0.05])
    assert np.allclose(w["wexp"][:2], [0.15, -0.05])


def test_scheme_names_follow_the_reference_matching():
    assert C.scheme_id("sbdf22") == C.SBDF2
    assert C.scheme_id("BDF3AB3") == C.SBDF3
    assert C.scheme_id("MODCNAB") == C.MODCNAB
    assert C.scheme_id("dp1-242") == C.DP1_242
    assert C.scheme_id("ARS343") == C.ARS343
    assert C.scheme_id("kc785") == C.KC785
    # initialize() tests 'ARS343' before 'MARS343' (dirk_schemes.f90:183-192)
    # and 'ARS343' is a substring of 'MARS343': the name selects ARS343
    assert C.scheme_id("MARS343") == C.ARS343


# DIRK schemes with an assembly stage and the order their tables reach as
# written in the reference (dirk_schemes.f90:299-711): (implicit, explicit).
# SSP433 writes alpha, beta, eta as single-precision literals, so its
# third-order conditions hold only to ~1e-8; MARS343's explicit assembly
# weights (= its implicit ones) meet only the first-order condition.
ASSEMBLY_ORDERS = {"SSP222": (2, 2), "LSDIRK222": (2, 2), "ARS232": (2, 2), "ARS233": (3, 3),
                   "SSP322": (2, 2), "SSP332": (2, 3), "SSP433": (2, 3), "ARS343": (3, 3),
                   "MARS343": (3, 1), "KC343": (3, 3), "KC564": (4, 4), "KC785": (4, 4), "DBM453": (3, 3)}


def _order(A, b, tol=1e-9):
    c = A.sum(1)
    conds = [[b.sum() - 1], [b @ c - 0.5], [b @ c**2 - 1 / 3, b @ A @ c - 1 / 6],
             [b @ c**3 - 0.25, b @ (c * (A @ c)) - 1 / 8, b @ A @ c**2 - 1 / 12, b @ A @ A @ c - 1 / 24]]
    p = 0
    for cs in conds:
        if max(abs(x) for x in cs) > tol:
            break
        p += 1
    return p


@pytest.mark.parametrize("name", sorted(ASSEMBLY_ORDERS))
def test_assembly_tables_order_conditions(name):
    """Tables of the assembly-stage DIRK schemes: SDIRK diagonal equal to the
    implicit weight, the order conditions each table meets, and for SSP433
    the third-order conditions to single precision."""
    from oracle.oracle import dirk_tables
    n, l_ass, w0, AI, AE, bI, bE = dirk_tables(getattr(C, name))
    assert l_ass == 1
    assert np.allclose(np.diag(AI)[1:], w0, rtol=0, atol=1e-15)
    assert np.all(np.triu(AE) == 0)          # explicit table strictly lower
    assert np.all(np.triu(AI, 1) == 0)       # implicit table lower
    oi, oe = ASSEMBLY_ORDERS[name]
    assert _order(AI, bI) == oi and _order(AE, bE) == oe, (_order(AI, bI), _order(AE, bE))
    if name == "SSP433":
        assert _order(AI, bI, tol=1e-7) == 3


def _coll_oracle(n_r=17, m_max=8, **kw):
    from oracle.oracle import Oracle
    from pizza_amd.config import default_config
    cfg = default_config(n_r_max=n_r, n_cheb_max=n_r, m_max=m_max, time_scheme="ARS343", cheb_method="colloc",
                         ra=1e5, ek=1e-3, pr=1.0, radratio=0.35, **kw)
    return Oracle(cfg)


def test_robin_bc_recovers_polynomial_boundary_values():
    """robin_bc (chebyshev.f90:385-445): for a polynomial of degree n_r - 1 the
    boundary-derivative rows are exact, so two Robin conditions built from
    its analytic derivative give back its boundary values."""
    from numpy.polynomial import chebyshev as C
    o = _coll_oracle()
    r = o.radial()[0]
    N = len(r)
    x = np.cos(np.pi * np.arange(N) / (N - 1))
    drx = 2.0 / (r[0] - r[-1])
    rng = np.random.default_rng(3)
    for trial in range(4):
        a = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        f = C.chebval(x, a)
        fr = drx * C.chebval(x, C.chebder(a))
        at, bt, ab, bb = rng.uniform(0.2, 2.0, 4) * np.array([1, 1, 1, 1])
        if trial == 1:
            at, ab = 0.0, 0.0           # Dirichlet on both walls
        if trial == 2:
            bt, bb = 0.0, 0.0           # Neumann on both walls
        rt = at * fr[0] + bt * f[0]
        rb = ab * fr[-1] + bb * f[-1]
        g = f.copy()
        g[0] = g[-1] = 7.0 + 3.0j
        out = o.robin(at, bt, rt, ab, bb, rb, g)
        assert np.allclose(out[1:-1], f[1:-1], rtol=0, atol=0)
        scale = np.abs(f).max()
        assert abs(out[0] - f[0]) < 1e-11 * scale and abs(out[-1] - f[-1]) < 1e-11 * scale, (trial, out[0], f[0])


def test_assembly_matrix_recovers_psi():
    """get_assMat (update_psi_coll_smat.f90:839-917): psi = (1 - x^2)^2 q(x)
    meets psi = dpsi/dr = 0 at both walls; its omega = -(psi'' + (1/r +
    beta) psi' + (beta/r + beta' - m^2/r^2) psi) on the interior rows gives
    psi back through the assembly matrix."""
    from numpy.polynomial import chebyshev as C
    o = _coll_oracle(n_r=25, m_max=8)
    r, or1, or2, beta, dbeta = o.radial()
    N = len(r)
    x = np.cos(np.pi * np.arange(N) / (N - 1))
    drx = 2.0 / (r[0] - r[-1])
    rng = np.random.default_rng(4)
    w4 = C.chebpow(C.chebsub([1.0], C.chebmulx([0.0, 1.0])), 2)   # (1 - x^2)^2
    for mi in (1, 3, 8):
        q = rng.standard_normal(N - 6) + 1j * rng.standard_normal(N - 6)
        cpsi = C.chebmul(w4, q)
        psi = C.chebval(x, cpsi)
        p1 = drx * C.chebval(x, C.chebder(cpsi))
        p2 = drx**2 * C.chebval(x, C.chebder(cpsi, 2))
        om = -(p2 + (or1 + beta) * p1 + (or1 * beta + dbeta - mi**2 * or2) * psi)
        om[[0, 1, N - 2, N - 1]] = 0.0
        got = o.ass_solve(mi, om)
        assert np.abs(got - psi).max() < 1e-9 * np.abs(psi).max(), (mi, np.abs(got - psi).max())


def test_fd_elliptic_solve_recovers_quadratic_psi():
    """get_elliptic_mat_Rdist (update_psi_fd.f90:838-888), the psi solve of
    the finite-difference assembly stage: the three-point stencils are exact
    for quadratics on any grid, so psi = c (r - r_o)(r - r_i), zero at both
    walls, comes back from omega = -(psi'' + (beta + 1/r) psi' + (beta' +
    beta/r - m^2/r^2) psi) on the stretched grid to rounding."""
    from oracle.oracle import Oracle
    from pizza_amd.config import default_config
    cfg = default_config(n_r_max=33, n_cheb_max=33, m_max=8, time_scheme="ARS343", radial_scheme="FD",
                         fd_ratio=0.1, ra=1e5, ek=1e-3, pr=1.0, radratio=0.35, buo_term="explicit")
    o = Oracle(cfg)
    r, or1, or2, beta, dbeta = o.radial()
    assert np.ptp(np.diff(r)) > 1e-3 * abs(r[1] - r[0])          # the grid is stretched
    c = 0.7 - 1.3j
    psi = c * (r - r[0]) * (r - r[-1])
    p1 = c * (2 * r - r[0] - r[-1])
    for mi in (1, 3, 8):
        om = -(2 * c + (beta + or1) * p1 + (dbeta + beta * or1 - mi**2 * or2) * psi)
        om[[0, -1]] = 5.0 + 2.0j                                  # ignored: psi = 0 rows
        got = o.fd_ell_solve(mi, om)
        assert got[0] == 0 and got[-1] == 0
        assert np.abs(got - psi).max() < 1e-12 * np.abs(psi).max(), (mi, np.abs(got - psi).max())

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
    with pytest.raises(ValueError):
        C.scheme_id("ARS343")
    with pytest.raises(ValueError):
        C.scheme_id("MARS343")

"""CPU tests that pin the parts of the oracle's production physics (the
integration method with Ekman pumping and the compositional equation) that
no golden series of the reference covers.

1. Cross-method test of the integration method's Ekman pumping.  The
   integration method treats pumping explicitly in the finish of the psi
   equation (update_psi_integ_smat.f90:895-913); the collocation method
   treats it implicitly (update_psi_coll_smat.f90:475-488, 756-791), and the
   collocation method with pumping is pinned by the reference's
   RotE1e3EkPump golden series.  Both methods discretise the same
   equations, so after the start-up transient has decayed their fields
   agree to the level the two truncations allow, and that level shrinks as
   n_r grows.  A planted sign flip, a dropped factor r or a dropped u_s part
   of the integration pumping term gives errors 10 to 500 times larger.
2. Exchange of temperature and composition.  With the same coefficients
   (ra = raxi, pr = sc), no conducting profiles and vertical advection of T
   (which removes the beta term the T equation alone carries,
   radial_loop.f90:365-372), a run driven by xi alone equals the run driven
   by T alone bit for bit: the xi equation (ChemFac, XidiffFac, the xi
   boundary rows and solves, the compositional buoyancy) is the T equation
   with its own names.  It runs with explicit buoyancy, with implicit
   buoyancy on the ARS schemes, whose implicit tableau has a zero first
   column, and on collocation grids.  With implicit buoyancy and a scheme
   that uses the first implicit stage, the integration start-up evaluates
   the buoyancy with psi in the xi slot (init_fields.f90:250-251), so there
   the two runs differ, as in the reference.  A planted sign flip of the
   compositional buoyancy or a doubled xi diffusivity breaks the equality.
3. The xi conducting profile is the T profile under the same boundary
   settings (one routine, radial.f90:284-397, for both).
"""
import numpy as np
import pytest

from helpers import rel
from oracle.oracle import Oracle, planted
from pizza_amd.config import default_config

FIELDS = ("us", "up", "temp", "xi", "psi")


def _pumping_run(method, n_r, plant=0, K=1000):
    """E = 1e-3, Ra = 2.77e4, Ra_xi = 2e4, Sc = 3 (the RotE1e3EkPump physics
    plus composition), rigid walls, Ekman pumping, CNAB2 dt = 1e-4; single
    modes m = 4 in T and m = 3 in xi (init_fields.f90:260-360).  Returns the
    fields after K steps, psi as h^2 psi on the integration method (its
    streamfunction is reduced by h^2 = r_o^2 - r^2) and omega away from the
    walls (distance > 0.05), where the methods' wall treatments differ."""
    cfg = default_config(n_r_max=n_r, n_cheb_max=n_r, m_max=32, minc=1, time_scheme="CNAB2",
                         cheb_method=method, matrix_solve="direct", dtMax=1.0e-4,
                         ra=2.77e4, ek=1.0e-3, pr=1.0, raxi=2.0e4, sc=3.0, radratio=0.35,
                         ktopt=1, kbott=1, ktopxi=1, kbotxi=1, ktopv=2, kbotv=2, l_ek_pump=True, g1=1.0)
    with planted(plant):
        o = Oracle(cfg)
        o.init_fields(init_t=4, amp_t=2.5e-4, init_xi=3, amp_xi=2.5e-4)
        o.start()
        o.step(K)
        r, _ = o.grid()
        F = {f: o.get_field(f) for f in FIELDS + ("om",)}
    if method == "integ":
        F["psi"] = F["psi"] * (r[0] ** 2 - r ** 2)[None, :]
    inner = (r[0] - r > 0.05) & (r - r[-1] > 0.05)
    F["om"] = F["om"][:, inner]
    return F


def _errs(a, b):
    return {f: rel(a[f], b[f]) for f in b}


def test_integration_ekman_pumping_against_collocation():
    """Measured after 1000 steps: n_r 33: u_s 9.2e-3, u_phi 1.4e-2, T 8.3e-3,
    xi 8.0e-3, h^2 psi 8.5e-3; n_r 65: 1.7e-3, 4.9e-3, 1.6e-3, 1.5e-3,
    1.6e-3.  Planted at n_r 33: sign flip 5.0 (u_s), no factor r 0.14, no
    u_s part 0.42."""
    errs = {}
    for n_r in (33, 65):
        c, i = _pumping_run("colloc", n_r), _pumping_run("integ", n_r)
        errs[n_r] = _errs(i, c)
        print(n_r, {k: f"{v:.2e}" for k, v in errs[n_r].items()}, flush=True)
        tol = 2.5e-2 if n_r == 33 else 8e-3
        for f in FIELDS:
            assert errs[n_r][f] < tol, (n_r, f, errs[n_r][f])
        assert errs[n_r]["om"] < 6e-2, (n_r, errs[n_r]["om"])
        if n_r == 33:
            for plant in (1, 2, 3):
                e = _errs(_pumping_run("integ", n_r, plant), c)
                print("planted", plant, {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
                assert min(e[f] for f in FIELDS) > 0.1, (plant, e)
        else:
            e = _errs(_pumping_run("integ", n_r, 2), c)
            assert min(e[f] for f in FIELDS) > 0.1, e
    # both methods converge to the same solution
    for f in ("us", "temp", "xi", "psi"):
        assert errs[65][f] < errs[33][f] / 3.0, (f, errs[33][f], errs[65][f])


def _exchange_run(which, method, scheme, buo_imp, plant=0, K=20):
    X, S = 3.0e7, 3.0
    kw = dict(ra=X, pr=S, raxi=0.0, sc=1.0) if which == "T" else dict(ra=0.0, pr=1.0, raxi=X, sc=S)
    cfg = default_config(n_r_max=33, n_cheb_max=33, m_max=32, minc=1, time_scheme=scheme,
                         cheb_method=method, matrix_solve="direct", dtMax=5.0e-6, ek=1.0e-5,
                         radratio=0.35, ktopt=1, kbott=1, ktopxi=1, kbotxi=1, ktopv=2, kbotv=2,
                         l_ek_pump=True, g1=1.0, buo_term="implicit" if buo_imp else "explicit",
                         l_temp_advz=True, tcond_fac=0.0, xicond_fac=0.0, **kw)
    from pizza_amd import workloads as W
    with planted(plant):
        o = Oracle(cfg)
        r, r_icb = W.radial_grid(cfg)
        o.set_field("temp" if which == "T" else "xi", W.seeded_field(o.n_m, r, r_icb, 1.0e-4))
        o.start()
        o.step(K)
        out = {f: o.get_field(f) for f in ("us", "up", "om", "psi")}
        out["s"] = o.get_field("temp" if which == "T" else "xi")
    return out


@pytest.mark.parametrize("method,scheme,buo_imp", [("integ", "BPR353", 0), ("integ", "ARS222", 1),
                                                    ("integ", "ARS443", 1), ("colloc", "BPR353", 1),
                                                    ("colloc", "CNAB2", 1)])
def test_composition_equals_temperature_under_exchange(method, scheme, buo_imp):
    a = _exchange_run("T", method, scheme, buo_imp)
    b = _exchange_run("xi", method, scheme, buo_imp)
    for f in a:
        assert np.array_equal(a[f], b[f]), f


def test_exchange_start_up_quirk_and_planted_mistakes():
    """The integration start-up with implicit buoyancy and BPR353 passes psi
    in the xi slot (init_fields.f90:250-251): the xi run then differs; and
    planted mistakes of the compositional physics break the equality."""
    a = _exchange_run("T", "integ", "BPR353", 1)
    b = _exchange_run("xi", "integ", "BPR353", 1)
    assert rel(b["us"], a["us"]) > 1e-3
    a = _exchange_run("T", "integ", "BPR353", 0)
    for plant in (4, 5):
        b = _exchange_run("xi", "integ", "BPR353", 0, plant)
        assert rel(b["us"], a["us"]) > 1e-3, plant


@pytest.mark.parametrize("k", [(1, 1), (1, 2), (2, 1), (2, 2)])
def test_composition_conducting_profile(k):
    """Same wall types and the same m = 0 wall values or fluxes for T and
    xi: the same conducting profile."""
    ktop, kbot = k
    cfg = default_config(n_r_max=33, n_cheb_max=33, m_max=8, ra=1e5, raxi=3e5, sc=3.0, ktopt=ktop, kbott=kbot,
                         ktopxi=ktop, kbotxi=kbot, cheb_method="integ")
    for a in (cfg.t_top, cfg.xi_top):
        a[0], a[1] = 0, 0.3
    for a in (cfg.t_bot, cfg.xi_bot):
        a[0], a[1] = 0, 0.5
    prof = Oracle(cfg).cond()
    assert np.abs(prof[0]).max() > 0 or np.abs(prof[1]).max() > 0
    assert np.array_equal(prof[0], prof[2]) and np.array_equal(prof[1], prof[3])


def _dmat_run(matrix_solve, dt, K, plant=0):
    from pizza_amd import workloads as W
    cfg = default_config(n_r_max=33, n_cheb_max=33, m_max=16, minc=1, time_scheme="CNAB2", cheb_method="colloc",
                         matrix_solve=matrix_solve, dtMax=dt, ra=1e5, ek=1e-3, pr=1.0, raxi=3e4, sc=3.0,
                         radratio=0.35, l_ek_pump=True, corio_term="explicit")
    with planted(plant):
        o = Oracle(cfg)
        W.seed_run(o, cfg, amp_t=1e-3, amp_xi=1e-3)
        o.start()
        o.step(K)
        return {f: o.get_field(f) for f in FIELDS}


def test_influence_matrix_converges_to_direct_solve():
    """The collocation influence-matrix method (update_psi_coll_dmat.f90)
    treats Coriolis and Ekman pumping explicitly where the direct solve
    (update_psi_coll_smat.f90) treats pumping implicitly: both CNAB2 runs
    are second order in time, so their difference at t = 2e-3 shrinks four
    times per halving of dt.  Measured: dt 1e-5: u_s 1.5e-6, u_phi 4.0e-6,
    psi 1.4e-6, T 2.9e-7, xi 1.9e-7; dt 5e-6: 3.8e-7, 9.9e-7, 3.6e-7,
    7.1e-8, 4.8e-8.  A planted sign flip of the explicit pumping term gives
    a difference that does not converge."""
    d1 = _errs(_dmat_run("influence", 1e-5, 200), _dmat_run("direct", 1e-5, 200))
    d2 = _errs(_dmat_run("influence", 5e-6, 400), _dmat_run("direct", 5e-6, 400))
    for f in FIELDS:
        assert d1[f] < 1e-5 and d2[f] < d1[f] / 3.0, (f, d1[f], d2[f])
    p2 = _errs(_dmat_run("influence", 5e-6, 400, plant=6), _dmat_run("direct", 5e-6, 400))
    assert max(p2.values()) > 10 * max(d2.values()), (p2, d2)

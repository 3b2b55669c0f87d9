"""The benchmark configurations at their own sizes, on the GPU, through the
C ABI: the production integration case (n_r 97, n_phi 384, BPR353, T + xi,
Ekman pumping) over 100 steps, the large finite-difference case (n_r 257,
n_phi 2046, SBDF3), the adjoint case (collocation n_r 65, n_phi 255, 400
CNAB2 steps, snapshots every 20) and the full 64-run uneven ensemble.

Field errors are max|gpu - oracle| / max|oracle| per field; every test
prints the measured values so that a run's log records them.
"""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def _fields(cfg):
    f = ["us", "up", "om", "psi"]
    if cfg.ra != 0:
        f.append("temp")
    if cfg.raxi != 0:
        f.append("xi")
    return f


def _report(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def test_production_integration_100_steps():
    """n_r 97, m_max 128 (n_phi 384), BPR353, T + xi, Ekman pumping, dt
    5e-6: GPU against the oracle after 1 step (1e-11 for T and xi, whose
    conditioning floor is 2e-14; 10 x the measured floor for the velocity
    fields, whose second radial derivatives amplify rounding: omega 5.7e-9,
    u_phi 1e-9) and after 100 steps (1e-8 for every field)."""
    from oracle.oracle import Oracle
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    cfg = W.c2_production()
    print("case: production integration n_r 97 m_max 128, 100 steps", flush=True)
    g, o = Run(cfg), Oracle(cfg)
    for x in (g, o):
        W.seed_run(x, cfg)
        x.start()
    g.step(1)
    o.step(1)
    e1 = {f: rel(g.get_field(f), o.get_field(f)) for f in _fields(cfg)}
    _report("after 1 step:", e1)
    tol1 = {"temp": 1e-11, "xi": 1e-11, "psi": 2e-10, "us": 1e-9, "up": 1e-8, "om": 6e-8}
    for f, e in e1.items():
        assert e < tol1[f], (f, e)
    g.step(99)
    o.step(99)
    assert abs(g.time()[0] - o.time()[0]) < 1e-15
    e100 = {f: rel(g.get_field(f), o.get_field(f)) for f in _fields(cfg) + ["temp_hat", "xi_hat", "psi_hat"]}
    _report("after 100 steps:", e100)
    for f, e in e100.items():
        assert e < 1e-8, (f, e)
    np.testing.assert_allclose(g.nusselt(), o.nusselt(), rtol=1e-10)
    np.testing.assert_allclose(g.ekin_3d(), o.ekin_3d(), rtol=1e-8)


def test_large_finite_difference_case():
    """n_r 257 on the stretched grid, m_max 682 (n_phi 2046), SBDF3 with its
    CNAB2 start-up steps, Ekman 1e-6, Ra 1e9, Pr 0.1, dt 1e-7: GPU against
    the oracle after 1 and 6 steps."""
    from oracle.oracle import Oracle
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    cfg = W.c5_fd()
    print("case: finite differences n_r 257 m_max 682, 6 steps", flush=True)
    g, o = Run(cfg), Oracle(cfg)
    for x in (g, o):
        W.seed_run(x, cfg, amp_t=1e-3)
        x.start()
    for n, tol in ((1, 1e-11), (5, 1e-10)):
        g.step(n)
        o.step(n)
        assert abs(g.time()[0] - o.time()[0]) < 1e-18
        e = {f: rel(g.get_field(f), o.get_field(f)) for f in _fields(cfg)}
        _report(f"after {n} more steps:", e)
        for f, v in e.items():
            assert v < tol, (f, n, v)
    np.testing.assert_allclose(g.ekin_3d(), o.ekin_3d(), rtol=1e-9, atol=1e-30)
    np.testing.assert_allclose(g.nusselt(), o.nusselt(), rtol=1e-10)


def _c4_start(run, cfg, T0):
    run.set_field("temp", T0)
    run.start()


def test_adjoint_case_forward_100_steps():
    """Collocation n_r 65, m_max 85 (n_phi 255), CNAB2, dt 1e-5: the forward
    step against the oracle after 1 and 100 steps."""
    from oracle.oracle import Oracle
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    cfg = W.c4_adjoint()
    print("case: collocation n_r 65 m_max 85, 100 steps", flush=True)
    g, o = Run(cfg), Oracle(cfg)
    for x in (g, o):
        W.seed_member(x, cfg, 0)
        x.start()
    for n, tol in ((1, 1e-11), (99, 1e-8)):
        g.step(n)
        o.step(n)
        e = {f: rel(g.get_field(f), o.get_field(f)) for f in _fields(cfg)}
        _report(f"after {n} more steps:", e)
        for f, v in e.items():
            assert v < tol, (f, n, v)


def _c4_setup():
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    cfg = W.c4_adjoint()
    run = Run(cfg)
    r, r_icb = W.radial_grid(cfg)
    T0 = W.seeded_field(run.n_m, r, r_icb, 1.0e-3, W.SEED, 0)
    return cfg, run, T0, r, r_icb


def _dot(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def test_adjoint_case_two_step_dot_product():
    """At the adjoint case's size: J = <W, T(2)> is a quadratic polynomial
    of the initial temperature, so central differences of any step equal
    <grad, d> up to rounding (1e-9 relative)."""
    from pizza_amd import workloads as W
    cfg, run, T0, r, r_icb = _c4_setup()
    print("case: adjoint dot product n_r 65 m_max 85", flush=True)
    rng = np.random.default_rng(11)
    Wt = rng.standard_normal((run.n_m, run.n_r)) + 1j * rng.standard_normal((run.n_m, run.n_r))

    def J(T):
        run.set_field("us", np.zeros_like(T))
        run.set_field("up", np.zeros_like(T))
        _c4_start(run, cfg, T)
        return run.adjoint_forward("temp", 2, 20, Wt)

    J(T0)
    g, _ = run.adjoint_backward()
    for k in range(3):
        d = W.seeded_field(run.n_m, r, r_icb, 1.0e-3, W.SEED + 50 + k, 0)
        fd = (J(T0 + d) - J(T0 - d)) / 2.0
        ad = _dot(g, d)
        print(f"direction {k}: central difference {fd:.15e} adjoint {ad:.15e}", flush=True)
        assert abs(fd - ad) <= 1e-9 * max(abs(fd), abs(ad)), (k, fd, ad)


def test_adjoint_case_nusselt_gradient_400_steps():
    """The adjoint case itself: the gradient of the outer-wall Nusselt number
    averaged over 400 CNAB2 steps (snapshots every 20) with respect to the
    initial temperature, against central differences (step 1e-6) of the GPU
    forward on 16 seeded directions and of the oracle's forward on 2 of
    them, 1e-6 relative."""
    from oracle.oracle import Oracle
    from pizza_amd import workloads as W
    cfg, run, T0, r, r_icb = _c4_setup()
    K = 400
    print("case: adjoint Nusselt gradient n_r 65 m_max 85, 400 steps", flush=True)

    z = np.zeros_like(T0)

    def J(T):
        # every window starts from rest: the handle ends a window at its last state
        run.set_field("us", z)
        run.set_field("up", z)
        _c4_start(run, cfg, T)
        return run.adjoint_forward("nutop", K, 20)

    J0 = J(T0)
    g, _ = run.adjoint_backward()

    def j_oracle(T):
        o = Oracle(cfg)
        o.set_field("temp", T)
        o.start()
        nu = 0.0
        for _ in range(K):
            o.step(1)
            nu += o.nusselt()[0]
        return nu / K

    # directions of unit amplitude (the seeded profile with unit Gaussian
    # coefficients), step 1e-6: the change of J is then about 1e-7, far above
    # the rounding of J (1e-16), and the perturbation (1e-6 against the
    # 1e-3 initial field) stays in the linear range over the 400 steps
    h = 1.0e-6
    for k in range(16):
        d = W.seeded_field(run.n_m, r, r_icb, 1.0, W.SEED + 100 + k, 0)
        fd = (J(T0 + h * d) - J(T0 - h * d)) / (2 * h)
        ad = _dot(g, d)
        print(f"direction {k}: GPU central difference {fd:.12e} adjoint {ad:.12e} "
              f"rel {abs(fd - ad) / abs(ad):.2e}", flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (k, fd, ad, J0)
        if k < 2:
            fo = (j_oracle(T0 + h * d) - j_oracle(T0 - h * d)) / (2 * h)
            print(f"direction {k}: oracle central difference {fo:.12e} rel {abs(fo - ad) / abs(ad):.2e}", flush=True)
            assert abs(fo - ad) <= 1e-6 * abs(ad), (k, fo, ad)


def test_full_uneven_ensemble_equals_single_runs():
    """All 64 members of the uneven ensemble (n_r in {33, 65, 97, 129},
    n_phi in {96, 192, 384, 510}, Ra log-uniform in [1e5, 1e8], Ekman 1e-4
    or 1e-5) under the autotuned m-loop shape, after 12 steps, against the
    same runs advanced one by one with pz_step: bitwise under the shapes
    that keep the single-run summation order (0 and 2), 1e-12 under the
    others."""
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Ensemble, Run
    cfgs = W.c3_ensemble(64)
    print("case: 64-run ensemble, 12 steps", flush=True)

    def members():
        out = []
        for k, c in enumerate(cfgs):
            x = Run(c)
            W.seed_member(x, c, k)
            x.start()
            out.append(x)
        return out

    alone = members()
    for x in alone:
        x.step(12)
    batch = members()
    ens = Ensemble(batch)
    ens.step(12)
    pol, ms = ens.policy()
    ens.close()
    print(f"policy {pol}, trial ms {ms}", flush=True)
    worst = 0.0
    for k, (a, b) in enumerate(zip(alone, batch)):
        assert a.time() == b.time(), k
        for f in ("temp", "us", "up", "om", "psi"):
            fa, fb = a.get_field(f), b.get_field(f)
            if pol in (0, 2):
                assert np.array_equal(fa, fb), (k, f, pol)
            else:
                err = rel(fb, fa)
                worst = max(worst, err)
                assert err <= 1e-12, (k, f, pol, err)
    print(f"largest relative difference {worst:.2e}", flush=True)


@pytest.mark.parametrize("member", [19, 5])
def test_ensemble_members_of_the_uneven_set_match_oracle(member):
    """Two members of the 64-run set against the oracle after 500 steps
    (1e-8 per field): member 19 is the largest grid (n_r 129, m_max 170),
    member 5 a small one."""
    from oracle.oracle import Oracle
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Ensemble, Run
    cfgs = W.c3_ensemble(64)
    c = cfgs[member]
    print(f"case: ensemble member {member} (n_r {c.n_r_max}, m_max {c.m_max}) vs oracle, 500 steps", flush=True)
    runs = []
    for k in (member, (member + 1) % 64):
        x = Run(cfgs[k])
        W.seed_member(x, cfgs[k], k)
        x.start()
        runs.append(x)
    ens = Ensemble(runs)
    ens.step(500)
    ens.close()
    o = Oracle(c)
    W.seed_member(o, c, member)
    o.start()
    o.step(500)
    g = runs[0]
    assert abs(g.time()[0] - o.time()[0]) < 1e-12 and g.time()[1] == o.time()[1]
    e = {f: rel(g.get_field(f), o.get_field(f)) for f in ("temp", "us", "up", "om", "psi")}
    _report(f"member {member} after 500 steps:", e)
    for f, v in e.items():
        assert v < 1e-8, (f, v)

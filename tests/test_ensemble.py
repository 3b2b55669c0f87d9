"""Ensembles (pz_ensemble_*): uneven runs advanced together with one batched
launch per kernel and stage.  Each member must equal the same run advanced
alone through pz_step — the batched kernels run the very same per-block
code — so the comparison is bitwise; and the members match the oracle
after 500 steps to 1e-8."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("temp", "us", "up", "om", "psi")


def _members(cfgs, seed_fn):
    from pizza_amd.runtime import Run
    runs = []
    for k, c in enumerate(cfgs):
        r = Run(c)
        seed_fn(r, c, k)
        r.start()
        runs.append(r)
    return runs


def _check_equal(cfgs, seed_fn, nsteps, tol=0.0):
    """tol = 0: bitwise; else the largest relative difference per field."""
    from pizza_amd.runtime import Ensemble
    print(f"case: ensemble of {len(cfgs)} runs, {nsteps} steps", flush=True)
    alone = _members(cfgs, seed_fn)
    for r in alone:
        r.step(nsteps)
    batch = _members(cfgs, seed_fn)
    ens = Ensemble(batch)
    ens.step(nsteps // 2)
    ens.step(nsteps - nsteps // 2)
    pol = ens.policy()
    ens.close()
    for k, (a, b) in enumerate(zip(alone, batch)):
        assert a.time() == b.time(), k
        for f in FIELDS:
            fa, fb = a.get_field(f), b.get_field(f)
            if tol == 0.0:
                assert np.array_equal(fa, fb), (k, f, pol)
            else:
                assert np.abs(fa - fb).max() <= tol * max(np.abs(fa).max(), 1e-300), (k, f, pol)
    return pol


def _c3_pick():
    from pizza_amd import workloads as W
    cfgs = W.c3_ensemble(nruns=64)
    pick = []
    for c in cfgs:   # one member of each n_r, plus the largest n_phi
        if c.n_r_max not in [p.n_r_max for p in pick] or (c.m_max == 170 and all(p.m_max != 170 for p in pick)):
            pick.append(c)
    assert len({c.n_r_max for c in pick}) == 4
    return pick


@pytest.mark.parametrize("policy", [0, 1, 2, 3])
def test_uneven_collocation_ensemble_equals_single_runs(policy, monkeypatch):
    """Every m-loop policy of the autotuner: policies 0 and 2 (radial rows in
    LDS or read through L2) run the single-run arithmetic bit for bit;
    1 and 3 (256 threads) sum the dense products in another order."""
    from pizza_amd import workloads as W
    monkeypatch.setenv("PZ_ENSEMBLE_POLICY", str(policy))
    pol = _check_equal(_c3_pick(), W.seed_member, 6, tol=0.0 if policy in (0, 2) else 1e-12)
    assert pol[0] == policy


def test_autotuned_ensemble_matches_single_runs(monkeypatch):
    """The autotuner times every policy from the same member states and
    restores them: the tuned ensemble still matches the single runs."""
    from pizza_amd import workloads as W
    monkeypatch.delenv("PZ_ENSEMBLE_POLICY", raising=False)
    pol, ms = _check_equal(_c3_pick(), W.seed_member, 6, tol=1e-12)
    assert np.all(ms > 0) and ms[pol] == ms.min()


def test_integration_ensemble_equals_single_runs():
    from pizza_amd import workloads as W
    cfgs = [W.small_bpr353(33, 32), W.small_bpr353(49, 48), W.small_bpr353(33, 64)]
    cfgs[1].ra = 1.0e7
    _check_equal(cfgs, lambda r, c, k: W.seed_run(r, c, seed=W.SEED + k), 4)


def test_ensemble_rejects_mixed_methods():
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Ensemble, PizzaError, Run
    a, b = Run(W.c1_collocation()), Run(W.c1_collocation_shape())
    for r in (a, b):
        r.start()
    with pytest.raises(PizzaError):
        Ensemble([a, b])


def test_ensemble_members_match_oracle_over_500_steps():
    """Each member of an uneven collocation ensemble against the oracle after
    500 steps, 1e-8 relative per field (one member changes dt under the
    Courant rule, so its matrices are rebuilt inside the batched step)."""
    from helpers import rel
    from oracle.oracle import Oracle
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Ensemble
    cfgs = []
    for n_r, m_max, ra in ((33, 32, 1e5), (33, 16, 3e5), (49, 21, 1e5)):
        c = W.c1_collocation()
        c.n_r_max = c.n_cheb_max = n_r
        c.m_max = m_max
        c.ra = ra
        cfgs.append(c)
    print("case: ensemble of 3 runs against the oracle, 500 steps", flush=True)
    runs = _members(cfgs, W.seed_member)
    ens = Ensemble(runs)
    ens.step(500)
    ens.close()
    for k, (c, r) in enumerate(zip(cfgs, runs)):
        o = Oracle(c)
        W.seed_member(o, c, k)
        o.start()
        o.step(500)
        assert abs(r.time()[0] - o.time()[0]) < 1e-12 and r.time()[1] == o.time()[1], k
        for f in FIELDS:
            err = rel(r.get_field(f), o.get_field(f))
            assert err < 1e-8, (k, f, err)

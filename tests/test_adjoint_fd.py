"""The discrete adjoint of the finite-difference step (k_fd_mloop_adj,
k_fd_start_adj; the radial loop's transpose is shared with collocation).

The reference has no adjoint; as for collocation (tests/test_adjoint.py)
the transposed kernels are pinned by properties of the step:
- over two steps <W, T(2)> is a quadratic polynomial of the initial T, so
  central differences equal <grad, d> for any step size: an exact
  dot-product test of every transposed piece (right-hand sides and slots,
  banded substitutions, ghost points, implicit terms and stencils), for
  CNAB2, SBDF2 and the CNAB2 start-up steps of SBDF3, tolerance 1e-9;
- the kinetic energy averaged over 40 SBDF3 steps against central
  differences of the GPU forward and of the oracle's forward, 1e-6;
- the gradients with respect to wall patterns (l_fd_bc_pattern): a fixed
  value (exact over two steps) and an outer heat-flux pattern (the C5
  objective, the kinetic energy), against central differences of the
  pattern coefficients.
"""
import numpy as np
import pytest

from test_adjoint import _J, _dot, _setup

pytestmark = pytest.mark.gpu


def _cfg(scheme="CNAB2", n_r=33, m_max=32):
    from pizza_amd import workloads as W
    from pizza_amd.config import scheme_id
    cfg = W.c1_collocation()
    cfg.n_r_max = cfg.n_cheb_max = n_r
    cfg.m_max = m_max
    cfg.l_cheb_coll = 0
    cfg.l_finite_diff = 1
    cfg.time_scheme = scheme_id(scheme)
    return cfg


@pytest.mark.parametrize("scheme,amp_u", [("CNAB2", 1.0e-3), ("CNAB2", 5.0e-2), ("SBDF2", 1.0e-3),
                                          ("SBDF3", 5.0e-2)])
def test_fd_two_step_dot_product(scheme, amp_u):
    cfg = _cfg(scheme)
    print(f"case: FD adjoint dot product {scheme} amp_u={amp_u}", flush=True)
    run, T0, us, up, field = _setup(cfg, amp_u=amp_u)
    rng = np.random.default_rng(11)
    Wt = rng.standard_normal((run.n_m, run.n_r)) + 1j * rng.standard_normal((run.n_m, run.n_r))
    _J(run, T0, us, up, "temp", 2, Wt)
    g, _ = run.adjoint_backward()
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    for k in range(3):
        d = field(1.0e-3)
        jp = _J(run, T0 + d, us, up, "temp", 2, Wt)
        jm = _J(run, T0 - d, us, up, "temp", 2, Wt)
        fd = (jp - jm) / 2.0
        ad = _dot(g, d)
        assert abs(fd - ad) <= 1e-9 * max(abs(fd), abs(ad)), (k, fd, ad)


def test_fd_nusselt_objective_matches_nusselt_series():
    cfg = _cfg()
    run, T0, us, up, _ = _setup(cfg)
    J = _J(run, T0, us, up, "nutop", 5, ckpt=2)
    run.set_field("us", us)
    run.set_field("up", up)
    run.set_field("temp", T0)
    run.start()
    nu = []
    for _ in range(5):
        run.step(1)
        nu.append(run.nusselt()[0])
    assert abs(J - np.mean(nu)) < 1e-12


def test_fd_ekin_gradient_vs_central_differences():
    """dJ/dT0 of the kinetic energy averaged over 40 SBDF3 steps (the
    start-up steps and the scheme's own) against central differences of
    the GPU forward and of the oracle's forward."""
    from oracle.oracle import Oracle
    cfg = _cfg("SBDF3")
    K = 40
    print("case: FD adjoint ekin vs central differences", flush=True)
    run, T0, us, up, field = _setup(cfg, amp_u=0.0)
    zero = np.zeros_like(T0)
    _J(run, T0, zero, zero, "ekin", K, ckpt=10)
    g, _ = run.adjoint_backward()

    def j_oracle(T):
        o = Oracle(cfg)
        o.set_field("temp", T)
        o.start()
        e = []
        for _ in range(K):
            o.step(1)
            v = o.ekin_3d()
            e.append(v[0] + v[1] + v[2])
        return float(np.mean(e))

    for k in range(3):
        d = field(1.0e-3)
        h = 1.0e-5
        fd_gpu = (_J(run, T0 + h * d, zero, zero, "ekin", K) - _J(run, T0 - h * d, zero, zero, "ekin", K)) / (2 * h)
        ad = _dot(g, d)
        assert abs(fd_gpu - ad) <= 1e-6 * abs(ad), (k, fd_gpu, ad)
        if k < 2:
            fd_or = (j_oracle(T0 + h * d) - j_oracle(T0 - h * d)) / (2 * h)
            assert abs(fd_or - ad) <= 1e-6 * abs(ad), (k, fd_or, ad)


@pytest.mark.parametrize("ktopt,objective", [(1, "temp"), (2, "temp"), (2, "ekin")])
def test_fd_wall_pattern_gradient(ktopt, objective):
    """Gradient with respect to outer-wall patterns of T (l_fd_bc_pattern:
    a fixed value, ktopt = 1, or a heat flux, ktopt = 2; modes m = 2 and
    5) over two steps against central differences of the coefficients:
    exact for <W, T(2)>, 1e-6 for the kinetic energy."""
    from pizza_amd.runtime import Run
    print(f"case: FD adjoint wall pattern ktopt={ktopt} {objective}", flush=True)
    base = [(2, 0.05, 0.02), (5, 0.01, -0.03)]

    def make(pattern):
        cfg = _cfg()
        cfg.ktopt = ktopt
        cfg.l_fd_bc_pattern = 1
        for n, (m, tr, ti) in enumerate(pattern):
            cfg.t_top[3 * n], cfg.t_top[3 * n + 1], cfg.t_top[3 * n + 2] = m, tr, ti
        return cfg

    run, T0, us, up, _ = _setup(make(base), amp_u=5.0e-2)
    rng = np.random.default_rng(5)
    Wt = rng.standard_normal((run.n_m, run.n_r)) + 1j * rng.standard_normal((run.n_m, run.n_r))
    W = Wt if objective == "temp" else None
    _J(run, T0, us, up, objective, 2, W)
    _, gbc = run.adjoint_backward()
    tol = 1e-9 if objective == "temp" else 1e-6
    for n, (m, tr, ti) in enumerate(base):
        for part in (1, 2):
            h = 0.01 if objective == "temp" else 1.0e-4
            js = []
            for sgn in (1, -1):
                pat = [list(e) for e in base]
                pat[n][part] += sgn * h
                r2 = Run(make(pat))
                js.append(_J(r2, T0, us, up, objective, 2, W))
                r2.close()
            fd = (js[0] - js[1]) / (2 * h)
            ad = gbc[m, 0].real if part == 1 else gbc[m, 0].imag
            assert abs(fd - ad) <= tol * max(abs(fd), abs(ad), 1e-30), (m, part, fd, ad)


def test_fd_long_grid_two_step_dot_product():
    """The long azimuthal grid (n_phi 2046 = 2 3 11 31, the C5 grid) takes
    the transposed in-place radial loop (k_radial_adj_inplace after
    k_gather_r on the replayed state): exact two-step dot-product test."""
    cfg = _cfg("SBDF3", n_r=33, m_max=682)
    print("case: FD adjoint dot product, n_phi 2046", flush=True)
    run, T0, us, up, field = _setup(cfg, amp_u=1.0e-2)
    rng = np.random.default_rng(13)
    Wt = rng.standard_normal((run.n_m, run.n_r)) + 1j * rng.standard_normal((run.n_m, run.n_r))
    _J(run, T0, us, up, "temp", 2, Wt)
    g, _ = run.adjoint_backward()
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    for k in range(2):
        d = field(1.0e-3)
        jp = _J(run, T0 + d, us, up, "temp", 2, Wt)
        jm = _J(run, T0 - d, us, up, "temp", 2, Wt)
        fd = (jp - jm) / 2.0
        ad = _dot(g, d)
        assert abs(fd - ad) <= 1e-9 * max(abs(fd), abs(ad)), (k, fd, ad)


def test_c5_flux_pattern_gradient():
    """The C5 gradient at its own size: n_r 257 (stretched), n_phi 2046,
    E 1e-6, Ra 1e9, Pr 0.1, SBDF3 dt 1e-7, 100 steps with checkpoints every
    10; the kinetic energy averaged over the window with respect to the
    seeded outer-wall heat-flux pattern (16 modes, modulus 0.1), against
    fourth-order central differences of the GPU forward along 8 seeded
    directions of the 16 complex coefficients (steps 3.2e-2 and 6.4e-2;
    tolerance 1e-6)."""
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    base = W.c5_flux_pattern()
    K, ck = 100, 10
    print("case: C5 flux-pattern gradient, 100 steps", flush=True)

    def forward(pattern, keep=False):
        cfg = W.c5_gradient(pattern)
        r = Run(cfg)
        W.seed_run(r, cfg, amp_t=1.0e-3)
        r.start()
        J = r.adjoint_forward("ekin", K, ck)
        if keep:
            return r, J
        r.close()
        return J

    run, J0 = forward(base, keep=True)
    _, gbc = run.adjoint_backward()
    run.close()
    g = np.array([gbc[m, 0] for m, _, _ in base])
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    rng = np.random.default_rng(21)

    def central(d, h):
        js = []
        for sgn in (1, -1):
            pat = [(m, tr + sgn * h * dd.real, ti + sgn * h * dd.imag) for (m, tr, ti), dd in zip(base, d)]
            js.append(forward(pat))
        return (js[0] - js[1]) / (2 * h)

    for k in range(8):
        d = rng.standard_normal(len(base)) + 1j * rng.standard_normal(len(base))
        # J is about 3.3e6 while its change along d is a few times 10 h: the
        # rounding of J (a few 1e-8) limits a central difference to about
        # 3e-8 / h, and the h^2 term grows with h, so the fourth-order
        # combination (4 D(h) - D(2h)) / 3 of two large steps is used
        # (h = 3.2e-2 of the modulus-0.1 coefficients).  Measured over the 8
        # directions: 1e-9 .. 9e-8 relative; the plain central differences
        # stray by up to 2e-6 at h = 8e-3 and 1.2e-5 at 3.2e-2.
        h = 3.2e-2
        fd = (4.0 * central(d, h) - central(d, 2.0 * h)) / 3.0
        ad = float(np.sum(g.real * d.real + g.imag * d.imag))
        print(f"direction {k}: fourth-order central difference {fd:.10e}, adjoint {ad:.10e}", flush=True)
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(ad)), (k, fd, ad, J0)


def test_c5_two_step_pattern_gradient_exact():
    """At the C5 size, over two steps the objective <W, T(2)> is quadratic
    in the pattern as in T0: the central difference along a seeded
    direction equals the pattern gradient to 1e-9."""
    from pizza_amd import workloads as W
    from pizza_amd.runtime import Run
    base = W.c5_flux_pattern()
    print("case: C5 two-step pattern gradient", flush=True)
    rng = np.random.default_rng(3)

    def forward(pattern, Wt, keep=False):
        cfg = W.c5_gradient(pattern)
        r = Run(cfg)
        W.seed_run(r, cfg, amp_t=1.0e-3)
        r.start()
        if Wt is None:
            Wt = rng.standard_normal((r.n_m, r.n_r)) + 1j * rng.standard_normal((r.n_m, r.n_r))
        J = r.adjoint_forward("temp", 2, 10, Wt)
        if keep:
            return r, J, Wt
        r.close()
        return J

    run, _, Wt = forward(base, None, keep=True)
    _, gbc = run.adjoint_backward()
    run.close()
    g = np.array([gbc[m, 0] for m, _, _ in base])
    d = rng.standard_normal(len(base)) + 1j * rng.standard_normal(len(base))
    h = 1.0e-3
    js = [forward([(m, tr + s * h * dd.real, ti + s * h * dd.imag) for (m, tr, ti), dd in zip(base, d)], Wt)
          for s in (1, -1)]
    fd = (js[0] - js[1]) / (2 * h)
    ad = float(np.sum(g.real * d.real + g.imag * d.imag))
    assert abs(fd - ad) <= 1e-9 * max(abs(fd), abs(ad)), (fd, ad)

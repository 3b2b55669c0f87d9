"""Named workloads of the IMEX step and their seeded initial conditions.

``c2_production`` is the headline configuration: sparse-Chebyshev
integration, n_r_max = 97, m_max = 128 (n_phi = 384), Ekman 1e-5, Ra 3e7,
Pr 1, thermal + compositional (Sc 3), DIRK BPR353, dt 5e-6.
``c1_collocation`` is the reference's small collocation case (n_r 33,
n_phi 96, Ekman 1e-3, Ra 1e5, Ekman pumping, CNAB2 dt 1e-4);
``c1_collocation_shape`` keeps those parameters on the integration method.

Seeded perturbation (both runs get the same fields through set_field):
    f(m, r) = A * g_m * sin(pi (r - r_icb)),  m = 1..m_max,
g_m complex Gaussian drawn from numpy's PCG64 with a fixed seed.
"""
import numpy as np

from .config import BPR353, CNAB2, default_config

SEED = 20261202


def c2_production(n_r=97, m_max=128, raxi=3.0e7):
    return default_config(
        n_r_max=n_r, n_cheb_max=n_r, m_max=m_max, minc=1, time_scheme="BPR353",
        cheb_method="integ", matrix_solve="direct", dtMax=5.0e-6,
        ra=3.0e7, ek=1.0e-5, pr=1.0, raxi=raxi, sc=3.0, radratio=0.35,
        ktopt=1, kbott=1, ktopxi=1, kbotxi=1, ktopv=2, kbotv=2, l_ek_pump=True,
        g1=1.0)


def c1_collocation():
    c = default_config(
        n_r_max=33, n_cheb_max=33, m_max=32, minc=1, time_scheme="CNAB2",
        cheb_method="colloc", matrix_solve="direct", dtMax=1.0e-4,
        ra=1.0e5, ek=1.0e-3, pr=1.0, radratio=0.35, l_ek_pump=True)
    assert c.time_scheme == CNAB2 and c.l_cheb_coll == 1
    return c


def c1_collocation_shape():
    c = default_config(
        n_r_max=33, n_cheb_max=33, m_max=32, minc=1, time_scheme="CNAB2",
        cheb_method="integ", matrix_solve="direct", dtMax=1.0e-4,
        ra=1.0e5, ek=1.0e-3, pr=1.0, radratio=0.35, l_ek_pump=True)
    assert c.time_scheme == CNAB2
    return c


def small_bpr353(n_r=33, m_max=32):
    """A CPU-sized version of the production physics (parity tests)."""
    c = c2_production(n_r=n_r, m_max=m_max)
    assert c.time_scheme == BPR353
    return c


def seeded_field(n_m, r, r_icb, amp, seed=SEED, stream=0):
    """complex (n_m, n_r) perturbation A g_m sin(pi (r - r_icb)), m >= 1."""
    rng = np.random.Generator(np.random.PCG64([seed, stream]))
    g = rng.standard_normal(n_m) + 1j * rng.standard_normal(n_m)
    g[0] = 0.0
    prof = np.sin(np.pi * (np.asarray(r) - r_icb))
    return amp * g[:, None] * prof[None, :]


def fd_grid(n, r_cmb, r_icb, ratio1, ratio2):
    """Stretched finite-difference radii (finite_differences.f90:87-188),
    outer boundary first."""
    r = np.zeros(n)
    r[0] = r_cmb
    if ratio2 == 1.0:
        drb = 1.0 / (n - 1.0)
        for k in range(1, n):
            r[k] = r[k - 1] - drb
    else:
        nbp = int((n - 1) / (2.0 * (1.0 + ratio1)))
        nbulk = n - 1 - 2 * nbp
        dra = np.exp(np.log(ratio2) / nbp)
        drb = dra ** nbp
        drb = 1.0 / (nbulk + 2.0 * dra * ((1.0 - drb) / (1.0 - dra)))
        drb *= dra ** nbp
        for k in range(1, nbp + 1):
            r[k] = r[k - 1] - drb
            drb /= dra
        for k in range(1, nbulk + 1):
            r[k + nbp] = r[k + nbp - 1] - drb
        for k in range(1, nbp + 1):
            drb *= dra
            r[k + nbp + nbulk] = r[k + nbp + nbulk - 1] - drb
    r[n - 1] = r_icb
    return r


def radial_grid(cfg):
    """Radii r_k, outer boundary first: Gauss-Lobatto points
    (chebyshev.f90:188-240) or the finite-difference grid."""
    n = cfg.n_r_max
    r_cmb = 1.0 / (1.0 - cfg.radratio)
    r_icb = r_cmb - 1.0
    if getattr(cfg, "l_finite_diff", 0):
        return fd_grid(n, r_cmb, r_icb, cfg.fd_stretch, cfg.fd_ratio), r_icb
    x = np.cos(np.pi * np.arange(n) / (n - 1))
    return 0.5 * (r_cmb - r_icb) * x + 0.5 * (r_cmb + r_icb), r_icb


def seed_run(run, cfg, amp_t=1.0e-4, amp_xi=1.0e-4, seed=SEED):
    """Injects the seeded T and xi perturbations into a Run or an oracle."""
    r, r_icb = radial_grid(cfg)
    n_m = cfg.m_max // cfg.minc + 1
    if cfg.ra != 0.0:
        run.set_field("temp", seeded_field(n_m, r, r_icb, amp_t, seed, 0))
    if cfg.raxi != 0.0:
        run.set_field("xi", seeded_field(n_m, r, r_icb, amp_xi, seed, 1))


C3_N_R = (33, 65, 97, 129)
C3_M_MAX = (32, 64, 128, 170)  # n_phi 96, 192, 384, 510 (512 is not reachable)


def c3_ensemble(nruns=64, seed=SEED):
    """The uneven ensemble: per run, n_r and n_phi drawn from the sets above,
    Ra log-uniform in [1e5, 1e8], Ekman 1e-4 or 1e-5, collocation, CNAB2,
    dtMax = Ek / 10.  Returns the configurations in run order."""
    rng = np.random.Generator(np.random.PCG64([seed, 99]))
    cfgs = []
    for _ in range(nruns):
        n_r = int(rng.choice(C3_N_R))
        m_max = int(rng.choice(C3_M_MAX))
        ra = float(10.0 ** rng.uniform(5.0, 8.0))
        ek = float(rng.choice((1.0e-4, 1.0e-5)))
        cfgs.append(default_config(
            n_r_max=n_r, n_cheb_max=n_r, m_max=m_max, minc=1, time_scheme="CNAB2",
            cheb_method="colloc", matrix_solve="direct", dtMax=0.1 * ek,
            ra=ra, ek=ek, pr=1.0, radratio=0.35))
    return cfgs


def seed_member(run, cfg, k, amp_t=1.0e-3, seed=SEED):
    """Seeded T perturbation of ensemble member k (seed base + member id)."""
    r, r_icb = radial_grid(cfg)
    n_m = cfg.m_max // cfg.minc + 1
    run.set_field("temp", seeded_field(n_m, r, r_icb, amp_t, seed + 1000 + k, 0))


def c4_adjoint(n_r=65, m_max=85):
    """The adjoint case: collocation, n_r 65, n_phi 256 (m_max 85), Ekman
    1e-4, Ra 1e6, Pr 1, CNAB2, dtMax = Ek / 10; objective: the outer-wall
    Nusselt number averaged over 400 steps, checkpoints every 20 steps."""
    return default_config(
        n_r_max=n_r, n_cheb_max=n_r, m_max=m_max, minc=1, time_scheme="CNAB2",
        cheb_method="colloc", matrix_solve="direct", dtMax=1.0e-5,
        ra=1.0e6, ek=1.0e-4, pr=1.0, radratio=0.35)


def c5_fd(n_r=257, m_max=682, steps_scheme="SBDF3"):
    """The large finite-difference case: n_r 257 on the stretched grid,
    n_phi 2046 (m_max 682), Ekman 1e-6, Ra 1e9, Pr 0.1, SBDF3, dt 1e-7.
    The reference's finite differences are second order."""
    return default_config(
        n_r_max=n_r, n_cheb_max=n_r, m_max=m_max, minc=1, time_scheme=steps_scheme,
        radial_scheme="FD", dtMax=1.0e-7, ra=1.0e9, ek=1.0e-6, pr=0.1, radratio=0.35)


C5_PATTERN_MODES = 16


def c5_flux_pattern(seed=SEED, amp=0.1, modes=C5_PATTERN_MODES):
    """The seeded outer-wall heat-flux pattern of the C5 gradient: modes
    m = 1 .. 16, complex coefficients of modulus amp with random phases."""
    rng = np.random.Generator(np.random.PCG64([seed, 55]))
    ph = rng.uniform(0.0, 2.0 * np.pi, modes)
    return [(m, amp * np.cos(p), amp * np.sin(p)) for m, p in zip(range(1, modes + 1), ph)]


def c5_gradient(pattern=None, n_r=257, m_max=682):
    """C5 with the heat-flux pattern on the outer wall (ktopt = 2, the
    finite-difference wall patterns switched on): the configuration of the
    gradient of the mean kinetic energy with respect to that pattern."""
    c = c5_fd(n_r=n_r, m_max=m_max)
    c.ktopt = 2
    c.l_fd_bc_pattern = 1
    for n, (m, tr, ti) in enumerate(pattern if pattern is not None else c5_flux_pattern()):
        c.t_top[3 * n], c.t_top[3 * n + 1], c.t_top[3 * n + 2] = m, tr, ti
    return c

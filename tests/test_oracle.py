"""CPU tests that pin the oracle (oracle/pz_oracle.c) before it is trusted.

1. End-to-end against the reference's own golden time series
   (samples/RotInteg/reference.out, first run: integration method, CNAB2,
   rotating, implicit Coriolis and buoyancy, l_rerror_fix, n_cheb_max <
   n_r_max), compared with the reference test's rtol 1e-8
   (samples/RotInteg/unitTest.py:116-120).
2. Closed forms for the transforms (DCT-I involution and Chebyshev-node
   identities, single-bin azimuthal FFT, exact derivatives of polynomials).
3. The sparse-Chebyshev stencils against an independent construction with
   numpy.polynomial.chebyshev (apply the differential operator to T_j,
   integrate J times, read the coefficient of T_n).
4. The collocation method end to end against the reference's golden series
   samples/RotE1e3EkPump/reference.out, first run (input.nml: collocation,
   CNAB2, Ekman pumping, n_r = 32, single-mode start), at rtol 1e-8
   (samples/RotE1e3EkPump/unitTest.py:116-120).
"""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

from helpers import golden_rows, rel, rotinteg_config
from oracle.oracle import Oracle


@pytest.fixture(scope="module")
def rot():
    cfg, start, _ = rotinteg_config()
    return cfg, start


def test_rotinteg_golden_series(rot):
    """1000 CNAB2 steps (t = 0..100) of the RotInteg start run."""
    cfg, start = rot
    ref = golden_rows(51)
    o = Oracle(cfg)
    o.init_fields(init_t=start["init_t"], amp_t=start["amp_t"])
    o.start()
    rows = []
    for k in range(51):
        if k:
            o.step(20)
        t, _ = o.time()
        rows.append(np.r_[t, o.ekin_3d()])
    rows = np.array(rows)
    np.testing.assert_allclose(rows, ref, rtol=1e-8, atol=1e-20)


def test_rotekpump_collocation_golden_series():
    """1000 CNAB2 steps (t = 0..0.1) of the collocation start run."""
    import os
    from helpers import GOLDEN
    from pizza_amd.config import config_from_namelist
    cfg, start, _ = config_from_namelist(os.path.join(GOLDEN, "rotekpump_input_start.nml"))
    assert cfg.l_cheb_coll == 1
    ref = np.loadtxt(os.path.join(GOLDEN, "rotekpump_start_e_kin_3D.txt"))
    o = Oracle(cfg)
    o.init_fields(init_t=start["init_t"], amp_t=start["amp_t"])
    o.start()
    rows = []
    for k in range(len(ref)):
        if k:
            o.step(100)
        t, _ = o.time()
        rows.append(np.r_[t, o.ekin_3d()])
    np.testing.assert_allclose(np.array(rows), ref, rtol=1e-8, atol=1e-20)


def test_dirkschemes_ars222_collocation_golden_series():
    """100 ARS222 steps of the reference's DIRKSchemes start run
    (collocation, n_r = 33; samples/DIRKSchemes/reference.out rows 1-11)."""
    from helpers import sample_config, sample_series
    cfg, _, _ = sample_config("dirkschemes")
    rows, ref = sample_series(Oracle(cfg), "dirkschemes", 10)
    np.testing.assert_allclose(rows, ref, rtol=1e-8, atol=1e-20)


def test_multistep_sbdf3_collocation_golden_series():
    """200 SBDF3 steps of the reference's MultistepSchemes start run
    (collocation, n_r = 32, CNAB2 start-up steps; every step logged;
    samples/MultistepSchemes/reference.out rows 1-201)."""
    from helpers import sample_config, sample_series
    cfg, _, _ = sample_config("multistep")
    rows, ref = sample_series(Oracle(cfg), "multistep", 1)
    np.testing.assert_allclose(rows, ref, rtol=1e-8, atol=1e-20)


def _oracle(n_r=33, m_max=32):
    from pizza_amd.workloads import small_bpr353
    return Oracle(small_bpr353(n_r, m_max))


@pytest.mark.parametrize("n_r", [33, 194])
def test_costf_closed_forms(n_r):
    """DCT-I closed forms; at n_r = 194 the transform length 386 = 2 x 193
    runs a prime radix of 193 in the oracle's FFT."""
    o = _oracle(n_r)
    N = o.n_r
    r, x = o.grid()
    # DCT-I of T_k at the Gauss-Lobatto nodes is a scaled unit vector
    for k in (0, 1, 5, N - 2, N - 1):
        f = np.cos(k * np.arccos(np.clip(x, -1, 1)))[None, :].astype(complex)
        c = o.costf(f)[0]
        expect = np.zeros(N)
        expect[k] = (N - 1) * np.sqrt(0.5 / (N - 1)) * (2.0 if k in (0, N - 1) else 1.0)
        assert np.max(np.abs(c - expect)) < 1e-13 * (N - 1)
    # involution
    rng = np.random.default_rng(3)
    f = rng.standard_normal((4, N)) + 1j * rng.standard_normal((4, N))
    assert rel(o.costf(o.costf(f)), f) < 1e-14
    # brute-force cosine sum
    j = np.arange(N)
    w = np.where((j == 0) | (j == N - 1), 1.0, 2.0)
    M = np.sqrt(0.5 / (N - 1)) * w[None, :] * np.cos(np.pi * np.outer(j, j) / (N - 1))
    # both sides round length-N sums (a radix-193 DFT at N = 194)
    assert rel(o.costf(f), f @ M.T) < 1e-14 * max(1.0, N / 33)


def test_azimuthal_fft_closed_forms():
    o = _oracle(17, 32)
    n = o.n_phi
    phi = 2 * np.pi * np.arange(n) / n
    g = np.stack([np.cos(5 * phi + 0.3), 2.0 + 0 * phi, np.sin(32 * phi)], axis=1)
    s = o.fft(g)
    expect = np.zeros((o.n_m, 3), dtype=complex)
    expect[5, 0] = 0.5 * np.exp(0.3j)
    expect[0, 1] = 2.0
    expect[32, 2] = -0.5j
    assert np.max(np.abs(s - expect)) < 1e-15 * n
    rng = np.random.default_rng(4)
    sp = rng.standard_normal((o.n_m, 3)) + 1j * rng.standard_normal((o.n_m, 3))
    sp[0] = sp[0].real
    assert rel(o.fft(o.ifft(sp)), sp) < 1e-14


def test_radial_derivatives_exact_on_polynomials():
    o = _oracle(33)
    r, _ = o.grid()
    f = np.stack([r ** 5, 3 * r ** 2 - r ** 7]).astype(complex)
    d1, d2 = o.get_ddr(f)
    e1 = np.stack([5 * r ** 4, 6 * r - 7 * r ** 6])
    e2 = np.stack([20 * r ** 3, 6 - 42 * r ** 5])
    assert rel(d1, e1) < 1e-12
    assert rel(d2, e2) < 1e-10
    assert rel(o.get_dr(f), e1) < 1e-12


# ---- stencils -------------------------------------------------------------
def _cheb_of_poly_in_r(pcoef_r, a, b):
    """Chebyshev series (in x) of sum_k p_k r^k with r = a x + b."""
    rx = np.array([b, a])
    out = np.zeros(1)
    powk = np.array([1.0])
    for k, pk in enumerate(pcoef_r):
        if k:
            powk = C.chebmul(powk, rx)
        if pk:
            out = C.chebadd(out, pk * powk)
    return out


def _apply(poly_ops, J, j, a, b):
    """I_r^J [sum_k p_k(r) D_r^k] T_j (T_0 counted with weight 1/2)."""
    basis = np.zeros(j + 1)
    basis[j] = 0.5 if j == 0 else 1.0
    acc = np.zeros(1)
    for k, p in poly_ops.items():
        d = C.chebder(basis, m=k) / a ** k if k else basis
        acc = C.chebadd(acc, C.chebmul(_cheb_of_poly_in_r(p, a, b), d))
    return C.chebint(acc, m=J) * a ** J


def _stencil_ref(poly_ops, J, idx, K, a, b):
    """Folded stencil of row idx in the reference's band order."""
    st = np.zeros(2 * K + 1)
    for nb in range(2 * K + 1):
        col = idx + K - nb
        if col < 0:
            continue
        g = _apply(poly_ops, J, col, a, b)
        st[nb] = g[idx] if idx < len(g) else 0.0
    return st


@pytest.mark.parametrize("which,J,K,m", [(0, 2, 4, 0), (1, 2, 3, 0), (2, 2, 2, 7), (5, 4, 8, 0),
                                         (6, 4, 7, 0), (7, 4, 8, 5), (8, 4, 8, 5), (4, 1, 2, 0)])
def test_sparse_stencils(which, J, K, m):
    o = _oracle(25, 16)
    r, _ = o.grid()
    r_cmb, r_icb = r[0], r[-1]
    a, b = 0.5 * (r_cmb - r_icb), 0.5 * (r_cmb + r_icb)
    rc2 = r_cmb ** 2
    m2 = m * m
    # polynomial coefficient lists (ascending powers of r) per derivative order
    if which == 0:
        ops = {0: [0, 0, 1]}
    elif which == 1:
        ops = {0: [0, 1]}
    elif which == 2:  # r^2 Delta_m
        ops = {2: [0, 0, 1], 1: [0, 1], 0: [-m2]}
    elif which == 4:
        ops = {0: [0, 1]}
    elif which == 5:
        ops = {0: [0, 0, 0, 0, 1]}
    elif which == 6:
        ops = {0: [0, 0, 0, 1]}
    else:
        # q = r^2 omega(psi) = -r^2 h^2 psi'' - (rc^2 r - 6 r^3) psi' + (6 r^2 + m^2 h^2) psi
        q = {2: [0, 0, -rc2, 0, 1], 1: [0, -rc2, 0, 6], 0: [m2 * rc2, 0, 6 - m2]}
        if which == 7:  # r^4 omega = r^2 q
            ops = {k: [0, 0] + v for k, v in q.items()}
        else:  # r^4 Delta_m omega = r^2 q'' - 3 r q' + (4 - m^2) q
            P = np.polynomial.polynomial
            ops = {}

            def add(k, p):
                cur = ops.get(k, [0])
                ops[k] = list(P.polyadd(cur, p))
            for k, p in q.items():
                dp, ddp = P.polyder(p), P.polyder(p, 2)
                # q'' = p'' psi^(k) + 2 p' psi^(k+1) + p psi^(k+2)
                add(k, P.polymul([0, 0, 1], ddp))
                add(k + 1, P.polymul([0, 0, 2], dp))
                add(k + 2, P.polymul([0, 0, 1], p))
                # -3 r q' = -3 r (p' psi^(k) + p psi^(k+1))
                add(k, P.polymul([0, -3], dp))
                add(k + 1, P.polymul([0, -3], p))
                add(k, P.polymul([4 - m2], p))
    for idx in range(J, 25):
        got = o.stencil(which, m, idx, K)
        ref = _stencil_ref(ops, J, idx, K, a, b)
        cols = idx + K - np.arange(2 * K + 1)
        got = np.where(cols >= 0, got, 0.0)  # unfolded negative-column leftovers
        scale = max(np.max(np.abs(ref)), 1e-300)
        assert np.max(np.abs(got - ref)) / scale < 1e-12, (which, idx, got, ref)


def _fd_cfg(n_r, fd, **kw):
    from pizza_amd.config import default_config
    base = dict(n_r_max=n_r, n_cheb_max=n_r, m_max=8, minc=1, time_scheme="CNAB2",
                radial_scheme="FD" if fd else "CHEB", cheb_method="colloc", dtMax=1e-4,
                ra=1e5, ek=1e-3, pr=1.0, radratio=0.35, buo_term="explicit")
    base.update(kw)
    return default_config(**base)


def test_fd_grid_matches_stretched_grid():
    """The oracle's FD radii equal the stretched grid restated on the host
    side (finite_differences.f90:87-188) for stretched and uniform grids."""
    from pizza_amd import workloads as W
    for ratio in (0.1, 1.0):
        cfg = _fd_cfg(129, True, fd_ratio=ratio)
        r, _ = Oracle(cfg).grid()
        assert np.array_equal(r, W.radial_grid(cfg)[0])
        assert r[0] == 1.0 / 0.65 and r[-1] == 1.0 / 0.65 - 1.0
        assert np.all(np.diff(r) < 0)


def test_fd_converges_to_chebyshev_solution():
    """Finite differences (second order) against the pinned Chebyshev
    collocation method on the same physics after 10 steps: T, u_s and u_phi
    differences shrink by about 4 per halving of the grid spacing, and omega
    converges away from the outer wall (where beta is singular)."""
    from numpy.polynomial import chebyshev as C
    from pizza_amd import workloads as W

    def run(cfg):
        o = Oracle(cfg)
        W.seed_run(o, cfg, amp_t=1e-3)
        o.start()
        o.step(10)
        return o

    oc = run(_fd_cfg(49, False))
    rc, xc = oc.grid()
    V = C.chebvander(xc, len(xc) - 1)
    errs = {}
    for n_r in (65, 129):
        of = run(_fd_cfg(n_r, True))
        rf, _ = of.grid()
        xq = (2 * rf - rc[0] - rc[-1]) / (rc[0] - rc[-1])
        for f in ("temp", "us", "up", "om"):
            coef = np.linalg.solve(V, oc.get_field(f).T).T
            ci = np.array([C.chebval(xq, c) for c in coef])
            d = np.abs(of.get_field(f) - ci) / np.max(np.abs(ci))
            errs[(n_r, f)] = d[:, 1:].max() if f == "om" else d.max()
    for f in ("temp", "us", "up"):
        assert errs[(65, f)] < 1e-2, (f, errs[(65, f)])
        assert errs[(65, f)] / errs[(129, f)] > 3.0, (f, errs[(65, f)], errs[(129, f)])
    assert errs[(129, "om")] < 2e-2 and errs[(129, "om")] < errs[(65, "om")]


@pytest.mark.parametrize("scheme", ["ARS343", "KC564"])
def test_fd_assembly_converges_to_chebyshev_assembly(scheme):
    """The assembly stage on finite differences (assemble_stage_Rdist,
    mloop_fd.f90:241-274) against the same scheme on the collocation method,
    whose assembly is pinned separately, after 10 steps of the same dt: the
    time errors agree to leading order, so T, u_s and u_phi differences
    shrink by about 4 per halving of the grid spacing."""
    from numpy.polynomial import chebyshev as C
    from pizza_amd import workloads as W

    def run(cfg):
        o = Oracle(cfg)
        W.seed_run(o, cfg, amp_t=1e-3)
        o.start()
        o.step(10)
        return o

    oc = run(_fd_cfg(49, False, time_scheme=scheme))
    rc, xc = oc.grid()
    V = C.chebvander(xc, len(xc) - 1)
    errs = {}
    for n_r in (65, 129):
        of = run(_fd_cfg(n_r, True, time_scheme=scheme))
        rf, _ = of.grid()
        xq = (2 * rf - rc[0] - rc[-1]) / (rc[0] - rc[-1])
        for f in ("temp", "us", "up"):
            coef = np.linalg.solve(V, oc.get_field(f).T).T
            ci = np.array([C.chebval(xq, c) for c in coef])
            errs[(n_r, f)] = (np.abs(of.get_field(f) - ci) / np.max(np.abs(ci))).max()
    for f in ("temp", "us", "up"):
        assert errs[(65, f)] < 1e-2, (f, errs[(65, f)])
        assert errs[(65, f)] / errs[(129, f)] > 3.0, (f, errs[(65, f)], errs[(129, f)])


def test_fd_stencils_exact_on_polynomials():
    """Fornberg weights on the stretched grid (finite_differences.f90:501-546):
    the 3-point first and second derivatives are exact for quadratics, the
    5-point third and fourth derivatives for quartics, the one-sided wall
    derivatives for quadratics; the wall rows are the symmetric formulas
    with ghost points at the mirrored spacing."""
    o = Oracle(_fd_cfg(65, True))
    r, _ = o.grid()
    N = len(r)
    dr, ddr, d3, d4, top, bot = o.fd_stencils()
    for p_ in range(3):
        for k in range(1, N - 1):
            f = (r[k - 1:k + 2] - r[k]) ** p_
            assert abs(dr[k] @ f - (1.0 if p_ == 1 else 0.0)) < 1e-9
            assert abs(ddr[k] @ f - (2.0 if p_ == 2 else 0.0)) < 1e-6
        assert abs(top @ (r[:3] - r[0]) ** p_ - (1.0 if p_ == 1 else 0.0)) < 1e-9
        assert abs(bot @ (r[[N - 1, N - 2, N - 3]] - r[-1]) ** p_ - (1.0 if p_ == 1 else 0.0)) < 1e-9
    for p_ in range(5):
        for k in range(2, N - 2):
            f = (r[k - 2:k + 3] - r[k]) ** p_
            assert abs(d3[k] @ f - (6.0 if p_ == 3 else 0.0)) < 1e-3 * max(1.0, np.abs(d3[k]).max() * 1e-9)
            assert abs(d4[k] @ f - (24.0 if p_ == 4 else 0.0)) < 1e-3 * max(1.0, np.abs(d4[k]).max() * 1e-9)
    h0 = r[1] - r[0]
    assert np.allclose(dr[0], [-0.5 / h0, 0.0, 0.5 / h0]) and np.allclose(ddr[0], np.array([1, -2, 1]) / h0**2)
    assert np.all(d3[0] == 0) and np.all(d4[-1] == 0)


@pytest.mark.parametrize("method", ["integ", "coll", "fd"])
def test_nusselt_closed_form(method):
    """NuTop / NuBot (outputs.f90:729-755) of a conducting state plus the
    axisymmetric perturbation eps sin(pi (r - r_icb)): 1 - eps pi / dTc/dr
    at the outer wall and 1 + eps pi / dTc/dr at the inner one (exactly 1
    without the perturbation)."""
    from pizza_amd.config import default_config
    n_r = 65 if method == "fd" else 33
    cfg = default_config(n_r_max=n_r, n_cheb_max=n_r, m_max=8, time_scheme="CNAB2", ra=1e5, ek=1e-3,
                         radratio=0.35, cheb_method="colloc" if method == "coll" else "integ",
                         radial_scheme="FD" if method == "fd" else "CHEB")
    o = Oracle(cfg)
    r, _ = o.grid()
    o.start()
    assert np.array_equal(o.nusselt(), [1.0, 1.0])
    eps = 1e-3
    T = np.zeros((o.n_m, o.n_r), complex)
    T[0] = eps * np.sin(np.pi * (r - r[-1]))
    o2 = Oracle(cfg)
    o2.set_field("temp", T)
    o2.start()
    # conducting state (radial.f90:284-397): dTc/dr = -r_icb r_cmb / r^2 / ln(r_cmb / r_icb) / r
    rc, ri = r[0], r[-1]
    dtc = np.array([-1.0 / (np.log(rc / ri) * rc), -1.0 / (np.log(rc / ri) * ri)])
    expect = 1.0 + np.array([-eps * np.pi, eps * np.pi]) / dtc
    tol = 3e-3 if method == "fd" else 1e-10
    assert np.allclose(o2.nusselt(), expect, rtol=tol), (o2.nusselt(), expect)

// Device code of libpizza_amd, part 4: the finite-difference radial method
// (radial_scheme = 'FD'): second-order stencils with ghost zones,
// tridiagonal T / xi / u_phi(m = 0) systems and a pentadiagonal psi system
// per m (update_temp_fd.f90, update_xi_fd.f90, update_psi_fd.f90,
// parallel_solver.f90, mloop_fd.f90).  Included once by pz_api.hip.
#pragma once
#include "pz_coll.h"

// Stencils of the finite-difference scheme (finite_differences.f90:190-446):
// dr / ddr on points k-1, k, k+1; dddr / ddddr on k-2 .. k+2; one-sided
// first derivatives at the two walls.
struct FdView {
    const double *c;
    int N;
    __device__ FdView(const double *fdc, int n) : c(fdc), N(n) {}
    __device__ double dr(int k, int od) const { return c[3 * k + od]; }
    __device__ double ddr(int k, int od) const { return c[3 * N + 3 * k + od]; }
    __device__ double d3(int k, int od) const { return c[6 * N + 5 * k + od]; }
    __device__ double d4(int k, int od) const { return c[11 * N + 5 * k + od]; }
    __device__ double top(int od) const { return c[16 * N + od]; }
    __device__ double bot(int od) const { return c[16 * N + 3 + od]; }
};

__device__ __forceinline__ cd cadd3(cd a, cd b, cd c) { return cadd(cadd(a, b), c); }

// get_dr_Rloc (radial_derivatives.f90:669-760) at point k of a column f
// (zero beyond the walls in the bulk formula, one-sided at the walls)
__device__ __forceinline__ cd fd_dr_at(const FdView &S, const cd *f, int k, int N)
{
    if (k == 0) return cadd3(cscale(S.top(0), f[0]), cscale(S.top(1), f[1]), cscale(S.top(2), f[2]));
    if (k == N - 1) return cadd3(cscale(S.bot(0), f[N - 1]), cscale(S.bot(1), f[N - 2]), cscale(S.bot(2), f[N - 3]));
    return cadd3(cscale(S.dr(k, 0), f[k - 1]), cscale(S.dr(k, 1), f[k]), cscale(S.dr(k, 2), f[k + 1]));
}

// Forward and back substitution of the factorised systems (solver_up_3 /
// solver_dn_3 and solver_up_5 / solver_dn_5, parallel_solver.f90:367-593)
// by one whole wave (every lane calls them): lane l owns rows [l C, l C + C),
// C = ceil(N / 64).  Each sweep is a linear recurrence of first
// (tridiagonal) or second (pentadiagonal) order; every lane reduces its rows
// to one affine map of the state entering them, an inclusive scan over the
// lanes composes the maps (lower lanes first going down the column, higher
// lanes first coming back), and each lane re-runs its rows from the state
// that the scan hands it.  The factors and the row arithmetic are the
// reference's; only the order in which the chunks' contributions are
// summed differs from its serial loops.
__device__ __forceinline__ double lane_up(double v, int s) { return __shfl_up(v, s, 64); }
__device__ __forceinline__ double lane_dn(double v, int s) { return __shfl_down(v, s, 64); }
__device__ __forceinline__ cd lane_up(cd v, int s) { return mk(lane_up(v.x, s), lane_up(v.y, s)); }
__device__ __forceinline__ cd lane_dn(cd v, int s) { return mk(lane_dn(v.x, s), lane_dn(v.y, s)); }
__device__ __forceinline__ cd cfma(cd a, cd b, cd c) { return cadd(cmul(a, b), c); }
__device__ __forceinline__ cd cneg(cd a) { return mk(-a.x, -a.y); }

// y -> a y + g, composed over lanes: UP takes lanes below first
template <bool UP> __device__ __forceinline__ void lane_scan1(double &a, cd &g)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const double ap = UP ? lane_up(a, s) : lane_dn(a, s);
        const cd gp = UP ? lane_up(g, s) : lane_dn(g, s);
        if (UP ? lane >= s : lane + s < 64) {
            g = cadd(cscale(a, gp), g);
            a *= ap;
        }
    }
}

// (p, q) -> M (p, q) + g with M = [[m00, m01], [m10, m11]]
struct LaneAff2 {
    cd m00, m01, m10, m11, g0, g1;
};
__device__ __forceinline__ LaneAff2 lane_shift(const LaneAff2 &t, int s, bool up)
{
    LaneAff2 o;
    o.m00 = up ? lane_up(t.m00, s) : lane_dn(t.m00, s);
    o.m01 = up ? lane_up(t.m01, s) : lane_dn(t.m01, s);
    o.m10 = up ? lane_up(t.m10, s) : lane_dn(t.m10, s);
    o.m11 = up ? lane_up(t.m11, s) : lane_dn(t.m11, s);
    o.g0 = up ? lane_up(t.g0, s) : lane_dn(t.g0, s);
    o.g1 = up ? lane_up(t.g1, s) : lane_dn(t.g1, s);
    return o;
}
template <bool UP> __device__ __forceinline__ void lane_scan2(LaneAff2 &t)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const LaneAff2 o = lane_shift(t, s, UP);
        if (UP ? lane >= s : lane + s < 64) {
            // t after o: M = Mt Mo, g = Mt go + gt
            const cd g0 = cfma(t.m00, o.g0, cfma(t.m01, o.g1, t.g0));
            const cd g1 = cfma(t.m10, o.g0, cfma(t.m11, o.g1, t.g1));
            const cd n00 = cfma(t.m00, o.m00, cmul(t.m01, o.m10)), n01 = cfma(t.m00, o.m01, cmul(t.m01, o.m11));
            const cd n10 = cfma(t.m10, o.m00, cmul(t.m11, o.m10)), n11 = cfma(t.m10, o.m01, cmul(t.m11, o.m11));
            t.m00 = n00; t.m01 = n01; t.m10 = n10; t.m11 = n11; t.g0 = g0; t.g1 = g1;
        }
    }
}

// tridiagonal: x[k + 1] for k in [0, N), ghosts x[0] and x[N + 1]; d holds
// 1/pivot
__device__ __forceinline__ void fd_tri_solve_wave(int N, const double *__restrict__ d, const double *__restrict__ lo,
                                                  const double *__restrict__ up, cd *x)
{
    const int lane = threadIdx.x & 63, C = (N + 63) >> 6;
    const int k0 = min(N, lane * C), k1 = min(N, k0 + C);
    // forward: y_k = x[k + 1] - c_k y_{k - 1}, c_0 = lo_0, c_k = d_{k-1} lo_k, y_{-1} = x[0]
    double a = 1.0;
    cd g = mk(0.0, 0.0);
    for (int k = k0; k < k1; k++) {
        const double c = k == 0 ? lo[0] : d[k - 1] * lo[k];
        g = csub(x[k + 1], cscale(c, g));
        a = -c * a;
    }
    lane_scan1<true>(a, g);
    const cd x0 = x[0];
    cd y = lane_up(cadd(cscale(a, x0), g), 1);
    if (lane == 0) y = x0;
    for (int k = k0; k < k1; k++) {
        const double c = k == 0 ? lo[0] : d[k - 1] * lo[k];
        y = csub(x[k + 1], cscale(c, y));
        x[k + 1] = y;
    }
    // backward: z_k = d_k (y_k - up_k z_{k + 1}), z_N = x[N + 1]
    a = 1.0;
    g = mk(0.0, 0.0);
    for (int k = k1 - 1; k >= k0; k--) {
        g = cscale(d[k], csub(x[k + 1], cscale(up[k], g)));
        a = -(d[k] * up[k]) * a;
    }
    lane_scan1<false>(a, g);
    const cd xn = x[N + 1];
    cd zv = lane_dn(cadd(cscale(a, xn), g), 1);
    if (lane == 63) zv = xn;
    for (int k = k1 - 1; k >= k0; k--) {
        zv = cscale(d[k], csub(x[k + 1], cscale(up[k], zv)));
        x[k + 1] = zv;
    }
}

// one sweep v_k = b_k - c1_k p - c2_k q of the pentadiagonal solve, the
// state (p, q) entering row k being (v_{k-1}, v_{k-2}) going down (FWD) or
// (v_{k+1}, v_{k+2}) coming back; b_k = d_k x[k + 2] going down, x[k + 2]
// coming back
template <bool FWD>
__device__ __forceinline__ void fd_penta_sweep_wave(int N, const cd *__restrict__ dg, const cd *__restrict__ c1,
                                                    const cd *__restrict__ c2, cd *x, cd p_in, cd q_in)
{
    const int lane = threadIdx.x & 63, C = (N + 63) >> 6;
    const int k0 = min(N, lane * C), k1 = min(N, k0 + C);
    const int n = k1 - k0;
    const cd z = mk(0.0, 0.0), one = mk(1.0, 0.0);
    LaneAff2 t{one, z, z, one, z, z};
    for (int i = 0; i < n; i++) {
        const int k = FWD ? k0 + i : k1 - 1 - i;
        const cd a1 = c1[k], a2 = c2[k];
        const cd b = FWD ? cmul(dg[k], x[k + 2]) : x[k + 2];
        const cd n0 = csub(csub(b, cmul(a1, t.g0)), cmul(a2, t.g1));
        t.g1 = t.g0;
        t.g0 = n0;
        const cd r00 = cneg(cfma(a1, t.m00, cmul(a2, t.m10))), r01 = cneg(cfma(a1, t.m01, cmul(a2, t.m11)));
        t.m10 = t.m00;
        t.m11 = t.m01;
        t.m00 = r00;
        t.m01 = r01;
    }
    lane_scan2<FWD>(t);
    cd p = cfma(t.m00, p_in, cfma(t.m01, q_in, t.g0)), q = cfma(t.m10, p_in, cfma(t.m11, q_in, t.g1));
    p = FWD ? lane_up(p, 1) : lane_dn(p, 1);
    q = FWD ? lane_up(q, 1) : lane_dn(q, 1);
    if (lane == (FWD ? 0 : 63)) { p = p_in; q = q_in; }
    for (int i = 0; i < n; i++) {
        const int k = FWD ? k0 + i : k1 - 1 - i;
        const cd b = FWD ? cmul(dg[k], x[k + 2]) : x[k + 2];
        const cd v = csub(csub(b, cmul(c1[k], p)), cmul(c2[k], q));
        x[k + 2] = v;
        q = p;
        p = v;
    }
}

// pentadiagonal: x[k + 2], ghosts two deep on each side
__device__ __forceinline__ void fd_penta_solve_wave(int N, const cd *d, const cd *l1, const cd *l2, const cd *u1,
                                                    const cd *u2, cd *x)
{
    fd_penta_sweep_wave<true>(N, d, l1, l2, x, x[1], x[0]);
    fd_penta_sweep_wave<false>(N, d, u1, u2, x, x[N + 2], x[N + 3]);
}

// prepare_mat_3 (parallel_solver.f90:135-187), one thread
__device__ void fd_lu3(int N, double *d, double *lo, double *up)
{
    lo[0] = 0.0;
    up[N - 1] = 0.0;
    double piv = 1.0 / d[0];
    d[0] = piv;
    for (int k = 1; k < N; k++) {
        piv = 1.0 / (d[k] - lo[k] * up[k - 1] * piv);
        d[k] = piv;
    }
}

// prepare_mat_5 (parallel_solver.f90:189-288), one thread
__device__ void fd_lu5(int N, cd *d, cd *l1, cd *l2, cd *u1, cd *u2)
{
    const cd z = mk(0.0, 0.0);
    l1[0] = z; l2[0] = z; l2[1] = z;
    u1[N - 1] = z; u2[N - 1] = z; u2[N - 2] = z;
    u1[1] = csub(u1[1], cdiv(cmul(l1[1], u2[0]), d[0]));
    d[1] = csub(d[1], cdiv(cmul(l1[1], u1[0]), d[0]));
    for (int k = 2; k < N; k++) {
        l1[k] = csub(l1[k], cdiv(cmul(l2[k], u1[k - 2]), d[k - 2]));
        u1[k] = csub(u1[k], cdiv(cmul(l1[k], u2[k - 1]), d[k - 1]));
        d[k] = csub(csub(d[k], cdiv(cmul(l1[k], u1[k - 1]), d[k - 1])), cdiv(cmul(l2[k], u2[k - 2]), d[k - 2]));
    }
    for (int k = 0; k < N; k++) {
        const cd dk = cdiv(mk(1.0, 0.0), d[k]);
        d[k] = dk;
        u1[k] = cmul(u1[k], dk); u2[k] = cmul(u2[k], dk);
        l1[k] = cmul(l1[k], dk); l2[k] = cmul(l2[k], dk);
    }
}

// LDS layout of the start, rebuild and assembly kernels of the
// finite-difference method (16-byte blocks): ghosted columns gT, gX
// (N + 2), gP (N + 4; u_phi of m = 0 at offset 1), six N-vectors, then the
// factorised systems of this m.
struct FdLayout {
    size_t gT, gX, gP, vec, mT, mX, mP, total;
    __host__ __device__ FdLayout(int N)
    {
        size_t o = 0;
        auto take = [&](size_t bytes) { size_t r = o; o += (bytes + 15) & ~(size_t)15; return r; };
        gT = take(sizeof(cd) * (N + 2));
        gX = take(sizeof(cd) * (N + 2));
        gP = take(sizeof(cd) * (N + 4));
        vec = take(sizeof(cd) * 6 * (size_t)N);
        mT = take(sizeof(double) * 3 * (size_t)N);
        mX = take(sizeof(double) * 3 * (size_t)N);
        mP = take(sizeof(cd) * 5 * (size_t)N);
        total = o;
    }
};

// fill_ghosts_temp / fill_ghosts_xi (update_temp_fd.f90:117-169): fixed
// value 2 f_0 - f_1; fixed flux, the mirror point f_1 plus 2 (r_0 - r_1) qt
// at the outer wall and minus 2 (r_{N-2} - r_{N-1}) qb at the inner one for
// a flux pattern (zero unless l_fd_bc_pattern: the reference comments these
// terms out at :142,162)
__device__ __forceinline__ void fd_ghost_scalar(cd *g, int N, int ktop, int kbot, const RadView &R, cd qt, cd qb)
{
    const double dt = R.r(0) - R.r(1), db = R.r(N - 2) - R.r(N - 1);
    g[0] = ktop == 1 ? csub(cscale(2.0, g[1]), g[2]) : cadd(g[2], cscale(2.0 * dt, qt));
    g[N + 1] = kbot == 1 ? csub(cscale(2.0, g[N]), g[N - 1]) : csub(g[N - 1], cscale(2.0 * db, qb));
}

// the wall pattern of mode mi when the finite-difference walls take them
__device__ __forceinline__ cd fd_wall(const PzDev &p, const cd *pat, int mi)
{
    return p.fd_pattern ? pat[mi] : mk(0.0, 0.0);
}

// a wall row of the finite-difference T / xi right-hand side (prepare_*_fd):
// the pattern value on a fixed-value wall (zero without patterns); on a
// fixed-flux wall the ghost point's flux term, moved to the right-hand side
// times minus the wall row's ghost coefficient (-w kappa (d2 + d/r) of the
// ghost)
__device__ __forceinline__ cd fd_wall_row(const PzDev &p, const RadView &R, const FdView &S, int k, int N, int ktop,
                                          int kbot, double diff, cd qt, cd qb, cd v)
{
    const double w = p.ss->wimp_lin[0];
    if (k == 0) {
        if (ktop == 1) return qt;
        if (p.fd_pattern) {
            const double lo = -w * diff * (S.ddr(0, 0) + R.or1(0) * S.dr(0, 0));
            v = csub(v, cscale(lo * 2.0 * (R.r(0) - R.r(1)), qt));
        }
    }
    if (k == N - 1) {
        if (kbot == 1) return qb;
        if (p.fd_pattern) {
            const double up = -w * diff * (S.ddr(N - 1, 2) + R.or1(N - 1) * S.dr(N - 1, 2));
            v = cadd(v, cscale(up * 2.0 * (R.r(N - 2) - R.r(N - 1)), qb));
        }
    }
    return v;
}

// fill_ghosts_psi (update_psi_fd.f90:140-215): u_phi(m = 0) at offset 1 or
// psi (m /= 0) at offset 2
__device__ void fd_ghost_psi(const PzDev &p, const RadView &R, int m, cd *g)
{
    const int N = p.N;
    const double dr = R.r(1) - R.r(0), db = R.r(N - 1) - R.r(N - 2);
    if (m == 0) {
        g[0] = p.ktopv == 1 ? csub(g[2], cscale(2.0 * dr * R.or1(0), g[1])) : csub(cscale(2.0, g[1]), g[2]);
        g[N + 1] = p.kbotv == 1 ? cadd(g[N - 1], cscale(2.0 * db * R.or1(N - 1), g[N]))
                                : csub(cscale(2.0, g[N]), g[N - 1]);
        return;
    }
    g[1] = p.ktopv == 1 ? cscale(-(1.0 - 0.5 * R.or1(0) * dr) / (1.0 + 0.5 * R.or1(0) * dr), g[3]) : g[3];
    g[0] = mk(0.0, 0.0);
    g[N + 2] = p.kbotv == 1 ? cscale(-(1.0 + 0.5 * R.or1(N - 1) * db) / (1.0 - 0.5 * R.or1(N - 1) * db), g[N])
                            : g[N];
    g[N + 3] = mk(0.0, 0.0);
}

// Implicit terms, old slots and the fields of one column from the ghosted
// solutions: get_temp_rhs_imp_ghost / get_xi_rhs_imp_ghost
// (update_temp_fd.f90:217-276) and get_psi_rhs_imp_ghost
// (update_psi_fd.f90:247-395).  up_lds (optional) receives u_phi.
__device__ void fd_implicit(const PzDev &p, const RadView &R, const FdView &S, int mi, int m, const cd *gT,
                            const cd *gX, const cd *gP, int inext, int lcalc, cd *up_lds)
{
    const int N = p.N, tid = threadIdx.x;
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N, si = (size_t)(inext - 1) * F;
    const double dm = (double)m, dm2 = dm * dm, nu = p.ViscFac, cf = p.CorFac;
    for (int k = tid; k < N; k += blockDim.x) {
        const size_t q = off + k;
        const double o1 = R.or1(k), o2 = R.or2(k);
        const double a0 = S.dr(k, 0), a1 = S.dr(k, 1), a2 = S.dr(k, 2);
        const double b0 = S.ddr(k, 0), b1 = S.ddr(k, 1), b2 = S.ddr(k, 2);
        if (p.l_heat) {
            const cd f0 = gT[k], f1 = gT[k + 1], f2 = gT[k + 2];
            const cd df = cadd3(cscale(a0, f0), cscale(a1, f1), cscale(a2, f2));
            const cd ddf = cadd3(cscale(b0, f0), cscale(b1, f1), cscale(b2, f2));
            p.temp[q] = f1;
            p.dtemp[q] = df;
            if (inext == 1) p.T_old[q] = f1;
            if (lcalc) p.T_impl[si + q] = cscale(p.TdiffFac, csub(cadd(ddf, cscale(o1, df)), cscale(dm2 * o2, f1)));
        }
        if (p.l_chem) {
            const cd f0 = gX[k], f1 = gX[k + 1], f2 = gX[k + 2];
            const cd df = cadd3(cscale(a0, f0), cscale(a1, f1), cscale(a2, f2));
            const cd ddf = cadd3(cscale(b0, f0), cscale(b1, f1), cscale(b2, f2));
            p.xi[q] = f1;
            if (inext == 1) p.X_old[q] = f1;
            if (lcalc) p.X_impl[si + q] = cscale(p.XidiffFac, csub(cadd(ddf, cscale(o1, df)), cscale(dm2 * o2, f1)));
        }
        if (m == 0) {
            const cd u0 = gP[k], u1 = gP[k + 1], u2 = gP[k + 2];
            const cd du = cadd3(cscale(a0, u0), cscale(a1, u1), cscale(a2, u2));
            const cd ddu = cadd3(cscale(b0, u0), cscale(b1, u1), cscale(b2, u2));
            if (inext == 1) p.P_old[q] = u1;
            if (lcalc)
                p.P_impl[si + q] = csub(cscale(nu, csub(cadd(ddu, cscale(o1, du)), cscale(o2, u1))),
                                        cscale(cf * R.ekpump(k), u1));
            p.us[q] = mk(0.0, 0.0);
            p.up[q] = u1;
            p.om[q] = cadd(du, cscale(o1, u1));
            p.psi[q] = mk(0.0, 0.0);
            if (up_lds) up_lds[k] = u1;
            continue;
        }
        const cd f0 = gP[k], f1 = gP[k + 1], f2 = gP[k + 2], f3 = gP[k + 3], f4 = gP[k + 4];
        const cd dp = cadd3(cscale(a0, f1), cscale(a1, f2), cscale(a2, f3));
        const cd ddp = cadd3(cscale(b0, f1), cscale(b1, f2), cscale(b2, f3));
        const cd ps = f2;
        const double be = R.beta(k), db = R.dbeta(k);
        if (inext == 1)
            p.P_old[q] = mk(-(ddp.x + (o1 + be) * dp.x + (o1 * be + db - dm2 * o2) * ps.x),
                            -(ddp.y + (o1 + be) * dp.y + (o1 * be + db - dm2 * o2) * ps.y));
        if (lcalc) {
            const cd dddp = cadd(cadd3(cscale(S.d3(k, 0), f0), cscale(S.d3(k, 1), f1), cscale(S.d3(k, 2), f2)),
                                 cadd(cscale(S.d3(k, 3), f3), cscale(S.d3(k, 4), f4)));
            const cd d4p = cadd(cadd3(cscale(S.d4(k, 0), f0), cscale(S.d4(k, 1), f1), cscale(S.d4(k, 2), f2)),
                                cadd(cscale(S.d4(k, 3), f3), cscale(S.d4(k, 4), f4)));
            const double o3 = R.or3(k), d2b = R.d2beta(k), d3b = R.d3beta(k);
            const double c3 = be + 2.0 * o1, c2 = 3.0 * db + 2.0 * be * o1 - o2 - 2.0 * dm2 * o2;
            const double c1 = 3.0 * d2b + 4.0 * db * o1 - be * o2 + o3 + dm2 * o2 * (2.0 * o1 - be);
            const double c0 = d3b + 2.0 * d2b * o1 - db * o2 + be * o3 - dm2 * o2 * (4.0 * o2 + be * o1 + db - dm2 * o2);
            cd v = cscale(-nu, cadd(cadd3(d4p, cscale(c3, dddp), cscale(c2, ddp)), cadd(cscale(c1, dp), cscale(c0, ps))));
            if (p.l_coriolis_imp) v = cadd(v, cmul(mk(0.0, cf * be * dm * o1), ps));
            if (p.l_ek_pump) {
                const double eu = R.ekp_up(k);
                // (db + be o1 - m^2 o2 + be eu) - i m o1 (ekp_us + i m ekp_dusdp)
                const cd P0 = mk(db + be * o1 - dm2 * o2 + be * eu + dm2 * o1 * R.ekp_dusdp(k), -dm * o1 * R.ekp_us(k));
                v = cadd(v, cscale(cf * R.ekpump(k), cadd3(ddp, cscale(o1 + be + eu, dp), cmul(P0, ps))));
            }
            p.P_impl[si + q] = v;
        }
        const cd us = cmul(mk(0.0, dm * o1), ps);
        const cd om = mk(-(ddp.x + (o1 + be) * dp.x + (be * o1 + db - dm2 * o2) * ps.x),
                         -(ddp.y + (o1 + be) * dp.y + (be * o1 + db - dm2 * o2) * ps.y));
        const cd up = csub(mk(-dp.x, -dp.y), cscale(be, ps));
        p.us[q] = us;
        p.up[q] = up;
        p.om[q] = om;
        p.psi[q] = ps;
        if (up_lds) up_lds[k] = up;
    }
}

// grid (n_m): get_tMat_Rdist (update_temp_fd.f90:387-483) and its xi twin,
// get_psimat_Rdist (update_psi_fd.f90:608-770) for m /= 0 and
// get_uphimat_Rdist (:772-836) for m = 0, each factorised in LDS by one thread
__device__ __forceinline__ void fd_build_body(const PzDev &p, int mi, int force)
{
    extern __shared__ double smem[];
    if (!force && p.ss->lMat == 0) return;
    const int N = p.N, m = p.idx2m[mi], tid = threadIdx.x;
    const FdLayout lay(N);
    char *lb = (char *)smem;
    double *mT = (double *)(lb + lay.mT), *mX = (double *)(lb + lay.mX);
    cd *mP = (cd *)(lb + lay.mP);
    double *mU = (double *)mP;
    const RadView R(p.radc, N);
    const FdView S(p.fdc, N);
    const double w = p.ss->wimp_lin[0], dm = (double)m, dm2 = dm * dm, nu = p.ViscFac, cf = p.CorFac;
    for (int k = tid; k < N; k += blockDim.x) {
        const double a0 = S.dr(k, 0), a1 = S.dr(k, 1), a2 = S.dr(k, 2);
        const double b0 = S.ddr(k, 0), b1 = S.ddr(k, 1), b2 = S.ddr(k, 2);
        const double o1 = R.or1(k), o2 = R.or2(k);
        for (int f = 0; f < 2; f++) {
            const double df = f == 0 ? p.TdiffFac : p.XidiffFac;
            double *M = f == 0 ? mT : mX;
            M[k] = 1.0 - w * df * (b1 + o1 * a1 - dm2 * o2);
            M[N + k] = -w * df * (b0 + o1 * a0);
            M[2 * N + k] = -w * df * (b2 + o1 * a2);
        }
        if (m == 0) {
            mU[k] = 1.0 - w * (nu * (b1 + o1 * a1 - o2) - cf * R.ekpump(k));
            mU[N + k] = -w * nu * (b0 + o1 * a0);
            mU[2 * N + k] = -w * nu * (b2 + o1 * a2);
            continue;
        }
        const double be = R.beta(k), db = R.dbeta(k), o3 = R.or3(k), d2b = R.d2beta(k), d3b = R.d3beta(k);
        const double ek = R.ekpump(k), eu = R.ekp_up(k);
        const double A3 = 2.0 * o1 + be, A2 = 3.0 * db + 2.0 * be * o1 - o2 - 2.0 * dm2 * o2;
        const double A1 = 3.0 * d2b + 4.0 * db * o1 - be * o2 + o3 + dm2 * o2 * (2.0 * o1 - be);
        const double A0 = d3b + 2.0 * d2b * o1 - db * o2 + be * o3 - dm2 * o2 * (4.0 * o2 + be * o1 + db - dm2 * o2);
        // db + be o1 - m^2 o2 + be eu - i m o1 (ekp_us + i m ekp_dusdp)
        const double P0r = db + be * o1 - dm2 * o2 + be * eu + dm2 * o1 * R.ekp_dusdp(k), P0i = -dm * o1 * R.ekp_us(k);
        const double s1 = S.d3(k, 1), s2 = S.d3(k, 2), s3 = S.d3(k, 3);
        const double q1 = S.d4(k, 1), q2 = S.d4(k, 2), q3 = S.d4(k, 3);
        const double dre = -(b1 + (o1 + be) * a1 + (be * o1 + db - dm2 * o2)) +
                           w * (nu * (q2 + A3 * s2 + A2 * b1 + A1 * a1 + A0) - cf * ek * (b1 + (be + o1 + eu) * a1 + P0r));
        const double dim = w * (-cf * ek * P0i - cf * be * dm * o1);
        mP[k] = mk(dre, dim);
        mP[N + k] = mk(-(b0 + (o1 + be) * a0) + w * (nu * (q1 + A3 * s1 + A2 * b0 + A1 * a0) - cf * ek * (b0 + (be + o1 + eu) * a0)), 0.0);
        mP[2 * N + k] = mk(w * nu * (S.d4(k, 0) + A3 * S.d3(k, 0)), 0.0);
        mP[3 * N + k] = mk(-(b2 + (o1 + be) * a2) + w * (nu * (q3 + A3 * s3 + A2 * b2 + A1 * a2) - cf * ek * (b2 + (be + o1 + eu) * a2)), 0.0);
        mP[4 * N + k] = mk(w * nu * (S.d4(k, 4) + A3 * S.d3(k, 4)), 0.0);
    }
    __syncthreads();
    if (tid == 0 || tid == 64) {
        const int f = tid == 0 ? 0 : 1;
        double *M = f == 0 ? mT : mX;
        const int ktop = f == 0 ? p.ktopt : p.ktopxi, kbot = f == 0 ? p.kbott : p.kbotxi;
        if (ktop == 1) { M[0] = 1.0; M[2 * N] = 0.0; M[N] = 0.0; }
        else M[2 * N] = M[2 * N] + M[N];
        if (kbot == 1) { M[N - 1] = 1.0; M[3 * N - 1] = 0.0; M[2 * N - 1] = 0.0; }
        else M[2 * N - 1] = M[3 * N - 1] + M[2 * N - 1];
        fd_lu3(N, M, M + N, M + 2 * N);
    } else if (tid == 128 && m == 0) {
        double *d = mU, *lo = mU + N, *up = mU + 2 * N;
        if (p.ktopv == 1) {
            up[0] = up[0] + lo[0];
            d[0] = d[0] - 2.0 * (R.r(1) - R.r(0)) * R.or1(0) * lo[0];
        } else { d[0] = 1.0; lo[0] = 0.0; up[0] = 0.0; }
        if (p.kbotv == 1) {
            lo[N - 1] = lo[N - 1] + up[N - 1];
            d[N - 1] = d[N - 1] + 2.0 * (R.r(N - 1) - R.r(N - 2)) * R.or1(N - 1) * up[N - 1];
        } else { d[N - 1] = 1.0; lo[N - 1] = 0.0; up[N - 1] = 0.0; }
        fd_lu3(N, d, lo, up);
    } else if (tid == 128) {
        cd *d = mP, *l1 = mP + N, *l2 = mP + 2 * N, *u1 = mP + 3 * N, *u2 = mP + 4 * N;
        const cd one = mk(1.0, 0.0), z = mk(0.0, 0.0);
        d[0] = one; l1[0] = z; l2[0] = z; u1[0] = z; u2[0] = z;
        if (p.ktopv == 1) {
            const double dr = R.r(1) - R.r(0), h = 0.5 * R.or1(0) * dr;
            d[1] = csub(d[1], cscale((1.0 - h) / (1.0 + h), l2[1]));
            l1[1] = cadd(l1[1], cscale(2.0 / (1.0 + h), l2[1]));
        } else {
            d[1] = cadd(d[1], l2[1]);
        }
        d[N - 1] = one; l1[N - 1] = z; l2[N - 1] = z; u1[N - 1] = z; u2[N - 1] = z;
        if (p.kbotv == 1) {
            const double dr = R.r(N - 1) - R.r(N - 2), h = 0.5 * R.or1(N - 1) * dr;
            d[N - 2] = csub(d[N - 2], cscale((1.0 + h) / (1.0 - h), u2[N - 2]));
            u1[N - 2] = cadd(u1[N - 2], cscale(2.0 / (1.0 - h), u2[N - 2]));
        } else {
            d[N - 2] = cadd(d[N - 2], u2[N - 2]);
        }
        fd_lu5(N, d, l1, l2, u1, u2);
    }
    __syncthreads();
    double *gT = p.fT + (size_t)mi * 3 * N, *gX = p.fX + (size_t)mi * 3 * N;
    for (int t = tid; t < 3 * N; t += blockDim.x) {
        if (p.l_heat) gT[t] = mT[t];
        if (p.l_chem) gX[t] = mX[t];
        if (m == 0) p.fU[t] = mU[t];
    }
    if (m != 0) {
        cd *gP = p.fP + (size_t)mi * 5 * N;
        for (int t = tid; t < 5 * N; t += blockDim.x) gP[t] = mP[t];
    }
}

__global__ void __launch_bounds__(256) k_fd_build(const PzDev *__restrict__ pd, int force)
{
    fd_build_body(*pd, blockIdx.x, force);
}

// LDS of the m-loop kernel (16-byte blocks): the ghosted columns gT, gX, gP
// of the right-hand sides (then of the solutions) and the factorised
// systems of this m; small enough for three workgroups per CU at n_r 257.
struct FdLoopLayout {
    size_t gT, gX, gP, mT, mX, mP, total;
    __host__ __device__ FdLoopLayout(int N)
    {
        size_t o = 0;
        auto take = [&](size_t bytes) { size_t r = o; o += (bytes + 15) & ~(size_t)15; return r; };
        gT = take(sizeof(cd) * (N + 2));
        gX = take(sizeof(cd) * (N + 2));
        gP = take(sizeof(cd) * (N + 4));
        mT = take(sizeof(double) * 3 * (size_t)N);
        mX = take(sizeof(double) * 3 * (size_t)N);
        mP = take(sizeof(cd) * 5 * (size_t)N);
        total = o;
    }
};

// imex_split of a multistep scheme at one point, the stage's right-hand
// side (P + w e) + Q, and rotate_imex (multistep_schemes.f90:462-498) of
// the point: every slot moves up by one from the values the split read,
// slot 0 of the explicit terms being e.  Multistep schemes keep at most
// four slots of each kind.
__device__ __forceinline__ cd fd_rhs_multistep(const PzDev &p, const PzScheme *s, cd *old, cd *impl, cd *expl,
                                               size_t q, double w, cd e)
{
    const size_t F = (size_t)p.nm * p.N;
    const cd z = mk(0.0, 0.0);
    cd vo[4], vi[4], ve[4];
#pragma unroll
    for (int o = 0; o < 4; o++) {
        vo[o] = o < p.nold ? gld(old, o * F + q) : z;
        vi[o] = o < p.nimp ? gld(impl, o * F + q) : z;
        ve[o] = o >= 1 && o < p.nexp ? gld(expl, o * F + q) : z;
    }
    cd r = cscale(s->wimp[0], vo[0]);
#pragma unroll
    for (int o = 1; o < 4; o++)
        if (o < p.nold) r = cadd(r, cscale(s->wimp[o], vo[o]));
#pragma unroll
    for (int o = 0; o < 4; o++)
        if (o < p.nimp) r = cadd(r, cscale(s->wimp_lin[o + 1], vi[o]));
    cd Q = z;
#pragma unroll
    for (int o = 1; o < 4; o++)
        if (o < p.nexp) {
            const cd t = cscale(s->wexp[o], ve[o]);
            Q = o == 1 ? t : cadd(Q, t);
        }
#pragma unroll
    for (int o = 3; o >= 1; o--) {
        if (o < p.nexp) expl[o * F + q] = o == 1 ? e : ve[o - 1];
        if (o < p.nold) old[o * F + q] = vo[o - 1];
        if (o < p.nimp) impl[o * F + q] = vi[o - 1];
    }
    return cadd(cadd(r, cscale(w, e)), Q);
}

// The right-hand side of one field at one point: multistep (with the slot
// rotation) or DIRK (imex_split)
__device__ __forceinline__ cd fd_rhs_point(const PzDev &p, const PzScheme *s, int istage, cd *old, cd *impl, cd *expl,
                                           size_t off, int k, double w, cd e)
{
    if (p.family == 0) return fd_rhs_multistep(p, s, old, impl, expl, off + k, w, e);
    cd a, b;
    imex_split(p, s, istage, old, impl, expl, off, k, a, b);
    return cadd(cadd(a, cscale(w, e)), b);
}

// One stage of mloop_Rdist (mloop_fd.f90:148-239) for one m, with
// finish_explicit_assembly_Rdist (mloop_fd.f90:78-110) when exp_calc.
// Buoyancy is explicit with finite differences (namelists.f90:459-466).
// One pass over the points forms the explicit terms (the radial stencils
// read their neighbours from memory), the right-hand sides and the slot
// rotation; one wave per field then solves its system.
__device__ __forceinline__ void fd_mloop_body(const PzDev &p, int mi, int istage, int exp_calc)
{
    extern __shared__ double smem[];
    const int N = p.N, m = p.idx2m[mi], tid = threadIdx.x;
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    const int is = istage - 1;
    const PzScheme *ss = p.ss;
    const FdLoopLayout lay(N);
    char *lb = (char *)smem;
    cd *gT = (cd *)(lb + lay.gT), *gX = (cd *)(lb + lay.gX), *gP = (cd *)(lb + lay.gP);
    double *mT = (double *)(lb + lay.mT), *mX = (double *)(lb + lay.mX);
    cd *mP = (cd *)(lb + lay.mP);
    const bool hT = p.l_heat, hX = p.l_chem;
    const RadView R(p.radc, N);
    const FdView S(p.fdc, N);
    if (hT) lds_copy(mT, p.fT + off * 3, sizeof(double) * 3 * N);
    if (hX) lds_copy(mX, p.fX + off * 3, sizeof(double) * 3 * N);
    if (m == 0) lds_copy(mP, p.fU, sizeof(double) * 3 * N);
    else lds_copy(mP, p.fP + off * 5, sizeof(cd) * 5 * N);
    const double wcur = imex_weight(p, ss, istage);
    const cd z = mk(0.0, 0.0);
    // explicit terms, right-hand sides with their boundary rows (prepare_*_fd)
    for (int k = tid; k < N; k += blockDim.x) {
        const size_t q = off + k, qe = is * F + q;
        const double o1 = R.or1(k);
        const cd us = exp_calc ? gld(p.us, q) : z;
        if (hT) {
            cd e = gld(p.T_expl, qe);
            if (exp_calc) {
                e = csub(csub(e, cscale(o1, fd_dr_at(S, p.dVsT + off, k, N))),
                         cscale(R.dtcond(k) - p.tadvz * R.beta(k) * R.tcond(k), us));
                p.T_expl[qe] = e;
            }
            cd v = fd_rhs_point(p, ss, istage, p.T_old, p.T_impl, p.T_expl, off, k, wcur, e);
            if (k == 0 || k == N - 1)
                v = fd_wall_row(p, R, S, k, N, p.ktopt, p.kbott, p.TdiffFac, fd_wall(p, p.topt, mi),
                                fd_wall(p, p.bott, mi), v);
            gT[k + 1] = v;
        }
        if (hX) {
            cd e = gld(p.X_expl, qe);
            if (exp_calc) {
                e = csub(csub(e, cscale(o1, fd_dr_at(S, p.dVsXi + off, k, N))), cscale(R.dxicond(k), us));
                p.X_expl[qe] = e;
            }
            cd v = fd_rhs_point(p, ss, istage, p.X_old, p.X_impl, p.X_expl, off, k, wcur, e);
            if (k == 0 || k == N - 1)
                v = fd_wall_row(p, R, S, k, N, p.ktopxi, p.kbotxi, p.XidiffFac, fd_wall(p, p.topxi, mi),
                                fd_wall(p, p.botxi, mi), v);
            gX[k + 1] = v;
        }
        cd e = gld(p.P_expl, qe);
        if (exp_calc && m != 0) {
            e = csub(e, cscale(o1, fd_dr_at(S, p.dVsOm + off, k, N)));
            if (!p.l_coriolis_imp) e = cadd(e, cscale(p.CorFac * R.beta(k), us));
            if (hT) e = csub(e, cmul(mk(0.0, p.BuoFac * R.rgrav(k) * o1 * (double)m), gld(p.temp, q)));
            if (hX) e = csub(e, cmul(mk(0.0, p.ChemFac * R.rgrav(k) * o1 * (double)m), gld(p.xi, q)));
            p.P_expl[qe] = e;
        }
        cd v = fd_rhs_point(p, ss, istage, p.P_old, p.P_impl, p.P_expl, off, k, wcur, e);
        if (m == 0) {
            if ((k == 0 && p.ktopv != 1) || (k == N - 1 && p.kbotv != 1)) v = z;
            gP[k + 1] = v;
        } else {
            if (k == 0 || k == N - 1) v = z;
            gP[k + 2] = v;
        }
    }
    if (tid == 0) {
        gT[0] = z; gT[N + 1] = z; gX[0] = z; gX[N + 1] = z;
        if (m == 0) { gP[0] = z; gP[N + 1] = z; }
        else { gP[0] = z; gP[1] = z; gP[N + 2] = z; gP[N + 3] = z; }
    }
    __syncthreads();
    // the solves, one wave each (blocks of 256 threads)
    const int wave = tid >> 6;
    if (wave == 0 && hT) fd_tri_solve_wave(N, mT, mT + N, mT + 2 * N, gT);
    if (wave == 1 && hX) fd_tri_solve_wave(N, mX, mX + N, mX + 2 * N, gX);
    if (wave == 2) {
        if (m == 0) fd_tri_solve_wave(N, (const double *)mP, (const double *)mP + N, (const double *)mP + 2 * N, gP);
        else fd_penta_solve_wave(N, mP, mP + N, mP + 2 * N, mP + 3 * N, mP + 4 * N, gP);
    }
    __syncthreads();
    if (tid == 0) {
        if (hT) fd_ghost_scalar(gT, N, p.ktopt, p.kbott, R, fd_wall(p, p.topt, mi), fd_wall(p, p.bott, mi));
        if (hX) fd_ghost_scalar(gX, N, p.ktopxi, p.kbotxi, R, fd_wall(p, p.topxi, mi), fd_wall(p, p.botxi, mi));
        fd_ghost_psi(p, R, m, gP);
    }
    __syncthreads();
    const int inext = (istage == p.nstages) ? 1 : istage + 1;
    // implicit terms for the next stage; also before a CNAB2 start-up step
    const int lcalc = p.l_imp_calc_rhs[inext - 1] || (p.bridge && ss->n_steps + 1 <= p.nold - 1);
    fd_implicit(p, R, S, mi, m, gT, gX, gP, inext, lcalc, nullptr);
}

__global__ void __launch_bounds__(256) k_fd_mloop(const PzDev *__restrict__ pd, int istage, int exp_calc)
{
    fd_mloop_body(*pd, blockIdx.x, istage, exp_calc);
}

// get_start_fields, finite-difference branch (init_fields.f90:129-186):
// ghosted columns from the fields (psi rebuilt from u_s), the first
// implicit state, then omega from u_phi and u_s.
__global__ void __launch_bounds__(256) k_fd_start(const PzDev *__restrict__ pd)
{
    const PzDev &p = *pd;
    extern __shared__ double smem[];
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x;
    const size_t off = (size_t)mi * N;
    const FdLayout lay(N);
    char *lb = (char *)smem;
    cd *gT = (cd *)(lb + lay.gT), *gX = (cd *)(lb + lay.gX), *gP = (cd *)(lb + lay.gP);
    cd *V = (cd *)(lb + lay.vec);
    cd *usv = V, *upv = V + N;
    const RadView R(p.radc, N);
    const FdView S(p.fdc, N);
    const int i0 = m == 0;
    for (int k = tid; k < N; k += blockDim.x) {
        const size_t q = off + k;
        if (p.l_heat) gT[k + 1] = p.temp[q];
        if (p.l_chem) gX[k + 1] = p.xi[q];
        if (i0) gP[k + 1] = p.up[q];
        else gP[k + 2] = cmul(mk(0.0, -R.r(k) / (double)m), p.us[q]); // us / (i m) r
    }
    __syncthreads();
    if (tid == 0) {
        if (p.l_heat) fd_ghost_scalar(gT, N, p.ktopt, p.kbott, R, fd_wall(p, p.topt, mi), fd_wall(p, p.bott, mi));
        if (p.l_chem)
            fd_ghost_scalar(gX, N, p.ktopxi, p.kbotxi, R, fd_wall(p, p.topxi, mi), fd_wall(p, p.botxi, mi));
        fd_ghost_psi(p, R, m, gP);
    }
    __syncthreads();
    fd_implicit(p, R, S, mi, m, gT, gX, gP, 1, 1, upv);
    for (int k = tid; k < N; k += blockDim.x) usv[k] = p.us[off + k];
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x)
        p.om[off + k] = csub(cadd(fd_dr_at(S, upv, k, N), cscale(R.or1(k), upv[k])),
                             cmul(mk(0.0, R.or1(k) * (double)m), usv[k]));
}

// ---------------------------------------------------------------------------
// Diagnostics of the time series (outputs.f90:418-830) and the m-spectra
// (spectra.f90:104-197)
// ---------------------------------------------------------------------------
// The radial derivative of a finite-difference field as the step computes
// it (get_temp_rhs_imp_ghost, update_temp_fd.f90:217-276): the ghost point
// of each wall from its boundary condition (fill_ghosts_temp,
// update_temp_fd.f90:117-169: fixed value 2 f_0 - f_1, fixed flux f_1),
// then the centred three-point stencil at every point.  One thread per
// (m, r).
__global__ void k_fd_dr_ghost(const PzDev p __attribute__((grid_constant)), const cd *f, cd *df, int ktop, int kbot,
                              const cd *pt, const cd *pb)
{
    const int N = p.N;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)p.nm * N) return;
    const int k = (int)(q % N), mi = (int)(q / N);
    const cd *c = f + (q - k);
    const FdView S(p.fdc, N);
    const RadView R(p.radc, N);
    const cd gt = cscale(2.0 * (R.r(0) - R.r(1)), fd_wall(p, pt, mi));
    const cd gb = cscale(2.0 * (R.r(N - 2) - R.r(N - 1)), fd_wall(p, pb, mi));
    const cd fm = k > 0 ? c[k - 1] : (ktop == 1 ? csub(cscale(2.0, c[0]), c[1]) : cadd(c[1], gt));
    const cd fp = k < N - 1 ? c[k + 1] : (kbot == 1 ? csub(cscale(2.0, c[N - 1]), c[N - 2]) : csub(c[N - 2], gb));
    df[q] = cadd3(cscale(S.dr(k, 0), fm), cscale(S.dr(k, 1), c[k]), cscale(S.dr(k, 2), fp));
}

__device__ __forceinline__ double cc2r(cd a, int m) { return m == 0 ? a.x * a.x : 2.0 * (a.x * a.x + a.y * a.y); }
__device__ __forceinline__ double cc22r(cd a, cd b, int m)
{
    return m == 0 ? a.x * b.x : 2.0 * (a.x * b.x + a.y * b.y);
}

// Radial profiles of get_time_series (outputs.f90:456-502), one thread per
// radius: out[j * N + r] for j = us2, up2, up2_axi, enstrophy, buoyancy
// power, chemical power, volume Nusselt and Sherwood integrands (before
// the r h factor), radial heat flux.
#define PZ_NPROF 9
__global__ void k_series_profiles(const PzDev p __attribute__((grid_constant)), const cd *dT, const cd *dX, double *out)
{
    const int nr = blockIdx.x * blockDim.x + threadIdx.x;
    if (nr >= p.N) return;
    const int N = p.N;
    double v[PZ_NPROF] = {};
    const double o2 = p.or2[nr];
    for (int i = 0; i < p.nm; i++) {
        const int m = p.idx2m[i];
        const size_t q = (size_t)i * N + nr;
        const cd us = p.us[q];
        v[0] += cc2r(us, m);
        v[1] += cc2r(p.up[q], m);
        v[2] += m == 0 ? cc2r(p.up[q], m) : 0.0;
        v[3] += cc2r(p.om[q], m);
        if (p.l_heat) {
            const cd t = p.temp[q];
            v[4] += cc22r(us, t, m);
            v[6] += cc2r(dT[q], m) + (double)m * (double)m * o2 * cc2r(t, m);
            v[8] += cc22r(us, t, m);
            if (m == 0) v[8] += -p.TdiffFac * dT[q].x - p.TdiffFac * p.dtcond[nr];
        }
        if (p.l_chem) {
            const cd x = p.xi[q];
            v[5] += cc22r(us, x, m);
            v[7] += cc2r(dX[q], m) + (double)m * (double)m * o2 * cc2r(x, m);
        }
    }
    for (int j = 0; j < PZ_NPROF; j++) out[(size_t)j * N + nr] = v[j];
}

// Per (m, r) squares of calculate_spectra (spectra.f90:131-153) before the
// pi r h factor and the radial integral: out[(j * nm + i) * N + r] for
// j = us, up, omega, T, xi.
__global__ void k_spectra_mr(const PzDev p __attribute__((grid_constant)), double *out)
{
    const size_t F = (size_t)p.nm * p.N;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= F) return;
    const int m = p.idx2m[q / p.N];
    out[q] = cc2r(p.us[q], m);
    out[F + q] = cc2r(p.up[q], m);
    out[2 * F + q] = cc2r(p.om[q], m);
    out[3 * F + q] = p.l_heat ? cc2r(p.temp[q], m) : 0.0;
    out[4 * F + q] = p.l_chem ? cc2r(p.xi[q], m) : 0.0;
}

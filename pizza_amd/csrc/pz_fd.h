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

// solver_up_3 / solver_dn_3 (parallel_solver.f90:367-537), one thread:
// x[k + 1] for k in [0, N), ghosts x[0] and x[N + 1]; d holds 1/pivot
__device__ void fd_tri_solve(int N, const double *d, const double *lo, const double *up, cd *x)
{
    cd prev = csub(x[1], cscale(lo[0], x[0]));
    x[1] = prev;
    for (int k = 1; k < N; k++) {
        prev = csub(x[k + 1], cscale(d[k - 1] * lo[k], prev));
        x[k + 1] = prev;
    }
    cd next = x[N + 1];
    for (int k = N - 1; k >= 0; k--) {
        next = cscale(d[k], csub(x[k + 1], cscale(up[k], next)));
        x[k + 1] = next;
    }
}

// solver_up_5 / solver_dn_5 (parallel_solver.f90:426-593), one thread:
// x[k + 2], ghosts two deep on each side
__device__ void fd_penta_solve(int N, const cd *d, const cd *l1, const cd *l2, const cd *u1, const cd *u2, cd *x)
{
    cd xm2 = x[0], xm1 = x[1];
    for (int k = 0; k < N; k++) {
        const cd v = csub(csub(cmul(d[k], x[k + 2]), cmul(l1[k], xm1)), cmul(l2[k], xm2));
        x[k + 2] = v;
        xm2 = xm1;
        xm1 = v;
    }
    cd xp1 = x[N + 2], xp2 = x[N + 3];
    for (int k = N - 1; k >= 0; k--) {
        const cd v = csub(csub(x[k + 2], cmul(u1[k], xp1)), cmul(u2[k], xp2));
        x[k + 2] = v;
        xp2 = xp1;
        xp1 = v;
    }
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

// LDS layout of the finite-difference kernels (16-byte blocks): ghosted
// columns gT, gX (N + 2), gP (N + 4; u_phi of m = 0 at offset 1), fifteen
// N-vectors, then the factorised systems of this m.
struct FdLayout {
    size_t gT, gX, gP, vec, mT, mX, mP, total;
    __host__ __device__ FdLayout(int N)
    {
        size_t o = 0;
        auto take = [&](size_t bytes) { size_t r = o; o += (bytes + 15) & ~(size_t)15; return r; };
        gT = take(sizeof(cd) * (N + 2));
        gX = take(sizeof(cd) * (N + 2));
        gP = take(sizeof(cd) * (N + 4));
        vec = take(sizeof(cd) * 15 * (size_t)N);
        mT = take(sizeof(double) * 3 * (size_t)N);
        mX = take(sizeof(double) * 3 * (size_t)N);
        mP = take(sizeof(cd) * 5 * (size_t)N);
        total = o;
    }
};

// fill_ghosts_temp / fill_ghosts_xi (update_temp_fd.f90:117-169)
__device__ __forceinline__ void fd_ghost_scalar(cd *g, int N, int ktop, int kbot)
{
    g[0] = ktop == 1 ? csub(cscale(2.0, g[1]), g[2]) : g[2];
    g[N + 1] = kbot == 1 ? csub(cscale(2.0, g[N]), g[N - 1]) : g[N - 1];
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

// One stage of mloop_Rdist (mloop_fd.f90:148-239) for one m, with
// finish_explicit_assembly_Rdist (mloop_fd.f90:78-110) when exp_calc.
// Buoyancy is explicit with finite differences (namelists.f90:459-466).
__device__ __forceinline__ void fd_mloop_body(const PzDev &p, int mi, int istage, int exp_calc)
{
    extern __shared__ double smem[];
    const int N = p.N, m = p.idx2m[mi], tid = threadIdx.x;
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    const int is = istage - 1;
    const PzScheme *ss = p.ss;
    const FdLayout lay(N);
    char *lb = (char *)smem;
    cd *gT = (cd *)(lb + lay.gT), *gX = (cd *)(lb + lay.gX), *gP = (cd *)(lb + lay.gP);
    cd *V = (cd *)(lb + lay.vec);
    cd *PT = V, *QT = V + N, *PX = V + 2 * N, *QX = V + 3 * N, *PP = V + 4 * N, *QP = V + 5 * N;
    cd *eT = V + 6 * N, *eX = V + 7 * N, *eP = V + 8 * N, *aT = V + 9 * N, *aX = V + 10 * N, *aP = V + 11 * N;
    cd *usv = V + 12 * N, *Tv = V + 13 * N, *Xv = V + 14 * N;
    double *mT = (double *)(lb + lay.mT), *mX = (double *)(lb + lay.mX);
    cd *mP = (cd *)(lb + lay.mP);
    const bool hT = p.l_heat, hX = p.l_chem;
    const RadView R(p.radc, N);
    const FdView S(p.fdc, N);
    if (hT) lds_copy(mT, p.fT + off * 3, sizeof(double) * 3 * N);
    if (hX) lds_copy(mX, p.fX + off * 3, sizeof(double) * 3 * N);
    if (m == 0) lds_copy(mP, p.fU, sizeof(double) * 3 * N);
    else lds_copy(mP, p.fP + off * 5, sizeof(cd) * 5 * N);
    for (int k = tid; k < N; k += blockDim.x) {
        const size_t q = off + k, qe = is * F + q;
        if (hT) eT[k] = gld(p.T_expl, qe);
        if (hX) eX[k] = gld(p.X_expl, qe);
        eP[k] = gld(p.P_expl, qe);
        if (exp_calc) {
            if (hT) { aT[k] = gld(p.dVsT, q); Tv[k] = gld(p.temp, q); }
            if (hX) { aX[k] = gld(p.dVsXi, q); Xv[k] = gld(p.xi, q); }
            aP[k] = gld(p.dVsOm, q);
            usv[k] = gld(p.us, q);
        }
        cd a, b;
        if (hT) { imex_split(p, ss, istage, p.T_old, p.T_impl, p.T_expl, off, k, a, b); PT[k] = a; QT[k] = b; }
        if (hX) { imex_split(p, ss, istage, p.X_old, p.X_impl, p.X_expl, off, k, a, b); PX[k] = a; QX[k] = b; }
        imex_split(p, ss, istage, p.P_old, p.P_impl, p.P_expl, off, k, a, b);
        PP[k] = a;
        QP[k] = b;
    }
    const double wcur = imex_weight(p, ss, istage);
    __syncthreads();
    if (exp_calc) {
        for (int k = tid; k < N; k += blockDim.x) {
            const double o1 = R.or1(k);
            const cd us = usv[k];
            if (hT) {
                const cd v = csub(csub(eT[k], cscale(o1, fd_dr_at(S, aT, k, N))),
                                  cscale(R.dtcond(k) - p.tadvz * R.beta(k) * R.tcond(k), us));
                eT[k] = v;
                p.T_expl[is * F + off + k] = v;
            }
            if (hX) {
                const cd v = csub(csub(eX[k], cscale(o1, fd_dr_at(S, aX, k, N))), cscale(R.dxicond(k), us));
                eX[k] = v;
                p.X_expl[is * F + off + k] = v;
            }
            if (m != 0) {
                cd v = csub(eP[k], cscale(o1, fd_dr_at(S, aP, k, N)));
                if (!p.l_coriolis_imp) v = cadd(v, cscale(p.CorFac * R.beta(k), us));
                if (hT) v = csub(v, cmul(mk(0.0, p.BuoFac * R.rgrav(k) * o1 * (double)m), Tv[k]));
                if (hX) v = csub(v, cmul(mk(0.0, p.ChemFac * R.rgrav(k) * o1 * (double)m), Xv[k]));
                eP[k] = v;
                p.P_expl[is * F + off + k] = v;
            }
        }
        __syncthreads();
    }
    // right-hand sides with their boundary rows (prepare_*_fd)
    const cd z = mk(0.0, 0.0);
    for (int k = tid; k < N; k += blockDim.x) {
        if (hT) {
            cd v = cadd(cadd(PT[k], cscale(wcur, eT[k])), QT[k]);
            if ((k == 0 && p.ktopt == 1) || (k == N - 1 && p.kbott == 1)) v = z;
            gT[k + 1] = v;
        }
        if (hX) {
            cd v = cadd(cadd(PX[k], cscale(wcur, eX[k])), QX[k]);
            if ((k == 0 && p.ktopxi == 1) || (k == N - 1 && p.kbotxi == 1)) v = z;
            gX[k + 1] = v;
        }
        cd v = cadd(cadd(PP[k], cscale(wcur, eP[k])), QP[k]);
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
    // the serial solves, one wave each
    if (tid == 0 && hT) fd_tri_solve(N, mT, mT + N, mT + 2 * N, gT);
    if (tid == 64 && hX) fd_tri_solve(N, mX, mX + N, mX + 2 * N, gX);
    if (tid == 128) {
        if (m == 0) fd_tri_solve(N, (const double *)mP, (const double *)mP + N, (const double *)mP + 2 * N, gP);
        else fd_penta_solve(N, mP, mP + N, mP + 2 * N, mP + 3 * N, mP + 4 * N, gP);
    }
    __syncthreads();
    if (tid == 0) {
        if (hT) fd_ghost_scalar(gT, N, p.ktopt, p.kbott);
        if (hX) fd_ghost_scalar(gX, N, p.ktopxi, p.kbotxi);
        fd_ghost_psi(p, R, m, gP);
    }
    if (hT) rotate_row(p, p.T_old, p.T_impl, p.T_expl, off);
    if (hX) rotate_row(p, p.X_old, p.X_impl, p.X_expl, off);
    rotate_row(p, p.P_old, p.P_impl, p.P_expl, off);
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
        if (p.l_heat) fd_ghost_scalar(gT, N, p.ktopt, p.kbott);
        if (p.l_chem) fd_ghost_scalar(gX, N, p.ktopxi, p.kbotxi);
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
__global__ void k_fd_dr_ghost(const PzDev p __attribute__((grid_constant)), const cd *f, cd *df, int ktop, int kbot)
{
    const int N = p.N;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)p.nm * N) return;
    const int k = (int)(q % N);
    const cd *c = f + (q - k);
    const FdView S(p.fdc, N);
    const cd fm = k > 0 ? c[k - 1] : (ktop == 1 ? csub(cscale(2.0, c[0]), c[1]) : c[1]);
    const cd fp = k < N - 1 ? c[k + 1] : (kbot == 1 ? csub(cscale(2.0, c[N - 1]), c[N - 2]) : c[N - 2]);
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

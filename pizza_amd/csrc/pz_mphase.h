// Device code of libpizza_amd, part 2: the per-m stage kernel (explicit
// assembly + implicit solves of the integration method), the start-up
// kernel, diagnostics and unit-operator kernels.  Included once by pz_api.hip.
#pragma once
#include "pz_device.h"

// set_imex_rhs for one m row (multistep_schemes.f90:366-413;
// dirk_schemes.f90:773-814), split around the current explicit slot: P sums
// (in the reference's order) every term before it, Q every term after it, and the
// right-hand side is (P + w cur) + Q with w = imex_weight().  Lets the
// global-memory terms be loaded and summed before cur is known.
__device__ __forceinline__ void imex_split(const PzDev &p, const PzScheme *s, int istage, const cd *old,
                                           const cd *impl, const cd *expl, size_t off, int k, cd &P, cd &Q)
{
    const size_t F = (size_t)p.nm * p.N;
    const int is = istage - 1;
    cd r, q = mk(0.0, 0.0);
    if (p.family == 0) {
        r = cscale(s->wimp[0], gld(old, off + k));
        for (int o = 1; o < p.nold; o++) r = cadd(r, cscale(s->wimp[o], gld(old, o * F + off + k)));
        for (int o = 0; o < p.nimp; o++) r = cadd(r, cscale(s->wimp_lin[o + 1], gld(impl, o * F + off + k)));
        for (int o = 0; o < is; o++) r = cadd(r, cscale(s->wexp[o], gld(expl, o * F + off + k)));
        for (int o = is + 1; o < p.nexp; o++) {
            const cd t = cscale(s->wexp[o], gld(expl, o * F + off + k));
            q = o == is + 1 ? t : cadd(q, t);
        }
    } else {
        r = gld(old, off + k);
        for (int st = 0; st < is; st++) r = cadd(r, cscale(s->bexp[istage][st], gld(expl, st * F + off + k)));
        for (int st = 0; st <= is; st++) {
            const cd t = cscale(s->bimp[istage][st], gld(impl, st * F + off + k));
            q = st == 0 ? t : cadd(q, t);
        }
    }
    P = r;
    Q = q;
}
__device__ __forceinline__ double imex_weight(const PzDev &p, const PzScheme *s, int istage)
{
    return p.family == 0 ? s->wexp[istage - 1] : s->bexp[istage][istage - 1];
}

// rotate_imex for one m row (multistep_schemes.f90:462-498); DIRK: no-op
__device__ __forceinline__ void rotate_row(const PzDev &p, cd *old, cd *impl, cd *expl, size_t off)
{
    if (p.family != 0) return;
    const size_t F = (size_t)p.nm * p.N;
    for (int k = threadIdx.x; k < p.N; k += blockDim.x) {
        for (int o = p.nexp - 1; o >= 1; o--) expl[o * F + off + k] = expl[(o - 1) * F + off + k];
        for (int o = p.nold - 1; o >= 1; o--) old[o * F + off + k] = old[(o - 1) * F + off + k];
        for (int o = p.nimp - 1; o >= 1; o--) impl[o * F + off + k] = impl[(o - 1) * F + off + k];
    }
}

// -BuoFac rgrav/r i m T - ChemFac rgrav/r i m X (update_psi_integ_smat.f90:286-304)
__device__ __forceinline__ cd buoyancy(const PzDev &p, const RadView &R, int m, int k, cd T, cd X)
{
    cd b = mk(0, 0);
    if (m == 0) return b;
    if (p.l_heat) {
        double s = -p.BuoFac * R.rgrav(k) * R.or1(k);
        b = cmul(mk(0.0, s * (double)m), T);
        if (p.l_chem) b = csub(b, cmul(mk(0.0, p.ChemFac * R.rgrav(k) * R.or1(k) * (double)m), X));
    } else if (p.l_chem) {
        double s = -p.ChemFac * R.rgrav(k) * R.or1(k);
        b = cmul(mk(0.0, s * (double)m), X);
    }
    return b;
}

// get_dcheb_real_1d with the rerror threshold (radial_derivatives.f90:143-185),
// one thread: Chebyshev coefficients of d/dx from those of f
__device__ __forceinline__ void dcheb_real(const PzDev &p, const double *__restrict__ f, double *__restrict__ df)
{
    const int N = p.N;
    // four independent maxima so that the loads overlap
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    int j = 0;
    for (; j + 3 < N; j += 4) {
        m0 = fmax(m0, fabs(f[j]));
        m1 = fmax(m1, fabs(f[j + 1]));
        m2 = fmax(m2, fabs(f[j + 2]));
        m3 = fmax(m3, fabs(f[j + 3]));
    }
    for (; j < N; j++) m0 = fmax(m0, fabs(f[j]));
    const double thr = fmax(fmax(m0, m1), fmax(m2, m3)) * p.thr;
    // the recurrence df(n-1) = df(n+1) + 2 n f(n) carried in registers:
    // up1 = df(n), up2 = df(n+1)
    df[N - 1] = 0.0;
    const int nc = N - 1;
    double up1 = fabs(f[nc]) >= thr ? (double)nc * f[nc] : 0.0, up2 = 0.0;
    df[nc - 1] = up1;
#pragma unroll 4
    for (int n = N - 2; n >= 1; n--) {
        const double fn = f[n];
        const double v = fabs(fn) >= thr ? up2 + (double)(2 * n) * fn : 0.0;
        df[n - 1] = v;
        up2 = up1;
        up1 = v;
    }
}

#define MPHASE_NVEC 24
#define PZ_MPHASE_THREADS 512
// threads of the m-loop kernels for n_r: 256 leave room for more
// workgroups per CU on small grids (the coupled collocation system needs
// 2 n_r <= threads)
__host__ __device__ inline int mloop_threads(int N) { return N <= 65 ? 256 : PZ_MPHASE_THREADS; }

// LDS layout of k_mphase: working vectors, transform tables and the
// reduction scratch of the dense solves (16-byte aligned blocks).
struct MpLayout {
    size_t vec, dtmp, cosT, sinT, rad, red, total;
    // rad_lds = false: the radial rows and the transform tables stay in
    // global memory (read through L2), which frees PZ_NRADC n_r + 2 L
    // doubles of LDS per workgroup (collocation m-loop only)
    __host__ __device__ __forceinline__ MpLayout(int N, int threads, bool rad_lds = true)
    {
        size_t o = 0;
        auto take = [&](size_t bytes) { size_t r = o; o += (bytes + 15) & ~(size_t)15; return r; };
        const int L = 2 * N - 2;
        vec = take(sizeof(cd) * MPHASE_NVEC * N);
        dtmp = take(sizeof(double) * 2 * N);
        cosT = rad_lds ? take(sizeof(double) * L) : 0;
        sinT = rad_lds ? take(sizeof(double) * L) : 0;
        rad = rad_lds ? take(sizeof(double) * PZ_NRADC * N) : 0;
        // FFT scratch of the Chebyshev transforms (6 L + 12), shared with the
        // reduction scratch of the dense solves: T and xi together (real),
        // psi (complex, n_r rows on the integration method, 2 n_r on the
        // collocation method)
        size_t mv = matvec_scratch(N, N, threads, 2, 2);
        const size_t m1 = matvec_scratch(N, N, threads, 1, 1), m2 = matvec_scratch(2 * N, N, threads, 1, 1);
        if (m1 > mv) mv = m1;
        if (m2 > mv) mv = m2;
        red = take(sizeof(cd) * ((size_t)(6 * L + 16) > mv ? (size_t)(6 * L + 16) : mv));
        total = o;
    }
};


// One stage of the m-loop for one azimuthal wavenumber per workgroup:
// finish_explicit_assembly (mloop.f90:169-229) when exp_calc, then
// update_temp_int / update_xi_int (update_temp_integ.f90:179-298) and
// update_psi_int_smat (update_psi_integ_smat.f90:247-573).
__device__ __forceinline__ void mphase_body(const PzDev &p, int mi, int istage, int exp_calc)
{
    extern __shared__ double smem[];
    const int N = p.N, m = p.idx2m[mi];
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    const int is = istage - 1;
    const PzScheme *ss = p.ss;
    const int tid = threadIdx.x;
    const MpLayout lay(N, blockDim.x);
    char *lb = (char *)smem;
    cd *V = (cd *)(lb + lay.vec);
    cd *aT = V, *cT = V + N, *eT = V + 2 * N;
    cd *aX = V + 3 * N, *cX = V + 4 * N, *eX = V + 5 * N;
    cd *aP = V + 6 * N, *cP = V + 7 * N, *eP = V + 8 * N;
    cd *psiv = V + 9 * N, *omv = V + 10 * N, *upv = V + 11 * N, *usv = V + 12 * N;
    cd *Tv = V + 13 * N, *Xv = V + 14 * N, *bv = V + 15 * N, *b2 = V + 16 * N, *wv = V + 17 * N;
    cd *PT = V + 18 * N, *QT = V + 19 * N, *PX = V + 20 * N, *QX = V + 21 * N, *PP = V + 22 * N, *QP = V + 23 * N;
    double *dtmp = (double *)(lb + lay.dtmp);
    double *cosT = (double *)(lb + lay.cosT), *sinT = (double *)(lb + lay.sinT);
    const bool hT = p.l_heat, hX = p.l_chem;
    cd *red = (cd *)(lb + lay.red), *fbuf = red;
    const int ldi = p.ldi;
    // stage the transform tables and the radial functions into LDS (the
    // first barrier below orders them before their first use)
    lds_copy(cosT, p.cosT, sizeof(double) * p.L);
    lds_copy(sinT, p.sinT, sizeof(double) * p.L);
    lds_copy(lb + lay.rad, p.radc, sizeof(double) * PZ_NRADC * N);
    const RadView R((const double *)(lb + lay.rad), N);

    // ---------------- global inputs of this m in one phase ----------------
    // explicit-assembly inputs (or the current explicit slot) and the
    // right-hand-side terms that do not depend on this stage's explicit part
    // four groups of N threads: T, xi, psi's right-hand side, the velocity
    // fields (so that four times as many loads are in flight)
    for (int t = tid; t < 4 * N; t += blockDim.x) {
        const int g = t / N, k = t - g * N;
        const size_t q = off + k, qe = is * F + q;
        cd a, b;
        if (g == 0) {
            if (hT) {
                lds_store(eT, k, gld(p.T_expl, qe));
                if (exp_calc) lds_store(aT, k, gld(p.dVsT, q));
                imex_split(p, ss, istage, p.T_old, p.T_impl, p.T_expl, off, k, a, b);
                lds_store(PT, k, a);
                lds_store(QT, k, b);
            }
            if (exp_calc && !p.l_buo_imp) lds_store(Tv, k, hT ? gld(p.temp, q) : mk(0, 0));
        } else if (g == 1) {
            if (hX) {
                lds_store(eX, k, gld(p.X_expl, qe));
                if (exp_calc) lds_store(aX, k, gld(p.dVsXi, q));
                imex_split(p, ss, istage, p.X_old, p.X_impl, p.X_expl, off, k, a, b);
                lds_store(PX, k, a);
                lds_store(QX, k, b);
            }
            if (exp_calc && !p.l_buo_imp) lds_store(Xv, k, hX ? gld(p.xi, q) : mk(0, 0));
        } else if (g == 2) {
            lds_store(eP, k, gld(p.P_expl, qe));
            if (exp_calc) lds_store(aP, k, gld(p.dVsOm, q));
            imex_split(p, ss, istage, p.P_old, p.P_impl, p.P_expl, off, k, a, b);
            lds_store(PP, k, a);
            lds_store(QP, k, b);
        } else if (exp_calc) {
            lds_store(psiv, k, gld(p.psi, q));
            lds_store(omv, k, gld(p.om, q));
            lds_store(upv, k, gld(p.up, q));
            lds_store(usv, k, gld(p.us, q));
        }
    }
    const double wcur = imex_weight(p, ss, istage);
    // the groups above wrote vectors that other threads read next, with or
    // without the explicit part
    __syncthreads();

    // ---------------- explicit assembly -----------------------------------
    if (exp_calc) {
        VecSet<const cd> din;
        VecSet<cd> dout;
        if (hT) { din.push(aT); dout.push(cT); }
        if (hX) { din.push(aX); dout.push(cX); }
        din.push(aP); dout.push(cP);
        dct_fft(p, din, dout, din.n, 1, cosT, sinT, fbuf);
        __syncthreads();
        // derivative of the advective flux (into the a-vectors)
        {
            VecSet<const cd> c3;
            VecSet<cd> d3;
            if (hT) { c3.push(cT); d3.push(aT); }
            if (hX) { c3.push(cX); d3.push(aX); }
            c3.push(cP); d3.push(aP);
            dr_fft(p, R, c3, d3, VecSet<cd>(), c3.n, cosT, sinT, fbuf);
        }
        __syncthreads();
        const double rc2 = p.r_cmb * p.r_cmb;
        for (int k = tid; k < N; k += blockDim.x) {
            const double rk = R.r(k), h2 = rc2 - rk * rk;
            if (hT) {
                // finish_exp_temp_int (update_temp_integ.f90:420-438)
                double s = h2 * R.dtcond(k) + p.tadvz * rk * R.tcond(k);
                cd t = cmul(mk(0.0, (double)m * s), psiv[k]);
                cd e = csub(cscale(rk, eT[k]), t);
                eT[k] = csub(e, aT[k]);
            }
            if (hX) {
                cd t = cmul(mk(0.0, (double)m * h2 * R.dxicond(k)), psiv[k]);
                cd e = csub(cscale(rk, eX[k]), t);
                eX[k] = csub(e, aX[k]);
            }
            // finish_exp_psi_int_smat (update_psi_integ_smat.f90:832-930)
            cd e = eP[k];
            if (m != 0) {
                e = cscale(rk, e);
                if (!p.l_coriolis_imp) e = csub(e, cmul(mk(0.0, p.CorFac * (double)m * rk), psiv[k]));
                if (!p.l_buo_imp) {
                    if (hT) e = csub(e, cmul(mk(0.0, p.BuoFac * R.rgrav(k) * (double)m), Tv[k]));
                    if (hX) e = csub(e, cmul(mk(0.0, p.ChemFac * R.rgrav(k) * (double)m), Xv[k]));
                }
            }
            if (p.l_ek_pump) {
                double ekp = p.CorFac * R.ekpump(k);
                if (m == 0) {
                    e = csub(e, cscale(ekp, upv[k]));
                } else {
                    cd t = cadd(omv[k], cscale(R.ekp_up(k), upv[k]));
                    t = cadd(t, cmul(mk(R.ekp_us(k), (double)m * R.ekp_dusdp(k)), usv[k]));
                    e = csub(e, cscale(ekp * rk, t));
                }
            }
            if (m != 0) e = csub(e, aP[k]);
            eP[k] = e;
        }
        __syncthreads();
        din = VecSet<const cd>();
        dout = VecSet<cd>();
        if (hT) { din.push(eT); dout.push(cT); }
        if (hX) { din.push(eX); dout.push(cX); }
        din.push(eP); dout.push(cP);
        dct_fft(p, din, dout, din.n, 1, cosT, sinT, fbuf);
        __syncthreads();
        if (hT) bandmv<3>(p, p.T_RHSE1, cT, eT, 2);
        if (hX) bandmv<3>(p, p.T_RHSE1, cX, eX, 2);
        if (m == 0) bandmv<4>(p, p.P_RHSE1, cP, eP, 2);
        else bandmv<7>(p, p.P_RHSE3, cP, eP, 4);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            if (hT) p.T_expl[is * F + off + k] = eT[k];
            if (hX) p.X_expl[is * F + off + k] = eX[k];
            p.P_expl[is * F + off + k] = eP[k];
        }
    }

    // ---------------- T and xi: rhs, bordered solve ------------------------
    for (int k = tid; k < N; k += blockDim.x) {
        if (hT) {
            cd r = cadd(cadd(PT[k], cscale(wcur, eT[k])), QT[k]);
            if (k == 0) r = p.topt[mi];
            if (k == 1) r = p.bott[mi];
            aT[k] = r;
        }
        if (hX) {
            cd r = cadd(cadd(PX[k], cscale(wcur, eX[k])), QX[k]);
            if (k == 0) r = p.topxi[mi];
            if (k == 1) r = p.botxi[mi];
            aX[k] = r;
        }
    }
    __syncthreads();
    // bordered solves of T and xi (update_temp_integ.f90:226-240) by their inverses
    if (hT && hX) {
        const VecSet<const double> inv(p.T_inv + off * ldi, p.X_inv + off * ldi);
        const VecSet<cd> xs(aT, aX);
        block_matvec<double, 2>(N, ldi, inv, xs, xs, red);
    } else if (hT || hX) {
        const VecSet<const double> inv(hT ? p.T_inv + off * ldi : p.X_inv + off * ldi);
        const VecSet<cd> xs(hT ? aT : aX);
        block_matvec<double, 1>(N, ldi, inv, xs, xs, red);
    }
    {
        VecSet<const cd> din;
        VecSet<cd> dout;
        if (hT) { din.push(aT); dout.push(Tv); }
        if (hX) { din.push(aX); dout.push(Xv); }
        if (din.n) dct_fft(p, din, dout, din.n, 0, cosT, sinT, fbuf);
    }
    __syncthreads();
    const int inext = (istage == p.nstages) ? 1 : istage + 1;
    // implicit terms for the next stage; also before a CNAB2 start-up step
    // of a multistep scheme, which recomputes them (step_time.f90:417-503)
    const int bridge_next = p.bridge && ss->n_steps + 1 <= p.nold - 1;
    const int lcalc = p.l_imp_calc_rhs[inext - 1] || bridge_next;
    if (hT) rotate_row(p, p.T_old, p.T_impl, p.T_expl, off);
    if (hX) rotate_row(p, p.X_old, p.X_impl, p.X_expl, off);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        if (hT) { p.temp_hat[off + k] = aT[k]; p.temp[off + k] = Tv[k]; }
        if (hX) { p.xi_hat[off + k] = aX[k]; p.xi[off + k] = Xv[k]; }
    }
    // get_temp_rhs_imp_int (update_temp_integ.f90:471-536)
    if (inext == 1) {
        if (hT) bandmv<4>(p, p.T_RHSE2, aT, p.T_old + off, 2);
        if (hX) bandmv<4>(p, p.T_RHSE2, aX, p.X_old + off, 2);
    }
    if (lcalc) {
        if (hT) bandmv<2>(p, p.T_RHSI + (size_t)mi * N * 5, aT, cT, 2);
        if (hX) bandmv<2>(p, p.T_RHSI + (size_t)mi * N * 5, aX, cX, 2);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            if (hT) p.T_impl[(inext - 1) * F + off + k] = cscale(p.TdiffFac * p.hdifT[mi], cT[k]);
            if (hX) p.X_impl[(inext - 1) * F + off + k] = cscale(p.XidiffFac * p.hdifXi[mi], cX[k]);
        }
    }
    __syncthreads();

    // ---------------- psi -------------------------------------------------
    if (p.l_buo_imp) {
        for (int k = tid; k < N; k += blockDim.x)
            bv[k] = buoyancy(p, R, m, k, hT ? Tv[k] : mk(0, 0), hX ? Xv[k] : mk(0, 0));
        __syncthreads();
        VecSet<const cd> din(bv);
        VecSet<cd> dout(wv);
        dct_fft(p, din, dout, 1, 0, cosT, sinT, fbuf);
        __syncthreads();
        if (m > 0) bandmv<8>(p, p.P_RHSE2, wv, b2, 4);
        __syncthreads();
    }
    const double wl = ss->wimp_lin[0];
    for (int k = tid; k < N; k += blockDim.x) {
        cd r = cadd(cadd(PP[k], cscale(wcur, eP[k])), QP[k]);
        if (p.l_buo_imp && m != 0) r = cadd(r, cscale(wl, b2[k]));
        aP[k] = r;
    }
    __syncthreads();
    {
        // bordered solve of psi (update_psi_integ_smat.f90:417-432) by its inverse
        VecSet<const cd> inv(p.P_inv + off * ldi);
        VecSet<cd> xs(aP);
        block_matvec<cd, 1>(N, ldi, inv, xs, xs, red);
    }
    const double rc2 = p.r_cmb * p.r_cmb;
    if (m == 0) {
        double *f = dtmp, *df = dtmp + N;
        for (int k = tid; k < N; k += blockDim.x) f[k] = aP[k].x;
        __syncthreads();
        if (tid == 0) dcheb_real(p, f, df);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) { cP[k] = mk(df[k], 0.0); eP[k] = mk(f[k], 0.0); }
        __syncthreads();
        VecSet<const cd> din(cP, eP);
        VecSet<cd> dout(omv, upv);
        dct_fft(p, din, dout, 2, 0, cosT, sinT, fbuf);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            double u0 = upv[k].x, om0 = R.drx(k) * omv[k].x;
            p.us[off + k] = mk(0, 0);
            p.up[off + k] = mk(u0, 0.0);
            p.om[off + k] = mk(om0 + R.or1(k) * u0, 0.0);
            eP[k] = mk(u0, 0.0);
        }
        __syncthreads();
        // work_r = costf(real(up(m=0))) for get_psi_rhs_imp_int_smat
        VecSet<const cd> di(eP);
        VecSet<cd> dd(wv);
        dct_fft(p, di, dd, 1, 0, cosT, sinT, fbuf);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) wv[k] = mk(wv[k].x, 0.0);
    } else {
        VecSet<const cd> c1(aP);
        VecSet<cd> d1(cP), d2(eP);
        // derivatives and grid values of psi from one FFT pass
        dr_fft(p, R, c1, d1, d2, 1, cosT, sinT, fbuf, psiv);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            const double rk = R.r(k), h2 = rc2 - rk * rk, or1 = R.or1(k);
            cd ps = psiv[k], dps = cP[k], d2ps = eP[k];
            p.psi_hat[off + k] = aP[k];
            p.psi[off + k] = ps;
            p.us[off + k] = cmul(mk(0.0, (double)m * or1 * h2), ps);
            p.up[off + k] = cadd(cscale(-h2, dps), cscale(3.0 * rk, ps));
            cd o = csub(cscale(-h2, d2ps), cscale(rc2 * or1 - 6.0 * rk, dps));
            p.om[off + k] = cadd(o, cscale(6.0 + (double)m * (double)m * R.or2(k) * h2, ps));
            wv[k] = aP[k];
        }
    }
    __syncthreads();
    rotate_row(p, p.P_old, p.P_impl, p.P_expl, off);
    __syncthreads();
    // get_psi_rhs_imp_int_smat (update_psi_integ_smat.f90:960-1115)
    if (inext == 1) {
        if (m == 0) bandmv<4>(p, p.P_RHSI + (size_t)mi * N * 17, wv, p.P_old + off, 2);
        else bandmv<8>(p, p.P_RHSI + (size_t)mi * N * 17, wv, p.P_old + off, 4);
    }
    if (lcalc) {
        if (m == 0) bandmv_c<2>(p, p.P_RHSIL + (size_t)mi * N * 17, wv, cT, 2);
        else bandmv_c<8>(p, p.P_RHSIL + (size_t)mi * N * 17, wv, cT, 4);
        cd *bimp = b2;
        if (bridge_next && p.l_buo_imp && hX && m != 0) {
            // the start-up recomputation passes psi in the composition slot
            // (step_time.f90:475-478): buoyancy of (T, psi) instead of (T, xi)
            for (int k = tid; k < N; k += blockDim.x)
                bv[k] = buoyancy(p, R, m, k, hT ? Tv[k] : mk(0, 0), psiv[k]);
            __syncthreads();
            dct_fft(p, VecSet<const cd>(bv), VecSet<cd>(aT), 1, 0, cosT, sinT, fbuf);
            __syncthreads();
            bandmv<8>(p, p.P_RHSE2, aT, aX, 4);
            bimp = aX;
        }
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            cd v = cT[k];
            if (p.l_buo_imp && m != 0) v = cadd(v, bimp[k]);
            p.P_impl[(inext - 1) * F + off + k] = v;
        }
    }
}

// the descriptor is read through a pointer (scalar loads): a by-value copy
// of this size would be materialised in scratch memory
__global__ void __launch_bounds__(PZ_MPHASE_THREADS) k_mphase(const PzDev *__restrict__ pd, int istage, int exp_calc)
{
    mphase_body(*pd, blockIdx.x, istage, exp_calc);
}
__global__ void __launch_bounds__(PZ_MPHASE_THREADS) k_mphase_batch(PzBatch b, int istage, int exp_calc)
{
    const int2 e = b.map[blockIdx.x];
    mphase_body(*b.runs[e.x], e.y, istage, exp_calc);
}

// get_start_fields, fresh start of the integration method
// (init_fields.f90:186-255): dT/dr, omega, hats and the first implicit state.
__global__ void __launch_bounds__(256) k_start(const PzDev p __attribute__((grid_constant)))
{
    extern __shared__ double smem[];
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi];
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    (void)F;
    const int tid = threadIdx.x;
    cd *V = (cd *)smem;
    cd *Tv = V, *Xv = V + N, *upv = V + 2 * N, *usv = V + 3 * N, *psv = V + 4 * N;
    cd *cT = V + 5 * N, *cX = V + 6 * N, *cU = V + 7 * N, *cPs = V + 8 * N;
    cd *dT = V + 9 * N, *dU = V + 10 * N, *bv = V + 11 * N, *wv = V + 12 * N, *b2 = V + 13 * N, *t1 = V + 14 * N;
    const bool hT = p.l_heat, hX = p.l_chem;
    for (int k = tid; k < N; k += blockDim.x) {
        Tv[k] = hT ? p.temp[off + k] : mk(0, 0);
        Xv[k] = hX ? p.xi[off + k] : mk(0, 0);
        upv[k] = p.up[off + k];
        usv[k] = p.us[off + k];
        psv[k] = p.psi[off + k];
    }
    __syncthreads();
    {
        VecSet<const cd> din(Tv, Xv, upv, psv);
        VecSet<cd> dout(cT, cX, cU, cPs);
        dct_many(p, din, dout, 4, 0);
    }
    __syncthreads();
    {
        VecSet<const cd> c2(cT, cU);
        VecSet<cd> d2(dT, dU);
        dr_many(p, RadView(p.radc, p.N), c2, d2, VecSet<cd>(), 2);
    }
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        if (hT) { p.dtemp[off + k] = dT[k]; p.temp_hat[off + k] = cT[k]; }
        if (hX) p.xi_hat[off + k] = cX[k];
        p.psi_hat[off + k] = cPs[k];
        cd o = cadd(dU[k], cscale(p.or1[k], upv[k]));
        p.om[off + k] = csub(o, cmul(mk(0.0, p.or1[k] * (double)m), usv[k]));
    }
    // scalar implicit states (istage = 1, l_calc_lin = .true.)
    if (hT) {
        bandmv<4>(p, p.T_RHSE2, cT, p.T_old + off, 2);
        bandmv<2>(p, p.T_RHSI + (size_t)mi * N * 5, cT, t1, 2);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) p.T_impl[off + k] = cscale(p.TdiffFac * p.hdifT[mi], t1[k]);
        __syncthreads();
    }
    if (hX) {
        bandmv<4>(p, p.T_RHSE2, cX, p.X_old + off, 2);
        bandmv<2>(p, p.T_RHSI + (size_t)mi * N * 5, cX, t1, 2);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) p.X_impl[off + k] = cscale(p.XidiffFac * p.hdifXi[mi], t1[k]);
        __syncthreads();
    }
    // psi: the reference passes psi in the composition slot (init_fields.f90:250-251)
    if (m == 0) {
        // work_r = costf(real(up(m=0)))
        for (int k = tid; k < N; k += blockDim.x) bv[k] = mk(upv[k].x, 0.0);
        __syncthreads();
        VecSet<const cd> di(bv);
        VecSet<cd> dd(wv);
        dct_many(p, di, dd, 1, 0);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) wv[k] = mk(wv[k].x, 0.0);
        __syncthreads();
        bandmv<4>(p, p.P_RHSI + (size_t)mi * N * 17, wv, p.P_old + off, 2);
        bandmv_c<2>(p, p.P_RHSIL + (size_t)mi * N * 17, wv, p.P_impl + off, 2);
    } else {
        bandmv<8>(p, p.P_RHSI + (size_t)mi * N * 17, cPs, p.P_old + off, 4);
        bandmv_c<8>(p, p.P_RHSIL + (size_t)mi * N * 17, cPs, t1, 4);
        if (p.l_buo_imp) {
            for (int k = tid; k < N; k += blockDim.x) bv[k] = buoyancy(p, RadView(p.radc, p.N), m, k, Tv[k], psv[k]);
            __syncthreads();
            VecSet<const cd> di(bv);
            VecSet<cd> dd(wv);
            dct_many(p, di, dd, 1, 0);
            __syncthreads();
            bandmv<8>(p, p.P_RHSE2, wv, b2, 4);
        }
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            cd v = t1[k];
            if (p.l_buo_imp) v = cadd(v, b2[k]);
            p.P_impl[off + k] = v;
        }
    }
}

// radial profiles for the kinetic energies (outputs.f90:458-467): per r,
// sum over m of cc2real(us), cc2real(up) and the m = 0 part of up
__global__ void k_ekin_profiles(const PzDev p __attribute__((grid_constant)), double *out)
{
    const int nr = blockIdx.x * blockDim.x + threadIdx.x;
    if (nr >= p.N) return;
    double us2 = 0.0, up2 = 0.0, upa = 0.0;
    for (int i = 0; i < p.nm; i++) {
        const int m = p.idx2m[i];
        cd u = p.us[(size_t)i * p.N + nr], v = p.up[(size_t)i * p.N + nr];
        double a = m == 0 ? u.x * u.x : 2.0 * (u.x * u.x + u.y * u.y);
        double b = m == 0 ? v.x * v.x : 2.0 * (v.x * v.x + v.y * v.y);
        us2 += a;
        up2 += b;
        if (m == 0) upa += b;
    }
    out[nr] = us2;
    out[p.N + nr] = up2;
    out[2 * p.N + nr] = upa;
}

// ---- unit-operator kernels on [vec][n_r] device arrays ----
__global__ void __launch_bounds__(256) k_costf_vecs(const PzDev p __attribute__((grid_constant)), cd *f)
{
    extern __shared__ double smem[];
    cd *a = (cd *)smem, *b = a + p.N, *buf = b + p.N;
    cd *x = f + (size_t)blockIdx.x * p.N;
    for (int k = threadIdx.x; k < p.N; k += blockDim.x) a[k] = x[k];
    __syncthreads();
    VecSet<const cd> din(a);
    VecSet<cd> dout(b);
    dct_fft(p, din, dout, 1, 0, p.cosT, p.sinT, buf);
    __syncthreads();
    for (int k = threadIdx.x; k < p.N; k += blockDim.x) x[k] = b[k];
}

__global__ void __launch_bounds__(256) k_dr_vecs(const PzDev p __attribute__((grid_constant)), const cd *f, cd *df, cd *ddf, int l_dct_in)
{
    extern __shared__ double smem[];
    cd *a = (cd *)smem, *c = a + p.N, *d1 = c + p.N, *d2 = d1 + p.N, *buf = d2 + p.N;
    const size_t off = (size_t)blockIdx.x * p.N;
    for (int k = threadIdx.x; k < p.N; k += blockDim.x) a[k] = f[off + k];
    __syncthreads();
    if (l_dct_in) {
        VecSet<const cd> din(a);
        VecSet<cd> dout(c);
        dct_fft(p, din, dout, 1, 0, p.cosT, p.sinT, buf);
    } else {
        for (int k = threadIdx.x; k < p.N; k += blockDim.x) c[k] = a[k];
    }
    __syncthreads();
    VecSet<const cd> cc(c);
    VecSet<cd> o1(d1), o2(d2);
    dr_fft(p, RadView(p.radc, p.N), cc, o1, ddf ? o2 : VecSet<cd>(), 1, p.cosT, p.sinT, buf);
    __syncthreads();
    for (int k = threadIdx.x; k < p.N; k += blockDim.x) {
        df[off + k] = d1[k];
        if (ddf) ddf[off + k] = d2[k];
    }
}

// azimuthal transforms, one radius per block: grid real [rad][n_phi],
// spectral [rad][nm]
__global__ void __launch_bounds__(256) k_phi_fft(const PzDev p __attribute__((grid_constant)), const double *grid, cd *spec)
{
    extern __shared__ double smem[];
    const int n = p.nphi, nm = p.nm;
    cd *a = (cd *)smem, *b = a + n;
    const double *g = grid + (size_t)blockIdx.x * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) a[j] = mk(g[j], 0.0);
    __syncthreads();
    cd *Z = fft_lds(p, a, b, n, 1, -1);
    for (int i = threadIdx.x; i < nm; i += blockDim.x) spec[(size_t)blockIdx.x * nm + i] = cdivr(Z[i], (double)n);
}

__global__ void __launch_bounds__(256) k_phi_ifft(const PzDev p __attribute__((grid_constant)), const cd *spec, double *grid)
{
    extern __shared__ double smem[];
    const int n = p.nphi, nm = p.nm;
    cd *a = (cd *)smem, *b = a + n;
    const cd *s = spec + (size_t)blockIdx.x * nm;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        cd z = mk(0, 0);
        if (k == 0) z = mk(s[0].x, 0.0);
        else if (k < nm) z = s[k];
        else if (n - k < nm) z = cconj(s[n - k]);
        a[k] = z;
    }
    __syncthreads();
    cd *Z = fft_lds(p, a, b, n, 1, +1);
    for (int j = threadIdx.x; j < n; j += blockDim.x) grid[(size_t)blockIdx.x * n + j] = Z[j].x;
}

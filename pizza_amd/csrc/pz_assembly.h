// Device code of libpizza_amd, part 4: the assembly stage of the DIRK
// schemes that end with one (assemble_stage, mloop.f90:105-160), for the
// collocation method.  The last stage evaluates only the explicit terms;
// the new fields are old + sum_s b_exp(s) expl(s) + b_imp(s) impl(s)
// (assemble_imex, dirk_schemes.f90:882-921), with the boundary values of
// T, xi and u_phi(m = 0) from two Robin conditions (robin_bc,
// chebyshev.f90:385-445) and psi recovered from the assembled omega
// (assemble_psi_coll_smat, update_psi_coll_smat.f90:529-672).
// Included once by pz_api.hip, after pz_coll.h.
#pragma once
#include "pz_coll.h"

// get_assMat (update_psi_coll_smat.f90:839-917): the time-independent
// matrix that gives psi from omega for one m /= 0 (non-penetration and
// no-slip or free-slip rows, boundary_fac columns), inverted by
// Gauss-Jordan into A_inv [m][col][ldi].  grid (n_m), built once.
__global__ void __launch_bounds__(1024) k_build_ass(const PzDev p __attribute__((grid_constant)))
{
    __shared__ double fcol[2 * 257];
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x, nc = p.ncheb, ldi = p.ldi;
    if (m == 0) return;
    const RadView R(p.radc, N);
    double *W = (double *)(p.coll_work + (size_t)mi * p.coll_ws); // N x 2N, row-major
    const double rn = p.rnorm, dm2 = (double)m * (double)m;
    for (int t = tid; t < N * 2 * N; t += blockDim.x) {
        const int r = t / (2 * N), c = t - r * 2 * N;
        double a;
        if (c >= N) {
            a = (c - N == r) ? 1.0 : 0.0;
        } else {
            const bool in = c < nc;
            if (r == 0) a = in ? rn * PZ_CM(0, c) : 0.0;
            else if (r == N - 1) a = in ? rn * PZ_CM(N - 1, c) : 0.0;
            else if (r == 1) a = !in ? 0.0 : (p.ktopv == 1 ? rn * (PZ_CD2M(0, c) - R.or1(0) * PZ_CDM(0, c)) : rn * PZ_CDM(0, c));
            else if (r == N - 2)
                a = !in ? 0.0
                        : (p.kbotv == 1 ? rn * (PZ_CD2M(N - 1, c) - R.or1(N - 1) * PZ_CDM(N - 1, c))
                                        : rn * PZ_CDM(N - 1, c));
            else
                a = -rn * (PZ_CD2M(r, c) + (R.or1(r) + R.beta(r)) * PZ_CDM(r, c) +
                           (R.or1(r) * R.beta(r) + R.dbeta(r) - dm2 * R.or2(r)) * PZ_CM(r, c));
            if (c == 0 || c == N - 1) a *= p.boundary_fac;
        }
        W[t] = a;
    }
    __syncthreads();
    const int info = gj_invert<double>(W, N, fcol);
    if (tid == 0 && info) p.info[mi] = info; // reported as a singular matrix
    if (info) return;
    double *out = p.A_inv + (size_t)mi * N * ldi;
    for (int t = tid; t < N * ldi; t += blockDim.x) {
        const int c = t / ldi, r = t - c * ldi;
        out[(size_t)c * ldi + r] = r < N ? W[(size_t)r * 2 * N + N + c] : 0.0;
    }
}

// old + sum_s b_exp(s) expl(s) + sum_s b_imp(s) impl(s) for element q
__device__ __forceinline__ cd assemble_elem(const PzDev &p, const PzScheme *s, const cd *old, const cd *impl,
                                            const cd *expl, size_t q)
{
    const size_t F = (size_t)p.nm * p.N;
    cd r = old[q];
    for (int st = 0; st < p.nstages; st++) r = cadd(r, cscale(s->bass_exp[st], expl[st * F + q]));
    for (int st = 0; st < p.nstages; st++) r = cadd(r, cscale(s->bass_imp[st], impl[st * F + q]));
    return r;
}

// robin_bc (chebyshev.f90:385-445) on an LDS vector f: the two boundary
// values from atop df/dr + btop f = rtop and abot df/dr + bbot f = rbot,
// with the boundary-derivative rows dr_top, dr_bot.  red: 4 + 4 * waves
// doubles of LDS.  Whole workgroup; ends with a barrier.
__device__ void robin_device(const PzDev &p, double at, double bt, cd rtop, double ab, double bb, cd rbot, cd *f,
                             double *red)
{
    const int N = p.N, lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const double *dt = p.dr_top, *db = p.dr_bot;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int k = 1 + threadIdx.x; k < N - 1; k += blockDim.x) {
        const cd v = f[k];
        s0 += dt[k] * v.x;
        s1 += dt[k] * v.y;
        s2 += db[k] * v.x;
        s3 += db[k] * v.y;
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    s3 = wave_sum(s3);
    __syncthreads();
    if (lane == 0) {
        red[4 + 4 * w] = s0;
        red[5 + 4 * w] = s1;
        red[6 + 4 * w] = s2;
        red[7 + 4 * w] = s3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double str = 0, sti = 0, sbr = 0, sbi = 0;
        for (int i = 0; i < nw; i++) {
            str += red[4 + 4 * i];
            sti += red[5 + 4 * i];
            sbr += red[6 + 4 * i];
            sbi += red[7 + 4 * i];
        }
        const double den = ab * at * db[0] * dt[N - 1] - (at * dt[0] + bt) * (ab * db[N - 1] + bb);
        const double vbr = (ab * db[0] * (rtop.x - at * str) - (at * dt[0] + bt) * (rbot.x - ab * sbr)) / den;
        const double vbi = (ab * db[0] * (rtop.y - at * sti) - (at * dt[0] + bt) * (rbot.y - ab * sbi)) / den;
        const double vtr = (rtop.x - at * (str + dt[N - 1] * vbr)) / (at * dt[0] + bt);
        const double vti = (rtop.y - at * (sti + dt[N - 1] * vbi)) / (at * dt[0] + bt);
        f[0] = mk(vtr, vti);
        f[N - 1] = mk(vbr, vbi);
    }
    __syncthreads();
}

#define ASM_NVEC 20
#define ASM_THREADS 256
// LDS of k_assemble_coll: ASM_NVEC vectors, the FFT scratch (6 L + 16) shared
// with the dense solve's reduction scratch (4 x threads), 3 n_r doubles and
// the reduction scratch of robin_device (4 + 4 x waves doubles)
__host__ __device__ inline int asm_scratch(int N) { return 6 * (2 * N - 2) + 16 > 4 * ASM_THREADS ? 6 * (2 * N - 2) + 16 : 4 * ASM_THREADS; }
__host__ __device__ inline size_t asm_lds_bytes(int N)
{
    return (size_t)(ASM_NVEC * N + asm_scratch(N)) * sizeof(cd) + (3 * (size_t)N + 4 + 64) * sizeof(double);
}

// ---------------------------------------------------------------------------
// The assembly stage for one m per workgroup: assemble_temp_coll /
// assemble_xi_coll (update_temp_coll.f90:238-307) and assemble_psi_coll_smat
// (update_psi_coll_smat.f90:529-672), each ending with the implicit terms of
// the next step's first stage (get_*_rhs_imp_coll with istage = 1).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(ASM_THREADS) k_assemble_coll(const PzDev *__restrict__ pd)
{
    extern __shared__ double smem[];
    const PzDev &p = *pd;
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x, nb = blockDim.x, nc = p.ncheb;
    const size_t off = (size_t)mi * N;
    const PzScheme *ss = p.ss;
    const RadView R(p.radc, N);
    const bool hT = p.l_heat, hX = p.l_chem;
    const int lcalc = p.l_imp_calc_rhs[0];
    const double dm = (double)m, dm2 = dm * dm;
    cd *V = (cd *)smem;
    cd *Tv = V, *Xv = V + N, *c1 = V + 2 * N, *d1 = V + 3 * N, *d2 = V + 4 * N;
    cd *wv = V + 5 * N, *z = V + 6 * N, *dps = V + 7 * N, *d2ps = V + 8 * N, *psg = V + 9 * N;
    cd *omv = V + 10 * N, *usv = V + 11 * N, *upv = V + 12 * N, *e1 = V + 13 * N, *e2 = V + 14 * N;
    cd *impl = V + 15 * N, *cX = V + 16 * N;
    cd *fbuf = V + ASM_NVEC * N;
    double *dt3 = (double *)(fbuf + asm_scratch(N)), *red = dt3 + 3 * N;
    const double *cosT = p.cosT, *sinT = p.sinT;

    // ---- explicit terms of the last stage: the finish of mcoll_body
    // (finish_exp_*_coll), which the skipped m-loop of that stage would do
    if (p.l_exp_calc[p.nstages - 1]) {
        const size_t F = (size_t)p.nm * N, qe = (size_t)(p.nstages - 1) * F + off;
        cd *aT = c1, *aX = d1, *aP = d2, *cT = dps, *cXe = d2ps, *cP = psg;
        cd *eT = omv, *eX = upv, *eP = e1, *Told = e2, *Xold = impl;
        for (int k = tid; k < N; k += nb) {
            if (hT) { eT[k] = p.T_expl[qe + k]; aT[k] = p.dVsT[off + k]; Told[k] = p.temp[off + k]; }
            if (hX) { eX[k] = p.X_expl[qe + k]; aX[k] = p.dVsXi[off + k]; Xold[k] = p.xi[off + k]; }
            eP[k] = p.P_expl[qe + k];
            aP[k] = p.dVsOm[off + k];
            usv[k] = p.us[off + k];
        }
        __syncthreads();
        {
            VecSet<const cd> din;
            VecSet<cd> dout;
            if (hT) { din.push(aT); dout.push(cT); }
            if (hX) { din.push(aX); dout.push(cXe); }
            din.push(aP); dout.push(cP);
            dct_fft(p, din, dout, din.n, 0, cosT, sinT, fbuf);
        }
        __syncthreads();
        {
            VecSet<const cd> c3;
            VecSet<cd> d3;
            if (hT) { c3.push(cT); d3.push(aT); }
            if (hX) { c3.push(cXe); d3.push(aX); }
            c3.push(cP); d3.push(aP);
            dr_fft(p, R, c3, d3, VecSet<cd>(), c3.n, cosT, sinT, fbuf);
        }
        __syncthreads();
        for (int k = tid; k < N; k += nb) {
            const double o1 = R.or1(k);
            if (hT)
                p.T_expl[qe + k] = csub(csub(eT[k], cscale(o1, aT[k])),
                                        cscale(R.dtcond(k) - p.tadvz * R.beta(k) * R.tcond(k), usv[k]));
            if (hX) p.X_expl[qe + k] = csub(csub(eX[k], cscale(o1, aX[k])), cscale(R.dxicond(k), usv[k]));
            if (m != 0) {
                cd e = csub(eP[k], cscale(o1, aP[k]));
                if (!p.l_coriolis_imp) e = cadd(e, cscale(p.CorFac * R.beta(k), usv[k]));
                if (!p.l_buo_imp) {
                    if (hT) e = csub(e, cmul(mk(0.0, p.BuoFac * R.rgrav(k) * o1 * dm), Told[k]));
                    if (hX) e = csub(e, cmul(mk(0.0, p.ChemFac * R.rgrav(k) * o1 * dm), Xold[k]));
                }
                p.P_expl[qe + k] = e;
            }
        }
        __syncthreads();
    }

    // ---- T and xi: interior from the assembled right-hand side ----------
    for (int k = tid; k < N; k += nb) {
        const size_t q = off + k;
        if (hT) {
            const cd r = assemble_elem(p, ss, p.T_old, p.T_impl, p.T_expl, q);
            Tv[k] = (k == 0 || k == N - 1) ? p.temp[q] : (m == 0 ? mk(r.x, 0.0) : r);
        }
        if (hX) {
            const cd r = assemble_elem(p, ss, p.X_old, p.X_impl, p.X_expl, q);
            Xv[k] = (k == 0 || k == N - 1) ? p.xi[q] : (m == 0 ? mk(r.x, 0.0) : r);
        }
        const cd r = assemble_elem(p, ss, p.P_old, p.P_impl, p.P_expl, q);
        wv[k] = r;
    }
    __syncthreads();
    // Dirichlet (k = 1): a = 0, b = 1; Neumann: a = 1, b = 0
    if (hT)
        robin_device(p, p.ktopt == 1 ? 0.0 : 1.0, p.ktopt == 1 ? 1.0 : 0.0, p.topt[mi], p.kbott == 1 ? 0.0 : 1.0,
                     p.kbott == 1 ? 1.0 : 0.0, p.bott[mi], Tv, red);
    if (hX)
        robin_device(p, p.ktopxi == 1 ? 0.0 : 1.0, p.ktopxi == 1 ? 1.0 : 0.0, p.topxi[mi], p.kbotxi == 1 ? 0.0 : 1.0,
                     p.kbotxi == 1 ? 1.0 : 0.0, p.botxi[mi], Xv, red);
    // first-stage implicit terms: get_ddr of the grid values
    if (hT) {
        if (lcalc) {
            dct_fft(p, VecSet<const cd>(Tv), VecSet<cd>(c1), 1, 0, cosT, sinT, fbuf);
            __syncthreads();
            dr_fft(p, R, VecSet<const cd>(c1), VecSet<cd>(d1), VecSet<cd>(d2), 1, cosT, sinT, fbuf);
            __syncthreads();
        }
        const double kap = p.TdiffFac * p.hdifT[mi];
        for (int k = tid; k < N; k += nb) {
            p.temp[off + k] = Tv[k];
            p.T_old[off + k] = Tv[k];
            if (lcalc) {
                p.dtemp[off + k] = d1[k];
                const cd v = csub(cadd(d2[k], cscale(R.or1(k), d1[k])), cscale(dm2 * R.or2(k), Tv[k]));
                p.T_impl[off + k] = cscale(kap, v);
            }
        }
        __syncthreads();
    }
    if (hX) {
        if (lcalc) {
            dct_fft(p, VecSet<const cd>(Xv), VecSet<cd>(cX), 1, 0, cosT, sinT, fbuf);
            __syncthreads();
            dr_fft(p, R, VecSet<const cd>(cX), VecSet<cd>(d1), VecSet<cd>(d2), 1, cosT, sinT, fbuf);
            __syncthreads();
        }
        const double kap = p.XidiffFac * p.hdifXi[mi];
        for (int k = tid; k < N; k += nb) {
            p.xi[off + k] = Xv[k];
            p.X_old[off + k] = Xv[k];
            if (lcalc) {
                const cd v = csub(cadd(d2[k], cscale(R.or1(k), d1[k])), cscale(dm2 * R.or2(k), Xv[k]));
                p.X_impl[off + k] = cscale(kap, v);
            }
        }
        __syncthreads();
    }

    // ---- psi / omega / u_phi ---------------------------------------------
    if (m == 0) {
        // u_phi(m = 0): interior from the assembly, boundary values rigid or
        // from the stress-free Robin rows (update_psi_coll_smat.f90:554-579)
        for (int k = tid; k < N; k += nb) upv[k] = (k == 0 || k == N - 1) ? mk(0, 0) : mk(wv[k].x, 0.0);
        __syncthreads();
        const int kt = p.ktopv, kb = p.kbotv;
        if (kt == 1 || kb == 1)
            robin_device(p, kt == 1 ? 1.0 : 0.0, kt == 1 ? -R.or1(0) : 1.0, mk(0, 0), kb == 1 ? 1.0 : 0.0,
                         kb == 1 ? -R.or1(N - 1) : 1.0, mk(0, 0), upv, red);
        // om0 = get_dr(uphi0) (grid values in): costf, dcheb, costf, drx
        dct_fft(p, VecSet<const cd>(upv), VecSet<cd>(c1), 1, 0, cosT, sinT, fbuf);
        __syncthreads();
        double *f = dt3, *df = dt3 + N;
        for (int k = tid; k < N; k += nb) f[k] = c1[k].x;
        __syncthreads();
        if (tid == 0) dcheb_real(p, f, df);
        __syncthreads();
        for (int k = tid; k < N; k += nb) d1[k] = mk(df[k], 0.0);
        __syncthreads();
        dct_fft(p, VecSet<const cd>(d1), VecSet<cd>(e1), 1, 0, cosT, sinT, fbuf);
        __syncthreads();
        for (int k = tid; k < N; k += nb) {
            const double u0 = upv[k].x;
            const cd om = mk(R.drx(k) * e1[k].x + R.or1(k) * u0, 0.0);
            p.us[off + k] = mk(0, 0);
            p.up[off + k] = upv[k];
            p.om[off + k] = om;
            p.psi[off + k] = mk(0, 0);
            p.P_old[off + k] = upv[k];
        }
        __syncthreads();
        if (lcalc) {
            coll_uphi0_impl(p, R, upv, e1, e2, dt3, cosT, sinT, fbuf, impl, p.hdifV[mi]);
            __syncthreads();
            for (int k = tid; k < N; k += nb) p.P_impl[off + k] = impl[k];
        }
        return;
    }
    // psi coefficients from omega: rows 0, 1, N-2, N-1 hold the boundary
    // conditions (zero right-hand side)
    for (int k = tid; k < N; k += nb) z[k] = (k < 2 || k > N - 3) ? mk(0, 0) : wv[k];
    __syncthreads();
    block_matvec<double, 1>(N, p.ldi, VecSet<const double>(p.A_inv + off * p.ldi), VecSet<cd>(z), VecSet<cd>(z),
                            (cd *)fbuf);
    for (int k = nc + tid; k < N; k += nb) z[k] = mk(0, 0);
    __syncthreads();
    dr_fft(p, R, VecSet<const cd>(z), VecSet<cd>(dps), VecSet<cd>(d2ps), 1, cosT, sinT, fbuf, psg);
    __syncthreads();
    for (int k = tid; k < N; k += nb) {
        const double o1 = R.or1(k), bk = R.beta(k);
        const cd ps = psg[k];
        const cd us = cmul(mk(0.0, dm * o1), ps);
        const cd up = csub(mk(-dps[k].x, -dps[k].y), cscale(bk, ps));
        const cd om = mk(-(d2ps[k].x + (o1 + bk) * dps[k].x + (o1 * bk + R.dbeta(k) - dm2 * R.or2(k)) * ps.x),
                         -(d2ps[k].y + (o1 + bk) * dps[k].y + (o1 * bk + R.dbeta(k) - dm2 * R.or2(k)) * ps.y));
        usv[k] = us;
        upv[k] = up;
        omv[k] = om;
        p.psi[off + k] = ps;
        p.us[off + k] = us;
        p.up[off + k] = up;
        p.om[off + k] = om;
        p.P_old[off + k] = om;
    }
    __syncthreads();
    if (!lcalc) return;
    // implicit terms of omega (get_psi_rhs_imp_coll_smat, grid values in)
    dct_fft(p, VecSet<const cd>(omv), VecSet<cd>(c1), 1, 0, cosT, sinT, fbuf);
    __syncthreads();
    dr_fft(p, R, VecSet<const cd>(c1), VecSet<cd>(e1), VecSet<cd>(e2), 1, cosT, sinT, fbuf);
    __syncthreads();
    for (int k = tid; k < N; k += nb) {
        const double o1 = R.or1(k), ekp = p.CorFac * R.ekpump(k);
        cd v = cscale(p.ViscFac * p.hdifV[mi], csub(cadd(e2[k], cscale(o1, e1[k])), cscale(dm2 * R.or2(k), omv[k])));
        cd pump = cadd(omv[k], cscale(R.ekp_up(k), upv[k]));
        pump = cadd(pump, cmul(mk(R.ekp_us(k), dm * R.ekp_dusdp(k)), usv[k]));
        v = csub(v, cscale(ekp, pump));
        if (p.l_coriolis_imp) v = cadd(v, cscale(p.CorFac * R.beta(k), usv[k]));
        if (p.l_buo_imp) {
            if (hT) v = csub(v, cmul(mk(0.0, R.rgrav(k) * o1 * p.BuoFac * dm), Tv[k]));
            if (hX) v = csub(v, cmul(mk(0.0, R.rgrav(k) * o1 * p.ChemFac * dm), Xv[k]));
        }
        p.P_impl[off + k] = v;
    }
}

// ---------------------------------------------------------------------------
// The assembly stage of the finite-difference method for one m per
// workgroup (assemble_stage_Rdist, mloop_fd.f90:241-274):
// assemble_temp_Rdist / assemble_xi_Rdist (update_temp_fd.f90:317-385,
// update_xi_fd.f90:293-361) take the assembled right-hand side as the new
// field (real at m = 0, zero on a fixed-value wall, or its pattern with
// l_fd_bc_pattern); assemble_psi_Rloc
// (update_psi_fd.f90:472-606) takes it as u_phi for m = 0 (zero on a
// no-slip wall) and solves -L psi = omega with psi = 0 at both walls for
// m /= 0 (get_elliptic_mat_Rdist, :838-888: tridiagonal, factorised here
// by one thread and solved by one wave).  Ghost zones and the first-stage
// implicit terms follow as in the m-loop.  LDS: FdLayout(N).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_assemble_fd(const PzDev *__restrict__ pd)
{
    extern __shared__ double smem[];
    const PzDev &p = *pd;
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x;
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    const PzScheme *ss = p.ss;
    const FdLayout lay(N);
    char *lb = (char *)smem;
    cd *gT = (cd *)(lb + lay.gT), *gX = (cd *)(lb + lay.gX), *gP = (cd *)(lb + lay.gP);
    cd *V = (cd *)(lb + lay.vec);
    cd *aT = V, *aX = V + N, *aP = V + 2 * N, *usv = V + 3 * N, *Tv = V + 4 * N, *Xv = V + 5 * N;
    double *eD = (double *)(lb + lay.mT), *eL = eD + N, *eU = eD + 2 * N;
    const bool hT = p.l_heat, hX = p.l_chem;
    const RadView R(p.radc, N);
    const FdView S(p.fdc, N);

    // explicit terms of the last stage (finish_explicit_assembly_Rdist,
    // mloop_fd.f90:78-110), which the skipped m-loop of that stage would do
    if (p.l_exp_calc[p.nstages - 1]) {
        const size_t qe = (size_t)(p.nstages - 1) * F + off;
        for (int k = tid; k < N; k += blockDim.x) {
            if (hT) { aT[k] = p.dVsT[off + k]; Tv[k] = p.temp[off + k]; }
            if (hX) { aX[k] = p.dVsXi[off + k]; Xv[k] = p.xi[off + k]; }
            aP[k] = p.dVsOm[off + k];
            usv[k] = p.us[off + k];
        }
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            const double o1 = R.or1(k);
            const cd us = usv[k];
            if (hT)
                p.T_expl[qe + k] = csub(csub(p.T_expl[qe + k], cscale(o1, fd_dr_at(S, aT, k, N))),
                                        cscale(R.dtcond(k) - p.tadvz * R.beta(k) * R.tcond(k), us));
            if (hX)
                p.X_expl[qe + k] = csub(csub(p.X_expl[qe + k], cscale(o1, fd_dr_at(S, aX, k, N))),
                                        cscale(R.dxicond(k), us));
            if (m != 0) {
                cd v = csub(p.P_expl[qe + k], cscale(o1, fd_dr_at(S, aP, k, N)));
                if (!p.l_coriolis_imp) v = cadd(v, cscale(p.CorFac * R.beta(k), us));
                if (hT) v = csub(v, cmul(mk(0.0, p.BuoFac * R.rgrav(k) * o1 * (double)m), Tv[k]));
                if (hX) v = csub(v, cmul(mk(0.0, p.ChemFac * R.rgrav(k) * o1 * (double)m), Xv[k]));
                p.P_expl[qe + k] = v;
            }
        }
        __syncthreads();
    }

    // assembled fields and right-hand sides with their wall rows
    const cd z = mk(0.0, 0.0);
    const double dm2 = (double)m * (double)m;
    for (int k = tid; k < N; k += blockDim.x) {
        const size_t q = off + k;
        if (hT) {
            cd v = assemble_elem(p, ss, p.T_old, p.T_impl, p.T_expl, q);
            if (m == 0) v.y = 0.0;
            if (k == 0 && p.ktopt == 1) v = fd_wall(p, p.topt, mi);
            if (k == N - 1 && p.kbott == 1) v = fd_wall(p, p.bott, mi);
            gT[k + 1] = v;
        }
        if (hX) {
            cd v = assemble_elem(p, ss, p.X_old, p.X_impl, p.X_expl, q);
            if (m == 0) v.y = 0.0;
            if (k == 0 && p.ktopxi == 1) v = fd_wall(p, p.topxi, mi);
            if (k == N - 1 && p.kbotxi == 1) v = fd_wall(p, p.botxi, mi);
            gX[k + 1] = v;
        }
        cd v = assemble_elem(p, ss, p.P_old, p.P_impl, p.P_expl, q);
        if (m == 0) {
            if ((k == 0 && p.ktopv != 1) || (k == N - 1 && p.kbotv != 1)) v = z;
            gP[k + 1] = v;
        } else {
            if (k == 0 || k == N - 1) v = z;
            gP[k + 2] = v;
            // the elliptic matrix row k (identity rows at the walls)
            const double c1 = R.beta(k) + R.or1(k);
            const bool wall = k == 0 || k == N - 1;
            eD[k] = wall ? 1.0 : -(S.ddr(k, 1) + c1 * S.dr(k, 1) + (R.dbeta(k) + R.beta(k) * R.or1(k) - dm2 * R.or2(k)));
            eU[k] = wall ? 0.0 : -(S.ddr(k, 2) + c1 * S.dr(k, 2));
            eL[k] = wall ? 0.0 : -(S.ddr(k, 0) + c1 * S.dr(k, 0));
        }
    }
    if (tid == 0) {
        gT[0] = z; gT[N + 1] = z; gX[0] = z; gX[N + 1] = z;
        if (m == 0) { gP[0] = z; gP[N + 1] = z; }
        else { gP[0] = z; gP[1] = z; gP[N + 2] = z; gP[N + 3] = z; }
    }
    __syncthreads();
    if (tid == 0 && m != 0) fd_lu3(N, eD, eL, eU);
    __syncthreads();
    if (tid < 64 && m != 0) fd_tri_solve_wave(N, eD, eL, eU, gP + 1);
    __syncthreads();
    if (tid == 0) {
        if (hT) fd_ghost_scalar(gT, N, p.ktopt, p.kbott, R, fd_wall(p, p.topt, mi), fd_wall(p, p.bott, mi));
        if (hX) fd_ghost_scalar(gX, N, p.ktopxi, p.kbotxi, R, fd_wall(p, p.topxi, mi), fd_wall(p, p.botxi, mi));
        fd_ghost_psi(p, R, m, gP);
    }
    __syncthreads();
    fd_implicit(p, R, S, mi, m, gT, gX, gP, 1, p.l_imp_calc_rhs[0], nullptr);
}

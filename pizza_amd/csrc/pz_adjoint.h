// Device code of libpizza_amd, part 5: the discrete adjoint of the IMEX
// step for the collocation and finite-difference methods and the multistep
// schemes.  The reference
// has no adjoint; every kernel here is the exact transpose of the forward
// kernel it names, so that <J dx, y> = <dx, J^T y> holds to rounding:
//   k_adj_probe       the forward's radial operators as dense matrices
//   k_mcoll_adj       transpose of k_mcoll (stage 1, multistep family)
//   k_radial_adj      transpose of k_radial_loop, linearised at the
//                     state the step started from
//   k_start_coll_adj  transpose of k_start_coll with respect to T
//   k_fd_mloop_adj    transpose of k_fd_mloop (stage 1, multistep family)
//   k_fd_start_adj    transpose of k_fd_start with respect to T
// Inner product: the real one, <a, b> = sum Re(conj(a) b) over every
// complex entry, so complex-linear maps transpose to their conjugates.
// Adjoint variables share the state layout (arena_layout), addressed
// through a PzDev view of their own.  Included once by pz_api.hip.
#pragma once
#include "pz_fd.h"

struct PzAdjDev {
    // dense operators, entry (i, j) = component i of the image of the unit
    // vector e_j, stored at [i N + j] (rows contiguous for the transposed
    // products, one thread per output):
    // C   costf (dct_fft), grid <-> Chebyshev coefficients
    // G   grid values from coefficients as dr_fft returns them
    // D1, D2  d/dr, d^2/dr^2 on the grid from coefficients (dr_fft)
    // E1  m = 0: grid values of the dcheb derivative series from coefficients
    // Uimp  m = 0: implicit term of u_phi from grid values (coll_uphi0_impl)
    const double *C, *G, *D1, *D2, *E1, *Uimp;
    const double *lnu; // d/dr at r_cmb of m = 0 grid values (Nusselt numbers)
    // kinetic energy weights per radius: pi times the rInt_R quadrature
    // weight times r h (u_phi) and r h + beta^2 r h^3 / 12 (u_s)
    const double *kus, *kup;
    // adjoints of the radial loop outputs T_expl(1), dVsT, xi, omega
    cd *eT0, *aT0, *eX0, *aX0, *eP0, *aP0;
    // accumulated adjoints of the boundary rows of the T and xi systems
    cd *gtop, *gbot, *gtopxi, *gbotxi;
    int *flag; // set to 1 when the window rebuilds its matrices
    // copies of the inverses the window's steps ran with: _s those of the
    // scheme's own weights, _b those of the start-up steps of a multistep
    // scheme bridged with CNAB2 (their implicit weight differs); the live
    // inverses may be rebuilt by a replay running beside the adjoint
    const double *T_inv_s, *X_inv_s, *U_inv_s, *T_inv_b, *X_inv_b, *U_inv_b;
    const cd *PC_inv_s, *PC_inv_b;
    // finite differences: copies of the factorised systems (fT, fX, fU, fP)
    const double *fT_s, *fX_s, *fU_s, *fT_b, *fX_b, *fU_b;
    const cd *fP_s, *fP_b;
    cd *gfs; // long grids: the forward grid columns by radius, 2 n_phi each
};

// out_f[j] (+)= sum_{i<n} M[j ld + i] y_f[i] for j < n: one wave per output
// column, lanes over i, so the column reads are coalesced.  Writers of
// output j are always the same wave, so successive calls may accumulate
// without a barrier; callers still synchronise before reading out.
template <int NF>
__device__ void matT(int n, int ld, const double *__restrict__ M, VecSet<const cd> y, VecSet<cd> out, bool acc)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j = w; j < n; j += nw) {
        const double *col = M + (size_t)j * ld;
        double sx[NF], sy[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) sx[f] = sy[f] = 0.0;
        for (int i = lane; i < n; i += 64) {
            const double a = col[i];
#pragma unroll
            for (int f = 0; f < NF; f++) {
                const cd v = y[f][i];
                sx[f] += a * v.x;
                sy[f] += a * v.y;
            }
        }
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const double rx = wave_sum(sx[f]), ry = wave_sum(sy[f]);
            if (lane == 0) out[f][j] = acc ? mk(out[f][j].x + rx, out[f][j].y + ry) : mk(rx, ry);
        }
    }
}

// out_f[j] (+)= sum_{i<n} M[i n + j] y_f[i] for j < n (the transpose of the
// n x n matrix M stored by rows): one thread per output, so the reads of a
// row are coalesced across the wave and y is a broadcast LDS read
template <int NF>
__device__ void matTr(int n, const double *__restrict__ M, VecSet<const cd> y, VecSet<cd> out, bool acc)
{
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double sx[NF], sy[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) sx[f] = sy[f] = 0.0;
#pragma unroll 8
        for (int i = 0; i < n; i++) {
            const double a = M[(size_t)i * n + j];
#pragma unroll
            for (int f = 0; f < NF; f++) {
                const cd v = y[f][i];
                sx[f] += a * v.x;
                sy[f] += a * v.y;
            }
        }
#pragma unroll
        for (int f = 0; f < NF; f++) out[f][j] = acc ? mk(out[f][j].x + sx[f], out[f][j].y + sy[f]) : mk(sx[f], sy[f]);
    }
}

// out[j] = sum_{i<rows} conj(M[j ld + i]) y[i] for j < ncol (complex M)
__device__ void matH(int rows, int ncol, int ld, const cd *__restrict__ M, const cd *y, cd *out)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j = w; j < ncol; j += nw) {
        const cd *col = M + (size_t)j * ld;
        double sx = 0.0, sy = 0.0;
        for (int i = lane; i < rows; i += 64) {
            const cd a = col[i], v = y[i];
            sx += a.x * v.x + a.y * v.y;
            sy += a.x * v.y - a.y * v.x;
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        if (lane == 0) out[j] = mk(sx, sy);
    }
}

// i s a for real s
__device__ __forceinline__ cd ci(double s, cd a) { return mk(-s * a.y, s * a.x); }

// ---------------------------------------------------------------------------
// The forward's radial operators applied to the unit vectors: block j
// writes column j of every matrix through the same device routines the
// step uses, so the transposes below are those of the step's own
// arithmetic.  Linear only without the rerror threshold (checked on the
// host).  LDS: 6 N + 6 L + 16 complex and 3 N doubles.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_probe(const PzDev *__restrict__ pd, double *C,
                                                   double *G, double *D1, double *D2, double *E1, double *Uimp,
                                                   double *lnu)
{
    const PzDev &p = *pd;
    extern __shared__ double smem[];
    const int j = blockIdx.x, N = p.N, tid = threadIdx.x;
    const RadView R(p.radc, N);
    cd *e = (cd *)smem, *c = e + N, *d1 = e + 2 * N, *d2 = e + 3 * N, *g = e + 4 * N, *t = e + 5 * N;
    cd *fbuf = e + 6 * N;
    double *dt3 = (double *)(fbuf + 6 * p.L + 16);
    // entry (k, j) at k N + j
    for (int k = tid; k < N; k += blockDim.x) e[k] = mk(k == j ? 1.0 : 0.0, 0.0);
    __syncthreads();
    dct_fft(p, VecSet<const cd>(e), VecSet<cd>(c), 1, 0, p.cosT, p.sinT, fbuf);
    __syncthreads();
    dr_fft(p, R, VecSet<const cd>(e), VecSet<cd>(d1), VecSet<cd>(d2), 1, p.cosT, p.sinT, fbuf, g);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        C[(size_t)k * N + j] = c[k].x;
        G[(size_t)k * N + j] = g[k].x;
        D1[(size_t)k * N + j] = d1[k].x;
        D2[(size_t)k * N + j] = d2[k].x;
    }
    __syncthreads();
    // d/dr at the outer wall of the grid vector e_j: costf, then get_dr
    dr_fft(p, R, VecSet<const cd>(c), VecSet<cd>(t), VecSet<cd>(), 1, p.cosT, p.sinT, fbuf);
    __syncthreads();
    if (tid == 0) lnu[j] = t[0].x;
    // m = 0: dcheb series of the coefficient vector e_j, back to the grid
    double *f = dt3, *df = dt3 + N;
    for (int k = tid; k < N; k += blockDim.x) f[k] = k == j ? 1.0 : 0.0;
    __syncthreads();
    if (tid == 0) dcheb_real(p, f, df);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) d1[k] = mk(df[k], 0.0);
    __syncthreads();
    dct_fft(p, VecSet<const cd>(d1), VecSet<cd>(t), 1, 0, p.cosT, p.sinT, fbuf);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) E1[(size_t)k * N + j] = t[k].x;
    __syncthreads();
    // m = 0 implicit term of the grid vector e_j
    coll_uphi0_impl(p, R, e, c, d1, dt3, p.cosT, p.sinT, fbuf, t, p.hdifV[0]);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) Uimp[(size_t)k * N + j] = t[k].x;
}

// Shifts the adjoints of one time array back through rotate_imex
// (multistep_schemes.f90:462-498) for element k: slot o of the step's
// input feeds slot o + 1 of its output; slot 0 of old and impl is
// rewritten by the step (impl only when lcalc), slot 0 of expl always.
__device__ __forceinline__ void unrotate(const PzDev &p, cd *old, cd *impl, cd *expl, size_t off, int k, int lcalc)
{
    const size_t F = (size_t)p.nm * p.N;
    for (int o = 0; o + 1 < p.nold; o++) old[o * F + off + k] = old[(o + 1) * F + off + k];
    old[(p.nold - 1) * F + off + k] = mk(0, 0);
    if (!lcalc && p.nimp > 1) impl[off + k] = cadd(impl[off + k], impl[F + off + k]);
    for (int o = lcalc ? 0 : 1; o + 1 < p.nimp; o++) impl[o * F + off + k] = impl[(o + 1) * F + off + k];
    if (lcalc || p.nimp > 1) impl[(p.nimp - 1) * F + off + k] = mk(0, 0);
    for (int o = 1; o + 1 < p.nexp; o++) expl[o * F + off + k] = expl[(o + 1) * F + off + k];
    if (p.nexp > 1) expl[(p.nexp - 1) * F + off + k] = mk(0, 0);
    expl[off + k] = mk(0, 0);
}

// the right-hand side rhs = sum wimp old + sum wimp_lin impl + sum wexp expl
// (set_imex_rhs, multistep_schemes.f90:366-413) transposed, for the slots
// other than the current explicit one
__device__ __forceinline__ void rhs_to_slots(const PzDev &p, const PzScheme *s, cd *old, cd *impl, cd *expl,
                                             size_t off, int k, cd rb)
{
    const size_t F = (size_t)p.nm * p.N;
    for (int o = 0; o < p.nold; o++) old[o * F + off + k] = cadd(old[o * F + off + k], cscale(s->wimp[o], rb));
    for (int o = 0; o < p.nimp; o++)
        impl[o * F + off + k] = cadd(impl[o * F + off + k], cscale(s->wimp_lin[o + 1], rb));
    for (int o = 1; o < p.nexp; o++) expl[o * F + off + k] = cadd(expl[o * F + off + k], cscale(s->wexp[o], rb));
}

#define ADJ_NVEC 34 // LDS vectors of k_mcoll_adj
// ---------------------------------------------------------------------------
// Transpose of mcoll_body (stage 1 of a multistep scheme) for one m: reads
// the adjoints of the step's outputs from lam (in place, row m), writes the
// adjoints of its inputs there, and the adjoints of the radial loop's
// outputs to ad.  fw holds the scheme state the step ran with (weights,
// step count); the m-loop is linear, so no field of fw is read.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(512) k_mcoll_adj(const PzDev *__restrict__ pfw, const PzDev *__restrict__ plam,
                                                   const PzAdjDev *__restrict__ pad)
{
    extern __shared__ double smem[];
    const PzDev &p = *pfw;
    const PzDev &L = *plam;
    const PzAdjDev &A = *pad;
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x, nb = blockDim.x;
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    const PzScheme *ss = p.ss;
    const RadView R(p.radc, N);
    const bool hT = p.l_heat, hX = p.l_chem;
    const int nc = p.ncheb, ldi = p.ldi;
    const double dm = (double)m, dm2 = dm * dm;
    const int lcalc = p.l_imp_calc_rhs[0] || (p.bridge && ss->n_steps + 1 <= p.nold - 1);
    const double wcur = ss->wexp[0], wl = ss->wimp_lin[0];
    // the CNAB2 start-up steps ran with their own matrices; the matrices
    // may change only at the first step and at the end of the start-up
    const bool brg = p.bridge && ss->n_steps <= p.nold - 1;
    const double *Tinv = brg ? A.T_inv_b : A.T_inv_s, *Xinv = brg ? A.X_inv_b : A.X_inv_s;
    const double *Uinv = brg ? A.U_inv_b : A.U_inv_s;
    const cd *PCinv = brg ? A.PC_inv_b : A.PC_inv_s;
    if (mi == 0 && tid == 0 &&
        (ss->l_new_dt || (ss->lMat && ss->n_steps > 1 && !(p.bridge && ss->n_steps <= p.nold))))
        *A.flag = 1;

    cd *V = (cd *)smem;
    cd *gT = V, *iT = V + N, *d1T = V + 2 * N, *d2T = V + 3 * N, *xT = V + 4 * N, *rT = V + 5 * N, *eT = V + 6 * N;
    cd *gX = V + 7 * N, *iX = V + 8 * N, *d1X = V + 9 * N, *d2X = V + 10 * N, *xX = V + 11 * N, *rX = V + 12 * N;
    cd *eX = V + 13 * N;
    cd *Lom = V + 14 * N, *Lpo = V + 15 * N, *Lps = V + 16 * N, *Lus = V + 17 * N, *Lup = V + 18 * N;
    cd *iP = V + 19 * N, *d1w = V + 20 * N, *d2w = V + 21 * N, *dps = V + 22 * N, *Lw = V + 23 * N;
    cd *zc = V + 24 * N; // 2 N: [omega; psi] coefficients
    cd *rP = V + 26 * N, *eP = V + 27 * N;
    cd *Lus0 = V + 28 * N, *LT0 = V + 29 * N, *LX0 = V + 30 * N; // adjoints of the step's input fields
    cd *t1 = V + 31 * N, *t2 = V + 32 * N, *t3 = V + 33 * N;

    // ---- adjoints of the step's outputs; time arrays shifted back -------
    for (int k = tid; k < N; k += nb) {
        if (hT) {
            gT[k] = cadd(L.temp[off + k], L.T_old[off + k]);
            iT[k] = lcalc ? L.T_impl[off + k] : mk(0, 0);
            d1T[k] = lcalc ? L.dtemp[off + k] : mk(0, 0);
            eT[k] = L.T_expl[off + k];
            if (p.nexp > 1) eT[k] = cadd(eT[k], L.T_expl[F + off + k]);
            unrotate(p, L.T_old, L.T_impl, L.T_expl, off, k, lcalc);
        }
        if (hX) {
            gX[k] = cadd(L.xi[off + k], L.X_old[off + k]);
            iX[k] = lcalc ? L.X_impl[off + k] : mk(0, 0);
            d1X[k] = mk(0, 0);
            eX[k] = L.X_expl[off + k];
            if (p.nexp > 1) eX[k] = cadd(eX[k], L.X_expl[F + off + k]);
            unrotate(p, L.X_old, L.X_impl, L.X_expl, off, k, lcalc);
        }
        Lom[k] = L.om[off + k];
        Lpo[k] = L.P_old[off + k];
        Lps[k] = L.psi[off + k];
        Lus[k] = L.us[off + k];
        Lup[k] = L.up[off + k];
        iP[k] = lcalc ? L.P_impl[off + k] : mk(0, 0);
        eP[k] = L.P_expl[off + k];
        if (p.nexp > 1) eP[k] = cadd(eP[k], L.P_expl[F + off + k]);
        unrotate(p, L.P_old, L.P_impl, L.P_expl, off, k, lcalc);
        Lus0[k] = mk(0, 0);
        LT0[k] = mk(0, 0);
        LX0[k] = mk(0, 0);
    }
    __syncthreads();

    // ---- psi / omega / u_phi (before T: the psi system reads the new T) --
    if (m == 0) {
        // u0 = up = Re(C f); om = drx Re(E1 f) + or1 u0; old = u0; impl = Uimp u0
        for (int k = tid; k < N; k += nb) {
            const double lo = Lom[k].x;
            t1[k] = mk(Lup[k].x + Lpo[k].x + R.or1(k) * lo, 0.0); // adjoint of u0 (grid)
            t2[k] = mk(R.drx(k) * lo, 0.0);                       // adjoint of E1 f
            t3[k] = mk(iP[k].x, 0.0);
        }
        __syncthreads();
        if (lcalc) {
            matTr<1>(N, A.Uimp, VecSet<const cd>(t3), VecSet<cd>(t1), true);
            __syncthreads();
        }
        matTr<1>(N, A.C, VecSet<const cd>(t1), VecSet<cd>(zc), false);
        matTr<1>(N, A.E1, VecSet<const cd>(t2), VecSet<cd>(zc), true);
        __syncthreads();
        for (int k = nc + tid; k < N; k += nb) zc[k] = mk(0, 0);
        __syncthreads();
        matT<1>(N, ldi, Uinv, VecSet<const cd>(zc), VecSet<cd>(rP), false);
        __syncthreads();
        for (int k = tid; k < N; k += nb) {
            const cd rb = (k == 0 || k == N - 1) ? mk(0, 0) : mk(rP[k].x, 0.0);
            eP[k] = cadd(eP[k], cscale(wcur, rb));
            rhs_to_slots(p, ss, L.P_old, L.P_impl, L.P_expl, off, k, rb);
            A.eP0[off + k] = eP[k];
            A.aP0[off + k] = mk(0, 0);
        }
    } else {
        const double VH = p.ViscFac * p.hdifV[mi];
        for (int k = tid; k < N; k += nb) {
            const double o1 = R.or1(k);
            cd lw = cadd(Lom[k], Lpo[k]), lus = Lus[k], lup = Lup[k];
            cd lps = Lps[k];
            cd a1 = mk(0, 0), a2 = mk(0, 0);
            if (lcalc) {
                const cd v = iP[k];
                const double ekp = p.CorFac * R.ekpump(k);
                a2 = cscale(VH, v);
                a1 = cscale(VH * o1, v);
                lw = cadd(lw, cscale(-VH * dm2 * R.or2(k) - ekp, v));
                lup = cadd(lup, cscale(-ekp * R.ekp_up(k), v));
                // us coefficient: -ekp (ekp_us + i m ekp_dusdp) (+ CorFac beta)
                const double cr = -ekp * R.ekp_us(k) + (p.l_coriolis_imp ? p.CorFac * R.beta(k) : 0.0);
                const double cim = -ekp * dm * R.ekp_dusdp(k);
                lus = cadd(lus, mk(cr * v.x + cim * v.y, cr * v.y - cim * v.x));
                if (p.l_buo_imp) {
                    if (hT) gT[k] = cadd(gT[k], ci(dm * R.rgrav(k) * o1 * p.BuoFac, v));
                    if (hX) gX[k] = cadd(gX[k], ci(dm * R.rgrav(k) * o1 * p.ChemFac, v));
                }
            }
            // us = i m or1 psi, up = -dpsi - beta psi
            lps = cadd(lps, ci(-dm * o1, lus));
            lps = csub(lps, cscale(R.beta(k), lup));
            Lw[k] = lw;
            Lps[k] = lps;
            dps[k] = mk(-lup.x, -lup.y);
            d1w[k] = a1;
            d2w[k] = a2;
        }
        __syncthreads();
        cd *zo = zc, *zp = zc + N;
        matTr<2>(N, A.D1, VecSet<const cd>(dps, d1w), VecSet<cd>(zp, zo), false);
        __syncthreads();
        matTr<1>(N, A.G, VecSet<const cd>(Lps), VecSet<cd>(zp), true);
        if (lcalc) {
            matTr<1>(N, A.G, VecSet<const cd>(Lw), VecSet<cd>(zo), true);
            matTr<1>(N, A.D2, VecSet<const cd>(d2w), VecSet<cd>(zo), true);
        } else {
            matTr<1>(N, A.C, VecSet<const cd>(Lw), VecSet<cd>(zo), false);
        }
        __syncthreads();
        for (int k = nc + tid; k < N; k += nb) { zo[k] = mk(0, 0); zp[k] = mk(0, 0); }
        __syncthreads();
        matH(2 * N, N, 2 * N, PCinv + off * 2 * N, zc, rP);
        __syncthreads();
        for (int k = tid; k < N; k += nb) {
            const double o1 = R.or1(k);
            const cd rb = (k == 0 || k == N - 1) ? mk(0, 0) : rP[k];
            if (p.l_buo_imp) {
                if (hT) gT[k] = cadd(gT[k], ci(wl * dm * p.BuoFac * R.rgrav(k) * o1, rb));
                if (hX) gX[k] = cadd(gX[k], ci(wl * dm * p.ChemFac * R.rgrav(k) * o1, rb));
            }
            const cd e = cadd(eP[k], cscale(wcur, rb));
            rhs_to_slots(p, ss, L.P_old, L.P_impl, L.P_expl, off, k, rb);
            A.eP0[off + k] = e;
            t1[k] = cscale(-o1, e);
            if (!p.l_coriolis_imp) Lus0[k] = cadd(Lus0[k], cscale(p.CorFac * R.beta(k), e));
            if (!p.l_buo_imp) {
                if (hT) LT0[k] = cadd(LT0[k], ci(dm * p.BuoFac * R.rgrav(k) * o1, e));
                if (hX) LX0[k] = cadd(LX0[k], ci(dm * p.ChemFac * R.rgrav(k) * o1, e));
            }
        }
        __syncthreads();
        matTr<1>(N, A.D1, VecSet<const cd>(t1), VecSet<cd>(t2), false);
        __syncthreads();
        matTr<1>(N, A.C, VecSet<const cd>(t2), VecSet<cd>(A.aP0 + off), false);
    }
    __syncthreads();

    // ---- T and xi ----------------------------------------------------------
    if (hT || hX) {
        const double kT = p.TdiffFac * p.hdifT[mi], kX = p.XidiffFac * p.hdifXi[mi];
        for (int k = tid; k < N; k += nb) {
            const double o1 = R.or1(k), o2 = R.or2(k);
            if (hT && lcalc) {
                const cd v = iT[k];
                d2T[k] = cscale(kT, v);
                d1T[k] = cadd(d1T[k], cscale(kT * o1, v));
                gT[k] = csub(gT[k], cscale(kT * dm2 * o2, v));
            }
            if (hX && lcalc) {
                const cd v = iX[k];
                d2X[k] = cscale(kX, v);
                d1X[k] = cscale(kX * o1, v);
                gX[k] = csub(gX[k], cscale(kX * dm2 * o2, v));
            }
        }
        __syncthreads();
        if (lcalc) {
            if (hT && hX) {
                matTr<2>(N, A.G, VecSet<const cd>(gT, gX), VecSet<cd>(xT, xX), false);
                matTr<2>(N, A.D1, VecSet<const cd>(d1T, d1X), VecSet<cd>(xT, xX), true);
                matTr<2>(N, A.D2, VecSet<const cd>(d2T, d2X), VecSet<cd>(xT, xX), true);
            } else {
                cd *g = hT ? gT : gX, *a = hT ? d1T : d1X, *b = hT ? d2T : d2X, *x = hT ? xT : xX;
                matTr<1>(N, A.G, VecSet<const cd>(g), VecSet<cd>(x), false);
                matTr<1>(N, A.D1, VecSet<const cd>(a), VecSet<cd>(x), true);
                matTr<1>(N, A.D2, VecSet<const cd>(b), VecSet<cd>(x), true);
            }
        } else {
            if (hT) matTr<1>(N, A.C, VecSet<const cd>(gT), VecSet<cd>(xT), false);
            if (hX) matTr<1>(N, A.C, VecSet<const cd>(gX), VecSet<cd>(xX), false);
        }
        __syncthreads();
        for (int k = nc + tid; k < N; k += nb) { xT[k] = mk(0, 0); xX[k] = mk(0, 0); }
        __syncthreads();
        if (hT) matT<1>(N, ldi, Tinv + off * ldi, VecSet<const cd>(xT), VecSet<cd>(rT), false);
        if (hX) matT<1>(N, ldi, Xinv + off * ldi, VecSet<const cd>(xX), VecSet<cd>(rX), false);
        __syncthreads();
        if (tid == 0) {
            if (hT) { A.gtop[mi] = cadd(A.gtop[mi], rT[0]); A.gbot[mi] = cadd(A.gbot[mi], rT[N - 1]); }
            if (hX) { A.gtopxi[mi] = cadd(A.gtopxi[mi], rX[0]); A.gbotxi[mi] = cadd(A.gbotxi[mi], rX[N - 1]); }
        }
        for (int k = tid; k < N; k += nb) {
            const double o1 = R.or1(k);
            if (hT) {
                const cd rb = (k == 0 || k == N - 1) ? mk(0, 0) : rT[k];
                const cd e = cadd(eT[k], cscale(wcur, rb));
                rhs_to_slots(p, ss, L.T_old, L.T_impl, L.T_expl, off, k, rb);
                A.eT0[off + k] = e;
                t1[k] = cscale(-o1, e);
                Lus0[k] = csub(Lus0[k], cscale(R.dtcond(k) - p.tadvz * R.beta(k) * R.tcond(k), e));
            }
            if (hX) {
                const cd rb = (k == 0 || k == N - 1) ? mk(0, 0) : rX[k];
                const cd e = cadd(eX[k], cscale(wcur, rb));
                rhs_to_slots(p, ss, L.X_old, L.X_impl, L.X_expl, off, k, rb);
                A.eX0[off + k] = e;
                t3[k] = cscale(-o1, e);
                Lus0[k] = csub(Lus0[k], cscale(R.dxicond(k), e));
            }
        }
        __syncthreads();
        if (hT) matTr<1>(N, A.D1, VecSet<const cd>(t1), VecSet<cd>(t2), false);
        if (hX) matTr<1>(N, A.D1, VecSet<const cd>(t3), VecSet<cd>(xX), false);
        __syncthreads();
        if (hT) matTr<1>(N, A.C, VecSet<const cd>(t2), VecSet<cd>(A.aT0 + off), false);
        if (hX) matTr<1>(N, A.C, VecSet<const cd>(xX), VecSet<cd>(A.aX0 + off), false);
    }
    __syncthreads();
    // ---- adjoints of the step's input fields (the radial loop adds its own)
    for (int k = tid; k < N; k += nb) {
        L.us[off + k] = Lus0[k];
        L.up[off + k] = mk(0, 0);
        L.om[off + k] = mk(0, 0);
        L.psi[off + k] = mk(0, 0);
        L.temp[off + k] = LT0[k];
        L.xi[off + k] = LX0[k];
        if (lcalc) L.dtemp[off + k] = mk(0, 0); // rewritten only with the implicit terms
    }
}

// Hermitian spectrum of z = a + i b for the complex FFT of two real grid
// columns (the packing of the radial loop): A, B hold m < nm, B may be null
__device__ __forceinline__ cd herm_pack(const cd *A, const cd *B, int k, int n, int nm)
{
    cd z = mk(0, 0);
    if (k == 0) {
        z = mk(A[0].x, B ? B[0].x : 0.0);
    } else if (k < nm) {
        const cd a = A[k], b = B ? B[k] : mk(0, 0);
        z = mk(a.x - b.y, a.y + b.x);
    } else if (n - k < nm) {
        const int kk = n - k;
        const cd a = cconj(A[kk]), b = B ? cconj(B[kk]) : mk(0, 0);
        z = mk(a.x - b.y, a.y + b.x);
    }
    return z;
}

// ---------------------------------------------------------------------------
// Transpose of radial_loop_body, linearised at the state the step started
// from (fw), for one radius: adjoints of the products from those of the
// explicit terms and dVs fields (ad), back through the FFTs to the spectra;
// adds to the adjoints of us, up, omega, T, xi in lam.
// LDS: (11 nm + 9 n) complex + 16 doubles.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_radial_adj(const PzDev *__restrict__ pfw, const PzDev *__restrict__ plam,
                                                    const PzAdjDev *__restrict__ pad)
{
    extern __shared__ double smem[];
    const PzDev &p = *pfw;
    const PzDev &L = *plam;
    const PzAdjDev &A = *pad;
    const int nr = blockIdx.x, n = p.nphi, nm = p.nm, N = p.N, tid = threadIdx.x, nb = blockDim.x;
    cd *spec = (cd *)smem;  // 5 nm: forward spectra us, up, om, T, xi
    cd *sb = spec + 5 * nm; // 6 nm: adjoint spectra (A, B) of each product
    cd *gf = sb + 6 * nm;   // 3 n: forward grid columns
    cd *bufa = gf + 3 * n, *bufb = bufa + 3 * n;
    const bool hT = p.l_heat, hX = p.l_chem;
    const int nin = hX ? 3 : 2, nout = 1 + (hT ? 1 : 0) + (hX ? 1 : 0);
    for (int i = tid; i < nm; i += nb) {
        const size_t q = (size_t)i * N + nr;
        spec[i] = p.us[q];
        spec[nm + i] = p.up[q];
        spec[2 * nm + i] = p.om[q];
        spec[3 * nm + i] = hT ? p.temp[q] : mk(0, 0);
        spec[4 * nm + i] = hX ? p.xi[q] : mk(0, 0);
    }
    const double or1 = p.or1[nr], rr = p.r[nr];
    const double fbeta = (1.0 - p.tadvz) * p.beta[nr] * or1;
    // adjoint spectra of the products, halved above m = 0 (Hermitian halves)
    for (int t = tid; t < nout * nm; t += nb) {
        const int tr = t / nm, i = t - tr * nm;
        const int m = p.idx2m[i];
        const size_t q = (size_t)i * N + nr;
        cd a, b;
        if (tr == 0) {
            a = m == 0 ? mk(0, 0) : ci(or1 * m, A.eP0[q]);
            b = A.aP0[q];
        } else if (tr == 1 && hT) {
            const cd e = A.eT0[q];
            a = ci(or1 * m, e);
            b = csub(A.aT0[q], cscale(fbeta, e));
        } else {
            a = ci(or1 * m, A.eX0[q]);
            b = A.aX0[q];
        }
        const double h = i == 0 ? 1.0 : 0.5;
        sb[2 * tr * nm + i] = cscale(h, a);
        sb[(2 * tr + 1) * nm + i] = cscale(h, b);
    }
    __syncthreads();
    // forward grid columns (us, up), (om, T), (xi, 0)
    for (int t = tid; t < nin * n; t += nb) {
        const int tr = t / n, k = t - tr * n;
        const cd *Bs = (2 * tr + 1) < 5 ? spec + (2 * tr + 1) * nm : nullptr;
        bufa[t] = herm_pack(spec + 2 * tr * nm, Bs, k, n, nm);
    }
    __syncthreads();
    cd *g = fft_lds(p, bufa, bufb, n, nin, +1);
    for (int t = tid; t < nin * n; t += nb) gf[t] = g[t];
    __syncthreads();
    // adjoints of the product columns: (1/n) Re sum_m Abar_m e^{+i m phi}
    for (int t = tid; t < nout * n; t += nb) {
        const int tr = t / n, k = t - tr * n;
        bufa[t] = herm_pack(sb + 2 * tr * nm, sb + (2 * tr + 1) * nm, k, n, nm);
    }
    __syncthreads();
    cd *s = fft_lds(p, bufa, bufb, n, nout, +1);
    cd *o = (s == bufa) ? bufb : bufa;
    const double inv_n = 1.0 / (double)n;
    const int tX = hT ? 2 : 1;
    for (int j = tid; j < n; j += nb) {
        const double us = gf[j].x, up = gf[j].y, om = gf[n + j].x, T = gf[n + j].y;
        const double X = hX ? gf[2 * n + j].x : 0.0;
        const double ao = s[j].x * inv_n, bo = s[j].y * inv_n;
        const double aT = hT ? s[n + j].x * inv_n : 0.0, bT = hT ? s[n + j].y * inv_n : 0.0;
        const double aX = hX ? s[tX * n + j].x * inv_n : 0.0, bX = hX ? s[tX * n + j].y * inv_n : 0.0;
        const double lup = ao * om + aT * T + aX * X;
        const double lus = rr * (bo * om + bT * T + bX * X);
        const double lom = ao * up + bo * rr * us;
        const double lT = aT * up + bT * rr * us;
        const double lX = aX * up + bX * rr * us;
        o[j] = mk(lus, lup);
        o[n + j] = mk(lom, lT);
        if (hX) o[2 * n + j] = mk(lX, 0.0);
    }
    __syncthreads();
    cd *other = (o == bufa) ? bufb : bufa;
    cd *Z = fft_lds(p, o, other, n, nin, -1);
    // back to the spectra: 2 sum_j g_j e^{-i m phi_j} above m = 0, the real
    // part at m = 0; plus the Reynolds-stress sum (radial_loop.f90:267-278)
    const double lsom = -A.eP0[nr].x; // P_expl(m = 0) = -usom
    const double cfac = 0.5 * p.ek * p.CorFac * p.ekpump[nr];
    for (int t = tid; t < nin * nm; t += nb) {
        const int tr = t / nm, i = t - tr * nm;
        const cd zk = Z[tr * n + i], zc = cconj(Z[tr * n + (i == 0 ? 0 : n - i)]);
        cd a = mk(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        cd b = mk(0.5 * (zk.y - zc.y), -0.5 * (zk.x - zc.x));
        if (i == 0) { a = mk(a.x, 0.0); b = mk(b.x, 0.0); }
        else { a = cscale(2.0, a); b = cscale(2.0, b); }
        const size_t q = (size_t)i * N + nr;
        if (tr == 0) {
            if (i == 0) b = mk(b.x + lsom * cfac * spec[2 * nm].x, b.y);
            else a = cadd(a, cscale(2.0 * lsom, spec[2 * nm + i]));
            L.us[q] = cadd(L.us[q], a);
            L.up[q] = cadd(L.up[q], b);
        } else if (tr == 1) {
            if (i == 0) a = mk(a.x + lsom * cfac * spec[nm].x, a.y);
            else a = cadd(a, cscale(2.0 * lsom, spec[i]));
            L.om[q] = cadd(L.om[q], a);
            if (hT) L.temp[q] = cadd(L.temp[q], b);
        } else {
            L.xi[q] = cadd(L.xi[q], a);
        }
    }
}

// ---------------------------------------------------------------------------
// Transpose of radial_loop_inplace_body (long azimuthal grids without
// composition), linearised at fw, for one radius; the steps of k_radial_adj
// with the spectra by radius from k_gather_r (run on fw, into fw's rrow)
// and the transforms of fft_inplace (two at a time) in one pair of LDS
// columns: the forward grid (us, up), (omega, T) is transformed there and
// parked in ad.gfs (2 n_phi per radius, read back by the same workgroup),
// then the adjoint product columns take the LDS.  67 KB of LDS and 512
// threads within 128 registers: two workgroups share a CU, as in the
// forward.
// ---------------------------------------------------------------------------
#define PZ_RADIAL_ADJ_THREADS 512
__global__ void __launch_bounds__(PZ_RADIAL_ADJ_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
k_radial_adj_inplace(const PzDev *__restrict__ pfw, const PzDev *__restrict__ plam, const PzAdjDev *__restrict__ pad)
{
    extern __shared__ double smem[];
    const PzDev &p = *pfw;
    const PzDev &L = *plam;
    const PzAdjDev &A = *pad;
    const int nr = blockIdx.x, n = p.nphi, nm = p.nm, N = p.N, tid = threadIdx.x, nb = blockDim.x;
    const size_t F = (size_t)nm * N;
    cd *sa = (cd *)smem, *t1 = sa + 2 * n, *t2 = t1 + 64;
    const int n2 = (n + 63) >> 6;
    cd *om = t2 + n2;
    cd *G = A.gfs + (size_t)nr * 2 * n;
    const bool hT = p.l_heat;
    const cd *Rus = p.rrow + (size_t)nr * nm, *Rup = Rus + F, *Rom = Rus + 2 * F, *RT = Rus + 3 * F;
    for (int k = tid; k < 64 + n2; k += nb) t1[k] = p.twid[k < 64 ? k : 64 * (k - 64)];
    auto fwd = [=](int k) { return cmul(t1[k & 63], t2[k >> 6]); };
    auto bwd = [=](int k) {
        const cd w = cmul(t1[k & 63], t2[k >> 6]);
        return mk(w.x, -w.y);
    };
    // the forward grid columns (us, up), (om, T), parked in G
    for (int t = tid; t < 2 * n; t += nb) {
        const int tr = t >= n, k = t - tr * n;
        sa[t] = herm_pack(tr == 0 ? Rus : Rom, tr == 0 ? Rup : (hT ? RT : nullptr), k, n, nm);
    }
    __syncthreads();
    fft_inplace(sa, n, 2, +1, p.radices, p.nrad, bwd, om);
    for (int t = tid; t < 2 * n; t += nb) G[t] = sa[t];
    __syncthreads();
    // adjoint product spectra (halved above m = 0) packed the same way
    {
        const double or1 = p.or1[nr], fbeta = (1.0 - p.tadvz) * p.beta[nr] * or1;
        for (int t = tid; t < 2 * n; t += nb) {
            const int tr = t >= n, k = t - tr * n;
            const bool lo = k < nm, hi = !lo && n - k < nm;
            cd z = mk(0, 0);
            if (lo || hi) {
                const int i = lo ? k : n - k, m = p.idx2m[i];
                const size_t q = (size_t)i * N + nr;
                cd a, b;
                if (tr == 0) {
                    a = m == 0 ? mk(0, 0) : ci(or1 * m, A.eP0[q]);
                    b = A.aP0[q];
                } else if (hT) {
                    const cd e = A.eT0[q];
                    a = ci(or1 * m, e);
                    b = csub(A.aT0[q], cscale(fbeta, e));
                } else {
                    a = b = mk(0, 0);
                }
                const double h = i == 0 ? 1.0 : 0.5;
                a = cscale(h, a);
                b = cscale(h, b);
                if (k == 0) z = mk(a.x, b.x);
                else if (lo) z = mk(a.x - b.y, a.y + b.x);
                else z = mk(a.x + b.y, b.x - a.y);
            }
            sa[t] = z;
        }
    }
    __syncthreads();
    fft_inplace(sa, n, 2, +1, p.radices, p.nrad, bwd, om);
    {
        const double rr = p.r[nr], inv_n = 1.0 / (double)n;
        for (int j = tid; j < n; j += nb) {
            const cd g0 = G[j], g1 = G[n + j];
            const double us = g0.x, up = g0.y, omv = g1.x, T = g1.y;
            const double ao = sa[j].x * inv_n, bo = sa[j].y * inv_n;
            const double aT = hT ? sa[n + j].x * inv_n : 0.0, bT = hT ? sa[n + j].y * inv_n : 0.0;
            sa[j] = mk(rr * (bo * omv + bT * T), ao * omv + aT * T);
            sa[n + j] = mk(ao * up + bo * rr * us, aT * up + bT * rr * us);
        }
    }
    __syncthreads();
    fft_inplace(sa, n, 2, -1, p.radices, p.nrad, fwd, om);
    const double lsom = -A.eP0[nr].x; // P_expl(m = 0) = -usom
    const double cfac = 0.5 * p.ek * p.CorFac * p.ekpump[nr];
    for (int t = tid; t < 2 * nm; t += nb) {
        const int tr = t >= nm, i = t - tr * nm;
        const cd zk = sa[tr * n + i], zc = cconj(sa[tr * n + (i == 0 ? 0 : n - i)]);
        cd a = mk(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        cd b = mk(0.5 * (zk.y - zc.y), -0.5 * (zk.x - zc.x));
        if (i == 0) { a = mk(a.x, 0.0); b = mk(b.x, 0.0); }
        else { a = cscale(2.0, a); b = cscale(2.0, b); }
        const size_t q = (size_t)i * N + nr;
        if (tr == 0) {
            if (i == 0) b = mk(b.x + lsom * cfac * Rom[0].x, b.y);
            else a = cadd(a, cscale(2.0 * lsom, Rom[i]));
            L.us[q] = cadd(L.us[q], a);
            L.up[q] = cadd(L.up[q], b);
        } else {
            if (i == 0) a = mk(a.x + lsom * cfac * Rup[0].x, a.y);
            else a = cadd(a, cscale(2.0 * lsom, Rus[i]));
            L.om[q] = cadd(L.om[q], a);
            if (hT) L.temp[q] = cadd(L.temp[q], b);
        }
    }
}

// adjoint seeds of the objectives at the state after a step
// NuTop: lam(T, m = 0) += w lnu (real)
__global__ void __launch_bounds__(256) k_adj_seed_nu(const PzDev *__restrict__ plam, const PzAdjDev *__restrict__ pad,
                                                     double w)
{
    const PzDev &L = *plam;
    for (int k = threadIdx.x; k < L.N; k += blockDim.x) L.temp[k] = mk(L.temp[k].x + w * pad->lnu[k], L.temp[k].y);
}
// weights on T: lam(T) += W (device array in the field layout)
__global__ void __launch_bounds__(256) k_adj_seed_w(const PzDev *__restrict__ plam, const cd *__restrict__ W)
{
    const PzDev &L = *plam;
    const size_t F = (size_t)L.nm * L.N;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < F; q += (size_t)gridDim.x * blockDim.x)
        L.temp[q] = cadd(L.temp[q], W[q]);
}
// forward value of the objective, accumulated over the steps in acc[0]:
// NuTop: sum_k lnu_k Re T(m = 0, r_k); weights: sum Re(conj(W) T)
__global__ void __launch_bounds__(256) k_obj_acc(const PzDev *__restrict__ pd, const PzAdjDev *__restrict__ pad,
                                                 const cd *__restrict__ W, double *acc)
{
    __shared__ double red[8];
    const PzDev &p = *pd;
    double v = 0.0;
    if (W) {
        const size_t F = (size_t)p.nm * p.N;
        for (size_t q = threadIdx.x; q < F; q += blockDim.x) v += W[q].x * p.temp[q].x + W[q].y * p.temp[q].y;
    } else {
        for (int k = threadIdx.x; k < p.N; k += blockDim.x) v += pad->lnu[k] * p.temp[k].x;
    }
    v = block_reduce(v, red, 0);
    if (threadIdx.x == 0) acc[0] += v;
}

// kinetic energy of a state (outputs.f90:552-570), us2_3D + up2_3D + uz2_3D
// = sum over m and r of c_m (kus_r |u_s|^2 + kup_r |u_phi|^2), c_0 = 1 on
// the real parts only, c_m = 2 for m >= 1: one workgroup per m writes its
// row's sum to acc[1 + mi]; k_obj_fin adds the rows, in order, to acc[0]
__global__ void __launch_bounds__(256) k_obj_ke(const PzDev *__restrict__ pd, const PzAdjDev *__restrict__ pad,
                                                double *acc)
{
    __shared__ double red[8];
    const PzDev &p = *pd;
    const int mi = blockIdx.x, N = p.N;
    const size_t off = (size_t)mi * N;
    const bool m0 = p.idx2m[mi] == 0;
    double v = 0.0;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        const cd u = p.us[off + k], w = p.up[off + k];
        if (m0) v += pad->kus[k] * u.x * u.x + pad->kup[k] * w.x * w.x;
        else v += 2.0 * (pad->kus[k] * (u.x * u.x + u.y * u.y) + pad->kup[k] * (w.x * w.x + w.y * w.y));
    }
    v = block_reduce(v, red, 0);
    if (threadIdx.x == 0) acc[1 + mi] = v;
}
__global__ void __launch_bounds__(256) k_obj_fin(int nm, double *acc)
{
    __shared__ double red[8];
    double v = 0.0;
    for (int i = threadIdx.x; i < nm; i += blockDim.x) v += acc[1 + i];
    v = block_reduce(v, red, 0);
    if (threadIdx.x == 0) acc[0] += v;
}
// w times the derivative of that energy at the state (us, up):
// lam(u_s) += 2 w c_m kus_r u_s, lam(u_phi) += 2 w c_m kup_r u_phi (no
// imaginary parts at m = 0)
__global__ void __launch_bounds__(256) k_adj_seed_ke(const cd *__restrict__ us, const cd *__restrict__ up,
                                                     const PzDev *__restrict__ plam,
                                                     const PzAdjDev *__restrict__ pad, double w)
{
    const PzDev &L = *plam;
    const size_t F = (size_t)L.nm * L.N;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < F; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q / L.N), k = (int)(q % L.N);
        const bool m0 = L.idx2m[i] == 0;
        const double c = m0 ? 2.0 * w : 4.0 * w, a = c * pad->kus[k], b = c * pad->kup[k];
        const cd ls = L.us[q], lp = L.up[q], u = us[q], v = up[q];
        L.us[q] = mk(ls.x + a * u.x, m0 ? ls.y : ls.y + a * u.y);
        L.up[q] = mk(lp.x + b * v.x, m0 ? lp.y : lp.y + b * v.y);
    }
}

// ---------------------------------------------------------------------------
// Transpose of k_start_coll with respect to T: the initial T field feeds
// temp, T_old(1), T_impl(1) and, with implicit buoyancy, P_impl(1).
// grad = lam(temp) + lam(T_old) + L^T lam(T_impl) + buoyancy^T lam(P_impl).
// LDS: 6 N complex.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_start_coll_adj(const PzDev *__restrict__ pd, const PzDev *__restrict__ plam,
                                                        const PzAdjDev *__restrict__ pad, cd *grad)
{
    extern __shared__ double smem[];
    const PzDev &p = *pd;
    const PzDev &L = *plam;
    const PzAdjDev &A = *pad;
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x, nb = blockDim.x;
    const size_t off = (size_t)mi * N;
    const RadView R(p.radc, N);
    const double dm = (double)m, kT = p.TdiffFac * p.hdifT[mi];
    cd *g = (cd *)smem, *a1 = g + N, *a2 = g + 2 * N, *c = g + 3 * N, *t = g + 4 * N;
    for (int k = tid; k < N; k += nb) {
        const cd v = L.T_impl[off + k];
        cd gv = cadd(L.temp[off + k], L.T_old[off + k]);
        gv = csub(gv, cscale(kT * dm * dm * R.or2(k), v));
        if (p.l_buo_imp && m != 0) gv = cadd(gv, ci(dm * R.rgrav(k) * R.or1(k) * p.BuoFac, L.P_impl[off + k]));
        g[k] = gv;
        a1[k] = cscale(kT * R.or1(k), v);
        a2[k] = cscale(kT, v);
    }
    __syncthreads();
    matTr<1>(N, A.D1, VecSet<const cd>(a1), VecSet<cd>(c), false);
    matTr<1>(N, A.D2, VecSet<const cd>(a2), VecSet<cd>(c), true);
    __syncthreads();
    matTr<1>(N, A.C, VecSet<const cd>(c), VecSet<cd>(t), false);
    __syncthreads();
    for (int k = tid; k < N; k += nb) grad[off + k] = cadd(g[k], t[k]);
}

// ---------------------------------------------------------------------------
// Finite differences
// ---------------------------------------------------------------------------
// sum_k D(k, j) v[k]: the transpose of fd_dr_at (one-sided at the walls)
__device__ __forceinline__ cd fd_drT_at(const FdView &S, const cd *v, int j, int N)
{
    cd s = mk(0, 0);
    for (int k = max(1, j - 1); k <= min(N - 2, j + 1); k++) s = cadd(s, cscale(S.dr(k, j - k + 1), v[k]));
    if (j <= 2) s = cadd(s, cscale(S.top(j), v[0]));
    if (j >= N - 3) s = cadd(s, cscale(S.bot(N - 1 - j), v[N - 1]));
    return s;
}

// transposed substitutions of the factorised systems (the forward ones are
// fd_tri_solve_wave / fd_penta_solve_wave), on x[0..N), by one whole wave:
// lane l owns rows [l C, l C + C), C = ceil(N / 64); each sweep is a linear
// recurrence that every lane reduces to one affine map of the state entering
// its rows, the maps are composed by a scan over the lanes and each lane
// re-runs its rows from the state the scan hands it (as in the forward
// substitutions).  The states entering the first row are zero.
//
// first order, y_k = c(k) y_{k-+1} + b(k, x_k), x_k <- y_k
template <bool FWD, class CF, class BF>
__device__ __forceinline__ void wave_sweep1(int N, cd *x, CF cf, BF bf)
{
    const int lane = threadIdx.x & 63, C = (N + 63) >> 6;
    const int k0 = min(N, lane * C), k1 = min(N, k0 + C), n = k1 - k0;
    double a = 1.0;
    cd g = mk(0.0, 0.0);
    for (int i = 0; i < n; i++) {
        const int k = FWD ? k0 + i : k1 - 1 - i;
        const double c = cf(k);
        g = cadd(cscale(c, g), bf(k, x[k]));
        a *= c;
    }
    lane_scan1<FWD>(a, g);
    cd y = FWD ? lane_up(g, 1) : lane_dn(g, 1);
    if (lane == (FWD ? 0 : 63)) y = mk(0.0, 0.0);
    for (int i = 0; i < n; i++) {
        const int k = FWD ? k0 + i : k1 - 1 - i;
        y = cadd(cscale(cf(k), y), bf(k, x[k]));
        x[k] = y;
    }
}
// second order, v_k = x_k - c1(k) v_{k-+1} - c2(k) v_{k-+2}, x_k <- out(k, v_k)
template <bool FWD, class C1, class C2, class OF>
__device__ __forceinline__ void wave_sweep2(int N, cd *x, C1 c1f, C2 c2f, OF of)
{
    const int lane = threadIdx.x & 63, C = (N + 63) >> 6;
    const int k0 = min(N, lane * C), k1 = min(N, k0 + C), n = k1 - k0;
    const cd z = mk(0.0, 0.0), one = mk(1.0, 0.0);
    LaneAff2 t{one, z, z, one, z, z};
    for (int i = 0; i < n; i++) {
        const int k = FWD ? k0 + i : k1 - 1 - i;
        const cd a1 = c1f(k), a2 = c2f(k);
        const cd n0 = csub(csub(x[k], cmul(a1, t.g0)), cmul(a2, t.g1));
        t.g1 = t.g0;
        t.g0 = n0;
        const cd r00 = cneg(cfma(a1, t.m00, cmul(a2, t.m10))), r01 = cneg(cfma(a1, t.m01, cmul(a2, t.m11)));
        t.m10 = t.m00;
        t.m11 = t.m01;
        t.m00 = r00;
        t.m01 = r01;
    }
    lane_scan2<FWD>(t);
    cd pv = FWD ? lane_up(t.g0, 1) : lane_dn(t.g0, 1), qv = FWD ? lane_up(t.g1, 1) : lane_dn(t.g1, 1);
    if (lane == (FWD ? 0 : 63)) { pv = z; qv = z; }
    for (int i = 0; i < n; i++) {
        const int k = FWD ? k0 + i : k1 - 1 - i;
        const cd v = csub(csub(x[k], cmul(c1f(k), pv)), cmul(c2f(k), qv));
        x[k] = of(k, v);
        qv = pv;
        pv = v;
    }
}
// tridiagonal (real): A^-1 = U^-1 L^-1 with L unit lower (c_k = d_{k-1}
// lo_k), U = diag(1/d) + up on the superdiagonal; A^-T = L^-T U^-T:
// s_k = d_k (x_k - up_{k-1} s_{k-1}), then t_k = s_k - d_k lo_{k+1} t_{k+1}
__device__ __forceinline__ void fd_tri_solveT_wave(int N, const double *__restrict__ d,
                                                   const double *__restrict__ lo, const double *__restrict__ up,
                                                   cd *x)
{
    wave_sweep1<true>(N, x, [&](int k) { return k == 0 ? 0.0 : -d[k] * up[k - 1]; },
                      [&](int k, cd v) { return cscale(d[k], v); });
    wave_sweep1<false>(N, x, [&](int k) { return k == N - 1 ? 0.0 : -d[k] * lo[k + 1]; },
                       [&](int, cd v) { return v; });
}
// pentadiagonal (complex): A^-1 = U^-1 L^-1 D with L, U unit triangular
// (l1, l2 below, u1, u2 above); A^-H = conj(D) L^-H U^-H
__device__ __forceinline__ void fd_penta_solveH_wave(int N, const cd *__restrict__ d, const cd *__restrict__ l1,
                                                     const cd *__restrict__ l2, const cd *__restrict__ u1,
                                                     const cd *__restrict__ u2, cd *x)
{
    const cd z = mk(0.0, 0.0);
    wave_sweep2<true>(N, x, [&](int k) { return k >= 1 ? cconj(u1[k - 1]) : z; },
                      [&](int k) { return k >= 2 ? cconj(u2[k - 2]) : z; }, [&](int, cd v) { return v; });
    wave_sweep2<false>(N, x, [&](int k) { return k + 1 < N ? cconj(l1[k + 1]) : z; },
                       [&](int k) { return k + 2 < N ? cconj(l2[k + 2]) : z; },
                       [&](int k, cd v) { return cmul(cconj(d[k]), v); });
}

// the adjoints of a T or xi column of fd_implicit (outputs temp, T_old(1)
// in gi, dtemp in gd, T_impl(1) in gm) on the ghosted layout G[0..N+2):
// G[j] = gi[j-1] + sum over the 3-point stencils of rows j-2 .. j
__device__ __forceinline__ cd fd_scalar_outT(const RadView &R, const FdView &S, int N, double kap, double dm2,
                                             const cd *gi, const cd *gd, const cd *gm, int j)
{
    cd s = (j >= 1 && j <= N) ? gi[j - 1] : mk(0, 0);
    for (int k = max(0, j - 2); k <= min(N - 1, j); k++) {
        const int c = j - k;
        double im = S.ddr(k, c) + R.or1(k) * S.dr(k, c);
        if (c == 1) im -= dm2 * R.or2(k);
        s = cadd(s, cadd(cscale(S.dr(k, c), gd[k]), cscale(kap * im, gm[k])));
    }
    return s;
}

// ghost points of a T / xi column (fd_ghost_scalar) transposed in place;
// returns the adjoints of the flux patterns (outer, inner)
__device__ __forceinline__ void fd_ghost_scalarT(cd *G, int N, int ktop, int kbot, const RadView &R, cd &qt, cd &qb)
{
    const cd g0 = G[0], gn = G[N + 1];
    if (ktop == 1) { G[1] = cadd(G[1], cscale(2.0, g0)); G[2] = csub(G[2], g0); }
    else { G[2] = cadd(G[2], g0); qt = cadd(qt, cscale(2.0 * (R.r(0) - R.r(1)), g0)); }
    if (kbot == 1) { G[N] = cadd(G[N], cscale(2.0, gn)); G[N - 1] = csub(G[N - 1], gn); }
    else { G[N - 1] = cadd(G[N - 1], gn); qb = csub(qb, cscale(2.0 * (R.r(N - 2) - R.r(N - 1)), gn)); }
    G[0] = mk(0, 0);
    G[N + 1] = mk(0, 0);
}

// ---------------------------------------------------------------------------
// Transpose of fd_mloop_body (stage 1 of a multistep scheme, explicit part
// included) for one m, in the layout of k_mcoll_adj: reads the adjoints of
// the step's outputs from lam (row m) and replaces them by those of its
// inputs; the adjoints of the radial loop's outputs go to ad, those of the
// wall patterns (l_fd_bc_pattern) are added to ad.gtop / gbot (T) and
// ad.gtopxi / gbotxi (xi).  The factorised systems are the copies the
// window's steps ran with.  LDS: 3 (N + 4) + 10 N complex.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fd_mloop_adj(const PzDev *__restrict__ pfw, const PzDev *__restrict__ plam,
                                                      const PzAdjDev *__restrict__ pad)
{
    extern __shared__ double smem[];
    const PzDev &p = *pfw;
    const PzDev &L = *plam;
    const PzAdjDev &A = *pad;
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x, nb = blockDim.x;
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    const PzScheme *ss = p.ss;
    const RadView R(p.radc, N);
    const FdView S(p.fdc, N);
    const bool hT = p.l_heat, hX = p.l_chem;
    const double dm = (double)m, dm2 = dm * dm, nu = p.ViscFac, cf = p.CorFac;
    const int lcalc = p.l_imp_calc_rhs[0] || (p.bridge && ss->n_steps + 1 <= p.nold - 1);
    const double wcur = ss->wexp[0], w = ss->wimp_lin[0];
    const bool brg = p.bridge && ss->n_steps <= p.nold - 1;
    const double *fT = (brg ? A.fT_b : A.fT_s) + off * 3, *fX = (brg ? A.fX_b : A.fX_s) + off * 3;
    const double *fU = brg ? A.fU_b : A.fU_s;
    const cd *fP = (brg ? A.fP_b : A.fP_s) + off * 5;
    if (mi == 0 && tid == 0 &&
        (ss->l_new_dt || (ss->lMat && ss->n_steps > 1 && !(p.bridge && ss->n_steps <= p.nold))))
        *A.flag = 1;

    cd *GT = (cd *)smem, *GX = GT + (N + 4), *GP = GX + (N + 4);
    cd *V = GP + (N + 4);
    cd *eT = V, *eX = V + N, *eP = V + 2 * N, *Lus0 = V + 3 * N, *LT0 = V + 4 * N, *LX0 = V + 5 * N;
    cd *a0 = V + 6 * N, *a1 = V + 7 * N, *a2 = V + 8 * N, *a3 = V + 9 * N; // per-row adjoints, reused
    __shared__ cd s_q[4]; // adjoints of the patterns: T outer, inner, xi outer, inner
    if (tid < 4) s_q[tid] = mk(0, 0);

    // ---- T and xi: adjoints of the outputs on the ghosted columns ----------
    for (int f = 0; f < 2; f++) {
        if ((f == 0 && !hT) || (f == 1 && !hX)) continue;
        cd *G = f == 0 ? GT : GX, *e = f == 0 ? eT : eX;
        cd *old = f == 0 ? L.T_old : L.X_old, *impl = f == 0 ? L.T_impl : L.X_impl, *expl = f == 0 ? L.T_expl : L.X_expl;
        const cd *fld = f == 0 ? L.temp : L.xi;
        for (int k = tid; k < N; k += nb) {
            const size_t q = off + k;
            a0[k] = cadd(fld[q], old[q]);
            a1[k] = f == 0 ? L.dtemp[q] : mk(0, 0);
            a2[k] = lcalc ? impl[q] : mk(0, 0);
            cd ev = expl[q];
            if (p.nexp > 1) ev = cadd(ev, expl[F + q]);
            e[k] = ev;
            unrotate(p, old, impl, expl, off, k, lcalc);
        }
        __syncthreads();
        const double kap = f == 0 ? p.TdiffFac : p.XidiffFac;
        for (int j = tid; j < N + 2; j += nb) G[j] = fd_scalar_outT(R, S, N, kap, dm2, a0, a1, a2, j);
        __syncthreads();
        if (tid == 0) {
            cd qt = mk(0, 0), qb = mk(0, 0);
            fd_ghost_scalarT(G, N, f == 0 ? p.ktopt : p.ktopxi, f == 0 ? p.kbott : p.kbotxi, R, qt, qb);
            s_q[2 * f] = cadd(s_q[2 * f], qt);
            s_q[2 * f + 1] = cadd(s_q[2 * f + 1], qb);
        }
        __syncthreads();
        if (tid < 64) {
            const double *M = f == 0 ? fT : fX;
            fd_tri_solveT_wave(N, M, M + N, M + 2 * N, G + 1);
        }
        __syncthreads();
    }

    // ---- psi (m /= 0) or u_phi (m = 0): adjoints of the outputs ------------
    for (int k = tid; k < N; k += nb) {
        const size_t q = off + k;
        const cd lold = L.P_old[q], lim = lcalc ? L.P_impl[q] : mk(0, 0);
        cd ev = L.P_expl[q];
        if (p.nexp > 1) ev = cadd(ev, L.P_expl[F + q]);
        eP[k] = ev;
        const double o1 = R.or1(k), o2 = R.or2(k);
        const cd lom = cadd(L.om[q], m == 0 ? mk(0, 0) : lold), lus = L.us[q], lup = L.up[q], lps = L.psi[q];
        if (m == 0) {
            // u1 direct, du, ddu
            a0[k] = cadd(cadd(lold, lup), cadd(cscale(o1, L.om[q]), cscale(-nu * o2 - cf * R.ekpump(k), lim)));
            a1[k] = cadd(L.om[q], cscale(nu * o1, lim));
            a2[k] = cscale(nu, lim);
        } else {
            const double be = R.beta(k), db = R.dbeta(k), o3 = R.or3(k), d2b = R.d2beta(k), d3b = R.d3beta(k);
            const double c3 = be + 2.0 * o1, c2 = 3.0 * db + 2.0 * be * o1 - o2 - 2.0 * dm2 * o2;
            const double c1 = 3.0 * d2b + 4.0 * db * o1 - be * o2 + o3 + dm2 * o2 * (2.0 * o1 - be);
            const double c0 = d3b + 2.0 * d2b * o1 - db * o2 + be * o3 - dm2 * o2 * (4.0 * o2 + be * o1 + db - dm2 * o2);
            const double ek = p.l_ek_pump ? cf * R.ekpump(k) : 0.0, eu = p.l_ek_pump ? R.ekp_up(k) : 0.0;
            // lddp, ldp, lps, ld3, ld4
            const cd lddp = cadd(cscale(-1.0, lom), cscale(-nu * c2 + ek, lim));
            const cd ldp = csub(cadd(cscale(-(o1 + be), lom), cscale(-nu * c1 + ek * (o1 + be + eu), lim)), lup);
            cd lp = csub(csub(lps, cscale(be, lup)), cscale(be * o1 + db - dm2 * o2, lom));
            lp = cadd(lp, ci(-dm * o1, lus));
            lp = cadd(lp, cscale(-nu * c0, lim));
            if (p.l_coriolis_imp) lp = cadd(lp, ci(-cf * be * dm * o1, lim));
            if (p.l_ek_pump) {
                const cd P0 = mk(db + be * o1 - dm2 * o2 + be * R.ekp_up(k) + dm2 * o1 * R.ekp_dusdp(k),
                                 -dm * o1 * R.ekp_us(k));
                lp = cadd(lp, cscale(ek, cmul(cconj(P0), lim)));
            }
            a0[k] = lp;
            a1[k] = ldp;
            a2[k] = lddp;
            a3[k] = cscale(-nu, lim);        // d4 weight 1
            Lus0[k] = cscale(-nu * c3, lim); // d3 (weight c3), kept in Lus0 until the solve
        }
        unrotate(p, L.P_old, L.P_impl, L.P_expl, off, k, lcalc);
    }
    __syncthreads();
    if (m == 0) {
        for (int j = tid; j < N + 2; j += nb) {
            cd s = (j >= 1 && j <= N) ? a0[j - 1] : mk(0, 0);
            for (int k = max(0, j - 2); k <= min(N - 1, j); k++) {
                const int c = j - k;
                s = cadd(s, cadd(cscale(S.dr(k, c), a1[k]), cscale(S.ddr(k, c), a2[k])));
            }
            GP[j] = s;
        }
    } else {
        for (int j = tid; j < N + 4; j += nb) {
            cd s = (j >= 2 && j <= N + 1) ? a0[j - 2] : mk(0, 0);
            for (int k = max(0, j - 4); k <= min(N - 1, j); k++) {
                const int c = j - k; // f_c of row k
                if (c >= 1 && c <= 3) s = cadd(s, cadd(cscale(S.dr(k, c - 1), a1[k]), cscale(S.ddr(k, c - 1), a2[k])));
                s = cadd(s, cadd(cscale(S.d3(k, c), Lus0[k]), cscale(S.d4(k, c), a3[k])));
            }
            GP[j] = s;
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (m == 0) {
            const double dr = R.r(1) - R.r(0), db = R.r(N - 1) - R.r(N - 2);
            const cd g0 = GP[0], gn = GP[N + 1];
            if (p.ktopv == 1) { GP[2] = cadd(GP[2], g0); GP[1] = cadd(GP[1], cscale(-2.0 * dr * R.or1(0), g0)); }
            else { GP[2] = csub(GP[2], g0); GP[1] = cadd(GP[1], cscale(2.0, g0)); }
            if (p.kbotv == 1) { GP[N - 1] = cadd(GP[N - 1], gn); GP[N] = cadd(GP[N], cscale(2.0 * db * R.or1(N - 1), gn)); }
            else { GP[N - 1] = csub(GP[N - 1], gn); GP[N] = cadd(GP[N], cscale(2.0, gn)); }
        } else {
            const double dr = R.r(1) - R.r(0), db = R.r(N - 1) - R.r(N - 2);
            const double ct = p.ktopv == 1 ? -(1.0 - 0.5 * R.or1(0) * dr) / (1.0 + 0.5 * R.or1(0) * dr) : 1.0;
            const double cb = p.kbotv == 1 ? -(1.0 + 0.5 * R.or1(N - 1) * db) / (1.0 - 0.5 * R.or1(N - 1) * db) : 1.0;
            GP[3] = cadd(GP[3], cscale(ct, GP[1]));
            GP[N] = cadd(GP[N], cscale(cb, GP[N + 2]));
        }
    }
    __syncthreads();
    if (tid < 64) {
        if (m == 0) fd_tri_solveT_wave(N, fU, fU + N, fU + 2 * N, GP + 1);
        else fd_penta_solveH_wave(N, fP, fP + N, fP + 2 * N, fP + 3 * N, fP + 4 * N, GP + 2);
    }
    __syncthreads();

    // ---- right-hand sides back to the slots and the explicit terms --------
    for (int k = tid; k < N; k += nb) {
        Lus0[k] = mk(0, 0);
        LT0[k] = mk(0, 0);
        LX0[k] = mk(0, 0);
    }
    __syncthreads();
    if (tid == 0) {
        // wall rows of T and xi: pattern values or flux terms
        for (int f = 0; f < 2; f++) {
            if ((f == 0 && !hT) || (f == 1 && !hX)) continue;
            cd *G = f == 0 ? GT : GX;
            const int ktop = f == 0 ? p.ktopt : p.ktopxi, kbot = f == 0 ? p.kbott : p.kbotxi;
            const double kap = f == 0 ? p.TdiffFac : p.XidiffFac;
            if (ktop == 1) { s_q[2 * f] = cadd(s_q[2 * f], G[1]); G[1] = mk(0, 0); }
            else {
                const double lo = -w * kap * (S.ddr(0, 0) + R.or1(0) * S.dr(0, 0));
                s_q[2 * f] = cadd(s_q[2 * f], cscale(-lo * 2.0 * (R.r(0) - R.r(1)), G[1]));
            }
            if (kbot == 1) { s_q[2 * f + 1] = cadd(s_q[2 * f + 1], G[N]); G[N] = mk(0, 0); }
            else {
                const double up = -w * kap * (S.ddr(N - 1, 2) + R.or1(N - 1) * S.dr(N - 1, 2));
                s_q[2 * f + 1] = cadd(s_q[2 * f + 1], cscale(up * 2.0 * (R.r(N - 2) - R.r(N - 1)), G[N]));
            }
        }
        if (m == 0) {
            if (p.ktopv != 1) GP[1] = mk(0, 0);
            if (p.kbotv != 1) GP[N] = mk(0, 0);
        } else {
            GP[2] = mk(0, 0);
            GP[N + 1] = mk(0, 0);
        }
        if (p.fd_pattern) {
            A.gtop[mi] = cadd(A.gtop[mi], s_q[0]);
            A.gbot[mi] = cadd(A.gbot[mi], s_q[1]);
            A.gtopxi[mi] = cadd(A.gtopxi[mi], s_q[2]);
            A.gbotxi[mi] = cadd(A.gbotxi[mi], s_q[3]);
        }
    }
    __syncthreads();
    for (int k = tid; k < N; k += nb) {
        const double o1 = R.or1(k);
        if (hT) {
            const cd rb = GT[k + 1];
            eT[k] = cadd(eT[k], cscale(wcur, rb));
            rhs_to_slots(p, ss, L.T_old, L.T_impl, L.T_expl, off, k, rb);
            A.eT0[off + k] = eT[k];
            Lus0[k] = csub(Lus0[k], cscale(R.dtcond(k) - p.tadvz * R.beta(k) * R.tcond(k), eT[k]));
        }
        if (hX) {
            const cd rb = GX[k + 1];
            eX[k] = cadd(eX[k], cscale(wcur, rb));
            rhs_to_slots(p, ss, L.X_old, L.X_impl, L.X_expl, off, k, rb);
            A.eX0[off + k] = eX[k];
            Lus0[k] = csub(Lus0[k], cscale(R.dxicond(k), eX[k]));
        }
        const cd rb = m == 0 ? GP[k + 1] : GP[k + 2];
        eP[k] = cadd(eP[k], cscale(wcur, rb));
        rhs_to_slots(p, ss, L.P_old, L.P_impl, L.P_expl, off, k, rb);
        A.eP0[off + k] = eP[k];
        if (m != 0) {
            if (!p.l_coriolis_imp) Lus0[k] = cadd(Lus0[k], cscale(cf * R.beta(k), eP[k]));
            if (hT) LT0[k] = ci(dm * p.BuoFac * R.rgrav(k) * o1, eP[k]);
            if (hX) LX0[k] = ci(dm * p.ChemFac * R.rgrav(k) * o1, eP[k]);
        }
    }
    __syncthreads();
    // adjoints of the dV fields: -o1 times the transposed one-sided stencil
    for (int k = tid; k < N; k += nb) {
        a0[k] = cscale(-R.or1(k), eT[k]);
        a1[k] = cscale(-R.or1(k), eX[k]);
        a2[k] = m == 0 ? mk(0, 0) : cscale(-R.or1(k), eP[k]);
    }
    __syncthreads();
    for (int j = tid; j < N; j += nb) {
        const size_t q = off + j;
        if (hT) A.aT0[q] = fd_drT_at(S, a0, j, N);
        if (hX) A.aX0[q] = fd_drT_at(S, a1, j, N);
        A.aP0[q] = fd_drT_at(S, a2, j, N);
        L.us[q] = Lus0[j];
        L.up[q] = mk(0, 0);
        L.om[q] = mk(0, 0);
        L.psi[q] = mk(0, 0);
        L.temp[q] = LT0[j];
        L.xi[q] = LX0[j];
        L.dtemp[q] = mk(0, 0);
    }
}

// ---------------------------------------------------------------------------
// Transpose of k_fd_start with respect to T: the initial T field fills the
// ghosted column, whose fd_implicit outputs are temp, dtemp, T_old(1) and
// T_impl(1); the flux patterns' adjoints of the ghost points go to
// ad.gtop / gbot.  LDS: 4 N + 4 complex.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fd_start_adj(const PzDev *__restrict__ pd, const PzDev *__restrict__ plam,
                                                      const PzAdjDev *__restrict__ pad, cd *grad)
{
    extern __shared__ double smem[];
    const PzDev &p = *pd;
    const PzDev &L = *plam;
    const PzAdjDev &A = *pad;
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x, nb = blockDim.x;
    const size_t off = (size_t)mi * N;
    const RadView R(p.radc, N);
    const FdView S(p.fdc, N);
    cd *G = (cd *)smem, *gi = G + N + 4, *gd = gi + N, *gm = gd + N;
    for (int k = tid; k < N; k += nb) {
        gi[k] = cadd(L.temp[off + k], L.T_old[off + k]);
        gd[k] = L.dtemp[off + k];
        gm[k] = L.T_impl[off + k];
    }
    __syncthreads();
    for (int j = tid; j < N + 2; j += nb)
        G[j] = fd_scalar_outT(R, S, N, p.TdiffFac, (double)m * (double)m, gi, gd, gm, j);
    __syncthreads();
    if (tid == 0) {
        cd qt = mk(0, 0), qb = mk(0, 0);
        fd_ghost_scalarT(G, N, p.ktopt, p.kbott, R, qt, qb);
        if (p.fd_pattern) {
            A.gtop[mi] = cadd(A.gtop[mi], qt);
            A.gbot[mi] = cadd(A.gbot[mi], qb);
        }
    }
    __syncthreads();
    for (int k = tid; k < N; k += nb) grad[off + k] = G[k + 1];
}

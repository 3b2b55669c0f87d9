/*
 * ORACLE — test infrastructure only (see pz_oracle.h).
 *
 * CPU restatement of the reference's IMEX step for the sparse-Chebyshev
 * (integration) method with direct solves.  Each function names the
 * reference routine (file:line, relative to the reference source tree) whose
 * algorithm it follows.  Arithmetic is FP64 throughout; the operator
 * stencils are composed in long double and rounded once.
 */
#include "pz_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;
typedef long double ld;

#define PI_F (4.0 * atan(1.0)) /* constants.f90:5 */
#ifdef _OPENMP
#include <omp.h>
#endif

/* Scratch buffers owned by the calling thread: the per-radius and per-m
 * loops of the step run on OpenMP threads, and every transform below takes
 * its work arrays from here instead of from the handle.  Buffer k grows on
 * demand; a thread uses one handle at a time. */
static cplx *tl_cbuf(int k, size_t n)
{
    static _Thread_local cplx *buf[6];
    static _Thread_local size_t cap[6];
    if (cap[k] < n) {
        free(buf[k]);
        buf[k] = malloc(n * sizeof(cplx));
        cap[k] = n;
    }
    return buf[k];
}
/* Planted mistakes for the tests that show a check can fail (0 = none):
 * 1 flips the sign of the integration method's explicit Ekman-pumping term
 *   (update_psi_integ_smat.f90:895-913),
 * 2 drops its factor r for m /= 0,
 * 3 drops its u_s part (ekp_dusdp, ekp_us),
 * 4 flips the sign of the compositional buoyancy (ChemFac),
 * 5 doubles the compositional diffusivity (XidiffFac),
 * 6 flips the sign of the collocation influence-matrix method's explicit
 *   Ekman-pumping term (update_psi_coll_dmat.f90:527-549),
 * 7 drops the flux pattern's part of the finite-difference wall row,
 * 8 flips the sign of the flux pattern in the finite-difference ghost point.
 * Read when a run is created (4, 5), in the explicit finish (1-3, 6) or in
 * the finite-difference boundary handling (7, 8). */
int or_plant = 0;

static double *tl_dbuf(size_t n)
{
    static _Thread_local double *buf;
    static _Thread_local size_t cap;
    if (cap < n) {
        free(buf);
        buf = malloc(n * sizeof(double));
        cap = n;
    }
    return buf;
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex FFT (unnormalised; dir=-1 forward, +1 backward).       */
/* Stands in for the FFTW plans of fft_fftw.f90:21-35 and dct_fftw.f90:77-84. */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    cplx *w;    /* w[k] = exp(-2 pi i k / n) */
} fftp;

static fftp *fft_new(int n)
{
    fftp *p = calloc(1, sizeof(fftp));
    p->n = n;
    p->w = malloc(sizeof(cplx) * n);
    for (int k = 0; k < n; k++) {
        ld ang = 2.0L * 3.14159265358979323846264338327950288L * (ld)k / (ld)n;
        p->w[k] = (double)cosl(ang) - I * (double)sinl(ang);
    }
    return p;
}

static void fft_free(fftp *p)
{
    if (!p) return;
    free(p->w);
    free(p);
}

static int small_factor(int n)
{
    if (n % 4 == 0) return 4;
    for (int f = 2; f * f <= n; f++)
        if (n % f == 0) return f;
    return n;
}

static inline cplx twid(const fftp *p, long idx, int dir)
{
    cplx w = p->w[idx % p->n];
    return dir < 0 ? w : conj(w);
}

/* y[0..n) = DFT of x[0], x[is], ..., decimation in time, radix f */
static void fft_rec(const fftp *p, const cplx *x, long is, cplx *y, int n, int dir)
{
    if (n == 1) {
        y[0] = x[0];
        return;
    }
    int f = small_factor(n), q = n / f;
    for (int r = 0; r < f; r++) fft_rec(p, x + r * is, is * f, y + r * q, q, dir);
    long step = p->n / n;
    /* radix f: a prime factor of n, any size (e.g. 193 in 2 n_r - 2 = 386) */
    cplx tbuf[64];
    cplx *t = f <= 64 ? tbuf : malloc(sizeof(cplx) * f);
    cplx *out = tl_cbuf(1, 0);
    for (int k = 0; k < q; k++) {
        for (int r = 0; r < f; r++) t[r] = y[r * q + k] * twid(p, (long)r * k * step, dir);
        for (int s = 0; s < f; s++) {
            cplx acc = 0;
            for (int r = 0; r < f; r++) acc += t[r] * twid(p, ((long)r * s % f) * q * step, dir);
            out[k + s * q] = acc;
        }
    }
    if (t != tbuf) free(t);
    memcpy(y, out, sizeof(cplx) * n);
}

static void fft_exec(fftp *p, const cplx *in, cplx *out, int dir)
{
    cplx *x = tl_cbuf(0, p->n);
    (void)tl_cbuf(1, p->n); /* sized before the recursion reuses it */
    memcpy(x, in, sizeof(cplx) * p->n);
    fft_rec(p, x, 1, out, p->n, dir);
}

/* ------------------------------------------------------------------------ */
/* Handle                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct { int kl, ku, n; double *d; } bandr;   /* dgbmv storage */
typedef struct { int kl, ku, n; cplx *d; } bandc;     /* zgbmv storage */

typedef struct {          /* bordered_matrix.f90:15-28 (real) */
    int kl, ku, ntau, nb; /* nb = nlines_band */
    double *A1, *A2, *A3, *A4; /* A1 ntau x ntau, A2 ntau x nb, A3 nb x ntau, A4 (2kl+ku+1) x nb */
    int *p1, *p4;
} bordr;

typedef struct {          /* bordered_matrix.f90:30-43 (complex) */
    int kl, ku, ntau, nb;
    cplx *A1, *A2, *A3, *A4;
    int *p1, *p4;
} bordc;

typedef struct {          /* time_array.f90:11-24 */
    int nold, nimp, nexp;
    cplx *old, *impl, *expl; /* [slot][n_m][n_r] */
} tarray;

typedef struct {          /* time_schemes.f90:10-40 + children */
    int family;           /* 0 multistep, 1 dirk */
    int scheme;
    int nstages, istage, nexp, nold, nimp;
    int l_assembly;
    double courfac;
    double dt[8], wimp[8], wimp_lin[8], wexp[8];
    double bimp[8][8], bexp[8][8];   /* scaled Butcher tables (row, col) 0-based */
    double bimp0[8][8], bexp0[8][8]; /* unscaled */
    double bass_imp[8], bass_exp[8];   /* assembly weights (dt-scaled) */
    double bass_imp0[8], bass_exp0[8];
    double wimp_lin0;
    int l_exp_calc[8], l_imp_calc_rhs[8];
} tscheme;

struct or_handle {
    pz_config c;
    int N, nm, nphi, ncheb, mmax, minc;
    int *idx2m;
    int l_heat, l_chem;
    double CorFac, BuoFac, TdiffFac, ChemFac, XidiffFac, ViscFac, tadvz;
    double r_cmb, r_icb, ek;
    double *r, *x, *drx, *ddrx, *or1, *or2, *rgrav, *beta, *ekpump, *ekp_up, *ekp_us,
        *ekp_dusdp, *height, *oheight, *tcond, *dtcond, *xicond, *dxicond, *delxr2,
        *delxh2;
    int *mR;
    double rnorm, cheb_fac, boundary_fac, thr;
    fftp *fdct, *fphi;
    int *der, *der2;
    double *rMat, *drMat, *d2rMat; /* tau lines: [2][N] */
    double *hdifT, *hdifXi, *hdifV;
    cplx *topt, *bott, *topxi, *botxi;
    /* fields [n_m][n_r] */
    cplx *us, *up, *om, *psi, *temp, *xi, *temp_hat, *xi_hat, *psi_hat, *dtemp, *dxi,
        *work, *dVsT, *dVsXi, *dVsOm, *buo;
    tarray dTdt, dxidt, dpsidt;
    /* operator matrices */
    bandr T_RHSE[2], *T_RHSI;
    bordr *T_LHS, *XI_LHS;
    double *tempfac, *xifac;
    bandr P_RHSE[3], *P_RHSI;
    bandc *P_RHSIL;
    bordc *P_LHS;
    double *psifac;
    struct optabs_s *tabs;
    tscheme ts;
    double time;
    int lMatNext, n_time_step, started;
    /* restart (read_checkpoint, checkpoints.f90:185-707): l_bridge_step and
     * l_AB1 of namelists.f90:364-365, the time and dt array of the file */
    int l_bridge, l_ab1, restart;
    double rst_time, rst_dt[8];
    double *dtr, *dth;
    /* finite differences (radial_scheme = 'FD', finite_differences.f90) */
    int fd;
    double *or3, *d2beta, *d3beta;
    double *fdr, *fddr;     /* [N][3]: stencil of points n-1, n, n+1 */
    double *fdddr, *fddddr; /* [N][5]: points n-2 .. n+2 */
    double fdr_top[3], fdr_bot[3];
    double *tTd, *tTl, *tTu, *tXd, *tXl, *tXu; /* tridiagonal T, xi per m: [m][N] */
    double *uD, *uL, *uU;                      /* tridiagonal u_phi (m = 0) */
    cplx *pD, *pL1, *pL2, *pU1, *pU2;          /* pentadiagonal psi per m: [m][N] */
    cplx *tg, *xg, *pg, *u0g; /* ghosted T, xi [m][N+2], psi [m][N+4], u_phi(m=0) [N+2] */
    /* collocation method (l_cheb_coll) */
    int coll;
    double *dbeta;
    double *cM, *cdM, *cd2M; /* T_n(x_k), d/dr, d2/dr2 at point k: [k * N + n] */
    double *cTmat, *cXmat, *cTfac, *cXfac, *cUmat;
    double *dr_top, *dr_bot; /* boundary-derivative rows (chebyshev.f90:338-360) */
    cplx *cAss;              /* LU of the psi assembly matrices, [m][N][N] */
    int *cAssPiv, lAss;
    int *cTpiv, *cXpiv, *cUpiv, *cPpiv;
    cplx *cPmat;
    double *cPfac;
    /* influence-matrix psi solve (update_psi_coll_dmat.f90): per m the
     * omega and psi matrices with their row/column scalings, the two
     * homogeneous solutions of each (Chebyshev space) and the 2 x 2
     * influence matrix */
    double *dOmat, *dOfac, *dPmat, *dPfac, *dom2, *dom3, *dpsi2, *dpsi3, *dinfl;
    int *dOpiv, *dPpiv;
    /* scratch */
    int wbig; /* length of the per-thread transform buffers (tl_cbuf) */
};

/* ------------------------------------------------------------------------ */
/* Config defaults: namelists.f90:530-672                                    */
/* ------------------------------------------------------------------------ */
void or_config_default(pz_config *c)
{
    memset(c, 0, sizeof(*c));
    c->n_r_max = 32;
    c->n_cheb_max = 32;
    c->m_max = 32;
    c->minc = 1;
    c->time_scheme = PZ_CNAB2;
    c->l_cheb_coll = 1;
    c->l_direct_solve = 1;
    c->l_coriolis_imp = 1;
    c->l_buo_imp = 1;
    c->l_non_rot = 0;
    c->time_scale_rot = 0;
    c->l_rerror_fix = 0;
    c->rerror_fac = 500.0;
    c->alpha = 0.5;
    c->dtMax = 1.0e-5;
    c->courfac = 1.0e3;
    c->dt_fac = 2.0;
    c->ra = 1.0e5;
    c->ek = 1.0e-3;
    c->pr = 1.0;
    c->raxi = 0.0;
    c->sc = 0.0;
    c->radratio = 0.35;
    c->g0 = 0.0;
    c->g1 = 1.0;
    c->g2 = 0.0;
    c->ktopt = c->kbott = 1;
    c->ktopxi = c->kbotxi = 1;
    c->ktopv = c->kbotv = 2;
    c->l_ek_pump = 0;
    c->l_temp_advz = 0;
    c->tcond_fac = 1.0;
    c->xicond_fac = 1.0;
    c->fd_stretch = 0.3; /* namelists.f90:544-545 */
    c->fd_ratio = 0.1;
    c->l_finite_diff = 0;
}

/* ------------------------------------------------------------------------ */
/* Azimuthal transforms: fourier::fft / ifft (fft_fftw.f90:44-90)             */
/* ------------------------------------------------------------------------ */
static void phi_fft(or_handle *h, const double *grid, cplx *spec, int m_R)
{
    int n = h->nphi;
    cplx *a = tl_cbuf(2, h->wbig), *b = tl_cbuf(3, h->wbig);
    for (int j = 0; j < n; j++) a[j] = grid[j];
    fft_exec(h->fphi, a, b, -1);
    for (int nm = 0; nm < h->nm; nm++) {
        if (h->idx2m[nm] <= m_R) spec[nm] = b[nm] / (double)n;
        else spec[nm] = 0;
    }
}

/* c2r semantics: the imaginary part of the zero bin is not used */
static void phi_ifft(or_handle *h, const cplx *spec, double *grid, int m_R)
{
    int n = h->nphi;
    cplx *a = tl_cbuf(2, h->wbig), *b = tl_cbuf(3, h->wbig);
    for (int k = 0; k < n; k++) a[k] = 0;
    for (int nm = 0; nm < h->nm; nm++) {
        cplx v = (h->idx2m[nm] <= m_R) ? spec[nm] : 0;
        if (nm == 0) {
            a[0] = creal(v);
        } else {
            a[nm] = v;
            a[n - nm] = conj(v);
        }
    }
    fft_exec(h->fphi, a, b, +1);
    for (int j = 0; j < n; j++) grid[j] = creal(b[j]);
}

/* ------------------------------------------------------------------------ */
/* Chebyshev transforms: dct_fftw.f90 (DCT_VERSION = dft_loop)               */
/* ------------------------------------------------------------------------ */

/* costf_complex_2d for one m (dct_fftw.f90:184-197): DCT-I as the forward
 * FFT of the even extension, scaled by cheb_fac. */
static void costf_vec(or_handle *h, cplx *f)
{
    int N = h->N, L = 2 * N - 2;
    cplx *a = tl_cbuf(2, h->wbig), *b = tl_cbuf(3, h->wbig);
    for (int j = 0; j < N; j++) a[j] = f[j];
    for (int j = N; j < L; j++) a[j] = f[L - j];
    fft_exec(h->fdct, a, b, -1);
    for (int k = 0; k < N; k++) f[k] = h->cheb_fac * b[k];
}

/* costf_real_1d (dct_fftw.f90:231-246): FFTW REDFT00 then cheb_fac */
static void costf_real(or_handle *h, double *f)
{
    int N = h->N, L = 2 * N - 2;
    cplx *a = tl_cbuf(2, h->wbig), *b = tl_cbuf(3, h->wbig);
    for (int j = 0; j < N; j++) a[j] = f[j];
    for (int j = N; j < L; j++) a[j] = f[L - j];
    fft_exec(h->fdct, a, b, -1);
    for (int k = 0; k < N; k++) f[k] = h->cheb_fac * creal(b[k]);
}

static void costf_field(or_handle *h, cplx *F)
{
#pragma omp parallel for schedule(static)
    for (int nm = 0; nm < h->nm; nm++) costf_vec(h, F + (size_t)nm * h->N);
}

/* get_dr_fft (dct_fftw.f90:258-316) followed by the drx map
 * (radial_derivatives.f90:388-391), one m. */
static void get_dr_vec(or_handle *h, const cplx *f, cplx *df, int l_dct_in)
{
    int N = h->N, L = 2 * N - 2;
    cplx *a = tl_cbuf(2, h->wbig), *o = tl_cbuf(4, h->wbig), *b = tl_cbuf(3, h->wbig);
    if (l_dct_in) {
        for (int j = 0; j < N; j++) a[j] = f[j];
        for (int j = N; j < L; j++) a[j] = f[L - j];
        fft_exec(h->fdct, a, o, -1);
    } else {
        for (int j = 0; j < N; j++) o[j] = f[j] / h->cheb_fac;
        for (int j = N; j < L; j++) o[j] = f[L - j] / h->cheb_fac;
    }
    o[N - 1] = 0.5 * o[N - 1];
    cplx tot = 0;
    for (int k = 1; k <= N - 1; k++) tot += (double)(k * k) * o[k];
    df[0] = tot / (double)(N - 1);
    tot = 0;
    for (int k = 1; k <= N - 1; k++) tot += (double)(((k + 1) % 2 == 0 ? 1 : -1) * k * k) * o[k];
    df[N - 1] = tot / (double)(N - 1);
    o[N - 1] = 2.0 * o[N - 1];
    for (int p = 0; p < L; p++) o[p] = I * (double)h->der[p] * o[p];
    fft_exec(h->fdct, o, b, +1);
    for (int j = 1; j < N - 1; j++)
        df[j] = -b[j] / sqrt(1.0 - h->x[j] * h->x[j]) / (double)(2 * N - 2);
    for (int j = 0; j < N; j++) df[j] = h->drx[j] * df[j];
}

/* get_ddr_fft (dct_fftw.f90:318-397) + map (radial_derivatives.f90:544-549) */
static void get_ddr_vec(or_handle *h, const cplx *f, cplx *df, cplx *ddf, int l_dct_in)
{
    int N = h->N, L = 2 * N - 2;
    cplx *a = tl_cbuf(2, h->wbig), *t = tl_cbuf(4, h->wbig), *b = tl_cbuf(3, h->wbig);
    cplx *w2 = malloc(sizeof(cplx) * L), *w1 = malloc(sizeof(cplx) * L);
    if (l_dct_in) {
        for (int j = 0; j < N; j++) a[j] = f[j];
        for (int j = N; j < L; j++) a[j] = f[L - j];
        fft_exec(h->fdct, a, t, -1);
    } else {
        for (int j = 0; j < N; j++) t[j] = f[j] / h->cheb_fac;
        for (int j = N; j < L; j++) t[j] = f[L - j] / h->cheb_fac;
    }
    t[N - 1] = 0.5 * t[N - 1];
    cplx tot = 0;
    for (int k = 1; k <= N - 1; k++) tot += (double)(k * k) * t[k];
    df[0] = tot / (double)(N - 1);
    tot = 0;
    for (int k = 2; k <= N - 1; k++) tot += (double)((long)k * k * (k * k - 1)) * t[k];
    ddf[0] = (1.0 / 3.0) * tot / (double)(N - 1);
    tot = 0;
    for (int k = 1; k <= N - 1; k++) tot += (double)(((k + 1) % 2 == 0 ? 1 : -1) * k * k) * t[k];
    df[N - 1] = tot / (double)(N - 1);
    tot = 0;
    for (int k = 2; k <= N - 1; k++)
        tot += (double)((k % 2 == 0 ? 1 : -1) * (long)k * k * (k * k - 1)) * t[k];
    ddf[N - 1] = (1.0 / 3.0) * tot / (double)(N - 1);
    t[N - 1] = 2.0 * t[N - 1];
    for (int p = 0; p < L; p++) w2[p] = I * (double)h->der[p] * t[p];
    fft_exec(h->fdct, w2, w1, +1);
    for (int p = 0; p < L; p++) t[p] = -(double)h->der2[p] * t[p];
    fft_exec(h->fdct, t, w2, +1);
    for (int j = 1; j < N - 1; j++) {
        double xj = h->x[j], s = 1.0 - xj * xj;
        df[j] = -w1[j] / sqrt(s) / (double)(2 * N - 2);
        ddf[j] = -w1[j] * xj / pow(s, 1.5) / (double)(2 * N - 2) + w2[j] / s / (double)(2 * N - 2);
    }
    for (int j = 0; j < N; j++) {
        ddf[j] = h->ddrx[j] * df[j] + h->drx[j] * h->drx[j] * ddf[j];
        df[j] = h->drx[j] * df[j];
    }
    (void)b;
    free(w1);
    free(w2);
}

/* get_dcheb_real_1d (radial_derivatives.f90:143-185) with the rerror
 * threshold, used by get_dr_real_1d (:420-485) for u_phi(m=0). */
static void get_dr_real(or_handle *h, const double *f_in, double *df, int l_dct)
{
    int N = h->N;
    double *f = malloc(sizeof(double) * N);
    memcpy(f, f_in, sizeof(double) * N);
    if (l_dct) costf_real(h, f);
    double mx = 0;
    for (int j = 0; j < N; j++) mx = fmax(mx, fabs(f[j]));
    double threshold = mx * h->thr;
    /* 1-based: df(n_cheb_max..n_r_max)=0 with n_cheb_max = n_r_max */
    df[N - 1] = 0.0;
    int nc = N - 1; /* n_cheb = n_cheb_max-1 (1-based) */
    double fac = (double)nc;
    if (fabs(f[nc]) >= threshold) df[nc - 1] = fac * f[nc];
    else df[nc - 1] = 0.0;
    for (int n = N - 2; n >= 1; n--) {
        fac = (double)(2 * n);
        if (fabs(f[n]) >= threshold) df[n - 1] = df[n + 1] + fac * f[n];
        else df[n - 1] = 0.0;
    }
    costf_real(h, df);
    for (int j = 0; j < N; j++) df[j] = h->drx[j] * df[j];
    free(f);
}

/* ------------------------------------------------------------------------ */
/* Sparse-Chebyshev operators (chebsparselib.f90): rows of I^J [p(r) D^k]    */
/* composed from multiplication by r = a x + b and the integration operator  */
/* in the Chebyshev basis with halved end coefficients.  The folding of       */
/* negative columns follows mirror_stencil (chebsparselib.f90:997-1020).     */
/* ------------------------------------------------------------------------ */
#define LP_OFF 8
#define LP_LEN 32
typedef struct { ld c[LP_LEN]; } lpoly; /* Laurent polynomial, exponent e at c[e+LP_OFF] */
#define DOP_K 7
typedef struct { lpoly a[DOP_K]; } dop; /* sum_k a_k(r) D^k */

static lpoly lp_zero(void) { lpoly p; memset(&p, 0, sizeof(p)); return p; }
static lpoly lp_mono(ld c, int e) { lpoly p = lp_zero(); p.c[e + LP_OFF] = c; return p; }
static lpoly lp_add(lpoly a, lpoly b, ld s)
{
    for (int i = 0; i < LP_LEN; i++) a.c[i] += s * b.c[i];
    return a;
}
static lpoly lp_mul(lpoly a, lpoly b)
{
    lpoly p = lp_zero();
    for (int i = 0; i < LP_LEN; i++)
        if (a.c[i] != 0)
            for (int j = 0; j < LP_LEN; j++)
                if (b.c[j] != 0) {
                    int e = (i - LP_OFF) + (j - LP_OFF) + LP_OFF;
                    if (e < 0 || e >= LP_LEN) { fprintf(stderr, "lpoly overflow\n"); abort(); }
                    p.c[e] += a.c[i] * b.c[j];
                }
    return p;
}
static lpoly lp_der(lpoly a)
{
    lpoly p = lp_zero();
    for (int i = 0; i < LP_LEN; i++) {
        int e = i - LP_OFF;
        if (a.c[i] != 0 && e != 0) p.c[i - 1] += (ld)e * a.c[i];
    }
    return p;
}
static dop dop_zero(void) { dop d; memset(&d, 0, sizeof(d)); return d; }
/* (A o B) = sum_k a_k D^k sum_l b_l D^l = sum a_k C(k,i) b_l^(i) D^(k-i+l) */
static dop dop_compose(const dop *A, const dop *B)
{
    dop R = dop_zero();
    for (int k = 0; k < DOP_K; k++) {
        int nz = 0;
        for (int i = 0; i < LP_LEN; i++) nz |= A->a[k].c[i] != 0;
        if (!nz) continue;
        for (int l = 0; l < DOP_K; l++) {
            int nzb = 0;
            for (int i = 0; i < LP_LEN; i++) nzb |= B->a[l].c[i] != 0;
            if (!nzb) continue;
            lpoly bl = B->a[l];
            ld binom = 1;
            for (int i = 0; i <= k; i++) {
                if (i > 0) { binom = binom * (ld)(k - i + 1) / (ld)i; bl = lp_der(bl); }
                int o = k - i + l;
                if (o >= DOP_K) { fprintf(stderr, "dop order overflow\n"); abort(); }
                R.a[o] = lp_add(R.a[o], lp_mul(A->a[k], bl), binom);
            }
        }
    }
    return R;
}
static dop dop_mulp(lpoly p, const dop *A)
{
    dop R = dop_zero();
    for (int k = 0; k < DOP_K; k++) R.a[k] = lp_mul(p, A->a[k]);
    return R;
}

typedef struct { int E, L; ld a, b; ld *buf[24]; } seqws; /* sequences on [-E, L) */

static void seq_mulr(const seqws *w, const ld *y, ld *o)
{
    int T = w->E + w->L;
    for (int i = 0; i < T; i++) {
        ld ym = i > 0 ? y[i - 1] : 0, yp = i + 1 < T ? y[i + 1] : 0;
        o[i] = w->a * (ym + yp) / 2 + w->b * y[i];
    }
}
static void seq_integ(const seqws *w, const ld *y, ld *o)
{
    int T = w->E + w->L;
    for (int i = 0; i < T; i++) {
        int n = i - w->E;
        if (n == 0) { o[i] = 0; continue; }
        ld ym = i > 0 ? y[i - 1] : 0, yp = i + 1 < T ? y[i + 1] : 0;
        o[i] = w->a * (ym - yp) / (2 * (ld)n);
    }
}
/* o = I^j [ p(r) y ], p a polynomial (no negative powers) */
static void seq_Ip(const seqws *w, int j, const lpoly *p, const ld *y, ld *o, int depth)
{
    int T = w->E + w->L;
    ld *acc = w->buf[2 * depth], *tmp = w->buf[2 * depth + 1];
    int deg = -1;
    for (int e = 0; e < LP_LEN - LP_OFF; e++)
        if (p->c[e + LP_OFF] != 0) deg = e;
    for (int e = -LP_OFF; e < 0; e++)
        if (p->c[e + LP_OFF] != 0) { fprintf(stderr, "negative power in operator\n"); abort(); }
    for (int i = 0; i < T; i++) acc[i] = 0;
    if (deg >= 0) {
        for (int i = 0; i < T; i++) acc[i] = p->c[deg + LP_OFF] * y[i];
        for (int e = deg - 1; e >= 0; e--) {
            seq_mulr(w, acc, tmp);
            for (int i = 0; i < T; i++) acc[i] = tmp[i] + p->c[e + LP_OFF] * y[i];
        }
    }
    for (int q = 0; q < j; q++) {
        seq_integ(w, acc, tmp);
        memcpy(acc, tmp, sizeof(ld) * T);
    }
    memcpy(o, acc, sizeof(ld) * T);
}
/* o = I^j [ p D^k ] y via I(p D g) = p g - I(p' g) */
static void seq_term(const seqws *w, int j, lpoly p, int k, const ld *y, ld *o, int depth)
{
    int T = w->E + w->L;
    if (k == 0) { seq_Ip(w, j, &p, y, o, depth); return; }
    if (j < k) { fprintf(stderr, "operator not integrable\n"); abort(); }
    ld *t1 = w->buf[2 * depth + 12], *t2 = w->buf[2 * depth + 13];
    seq_term(w, j - 1, p, k - 1, y, t1, depth + 1);
    memcpy(o, t1, sizeof(ld) * T);
    seq_term(w, j, lp_der(p), k - 1, y, t2, depth + 1);
    for (int i = 0; i < T; i++) o[i] -= t2[i];
}

/*
 * ext[n*(2K+1) + (K+d)] = row n, column n+d of I^J[op] on the extended
 * (unfolded) index set, rows n in [0, N), d in [-K, K].
 */
static void op_ext(const dop *op, int J, int N, int K, ld a, ld b, ld *ext)
{
    seqws w;
    w.E = K + 24;
    w.L = N + K + 24;
    w.a = a;
    w.b = b;
    int T = w.E + w.L;
    for (int i = 0; i < 24; i++) w.buf[i] = calloc(T, sizeof(ld));
    ld *e = calloc(T, sizeof(ld)), *o = calloc(T, sizeof(ld)), *acc = calloc(T, sizeof(ld));
    memset(ext, 0, sizeof(ld) * (size_t)N * (2 * K + 1));
    for (int col = -K; col < N + K; col++) {
        memset(e, 0, sizeof(ld) * T);
        e[col + w.E] = 1;
        memset(acc, 0, sizeof(ld) * T);
        for (int k = 0; k < DOP_K; k++) {
            int nz = 0;
            for (int i = 0; i < LP_LEN; i++) nz |= op->a[k].c[i] != 0;
            if (!nz) continue;
            seq_term(&w, J, op->a[k], k, e, o, 0);
            for (int i = 0; i < T; i++) acc[i] += o[i];
        }
        for (int n = 0; n < N; n++) {
            int d = col - n;
            if (d < -K || d > K) continue;
            ext[(size_t)n * (2 * K + 1) + (K + d)] = acc[n + w.E];
        }
    }
    for (int i = 0; i < 24; i++) free(w.buf[i]);
    free(e);
    free(o);
    free(acc);
}

/* reference stencil order (n_band = 1..2K+1 <-> column idx+K+1-n_band),
 * folded as mirror_stencil does; the unfolded negative-column entries stay */
static void ext_to_stencil(const ld *ext_row, int idx, int K, ld *st)
{
    for (int nb = 0; nb < 2 * K + 1; nb++) {
        int d = K - nb; /* column = idx + d */
        st[nb] = ext_row[K + d];
    }
    if (idx < K) {
        for (int n = 1; n <= K - idx; n++) {
            /* column n <- column -n */
            int nb_pos = K - (n - idx), nb_neg = K - (-n - idx);
            st[nb_pos] += st[nb_neg];
        }
    }
}

/* Operator library: I2R2, I2R1, I2R2LAPL, R2, I1R1, I4R4, I4R3, LAPLROT, LAPLROT2 */
typedef struct { int J; dop op; } opdef;

static dop dop_poly(lpoly p) { dop d = dop_zero(); d.a[0] = p; return d; }

/* r^2 Delta_m = r^2 D^2 + r D - m^2 (split in m-independent / m^2 parts) */
static void def_lapl(dop *d0, dop *d2)
{
    *d0 = dop_zero();
    *d2 = dop_zero();
    d0->a[2] = lp_mono(1, 2);
    d0->a[1] = lp_mono(1, 1);
    d2->a[0] = lp_mono(-1, 0);
}

/* omega = L_rot psi = -h^2 psi'' - (rc^2/r - 6 r) psi' + (6 + m^2 h^2 / r^2) psi,
 * h^2 = rc^2 - r^2 (from the u_s, u_phi, omega reconstruction of
 * update_psi_integ_smat.f90:504-523). Returned as Lr0 + m^2 Lr2. */
static void def_lrot(ld rc, dop *L0, dop *L2)
{
    lpoly h2 = lp_add(lp_mono(rc * rc, 0), lp_mono(1, 2), -1);
    *L0 = dop_zero();
    *L2 = dop_zero();
    L0->a[2] = lp_add(lp_zero(), h2, -1);
    L0->a[1] = lp_add(lp_mono(-rc * rc, -1), lp_mono(6, 1), 1);
    L0->a[0] = lp_mono(6, 0);
    L2->a[0] = lp_mul(h2, lp_mono(1, -2));
}

/* ------------------------------------------------------------------------ */
/* Band / bordered LU with LAPACK semantics (dgbtf2/zgbtf2, dgbtrs/zgbtrs,   */
/* dgetf2, dgemm/dgemv) as called by algebra_lapack.f90:176-416.             */
/* LU storage: ab[(kl+ku+i-j) + j*ld], ld = 2kl+ku+1, 0-based i, j.          */
/* ------------------------------------------------------------------------ */
static void gbtrf_r(int n, int kl, int ku, double *ab, int *ipiv, int *info)
{
    int ld = 2 * kl + ku + 1, kv = ku + kl;
#define AB(i, j) ab[(i) + (size_t)(j) * ld] /* 0-based storage row i, col j */
    *info = 0;
    for (int j = ku + 1; j < (kv < n ? kv : n); j++)
        for (int i = kv - j; i < kl; i++) AB(i, j) = 0;
    int ju = 0;
    for (int j = 0; j < n; j++) {
        if (j + kv < n)
            for (int i = 0; i < kl; i++) AB(i, j + kv) = 0;
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        int jp = 0;
        double mx = fabs(AB(kv, j));
        for (int i = 1; i <= km; i++)
            if (fabs(AB(kv + i, j)) > mx) { mx = fabs(AB(kv + i, j)); jp = i; }
        ipiv[j] = jp + j;
        if (AB(kv + jp, j) != 0) {
            int t = j + ku + jp;
            if (t > n - 1) t = n - 1;
            if (t > ju) ju = t;
            if (jp != 0)
                for (int c = 0; c <= ju - j; c++) {
                    double tmp = AB(kv + jp - c, j + c);
                    AB(kv + jp - c, j + c) = AB(kv - c, j + c);
                    AB(kv - c, j + c) = tmp;
                }
            if (km > 0) {
                double rp = 1.0 / AB(kv, j);
                for (int i = 1; i <= km; i++) AB(kv + i, j) *= rp;
                for (int c = 1; c <= ju - j; c++) {
                    double temp = -AB(kv - c, j + c);
                    if (temp != 0)
                        for (int i = 1; i <= km; i++) AB(kv + i - c, j + c) += AB(kv + i, j) * temp;
                }
            }
        } else if (*info == 0) {
            *info = j + 1;
        }
    }
#undef AB
}

static inline double cabs1(cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

static void gbtrf_c(int n, int kl, int ku, cplx *ab, int *ipiv, int *info)
{
    int ld = 2 * kl + ku + 1, kv = ku + kl;
#define AB(i, j) ab[(i) + (size_t)(j) * ld]
    *info = 0;
    for (int j = ku + 1; j < (kv < n ? kv : n); j++)
        for (int i = kv - j; i < kl; i++) AB(i, j) = 0;
    int ju = 0;
    for (int j = 0; j < n; j++) {
        if (j + kv < n)
            for (int i = 0; i < kl; i++) AB(i, j + kv) = 0;
        int km = kl < n - 1 - j ? kl : n - 1 - j;
        int jp = 0;
        double mx = cabs1(AB(kv, j));
        for (int i = 1; i <= km; i++)
            if (cabs1(AB(kv + i, j)) > mx) { mx = cabs1(AB(kv + i, j)); jp = i; }
        ipiv[j] = jp + j;
        if (AB(kv + jp, j) != 0) {
            int t = j + ku + jp;
            if (t > n - 1) t = n - 1;
            if (t > ju) ju = t;
            if (jp != 0)
                for (int c = 0; c <= ju - j; c++) {
                    cplx tmp = AB(kv + jp - c, j + c);
                    AB(kv + jp - c, j + c) = AB(kv - c, j + c);
                    AB(kv - c, j + c) = tmp;
                }
            if (km > 0) {
                cplx rp = 1.0 / AB(kv, j);
                for (int i = 1; i <= km; i++) AB(kv + i, j) *= rp;
                for (int c = 1; c <= ju - j; c++) {
                    cplx temp = -AB(kv - c, j + c);
                    if (temp != 0)
                        for (int i = 1; i <= km; i++) AB(kv + i - c, j + c) += AB(kv + i, j) * temp;
                }
            }
        } else if (*info == 0) {
            *info = j + 1;
        }
    }
#undef AB
}

static void gbtrs_r(int n, int kl, int ku, const double *ab, const int *ipiv, double *b)
{
    int ld = 2 * kl + ku + 1, kd = ku + kl; /* diag row of U */
    if (kl > 0)
        for (int j = 0; j < n - 1; j++) {
            int lm = kl < n - 1 - j ? kl : n - 1 - j;
            int l = ipiv[j];
            if (l != j) { double t = b[l]; b[l] = b[j]; b[j] = t; }
            double bj = -b[j];
            for (int i = 1; i <= lm; i++) b[j + i] += ab[kd + i + (size_t)j * ld] * bj;
        }
    int k = kl + ku;
    for (int j = n - 1; j >= 0; j--) {
        if (b[j] != 0) {
            b[j] = b[j] / ab[kd + (size_t)j * ld];
            double temp = b[j];
            int i0 = j - k > 0 ? j - k : 0;
            for (int i = j - 1; i >= i0; i--) b[i] -= temp * ab[kd + i - j + (size_t)j * ld];
        }
    }
}

static void gbtrs_c(int n, int kl, int ku, const cplx *ab, const int *ipiv, cplx *b)
{
    int ld = 2 * kl + ku + 1, kd = ku + kl;
    if (kl > 0)
        for (int j = 0; j < n - 1; j++) {
            int lm = kl < n - 1 - j ? kl : n - 1 - j;
            int l = ipiv[j];
            if (l != j) { cplx t = b[l]; b[l] = b[j]; b[j] = t; }
            cplx bj = -b[j];
            for (int i = 1; i <= lm; i++) b[j + i] += ab[kd + i + (size_t)j * ld] * bj;
        }
    int k = kl + ku;
    for (int j = n - 1; j >= 0; j--) {
        if (b[j] != 0) {
            b[j] = b[j] / ab[kd + (size_t)j * ld];
            cplx temp = b[j];
            int i0 = j - k > 0 ? j - k : 0;
            for (int i = j - 1; i >= i0; i--) b[i] -= temp * ab[kd + i - j + (size_t)j * ld];
        }
    }
}

/* small dense LU, column major a[i + j*n], partial pivoting */
static void getrf_r(int n, double *a, int *piv, int *info)
{
    *info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        for (int i = j + 1; i < n; i++)
            if (fabs(a[i + j * n]) > fabs(a[p + j * n])) p = i;
        piv[j] = p;
        if (a[p + j * n] != 0) {
            if (p != j)
                for (int c = 0; c < n; c++) { double t = a[j + c * n]; a[j + c * n] = a[p + c * n]; a[p + c * n] = t; }
            double rp = 1.0 / a[j + j * n];
            for (int i = j + 1; i < n; i++) a[i + j * n] *= rp;
        } else if (*info == 0) *info = j + 1;
        for (int c = j + 1; c < n; c++) {
            double t = -a[j + c * n];
            for (int i = j + 1; i < n; i++) a[i + c * n] += a[i + j * n] * t;
        }
    }
}
static void getrf_c(int n, cplx *a, int *piv, int *info)
{
    *info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        for (int i = j + 1; i < n; i++)
            if (cabs1(a[i + j * n]) > cabs1(a[p + j * n])) p = i;
        piv[j] = p;
        if (a[p + j * n] != 0) {
            if (p != j)
                for (int c = 0; c < n; c++) { cplx t = a[j + c * n]; a[j + c * n] = a[p + c * n]; a[p + c * n] = t; }
            cplx rp = 1.0 / a[j + j * n];
            for (int i = j + 1; i < n; i++) a[i + j * n] *= rp;
        } else if (*info == 0) *info = j + 1;
        for (int c = j + 1; c < n; c++) {
            cplx t = -a[j + c * n];
            for (int i = j + 1; i < n; i++) a[i + c * n] += a[i + j * n] * t;
        }
    }
}
static void getrs_r(int n, const double *a, const int *piv, double *b)
{
    for (int i = 0; i < n; i++)
        if (piv[i] != i) { double t = b[i]; b[i] = b[piv[i]]; b[piv[i]] = t; }
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) b[i] -= b[j] * a[i + j * n];
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= a[j + j * n];
        for (int i = 0; i < j; i++) b[i] -= b[j] * a[i + j * n];
    }
}
static void getrs_c(int n, const cplx *a, const int *piv, cplx *b)
{
    for (int i = 0; i < n; i++)
        if (piv[i] != i) { cplx t = b[i]; b[i] = b[piv[i]]; b[piv[i]] = t; }
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) b[i] -= b[j] * a[i + j * n];
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= a[j + j * n];
        for (int i = 0; i < j; i++) b[i] -= b[j] * a[i + j * n];
    }
}

/* prepare_bordered_mat_real (algebra_lapack.f90:216-250) */
static int bord_prep_r(bordr *B)
{
    int info, nt = B->ntau, nb = B->nb;
    gbtrf_r(nb, B->kl, B->ku, B->A4, B->p4, &info);
    if (info) return info;
    for (int c = 0; c < nt; c++) gbtrs_r(nb, B->kl, B->ku, B->A4, B->p4, B->A3 + (size_t)c * nb);
    /* A1 <- A1 - A2*A3 (dgemm loop order: j, l, i) */
    for (int j = 0; j < nt; j++)
        for (int l = 0; l < nb; l++) {
            double temp = -B->A3[l + (size_t)j * nb];
            if (temp != 0)
                for (int i = 0; i < nt; i++) B->A1[i + j * nt] += temp * B->A2[i + (size_t)l * nt];
        }
    getrf_r(nt, B->A1, B->p1, &info);
    return info;
}
static int bord_prep_c(bordc *B)
{
    int info, nt = B->ntau, nb = B->nb;
    gbtrf_c(nb, B->kl, B->ku, B->A4, B->p4, &info);
    if (info) return info;
    for (int c = 0; c < nt; c++) gbtrs_c(nb, B->kl, B->ku, B->A4, B->p4, B->A3 + (size_t)c * nb);
    for (int j = 0; j < nt; j++)
        for (int l = 0; l < nb; l++) {
            cplx temp = -B->A3[l + (size_t)j * nb];
            if (temp != 0)
                for (int i = 0; i < nt; i++) B->A1[i + j * nt] += temp * B->A2[i + (size_t)l * nt];
        }
    getrf_c(nt, B->A1, B->p1, &info);
    return info;
}
/* solve_bordered_mat_real (algebra_lapack.f90:332-373) */
static void bord_solve_r(const bordr *B, double *rhs)
{
    int nt = B->ntau, nb = B->nb;
    double *r2 = rhs + nt;
    gbtrs_r(nb, B->kl, B->ku, B->A4, B->p4, r2);
    for (int j = 0; j < nb; j++) { /* dgemv N: rhs1 -= A2*r2 */
        double temp = -r2[j];
        if (temp != 0)
            for (int i = 0; i < nt; i++) rhs[i] += temp * B->A2[i + (size_t)j * nt];
    }
    getrs_r(nt, B->A1, B->p1, rhs);
    for (int j = 0; j < nt; j++) {
        double temp = -rhs[j];
        if (temp != 0)
            for (int i = 0; i < nb; i++) r2[i] += temp * B->A3[i + (size_t)j * nb];
    }
}
static void bord_solve_c(const bordc *B, cplx *rhs)
{
    int nt = B->ntau, nb = B->nb;
    cplx *r2 = rhs + nt;
    gbtrs_c(nb, B->kl, B->ku, B->A4, B->p4, r2);
    for (int j = 0; j < nb; j++) {
        cplx temp = -r2[j];
        if (temp != 0)
            for (int i = 0; i < nt; i++) rhs[i] += temp * B->A2[i + (size_t)j * nt];
    }
    getrs_c(nt, B->A1, B->p1, rhs);
    for (int j = 0; j < nt; j++) {
        cplx temp = -rhs[j];
        if (temp != 0)
            for (int i = 0; i < nb; i++) r2[i] += temp * B->A3[i + (size_t)j * nb];
    }
}

/* dgbmv 'N' with alpha=1, beta=0 on re and im separately
 * (band_matrix.f90:263-294) */
static void band_mv_r(const bandr *B, cplx *v)
{
    int n = B->n, ld = B->kl + B->ku + 1;
    double *xr = malloc(sizeof(double) * n * 4), *xi = xr + n, *yr = xr + 2 * n, *yi = xr + 3 * n;
    for (int i = 0; i < n; i++) { xr[i] = creal(v[i]); xi[i] = cimag(v[i]); yr[i] = 0; yi[i] = 0; }
    for (int j = 0; j < n; j++) {
        double tr = xr[j], ti = xi[j];
        int i0 = j - B->ku > 0 ? j - B->ku : 0, i1 = j + B->kl < n - 1 ? j + B->kl : n - 1;
        if (tr != 0)
            for (int i = i0; i <= i1; i++) yr[i] += tr * B->d[B->ku + i - j + (size_t)j * ld];
        if (ti != 0)
            for (int i = i0; i <= i1; i++) yi[i] += ti * B->d[B->ku + i - j + (size_t)j * ld];
    }
    for (int i = 0; i < n; i++) v[i] = yr[i] + I * yi[i];
    free(xr);
}
/* zgbmv (band_matrix.f90:345-366) */
static void band_mv_c(const bandc *B, cplx *v)
{
    int n = B->n, ld = B->kl + B->ku + 1;
    cplx *x = malloc(sizeof(cplx) * n);
    memcpy(x, v, sizeof(cplx) * n);
    for (int i = 0; i < n; i++) v[i] = 0;
    for (int j = 0; j < n; j++) {
        cplx t = x[j];
        if (t == 0) continue;
        int i0 = j - B->ku > 0 ? j - B->ku : 0, i1 = j + B->kl < n - 1 ? j + B->kl : n - 1;
        for (int i = i0; i <= i1; i++) v[i] += t * B->d[B->ku + i - j + (size_t)j * ld];
    }
    free(x);
}

/* ------------------------------------------------------------------------ */
/* Matrix assembly                                                          */
/* ------------------------------------------------------------------------ */

/* Fills rows i_r = n_r + nbounds of an RHS band matrix from stencils
 * (get_rhs_exp_mat, update_temp_integ.f90:778-812): row idx (0-based) =
 * i_r-1 holds rnorm*stencil; rows < nbounds stay zero. */
static void fill_rhs_band_r(or_handle *h, bandr *B, int nbounds, const ld *ext, int K)
{
    int N = h->N, ld_ = B->kl + B->ku + 1;
    ld *st = malloc(sizeof(ld) * (2 * K + 1));
    memset(B->d, 0, sizeof(double) * (size_t)ld_ * N);
    for (int idx = nbounds; idx < N; idx++) {
        ext_to_stencil(ext + (size_t)idx * (2 * K + 1), idx, K, st);
        for (int nb = 0; nb < 2 * K + 1; nb++) {
            int col = idx + K - nb;
            if (col < 0 || col >= N) continue;
            B->d[B->ku + idx - col + (size_t)col * ld_] = h->rnorm * (double)st[nb];
        }
    }
    free(st);
}

static void band_alloc_r(bandr *B, int kl, int ku, int n)
{
    B->kl = kl; B->ku = ku; B->n = n;
    B->d = calloc((size_t)(kl + ku + 1) * n, sizeof(double));
}
static void band_alloc_c(bandc *B, int kl, int ku, int n)
{
    B->kl = kl; B->ku = ku; B->n = n;
    B->d = calloc((size_t)(kl + ku + 1) * n, sizeof(cplx));
}
static void bord_alloc_r(bordr *B, int kl, int ku, int ntau, int n)
{
    B->kl = kl; B->ku = ku; B->ntau = ntau; B->nb = n - ntau;
    B->A1 = calloc(ntau * ntau, sizeof(double));
    B->A2 = calloc((size_t)ntau * B->nb, sizeof(double));
    B->A3 = calloc((size_t)ntau * B->nb, sizeof(double));
    B->A4 = calloc((size_t)(2 * kl + ku + 1) * B->nb, sizeof(double));
    B->p1 = calloc(ntau, sizeof(int));
    B->p4 = calloc(B->nb, sizeof(int));
}
static void bord_alloc_c(bordc *B, int kl, int ku, int ntau, int n)
{
    B->kl = kl; B->ku = ku; B->ntau = ntau; B->nb = n - ntau;
    B->A1 = calloc(ntau * ntau, sizeof(cplx));
    B->A2 = calloc((size_t)ntau * B->nb, sizeof(cplx));
    B->A3 = calloc((size_t)ntau * B->nb, sizeof(cplx));
    B->A4 = calloc((size_t)(2 * kl + ku + 1) * B->nb, sizeof(cplx));
    B->p1 = calloc(ntau, sizeof(int));
    B->p4 = calloc(B->nb, sizeof(int));
}

/* ext tables of the m-independent operator parts */
typedef struct {
    int K;
    ld *e;
} exttab;

typedef struct optabs_s {
    exttab I2R2, I2R1, L2_0, L2_2, R2, I1R1, I4R4, I4R3, LR_0, LR_2, LR2_0, LR2_2, LR2_4;
} optabs;

static void make_ext(or_handle *h, exttab *t, const dop *op, int J, int K)
{
    ld a = 0.5L * ((ld)h->r_cmb - (ld)h->r_icb), b = 0.5L * ((ld)h->r_cmb + (ld)h->r_icb);
    t->K = K;
    t->e = malloc(sizeof(ld) * (size_t)h->N * (2 * K + 1));
    op_ext(op, J, h->N, K, a, b, t->e);
}

static optabs *build_optabs(or_handle *h)
{
    optabs *T = calloc(1, sizeof(optabs));
    dop d;
    d = dop_poly(lp_mono(1, 2)); make_ext(h, &T->I2R2, &d, 2, 4);
    d = dop_poly(lp_mono(1, 1)); make_ext(h, &T->I2R1, &d, 2, 3);
    dop l0, l2;
    def_lapl(&l0, &l2);
    make_ext(h, &T->L2_0, &l0, 2, 2);
    make_ext(h, &T->L2_2, &l2, 2, 2);
    d = dop_poly(lp_mono(1, 2)); make_ext(h, &T->R2, &d, 0, 2);
    d = dop_poly(lp_mono(1, 1)); make_ext(h, &T->I1R1, &d, 1, 2);
    d = dop_poly(lp_mono(1, 4)); make_ext(h, &T->I4R4, &d, 4, 8);
    d = dop_poly(lp_mono(1, 3)); make_ext(h, &T->I4R3, &d, 4, 7);
    dop L0, L2;
    def_lrot((ld)h->r_cmb, &L0, &L2);
    dop r4L0 = dop_mulp(lp_mono(1, 4), &L0), r4L2 = dop_mulp(lp_mono(1, 4), &L2);
    make_ext(h, &T->LR_0, &r4L0, 4, 8);
    make_ext(h, &T->LR_2, &r4L2, 4, 8);
    /* r^4 Delta_m L_rot = (r^4 D^2 + r^3 D) L_rot - m^2 r^2 L_rot */
    dop Dm = dop_zero();
    Dm.a[2] = lp_mono(1, 4);
    Dm.a[1] = lp_mono(1, 3);
    dop A0 = dop_compose(&Dm, &L0), A2 = dop_compose(&Dm, &L2);
    dop B2 = dop_mulp(lp_mono(-1, 2), &L0), B4 = dop_mulp(lp_mono(-1, 2), &L2);
    dop C2 = A2;
    for (int k = 0; k < DOP_K; k++) C2.a[k] = lp_add(A2.a[k], B2.a[k], 1);
    make_ext(h, &T->LR2_0, &A0, 4, 8);
    make_ext(h, &T->LR2_2, &C2, 4, 8);
    make_ext(h, &T->LR2_4, &B4, 4, 8);
    return T;
}

static void free_optabs(optabs *T)
{
    exttab *list[] = {&T->I2R2, &T->I2R1, &T->L2_0, &T->L2_2, &T->R2, &T->I1R1, &T->I4R4,
                      &T->I4R3, &T->LR_0, &T->LR_2, &T->LR2_0, &T->LR2_2, &T->LR2_4};
    for (size_t i = 0; i < sizeof(list) / sizeof(list[0]); i++) free(list[i]->e);
    free(T);
}

/* stencil (reference order, length 2K+1) of a combination sum_i c_i * tab_i
 * embedded at half-bandwidth K */
static void comb_stencil(int idx, int K, int nt, const exttab **tabs, const ld *coef, ld *st)
{
    ld *row = calloc(2 * K + 1, sizeof(ld));
    for (int t = 0; t < nt; t++) {
        const exttab *E = tabs[t];
        for (int d = -E->K; d <= E->K; d++) row[K + d] += coef[t] * E->e[(size_t)idx * (2 * E->K + 1) + E->K + d];
    }
    ext_to_stencil(row, idx, K, st);
    free(row);
}

/* get_lhs_mat_tau for T and xi (update_temp_integ.f90:631-776) */
static int build_lhs_scalar(or_handle *h, optabs *T, bordr *A, double *fac, int m, double wimp,
                            double diff, int ktop, int kbot)
{
    int N = h->N, K = 4, nt = 2, nb = N - 2, ld4 = 2 * A->kl + A->ku + 1;
    ld st[9];
    memset(A->A1, 0, sizeof(double) * nt * nt);
    memset(A->A2, 0, sizeof(double) * nt * nb);
    memset(A->A3, 0, sizeof(double) * nt * nb);
    memset(A->A4, 0, sizeof(double) * ld4 * nb);
    const exttab *tabs[3] = {&T->I2R2, &T->L2_0, &T->L2_2};
    ld coef[3] = {1, -(ld)wimp * (ld)diff, -(ld)wimp * (ld)diff * (ld)m * (ld)m};
    for (int n_r = 1; n_r <= nb; n_r++) {
        int idx = n_r + nt - 1;
        comb_stencil(idx, K, 3, tabs, coef, st);
        double mx = 0;
        for (int nbnd = 1; nbnd <= 2 * K + 1; nbnd++) {
            double v = (double)st[nbnd - 1];
            int c = n_r + A->ku + 1 - nbnd;
            if (c <= nb && c >= 1) A->A4[(A->kl + nbnd - 1) + (size_t)(c - 1) * ld4] = h->rnorm * v;
            if (fabs(v) > mx) mx = fabs(v);
        }
        fac[n_r + nt - 1] = 1.0 / h->rnorm / mx;
        for (int nbnd = 1; nbnd <= A->kl; nbnd++)
            if (n_r <= nbnd && n_r + nt - nbnd >= 1)
                A->A3[(n_r - 1) + (size_t)(n_r + nt - nbnd - 1) * nb] = h->rnorm * (double)st[A->ku + nbnd];
    }
    for (int n = 0; n < N; n++) {
        double v1 = ktop == 1 ? h->rMat[0 * N + n] : h->drMat[0 * N + n];
        double v2 = kbot == 1 ? h->rMat[1 * N + n] : h->drMat[1 * N + n];
        if (n < nt) { A->A1[0 + n * nt] = h->rnorm * v1; A->A1[1 + n * nt] = h->rnorm * v2; }
        else { A->A2[0 + (n - nt) * nt] = h->rnorm * v1; A->A2[1 + (n - nt) * nt] = h->rnorm * v2; }
    }
    for (int i = 0; i < nt; i++) A->A1[i] = h->boundary_fac * A->A1[i];
    for (int i = 0; i < nt; i++) {
        double mx = 0;
        for (int c = 0; c < nb; c++) mx = fmax(mx, fabs(A->A2[i + c * nt]));
        fac[i] = 1.0 / mx;
    }
    for (int i = 0; i < nt; i++) {
        for (int c = 0; c < nt; c++) A->A1[i + c * nt] *= fac[i];
        for (int c = 0; c < nb; c++) A->A2[i + c * nt] *= fac[i];
    }
    for (int n = nt; n < N; n++)
        for (int c = 0; c < nt; c++) A->A3[(n - nt) + (size_t)c * nb] *= fac[n];
    for (int n_r = 1; n_r <= nb; n_r++)
        for (int nbnd = 1; nbnd <= 2 * K + 1; nbnd++) {
            int c = n_r + A->ku + 1 - nbnd;
            if (c <= nb && c >= 1) A->A4[(A->kl + nbnd - 1) + (size_t)(c - 1) * ld4] *= fac[n_r + nt - 1];
        }
    return bord_prep_r(A);
}

/* get_lhs_mat_tau for psi (update_psi_integ_smat.f90:1286-1581),
 * rotating container 'sphere', no full disk */
static int build_lhs_psi(or_handle *h, optabs *T, bordc *A, double *fac, int m, double wimp)
{
    int N = h->N, nt = A->ntau, nb = N - nt, K = A->ku, ld4 = 2 * A->kl + A->ku + 1;
    double nu = h->ViscFac * h->hdifV[0];
    ld *st = malloc(sizeof(ld) * (2 * K + 1)), *sc = malloc(sizeof(ld) * (2 * K + 1));
    memset(A->A1, 0, sizeof(cplx) * nt * nt);
    memset(A->A2, 0, sizeof(cplx) * nt * nb);
    memset(A->A3, 0, sizeof(cplx) * nt * nb);
    memset(A->A4, 0, sizeof(cplx) * ld4 * nb);
    const exttab *tabsA[4];
    ld cA[4];
    int ntA;
    const exttab *tabC[1] = {&T->I4R4};
    ld cC[1] = {0};
    ld M2 = (ld)m * (ld)m;
    if (m == 0) {
        tabsA[0] = &T->I2R2; tabsA[1] = &T->R2; tabsA[2] = &T->I1R1;
        cA[0] = 1; cA[1] = -(ld)wimp * (ld)nu; cA[2] = 3.0L * (ld)wimp * (ld)nu;
        ntA = 3;
    } else {
        /* LAPLROT - wimp*nu*LAPLROT2 */
        static const exttab *dummy;
        (void)dummy;
        tabsA[0] = &T->LR_0; tabsA[1] = &T->LR_2; tabsA[2] = &T->LR2_0; tabsA[3] = &T->LR2_2;
        cA[0] = 1; cA[1] = M2; cA[2] = -(ld)wimp * (ld)nu; cA[3] = -(ld)wimp * (ld)nu * M2;
        ntA = 4;
        if (h->c.l_coriolis_imp) cC[0] = (ld)wimp * (ld)h->CorFac * (ld)m;
    }
    for (int n_r = 1; n_r <= nb; n_r++) {
        int idx = n_r + nt - 1;
        comb_stencil(idx, K, ntA, tabsA, cA, st);
        if (m != 0) {
            /* the m^4 part of LAPLROT2 */
            ld tmp[17];
            const exttab *t4[1] = {&T->LR2_4};
            ld c4[1] = {-(ld)wimp * (ld)nu * M2 * M2};
            comb_stencil(idx, K, 1, t4, c4, tmp);
            for (int i = 0; i < 2 * K + 1; i++) st[i] += tmp[i];
            comb_stencil(idx, K, 1, tabC, cC, sc);
        } else {
            for (int i = 0; i < 2 * K + 1; i++) sc[i] = 0;
        }
        double mx = 0;
        for (int nbnd = 1; nbnd <= 2 * K + 1; nbnd++) {
            cplx v = (double)st[nbnd - 1] + I * (double)sc[nbnd - 1];
            int c = n_r + A->ku + 1 - nbnd;
            if (c <= nb && c >= 1) A->A4[(A->kl + nbnd - 1) + (size_t)(c - 1) * ld4] = h->rnorm * v;
            if (cabs(v) > mx) mx = cabs(v);
        }
        fac[n_r + nt - 1] = 1.0 / h->rnorm / mx;
        for (int nbnd = 1; nbnd <= A->kl; nbnd++)
            if (n_r <= nbnd && n_r + nt - nbnd >= 1)
                A->A3[(n_r - 1) + (size_t)(n_r + nt - nbnd - 1) * nb] =
                    h->rnorm * ((double)st[A->ku + nbnd] + I * (double)sc[A->ku + nbnd]);
    }
    /* tau lines */
    for (int n = 0; n < N; n++) {
        double dn2 = (double)n * (double)n;
        double d3top = 8.0 / 15.0 * dn2 * (dn2 - 1.0) * (dn2 - 4.0);
        double v[4];
        int nv;
        if (m == 0) {
            v[0] = h->c.ktopv == 1 ? (h->drMat[n] - h->or1[0] * h->rMat[n]) : h->rMat[n];
            v[1] = h->c.kbotv == 1 ? (h->drMat[N + n] - h->or1[N - 1] * h->rMat[N + n]) : h->rMat[N + n];
            nv = 2;
        } else {
            v[0] = h->rMat[n];
            v[1] = h->rMat[N + n];
            v[2] = h->c.ktopv == 1 ? (h->d2rMat[n] - h->or1[0] * h->drMat[n])
                                   : (h->c.l_non_rot ? h->drMat[n] : d3top);
            v[3] = h->c.kbotv == 1 ? (h->d2rMat[N + n] - h->or1[N - 1] * h->drMat[N + n]) : h->drMat[N + n];
            nv = 4;
        }
        for (int i = 0; i < nv; i++) {
            if (n < nt) A->A1[i + n * nt] = h->rnorm * v[i];
            else A->A2[i + (n - nt) * nt] = h->rnorm * v[i];
        }
    }
    for (int i = 0; i < nt; i++) A->A1[i] = h->boundary_fac * A->A1[i];
    for (int i = 0; i < nt; i++) {
        double mx = 0;
        for (int c = 0; c < nb; c++) mx = fmax(mx, cabs(A->A2[i + c * nt]));
        fac[i] = 1.0 / mx;
    }
    for (int i = 0; i < nt; i++) {
        for (int c = 0; c < nt; c++) A->A1[i + c * nt] *= fac[i];
        for (int c = 0; c < nb; c++) A->A2[i + c * nt] *= fac[i];
    }
    for (int n = nt; n < N; n++)
        for (int c = 0; c < nt; c++) A->A3[(n - nt) + (size_t)c * nb] *= fac[n];
    for (int n_r = 1; n_r <= nb; n_r++)
        for (int nbnd = 1; nbnd <= 2 * K + 1; nbnd++) {
            int c = n_r + A->ku + 1 - nbnd;
            if (c <= nb && c >= 1) A->A4[(A->kl + nbnd - 1) + (size_t)(c - 1) * ld4] *= fac[n_r + nt - 1];
        }
    free(st);
    free(sc);
    return bord_prep_c(A);
}

/* ------------------------------------------------------------------------ */
/* Time schemes: multistep_schemes.f90, dirk_schemes.f90                    */
/* ------------------------------------------------------------------------ */
static void ts_butcher(tscheme *t, int nst, const double *bi, const double *be)
{
    int n = nst + 1;
    t->nimp = t->nexp = t->nstages = nst;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) { t->bimp0[i][j] = bi[i * n + j]; t->bexp0[i][j] = be[i * n + j]; }
    for (int s = 0; s < nst; s++) { t->l_exp_calc[s] = 1; t->l_imp_calc_rhs[s] = 1; }
}

/* tables of a DIRK scheme with an assembly stage: nst x nst Butcher tables
 * and the assembly rows (dirk_schemes.f90:239-243: sizet = nstages) */
static void ts_butcher_ass(tscheme *t, int nst, const double *bi, const double *be, const double *bai,
                           const double *bae)
{
    t->nimp = t->nexp = t->nstages = nst;
    for (int i = 0; i < nst; i++) {
        for (int j = 0; j < nst; j++) { t->bimp0[i][j] = bi[i * nst + j]; t->bexp0[i][j] = be[i * nst + j]; }
        t->bass_imp0[i] = bai[i];
        t->bass_exp0[i] = bae[i];
    }
    for (int s = 0; s < nst; s++) { t->l_exp_calc[s] = 1; t->l_imp_calc_rhs[s] = 1; }
    t->l_assembly = 1;
}

/* initialize (multistep_schemes.f90:35-125; dirk_schemes.f90:34-263) and
 * the DIRK tables of set_weights (dirk_schemes.f90:286-525) */
static int ts_init(tscheme *t, int scheme, double courfac_nml)
{
    memset(t, 0, sizeof(*t));
    t->scheme = scheme;
    t->nold = 1;
    t->istage = 1;
    double cf;
    if (scheme >= PZ_CNAB2 && scheme <= PZ_TVB33) {
        t->family = 0;
        t->nstages = 1;
        t->l_exp_calc[0] = 1; t->l_imp_calc_rhs[0] = 1;
        switch (scheme) {
        case PZ_CNAB2: t->nimp = 1; t->nold = 1; t->nexp = 2; cf = 3.0; break;           /* :58-63 */
        case PZ_MODCNAB: t->nold = 2; t->nimp = 2; t->nexp = 2; cf = 2.9; break;         /* :64-69 */
        case PZ_CNLF: t->nold = 2; t->nimp = 2; t->nexp = 2; cf = 3.0; break;            /* :70-75 */
        case PZ_SBDF2: t->nold = 2; t->nimp = 1; t->nexp = 2; t->l_imp_calc_rhs[0] = 0; cf = 2.8; break;
        case PZ_SBDF3: t->nold = 3; t->nimp = 1; t->nexp = 3; t->l_imp_calc_rhs[0] = 0; cf = 4.0; break;
        case PZ_TVB33: t->nold = 3; t->nimp = 3; t->nexp = 3; cf = 4.0; break;           /* :90-95 */
        default: t->nold = 4; t->nimp = 1; t->nexp = 4; t->l_imp_calc_rhs[0] = 0; cf = 5.5; break; /* SBDF4 */
        }
    } else {
        t->family = 1;
        double g, r2 = sqrt(2.0);
        switch (scheme) {
        case PZ_ARS222: { /* :287-299 */
            g = 0.5 * (2.0 - r2);
            double del = 1.0 - 0.5 / g;
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - g, g}, be[9] = {0, 0, 0, g, 0, 0, del, 1 - del, 0};
            ts_butcher(t, 2, bi, be);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_ARS443: { /* :464-480 */
            double bi[25] = {0, 0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1.0 / 6.0, 0.5, 0, 0,
                             0, -0.5, 0.5, 0.5, 0, 0, 1.5, -1.5, 0.5, 0.5};
            double be[25] = {0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 11.0 / 18.0, 1.0 / 18.0, 0, 0, 0,
                             5.0 / 6.0, -5.0 / 6.0, 0.5, 0, 0, 0.25, 1.75, 0.75, -1.75, 0};
            ts_butcher(t, 4, bi, be);
            t->wimp_lin0 = 0.5; t->l_imp_calc_rhs[0] = 0; cf = 0.9;
            break;
        }
        case PZ_BPR353: { /* :481-497 */
            double bi[25] = {0, 0, 0, 0, 0, 0.5, 0.5, 0, 0, 0, 5.0 / 18.0, -1.0 / 9.0, 0.5, 0, 0,
                             0.5, 0, 0, 0.5, 0, 0.25, 0, 0.75, -0.5, 0.5};
            double be[25] = {0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 4.0 / 9.0, 2.0 / 9.0, 0, 0, 0,
                             0.25, 0, 0.75, 0, 0, 0.25, 0, 0.75, 0, 0};
            ts_butcher(t, 4, bi, be);
            t->wimp_lin0 = 0.5; t->l_exp_calc[3] = 0; cf = 1.0;
            break;
        }
        case PZ_DP1_242: { /* :393-409 */
            double bi[25] = {0, 0, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1.0 / 6.0, 0.5, 0, 0,
                             0, -0.5, 0.5, 0.5, 0, 0, 1.5, -1.5, 0.5, 0.5};
            double be[25] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0 / 3.0, 0, 0, 0,
                             0, 1, 0, 0, 0, 0, 0.5, 0, 0.5, 0};
            ts_butcher(t, 4, bi, be);
            t->wimp_lin0 = 0.5; t->l_imp_calc_rhs[0] = 0; t->l_exp_calc[0] = 0; t->l_exp_calc[2] = 0;
            cf = 1.35;
            break;
        }
        case PZ_DP2_242: { /* :410-426 */
            g = 1.0 / 3.0;
            double bi[25] = {0, 0, 0, 0, 0, 0, g, 0, 0, 0, 0, -g, g, 0, 0,
                             0, 0, 1 - g, g, 0, 0, 0, 0.5, 0.5 - g, g};
            double be[25] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                             0, 0, 1, 0, 0, 0, 0, 0.5, 0.5, 0};
            ts_butcher(t, 4, bi, be);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; t->l_exp_calc[0] = 0; cf = 1.35;
            break;
        }
        case PZ_LZ232: { /* :442-451 */
            double bi[9] = {0, 0, 0, -0.25, 0.5, 0, 0.5, 0, 0.5}, be[9] = {0, 0, 0, 0.25, 0, 0, -1, 2, 0};
            ts_butcher(t, 2, bi, be);
            t->wimp_lin0 = 0.5; cf = 1.25;
            break;
        }
        case PZ_CK232: { /* :452-463 */
            g = 1.0 - 0.5 * r2;
            double bi[9] = {0, 0, 0, -1.0 / 3.0 + 0.5 * r2, g, 0, 0.75 - 0.25 * r2, -0.75 + 0.75 * r2, g};
            double be[9] = {0, 0, 0, 2.0 / 3.0, 0, 0, 0.25, 0.75, 0};
            ts_butcher(t, 2, bi, be);
            t->wimp_lin0 = g; cf = 1.25;
            break;
        }
        case PZ_LZ453: { /* :498-513 */
            double bi[25] = {0, 0, 0, 0, 0, -44.0 / 45.0, 1.2, 0, 0, 0, -47.0 / 300.0, -0.71, 1.2, 0, 0,
                             3.375, -3.25, -59.0 / 120.0, 1.2, 0, 89.0 / 50.0, -486.0 / 55.0, 8.9, -562.0 / 275.0, 1.2};
            double be[25] = {0, 0, 0, 0, 0, 2.0 / 9.0, 0, 0, 0, 0, 71.0 / 420.0, 23.0 / 140.0, 0, 0, 0,
                             -281.0 / 336.0, 187.0 / 112.0, 0, 0, 0, 0.1, 0, 0.5, 0.4, 0};
            ts_butcher(t, 4, bi, be);
            t->wimp_lin0 = 1.2; cf = 1.15;
            break;
        }
        case PZ_PC2: { /* :514-525 */
            double bi[16] = {0, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0, 0.5, 0, 0, 0.5};
            double be[16] = {0, 0, 0, 0, 1, 0, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0, 0.5, 0};
            ts_butcher(t, 3, bi, be);
            t->wimp_lin0 = 0.5; cf = 0.8;
            break;
        }
        case PZ_SSP222: { /* :299-311 */
            g = 1.0 - 1.0 / r2;
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - 2 * g, g}, be[9] = {0, 0, 0, 0, 0, 0, 0, 1, 0};
            double ba[3] = {0, 0.5, 0.5};
            ts_butcher_ass(t, 3, bi, be, ba, ba);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_LSDIRK222: { /* :312-326 */
            g = 1.0 - 1.0 / r2;
            double chat = 1.0 / (2.0 * g);
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - g, g}, be[9] = {0, 0, 0, 0, 0, 0, 0, chat, 0};
            double ba[3] = {0, 1 - g, g};
            ts_butcher_ass(t, 3, bi, be, ba, ba);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_ARS232: { /* :427-441 */
            g = 0.5 * (2.0 - r2);
            double del = -2.0 * r2 / 3.0;
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - g, g}, be[9] = {0, 0, 0, g, 0, 0, del, 1 - del, 0};
            double ba[3] = {0, 1 - g, g};
            ts_butcher_ass(t, 3, bi, be, ba, ba);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_ARS233: { /* :327-341 */
            g = (3.0 + sqrt(3.0)) / 6.0;
            double bi[9] = {0, 0, 0, 0, g, 0, 0, 1 - 2 * g, g}, be[9] = {0, 0, 0, g, 0, 0, g - 1, 2 * (1 - g), 0};
            double ba[3] = {0, 0.5, 0.5};
            ts_butcher_ass(t, 3, bi, be, ba, ba);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_SSP322: { /* :342-356 */
            double bi[16] = {0, 0, 0, 0, 0, 0.5, 0, 0, 0, -0.5, 0.5, 0, 0, 0, 0.5, 0.5};
            double be[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
            double ba[4] = {0, 0, 0.5, 0.5};
            ts_butcher_ass(t, 4, bi, be, ba, ba);
            t->wimp_lin0 = 0.5; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_SSP332: { /* :357-372 */
            g = 1.0 - 1.0 / r2;
            double bi[16] = {0, 0, 0, 0, 0, g, 0, 0, 0, 1 - 2 * g, g, 0, 0, 0.5 - g, 0, g};
            double be[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0.25, 0.25, 0};
            double ba[4] = {0, 1.0 / 6.0, 1.0 / 6.0, 2.0 * (1.0 / 3.0)};
            ts_butcher_ass(t, 4, bi, be, ba, ba);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_SSP433: { /* :373-392; alpha, beta, eta are default-real
                             (single precision) literals in the reference */
            const double al = (double)0.24169426078821f, be_ = (double)0.06042356519705f;
            const double et = (double)0.12915286960590f;
            double bi[25] = {0, 0, 0, 0, 0, 0, al, 0, 0, 0, 0, -al, al, 0, 0,
                             0, 0, 1 - al, al, 0, 0, be_, et, 0.5 - be_ - et - al, al};
            double be[25] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                             0, 0, 1, 0, 0, 0, 0, 0.25, 0.25, 0};
            double ba[5] = {0, 0, 1.0 / 6.0, 1.0 / 6.0, 2.0 * (1.0 / 3.0)};
            ts_butcher_ass(t, 5, bi, be, ba, ba);
            t->wimp_lin0 = al; t->l_imp_calc_rhs[0] = 0; cf = 1.35;
            break;
        }
        case PZ_ARS343: { /* :526-545 */
            g = 0.435866521508459;
            const double b1 = -1.5 * g * g + 4.0 * g - 0.25, b2 = 1.5 * g * g - 5.0 * g + 1.25;
            double bi[16] = {0, 0, 0, 0, 0, g, 0, 0, 0, 0.5 * (1 - g), g, 0, 0, b1, b2, g};
            double be[16] = {0, 0, 0, 0, g, 0, 0, 0, 0.3212788860286278, 0.3966543747256017, 0, 0,
                             -0.1058582960718797, 0.5529291480359398, 0.5529291480359398, 0};
            double ba[4] = {0, b1, b2, g};
            ts_butcher_ass(t, 4, bi, be, ba, ba);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 0.7;
            break;
        }
        case PZ_MARS343: { /* :571-590 */
            g = 0.435866521508458;
            const double b1 = 1.20849664917601, b2 = -0.64436317068446;
            double bi[16] = {0, 0, 0, 0, 0, g, 0, 0, 0, 0.5 * (1 - g), g, 0, 0, b1, b2, g};
            double be[16] = {0, 0, 0, 0, 0.877173304301691, 0, 0, 0, 0.535396540307354, 0.182536720446875, 0, 0,
                             0.63041255815287, -0.83193390106308, 1.20152134291021, 0};
            double ba[4] = {0, b1, b2, g};
            ts_butcher_ass(t, 4, bi, be, ba, ba);
            t->wimp_lin0 = g; t->l_imp_calc_rhs[0] = 0; cf = 0.5;
            break;
        }
        case PZ_KC343: { /* :546-570 */
            g = 1767732205903.0 / 4055673282236.0;
            const double a30 = 1471266399579.0 / 7840856788654.0, a31 = -4482444167858.0 / 7529755066697.0;
            const double a32 = 11266239266428.0 / 11593286722821.0;
            double bi[16] = {0, 0, 0, 0, g, g, 0, 0, 2746238789719.0 / 10658868560708.0,
                             -640167445237.0 / 6845629431997.0, g, 0, a30, a31, a32, g};
            double be[16] = {0, 0, 0, 0, 1767732205903.0 / 2027836641118.0, 0, 0, 0,
                             5535828885825.0 / 10492691773637.0, 788022342437.0 / 10882634858940.0, 0, 0,
                             6485989280629.0 / 16251701735622.0, -4246266847089.0 / 9704473918619.0,
                             10755448449292.0 / 10357097424841.0, 0};
            double ba[4] = {a30, a31, a32, g};
            ts_butcher_ass(t, 4, bi, be, ba, ba);
            t->wimp_lin0 = g; cf = 0.5;
            break;
        }
        case PZ_KC564: { /* :591-625 */
            double bi[36] = {0, 0, 0, 0, 0, 0,
                             0.25, 0.25, 0, 0, 0, 0,
                             8611.0 / 62500.0, -1743.0 / 31250.0, 0.25, 0, 0, 0,
                             5012029.0 / 34652500.0, -654441.0 / 2922500.0, 174375.0 / 388108.0, 0.25, 0, 0,
                             15267082809.0 / 155376265600.0, -71443401.0 / 120774400.0, 730878875.0 / 902184768.0,
                             2285395.0 / 8070912.0, 0.25, 0,
                             82889.0 / 524892.0, 0, 15625.0 / 83664.0, 69875.0 / 102672.0, -2260.0 / 8211.0, 0.25};
            double be[36] = {0, 0, 0, 0, 0, 0,
                             0.5, 0, 0, 0, 0, 0,
                             13861.0 / 62500.0, 6889.0 / 62500.0, 0, 0, 0, 0,
                             -116923316275.0 / 2393684061468.0, -2731218467317.0 / 15368042101831.0,
                             9408046702089.0 / 11113171139209.0, 0, 0, 0,
                             -451086348788.0 / 2902428689909.0, -2682348792572.0 / 7519795681897.0,
                             12662868775082.0 / 11960479115383.0, 3355817975965.0 / 11060851509271.0, 0, 0,
                             647845179188.0 / 3216320057751.0, 73281519250.0 / 8382639484533.0,
                             552539513391.0 / 3454668386233.0, 3354512671639.0 / 8306763924573.0,
                             4040.0 / 17871.0, 0};
            double ba[6] = {82889.0 / 524892.0, 0, 15625.0 / 83664.0, 69875.0 / 102672.0, -2260.0 / 8211.0, 0.25};
            ts_butcher_ass(t, 6, bi, be, ba, ba);
            t->wimp_lin0 = 0.25; cf = 0.5;
            break;
        }
        case PZ_DBM453: { /* :626-651 */
            const double w = 0.32591194130117247;
            double bi[25] = {0, 0, 0, 0, 0,
                             -0.22284985318525410, w, 0, 0, 0,
                             -0.46801347074080545, 0.86349284225716961, w, 0, 0,
                             -0.46509906651927421, 0.81063103116959553, 0.61036726756832357, w, 0,
                             0.87795339639076675, -0.72692641526151547, 0.75204137157372720, -0.22898029400415088, w};
            double be[25] = {0, 0, 0, 0, 0,
                             0.10306208811591838, 0, 0, 0, 0,
                             -0.94124866143519894, 1.6626399742527356, 0, 0, 0,
                             -1.3670975201437765, 1.3815852911016873, 1.2673234025619065, 0, 0,
                             -0.81287582068772448, 0.81223739060505738, 0.90644429603699305, 0.094194134045674111, 0};
            double ba[5] = {0.87795339639076672, -0.72692641526151549, 0.7520413715737272, -0.22898029400415090, w};
            ts_butcher_ass(t, 5, bi, be, ba, ba);
            t->wimp_lin0 = w; cf = 0.5;
            break;
        }
        case PZ_KC785: { /* :652-711 */
            const double w = 2.0 / 9.0;
            const double c2 = 2366667076620.0 / 8822750406821.0, c3 = -257962897183.0 / 4451812247028.0;
            const double c4 = -486229321650.0 / 11227943450093.0, c5 = 621307788657.0 / 4714163060173.0;
            const double c6 = 2036305566805.0 / 6583108094622.0;
            const double b2 = 3517720773327.0 / 20256071687669.0, b3 = 4569610470461.0 / 17934693873752.0;
            const double b4 = 2819471173109.0 / 11655438449929.0, b5 = 3296210113763.0 / 10722700128969.0;
            const double b6 = -1142099968913.0 / 5710983926999.0;
            double bi[64] = {0, 0, 0, 0, 0, 0, 0, 0,
                             w, w, 0, 0, 0, 0, 0, 0,
                             c2, c2, w, 0, 0, 0, 0, 0,
                             c3, c3, 128530224461.0 / 14379561246022.0, w, 0, 0, 0, 0,
                             c4, c4, -225633144460.0 / 6633558740617.0, 1741320951451.0 / 6824444397158.0, w, 0, 0, 0,
                             c5, c5, -125196015625.0 / 3866852212004.0, 940440206406.0 / 7593089888465.0,
                             961109811699.0 / 6734810228204.0, w, 0, 0,
                             c6, c6, -3039402635899.0 / 4450598839912.0, -1829510709469.0 / 31102090912115.0,
                             -286320471013.0 / 6931253422520.0, 8651533662697.0 / 9642993110008.0, w, 0,
                             0, 0, b2, b3, b4, b5, b6, w};
            const double e7 = 2193717860234.0 / 3570523412979.0;
            double be[64] = {0, 0, 0, 0, 0, 0, 0, 0,
                             4.0 / 9.0, 0, 0, 0, 0, 0, 0, 0,
                             1.0 / 9.0, 1183333538310.0 / 1827251437969.0, 0, 0, 0, 0, 0, 0,
                             895379019517.0 / 9750411845327.0, 477606656805.0 / 13473228687314.0,
                             -112564739183.0 / 9373365219272.0, 0, 0, 0, 0, 0,
                             -4458043123994.0 / 13015289567637.0, -2500665203865.0 / 9342069639922.0,
                             983347055801.0 / 8893519644487.0, 2185051477207.0 / 2551468980502.0, 0, 0, 0, 0,
                             -167316361917.0 / 17121522574472.0, 1605541814917.0 / 7619724128744.0,
                             991021770328.0 / 13052792161721.0, 2342280609577.0 / 11279663441611.0,
                             3012424348531.0 / 12792462456678.0, 0, 0, 0,
                             6680998715867.0 / 14310383562358.0, 5029118570809.0 / 3897454228471.0,
                             2415062538259.0 / 6382199904604.0, -3924368632305.0 / 6964820224454.0,
                             -4331110370267.0 / 15021686902756.0, -3944303808049.0 / 11994238218192.0, 0, 0,
                             e7, e7, 5952760925747.0 / 18750164281544.0, -4412967128996.0 / 6196664114337.0,
                             4151782504231.0 / 36106512998704.0, 572599549169.0 / 6265429158920.0,
                             -457874356192.0 / 11306498036315.0, 0};
            double ba[8] = {0, 0, b2, b3, b4, b5, b6, w};
            ts_butcher_ass(t, 8, bi, be, ba, ba);
            t->wimp_lin0 = w; cf = 0.5;
            break;
        }
        default:
            return -1;
        }
    }
    t->courfac = fabs(courfac_nml) >= 1.0e3 ? cf : courfac_nml;
    return 0;
}

/* multistep weights for the dt history (multistep_schemes.f90:137-305) */
static void ms_weights(int scheme, double alpha, const double *dt, double *wimp, double *wl, double *wexp)
{
    double d0 = dt[0];
    if (scheme == PZ_CNLF) { /* :158-167 */
        double de = dt[0] / dt[1];
        wimp[0] = (1.0 - de) * (1.0 + de);
        wimp[1] = de * de;
        wl[0] = 0.5 * (1.0 + de) / de * d0;
        wl[1] = 0.5 * (1.0 + de) * (de - 1.0) / de * d0;
        wl[2] = 0.5 * (1.0 + de) * d0;
        wexp[0] = (1.0 + de) * d0;
        wexp[1] = 0.0;
    } else if (scheme == PZ_MODCNAB) { /* :168-177 */
        double de = dt[0] / dt[1];
        wimp[0] = 1.0;
        wimp[1] = 0.0;
        wl[0] = (0.5 + 1.0 / de / 16.0) * d0;
        wl[1] = (7.0 / 16.0 - 1.0 / de / 16.0) * d0;
        wl[2] = 1.0 / 16.0 * d0;
        wexp[0] = (1.0 + 0.5 * de) * d0;
        wexp[1] = -0.5 * de * d0;
    } else if (scheme == PZ_SBDF2) { /* :178-186 */
        double de = dt[0] / dt[1];
        wl[0] = (1.0 + de) / (1.0 + 2.0 * de) * d0;
        wl[1] = 0.0;
        wimp[0] = (1.0 + de) * (1.0 + de) / (1.0 + 2.0 * de);
        wimp[1] = -de * de / (1.0 + 2.0 * de);
        wexp[0] = (1.0 + de) * (1.0 + de) * d0 / (1.0 + 2.0 * de);
        wexp[1] = -de * (1.0 + de) * d0 / (1.0 + 2.0 * de);
    } else if (scheme == PZ_SBDF3) { /* :187-212 */
        double dn = dt[1] / dt[0], dn1 = dt[2] / dt[0];
        double a0 = 1.0 + 1.0 / (1.0 + dn) + 1.0 / (1.0 + dn + dn1);
        double a1 = (1.0 + dn) * (1.0 + dn + dn1) / (dn * (dn + dn1));
        double a2 = -(1.0 + dn + dn1) / (dn * dn1 * (1.0 + dn));
        double a3 = (1.0 + dn) / (dn1 * (dn + dn1) * (1.0 + dn + dn1));
        double b0 = (1.0 + dn) * (1.0 + dn + dn1) / (dn * (dn + dn1));
        double b1 = -(1.0 + dn + dn1) / (dn * dn1);
        double b2 = (1.0 + dn) / (dn1 * (dn + dn1));
        wl[0] = 1.0 / a0 * d0; wl[1] = 0.0;
        wimp[0] = a1 / a0; wimp[1] = a2 / a0; wimp[2] = a3 / a0;
        wexp[0] = b0 / a0 * d0; wexp[1] = b1 / a0 * d0; wexp[2] = b2 / a0 * d0;
    } else if (scheme == PZ_TVB33) { /* :213-263 */
        double gam = (2619.0 - sqrt(5995281.0)) / 345.0;
        double th = (3525.0 * gam * gam - 3118.0 * gam - 5084.0) / 5468.0;
        double c = (2178.0 * th - 959.0 * gam * gam + 130.0 * gam + 726.0) / 4096.0;
        double dn = dt[0] / dt[1], dn1 = dt[1] / dt[2];
        double a3 = -pow(dn1, 3) * dn * dn * (3.0 * gam * gam * dn + 2.0 * gam * (1.0 - dn) - 1.0) / (1.0 + dn1) /
                    (1.0 + dn1 * (1.0 + dn));
        double a2 = -dn * dn * (gam * dn * dn1 * (2.0 - 3.0 * gam) + (1.0 - 2.0 * gam) * (1.0 + dn1)) / (1.0 + dn);
        double a1 = -(dn1 * (1.0 + gam * dn) * (1.0 + dn * (3.0 * gam - 2.0)) + dn * (2.0 * gam - 1.0) + 1.0) /
                        (1.0 + dn1) - th;
        double a0 = th + (1.0 + 2.0 * gam * dn + dn1 * (1.0 + gam * dn) * (1.0 + 3.0 * gam * dn)) / (1.0 + dn) /
                             (1.0 + dn1 * (1.0 + dn));
        double b2 = dn1 * dn1 * dn * (6.0 * gam * (1.0 + gam * dn) + th * (3.0 + 2.0 * dn)) / (6.0 * (1.0 + dn1));
        double b1 = -gam * dn * (1.0 + dn1 * (1.0 + gam * dn)) - th / 6.0 * dn * (3.0 + dn1 * (3.0 + 2.0 * dn));
        double b0 = (1.0 + gam * dn) * (1.0 + dn1 * (1.0 + gam * dn)) / (1.0 + dn1) +
                    th * (1.0 + 0.5 * dn + dn1 * dn * (3.0 + 2.0 * dn) / (6.0 * (1.0 + dn1)));
        double c3 = th * dn1 * dn1 * dn * (3.0 + 2.0 * dn) / (6.0 * (1.0 + dn1)) - c;
        double c2 = (c * (1.0 + dn1) * (1.0 + dn1 * (1.0 + dn)) - dn1 * dn1 * dn * dn * gam * (1.0 - gam)) /
                        (dn1 * dn1 * (1.0 + dn)) - th * dn * (3.0 + dn1 * (3.0 + 2.0 * dn)) / 6.0;
        double c1 = (dn1 * dn1 * dn * (1.0 - gam) * (1.0 + gam * dn) - c * (1.0 + dn1 * (1.0 + dn))) / (dn1 * dn1) / dn +
                    th * (1.0 + 0.5 * dn + dn1 * dn * (3.0 + 2.0 * dn) / (6.0 * (1.0 + dn1)));
        double c0 = (dn1 * dn1 * dn * gam * (1.0 + gam * dn) + c * (1.0 + dn1)) / (dn1 * dn1 * dn * (1.0 + dn));
        wl[0] = c0 / a0 * d0; wl[1] = c1 / a0 * d0; wl[2] = c2 / a0 * d0; wl[3] = c3 / a0 * d0;
        wimp[0] = -a1 / a0; wimp[1] = -a2 / a0; wimp[2] = -a3 / a0;
        wexp[0] = b0 / a0 * d0; wexp[1] = b1 / a0 * d0; wexp[2] = b2 / a0 * d0;
    } else if (scheme == PZ_SBDF4) { /* :264-302 */
        double dn = dt[0] / dt[1], dn1 = dt[1] / dt[2], dn2 = dt[2] / dt[3];
        double c1 = 1.0 + dn2 * (1.0 + dn1), c2 = 1.0 + dn1 * (1.0 + dn), c3 = 1.0 + dn2 * c2;
        double a4 = (1.0 + dn) / (1.0 + dn2) * c2 / c1 / c3 * (pow(dn2, 4) * pow(dn1, 3) * dn * dn);
        double a3 = -pow(dn1, 3) * dn * dn * (1.0 + dn) / (1.0 + dn1) * c3 / c2;
        double a2 = dn * (dn / (1.0 + dn) + dn1 * dn * (c3 + dn2) / (1.0 + dn2));
        double a1 = -1.0 - dn * (1.0 + dn1 * (1.0 + dn) * (1.0 + dn2 * c2 / c1) / (1.0 + dn1));
        double a0 = 1.0 + dn / (1.0 + dn) + dn1 * dn / c2 + dn2 * dn1 * dn / c3;
        double b3 = -pow(dn2, 3) * dn1 * dn1 * dn * (1.0 + dn) / (1.0 + dn2) * c2 / c1;
        double b2 = dn1 * dn1 * dn * (1.0 + dn) / (1.0 + dn1) * c3;
        double b1 = -c2 * c3 * dn / (1.0 + dn2);
        double b0 = dn1 * (1.0 + dn) / (1.0 + dn1) * ((1.0 + dn) * (c3 + dn2) + (1.0 + dn2) / dn1) / c1;
        wl[0] = 1.0 / a0 * d0; wl[1] = 0.0;
        wimp[0] = -a1 / a0; wimp[1] = -a2 / a0; wimp[2] = -a3 / a0; wimp[3] = -a4 / a0;
        wexp[0] = b0 / a0 * d0; wexp[1] = b1 / a0 * d0; wexp[2] = b2 / a0 * d0; wexp[3] = b3 / a0 * d0;
    } else { /* CNAB2, :151-157 */
        wimp[0] = 1.0;
        wl[0] = alpha * d0;
        wl[1] = (1 - alpha) * d0;
        wexp[0] = (1.0 + 0.5 * dt[0] / dt[1]) * d0;
        wexp[1] = -0.5 * dt[0] * dt[0] / dt[1];
    }
}

/* set_weights (multistep_schemes.f90:137-305; dirk_schemes.f90:275-713) */
static void ts_set_weights(tscheme *t, double alpha, int *lMatNext)
{
    double wimp_old = t->wimp_lin[0];
    if (t->family == 0) {
        ms_weights(t->scheme, alpha, t->dt, t->wimp, t->wimp_lin, t->wexp);
        if (t->wimp_lin[0] != wimp_old) *lMatNext = 1;
    } else {
        t->wimp_lin[0] = t->dt[0] * t->wimp_lin0;
        for (int i = 0; i < 8; i++) {
            for (int j = 0; j < 8; j++) {
                t->bimp[i][j] = t->dt[0] * t->bimp0[i][j];
                t->bexp[i][j] = t->dt[0] * t->bexp0[i][j];
            }
            t->bass_imp[i] = t->dt[0] * t->bass_imp0[i];
            t->bass_exp[i] = t->dt[0] * t->bass_exp0[i];
        }
    }
}

/* bridge_with_cnab2 (multistep_schemes.f90:500-524): CNAB2 weights, the
 * remaining ones zero */
static void ts_bridge(tscheme *t, double alpha)
{
    int save = t->scheme;
    t->scheme = PZ_CNAB2;
    ms_weights(PZ_CNAB2, alpha, t->dt, t->wimp, t->wimp_lin, t->wexp);
    t->scheme = save;
    for (int i = 1; i < 8; i++) t->wimp[i] = 0.0;
    for (int i = 2; i < 8; i++) { t->wimp_lin[i] = 0.0; t->wexp[i] = 0.0; }
}

/* the unscaled tables of a DIRK scheme: nstages, whether it has an
 * assembly stage, the Butcher tables (8 x 8 row-major) and the assembly
 * rows (8); tests check them against the order conditions */
int or_dirk_tables(int scheme, int *nstages, int *l_assembly, double *wimp_lin0, double *bimp, double *bexp,
                   double *bass_imp, double *bass_exp)
{
    tscheme t;
    if (ts_init(&t, scheme, 1.0e4) || t.family != 1) return -1;
    *nstages = t.nstages;
    *l_assembly = t.l_assembly;
    *wimp_lin0 = t.wimp_lin0;
    for (int i = 0; i < 8; i++) {
        for (int j = 0; j < 8; j++) { bimp[i * 8 + j] = t.bimp0[i][j]; bexp[i * 8 + j] = t.bexp0[i][j]; }
        bass_imp[i] = t.bass_imp0[i];
        bass_exp[i] = t.bass_exp0[i];
    }
    return 0;
}

int or_scheme_weights(int scheme, double alpha, const double *dt, int *family, int *nstages,
                      double *wimp, double *wimp_lin, double *wexp, double *bimp, double *bexp)
{
    tscheme t;
    if (ts_init(&t, scheme, 1.0e4)) return -1;
    for (int i = 0; i < 4; i++) t.dt[i] = dt[i];
    int dummy = 0;
    ts_set_weights(&t, alpha, &dummy);
    *family = t.family;
    *nstages = t.nstages;
    for (int i = 0; i < 8; i++) { wimp[i] = t.wimp[i]; wimp_lin[i] = t.wimp_lin[i]; wexp[i] = t.wexp[i]; }
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) { bimp[i * 8 + j] = t.bimp[i][j]; bexp[i * 8 + j] = t.bexp[i][j]; }
    return 0;
}

/* set_dt_array (multistep_schemes.f90:307-348; dirk_schemes.f90:715-771) */
static int ts_set_dt_array(tscheme *t, double dt_new, double dt_min)
{
    if (t->family == 0) {
        double last = t->dt[t->nexp - 1];
        for (int i = t->nexp - 1; i > 0; i--) t->dt[i] = t->dt[i - 1];
        (void)last;
    }
    t->dt[0] = dt_new;
    if (dt_new < dt_min) return -2;
    return 0;
}

/* set_imex_rhs (multistep_schemes.f90:366-413; dirk_schemes.f90:773-814) */
static void ts_set_imex_rhs(or_handle *h, const tscheme *t, const tarray *a, cplx *rhs)
{
    size_t F = (size_t)h->nm * h->N;
    if (t->family == 0) {
        for (size_t i = 0; i < F; i++) rhs[i] = t->wimp[0] * a->old[i];
        for (int o = 1; o < t->nold; o++)
            for (size_t i = 0; i < F; i++) rhs[i] += t->wimp[o] * a->old[o * F + i];
        for (int o = 0; o < t->nimp; o++)
            for (size_t i = 0; i < F; i++) rhs[i] += t->wimp_lin[o + 1] * a->impl[o * F + i];
        for (int o = 0; o < t->nexp; o++)
            for (size_t i = 0; i < F; i++) rhs[i] += t->wexp[o] * a->expl[o * F + i];
    } else {
        int is = t->istage;
        for (size_t i = 0; i < F; i++) rhs[i] = a->old[i];
        for (int s = 0; s < is; s++)
            for (size_t i = 0; i < F; i++) rhs[i] += t->bexp[is][s] * a->expl[s * F + i];
        for (int s = 0; s < is; s++)
            for (size_t i = 0; i < F; i++) rhs[i] += t->bimp[is][s] * a->impl[s * F + i];
    }
}

/* rotate_imex (multistep_schemes.f90:462-498); a no-op for DIRK */
static void ts_rotate(or_handle *h, const tscheme *t, tarray *a)
{
    if (t->family != 0) return;
    size_t F = (size_t)h->nm * h->N;
    for (int o = t->nexp - 1; o >= 1; o--) memcpy(a->expl + o * F, a->expl + (o - 1) * F, sizeof(cplx) * F);
    for (int o = t->nold - 1; o >= 1; o--) memcpy(a->old + o * F, a->old + (o - 1) * F, sizeof(cplx) * F);
    for (int o = t->nimp - 1; o >= 1; o--) memcpy(a->impl + o * F, a->impl + (o - 1) * F, sizeof(cplx) * F);
}

static void tarray_alloc(or_handle *h, tarray *a, const tscheme *t)
{
    size_t F = (size_t)h->nm * h->N;
    /* dt_fieldsLast.f90:74-80 with the argument order of time_array.f90:28 */
    a->nold = t->nold; a->nimp = t->nimp; a->nexp = t->nexp;
    a->old = calloc(F * t->nold, sizeof(cplx));
    a->impl = calloc(F * t->nimp, sizeof(cplx));
    a->expl = calloc(F * t->nexp, sizeof(cplx));
}

/* ------------------------------------------------------------------------ */
/* Setup                                                                     */
/* ------------------------------------------------------------------------ */
static double *dalloc(int n) { return calloc(n, sizeof(double)); }

/* get_conducting_state, 2-D branch (radial.f90:284-397) */
static void conducting_state(or_handle *h, double cond_fac, int ktop, int kbot, const double *t_top,
                             const double *t_bot, double *tc, double *dtc)
{
    int N = h->N;
    double f_bot = 0, f_top = 0, epsc0 = 0.0;
    for (int n = 0; n < PZ_N_T_BOUNDS; n++) {
        int m_bot = (int)t_bot[3 * n], m_top = (int)t_top[3 * n];
        double tr_bot = t_bot[3 * n + 1], tr_top = t_top[3 * n + 1];
        if (m_bot == 0 && fabs(tr_bot) > 10.0 * 2.220446049250313e-16) f_bot = tr_bot;
        if (m_top == 0 && fabs(tr_top) > 10.0 * 2.220446049250313e-16) f_top = tr_top;
    }
    double rc = h->r_cmb, ri = h->r_icb, rr = h->c.radratio;
    for (int n = 0; n < N; n++) {
        double r = h->r[n], or1 = h->or1[n];
        if (ktop == 1 && kbot == 1) {
            tc[n] = cond_fac * (log(r / rc) / log(rr));
            dtc[n] = cond_fac * (or1 / log(rr));
        } else if (ktop == 2 && kbot == 1) {
            tc[n] = cond_fac * (f_bot + f_top * rc * log(r / ri) + 0.25 * epsc0 * (ri * ri - r * r + 2.0 * rc * rc * log(r / ri)));
            dtc[n] = cond_fac * (f_top * rc * or1 + 0.5 * epsc0 * (rc * rc * or1 - r));
        } else if (ktop == 1 && kbot == 2) {
            tc[n] = cond_fac * (f_bot * ri * log(r / rc) + 0.25 * epsc0 * (rc * rc - r * r + 2.0 * ri * ri * log(r / rc)) + f_top);
            dtc[n] = cond_fac * (f_bot * ri * or1 + 0.5 * epsc0 * (ri * ri * or1 - r));
        } else {
            double e0 = -2.0 * (rc * f_top - ri * f_bot) / (rc * rc - ri * ri);
            if (fabs(e0) < 10.0 * 2.220446049250313e-16) e0 = 0.0;
            tc[n] = cond_fac * (0.5 * e0 * (rc * rc * log(r / rc) - 0.5 * (r * r - rc * rc)) + rc * f_top * log(r / rc));
            dtc[n] = cond_fac * (f_top * rc * or1 + 0.5 * e0 * (rc * rc * or1 - r));
        }
    }
}

static void bc_pattern(or_handle *h, const double *tb, const double *tt, cplx *bot, cplx *top)
{
    /* horizontal.f90:149-173 */
    for (int nm = 0; nm < h->nm; nm++) {
        int m = h->idx2m[nm];
        bot[nm] = 0;
        top[nm] = 0;
        for (int n = 0; n < PZ_N_T_BOUNDS; n++) {
            int m_bot = (int)tb[3 * n], m_top = (int)tt[3 * n];
            double tr_bot = tb[3 * n + 1], ti_bot = tb[3 * n + 2];
            double tr_top = tt[3 * n + 1], ti_top = tt[3 * n + 2];
            if (m_bot == m && (tr_bot != 0 || ti_bot != 0)) { if (m == 0) ti_bot = 0; bot[nm] = tr_bot + I * ti_bot; }
            if (m_top == m && (tr_top != 0 || ti_top != 0)) { if (m == 0) ti_top = 0; top[nm] = tr_top + I * ti_top; }
        }
    }
}

static int build_matrices(or_handle *h, int lhs_only);
static void coll_alloc(or_handle *h);
static int fd_grid(or_handle *h, double ratio1, double ratio2);
static void fd_alloc(or_handle *h);

int or_create(const pz_config *cfg, or_handle **out)
{
    *out = NULL;
    /* the influence-matrix psi solve is restated for collocation */
    if (!cfg->l_direct_solve && (!cfg->l_cheb_coll || cfg->l_finite_diff)) return -11;
    /* non-rotating runs change the integration operators (h^2 = 1); they
       are restated for collocation and finite differences */
    if (cfg->l_non_rot && !cfg->l_cheb_coll && !cfg->l_finite_diff) return -12;
    if (cfg->radratio == 0.0) return -13;
    or_handle *h = calloc(1, sizeof(or_handle));
    h->c = *cfg;
    /* namelists.f90:448-457: the influence-matrix method treats Coriolis
       explicitly */
    if (!cfg->l_direct_solve && !cfg->l_non_rot) h->c.l_coriolis_imp = 0;
    h->l_bridge = 1; /* namelists.f90:365: a fresh start bridges with CNAB2 */
    int N = h->N = cfg->n_r_max;
    h->mmax = cfg->m_max;
    h->minc = cfg->minc;
    h->nm = cfg->m_max / cfg->minc + 1;
    h->nphi = 3 * cfg->m_max / cfg->minc;
    h->ncheb = cfg->n_cheb_max;
    h->idx2m = calloc(h->nm, sizeof(int));
    for (int i = 0; i < h->nm; i++) h->idx2m[i] = i * cfg->minc;
    h->coll = cfg->l_cheb_coll != 0;
    h->fd = cfg->l_finite_diff != 0;
    if (h->fd) h->coll = 0;
    if (ts_init(&h->ts, cfg->time_scheme, cfg->courfac)) { free(h); return -14; }
    /* the assembly stage is restated for collocation and finite differences */
    if (h->ts.l_assembly && !h->coll && !h->fd) { free(h); return -15; }
    /* the assembly stage of the influence-matrix method is not restated */
    if (h->ts.l_assembly && !cfg->l_direct_solve) { free(h); return -15; }
    /* namelists.f90:398-526 */
    h->l_heat = cfg->ra != 0.0;
    h->l_chem = cfg->raxi != 0.0;
    h->tadvz = cfg->l_temp_advz ? 1.0 : 0.0;
    h->r_cmb = 1.0 / (1.0 - cfg->radratio);
    h->r_icb = h->r_cmb - 1.0;
    /* namelists.f90:312-315: ek = -1 flags a non-rotating run */
    h->ek = cfg->l_non_rot ? -1.0 : cfg->ek;
    if (cfg->l_non_rot) {
        /* namelists.f90:473-489 */
        h->CorFac = 0.0;
        h->BuoFac = h->l_heat ? cfg->ra / cfg->pr : 0.0;
        h->TdiffFac = h->l_heat ? 1.0 / cfg->pr : 0.0;
        h->ChemFac = h->l_chem ? cfg->raxi / cfg->sc : 0.0;
        h->XidiffFac = h->l_chem ? 1.0 / cfg->sc : 0.0;
        h->ViscFac = 1.0;
    } else if (cfg->time_scale_rot) {
        h->CorFac = 2.0;
        h->BuoFac = h->l_heat ? cfg->ra * cfg->ek * cfg->ek / cfg->pr : 0.0;
        h->TdiffFac = h->l_heat ? cfg->ek / cfg->pr : 0.0;
        h->ChemFac = h->l_chem ? cfg->raxi * cfg->ek * cfg->ek / cfg->sc : 0.0;
        h->XidiffFac = h->l_chem ? cfg->ek / cfg->sc : 0.0;
        h->ViscFac = cfg->ek;
    } else {
        h->CorFac = 2.0 / cfg->ek;
        h->BuoFac = h->l_heat ? cfg->ra / cfg->pr : 0.0;
        h->TdiffFac = h->l_heat ? 1.0 / cfg->pr : 0.0;
        h->ChemFac = h->l_chem ? cfg->raxi / cfg->sc : 0.0;
        h->XidiffFac = h->l_chem ? 1.0 / cfg->sc : 0.0;
        h->ViscFac = 1.0;
    }
    if (or_plant == 4) h->ChemFac = -h->ChemFac;
    if (or_plant == 5) h->XidiffFac = 2.0 * h->XidiffFac;
    /* chebyshev.f90:55-60, dct_fftw.f90:124 */
    h->rnorm = sqrt(2.0 / (double)(N - 1));
    h->boundary_fac = 0.5;
    h->cheb_fac = sqrt(0.5 / (double)(N - 1));
    /* radial_derivatives.f90:61-66 */
    h->thr = cfg->l_rerror_fix ? cfg->rerror_fac * 2.220446049250313e-16 : 0.0;
    /* dct_fftw.f90:79-89 (n_cheb_max = n_r_max in chebt) */
    int L = 2 * N - 2;
    h->der = calloc(L, sizeof(int));
    h->der2 = calloc(L, sizeof(int));
    for (int k = 1; k <= N - 1; k++) h->der[L - k] = -k;
    for (int k = 0; k < N; k++) h->der[k] = k;
    for (int p = 0; p < L; p++) h->der2[p] = h->der[p] * h->der[p];
    h->der[N - 1] = 0;
    h->fdct = fft_new(L);
    h->fphi = fft_new(h->nphi);
    int big = L > h->nphi ? L : h->nphi;
    h->wbig = big;
    /* grid: chebyshev.f90:188-240 */
    double *arr[] = {0};
    (void)arr;
    h->r = dalloc(N); h->x = dalloc(N); h->drx = dalloc(N); h->ddrx = dalloc(N);
    h->or1 = dalloc(N); h->or2 = dalloc(N); h->rgrav = dalloc(N); h->beta = dalloc(N);
    h->ekpump = dalloc(N); h->ekp_up = dalloc(N); h->ekp_us = dalloc(N); h->ekp_dusdp = dalloc(N);
    h->height = dalloc(N); h->oheight = dalloc(N); h->tcond = dalloc(N); h->dtcond = dalloc(N);
    h->xicond = dalloc(N); h->dxicond = dalloc(N); h->delxr2 = dalloc(N); h->delxh2 = dalloc(N);
    h->mR = calloc(N, sizeof(int));
    double pi = PI_F, bma = 0.5 * (h->r_cmb - h->r_icb), bpa = 0.5 * (h->r_icb + h->r_cmb);
    for (int k = 0; k < N; k++) {
        h->x[k] = cos(pi * (double)k / (double)(N - 1));
        h->r[k] = bma * h->x[k] + bpa;
        h->drx[k] = 2.0 / (h->r_cmb - h->r_icb);
        h->ddrx[k] = 0.0;
    }
    if (h->fd && fd_grid(h, cfg->fd_stretch, cfg->fd_ratio)) { free(h); return -15; }
    /* tau lines, chebyshev.f90:362-381 */
    h->rMat = dalloc(2 * N); h->drMat = dalloc(2 * N); h->d2rMat = dalloc(2 * N);
    for (int n = 0; n < N; n++) {
        double dn2 = (double)n * (double)n, monen = (n % 2 == 0) ? 1.0 : -1.0;
        h->rMat[n] = 1.0; h->rMat[N + n] = monen;
        h->drMat[n] = 2.0 * dn2; h->drMat[N + n] = 2.0 * dn2 * (-1.0) * monen;
        h->d2rMat[n] = 4.0 * (1.0 / 3.0) * dn2 * (dn2 - 1.0);
        h->d2rMat[N + n] = 4.0 * (1.0 / 3.0) * dn2 * (dn2 - 1.0) * monen;
    }
    /* radial.f90:115-282 */
    double ek_pump_fac = (cfg->l_ek_pump && !cfg->l_non_rot) ? 1.0 : 0.0;
    double rc = h->r_cmb;
    for (int n = 0; n < N; n++) {
        h->or1[n] = 1.0 / h->r[n];
        h->or2[n] = h->or1[n] * h->or1[n];
        h->rgrav[n] = cfg->g0 + cfg->g1 * h->r[n] / rc + cfg->g2 * rc * rc * h->or2[n];
        h->mR[n] = cfg->m_max;
    }
    h->delxh2[0] = pow(2.0 * pi / 3.0 / h->mR[0] * rc, 2);
    h->delxh2[N - 1] = pow(2.0 * pi / 3.0 / h->mR[N - 1] * h->r_icb, 2);
    h->delxr2[0] = pow(h->r[0] - h->r[1], 2);
    h->delxr2[N - 1] = pow(h->r[N - 2] - h->r[N - 1], 2);
    for (int n = 1; n < N - 1; n++) {
        h->delxh2[n] = pow(2.0 * pi / 3.0 / h->mR[n] * h->r[n], 2);
        double dm = fmin(h->r[n - 1] - h->r[n], h->r[n] - h->r[n + 1]);
        h->delxr2[n] = dm * dm;
    }
    if (h->l_heat) conducting_state(h, cfg->tcond_fac, cfg->ktopt, cfg->kbott, cfg->t_top, cfg->t_bot, h->tcond, h->dtcond);
    if (h->l_chem) conducting_state(h, cfg->xicond_fac, cfg->ktopxi, cfg->kbotxi, cfg->xi_top, cfg->xi_bot, h->xicond, h->dxicond);
    for (int n = 1; n < N; n++) {
        double s2 = rc * rc - h->r[n] * h->r[n];
        h->height[n] = 2.0 * sqrt(s2);
        h->oheight[n] = 0.5 / sqrt(s2);
        h->beta[n] = -h->r[n] / s2;
        h->ekpump[n] = 0.5 * ek_pump_fac * sqrt(cfg->ek * rc) / pow(s2, 0.75);
    }
    for (int n = 0; n < N; n++) {
        h->ekp_up[n] = -0.5 * h->beta[n];
        h->ekp_us[n] = -5.0 * rc * h->beta[n] * h->oheight[n];
        h->ekp_dusdp[n] = h->beta[n];
    }
    if (cfg->l_non_rot)
        /* radial.f90:197-207 */
        for (int n = 0; n < N; n++) {
            h->beta[n] = h->ekpump[n] = h->ekp_up[n] = h->ekp_us[n] = h->ekp_dusdp[n] = 0.0;
            h->height[n] = h->oheight[n] = 1.0;
        }
    /* horizontal.f90:69-173 (no hyperdiffusion) */
    h->hdifT = dalloc(h->nm); h->hdifXi = dalloc(h->nm); h->hdifV = dalloc(h->nm);
    for (int i = 0; i < h->nm; i++) h->hdifT[i] = h->hdifXi[i] = h->hdifV[i] = 1.0;
    h->topt = calloc(h->nm, sizeof(cplx)); h->bott = calloc(h->nm, sizeof(cplx));
    h->topxi = calloc(h->nm, sizeof(cplx)); h->botxi = calloc(h->nm, sizeof(cplx));
    bc_pattern(h, cfg->t_bot, cfg->t_top, h->bott, h->topt);
    /* the composition pattern is read from t_bot/t_top (horizontal.f90:176-198) */
    bc_pattern(h, cfg->t_bot, cfg->t_top, h->botxi, h->topxi);
    /* fields */
    size_t F = (size_t)h->nm * N;
    cplx **fl[] = {&h->us, &h->up, &h->om, &h->psi, &h->temp, &h->xi, &h->temp_hat, &h->xi_hat,
                   &h->psi_hat, &h->dtemp, &h->dxi, &h->work, &h->dVsT, &h->dVsXi, &h->dVsOm, &h->buo};
    for (size_t i = 0; i < sizeof(fl) / sizeof(fl[0]); i++) *fl[i] = calloc(F, sizeof(cplx));
    tarray_alloc(h, &h->dTdt, &h->ts);
    tarray_alloc(h, &h->dxidt, &h->ts);
    tarray_alloc(h, &h->dpsidt, &h->ts);
    h->dtr = dalloc(N);
    h->dth = dalloc(N);
    /* matrices */
    if (h->fd) {
        fd_alloc(h);
        h->lMatNext = 1;
        *out = h;
        return 0;
    }
    if (h->coll) {
        coll_alloc(h);
        h->lMatNext = 1;
        *out = h;
        return 0;
    }
    band_alloc_r(&h->T_RHSE[0], 3, 3, N);
    band_alloc_r(&h->T_RHSE[1], 4, 4, N);
    h->T_RHSI = calloc(h->nm, sizeof(bandr));
    h->T_LHS = calloc(h->nm, sizeof(bordr));
    h->XI_LHS = calloc(h->nm, sizeof(bordr));
    h->tempfac = dalloc(h->nm * N);
    h->xifac = dalloc(h->nm * N);
    band_alloc_r(&h->P_RHSE[0], 4, 4, N);
    band_alloc_r(&h->P_RHSE[1], 8, 8, N);
    band_alloc_r(&h->P_RHSE[2], 7, 7, N);
    h->P_RHSI = calloc(h->nm, sizeof(bandr));
    h->P_RHSIL = calloc(h->nm, sizeof(bandc));
    h->P_LHS = calloc(h->nm, sizeof(bordc));
    h->psifac = dalloc(h->nm * N);
    for (int nm = 0; nm < h->nm; nm++) {
        int m = h->idx2m[nm];
        band_alloc_r(&h->T_RHSI[nm], 2, 2, N);
        bord_alloc_r(&h->T_LHS[nm], 4, 4, 2, N);
        bord_alloc_r(&h->XI_LHS[nm], 4, 4, 2, N);
        if (m == 0) {
            band_alloc_r(&h->P_RHSI[nm], 4, 4, N);
            band_alloc_c(&h->P_RHSIL[nm], 2, 2, N);
            bord_alloc_c(&h->P_LHS[nm], 4, 4, 2, N);
        } else {
            band_alloc_r(&h->P_RHSI[nm], 8, 8, N);
            band_alloc_c(&h->P_RHSIL[nm], h->c.l_coriolis_imp ? 8 : 6, h->c.l_coriolis_imp ? 8 : 6, N);
            bord_alloc_c(&h->P_LHS[nm], 8, 8, 4, N);
        }
    }
    build_matrices(h, 0);
    h->lMatNext = 1;
    *out = h;
    return 0;
}

/* RHS matrices (get_rhs_exp_mat / get_rhs_imp_mat / get_rhs_imp_lin_mat) and,
 * when lhs_only, just the LHS matrices with the current wimp_lin(1). */
static int build_matrices(or_handle *h, int lhs_only)
{
    if (!h->tabs) h->tabs = build_optabs(h);
    optabs *T = h->tabs;
    int N = h->N;
    if (!lhs_only) {
        fill_rhs_band_r(h, &h->T_RHSE[0], 2, T->I2R1.e, T->I2R1.K);
        fill_rhs_band_r(h, &h->T_RHSE[1], 2, T->I2R2.e, T->I2R2.K);
        fill_rhs_band_r(h, &h->P_RHSE[0], 2, T->I2R2.e, T->I2R2.K);
        fill_rhs_band_r(h, &h->P_RHSE[1], 4, T->I4R4.e, T->I4R4.K);
        fill_rhs_band_r(h, &h->P_RHSE[2], 4, T->I4R3.e, T->I4R3.K);
        for (int nm = 0; nm < h->nm; nm++) {
            int m = h->idx2m[nm];
            ld M2 = (ld)m * m;
            /* T: I2R2LAPL(m) with K=2 */
            {
                int K = 2;
                ld *ext = calloc((size_t)N * (2 * K + 1), sizeof(ld));
                for (size_t i = 0; i < (size_t)N * (2 * K + 1); i++) ext[i] = T->L2_0.e[i] + M2 * T->L2_2.e[i];
                fill_rhs_band_r(h, &h->T_RHSI[nm], 2, ext, K);
                free(ext);
            }
            double nu = h->ViscFac * h->hdifV[nm];
            if (m == 0) {
                fill_rhs_band_r(h, &h->P_RHSI[nm], 2, T->I2R2.e, T->I2R2.K);
                /* nu*(R2 - 3 I1R1) */
                int K = 2;
                ld *ext = calloc((size_t)N * (2 * K + 1), sizeof(ld));
                for (size_t i = 0; i < (size_t)N * (2 * K + 1); i++)
                    ext[i] = (ld)nu * (T->R2.e[i] - 3.0L * T->I1R1.e[i]);
                bandr tmp;
                band_alloc_r(&tmp, 2, 2, N);
                fill_rhs_band_r(h, &tmp, 2, ext, K);
                for (size_t i = 0; i < (size_t)5 * N; i++) h->P_RHSIL[nm].d[i] = tmp.d[i];
                free(tmp.d);
                free(ext);
            } else {
                int K = 8;
                ld *ext = calloc((size_t)N * (2 * K + 1), sizeof(ld));
                for (size_t i = 0; i < (size_t)N * (2 * K + 1); i++) ext[i] = T->LR_0.e[i] + M2 * T->LR_2.e[i];
                fill_rhs_band_r(h, &h->P_RHSI[nm], 4, ext, K);
                /* nu*LAPLROT2 + i*(-CorFac*m)*I4R4 */
                int Kc = h->P_RHSIL[nm].ku;
                ld *er = calloc((size_t)N * (2 * Kc + 1), sizeof(ld));
                ld *ei = calloc((size_t)N * (2 * Kc + 1), sizeof(ld));
                for (int n = 0; n < N; n++)
                    for (int d = -Kc; d <= Kc; d++) {
                        size_t s8 = (size_t)n * 17 + 8 + d, sc = (size_t)n * (2 * Kc + 1) + Kc + d;
                        er[sc] = (ld)nu * (T->LR2_0.e[s8] + M2 * T->LR2_2.e[s8] + M2 * M2 * T->LR2_4.e[s8]);
                        ei[sc] = h->c.l_coriolis_imp ? -(ld)h->CorFac * (ld)m * T->I4R4.e[s8] : 0;
                    }
                bandr tr, ti;
                band_alloc_r(&tr, Kc, Kc, N);
                band_alloc_r(&ti, Kc, Kc, N);
                fill_rhs_band_r(h, &tr, 4, er, Kc);
                fill_rhs_band_r(h, &ti, 4, ei, Kc);
                for (size_t i = 0; i < (size_t)(2 * Kc + 1) * N; i++) h->P_RHSIL[nm].d[i] = tr.d[i] + I * ti.d[i];
                free(tr.d); free(ti.d); free(er); free(ei); free(ext);
            }
        }
    } else {
        double wimp = h->ts.wimp_lin[0];
        for (int nm = 0; nm < h->nm; nm++) {
            int m = h->idx2m[nm], info;
            if (h->l_heat) {
                info = build_lhs_scalar(h, T, &h->T_LHS[nm], h->tempfac + (size_t)nm * N, m, wimp,
                                        h->TdiffFac * h->hdifT[nm], h->c.ktopt, h->c.kbott);
                if (info) return info;
            }
            if (h->l_chem) {
                info = build_lhs_scalar(h, T, &h->XI_LHS[nm], h->xifac + (size_t)nm * N, m, wimp,
                                        h->XidiffFac * h->hdifXi[nm], h->c.ktopxi, h->c.kbotxi);
                if (info) return info;
            }
            info = build_lhs_psi(h, T, &h->P_LHS[nm], h->psifac + (size_t)nm * N, m, wimp);
            if (info) return info;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Physics                                                                   */
/* ------------------------------------------------------------------------ */
static inline double cc22real(cplx a, cplx b, int m)
{
    if (m == 0) return creal(a) * creal(b);
    return 2.0 * (creal(a) * creal(b) + cimag(a) * cimag(b));
}
static inline double cc2real(cplx a, int m)
{
    if (m == 0) return creal(a) * creal(a);
    return 2.0 * (creal(a) * creal(a) + cimag(a) * cimag(a));
}

/* rloop::radial_loop (radial_loop.f90:231-394) writing the stage slots */
static void radial_loop(or_handle *h, int is /* 0-based stage */)
{
    int N = h->N, nm = h->nm, np = h->nphi;
    size_t F = (size_t)nm * N;
    cplx *dpsidt = h->dpsidt.expl + is * F, *dTdt = h->dTdt.expl + is * F, *dxidt = h->dxidt.expl + is * F;
    /* the radii are independent: one OpenMP thread per radius, each with its
     * own grid and spectrum buffers */
#pragma omp parallel for schedule(static)
    for (int n_r = 0; n_r < N; n_r++) {
        double *gb = tl_dbuf(11 * (size_t)np);
        double *usg = gb, *upg = gb + np, *omg = gb + 2 * np, *tg = gb + 3 * np, *xg = gb + 4 * np;
        double *usT = gb + 5 * np, *upT = gb + 6 * np, *usX = gb + 7 * np, *upX = gb + 8 * np, *usO = gb + 9 * np,
               *upO = gb + 10 * np;
        cplx *sp = tl_cbuf(5, 2 * (size_t)nm), *o = sp + nm;
        double usom = 0;
        for (int i = 0; i < nm; i++) {
            int m = h->idx2m[i];
            cplx usf = (m == 0) ? 0.5 * h->ek * h->CorFac * h->ekpump[n_r] * h->up[(size_t)i * N + n_r]
                                : h->us[(size_t)i * N + n_r];
            usom += cc22real(usf, h->om[(size_t)i * N + n_r], m);
        }
#define GATHER(f) for (int i = 0; i < nm; i++) sp[i] = (f)[(size_t)i * N + n_r]
        GATHER(h->us); phi_ifft(h, sp, usg, h->mR[n_r]);
        GATHER(h->up); phi_ifft(h, sp, upg, h->mR[n_r]);
        GATHER(h->om); phi_ifft(h, sp, omg, h->mR[n_r]);
        if (h->l_heat) { GATHER(h->temp); phi_ifft(h, sp, tg, h->mR[n_r]); }
        if (h->l_chem) { GATHER(h->xi); phi_ifft(h, sp, xg, h->mR[n_r]); }
#undef GATHER
        if (h->ts.istage == 1) {
            /* courant_mod::courant (courant.f90:41-76) */
            double vr = 0, vh = 0, cf = h->ts.courfac;
            for (int j = 0; j < np; j++) { vr = fmax(vr, usg[j] * usg[j]); vh = fmax(vh, upg[j] * upg[j]); }
            vr = cf * cf * vr;
            vh = cf * cf * vh;
            h->dtr[n_r] = 1.0e10;
            h->dth[n_r] = 1.0e10;
            if (vr != 0.0) h->dtr[n_r] = fmin(h->dtr[n_r], sqrt(h->delxr2[n_r] / vr));
            if (vh != 0.0) h->dth[n_r] = fmin(h->dth[n_r], sqrt(h->delxh2[n_r] / vh));
        }
        double r = h->r[n_r];
        for (int j = 0; j < np; j++) {
            if (h->l_heat) { usT[j] = r * usg[j] * tg[j]; upT[j] = upg[j] * tg[j]; }
            if (h->l_chem) { usX[j] = r * usg[j] * xg[j]; upX[j] = upg[j] * xg[j]; }
            usO[j] = r * usg[j] * omg[j];
            upO[j] = upg[j] * omg[j];
        }
#define SCATTER(f) for (int i = 0; i < nm; i++) (f)[(size_t)i * N + n_r] = o[i]
        phi_fft(h, upO, o, h->mR[n_r]); SCATTER(dpsidt);
        phi_fft(h, usO, o, h->mR[n_r]); SCATTER(h->dVsOm);
        if (h->l_heat) {
            phi_fft(h, upT, o, h->mR[n_r]); SCATTER(dTdt);
            phi_fft(h, usT, o, h->mR[n_r]); SCATTER(h->dVsT);
        }
        if (h->l_chem) {
            phi_fft(h, upX, o, h->mR[n_r]); SCATTER(dxidt);
            phi_fft(h, usX, o, h->mR[n_r]); SCATTER(h->dVsXi);
        }
#undef SCATTER
        double or1 = h->or1[n_r];
        for (int i = 0; i < nm; i++) {
            int m = h->idx2m[i];
            size_t q = (size_t)i * N + n_r;
            cplx cim = -or1 * I * (double)m;
            if (h->l_heat) dTdt[q] = cim * dTdt[q] - (1.0 - h->tadvz) * h->beta[n_r] * or1 * h->dVsT[q];
            if (h->l_chem) dxidt[q] = cim * dxidt[q];
            if (i == 0) dpsidt[q] = -usom;
            else dpsidt[q] = cim * dpsidt[q];
        }
    }
}

static void dealias(or_handle *h, cplx *F)
{
    for (int i = 0; i < h->nm; i++)
        for (int k = h->ncheb; k < h->N; k++) F[(size_t)i * h->N + k] = 0;
}

/* finish_exp_temp_int / finish_exp_xi_int (update_temp_integ.f90:396-469,
 * update_xi_integ.f90:280-354) */
static void finish_exp_scalar(or_handle *h, cplx *dVs, cplx *exp_last, const double *dcond,
                              const double *cond, int is_temp)
{
    int N = h->N;
    costf_field(h, dVs);
    dealias(h, dVs);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) get_dr_vec(h, dVs + (size_t)i * N, h->work + (size_t)i * N, 0);
    for (int n_r = 0; n_r < N; n_r++) {
        double h2 = h->r_cmb * h->r_cmb - h->r[n_r] * h->r[n_r];
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            size_t q = (size_t)i * N + n_r;
            if (is_temp)
                exp_last[q] = h->r[n_r] * exp_last[q] -
                              I * (double)m * (h2 * dcond[n_r] + h->tadvz * h->r[n_r] * cond[n_r]) * h->psi[q];
            else
                exp_last[q] = h->r[n_r] * exp_last[q] - I * (double)m * h2 * dcond[n_r] * h->psi[q];
        }
    }
    for (size_t q = 0; q < (size_t)h->nm * N; q++) exp_last[q] -= h->work[q];
    costf_field(h, exp_last);
    dealias(h, exp_last);
    for (int i = 0; i < h->nm; i++) {
        cplx *v = exp_last + (size_t)i * N;
        band_mv_r(&h->T_RHSE[0], v);
        v[0] = 0;
        v[1] = 0;
    }
}

/* finish_exp_psi_int_smat (update_psi_integ_smat.f90:792-958) */
static void finish_exp_psi(or_handle *h, cplx *dpsi)
{
    int N = h->N;
    costf_field(h, h->dVsOm);
    dealias(h, h->dVsOm);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) get_dr_vec(h, h->dVsOm + (size_t)i * N, h->work + (size_t)i * N, 0);
    for (int n_r = 0; n_r < N; n_r++)
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            size_t q = (size_t)i * N + n_r;
            if (m != 0) {
                dpsi[q] = h->r[n_r] * dpsi[q];
                if (!h->c.l_coriolis_imp) dpsi[q] = dpsi[q] - h->CorFac * I * (double)m * h->r[n_r] * h->psi[q];
                if (!h->c.l_buo_imp) {
                    if (h->l_heat) dpsi[q] = dpsi[q] - h->BuoFac * h->rgrav[n_r] * I * (double)m * h->temp[q];
                    if (h->l_chem) dpsi[q] = dpsi[q] - h->ChemFac * h->rgrav[n_r] * I * (double)m * h->xi[q];
                }
            }
        }
    if (h->c.l_ek_pump) {
        const double sgn = or_plant == 1 ? -1.0 : 1.0, rfac = or_plant == 2 ? 0.0 : 1.0;
        const double usfac = or_plant == 3 ? 0.0 : 1.0;
        for (int n_r = 0; n_r < N; n_r++) {
            double ekp_fac = sgn * h->CorFac * h->ekpump[n_r];
            for (int i = 0; i < h->nm; i++) {
                int m = h->idx2m[i];
                size_t q = (size_t)i * N + n_r;
                if (m == 0) dpsi[q] = dpsi[q] - ekp_fac * h->up[q];
                else
                    dpsi[q] = dpsi[q] - ekp_fac * (rfac * h->r[n_r] + (1.0 - rfac)) *
                                            (h->om[q] + h->ekp_up[n_r] * h->up[q] +
                                             usfac * (I * (double)m * h->ekp_dusdp[n_r] + h->ekp_us[n_r]) * h->us[q]);
            }
        }
    }
    for (int i = 1; i < h->nm; i++)
        for (int n_r = 0; n_r < N; n_r++) dpsi[(size_t)i * N + n_r] -= h->work[(size_t)i * N + n_r];
    costf_field(h, dpsi);
    dealias(h, dpsi);
    for (int i = 0; i < h->nm; i++) {
        cplx *v = dpsi + (size_t)i * N;
        if (h->idx2m[i] == 0) band_mv_r(&h->P_RHSE[0], v);
        else { band_mv_r(&h->P_RHSE[2], v); v[2] = 0; v[3] = 0; }
        v[0] = 0;
        v[1] = 0;
    }
}

/* get_temp_rhs_imp_int / get_xi_rhs_imp_int (update_temp_integ.f90:471-536) */
static void rhs_imp_scalar(or_handle *h, const cplx *hat, tarray *a, int istage, int l_calc_lin,
                           double diff, const double *hdif)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    cplx *v = malloc(sizeof(cplx) * N);
    if (istage == 1)
        for (int i = 0; i < h->nm; i++) {
            memcpy(v, hat + (size_t)i * N, sizeof(cplx) * N);
            band_mv_r(&h->T_RHSE[1], v);
            v[0] = 0; v[1] = 0;
            memcpy(a->old + (size_t)i * N, v, sizeof(cplx) * N);
        }
    if (l_calc_lin)
        for (int i = 0; i < h->nm; i++) {
            memcpy(v, hat + (size_t)i * N, sizeof(cplx) * N);
            band_mv_r(&h->T_RHSI[i], v);
            v[0] = 0; v[1] = 0;
            for (int k = 0; k < N; k++) a->impl[(istage - 1) * F + (size_t)i * N + k] = diff * hdif[i] * v[k];
        }
    free(v);
}

/* get_psi_rhs_imp_int_smat (update_psi_integ_smat.f90:960-1115) */
static void rhs_imp_psi(or_handle *h, const cplx *psi_hat, const cplx *up, const cplx *temp,
                        const cplx *xi, tarray *a, int istage, int l_calc_lin)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    double *work_r = dalloc(N);
    for (int n = 0; n < N; n++) work_r[n] = creal(up[n]);
    costf_real(h, work_r);
    cplx *v = malloc(sizeof(cplx) * N);
    if (istage == 1)
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            for (int k = 0; k < N; k++) v[k] = m == 0 ? work_r[k] : psi_hat[(size_t)i * N + k];
            band_mv_r(&h->P_RHSI[i], v);
            v[0] = 0; v[1] = 0;
            if (m > 0) { v[2] = 0; v[3] = 0; }
            memcpy(a->old + (size_t)i * N, v, sizeof(cplx) * N);
        }
    if (l_calc_lin) {
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            for (int k = 0; k < N; k++) v[k] = m == 0 ? work_r[k] : psi_hat[(size_t)i * N + k];
            band_mv_c(&h->P_RHSIL[i], v);
            v[0] = 0; v[1] = 0;
            if (m > 0) { v[2] = 0; v[3] = 0; }
            memcpy(h->work + (size_t)i * N, v, sizeof(cplx) * N);
        }
        cplx *impl = a->impl + (size_t)(istage - 1) * F;
        memcpy(impl, h->work, sizeof(cplx) * F);
        if (h->c.l_buo_imp) {
            for (int n_r = 0; n_r < N; n_r++)
                for (int i = 0; i < h->nm; i++) {
                    int m = h->idx2m[i];
                    size_t q = (size_t)i * N + n_r;
                    if (m != 0) {
                        if (h->l_heat) {
                            h->work[q] = -h->BuoFac * h->rgrav[n_r] * h->or1[n_r] * I * (double)m * temp[q];
                            if (h->l_chem)
                                h->work[q] = h->work[q] - h->ChemFac * h->rgrav[n_r] * h->or1[n_r] * I * (double)m * xi[q];
                        } else if (h->l_chem) {
                            h->work[q] = -h->ChemFac * h->rgrav[n_r] * h->or1[n_r] * I * (double)m * xi[q];
                        }
                    }
                }
            costf_field(h, h->work);
            for (int i = 0; i < h->nm; i++) {
                if (h->idx2m[i] <= 0) continue;
                memcpy(v, h->work + (size_t)i * N, sizeof(cplx) * N);
                band_mv_r(&h->P_RHSE[1], v);
                v[0] = v[1] = v[2] = v[3] = 0;
                for (int k = 0; k < N; k++) impl[(size_t)i * N + k] += v[k];
            }
        }
    }
    free(v);
    free(work_r);
}

/* update_temp_int / update_xi_int (update_temp_integ.f90:179-298) */
static int update_scalar(or_handle *h, int lMat, int is_temp)
{
    int N = h->N;
    tscheme *t = &h->ts;
    bordr *LHS = is_temp ? h->T_LHS : h->XI_LHS;
    double *fac = is_temp ? h->tempfac : h->xifac;
    tarray *a = is_temp ? &h->dTdt : &h->dxidt;
    cplx *hat = is_temp ? h->temp_hat : h->xi_hat, *fld = is_temp ? h->temp : h->xi;
    cplx *top = is_temp ? h->topt : h->topxi, *bot = is_temp ? h->bott : h->botxi;
    (void)lMat;
    ts_set_imex_rhs(h, t, a, h->work);
    double *rr = dalloc(N), *ri = dalloc(N);
    for (int i = 0; i < h->nm; i++) {
        cplx *w = h->work + (size_t)i * N;
        w[0] = top[i];
        w[1] = bot[i];
        for (int k = 0; k < N; k++) w[k] = w[k] * fac[(size_t)i * N + k];
        for (int k = 0; k < N; k++) { rr[k] = creal(w[k]); ri[k] = cimag(w[k]); }
        bord_solve_r(&LHS[i], rr);
        bord_solve_r(&LHS[i], ri);
        for (int k = 0; k < N; k++) hat[(size_t)i * N + k] = rr[k] + I * ri[k];
    }
    free(rr);
    free(ri);
    memcpy(fld, hat, sizeof(cplx) * h->nm * N);
    costf_field(h, fld);
    ts_rotate(h, t, a);
    double diff = is_temp ? h->TdiffFac : h->XidiffFac;
    double *hdif = is_temp ? h->hdifT : h->hdifXi;
    if (t->istage == t->nstages) rhs_imp_scalar(h, hat, a, 1, t->l_imp_calc_rhs[0], diff, hdif);
    else rhs_imp_scalar(h, hat, a, t->istage + 1, t->l_imp_calc_rhs[t->istage], diff, hdif);
    return 0;
}

/* update_psi_int_smat (update_psi_integ_smat.f90:247-573) */
static int update_psi(or_handle *h)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    tscheme *t = &h->ts;
    if (h->c.l_buo_imp) {
        for (int n_r = 0; n_r < N; n_r++)
            for (int i = 0; i < h->nm; i++) {
                int m = h->idx2m[i];
                size_t q = (size_t)i * N + n_r;
                if (m != 0) {
                    if (h->l_heat) {
                        h->buo[q] = -h->BuoFac * h->rgrav[n_r] * h->or1[n_r] * I * (double)m * h->temp[q];
                        if (h->l_chem)
                            h->buo[q] = h->buo[q] - h->ChemFac * h->rgrav[n_r] * h->or1[n_r] * I * (double)m * h->xi[q];
                    } else if (h->l_chem) {
                        h->buo[q] = -h->ChemFac * h->rgrav[n_r] * h->or1[n_r] * I * (double)m * h->xi[q];
                    }
                } else {
                    h->buo[q] = 0;
                }
            }
        costf_field(h, h->buo);
        for (int i = 0; i < h->nm; i++) {
            if (h->idx2m[i] <= 0) continue;
            cplx *v = h->buo + (size_t)i * N;
            band_mv_r(&h->P_RHSE[1], v);
            v[0] = v[1] = v[2] = v[3] = 0;
        }
    }
    ts_set_imex_rhs(h, t, &h->dpsidt, h->work);
    double *uphi0 = dalloc(N), *om0 = dalloc(N);
    cplx *rhs = malloc(sizeof(cplx) * N);
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        for (int k = 0; k < N; k++) {
            rhs[k] = h->work[(size_t)i * N + k];
            if (h->c.l_buo_imp && m != 0) rhs[k] = rhs[k] + t->wimp_lin[0] * h->buo[(size_t)i * N + k];
        }
        for (int k = 0; k < N; k++) rhs[k] = rhs[k] * h->psifac[(size_t)i * N + k];
        bord_solve_c(&h->P_LHS[i], rhs);
        if (m == 0) for (int k = 0; k < N; k++) uphi0[k] = creal(rhs[k]);
        else memcpy(h->psi_hat + (size_t)i * N, rhs, sizeof(cplx) * N);
    }
    for (int i = 0; i < h->nm; i++)
        if (h->idx2m[i] != 0) memcpy(h->psi + (size_t)i * N, h->psi_hat + (size_t)i * N, sizeof(cplx) * N);
    get_dr_real(h, uphi0, om0, 0);
    costf_real(h, uphi0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++)
        get_ddr_vec(h, h->psi + (size_t)i * N, h->work + (size_t)i * N, h->om + (size_t)i * N, 0);
    costf_field(h, h->psi);
    double rc2 = h->r_cmb * h->r_cmb;
    for (int n_r = 0; n_r < N; n_r++) {
        double h2 = rc2 - h->r[n_r] * h->r[n_r];
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            size_t q = (size_t)i * N + n_r;
            if (m == 0) {
                h->us[q] = 0.0;
                h->up[q] = uphi0[n_r];
                h->om[q] = om0[n_r] + h->or1[n_r] * uphi0[n_r];
            } else {
                h->us[q] = I * (double)m * h->or1[n_r] * h2 * h->psi[q];
                h->up[q] = -h2 * h->work[q] + 3.0 * h->r[n_r] * h->psi[q];
                h->om[q] = -h2 * h->om[q] - (rc2 * h->or1[n_r] - 6.0 * h->r[n_r]) * h->work[q] +
                           (6.0 + (double)m * (double)m * h->or2[n_r] * h2) * h->psi[q];
            }
        }
    }
    ts_rotate(h, t, &h->dpsidt);
    if (t->istage == t->nstages)
        rhs_imp_psi(h, h->psi_hat, h->up, h->temp, h->xi, &h->dpsidt, 1, t->l_imp_calc_rhs[0]);
    else
        rhs_imp_psi(h, h->psi_hat, h->up, h->temp, h->xi, &h->dpsidt, t->istage + 1,
                    t->l_imp_calc_rhs[t->istage]);
    free(uphi0);
    free(om0);
    free(rhs);
    (void)F;
    return 0;
}

/* dt_courant (courant.f90:78-167) */
static void dt_courant(or_handle *h, int *l_new_dt, double *dt_new)
{
    double dtMax = h->c.dtMax, dt = h->ts.dt[0];
    double dt_r = 1000.0 * dtMax, dt_h = dt_r;
    for (int n = 0; n < h->N; n++) { dt_r = fmin(h->dtr[n], dt_r); dt_h = fmin(h->dth[n], dt_h); }
    double dt_rh = fmin(dt_r, dt_h);
    double dt_2 = fmin(0.5 * (1.0 / h->c.dt_fac + 1.0) * dt_rh, dtMax);
    if (dt > dtMax) { *l_new_dt = 1; *dt_new = dtMax; }
    else if (dt > dt_rh) { *l_new_dt = 1; *dt_new = dt_2; }
    else if (h->c.dt_fac * dt < dt_rh && dt < dtMax) { *l_new_dt = 1; *dt_new = dt_2; }
    else { *l_new_dt = 0; *dt_new = dt; }
}

/* one iteration of time_loop's stage loop (step_time.f90:215-348) */
/* ------------------------------------------------------------------------ */
/* Collocation method (cheb_method = 'colloc', matrix_solve = 'direct')      */
/* ------------------------------------------------------------------------ */

/* get_der_mat, collocation branch (chebyshev.f90:296-321): T_n at the grid
 * points by the three-term recursion, with the r-map drx/ddrx folded into
 * the derivatives; stored (point k, degree n) as after the transposition */
static void coll_der_mat(or_handle *h)
{
    int N = h->N;
    h->cM = dalloc(N * N); h->cdM = dalloc(N * N); h->cd2M = dalloc(N * N);
    for (int k = 0; k < N; k++) {
        double *r0 = h->cM + (size_t)k * N, *r1 = h->cdM + (size_t)k * N, *r2 = h->cd2M + (size_t)k * N;
        double xk = h->x[k], a = h->drx[k], b = h->ddrx[k];
        r0[0] = 1.0; r0[1] = xk;
        r1[0] = 0.0; r1[1] = a;
        r2[0] = 0.0; r2[1] = b;
        for (int n = 2; n < N; n++) {
            r0[n] = 2.0 * xk * r0[n - 1] - r0[n - 2];
            r1[n] = 2.0 * a * r0[n - 1] + 2.0 * xk * r1[n - 1] - r1[n - 2];
            r2[n] = 2.0 * b * r0[n - 1] + 4.0 * a * r1[n - 1] + 2.0 * xk * r2[n - 1] - r2[n - 2];
        }
    }
}
/* the boundary first-derivative rows of the collocation method
 * (chebyshev.f90:338-360): the flipped-row form at r_cmb, its mirror at
 * r_icb, both times the mapping */
static void coll_dr_rows(or_handle *h)
{
    int N = h->N;
    h->dr_top = dalloc(N);
    h->dr_bot = dalloc(N);
    h->dr_top[0] = (2.0 * (N - 1) * (N - 1) + 1.0) / 6.0;
    for (int k = 2; k <= N; k++) {
        double s = sin(0.5 * (k - 1) * M_PI / (N - 1));
        double diff = 2.0 * s * s;
        h->dr_top[k - 1] = 2.0 * ((k % 2 == 0) ? -1.0 : 1.0) / diff;
    }
    h->dr_top[N - 1] *= 0.5;
    for (int k = 0; k < N; k++) h->dr_bot[k] = -h->dr_top[N - 1 - k];
    for (int k = 0; k < N; k++) h->dr_top[k] *= h->drx[0];
    double db = h->drx[N - 1];
    for (int k = 0; k < N; k++) h->dr_bot[k] *= db;
}

/* robin_bc (chebyshev.f90:385-445): the two boundary values of f from
 * atop df/dr + btop f = rhs_top, abot df/dr + bbot f = rhs_bot */
static void robin_bc(or_handle *h, double atop, double btop, cplx rhs_top, double abot, double bbot, cplx rhs_bot,
                     cplx *f)
{
    int N = h->N;
    const double *dt = h->dr_top, *db = h->dr_bot;
    double str = 0, sti = 0, sbr = 0, sbi = 0;
    for (int k = 1; k < N - 1; k++) {
        str += dt[k] * creal(f[k]);
        sti += dt[k] * cimag(f[k]);
        sbr += db[k] * creal(f[k]);
        sbi += db[k] * cimag(f[k]);
    }
    double den = abot * atop * db[0] * dt[N - 1] - (atop * dt[0] + btop) * (abot * db[N - 1] + bbot);
    double vbr = (abot * db[0] * (creal(rhs_top) - atop * str) - (atop * dt[0] + btop) * (creal(rhs_bot) - abot * sbr)) / den;
    double vbi = (abot * db[0] * (cimag(rhs_top) - atop * sti) - (atop * dt[0] + btop) * (cimag(rhs_bot) - abot * sbi)) / den;
    double vtr = (creal(rhs_top) - atop * (str + dt[N - 1] * vbr)) / (atop * dt[0] + btop);
    double vti = (cimag(rhs_top) - atop * (sti + dt[N - 1] * vbi)) / (atop * dt[0] + btop);
    f[0] = vtr + I * vti;
    f[N - 1] = vbr + I * vbi;
}

#define CM(k, n) h->cM[(size_t)(k) * N + (n)]
#define CDM(k, n) h->cdM[(size_t)(k) * N + (n)]
#define CD2M(k, n) h->cd2M[(size_t)(k) * N + (n)]

/* get_tempMat / get_xiMat (update_temp_coll.f90:310-393, update_xi_coll.f90):
 * boundary rows, diffusion operator on the interior rows, boundary_fac on
 * the first and last columns, row preconditioning, dgetrf */
static int coll_scalar_mat(or_handle *h, int m, double diff, int ktop, int kbot, double *A, double *fac,
                           int *piv)
{
    int N = h->N, nc = h->ncheb;
    double wimp = h->ts.wimp_lin[0], dm2 = (double)m * (double)m;
    for (int c = 0; c < N; c++) {
        A[0 + c * N] = c < nc ? h->rnorm * (ktop == 1 ? CM(0, c) : CDM(0, c)) : 0.0;
        A[(N - 1) + c * N] = c < nc ? h->rnorm * (kbot == 1 ? CM(N - 1, c) : CDM(N - 1, c)) : 0.0;
        for (int r = 1; r < N - 1; r++)
            A[r + c * N] = h->rnorm * (CM(r, c) - wimp * diff *
                                                      (CD2M(r, c) + h->or1[r] * CDM(r, c) - dm2 * h->or2[r] * CM(r, c)));
    }
    for (int r = 0; r < N; r++) {
        A[r] *= h->boundary_fac;
        A[r + (N - 1) * N] *= h->boundary_fac;
    }
    for (int r = 0; r < N; r++) {
        double mx = 0.0;
        for (int c = 0; c < N; c++) mx = fmax(mx, fabs(A[r + c * N]));
        fac[r] = 1.0 / mx;
        for (int c = 0; c < N; c++) A[r + c * N] *= fac[r];
    }
    int info;
    getrf_r(N, A, piv, &info);
    return info;
}

/* get_psiMat (update_psi_coll_smat.f90:674-837): the coupled [omega; psi]
 * system of one m /= 0, row then column max-scaling, zgetrf */
static int coll_psi_mat(or_handle *h, int m, cplx *P, double *fac, int *piv)
{
    int N = h->N, nc = h->ncheb, N2 = 2 * N;
    double wimp = h->ts.wimp_lin[0], dm2 = (double)m * (double)m, rn = h->rnorm;
    cplx cim = I * (double)m;
    memset(P, 0, sizeof(cplx) * N2 * N2);
    for (int c = 0; c < nc; c++) {
        int cp = c + N;
        P[0 + (size_t)cp * N2] = rn * CM(0, c);
        P[(N - 1) + (size_t)cp * N2] = rn * CM(N - 1, c);
        P[N + (size_t)cp * N2] = h->c.ktopv == 1 ? rn * (CD2M(0, c) - h->or1[0] * CDM(0, c)) : rn * CDM(0, c);
        P[(N2 - 1) + (size_t)cp * N2] = h->c.kbotv == 1 ? rn * (CD2M(N - 1, c) - h->or1[N - 1] * CDM(N - 1, c))
                                                         : rn * CDM(N - 1, c);
    }
    for (int c = 0; c < N; c++) {
        int cp = c + N;
        for (int r = 1; r < N - 1; r++) {
            int rp = r + N;
            double ekp = h->CorFac * h->ekpump[r];
            P[r + (size_t)c * N2] = rn * (CM(r, c) - wimp * (h->ViscFac * (CD2M(r, c) + h->or1[r] * CDM(r, c) -
                                                                          dm2 * h->or2[r] * CM(r, c)) -
                                                              ekp * CM(r, c)));
            cplx v = -rn * wimp * (ekp * h->ekp_up[r] * CDM(r, c) +
                                   ekp * (h->ekp_up[r] * h->beta[r] - (h->ekp_dusdp[r] * cim + h->ekp_us[r]) * cim *
                                                                          h->or1[r]) * CM(r, c));
            if (h->c.l_coriolis_imp) v -= rn * wimp * h->CorFac * h->beta[r] * h->or1[r] * cim * CM(r, c);
            P[r + (size_t)cp * N2] = v;
            P[rp + (size_t)c * N2] = rn * CM(r, c);
            P[rp + (size_t)cp * N2] = rn * (CD2M(r, c) + (h->or1[r] + h->beta[r]) * CDM(r, c) +
                                            (h->or1[r] * h->beta[r] + h->dbeta[r] - dm2 * h->or2[r]) * CM(r, c));
        }
    }
    int bc[4] = {0, N - 1, N, N2 - 1};
    for (int r = 0; r < N2; r++)
        for (int q = 0; q < 4; q++) P[r + (size_t)bc[q] * N2] *= h->boundary_fac;
    for (int r = 0; r < N2; r++) {
        double mx = 0.0;
        for (int c = 0; c < N2; c++) mx = fmax(mx, cabs(P[r + (size_t)c * N2]));
        fac[r] = 1.0 / mx;
        for (int c = 0; c < N2; c++) P[r + (size_t)c * N2] *= fac[r];
    }
    for (int c = 0; c < N2; c++) {
        double mx = 0.0;
        for (int r = 0; r < N2; r++) mx = fmax(mx, cabs(P[r + (size_t)c * N2]));
        fac[N2 + c] = 1.0 / mx;
        for (int r = 0; r < N2; r++) P[r + (size_t)c * N2] *= fac[N2 + c];
    }
    int info;
    getrf_c(N2, P, piv, &info);
    return info;
}

/* get_uphiMat (update_psi_coll_smat.f90:919-987): m = 0, no preconditioning */
static int coll_uphi_mat(or_handle *h, double *A, int *piv)
{
    int N = h->N, nc = h->ncheb;
    double wimp = h->ts.wimp_lin[0], rn = h->rnorm;
    for (int c = 0; c < N; c++) {
        A[0 + c * N] = c >= nc ? 0.0 : (h->c.ktopv == 1 ? rn * (CDM(0, c) - h->or1[0] * CM(0, c)) : rn * CM(0, c));
        A[(N - 1) + c * N] = c >= nc ? 0.0
                                     : (h->c.kbotv == 1 ? rn * (CDM(N - 1, c) - h->or1[N - 1] * CM(N - 1, c))
                                                        : rn * CM(N - 1, c));
        for (int r = 1; r < N - 1; r++)
            A[r + c * N] = rn * (CM(r, c) - wimp * (h->ViscFac * (CD2M(r, c) + h->or1[r] * CDM(r, c) -
                                                                 h->or2[r] * CM(r, c)) -
                                                     h->CorFac * h->ekpump[r] * CM(r, c)));
    }
    for (int r = 0; r < N; r++) {
        A[r] *= h->boundary_fac;
        A[r + (N - 1) * N] *= h->boundary_fac;
    }
    int info;
    getrf_r(N, A, piv, &info);
    return info;
}

/* row then column max-scaling of a column-major n x n matrix
 * (update_psi_coll_dmat.f90:880-896): fac[0..n) rows, fac[n..2n) columns */
static void scale_rows_cols(int n, double *A, double *fac)
{
    for (int r = 0; r < n; r++) {
        double mx = 0.0;
        for (int c = 0; c < n; c++) mx = fmax(mx, fabs(A[r + (size_t)c * n]));
        fac[r] = 1.0 / mx;
        for (int c = 0; c < n; c++) A[r + (size_t)c * n] *= fac[r];
    }
    for (int c = 0; c < n; c++) {
        double mx = 0.0;
        for (int r = 0; r < n; r++) mx = fmax(mx, fabs(A[r + (size_t)c * n]));
        fac[n + c] = 1.0 / mx;
        for (int r = 0; r < n; r++) A[r + (size_t)c * n] *= fac[n + c];
    }
}

/* get_omMat of the influence-matrix method (update_psi_coll_dmat.f90:
 * 829-911): omega = 0 rows at both walls, the implicit viscous operator
 * inside, boundary_fac columns, row and column scalings, dgetrf */
static int coll_om_mat_dmat(or_handle *h, int i, double *A, double *fac, int *piv)
{
    int N = h->N, nc = h->ncheb, m = h->idx2m[i];
    double wimp = h->ts.wimp_lin[0], rn = h->rnorm, dm2 = (double)m * (double)m;
    for (int c = 0; c < N; c++) {
        A[0 + (size_t)c * N] = c < nc ? rn * CM(0, c) : 0.0;
        A[(N - 1) + (size_t)c * N] = c < nc ? rn * CM(N - 1, c) : 0.0;
        for (int r = 1; r < N - 1; r++)
            A[r + (size_t)c * N] = rn * (CM(r, c) - wimp * h->ViscFac * h->hdifV[i] *
                                                        (CD2M(r, c) + h->or1[r] * CDM(r, c) - dm2 * h->or2[r] * CM(r, c)));
    }
    for (int r = 0; r < N; r++) {
        A[r] *= h->boundary_fac;
        A[r + (size_t)(N - 1) * N] *= h->boundary_fac;
    }
    scale_rows_cols(N, A, fac);
    int info;
    getrf_r(N, A, piv, &info);
    return info;
}

/* get_psiMat of the influence-matrix method (update_psi_coll_dmat.f90:
 * 913-983): psi = 0 rows at both walls, the stream-function operator
 * (with beta) inside */
static int coll_psi_mat_dmat(or_handle *h, int i, double *A, double *fac, int *piv)
{
    int N = h->N, nc = h->ncheb, m = h->idx2m[i];
    double rn = h->rnorm, dm2 = (double)m * (double)m;
    for (int c = 0; c < N; c++) {
        A[0 + (size_t)c * N] = c < nc ? rn * CM(0, c) : 0.0;
        A[(N - 1) + (size_t)c * N] = c < nc ? rn * CM(N - 1, c) : 0.0;
        for (int r = 1; r < N - 1; r++)
            A[r + (size_t)c * N] = rn * (CD2M(r, c) + (h->or1[r] + h->beta[r]) * CDM(r, c) +
                                         (h->or1[r] * h->beta[r] + h->dbeta[r] - dm2 * h->or2[r]) * CM(r, c));
    }
    for (int r = 0; r < N; r++) {
        A[r] *= h->boundary_fac;
        A[r + (size_t)(N - 1) * N] *= h->boundary_fac;
    }
    scale_rows_cols(N, A, fac);
    int info;
    getrf_r(N, A, piv, &info);
    return info;
}

/* get_bc_influence_matrix (update_psi_coll_dmat.f90:1302-1452): the wall
 * values of the velocity condition (du_phi/dr - u_phi/r = 0 stress-free,
 * u_phi = 0 no-slip, in dpsi/dr form) of Chebyshev coefficients c */
static void coll_bc_dmat(or_handle *h, const cplx *c, cplx *top, cplx *bot)
{
    int N = h->N;
    cplx t = 0.0, b = 0.0;
    for (int n = 0; n < N; n++) {
        double fac = (n == 0 || n == N - 1) ? h->boundary_fac : 1.0;
        t += fac * (h->c.ktopv == 1 ? CD2M(0, n) - h->or1[0] * CDM(0, n) : CDM(0, n)) * c[n];
        b += fac * (h->c.kbotv == 1 ? CD2M(N - 1, n) - h->or1[N - 1] * CDM(N - 1, n) : CDM(N - 1, n)) * c[n];
    }
    *top = h->rnorm * t;
    *bot = h->rnorm * b;
}

/* one mode's influence-matrix set-up (update_psi_coll_dmat.f90:182-205,
 * 296-312, 344-347, 1136-1254): omega_2 (omega = 1 at the inner wall),
 * omega_3 (= 1 at the outer wall), psi_2,3 from L_psi psi = -omega_2,3 and
 * the inverse of the matrix of their wall velocity conditions */
static int coll_dmat_mode(or_handle *h, int i)
{
    int N = h->N, nc = h->ncheb;
    double *Om = h->dOmat + (size_t)i * N * N, *of = h->dOfac + (size_t)i * 2 * N;
    double *Pm = h->dPmat + (size_t)i * N * N, *pf = h->dPfac + (size_t)i * 2 * N;
    int *op = h->dOpiv + (size_t)i * N, *pp = h->dPpiv + (size_t)i * N;
    int info = coll_om_mat_dmat(h, i, Om, of, op);
    if (!info) info = coll_psi_mat_dmat(h, i, Pm, pf, pp);
    if (info) return info;
    double *rhs = tl_dbuf(2 * (size_t)N), *w = rhs + N;
    cplx bt[2], bb[2];
    for (int set = 0; set < 2; set++) {
        double *om = (set == 0 ? h->dom2 : h->dom3) + (size_t)i * N;
        double *ps = (set == 0 ? h->dpsi2 : h->dpsi3) + (size_t)i * N;
        memset(rhs, 0, sizeof(double) * N);
        if (set == 0) rhs[N - 1] = of[N - 1];
        else rhs[0] = of[0];
        getrs_r(N, Om, op, rhs);
        for (int k = 0; k < N; k++) om[k] = k < nc ? rhs[k] * of[N + k] : 0.0;
        /* get_intermediate_psi: omega to the grid, -omega inside */
        memcpy(w, om, sizeof(double) * N);
        costf_real(h, w);
        rhs[0] = rhs[N - 1] = 0.0;
        for (int k = 1; k < N - 1; k++) rhs[k] = -w[k];
        for (int k = 0; k < N; k++) rhs[k] *= pf[k];
        getrs_r(N, Pm, pp, rhs);
        for (int k = 0; k < N; k++) ps[k] = k < nc ? rhs[k] * pf[N + k] : 0.0;
        cplx *tmp = tl_cbuf(5, N);
        for (int k = 0; k < N; k++) tmp[k] = ps[k];
        coll_bc_dmat(h, tmp, &bt[set], &bb[set]);
    }
    double t2 = creal(bt[0]), b2 = creal(bb[0]), t3 = creal(bt[1]), b3 = creal(bb[1]);
    double det = t2 * b3 - t3 * b2;
    double *inf = h->dinfl + (size_t)i * 4;
    inf[0] = b3 / det;  /* (1,1) */
    inf[1] = -t3 / det; /* (1,2) */
    inf[2] = -b2 / det; /* (2,1) */
    inf[3] = t2 / det;  /* (2,2) */
    return 0;
}

/* the per-m LU factorisations of one lMat rebuild */
static int coll_build(or_handle *h)
{
    int N = h->N, N2 = 2 * N;
    /* the modes factorise independently; the first failing mode (in m
     * order) reports its info, as the serial loop would */
    int *infos = calloc(h->nm, sizeof(int));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i], info = 0;
        if (h->l_heat)
            info = coll_scalar_mat(h, m, h->TdiffFac * h->hdifT[i], h->c.ktopt, h->c.kbott, h->cTmat + (size_t)i * N * N,
                                   h->cTfac + (size_t)i * N, h->cTpiv + (size_t)i * N);
        if (!info && h->l_chem)
            info = coll_scalar_mat(h, m, h->XidiffFac * h->hdifXi[i], h->c.ktopxi, h->c.kbotxi,
                                   h->cXmat + (size_t)i * N * N, h->cXfac + (size_t)i * N, h->cXpiv + (size_t)i * N);
        if (!info) {
            if (m == 0) info = coll_uphi_mat(h, h->cUmat, h->cUpiv);
            else if (!h->c.l_direct_solve) info = coll_dmat_mode(h, i);
            else info = coll_psi_mat(h, m, h->cPmat + (size_t)i * N2 * N2, h->cPfac + (size_t)i * 2 * N2,
                                     h->cPpiv + (size_t)i * N2);
        }
        infos[i] = info;
    }
    int info = 0;
    for (int i = 0; i < h->nm && !info; i++) info = infos[i];
    free(infos);
    return info;
}

static void coll_alloc(or_handle *h)
{
    int N = h->N, N2 = 2 * N;
    size_t nm = h->nm;
    h->dbeta = dalloc(N);
    for (int n = 1; n < N && !h->c.l_non_rot; n++) {
        double s2 = h->r_cmb * h->r_cmb - h->r[n] * h->r[n];
        h->dbeta[n] = -(h->r_cmb * h->r_cmb + h->r[n] * h->r[n]) / (s2 * s2); /* radial.f90:222 */
    }
    coll_der_mat(h);
    coll_dr_rows(h);
    h->cTmat = calloc(nm * N * N, sizeof(double)); h->cXmat = calloc(nm * N * N, sizeof(double));
    h->cTfac = dalloc(nm * N); h->cXfac = dalloc(nm * N);
    h->cTpiv = calloc(nm * N, sizeof(int)); h->cXpiv = calloc(nm * N, sizeof(int));
    h->cUmat = dalloc(N * N); h->cUpiv = calloc(N, sizeof(int));
    if (h->c.l_direct_solve) {
        h->cPmat = calloc(nm * N2 * N2, sizeof(cplx));
        h->cPfac = dalloc(nm * 2 * N2);
        h->cPpiv = calloc(nm * N2, sizeof(int));
    } else {
        h->dOmat = dalloc(nm * N * N); h->dPmat = dalloc(nm * N * N);
        h->dOfac = dalloc(nm * 2 * N); h->dPfac = dalloc(nm * 2 * N);
        h->dOpiv = calloc(nm * N, sizeof(int)); h->dPpiv = calloc(nm * N, sizeof(int));
        h->dom2 = dalloc(nm * N); h->dom3 = dalloc(nm * N);
        h->dpsi2 = dalloc(nm * N); h->dpsi3 = dalloc(nm * N);
        h->dinfl = dalloc(nm * 4);
    }
}

/* get_ddr_real_1d with get_ddcheb_real_1d (radial_derivatives.f90:255-311,
 * 593-667): grid values in, derivatives out; below the rerror threshold the
 * recursion carries the previous value instead of adding the term */
static void get_ddr_real(or_handle *h, const double *f_in, double *df, double *ddf)
{
    int N = h->N;
    double *f = malloc(sizeof(double) * N);
    memcpy(f, f_in, sizeof(double) * N);
    costf_real(h, f);
    double mx = 0;
    for (int j = 0; j < N; j++) mx = fmax(mx, fabs(f[j]));
    double thr = mx * h->thr;
    df[N - 1] = 0.0;
    ddf[N - 1] = 0.0;
    int nc = N - 1;
    df[nc - 1] = fabs(f[nc]) >= thr ? (double)nc * f[nc] : 0.0;
    ddf[nc - 1] = 0.0;
    for (int n = N - 2; n >= 1; n--) {
        double fac = (double)(2 * n);
        df[n - 1] = fabs(f[n]) >= thr ? df[n + 1] + fac * f[n] : df[n + 1];
        ddf[n - 1] = fabs(df[n]) >= thr ? ddf[n + 1] + fac * df[n] : ddf[n + 1];
    }
    costf_real(h, df);
    costf_real(h, ddf);
    for (int j = 0; j < N; j++) {
        ddf[j] = h->ddrx[j] * df[j] + h->drx[j] * h->drx[j] * ddf[j];
        df[j] = h->drx[j] * df[j];
    }
    free(f);
}

/* finish_exp_temp_coll / finish_exp_xi_coll (update_temp_coll.f90:152-176,
 * update_xi_coll.f90:148-166) and finish_exp_psi_coll_smat
 * (update_psi_coll_smat.f90:323-378) */
static void coll_finish_exp(or_handle *h, int is)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    if (h->l_heat) {
        cplx *e = h->dTdt.expl + is * F;
    #pragma omp parallel for schedule(static)
        for (int i = 0; i < h->nm; i++) get_dr_vec(h, h->dVsT + (size_t)i * N, h->work + (size_t)i * N, 1);
        for (int i = 0; i < h->nm; i++)
            for (int k = 0; k < N; k++) {
                size_t q = (size_t)i * N + k;
                e[q] = e[q] - h->or1[k] * h->work[q] - h->us[q] * (h->dtcond[k] - h->tadvz * h->beta[k] * h->tcond[k]);
            }
    }
    if (h->l_chem) {
        cplx *e = h->dxidt.expl + is * F;
    #pragma omp parallel for schedule(static)
        for (int i = 0; i < h->nm; i++) get_dr_vec(h, h->dVsXi + (size_t)i * N, h->work + (size_t)i * N, 1);
        for (int i = 0; i < h->nm; i++)
            for (int k = 0; k < N; k++) {
                size_t q = (size_t)i * N + k;
                e[q] = e[q] - h->or1[k] * h->work[q] - h->us[q] * h->dxicond[k];
            }
    }
    cplx *e = h->dpsidt.expl + is * F;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) get_dr_vec(h, h->dVsOm + (size_t)i * N, h->work + (size_t)i * N, 1);
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        if (m == 0) continue;
        for (int k = 0; k < N; k++) {
            size_t q = (size_t)i * N + k;
            e[q] = e[q] - h->or1[k] * h->work[q];
            if (!h->c.l_coriolis_imp) e[q] += h->CorFac * h->beta[k] * h->us[q];
            if (!h->c.l_buo_imp) {
                if (h->l_heat) e[q] -= h->BuoFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->temp[q];
                if (h->l_chem) e[q] -= h->ChemFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->xi[q];
            }
            /* finish_exp_psi_coll_dmat (update_psi_coll_dmat.f90:527-549):
               the influence-matrix method treats pumping explicitly */
            if (!h->c.l_direct_solve && h->c.l_ek_pump)
                e[q] -= (or_plant == 6 ? -1.0 : 1.0) * h->CorFac * h->ekpump[k] *
                        (h->om[q] + h->ekp_up[k] * h->up[q] + (h->ekp_dusdp[k] * I * (double)m + h->ekp_us[k]) * h->us[q]);
        }
    }
}

/* get_temp_rhs_imp_coll / get_xi_rhs_imp_coll (update_temp_coll.f90:178-236):
 * fld holds Chebyshev coefficients when in_cheb (left as grid values) */
static void coll_rhs_imp_scalar(or_handle *h, cplx *fld, cplx *dfld, tarray *a, int istage, int l_calc_lin,
                                int in_cheb, double diff, const double *hdif)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    if (l_calc_lin) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < h->nm; i++)
            get_ddr_vec(h, fld + (size_t)i * N, dfld + (size_t)i * N, h->work + (size_t)i * N, !in_cheb);
    }
    if (in_cheb) costf_field(h, fld);
    if (istage == 1) memcpy(a->old, fld, sizeof(cplx) * F);
    if (l_calc_lin)
        for (int i = 0; i < h->nm; i++) {
            double dm2 = (double)h->idx2m[i] * (double)h->idx2m[i];
            for (int k = 0; k < N; k++) {
                size_t q = (size_t)i * N + k;
                a->impl[(istage - 1) * F + q] =
                    diff * hdif[i] * (h->work[q] + h->or1[k] * dfld[q] - dm2 * h->or2[k] * fld[q]);
            }
        }
}

/* get_psi_rhs_imp_coll_smat (update_psi_coll_smat.f90:380-527) */
static void coll_rhs_imp_psi(or_handle *h, tarray *a, int istage, int l_calc_lin, int in_cheb)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    int nm0 = -1;
    for (int i = 0; i < h->nm; i++)
        if (h->idx2m[i] == 0) nm0 = i;
    double *om0 = dalloc(N);
    if (nm0 >= 0)
        for (int k = 0; k < N; k++) om0[k] = creal(h->om[(size_t)nm0 * N + k]);
    cplx *dom = calloc(F, sizeof(cplx));
    if (l_calc_lin) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < h->nm; i++)
            get_ddr_vec(h, h->om + (size_t)i * N, dom + (size_t)i * N, h->work + (size_t)i * N, !in_cheb);
    }
    if (in_cheb) costf_field(h, h->om);
    if (nm0 >= 0)
        for (int k = 0; k < N; k++) h->om[(size_t)nm0 * N + k] = om0[k];
    if (istage == 1)
        for (int i = 0; i < h->nm; i++)
            memcpy(a->old + (size_t)i * N, (h->idx2m[i] == 0 ? h->up : h->om) + (size_t)i * N, sizeof(cplx) * N);
    if (l_calc_lin) {
        double *u0 = dalloc(N), *du0 = dalloc(N), *d2u0 = dalloc(N);
        if (nm0 >= 0) {
            for (int k = 0; k < N; k++) u0[k] = creal(h->up[(size_t)nm0 * N + k]);
            get_ddr_real(h, u0, du0, d2u0);
        }
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            double dm2 = (double)m * (double)m;
            for (int k = 0; k < N; k++) {
                size_t q = (size_t)i * N + k;
                cplx v;
                if (m == 0) {
                    v = h->ViscFac * h->hdifV[i] * (d2u0[k] + h->or1[k] * du0[k] - h->or2[k] * u0[k]) -
                        h->CorFac * h->ekpump[k] * u0[k];
                } else {
                    v = h->ViscFac * h->hdifV[i] * (h->work[q] + h->or1[k] * dom[q] - dm2 * h->or2[k] * h->om[q]);
                    /* get_psi_rhs_imp_coll_dmat (update_psi_coll_dmat.f90:641-667)
                       has no pumping term */
                    if (h->c.l_direct_solve)
                        v -= h->CorFac * h->ekpump[k] *
                             (h->om[q] + h->ekp_up[k] * h->up[q] +
                              (h->ekp_dusdp[k] * I * (double)m + h->ekp_us[k]) * h->us[q]);
                    if (h->c.l_coriolis_imp) v += h->CorFac * h->beta[k] * h->us[q];
                    if (h->c.l_buo_imp) {
                        if (h->l_heat) v -= h->rgrav[k] * h->or1[k] * h->BuoFac * I * (double)m * h->temp[q];
                        if (h->l_chem) v -= h->rgrav[k] * h->or1[k] * h->ChemFac * I * (double)m * h->xi[q];
                    }
                }
                a->impl[(istage - 1) * F + q] = v;
            }
        }
        free(u0); free(du0); free(d2u0);
    }
    free(dom);
    free(om0);
}

/* update_temp_co / update_xi_co (update_temp_coll.f90:66-150) */
static void coll_update_scalar(or_handle *h, int is_temp)
{
    int N = h->N;
    tscheme *t = &h->ts;
    tarray *a = is_temp ? &h->dTdt : &h->dxidt;
    cplx *fld = is_temp ? h->temp : h->xi, *dfld = is_temp ? h->dtemp : h->dxi;
    cplx *top = is_temp ? h->topt : h->topxi, *bot = is_temp ? h->bott : h->botxi;
    double *Am = is_temp ? h->cTmat : h->cXmat, *fac = is_temp ? h->cTfac : h->cXfac;
    int *piv = is_temp ? h->cTpiv : h->cXpiv;
    ts_set_imex_rhs(h, t, a, h->work);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) {
        double *rr = tl_dbuf(2 * (size_t)N), *ri = rr + N;
        cplx *w = h->work + (size_t)i * N;
        w[0] = top[i];
        w[N - 1] = bot[i];
        for (int k = 0; k < N; k++) {
            w[k] *= fac[(size_t)i * N + k];
            rr[k] = creal(w[k]);
            ri[k] = cimag(w[k]);
        }
        getrs_r(N, Am + (size_t)i * N * N, piv + (size_t)i * N, rr);
        getrs_r(N, Am + (size_t)i * N * N, piv + (size_t)i * N, ri);
        for (int k = 0; k < N; k++) fld[(size_t)i * N + k] = k < h->ncheb ? rr[k] + I * ri[k] : 0.0;
    }
    ts_rotate(h, t, a);
    double diff = is_temp ? h->TdiffFac : h->XidiffFac;
    double *hdif = is_temp ? h->hdifT : h->hdifXi;
    if (t->istage == t->nstages) coll_rhs_imp_scalar(h, fld, dfld, a, 1, t->l_imp_calc_rhs[0], 1, diff, hdif);
    else coll_rhs_imp_scalar(h, fld, dfld, a, t->istage + 1, t->l_imp_calc_rhs[t->istage], 1, diff, hdif);
}

/* update_om_coll_dmat (update_psi_coll_dmat.f90:126-466): omega from its
 * own equation with omega = 0 at the walls, psi from -L_psi psi = omega with
 * psi = 0 at the walls, then the multiples c1, c2 of the homogeneous pairs
 * (omega_2,3, psi_2,3) that make psi meet the velocity conditions */
static void coll_update_psi_dmat(or_handle *h)
{
    int N = h->N, nc = h->ncheb;
    tscheme *t = &h->ts;
    tarray *a = &h->dpsidt;
    double wimp = t->wimp_lin[0];
    ts_set_imex_rhs(h, t, a, h->work);
    double *u0 = dalloc(N), *om0 = dalloc(N);
    int nm0 = -1;
    for (int i = 0; i < h->nm; i++)
        if (h->idx2m[i] == 0) nm0 = i;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        cplx *w = h->work + (size_t)i * N;
        if (m == 0) {
            u0[0] = 0.0;
            u0[N - 1] = 0.0;
            for (int k = 1; k < N - 1; k++) u0[k] = creal(w[k]);
            getrs_r(N, h->cUmat, h->cUpiv, u0);
            for (int k = nc; k < N; k++) u0[k] = 0.0;
            continue;
        }
        const double *of = h->dOfac + (size_t)i * 2 * N, *pf = h->dPfac + (size_t)i * 2 * N;
        double *rr = tl_dbuf(2 * (size_t)N), *ri = rr + N;
        cplx *om = h->om + (size_t)i * N, *ps = h->psi + (size_t)i * N, *og = tl_cbuf(5, N);
        /* omega (Chebyshev coefficients) */
        rr[0] = ri[0] = rr[N - 1] = ri[N - 1] = 0.0;
        for (int k = 1; k < N - 1; k++) {
            size_t q = (size_t)i * N + k;
            cplx v = w[k];
            if (h->c.l_buo_imp && (h->l_heat || h->l_chem)) {
                if (h->l_heat) v -= wimp * h->BuoFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->temp[q];
                if (h->l_chem) v -= wimp * h->ChemFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->xi[q];
            }
            rr[k] = creal(v);
            ri[k] = cimag(v);
        }
        for (int k = 0; k < N; k++) { rr[k] *= of[k]; ri[k] *= of[k]; }
        getrs_r(N, h->dOmat + (size_t)i * N * N, h->dOpiv + (size_t)i * N, rr);
        getrs_r(N, h->dOmat + (size_t)i * N * N, h->dOpiv + (size_t)i * N, ri);
        for (int k = 0; k < N; k++) om[k] = k < nc ? (rr[k] + I * ri[k]) * of[N + k] : 0.0;
        /* omega on the grid, then psi_1 */
        memcpy(og, om, sizeof(cplx) * N);
        costf_vec(h, og);
        rr[0] = ri[0] = rr[N - 1] = ri[N - 1] = 0.0;
        for (int k = 1; k < N - 1; k++) { rr[k] = -creal(og[k]); ri[k] = -cimag(og[k]); }
        for (int k = 0; k < N; k++) { rr[k] *= pf[k]; ri[k] *= pf[k]; }
        getrs_r(N, h->dPmat + (size_t)i * N * N, h->dPpiv + (size_t)i * N, rr);
        getrs_r(N, h->dPmat + (size_t)i * N * N, h->dPpiv + (size_t)i * N, ri);
        for (int k = 0; k < N; k++) ps[k] = k < nc ? (rr[k] + I * ri[k]) * pf[N + k] : 0.0;
        /* solve_influence_matrix (:1256-1300) */
        cplx bt, bb;
        coll_bc_dmat(h, ps, &bt, &bb);
        const double *inf = h->dinfl + (size_t)i * 4;
        cplx c1 = -inf[0] * bt - inf[1] * bb, c2 = -inf[2] * bt - inf[3] * bb;
        const double *p2 = h->dpsi2 + (size_t)i * N, *p3 = h->dpsi3 + (size_t)i * N;
        const double *o2 = h->dom2 + (size_t)i * N, *o3 = h->dom3 + (size_t)i * N;
        for (int k = 0; k < N; k++) {
            ps[k] = ps[k] + c1 * p2[k] + c2 * p3[k];
            om[k] = om[k] + c1 * o2[k] + c2 * o3[k];
        }
    }
    if (nm0 >= 0) {
        get_dr_real(h, u0, om0, 0);
        costf_real(h, u0);
        for (int k = 0; k < N; k++) h->psi[(size_t)nm0 * N + k] = 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) get_dr_vec(h, h->psi + (size_t)i * N, h->work + (size_t)i * N, 0);
    costf_field(h, h->psi);
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        for (int k = 0; k < N; k++) {
            size_t q = (size_t)i * N + k;
            if (m == 0) {
                h->us[q] = 0.0;
                h->up[q] = u0[k];
                h->om[q] = om0[k] + h->or1[k] * u0[k];
            } else {
                h->us[q] = I * (double)m * h->or1[k] * h->psi[q];
                h->up[q] = -h->work[q] - h->beta[k] * h->psi[q];
            }
        }
    }
    ts_rotate(h, t, a);
    if (t->istage == t->nstages) coll_rhs_imp_psi(h, a, 1, t->l_imp_calc_rhs[0], 1);
    else coll_rhs_imp_psi(h, a, t->istage + 1, t->l_imp_calc_rhs[t->istage], 1);
    free(u0);
    free(om0);
}

/* update_om_coll_smat (update_psi_coll_smat.f90:88-321) */
static void coll_update_psi(or_handle *h)
{
    if (!h->c.l_direct_solve) {
        coll_update_psi_dmat(h);
        return;
    }
    int N = h->N, N2 = 2 * N;
    tscheme *t = &h->ts;
    tarray *a = &h->dpsidt;
    double wimp = t->wimp_lin[0];
    ts_set_imex_rhs(h, t, a, h->work);
    double *u0 = dalloc(N), *om0 = dalloc(N);
    int nm0 = -1;
    for (int i = 0; i < h->nm; i++)
        if (h->idx2m[i] == 0) nm0 = i;
    /* one mode per OpenMP thread; only the m = 0 iteration writes u0 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        cplx *w = h->work + (size_t)i * N;
        cplx *rhs = tl_cbuf(5, N2);
        if (m == 0) {
            u0[0] = 0.0;
            u0[N - 1] = 0.0;
            for (int k = 1; k < N - 1; k++) u0[k] = creal(w[k]);
            getrs_r(N, h->cUmat, h->cUpiv, u0);
            for (int k = h->ncheb; k < N; k++) u0[k] = 0.0;
            continue;
        }
        memset(rhs, 0, sizeof(cplx) * N2);
        for (int k = 1; k < N - 1; k++) {
            size_t q = (size_t)i * N + k;
            cplx v = w[k];
            if (h->c.l_buo_imp && (h->l_heat || h->l_chem)) {
                if (h->l_heat) v -= wimp * h->BuoFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->temp[q];
                if (h->l_chem) v -= wimp * h->ChemFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->xi[q];
            }
            rhs[k] = v;
        }
        const double *fac = h->cPfac + (size_t)i * 2 * N2;
        for (int k = 0; k < N2; k++) rhs[k] *= fac[k];
        getrs_c(N2, h->cPmat + (size_t)i * N2 * N2, h->cPpiv + (size_t)i * N2, rhs);
        for (int k = 0; k < N2; k++) rhs[k] *= fac[N2 + k];
        for (int k = 0; k < N; k++) {
            h->om[(size_t)i * N + k] = k < h->ncheb ? rhs[k] : 0.0;
            h->psi[(size_t)i * N + k] = k < h->ncheb ? rhs[N + k] : 0.0;
        }
    }
    if (nm0 >= 0) {
        get_dr_real(h, u0, om0, 0);
        costf_real(h, u0);
        for (int k = 0; k < N; k++) h->psi[(size_t)nm0 * N + k] = 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) get_dr_vec(h, h->psi + (size_t)i * N, h->work + (size_t)i * N, 0);
    costf_field(h, h->psi);
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        for (int k = 0; k < N; k++) {
            size_t q = (size_t)i * N + k;
            if (m == 0) {
                h->us[q] = 0.0;
                h->up[q] = u0[k];
                h->om[q] = om0[k] + h->or1[k] * u0[k];
            } else {
                h->us[q] = I * (double)m * h->or1[k] * h->psi[q];
                h->up[q] = -h->work[q] - h->beta[k] * h->psi[q];
            }
        }
    }
    ts_rotate(h, t, a);
    if (t->istage == t->nstages) coll_rhs_imp_psi(h, a, 1, t->l_imp_calc_rhs[0], 1);
    else coll_rhs_imp_psi(h, a, t->istage + 1, t->l_imp_calc_rhs[t->istage], 1);
    free(u0);
    free(om0);
}

/* assemble_imex (dirk_schemes.f90:882-921): old + sum_s b_exp(s) expl(s)
 * + sum_s b_imp(s) impl(s) over all stages */
static void ts_assemble(or_handle *h, const tscheme *t, const tarray *a, cplx *rhs)
{
    size_t F = (size_t)h->nm * h->N;
    for (size_t i = 0; i < F; i++) rhs[i] = a->old[i];
    for (int s = 0; s < t->nstages; s++)
        for (size_t i = 0; i < F; i++) rhs[i] += t->bass_exp[s] * a->expl[s * F + i];
    for (int s = 0; s < t->nstages; s++)
        for (size_t i = 0; i < F; i++) rhs[i] += t->bass_imp[s] * a->impl[s * F + i];
}

/* assemble_temp_coll / assemble_xi_coll (update_temp_coll.f90:238-307):
 * interior grid values from the assembled right-hand side (real at m = 0),
 * boundary values by robin_bc, then the first-stage implicit terms */
static void coll_assemble_scalar(or_handle *h, int is_temp)
{
    int N = h->N;
    tscheme *t = &h->ts;
    tarray *a = is_temp ? &h->dTdt : &h->dxidt;
    cplx *fld = is_temp ? h->temp : h->xi, *dfld = is_temp ? h->dtemp : h->dxi;
    cplx *top = is_temp ? h->topt : h->topxi, *bot = is_temp ? h->bott : h->botxi;
    int ktop = is_temp ? h->c.ktopt : h->c.ktopxi, kbot = is_temp ? h->c.kbott : h->c.kbotxi;
    ts_assemble(h, t, a, h->work);
    /* Dirichlet: (a, b) = (0, 1); Neumann: (1, 0) */
    double at = ktop == 1 ? 0.0 : 1.0, bt = ktop == 1 ? 1.0 : 0.0;
    double ab = kbot == 1 ? 0.0 : 1.0, bb = kbot == 1 ? 1.0 : 0.0;
    for (int i = 0; i < h->nm; i++) {
        cplx *f = fld + (size_t)i * N, *w = h->work + (size_t)i * N;
        for (int k = 1; k < N - 1; k++) f[k] = h->idx2m[i] == 0 ? creal(w[k]) : w[k];
        robin_bc(h, at, bt, top[i], ab, bb, bot[i], f);
    }
    double diff = is_temp ? h->TdiffFac : h->XidiffFac;
    double *hdif = is_temp ? h->hdifT : h->hdifXi;
    coll_rhs_imp_scalar(h, fld, dfld, a, 1, t->l_imp_calc_rhs[0], 0, diff, hdif);
}

/* get_assMat (update_psi_coll_smat.f90:839-917): psi from omega with the
 * non-penetration and no-slip / free-slip rows, boundary_fac columns, LU */
static int coll_ass_mat(or_handle *h)
{
    int N = h->N, nc = h->ncheb;
    double rn = h->rnorm;
    h->cAss = calloc((size_t)h->nm * N * N, sizeof(cplx));
    h->cAssPiv = calloc((size_t)h->nm * N, sizeof(int));
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        if (m == 0) continue;
        double dm2 = (double)m * (double)m;
        cplx *A = h->cAss + (size_t)i * N * N;
        for (int c = 0; c < N; c++) {
            int in = c < nc;
            A[0 + c * N] = in ? rn * CM(0, c) : 0.0;
            A[(N - 1) + c * N] = in ? rn * CM(N - 1, c) : 0.0;
            A[1 + c * N] = !in ? 0.0 : (h->c.ktopv == 1 ? rn * (CD2M(0, c) - h->or1[0] * CDM(0, c)) : rn * CDM(0, c));
            A[(N - 2) + c * N] = !in ? 0.0
                                     : (h->c.kbotv == 1 ? rn * (CD2M(N - 1, c) - h->or1[N - 1] * CDM(N - 1, c))
                                                        : rn * CDM(N - 1, c));
            for (int r = 2; r < N - 2; r++)
                A[r + c * N] = -rn * (CD2M(r, c) + (h->or1[r] + h->beta[r]) * CDM(r, c) +
                                      (h->or1[r] * h->beta[r] + h->dbeta[r] - dm2 * h->or2[r]) * CM(r, c));
        }
        for (int r = 0; r < N; r++) {
            A[r] *= h->boundary_fac;
            A[r + (size_t)(N - 1) * N] *= h->boundary_fac;
        }
        int info = 0;
        getrf_c(N, A, h->cAssPiv + (size_t)i * N, &info);
        if (info) return info;
    }
    h->lAss = 1;
    return 0;
}

/* assemble_psi_coll_smat (update_psi_coll_smat.f90:529-672) */
static int coll_assemble_psi(or_handle *h)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    tscheme *t = &h->ts;
    tarray *a = &h->dpsidt;
    if (!h->lAss) {
        int info = coll_ass_mat(h);
        if (info) return info;
    }
    ts_assemble(h, t, a, h->work);
    double *u0 = dalloc(N), *om0 = dalloc(N);
    cplx *rhs = calloc(N, sizeof(cplx)), *dpsi = calloc(F, sizeof(cplx)), *d2psi = calloc(F, sizeof(cplx));
    int nm0 = -1;
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        cplx *w = h->work + (size_t)i * N;
        if (m == 0) {
            nm0 = i;
            cplx *up0 = h->up + (size_t)i * N;
            for (int k = 1; k < N - 1; k++) up0[k] = creal(w[k]);
            int kt = h->c.ktopv, kb = h->c.kbotv;
            if (kt != 1 && kb != 1) {
                up0[0] = 0.0;
                up0[N - 1] = 0.0;
            } else if (kt == 1 && kb != 1) {
                robin_bc(h, 1.0, -h->or1[0], 0.0, 0.0, 1.0, 0.0, up0);
            } else if (kb == 1 && kt != 1) {
                robin_bc(h, 0.0, 1.0, 0.0, 1.0, -h->or1[N - 1], 0.0, up0);
            } else {
                robin_bc(h, 1.0, -h->or1[0], 0.0, 1.0, -h->or1[N - 1], 0.0, up0);
            }
            continue;
        }
        rhs[0] = rhs[1] = rhs[N - 2] = rhs[N - 1] = 0.0;
        for (int k = 2; k < N - 2; k++) rhs[k] = w[k];
        getrs_c(N, h->cAss + (size_t)i * N * N, h->cAssPiv + (size_t)i * N, rhs);
        for (int k = 0; k < N; k++) h->psi[(size_t)i * N + k] = k < h->ncheb ? rhs[k] : 0.0;
    }
    if (nm0 >= 0) {
        for (int k = 0; k < N; k++) u0[k] = creal(h->up[(size_t)nm0 * N + k]);
        get_dr_real(h, u0, om0, 1);
        for (int k = 0; k < N; k++) h->psi[(size_t)nm0 * N + k] = 0.0;
    }
    for (int i = 0; i < h->nm; i++)
        get_ddr_vec(h, h->psi + (size_t)i * N, dpsi + (size_t)i * N, d2psi + (size_t)i * N, 0);
    costf_field(h, h->psi);
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        double dm2 = (double)m * (double)m;
        for (int k = 0; k < N; k++) {
            size_t q = (size_t)i * N + k;
            if (m == 0) {
                h->us[q] = 0.0;
                h->om[q] = om0[k] + h->or1[k] * u0[k];
            } else {
                h->us[q] = I * (double)m * h->or1[k] * h->psi[q];
                h->up[q] = -dpsi[q] - h->beta[k] * h->psi[q];
                h->om[q] = -(d2psi[q] + (h->or1[k] + h->beta[k]) * dpsi[q] +
                             (h->or1[k] * h->beta[k] + h->dbeta[k] - dm2 * h->or2[k]) * h->psi[q]);
            }
        }
    }
    coll_rhs_imp_psi(h, a, 1, t->l_imp_calc_rhs[0], 0);
    free(u0); free(om0); free(rhs); free(dpsi); free(d2psi);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Finite differences in radius (radial_scheme = 'FD'): second-order        */
/* stencils on the stretched grid, tridiagonal T / xi / u_phi(m=0) and      */
/* pentadiagonal psi systems with ghost zones                               */
/* ------------------------------------------------------------------------ */

/* Fornberg's weights (Math. Comp. 51 (1988) 699-706), the algorithm of
 * populate_fd_weights (finite_differences.f90:501-546): c[j * (m + 1) + k] is
 * the weight of the value at x[j] in the k-th derivative at z. */
static void fd_weights(double z, const double *x, int nd, int m, double *c)
{
    const int w = m + 1;
    double c1 = 1.0, c4 = x[0] - z;
    for (int i = 0; i < (nd + 1) * w; i++) c[i] = 0.0;
    c[0] = 1.0;
    for (int i = 1; i <= nd; i++) {
        int mn = i < m ? i : m;
        double c2 = 1.0, c5 = c4;
        c4 = x[i] - z;
        for (int j = 0; j < i; j++) {
            double c3 = x[i] - x[j];
            c2 *= c3;
            if (j == i - 1) {
                for (int k = mn; k >= 1; k--)
                    c[i * w + k] = c1 * ((double)k * c[(i - 1) * w + k - 1] - c5 * c[(i - 1) * w + k]) / c2;
                c[i * w] = -c1 * c5 * c[(i - 1) * w] / c2;
            }
            for (int k = mn; k >= 1; k--) c[j * w + k] = (c4 * c[j * w + k] - (double)k * c[j * w + k - 1]) / c3;
            c[j * w] = c4 * c[j * w] / c3;
        }
        c1 = c2;
    }
}

/* get_FD_grid (finite_differences.f90:87-188), annulus (r_icb > 0):
 * r[0] = r_cmb, decreasing; ratio1 = fd_stretch, ratio2 = fd_ratio */
static int fd_grid(or_handle *h, double ratio1, double ratio2)
{
    int N = h->N;
    double *r = h->r, drb, dra;
    r[0] = h->r_cmb;
    if (ratio2 == 1.0) {
        drb = 1.0 / ((double)N - 1.0);
        for (int k = 1; k < N; k++) r[k] = r[k - 1] - drb;
    } else {
        int nbp = (int)((double)(N - 1) / (2.0 * (1.0 + ratio1)));
        int nbulk = N - 1 - 2 * nbp;
        dra = exp(log(ratio2) / (double)nbp);
        drb = 1.0;
        for (int i = 0; i < nbp; i++) drb *= dra;
        drb = 1.0 / ((double)nbulk + 2.0 * dra * ((1.0 - drb) / (1.0 - dra)));
        for (int i = 0; i < nbp; i++) drb *= dra;
        for (int k = 1; k <= nbp; k++) {
            r[k] = r[k - 1] - drb;
            drb /= dra;
        }
        for (int k = 1; k <= nbulk; k++) r[k + nbp] = r[k + nbp - 1] - drb;
        for (int k = 1; k <= nbp; k++) {
            drb *= dra;
            r[k + nbp + nbulk] = r[k + nbp + nbulk - 1] - drb;
        }
    }
    if (r[N - 1] != h->r_icb) {
        if (fabs(r[N - 1] - h->r_icb) > drb) return -1;
        r[N - 1] = h->r_icb;
    }
    return 0;
}

/* get_FD_coeffs, order 2 and boundary order 2 (finite_differences.f90:190-446):
 * the stencils the step uses */
static void fd_coeffs(or_handle *h)
{
    int N = h->N;
    const double *r = h->r;
    double x[5], c[25];
    for (int k = 0; k < N; k++) {
        double *d1 = h->fdr + 3 * k, *d2 = h->fddr + 3 * k;
        if (k == 0 || k == N - 1) {
            double drl = k == 0 ? r[1] - r[0] : r[N - 1] - r[N - 2];
            d1[0] = -0.5 / drl; d1[1] = 0.0; d1[2] = 0.5 / drl;
            d2[0] = 1.0 / drl / drl; d2[1] = -2.0 / drl / drl; d2[2] = 1.0 / drl / drl;
        } else {
            for (int od = 0; od < 3; od++) x[od] = r[k - 1 + od] - r[k];
            fd_weights(0.0, x, 2, 2, c);
            for (int od = 0; od < 3; od++) { d1[od] = c[od * 3 + 1]; d2[od] = c[od * 3 + 2]; }
        }
    }
    for (int od = 0; od < 3; od++) x[od] = r[od] - r[0];
    fd_weights(0.0, x, 2, 2, c);
    for (int od = 0; od < 3; od++) h->fdr_top[od] = c[od * 3 + 1];
    for (int od = 0; od < 3; od++) x[od] = r[N - 1 - od] - r[N - 1];
    fd_weights(0.0, x, 2, 2, c);
    for (int od = 0; od < 3; od++) h->fdr_bot[od] = c[od * 3 + 1];
    /* 3rd and 4th derivatives: symmetric ghost points beyond the walls */
    for (int k = 1; k < N - 1; k++) {
        for (int od = 0; od < 5; od++) {
            if (k == 1 && od == 0) x[od] = 2.0 * (r[0] - r[1]);
            else if (k == N - 2 && od == 4) x[od] = 2.0 * (r[N - 1] - r[N - 2]);
            else x[od] = r[k - 2 + od] - r[k];
        }
        fd_weights(0.0, x, 4, 4, c);
        for (int od = 0; od < 5; od++) { h->fdddr[5 * k + od] = c[od * 5 + 3]; h->fddddr[5 * k + od] = c[od * 5 + 4]; }
    }
}

#define TG(g, i, k) (g)[(size_t)(i) * (N + 2) + (k) + 1]
#define PG(g, i, k) (g)[(size_t)(i) * (N + 4) + (k) + 2]

static void fd_alloc(or_handle *h)
{
    int N = h->N, nm = h->nm;
    double rc = h->r_cmb;
    h->or3 = dalloc(N); h->dbeta = dalloc(N); h->d2beta = dalloc(N); h->d3beta = dalloc(N);
    for (int n = 0; n < N; n++) h->or3[n] = h->or2[n] * h->or1[n];
    /* radial.f90:212-226 (sphere container; zero at r_cmb) */
    for (int n = 1; n < N && !h->c.l_non_rot; n++) {
        double r = h->r[n], s2 = rc * rc - r * r, hh = h->height[n];
        h->dbeta[n] = -(rc * rc + r * r) / (s2 * s2);
        h->d2beta[n] = -128.0 * r * (3.0 * rc * rc + r * r) / pow(hh, 6);
        h->d3beta[n] = -1536.0 * (pow(rc, 4) + pow(r, 4) + 6.0 * r * r * rc * rc) / pow(hh, 8);
    }
    h->fdr = dalloc(3 * N); h->fddr = dalloc(3 * N); h->fdddr = dalloc(5 * N); h->fddddr = dalloc(5 * N);
    fd_coeffs(h);
    h->tTd = dalloc(nm * N); h->tTl = dalloc(nm * N); h->tTu = dalloc(nm * N);
    h->tXd = dalloc(nm * N); h->tXl = dalloc(nm * N); h->tXu = dalloc(nm * N);
    h->uD = dalloc(N); h->uL = dalloc(N); h->uU = dalloc(N);
    size_t F = (size_t)nm * N;
    h->pD = calloc(F, sizeof(cplx)); h->pL1 = calloc(F, sizeof(cplx)); h->pL2 = calloc(F, sizeof(cplx));
    h->pU1 = calloc(F, sizeof(cplx)); h->pU2 = calloc(F, sizeof(cplx));
    h->tg = calloc((size_t)nm * (N + 2), sizeof(cplx));
    h->xg = calloc((size_t)nm * (N + 2), sizeof(cplx));
    h->pg = calloc((size_t)nm * (N + 4), sizeof(cplx));
    h->u0g = calloc(N + 2, sizeof(cplx));
}

/* prepare_mat_3 (parallel_solver.f90:135-187): diag <- 1/pivot */
static void fd_lu3(int N, double *d, double *lo, double *up)
{
    lo[0] = 0.0;
    up[N - 1] = 0.0;
    d[0] = 1.0 / d[0];
    for (int k = 1; k < N; k++) d[k] = 1.0 / (d[k] - lo[k] * up[k - 1] * d[k - 1]);
}

/* prepare_mat_5 (parallel_solver.f90:189-288) */
static void fd_lu5(int N, cplx *d, cplx *l1, cplx *l2, cplx *u1, cplx *u2)
{
    l1[0] = 0.0; l2[0] = 0.0; l2[1] = 0.0;
    u1[N - 1] = 0.0; u2[N - 1] = 0.0; u2[N - 2] = 0.0;
    u1[1] = u1[1] - l1[1] * u2[0] / d[0];
    d[1] = d[1] - l1[1] * u1[0] / d[0];
    for (int k = 2; k < N; k++) {
        l1[k] = l1[k] - l2[k] * u1[k - 2] / d[k - 2];
        u1[k] = u1[k] - l1[k] * u2[k - 1] / d[k - 1];
        d[k] = d[k] - l1[k] * u1[k - 1] / d[k - 1] - l2[k] * u2[k - 2] / d[k - 2];
    }
    for (int k = 0; k < N; k++) {
        d[k] = 1.0 / d[k];
        u1[k] *= d[k]; u2[k] *= d[k]; l1[k] *= d[k]; l2[k] *= d[k];
    }
}

/* get_tMat_Rdist (update_temp_fd.f90:387-483) and its xi twin */
static void fd_scalar_mat(or_handle *h, int i, double diff, int ktop, int kbot, double *d, double *lo, double *up)
{
    int N = h->N, m = h->idx2m[i];
    double w = h->ts.wimp_lin[0], dm2 = (double)m * (double)m;
    for (int k = 0; k < N; k++) {
        const double *a = h->fdr + 3 * k, *b = h->fddr + 3 * k;
        d[k] = 1.0 - w * diff * (b[1] + h->or1[k] * a[1] - dm2 * h->or2[k]);
        lo[k] = -w * diff * (b[0] + h->or1[k] * a[0]);
        up[k] = -w * diff * (b[2] + h->or1[k] * a[2]);
    }
    if (ktop == 1) { d[0] = 1.0; up[0] = 0.0; lo[0] = 0.0; }
    else up[0] = up[0] + lo[0];
    if (kbot == 1) { d[N - 1] = 1.0; up[N - 1] = 0.0; lo[N - 1] = 0.0; }
    else lo[N - 1] = up[N - 1] + lo[N - 1];
    fd_lu3(N, d, lo, up);
}

/* get_uphimat_Rdist (update_psi_fd.f90:772-836) */
static void fd_uphi_mat(or_handle *h)
{
    int N = h->N;
    double w = h->ts.wimp_lin[0], nu = h->ViscFac;
    double *d = h->uD, *lo = h->uL, *up = h->uU;
    for (int k = 0; k < N; k++) {
        const double *a = h->fdr + 3 * k, *b = h->fddr + 3 * k;
        d[k] = 1.0 - w * (nu * (b[1] + h->or1[k] * a[1] - h->or2[k]) - h->CorFac * h->ekpump[k]);
        lo[k] = -w * nu * (b[0] + h->or1[k] * a[0]);
        up[k] = -w * nu * (b[2] + h->or1[k] * a[2]);
    }
    if (h->c.ktopv == 1) {
        up[0] = up[0] + lo[0];
        d[0] = d[0] - 2.0 * (h->r[1] - h->r[0]) * h->or1[0] * lo[0];
    } else { d[0] = 1.0; lo[0] = 0.0; up[0] = 0.0; }
    if (h->c.kbotv == 1) {
        lo[N - 1] = lo[N - 1] + up[N - 1];
        d[N - 1] = d[N - 1] + 2.0 * (h->r[N - 1] - h->r[N - 2]) * h->or1[N - 1] * up[N - 1];
    } else { d[N - 1] = 1.0; lo[N - 1] = 0.0; up[N - 1] = 0.0; }
    fd_lu3(N, d, lo, up);
}

/* get_psimat_Rdist (update_psi_fd.f90:608-770), one m /= 0 */
static void fd_psi_mat(or_handle *h, int i)
{
    int N = h->N, m = h->idx2m[i];
    size_t o = (size_t)i * N;
    double w = h->ts.wimp_lin[0], nu = h->ViscFac, cf = h->CorFac, dm = (double)m, dm2 = dm * dm;
    cplx *d = h->pD + o, *l1 = h->pL1 + o, *l2 = h->pL2 + o, *u1 = h->pU1 + o, *u2 = h->pU2 + o;
    for (int k = 0; k < N; k++) {
        const double *a = h->fdr + 3 * k, *b = h->fddr + 3 * k, *c3 = h->fdddr + 5 * k, *c4 = h->fddddr + 5 * k;
        double o1 = h->or1[k], o2 = h->or2[k], o3 = h->or3[k], be = h->beta[k], db = h->dbeta[k];
        double d2b = h->d2beta[k], d3b = h->d3beta[k], ek = h->ekpump[k], eu = h->ekp_up[k];
        double A3 = 2.0 * o1 + be;
        double A2 = 3.0 * db + 2.0 * be * o1 - o2 - 2.0 * dm2 * o2;
        double A1 = 3.0 * d2b + 4.0 * db * o1 - be * o2 + o3 + dm2 * o2 * (2.0 * o1 - be);
        double A0 = d3b + 2.0 * d2b * o1 - db * o2 + be * o3 - dm2 * o2 * (4.0 * o2 + be * o1 + db - dm2 * o2);
        cplx P0 = db + be * o1 - dm2 * o2 + be * eu - I * dm * o1 * (h->ekp_us[k] + I * dm * h->ekp_dusdp[k]);
        d[k] = -(b[1] + (o1 + be) * a[1] + (be * o1 + db - dm2 * o2)) +
               w * (nu * (c4[2] + A3 * c3[2] + A2 * b[1] + A1 * a[1] + A0) -
                    cf * ek * (b[1] + (be + o1 + eu) * a[1] + P0) - cf * be * I * dm * o1);
        l1[k] = -(b[0] + (o1 + be) * a[0]) +
                w * (nu * (c4[1] + A3 * c3[1] + A2 * b[0] + A1 * a[0]) - cf * ek * (b[0] + (be + o1 + eu) * a[0]));
        u1[k] = -(b[2] + (o1 + be) * a[2]) +
                w * (nu * (c4[3] + A3 * c3[3] + A2 * b[2] + A1 * a[2]) - cf * ek * (b[2] + (be + o1 + eu) * a[2]));
        l2[k] = w * nu * (c4[0] + A3 * c3[0]);
        u2[k] = w * nu * (c4[4] + A3 * c3[4]);
    }
    /* non-penetration rows and the second condition through the ghost point */
    d[0] = 1.0; l1[0] = 0.0; l2[0] = 0.0; u1[0] = 0.0; u2[0] = 0.0;
    if (h->c.ktopv == 1) {
        double dr = h->r[1] - h->r[0], fac = (1.0 - 0.5 * h->or1[0] * dr) / (1.0 + 0.5 * h->or1[0] * dr);
        d[1] = d[1] - fac * l2[1];
        l1[1] = l1[1] + 2.0 / (1.0 + 0.5 * h->or1[0] * dr) * l2[1];
    } else {
        d[1] = d[1] + l2[1];
    }
    d[N - 1] = 1.0; l1[N - 1] = 0.0; l2[N - 1] = 0.0; u1[N - 1] = 0.0; u2[N - 1] = 0.0;
    if (h->c.kbotv == 1) {
        double dr = h->r[N - 1] - h->r[N - 2];
        double fac = (1.0 + 0.5 * h->or1[N - 1] * dr) / (1.0 - 0.5 * h->or1[N - 1] * dr);
        d[N - 2] = d[N - 2] - fac * u2[N - 2];
        u1[N - 2] = u1[N - 2] + 2.0 * u2[N - 2] / (1.0 - 0.5 * h->or1[N - 1] * dr);
    } else {
        d[N - 2] = d[N - 2] + u2[N - 2];
    }
    fd_lu5(N, d, l1, l2, u1, u2);
}

static void fd_build(or_handle *h)
{
    int N = h->N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) {
        if (h->l_heat)
            fd_scalar_mat(h, i, h->TdiffFac, h->c.ktopt, h->c.kbott, h->tTd + (size_t)i * N, h->tTl + (size_t)i * N,
                          h->tTu + (size_t)i * N);
        if (h->l_chem)
            fd_scalar_mat(h, i, h->XidiffFac, h->c.ktopxi, h->c.kbotxi, h->tXd + (size_t)i * N,
                          h->tXl + (size_t)i * N, h->tXu + (size_t)i * N);
        if (h->idx2m[i] != 0) fd_psi_mat(h, i);
    }
    fd_uphi_mat(h);
}

/* solver_up_3 / solver_dn_3 (and solver_single): x[k + 1], ghosts x[0], x[N + 1] */
static void fd_solve3(int N, const double *d, const double *lo, const double *up, cplx *x)
{
    x[1] = x[1] - lo[0] * x[0];
    for (int k = 1; k < N; k++) x[k + 1] = x[k + 1] - d[k - 1] * lo[k] * x[k];
    for (int k = N - 1; k >= 0; k--) x[k + 1] = (x[k + 1] - up[k] * x[k + 2]) * d[k];
}

/* solver_up_5 / solver_dn_5: x[k + 2], ghosts two deep */
static void fd_solve5(int N, const cplx *d, const cplx *l1, const cplx *l2, const cplx *u1, const cplx *u2, cplx *x)
{
    for (int k = 0; k < N; k++) x[k + 2] = d[k] * x[k + 2] - l1[k] * x[k + 1] - l2[k] * x[k];
    for (int k = N - 1; k >= 0; k--) x[k + 2] = x[k + 2] - u1[k] * x[k + 3] - u2[k] * x[k + 4];
}

/* get_dr_Rloc (radial_derivatives.f90:669-760) of one radial column */
static void fd_dr_col(or_handle *h, const cplx *f, cplx *df)
{
    int N = h->N;
    for (int k = 0; k < N; k++) {
        const double *a = h->fdr + 3 * k;
        cplx fm = k > 0 ? f[k - 1] : 0.0, fp = k < N - 1 ? f[k + 1] : 0.0;
        df[k] = a[0] * fm + a[1] * f[k] + a[2] * fp;
    }
    df[0] = 0.0;
    for (int od = 0; od < 3; od++) df[0] += h->fdr_top[od] * f[od];
    df[N - 1] = 0.0;
    for (int od = 0; od < 3; od++) df[N - 1] += h->fdr_bot[od] * f[N - 1 - od];
}

/* finish_explicit_assembly_Rdist (mloop_fd.f90:78-110; update_temp_fd.f90:278-315,
 * update_xi_fd.f90:280-290, update_psi_fd.f90:397-470); buoyancy is explicit */
static void fd_finish_exp(or_handle *h, int is)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->nm; i++) {
        cplx *w = tl_cbuf(5, N);
        int m = h->idx2m[i];
        size_t o = (size_t)i * N;
        cplx *ep = h->dpsidt.expl + is * F + o;
        if (m != 0) {
            fd_dr_col(h, h->dVsOm + o, w);
            for (int k = 0; k < N; k++) {
                ep[k] = ep[k] - h->or1[k] * w[k];
                if (!h->c.l_coriolis_imp) ep[k] = ep[k] + h->CorFac * h->beta[k] * h->us[o + k];
                if (h->l_heat) ep[k] = ep[k] - h->BuoFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->temp[o + k];
                if (h->l_chem) ep[k] = ep[k] - h->ChemFac * h->rgrav[k] * h->or1[k] * I * (double)m * h->xi[o + k];
            }
        }
        if (h->l_heat) {
            cplx *et = h->dTdt.expl + is * F + o;
            fd_dr_col(h, h->dVsT + o, w);
            for (int k = 0; k < N; k++)
                et[k] = et[k] - h->or1[k] * w[k] - h->us[o + k] * (h->dtcond[k] - h->tadvz * h->beta[k] * h->tcond[k]);
        }
        if (h->l_chem) {
            cplx *ex = h->dxidt.expl + is * F + o;
            fd_dr_col(h, h->dVsXi + o, w);
            for (int k = 0; k < N; k++) ex[k] = ex[k] - h->or1[k] * w[k] - h->us[o + k] * h->dxicond[k];
        }
    }
}

/* the wall patterns (t_top, t_bot per m) of the finite-difference T
 * (is_temp) or xi equation when l_fd_bc_pattern is set, else NULL: the
 * reference holds the perturbation at zero on the walls */
static void fd_patterns(const or_handle *h, int is_temp, const cplx **top, const cplx **bot)
{
    *top = *bot = NULL;
    if (!h->c.l_fd_bc_pattern) return;
    *top = is_temp ? h->topt : h->topxi;
    *bot = is_temp ? h->bott : h->botxi;
}

/* fill_ghosts_temp / fill_ghosts_xi (update_temp_fd.f90:117-169): fixed
 * value, 2 f_0 - f_1; fixed flux, the mirror point f_1, plus
 * 2 (r_0 - r_1) q at the outer wall and minus 2 (r_{N-2} - r_{N-1}) q at
 * the inner one for a flux pattern q (the terms the reference comments out
 * at :142,162) */
static void fd_fill_ghosts_scalar(or_handle *h, cplx *g, int ktop, int kbot, const cplx *qt, const cplx *qb)
{
    int N = h->N;
    double dt = h->r[0] - h->r[1], db = h->r[N - 2] - h->r[N - 1];
    for (int i = 0; i < h->nm; i++) {
        TG(g, i, -1) = ktop == 1 ? 2.0 * TG(g, i, 0) - TG(g, i, 1)
                                 : TG(g, i, 1) + (qt ? (or_plant == 8 ? -2.0 : 2.0) * dt * qt[i] : 0.0);
        TG(g, i, N) = kbot == 1 ? 2.0 * TG(g, i, N - 1) - TG(g, i, N - 2)
                                : TG(g, i, N - 2) - (qb ? 2.0 * db * qb[i] : 0.0);
    }
}

/* the wall rows of a finite-difference T / xi right-hand side: the
 * pattern value on a fixed-value wall (zero without patterns), and on a
 * fixed-flux wall the inhomogeneous part of the ghost point moved to the
 * right-hand side, minus the wall row's ghost coefficient times it */
static void fd_wall_rows(or_handle *h, cplx *g, int i, int ktop, int kbot, double diff, const cplx *qt,
                         const cplx *qb)
{
    int N = h->N;
    double w = h->ts.wimp_lin[0];
    if (ktop == 1) TG(g, i, 0) = qt ? qt[i] : 0.0;
    else if (qt) {
        const double *a = h->fdr, *b = h->fddr;
        double lo = -w * diff * (b[0] + h->or1[0] * a[0]);
        if (or_plant != 7) TG(g, i, 0) -= lo * 2.0 * (h->r[0] - h->r[1]) * qt[i];
    }
    if (kbot == 1) TG(g, i, N - 1) = qb ? qb[i] : 0.0;
    else if (qb) {
        const double *a = h->fdr + 3 * (N - 1), *b = h->fddr + 3 * (N - 1);
        double up = -w * diff * (b[2] + h->or1[N - 1] * a[2]);
        TG(g, i, N - 1) += up * 2.0 * (h->r[N - 2] - h->r[N - 1]) * qb[i];
    }
}

/* fill_ghosts_psi (update_psi_fd.f90:140-215) */
static void fd_fill_ghosts_psi(or_handle *h)
{
    int N = h->N;
    cplx *u = h->u0g;
    double dr = h->r[1] - h->r[0];
    u[0] = h->c.ktopv == 1 ? u[2] - 2.0 * dr * h->or1[0] * u[1] : 2.0 * u[1] - u[2];
    double db = h->r[N - 1] - h->r[N - 2];
    u[N + 1] = h->c.kbotv == 1 ? u[N - 1] + 2.0 * db * h->or1[N - 1] * u[N] : 2.0 * u[N] - u[N - 1];
    for (int i = 0; i < h->nm; i++) {
        if (h->idx2m[i] == 0) continue;
        if (h->c.ktopv == 1)
            PG(h->pg, i, -1) = -(1.0 - 0.5 * h->or1[0] * dr) / (1.0 + 0.5 * h->or1[0] * dr) * PG(h->pg, i, 1);
        else
            PG(h->pg, i, -1) = PG(h->pg, i, 1);
        PG(h->pg, i, -2) = 0.0;
        if (h->c.kbotv == 1)
            PG(h->pg, i, N) = -(1.0 + 0.5 * h->or1[N - 1] * db) / (1.0 - 0.5 * h->or1[N - 1] * db) * PG(h->pg, i, N - 2);
        else
            PG(h->pg, i, N) = PG(h->pg, i, N - 2);
        PG(h->pg, i, N + 1) = 0.0;
    }
}

/* get_temp_rhs_imp_ghost (update_temp_fd.f90:217-276) and its xi twin */
static void fd_rhs_imp_scalar(or_handle *h, const cplx *g, cplx *dfld, tarray *a, int istage, int l_calc_lin,
                              double diff)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    for (int i = 0; i < h->nm; i++) {
        double dm2 = (double)h->idx2m[i] * (double)h->idx2m[i];
        for (int k = 0; k < N; k++) {
            const double *d1 = h->fdr + 3 * k, *d2 = h->fddr + 3 * k;
            cplx fm = TG(g, i, k - 1), f0 = TG(g, i, k), fp = TG(g, i, k + 1);
            cplx df = d1[0] * fm + d1[1] * f0 + d1[2] * fp, ddf = d2[0] * fm + d2[1] * f0 + d2[2] * fp;
            size_t q = (size_t)i * N + k;
            if (dfld) dfld[q] = df;
            if (istage == 1) a->old[q] = f0;
            if (l_calc_lin) a->impl[(istage - 1) * F + q] = diff * (ddf + h->or1[k] * df - dm2 * h->or2[k] * f0);
        }
    }
}

/* get_psi_rhs_imp_ghost (update_psi_fd.f90:247-395): old and implicit terms,
 * and u_s, u_phi, omega from psi (m /= 0) and u_phi (m = 0) */
static void fd_rhs_imp_psi(or_handle *h, tarray *a, int istage, int l_calc_lin)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    const cplx *u = h->u0g;
    double nu = h->ViscFac, cf = h->CorFac;
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        double dm = (double)m, dm2 = dm * dm;
        for (int k = 0; k < N; k++) {
            size_t q = (size_t)i * N + k;
            const double *d1 = h->fdr + 3 * k, *d2 = h->fddr + 3 * k;
            double o1 = h->or1[k], o2 = h->or2[k], o3 = h->or3[k], be = h->beta[k], db = h->dbeta[k];
            if (m == 0) {
                cplx du = d1[0] * u[k] + d1[1] * u[k + 1] + d1[2] * u[k + 2];
                cplx ddu = d2[0] * u[k] + d2[1] * u[k + 1] + d2[2] * u[k + 2];
                if (istage == 1) a->old[q] = u[k + 1];
                if (l_calc_lin)
                    a->impl[(istage - 1) * F + q] = nu * (ddu + o1 * du - o2 * u[k + 1]) - cf * h->ekpump[k] * u[k + 1];
                h->us[q] = 0.0;
                h->up[q] = u[k + 1];
                h->om[q] = du + o1 * u[k + 1];
                continue;
            }
            const double *c3 = h->fdddr + 5 * k, *c4 = h->fddddr + 5 * k;
            cplx f[5];
            for (int j = 0; j < 5; j++) f[j] = PG(h->pg, i, k - 2 + j);
            cplx dp = d1[0] * f[1] + d1[1] * f[2] + d1[2] * f[3];
            cplx ddp = d2[0] * f[1] + d2[1] * f[2] + d2[2] * f[3];
            cplx dddp = c3[0] * f[0] + c3[1] * f[1] + c3[2] * f[2] + c3[3] * f[3] + c3[4] * f[4];
            cplx d4p = c4[0] * f[0] + c4[1] * f[1] + c4[2] * f[2] + c4[3] * f[3] + c4[4] * f[4];
            cplx ps = f[2];
            if (istage == 1) a->old[q] = -(ddp + (o1 + be) * dp + (o1 * be + db - dm2 * o2) * ps);
            if (l_calc_lin) {
                double d2b = h->d2beta[k], d3b = h->d3beta[k];
                cplx v = -nu * (d4p + (be + 2.0 * o1) * dddp + (3.0 * db + 2.0 * be * o1 - o2 - 2.0 * dm2 * o2) * ddp +
                                (3.0 * d2b + 4.0 * db * o1 - be * o2 + o3 + dm2 * o2 * (2.0 * o1 - be)) * dp +
                                (d3b + 2.0 * d2b * o1 - db * o2 + be * o3 -
                                 dm2 * o2 * (4.0 * o2 + be * o1 + db - dm2 * o2)) * ps);
                if (h->c.l_coriolis_imp) v = v + cf * be * I * dm * o1 * ps;
                if (h->c.l_ek_pump)
                    v = v + cf * h->ekpump[k] *
                                (ddp + (o1 + be + h->ekp_up[k]) * dp +
                                 (db + be * o1 - dm2 * o2 + be * h->ekp_up[k] -
                                  I * dm * o1 * (h->ekp_us[k] + I * dm * h->ekp_dusdp[k])) * ps);
                a->impl[(istage - 1) * F + q] = v;
            }
            h->us[q] = I * dm * o1 * ps;
            h->om[q] = -(ddp + (o1 + be) * dp + (be * o1 + db - dm2 * o2) * ps);
            h->up[q] = -dp - be * ps;
            h->psi[q] = ps;
        }
    }
}

/* mloop_Rdist (mloop_fd.f90:148-239) for one stage */
static void fd_mloop(or_handle *h)
{
    tscheme *t = &h->ts;
    int N = h->N, nm = h->nm;
    size_t F = (size_t)nm * N;
    cplx *rhs = malloc(sizeof(cplx) * F);
    /* prepare_temp_fd / prepare_xi_fd: right-hand sides, boundary rows */
    for (int f = 0; f < 2; f++) {
        if ((f == 0 && !h->l_heat) || (f == 1 && !h->l_chem)) continue;
        cplx *g = f == 0 ? h->tg : h->xg;
        int ktop = f == 0 ? h->c.ktopt : h->c.ktopxi, kbot = f == 0 ? h->c.kbott : h->c.kbotxi;
        const cplx *qt, *qb;
        fd_patterns(h, f == 0, &qt, &qb);
        ts_set_imex_rhs(h, t, f == 0 ? &h->dTdt : &h->dxidt, rhs);
        for (int i = 0; i < nm; i++) {
            for (int k = 0; k < N; k++) TG(g, i, k) = rhs[(size_t)i * N + k];
            fd_wall_rows(h, g, i, ktop, kbot, f == 0 ? h->TdiffFac : h->XidiffFac, qt, qb);
            TG(g, i, -1) = 0.0;
            TG(g, i, N) = 0.0;
        }
    }
    /* prepare_psi_fd */
    ts_set_imex_rhs(h, t, &h->dpsidt, rhs);
    for (int i = 0; i < nm; i++) {
        int m = h->idx2m[i];
        for (int k = 0; k < N; k++) PG(h->pg, i, k) = rhs[(size_t)i * N + k];
        if (m == 0) {
            for (int k = 0; k < N; k++) h->u0g[k + 1] = rhs[(size_t)i * N + k];
            if (h->c.ktopv != 1) h->u0g[1] = 0.0;
            h->u0g[0] = 0.0;
            if (h->c.kbotv != 1) h->u0g[N] = 0.0;
            h->u0g[N + 1] = 0.0;
        } else {
            PG(h->pg, i, 0) = 0.0; PG(h->pg, i, -1) = 0.0; PG(h->pg, i, -2) = 0.0;
            PG(h->pg, i, N - 1) = 0.0; PG(h->pg, i, N) = 0.0; PG(h->pg, i, N + 1) = 0.0;
        }
    }
    free(rhs);
    /* solves */
    fd_solve3(N, h->uD, h->uL, h->uU, h->u0g);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nm; i++) {
        size_t o = (size_t)i * N;
        if (h->l_heat) fd_solve3(N, h->tTd + o, h->tTl + o, h->tTu + o, &TG(h->tg, i, -1));
        if (h->l_chem) fd_solve3(N, h->tXd + o, h->tXl + o, h->tXu + o, &TG(h->xg, i, -1));
        if (h->idx2m[i] != 0)
            fd_solve5(N, h->pD + o, h->pL1 + o, h->pL2 + o, h->pU1 + o, h->pU2 + o, &PG(h->pg, i, -2));
    }
    /* ghost zones, then the arrays of the next stage */
    {
        const cplx *qt, *qb;
        fd_patterns(h, 1, &qt, &qb);
        if (h->l_heat) fd_fill_ghosts_scalar(h, h->tg, h->c.ktopt, h->c.kbott, qt, qb);
        fd_patterns(h, 0, &qt, &qb);
        if (h->l_chem) fd_fill_ghosts_scalar(h, h->xg, h->c.ktopxi, h->c.kbotxi, qt, qb);
    }
    fd_fill_ghosts_psi(h);
    int inext = t->istage == t->nstages ? 1 : t->istage + 1;
    int lcalc = t->l_imp_calc_rhs[inext - 1];
    if (h->l_heat) {
        ts_rotate(h, t, &h->dTdt);
        fd_rhs_imp_scalar(h, h->tg, h->dtemp, &h->dTdt, inext, lcalc, h->TdiffFac);
        for (int i = 0; i < nm; i++)
            for (int k = 0; k < N; k++) h->temp[(size_t)i * N + k] = TG(h->tg, i, k);
    }
    if (h->l_chem) {
        ts_rotate(h, t, &h->dxidt);
        fd_rhs_imp_scalar(h, h->xg, h->dxi, &h->dxidt, inext, lcalc, h->XidiffFac);
        for (int i = 0; i < nm; i++)
            for (int k = 0; k < N; k++) h->xi[(size_t)i * N + k] = TG(h->xg, i, k);
    }
    ts_rotate(h, t, &h->dpsidt);
    fd_rhs_imp_psi(h, &h->dpsidt, inext, lcalc);
}

/* ghosted copies of the fields and the first implicit state: the FD branch
 * of get_start_fields (init_fields.f90:129-186), also the start-up
 * recomputation of start_from_another_scheme (step_time.f90:426-463) */
static void fd_from_fields(or_handle *h, int with_om)
{
    int N = h->N, nm = h->nm;
    if (h->l_heat) {
        for (int i = 0; i < nm; i++)
            for (int k = 0; k < N; k++) TG(h->tg, i, k) = h->temp[(size_t)i * N + k];
        const cplx *qt, *qb;
        fd_patterns(h, 1, &qt, &qb);
        fd_fill_ghosts_scalar(h, h->tg, h->c.ktopt, h->c.kbott, qt, qb);
        fd_rhs_imp_scalar(h, h->tg, h->dtemp, &h->dTdt, 1, 1, h->TdiffFac);
    }
    if (h->l_chem) {
        for (int i = 0; i < nm; i++)
            for (int k = 0; k < N; k++) TG(h->xg, i, k) = h->xi[(size_t)i * N + k];
        const cplx *qt, *qb;
        fd_patterns(h, 0, &qt, &qb);
        fd_fill_ghosts_scalar(h, h->xg, h->c.ktopxi, h->c.kbotxi, qt, qb);
        fd_rhs_imp_scalar(h, h->xg, h->dxi, &h->dxidt, 1, 1, h->XidiffFac);
    }
    /* psi is unknown: rebuild it from u_s */
    for (int i = 0; i < nm; i++) {
        int m = h->idx2m[i];
        for (int k = 0; k < N; k++) {
            size_t q = (size_t)i * N + k;
            PG(h->pg, i, k) = m == 0 ? 0.0 : h->us[q] / (I * (double)m) * h->r[k];
        }
    }
    int i0 = -1;
    for (int i = 0; i < nm; i++)
        if (h->idx2m[i] == 0) i0 = i;
    for (int k = 0; k < N; k++) h->u0g[k + 1] = i0 >= 0 ? h->up[(size_t)i0 * N + k] : 0.0;
    fd_fill_ghosts_psi(h);
    fd_rhs_imp_psi(h, &h->dpsidt, 1, 1);
    if (with_om) {
        /* omega from u_phi and u_s (get_dr_Rloc of u_phi) */
        cplx *w = malloc(sizeof(cplx) * N);
        for (int i = 0; i < nm; i++) {
            int m = h->idx2m[i];
            size_t o = (size_t)i * N;
            fd_dr_col(h, h->up + o, w);
            for (int k = 0; k < N; k++)
                h->om[o + k] = w[k] + h->or1[k] * h->up[o + k] - h->or1[k] * I * (double)m * h->us[o + k];
        }
        free(w);
    }
}

/* test hook: the finite-difference stencils, [dr 3N | ddr 3N | dddr 5N |
 * ddddr 5N | dr_top 3 | dr_bot 3]; -1 unless the run uses finite differences */
int or_fd_stencils(or_handle *h, double *out)
{
    if (!h->fd) return -1;
    int N = h->N;
    memcpy(out, h->fdr, sizeof(double) * 3 * N);
    memcpy(out + 3 * N, h->fddr, sizeof(double) * 3 * N);
    memcpy(out + 6 * N, h->fdddr, sizeof(double) * 5 * N);
    memcpy(out + 11 * N, h->fddddr, sizeof(double) * 5 * N);
    memcpy(out + 16 * N, h->fdr_top, sizeof(double) * 3);
    memcpy(out + 16 * N + 3, h->fdr_bot, sizeof(double) * 3);
    return 0;
}

/* rInt_R, finite-difference branch (integration.f90:60-96) */
static double rint_fd(or_handle *h, const double *f)
{
    int N = h->N;
    const double *r = h->r;
    double s = 0.0;
#define SIMPS(n)                                                                                     \
    do {                                                                                             \
        double h2 = r[(n) + 1] - r[(n)], h1 = r[(n)] - r[(n) - 1];                                   \
        s += (h1 + h2) / 6.0 * (f[(n) - 1] * (2.0 * h1 - h2) / h1 + f[(n)] * (h1 + h2) * (h1 + h2) / (h1 * h2) + \
                                f[(n) + 1] * (2.0 * h2 - h1) / h2);                                  \
    } while (0)
    if (N % 2 == 1) {
        for (int n = 1; n < N - 1; n += 2) SIMPS(n);
        return -s;
    }
    s = 0.5 * (r[1] - r[0]) * (f[1] + f[0]);
    for (int n = 2; n < N - 1; n += 2) SIMPS(n);
    s += 0.5 * (r[N - 1] - r[N - 2]) * (f[N - 1] + f[N - 2]);
    for (int n = 1; n < N - 1; n += 2) SIMPS(n);
#undef SIMPS
    return -0.5 * s;
}

/* assemble_temp_Rdist / assemble_xi_Rdist (update_temp_fd.f90:317-385,
 * update_xi_fd.f90:293-361): the assembled right-hand side is the new field
 * (real at m = 0), zero on a fixed-value wall; ghosts by fill_ghosts_*,
 * then the first-stage implicit terms */
static void fd_assemble_scalar(or_handle *h, int is_temp)
{
    int N = h->N;
    tscheme *t = &h->ts;
    tarray *a = is_temp ? &h->dTdt : &h->dxidt;
    cplx *g = is_temp ? h->tg : h->xg, *fld = is_temp ? h->temp : h->xi;
    int ktop = is_temp ? h->c.ktopt : h->c.ktopxi, kbot = is_temp ? h->c.kbott : h->c.kbotxi;
    const cplx *qt, *qb;
    fd_patterns(h, is_temp, &qt, &qb);
    ts_assemble(h, t, a, h->work);
    for (int i = 0; i < h->nm; i++) {
        for (int k = 0; k < N; k++) {
            cplx w = h->work[(size_t)i * N + k];
            TG(g, i, k) = h->idx2m[i] == 0 ? creal(w) : w;
        }
        if (ktop == 1) TG(g, i, 0) = qt ? qt[i] : 0.0;
        if (kbot == 1) TG(g, i, N - 1) = qb ? qb[i] : 0.0;
    }
    fd_fill_ghosts_scalar(h, g, ktop, kbot, qt, qb);
    fd_rhs_imp_scalar(h, g, is_temp ? h->dtemp : h->dxi, a, 1, t->l_imp_calc_rhs[0],
                      is_temp ? h->TdiffFac : h->XidiffFac);
    for (int i = 0; i < h->nm; i++)
        for (int k = 0; k < N; k++) fld[(size_t)i * N + k] = TG(g, i, k);
}

/* -L psi = omega for one m /= 0 with psi = 0 at both walls: the
 * tridiagonal matrix of get_elliptic_mat_Rdist (update_psi_fd.f90:838-888)
 * factorised by prepare_mat_3 and solved by solver_up_3 / solver_dn_3.
 * x: N + 2 values, x[k + 1] at point k; on entry the omega rows
 * (x[1] and x[N] are set to zero), on return psi */
static void fd_ell_solve(or_handle *h, int m, cplx *x)
{
    int N = h->N;
    double *d = dalloc(N), *lo = dalloc(N), *up = dalloc(N);
    double dm2 = (double)m * (double)m;
    for (int k = 0; k < N; k++) {
        const double *a1 = h->fdr + 3 * k, *a2 = h->fddr + 3 * k;
        double c1 = h->beta[k] + h->or1[k];
        d[k] = -(a2[1] + c1 * a1[1] + (h->dbeta[k] + h->beta[k] * h->or1[k] - dm2 * h->or2[k]));
        up[k] = -(a2[2] + c1 * a1[2]);
        lo[k] = -(a2[0] + c1 * a1[0]);
    }
    d[0] = 1.0; up[0] = 0.0; lo[0] = 0.0;          /* non-penetration rows */
    d[N - 1] = 1.0; up[N - 1] = 0.0; lo[N - 1] = 0.0;
    fd_lu3(N, d, lo, up);
    x[0] = 0.0; x[1] = 0.0; x[N] = 0.0; x[N + 1] = 0.0;
    fd_solve3(N, d, lo, up, x);
    free(d); free(lo); free(up);
}

/* assemble_psi_Rloc (update_psi_fd.f90:472-606): u_phi(m = 0) is the
 * assembled right-hand side (zero on a no-slip wall); for m /= 0 psi solves
 * -L psi = omega (fd_ell_solve); ghosts by fill_ghosts_psi, then u_s,
 * u_phi, omega and the first-stage implicit terms */
static void fd_assemble_psi(or_handle *h)
{
    int N = h->N;
    tscheme *t = &h->ts;
    ts_assemble(h, t, &h->dpsidt, h->work);
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        const cplx *w = h->work + (size_t)i * N;
        if (m == 0) {
            for (int k = 0; k < N; k++) h->u0g[k + 1] = w[k];
            if (h->c.ktopv != 1) h->u0g[1] = 0.0;
            if (h->c.kbotv != 1) h->u0g[N] = 0.0;
            h->u0g[0] = 0.0;
            h->u0g[N + 1] = 0.0;
            continue;
        }
        for (int k = -2; k <= N + 1; k++) PG(h->pg, i, k) = 0.0;
        for (int k = 0; k < N; k++) PG(h->pg, i, k) = w[k];
        fd_ell_solve(h, m, &PG(h->pg, i, -1));
    }
    fd_fill_ghosts_psi(h);
    fd_rhs_imp_psi(h, &h->dpsidt, 1, t->l_imp_calc_rhs[0]);
}
#undef TG
#undef PG

static int one_step(or_handle *h)
{
    tscheme *t = &h->ts;
    size_t F = (size_t)h->nm * h->N;
    int lMat = 0;
    t->istage = 1;
    for (int n_stage = 1; n_stage <= t->nstages; n_stage++) {
        int is = t->istage - 1;
        if (t->l_exp_calc[n_stage - 1] && h->fd) {
            radial_loop(h, is);
            fd_finish_exp(h, is);
        } else if (t->l_exp_calc[n_stage - 1] && h->coll) {
            radial_loop(h, is);
            coll_finish_exp(h, is);
        } else if (t->l_exp_calc[n_stage - 1]) {
            radial_loop(h, is);
            if (h->l_heat) finish_exp_scalar(h, h->dVsT, h->dTdt.expl + is * F, h->dtcond, h->tcond, 1);
            if (h->l_chem) finish_exp_scalar(h, h->dVsXi, h->dxidt.expl + is * F, h->dxicond, h->xicond, 0);
            finish_exp_psi(h, h->dpsidt.expl + is * F);
        }
        int l_new_dt = 0;
        if (t->istage == 1) {
            double dt_new;
            dt_courant(h, &l_new_dt, &dt_new);
            if (ts_set_dt_array(t, dt_new, h->c.dtMax / 1.0e6)) return -2;
            ts_set_weights(t, h->c.alpha, &h->lMatNext);
            h->time += t->dt[0];
        }
        lMat = h->lMatNext;
        if ((l_new_dt || lMat) && t->istage == 1) lMat = 1;
        h->lMatNext = 0;
        /* start_from_another_scheme (step_time.f90:417-503) on a fresh start
         * (l_bridge_step, namelists.f90:365): the first nold-1 steps of a
         * multistep scheme other than CNAB2 recompute the implicit terms and
         * take the CNAB2 weights */
        if (h->l_bridge && t->family == 0 && t->scheme != PZ_CNAB2 && h->n_time_step + 1 <= t->nold - 1) {
            if (h->fd) {
                fd_from_fields(h, 0);
            } else if (h->coll) {
                if (h->l_heat) coll_rhs_imp_scalar(h, h->temp, h->dtemp, &h->dTdt, 1, 1, 0, h->TdiffFac, h->hdifT);
                if (h->l_chem) coll_rhs_imp_scalar(h, h->xi, h->dxi, &h->dxidt, 1, 1, 0, h->XidiffFac, h->hdifXi);
                coll_rhs_imp_psi(h, &h->dpsidt, 1, 1, 0);
            } else {
                if (h->l_heat) rhs_imp_scalar(h, h->temp_hat, &h->dTdt, 1, 1, h->TdiffFac, h->hdifT);
                if (h->l_chem) rhs_imp_scalar(h, h->xi_hat, &h->dxidt, 1, 1, h->XidiffFac, h->hdifXi);
                /* psi in the composition slot, as at step_time.f90:475-478 */
                rhs_imp_psi(h, h->psi_hat, h->up, h->temp, h->psi, &h->dpsidt, 1, 1);
            }
            ts_bridge(t, h->c.alpha);
        }
        /* start_with_ab1 (multistep_schemes.f90:526-534) on the first step
         * after a restart that changed ek or the radial method
         * (step_time.f90:511-514) */
        if (h->l_ab1 && h->n_time_step == 0 && t->family == 0) {
            t->wexp[0] = t->dt[0];
            for (int o = 1; o < t->nexp; o++) t->wexp[o] = 0.0;
            h->l_ab1 = 0;
        }
        if (t->l_assembly && t->istage == t->nstages) break; /* step_time.f90:313 */
        if (lMat) {
            int info = 0;
            if (h->fd) fd_build(h);
            else info = h->coll ? coll_build(h) : build_matrices(h, 1);
            if (info) return info;
        }
        if (h->fd) {
            fd_mloop(h);
        } else if (h->coll) {
            if (h->l_heat) coll_update_scalar(h, 1);
            if (h->l_chem) coll_update_scalar(h, 0);
            coll_update_psi(h);
        } else {
            if (h->l_heat) update_scalar(h, lMat, 1);
            if (h->l_chem) update_scalar(h, lMat, 0);
            update_psi(h);
        }
        t->istage++;
    }
    if (t->l_assembly && h->fd) { /* assemble_stage_Rdist (step_time.f90:353-362, mloop_fd.f90:241-274) */
        if (h->l_heat) fd_assemble_scalar(h, 1);
        if (h->l_chem) fd_assemble_scalar(h, 0);
        fd_assemble_psi(h);
    } else if (t->l_assembly) { /* assemble_stage (step_time.f90:363-368, mloop.f90:105-160) */
        if (h->l_heat) coll_assemble_scalar(h, 1);
        if (h->l_chem) coll_assemble_scalar(h, 0);
        int info = coll_assemble_psi(h);
        if (info) return info;
    }
    h->n_time_step++;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* API                                                                       */
/* ------------------------------------------------------------------------ */
void or_destroy(or_handle *h)
{
    if (!h) return;
    /* process-lifetime test helper: buffers are released with the process */
    fft_free(h->fdct);
    fft_free(h->fphi);
    if (h->tabs) free_optabs(h->tabs);
    free(h);
}

int or_dims(const or_handle *h, int *nm, int *nr, int *np)
{
    *nm = h->nm;
    *nr = h->N;
    *np = h->nphi;
    return 0;
}

static cplx *field_ptr(or_handle *h, int f)
{
    switch (f) {
    case PZ_US: return h->us;
    case PZ_UP: return h->up;
    case PZ_OM: return h->om;
    case PZ_PSI: return h->psi;
    case PZ_TEMP: return h->temp;
    case PZ_XI: return h->xi;
    case PZ_TEMP_HAT: return h->temp_hat;
    case PZ_XI_HAT: return h->xi_hat;
    case PZ_PSI_HAT: return h->psi_hat;
    case PZ_DTEMP: return h->dtemp;
    default: return NULL;
    }
}

/* host layout: (n_m, n_r) m fastest, interleaved re/im */
int or_set_field(or_handle *h, int f, const double *host)
{
    cplx *p = field_ptr(h, f);
    if (!p) return -1;
    for (int i = 0; i < h->nm; i++)
        for (int k = 0; k < h->N; k++) {
            size_t s = 2 * ((size_t)k * h->nm + i);
            p[(size_t)i * h->N + k] = host[s] + I * host[s + 1];
        }
    return 0;
}

int or_get_field(or_handle *h, int f, double *host)
{
    cplx *p = field_ptr(h, f);
    if (!p) return -1;
    for (int i = 0; i < h->nm; i++)
        for (int k = 0; k < h->N; k++) {
            size_t s = 2 * ((size_t)k * h->nm + i);
            host[s] = creal(p[(size_t)i * h->N + k]);
            host[s + 1] = cimag(p[(size_t)i * h->N + k]);
        }
    return 0;
}

/* Gaspari-Cohn pieces (useful.f90:295-341) */
static double gc_center(double c, double L)
{
    double pi = PI_F;
    return pi * L * L * L * (1.0 - exp(-2.0 * c / L)) - 2.0 * pi * c * L * (c + L) * exp(-2.0 * c / L);
}
static double gc_middle(double r, double c, double L)
{
    double pi = PI_F;
    return pi / 3.0 * L * r * (r + 3.0 * L) * exp(-r / L) +
           (2.0 * pi * L * L * (c + L) * (c + L)) / r * exp(-2.0 * c / L) * (1.0 - (1.0 - r / (c + L)) * exp(r / L)) +
           pi * L * L * L * (exp(-r / L) - exp((r - 2.0 * c) / L));
}
static double gc_edge(double r, double c, double L)
{
    double pi = PI_F;
    return 2.0 * pi * L / r * exp(-r / L) *
           (0.5 * (r * (r + L) * (2.0 * c - r)) + 1.0 / 3.0 * (pow(r - c, 3) - c * c * c) - L * (c + L) * (r - c + L) +
            L * (c + L) * (c + L) * exp((r - 2.0 * c) / L));
}

/* initProp (init_fields.f90:260-360), deterministic branches */
static int init_prop(or_handle *h, cplx *prop, int init, double amp)
{
    int N = h->N;
    double pi = PI_F;
    if (init > 0) {
        if (init % h->minc != 0 || init > h->mmax) return -3;
        int idx = init / h->minc;
        for (int n = 0; n < N; n++) prop[(size_t)idx * N + n] += sin(pi * (h->r[n] - h->r_icb)) * amp;
    } else if (init == -1) {
        double sigma_r = 0.1 / sqrt(2.0);
        double rc = 0.5 * (h->r_cmb + h->r_icb);
        int ir = 0;
        double best = fabs(h->r[0] - rc);
        for (int n = 1; n < N; n++)
            if (fabs(h->r[n] - rc) < best) { best = fabs(h->r[n] - rc); ir = n; }
        rc = h->r[ir];
        double c1 = 0.5 * (h->r[0] - rc), c2 = 0.5 * (rc - h->r[N - 1]), L = 0.5 * sigma_r;
        double *gasp = dalloc(N);
        for (int n = 0; n < N; n++) {
            double r = h->r[n];
            if (n == ir) gasp[n] = 1.0;
            else if (r < rc - c2) gasp[n] = gc_edge(rc - r, c2, L);
            else if (r >= rc - c2 && r < rc) gasp[n] = gc_middle(rc - r, c2, L);
            else if (r > rc && r <= rc + c1) gasp[n] = gc_middle(r - rc, c1, L);
            else if (r > rc + c1) gasp[n] = gc_edge(r - rc, c1, L);
        }
        for (int n = 0; n < N; n++) {
            if (n < ir) gasp[n] = gasp[n] / gc_center(c1, L);
            else if (n > ir) gasp[n] = gasp[n] / gc_center(c2, L);
        }
        double phi0 = pi / h->minc;
        double *pf = dalloc(h->nphi);
        cplx *sp = malloc(sizeof(cplx) * h->nm);
        for (int n = 0; n < N; n++) {
            double c_r = amp * gasp[n];
            for (int j = 0; j < h->nphi; j++) {
                double phi = (double)j * 2.0 * pi / h->minc / (double)h->nphi;
                pf[j] = c_r * exp(-(phi - phi0) * (phi - phi0) / (0.2 * 0.2));
            }
            phi_fft(h, pf, sp, h->mR[n]);
            for (int i = 0; i < h->nm; i++) prop[(size_t)i * N + n] += sp[i];
        }
        free(gasp);
        free(pf);
        free(sp);
    } else {
        return -4; /* unseeded random branch is not reproducible */
    }
    return 0;
}

int or_init_fields(or_handle *h, int init_t, double amp_t, int init_xi, double amp_xi, int init_u,
                   double amp_u)
{
    int rc = 0;
    if (init_t != 0 && h->l_heat) rc |= init_prop(h, h->temp, init_t, amp_t);
    if (init_xi != 0 && h->l_chem) rc |= init_prop(h, h->xi, init_xi, amp_xi);
    if (init_u != 0) {
        /* initU (init_fields.f90:455-509) */
        int N = h->N;
        double pi = PI_F;
        if (init_u > 0) {
            if (init_u % h->minc != 0 || init_u > h->mmax) return -3;
            int idx = init_u / h->minc;
            for (int n = 0; n < N; n++) h->us[(size_t)idx * N + n] += sin(pi * (h->r[n] - h->r_icb)) * amp_u;
            for (int n = 0; n < N; n++)
                for (int i = 0; i < h->nm; i++) {
                    int m = h->idx2m[i];
                    h->om[(size_t)i * N + n] = -I * (double)m * h->or1[n] * h->us[(size_t)i * N + n];
                }
        } else {
            for (int n = 0; n < N; n++) h->up[n] += amp_u * sin(pi * (h->r[n] - h->r_icb));
        }
    }
    return rc;
}

/* get_start_fields, fresh start (init_fields.f90:96-112, 186-255) */
int or_start(or_handle *h)
{
    int N = h->N;
    size_t F = (size_t)h->nm * N;
    tscheme *t = &h->ts;
    h->time = h->restart ? h->rst_time : 0.0;
    h->n_time_step = 0;
    for (int i = 0; i < 8; i++) t->dt[i] = h->restart ? h->rst_dt[i] : h->c.dtMax;
    if (h->restart && !h->fd && !h->coll) {
        /* get_start_fields after read_checkpoint (init_fields.f90:67-95): the
         * integration method rebuilds psi(m /= 0) from the stored u_s */
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            for (int k = 0; k < N; k++) {
                double h2 = h->r_cmb * h->r_cmb - h->r[k] * h->r[k];
                h->psi[(size_t)i * N + k] = (m > 0 && k > 0) ? -I * h->r[k] / (double)m / h2 * h->us[(size_t)i * N + k] : 0.0;
            }
        }
    }
    h->restart = 0;
    int dummy = 0;
    ts_set_weights(t, h->c.alpha, &dummy);
    /* dummy Courant limits before the first radial loop (step_time.f90:124-126) */
    for (int n = 0; n < N; n++) { h->dtr[n] = 1.0e10; h->dth[n] = 1.0e10; }
    if (h->fd) {
        fd_from_fields(h, 1);
        h->started = 1;
        h->lMatNext = 1;
        return 0;
    }
    for (int i = 0; i < h->nm; i++) {
        if (h->l_heat) get_dr_vec(h, h->temp + (size_t)i * N, h->dtemp + (size_t)i * N, 1);
        if (h->l_chem) get_dr_vec(h, h->xi + (size_t)i * N, h->dxi + (size_t)i * N, 1);
        get_dr_vec(h, h->up + (size_t)i * N, h->work + (size_t)i * N, 1);
    }
    for (int n_r = 0; n_r < N; n_r++)
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            size_t q = (size_t)i * N + n_r;
            h->om[q] = h->work[q] + h->or1[n_r] * h->up[q] - h->or1[n_r] * I * (double)m * h->us[q];
        }
    if (h->coll) {
        /* get_start_fields, collocation branch (init_fields.f90:228-244) */
        if (h->l_heat) coll_rhs_imp_scalar(h, h->temp, h->dtemp, &h->dTdt, 1, 1, 0, h->TdiffFac, h->hdifT);
        if (h->l_chem) coll_rhs_imp_scalar(h, h->xi, h->dxi, &h->dxidt, 1, 1, 0, h->XidiffFac, h->hdifXi);
        coll_rhs_imp_psi(h, &h->dpsidt, 1, 1, 0);
        h->started = 1;
        h->lMatNext = 1;
        return 0;
    }
    memcpy(h->temp_hat, h->temp, sizeof(cplx) * F);
    memcpy(h->xi_hat, h->xi, sizeof(cplx) * F);
    memcpy(h->psi_hat, h->psi, sizeof(cplx) * F);
    if (h->l_heat) costf_field(h, h->temp_hat);
    if (h->l_chem) costf_field(h, h->xi_hat);
    costf_field(h, h->psi_hat);
    if (h->l_heat) rhs_imp_scalar(h, h->temp_hat, &h->dTdt, 1, 1, h->TdiffFac, h->hdifT);
    if (h->l_chem) rhs_imp_scalar(h, h->xi_hat, &h->dxidt, 1, 1, h->XidiffFac, h->hdifXi);
    /* the reference passes psi in the composition slot here (init_fields.f90:250-251) */
    rhs_imp_psi(h, h->psi_hat, h->up, h->temp, h->psi, &h->dpsidt, 1, 1);
    h->started = 1;
    h->lMatNext = 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Checkpoints (checkpoints.f90): version 7 writer and the reader with its    */
/* grid, truncation and scheme remap                                          */
/* ------------------------------------------------------------------------ */

/* costf1 of a real vector of length n (dct_fftw.f90:231-246) as the direct
 * cosine sum of the DCT-I, sqrt(0.5/(n-1)) scaling */
static void ck_costf(int n, double *f)
{
    double *g = malloc(sizeof(double) * n);
    const ld pi = 3.14159265358979323846264338327950288L;
    for (int k = 0; k < n; k++) {
        ld acc = (ld)f[0] + ((k & 1) ? -(ld)f[n - 1] : (ld)f[n - 1]);
        for (int j = 1; j < n - 1; j++) acc += 2.0L * (ld)f[j] * cosl(pi * (ld)(((long)j * k) % (2 * (n - 1))) / (ld)(n - 1));
        g[k] = (double)(acc * sqrtl(0.5L / (ld)(n - 1)));
    }
    memcpy(f, g, sizeof(double) * n);
    free(g);
}

/* the cubic through four points at xn (polynomial_interpolation,
 * useful.f90:99-179, there by Neville's recursion), in Lagrange form */
static cplx ck_cubic(const double *x, const cplx *y, double xn)
{
    cplx out = 0;
    for (int a = 0; a < 4; a++) {
        double w = 1.0;
        for (int b = 0; b < 4; b++)
            if (b != a) w *= (xn - x[b]) / (x[a] - x[b]);
        out += w * y[a];
    }
    return out;
}

typedef struct {
    int n_r, n_m, m_max, minc, cheb, order_boundary;
    double *r;
} ck_grid;

/* map_field / map_field_r (checkpoints.f90:709-866) of one field stored
 * (n_m_old, n_r_old), m fastest, into dst [m][r] of this run */
static void ck_map(or_handle *h, const ck_grid *g, const cplx *old, cplx *dst, int lBc, int l_phys)
{
    int N = h->N, cheb_new = !h->fd, ob_new = h->fd ? 2 : 0;
    int same = g->n_r == N && g->order_boundary == ob_new && g->cheb == cheb_new;
    int nmax = N > g->n_r ? N : g->n_r;
    cplx *col = malloc(sizeof(cplx) * nmax);
    double *re = malloc(sizeof(double) * nmax), *im = malloc(sizeof(double) * nmax);
    for (int i = 0; i < h->nm; i++) {
        int m = h->idx2m[i];
        int io = (m <= g->m_max && m % g->minc == 0) ? m / g->minc : -1;
        cplx *d = dst + (size_t)i * N;
        if (io < 0) {
            for (int k = 0; k < N; k++) d[k] = 0;
            continue;
        }
        for (int k = 0; k < g->n_r; k++) col[k] = old[(size_t)k * g->n_m + io];
        if (same) {
            for (int k = 0; k < N; k++) d[k] = col[k];
            continue;
        }
        if (lBc && l_phys) col[0] = 2.0 * col[1] - col[2];
        if (cheb_new && g->cheb && g->order_boundary == ob_new) {
            for (int k = 0; k < g->n_r; k++) { re[k] = creal(col[k]); im[k] = cimag(col[k]); }
            if (l_phys) { ck_costf(g->n_r, re); ck_costf(g->n_r, im); }
            for (int k = g->n_r; k < N; k++) re[k] = im[k] = 0.0;
            if (l_phys) { ck_costf(N, re); ck_costf(N, im); }
            double fac = sqrt(2.0 / (double)(g->n_r - 1)) / h->rnorm;
            for (int k = 0; k < N; k++) d[k] = fac * (re[k] + I * im[k]);
        } else {
            for (int k = 0; k < N; k++) {
                int j = 0;
                for (int q = 1; q < g->n_r; q++)
                    if (fabs(g->r[q] - h->r[k]) < fabs(g->r[j] - h->r[k])) j = q;
                if (j < 2) j = 2;
                if (j == g->n_r - 1) j = g->n_r - 2;
                d[k] = ck_cubic(g->r + j - 2, col + j - 2, h->r[k]);
            }
        }
    }
    free(col);
    free(re);
    free(im);
}

static int ck_read(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n ? 0 : -1; }

/* read_checkpoint (checkpoints.f90:185-707) into this run; or_start then
 * does the restart branch of get_start_fields */
int or_read_checkpoint(or_handle *h, const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int rc = -2, version = 0, l_coll_old = 1, nexp_old = 2, nimp_old = 2, nold_old = 2, fam_dirk = 0;
    int itmp[3], lh = 1, lc = 0, lp = 0;
    double time = 0, dt_old[16] = {0}, par[7] = {0};
    char fam[11] = {0}, ver[73] = {0};
    ck_grid g = {0};
    cplx *buf = NULL, *tmp = NULL;
    if (ck_read(f, &version, 4)) goto out;
    if (version == 1) {
        double d2[3];
        if (ck_read(f, d2, 24)) goto out;
        time = d2[0]; dt_old[1] = d2[1]; dt_old[0] = d2[2];
        nimp_old = nold_old = nexp_old = 2;
    } else if (version == 2) {
        if (ck_read(f, &time, 8) || ck_read(f, itmp, 8)) goto out;
        nexp_old = itmp[0]; nold_old = itmp[1]; nimp_old = nold_old;
        if (nexp_old > 16 || ck_read(f, dt_old, 8 * (size_t)nexp_old)) goto out;
    } else if (version == 3) {
        if (ck_read(f, &time, 8) || ck_read(f, &l_coll_old, 4) || ck_read(f, itmp, 8)) goto out;
        nexp_old = itmp[0]; nold_old = itmp[1]; nimp_old = nold_old;
        if (nexp_old > 16 || ck_read(f, dt_old, 8 * (size_t)nexp_old)) goto out;
    } else if (version == 4) {
        if (ck_read(f, &time, 8) || ck_read(f, &l_coll_old, 4) || ck_read(f, itmp, 12)) goto out;
        nexp_old = itmp[0]; nimp_old = itmp[1]; nold_old = itmp[2];
        if (nexp_old > 16 || ck_read(f, dt_old, 8 * (size_t)nexp_old)) goto out;
    } else if (version >= 5 && version <= 7) {
        if (ck_read(f, &time, 8) || ck_read(f, &l_coll_old, 4) || ck_read(f, fam, 10) || ck_read(f, itmp, 12)) goto out;
        nexp_old = itmp[0]; nimp_old = itmp[1]; nold_old = itmp[2];
        fam_dirk = strncmp(fam, "DIRK", 4) == 0;
        if (fam_dirk) { if (ck_read(f, dt_old, 8)) goto out; }
        else if (nexp_old > 16 || ck_read(f, dt_old, 8 * (size_t)nexp_old)) goto out;
    } else {
        goto out;
    }
    /* the dt array of the new scheme (checkpoints.f90:278-282, 357-366) */
    {
        tscheme *t = &h->ts;
        int nnew = t->family == 0 ? t->nexp : 1;
        int nold_arr = fam_dirk ? (nnew > 1 ? nnew : 1) : (nexp_old > nnew ? nexp_old : nnew);
        if (fam_dirk) for (int o = 1; o < nold_arr; o++) dt_old[o] = dt_old[0];
        else for (int o = nexp_old - 1; o < nnew; o++) dt_old[o] = dt_old[nexp_old - 1];
        for (int o = 0; o < 8; o++) h->rst_dt[o] = o < nold_arr ? dt_old[o] : dt_old[nold_arr - 1];
        if (t->family == 0) for (int o = nnew; o < 8; o++) h->rst_dt[o] = h->rst_dt[nnew - 1];
    }
    if (ck_read(f, par, version >= 7 ? 56 : 48)) goto out;
    if (version < 7) { par[6] = par[5]; par[5] = 0.0; } /* no stef */
    if (ck_read(f, itmp, 12)) goto out;
    g.n_r = itmp[0]; g.m_max = itmp[1]; g.minc = itmp[2];
    if (g.n_r < 4 || g.n_r > 100000 || g.m_max < 0 || g.minc < 1) goto out;
    g.n_m = g.m_max / g.minc + 1;
    {
        int nn[2];
        double al[2];
        if (ck_read(f, ver, 72) || ck_read(f, nn, 8) || ck_read(f, al, 16)) goto out;
        g.cheb = strncmp(ver, "cheb", 4) == 0;
        g.order_boundary = nn[1];
    }
    g.r = malloc(sizeof(double) * g.n_r);
    if (ck_read(f, g.r, 8 * (size_t)g.n_r) || ck_read(f, &lh, 4) || ck_read(f, &lc, 4)) goto out;
    if (version >= 7 && ck_read(f, &lp, 4)) goto out;
    if (version <= 5) { nimp_old--; nold_old--; }
    {
        /* restart flags (checkpoints.f90:398-434) */
        tscheme *t = &h->ts;
        double ek_old = par[4];
        h->l_ab1 = ek_old != h->ek;
        if (t->family == 0) {
            if (fam_dirk) h->l_bridge = 1;
            else h->l_bridge = (t->nold > nold_old || t->nimp > nimp_old);
            if ((l_coll_old != 0) != (h->c.l_cheb_coll != 0)) { h->l_ab1 = 1; h->l_bridge = 1; }
            if (version < 4 && nold_old > 2) { h->l_ab1 = 1; h->l_bridge = 1; }
        } else {
            h->l_bridge = 0;
        }
    }
    {
        size_t Fo = (size_t)g.n_m * g.n_r, F = (size_t)h->nm * h->N;
        buf = malloc(sizeof(cplx) * Fo);
        tmp = malloc(sizeof(cplx) * F);
        tscheme *t = &h->ts;
#define CK_FIELD(dst, lBc, phys)                                               \
        do {                                                                   \
            if (ck_read(f, buf, sizeof(cplx) * Fo)) goto out;                  \
            cplx *dd_ = (dst);                                                 \
            ck_map(h, &g, buf, tmp, lBc, phys);                                \
            if (dd_) memcpy(dd_, tmp, sizeof(cplx) * F);                       \
        } while (0)
#define CK_HIST(a)                                                                                 \
        do {                                                                                       \
            if (!fam_dirk) {                                                                       \
                for (int o = 2; o <= nexp_old; o++) CK_FIELD(o <= t->nexp && t->family == 0 ? (a)->expl + (o - 1) * F : NULL, 1, l_coll_old); \
                for (int o = 2; o <= nimp_old; o++) CK_FIELD(o <= t->nimp && t->family == 0 ? (a)->impl + (o - 1) * F : NULL, 1, l_coll_old); \
                if (version > 3)                                                                   \
                    for (int o = 2; o <= nold_old; o++) CK_FIELD(o <= t->nold && t->family == 0 ? (a)->old + (o - 1) * F : NULL, 0, l_coll_old); \
            }                                                                                      \
        } while (0)
        CK_FIELD(h->us, 0, 1);
        CK_FIELD(h->up, 0, 1);
        CK_HIST(&h->dpsidt);
        if (lh) { CK_FIELD(h->temp, 0, 1); CK_HIST(&h->dTdt); }
        if (lc) { CK_FIELD(h->xi, 0, 1); CK_HIST(&h->dxidt); }
#undef CK_HIST
#undef CK_FIELD
    }
    h->rst_time = time;
    h->restart = 1;
    rc = 0;
out:
    fclose(f);
    free(g.r);
    free(buf);
    free(tmp);
    return rc;
}

/* write_checkpoint_mloc (checkpoints.f90:33-183), version 7 */
int or_write_checkpoint(or_handle *h, const char *path)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    tscheme *t = &h->ts;
    int N = h->N, i4;
    size_t F = (size_t)h->nm * N;
    i4 = 7; fwrite(&i4, 4, 1, f);
    fwrite(&h->time, 8, 1, f);
    i4 = h->c.l_cheb_coll ? 1 : 0; fwrite(&i4, 4, 1, f);
    fwrite(t->family == 0 ? "MULTISTEP " : "DIRK      ", 1, 10, f);
    int nen[3] = {t->nexp, t->nimp, t->nold};
    fwrite(nen, 4, 3, f);
    fwrite(t->dt, 8, t->family == 0 ? (size_t)t->nexp : 1, f);
    double par[7] = {h->c.ra, h->c.pr, h->c.raxi, h->c.sc, h->ek, 0.0, h->c.radratio};
    fwrite(par, 8, 7, f);
    int dims[3] = {N, h->mmax, h->minc};
    fwrite(dims, 4, 3, f);
    char ver[72];
    memset(ver, ' ', 72);
    memcpy(ver, h->fd ? "fd" : "cheb", h->fd ? 2 : 4);
    fwrite(ver, 1, 72, f);
    int nn[2] = {h->fd ? N : h->ncheb, h->fd ? 2 : 0};
    fwrite(nn, 4, 2, f);
    double al[2] = {0.8, 0.0}; /* alph1, alph2 defaults (namelists.f90:556-557) */
    fwrite(al, 8, 2, f);
    fwrite(h->r, 8, N, f);
    int fl[3] = {h->l_heat, h->l_chem, 0};
    fwrite(fl, 4, 3, f);
    cplx *out = malloc(sizeof(cplx) * F);
#define CK_PUT(src)                                                                 \
    do {                                                                            \
        for (int i = 0; i < h->nm; i++)                                             \
            for (int k = 0; k < N; k++) out[(size_t)k * h->nm + i] = (src)[(size_t)i * N + k]; \
        fwrite(out, sizeof(cplx), F, f);                                            \
    } while (0)
#define CK_PUT_HIST(a)                                                              \
    do {                                                                            \
        if (t->family == 0) {                                                       \
            for (int o = 1; o < t->nexp; o++) CK_PUT((a)->expl + o * F);            \
            for (int o = 1; o < t->nimp; o++) CK_PUT((a)->impl + o * F);            \
            for (int o = 1; o < t->nold; o++) CK_PUT((a)->old + o * F);             \
        }                                                                           \
    } while (0)
    CK_PUT(h->us);
    CK_PUT(h->up);
    CK_PUT_HIST(&h->dpsidt);
    if (h->l_heat) { CK_PUT(h->temp); CK_PUT_HIST(&h->dTdt); }
    if (h->l_chem) { CK_PUT(h->xi); CK_PUT_HIST(&h->dxidt); }
#undef CK_PUT_HIST
#undef CK_PUT
    free(out);
    int bad = ferror(f);
    fclose(f);
    return bad ? -2 : 0;
}

int or_step(or_handle *h, int nsteps)
{
    if (!h->started) return -5;
    for (int s = 0; s < nsteps; s++) {
        int rc = one_step(h);
        if (rc) return rc;
    }
    return 0;
}

int or_time(or_handle *h, double *time, double *dt)
{
    *time = h->time;
    *dt = h->ts.dt[0];
    return 0;
}

/* rInt_R, Chebyshev branch (integration.f90:18-58) */
static double rint_r(or_handle *h, const double *f)
{
    if (h->fd) return rint_fd(h, f);
    int N = h->N;
    double *f2 = dalloc(N);
    for (int n = 0; n < N; n++) f2[n] = f[n] / h->drx[n];
    costf_real(h, f2);
    f2[0] = 0.5 * f2[0];
    f2[N - 1] = 0.5 * f2[N - 1];
    double s = f2[0];
    for (int nc = 3; nc <= N; nc += 2) s = s - 1.0 / (double)(nc * (nc - 2)) * f2[nc - 1];
    s = 2.0 * sqrt(2.0 / (double)(N - 1)) * s;
    free(f2);
    return s;
}

static double round_off(double v) { return fabs(v) < 10.0 * 2.220446049250313e-16 ? 0.0 : v; }

/* test hooks of the assembly stage: robin_bc on one complex column (f
 * interleaved re, im, n_r values), and psi on the grid from omega through
 * the assembly matrix of mode index mi; radial rows r, or1, or2, beta, dbeta */
int or_robin(or_handle *h, double at, double bt, const double *rtop, double ab, double bb, const double *rbot,
             double *f)
{
    if (!h->coll) return -1;
    robin_bc(h, at, bt, rtop[0] + I * rtop[1], ab, bb, rbot[0] + I * rbot[1], (cplx *)f);
    return 0;
}

int or_ass_solve(or_handle *h, int mi, const double *omega, double *psi)
{
    int N = h->N;
    if (!h->coll || mi <= 0 || mi >= h->nm) return -1;
    if (!h->lAss && coll_ass_mat(h)) return -2;
    cplx *rhs = calloc(N, sizeof(cplx));
    const cplx *w = (const cplx *)omega;
    for (int k = 2; k < N - 2; k++) rhs[k] = w[k];
    getrs_c(N, h->cAss + (size_t)mi * N * N, h->cAssPiv + (size_t)mi * N, rhs);
    for (int k = h->ncheb; k < N; k++) rhs[k] = 0.0;
    costf_vec(h, rhs);
    memcpy(psi, rhs, sizeof(cplx) * N);
    free(rhs);
    return 0;
}

int or_fd_ell_solve(or_handle *h, int mi, const double *omega, double *psi)
{
    int N = h->N;
    if (!h->fd || mi <= 0 || mi >= h->nm) return -1;
    cplx *x = calloc(N + 2, sizeof(cplx));
    memcpy(x + 1, omega, sizeof(cplx) * N);
    fd_ell_solve(h, h->idx2m[mi], x);
    memcpy(psi, x + 1, sizeof(cplx) * N);
    free(x);
    return 0;
}

/* OpenMP threads of the oracle's loops: n > 0 sets them; returns the
 * number in use (1 without OpenMP) */
int or_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int or_cond(or_handle *h, double *out)
{
    int N = h->N;
    for (int n = 0; n < N; n++) {
        out[n] = h->l_heat ? h->tcond[n] : 0.0;
        out[N + n] = h->l_heat ? h->dtcond[n] : 0.0;
        out[2 * N + n] = h->l_chem ? h->xicond[n] : 0.0;
        out[3 * N + n] = h->l_chem ? h->dxicond[n] : 0.0;
    }
    return 0;
}

int or_radial(or_handle *h, double *out)
{
    int N = h->N;
    for (int k = 0; k < N; k++) {
        out[k] = h->r[k];
        out[N + k] = h->or1[k];
        out[2 * N + k] = h->or2[k];
        out[3 * N + k] = h->beta[k];
        out[4 * N + k] = h->dbeta[k];
    }
    return 0;
}

/* classical Nusselt numbers (outputs.f90:729-755) */
int or_nusselt(or_handle *h, double out[2])
{
    int N = h->N;
    if (!h->l_heat) { out[0] = out[1] = 1.0; return 0; }
    cplx *d = calloc(N, sizeof(cplx));
    if (h->fd) memcpy(d, h->dtemp, sizeof(cplx) * N);
    else get_dr_vec(h, h->temp, d, 1);
    out[0] = round_off(fabs(h->dtcond[0]) > 10.0 * 2.220446049250313e-16 ? 1.0 + creal(d[0]) / h->dtcond[0] : 1.0);
    out[1] = round_off(fabs(h->dtcond[N - 1]) > 10.0 * 2.220446049250313e-16 ? 1.0 + creal(d[N - 1]) / h->dtcond[N - 1] : 1.0);
    free(d);
    return 0;
}

/* 3-D kinetic energies (outputs.f90:458-571) */
int or_ekin_3d(or_handle *h, double out[4])
{
    int N = h->N;
    double pi = PI_F;
    double *us2 = dalloc(N), *up2 = dalloc(N), *upa = dalloc(N), *tmp = dalloc(N);
    for (int n = 0; n < N; n++)
        for (int i = 0; i < h->nm; i++) {
            int m = h->idx2m[i];
            us2[n] += cc2real(h->us[(size_t)i * N + n], m);
            up2[n] += cc2real(h->up[(size_t)i * N + n], m);
            if (m == 0) upa[n] += cc2real(h->up[(size_t)i * N + n], m);
        }
    for (int n = 0; n < N; n++) tmp[n] = us2[n] * h->r[n] * h->height[n];
    out[0] = round_off(pi * rint_r(h, tmp));
    for (int n = 0; n < N; n++) tmp[n] = up2[n] * h->r[n] * h->height[n];
    out[1] = round_off(pi * rint_r(h, tmp));
    for (int n = 0; n < N; n++) tmp[n] = us2[n] * h->beta[n] * h->beta[n] * h->r[n] * pow(h->height[n], 3) / 12.0;
    out[2] = round_off(pi * rint_r(h, tmp));
    for (int n = 0; n < N; n++) tmp[n] = upa[n] * h->r[n] * h->height[n];
    out[3] = round_off(pi * rint_r(h, tmp));
    free(us2); free(up2); free(upa); free(tmp);
    return 0;
}

/* simps (integration.f90:102-150) on abscissae x */
static double simps_x(int n, const double *f, const double *x)
{
    double s = 0.0;
#define SX(k)                                                                                                \
    do {                                                                                                     \
        double h2 = x[(k) + 1] - x[(k)], h1 = x[(k)] - x[(k) - 1];                                           \
        s += (h1 + h2) / 6.0 *                                                                               \
             (f[(k) - 1] * (2.0 * h1 - h2) / h1 + f[(k)] * (h1 + h2) * (h1 + h2) / (h1 * h2) + f[(k) + 1] * (2.0 * h2 - h1) / h2); \
    } while (0)
    if (n % 2 == 1) {
        for (int k = 1; k < n - 1; k += 2) SX(k);
        return -s;
    }
    s = 0.5 * (x[1] - x[0]) * (f[1] + f[0]);
    for (int k = 2; k < n - 1; k += 2) SX(k);
    s += 0.5 * (x[n - 1] - x[n - 2]) * (f[n - 1] + f[n - 2]);
    for (int k = 1; k < n - 1; k += 2) SX(k);
#undef SX
    return -0.5 * s;
}

/* get_time_series (outputs.f90:418-830) with get_lengthscales (:897-962)
 * and calculate_spectra (spectra.f90:104-197): out = [time, e_kin_2D(3),
 * e_kin_3D(4), power_2D(5), power_3D(5), reynolds_2D(3), reynolds_3D(3),
 * length_scales(5), heat(7), chem(6)] (42 values); spec = us2_m, up2_m,
 * enst_m, temp2_m, xi2_m (5 n_m values, may be NULL) */
int or_time_series(or_handle *h, double *out, double *spec)
{
    int N = h->N, nm = h->nm;
    size_t F = (size_t)nm * N;
    double pi = PI_F, tiny = 10.0 * 2.220446049250313e-16;
    cplx *dT = calloc(F, sizeof(cplx)), *dX = calloc(F, sizeof(cplx));
    /* radial derivatives as the reference takes them */
    for (int i = 0; i < nm; i++) {
        const cplx *t = h->temp + (size_t)i * N, *x = h->xi + (size_t)i * N;
        if (h->fd) {
            for (int f = 0; f < 2; f++) {
                if ((f == 0 && !h->l_heat) || (f == 1 && !h->l_chem)) continue;
                const cplx *c = f == 0 ? t : x;
                int ktop = f == 0 ? h->c.ktopt : h->c.ktopxi, kbot = f == 0 ? h->c.kbott : h->c.kbotxi;
                cplx *d = (f == 0 ? dT : dX) + (size_t)i * N;
                const cplx *qt, *qb;
                fd_patterns(h, f == 0, &qt, &qb);
                cplx gt = qt ? 2.0 * (h->r[0] - h->r[1]) * qt[i] : 0.0;
                cplx gb = qb ? 2.0 * (h->r[N - 2] - h->r[N - 1]) * qb[i] : 0.0;
                for (int k = 0; k < N; k++) {
                    cplx fm = k > 0 ? c[k - 1] : (ktop == 1 ? 2.0 * c[0] - c[1] : c[1] + gt);
                    cplx fp = k < N - 1 ? c[k + 1] : (kbot == 1 ? 2.0 * c[N - 1] - c[N - 2] : c[N - 2] - gb);
                    const double *a = h->fdr + 3 * k;
                    d[k] = a[0] * fm + a[1] * c[k] + a[2] * fp;
                }
            }
        } else {
            if (h->l_heat) get_dr_vec(h, t, dT + (size_t)i * N, 1);
            if (h->l_chem) get_dr_vec(h, x, dX + (size_t)i * N, 1);
        }
    }
    double *P = calloc(9 * (size_t)N, sizeof(double));
    double *us2 = P, *up2 = P + N, *upa = P + 2 * N, *en = P + 3 * N, *buo = P + 4 * N, *chp = P + 5 * N;
    double *nuv = P + 6 * N, *shv = P + 7 * N, *flux = P + 8 * N;
    for (int n = 0; n < N; n++) {
        for (int i = 0; i < nm; i++) {
            int m = h->idx2m[i];
            size_t q = (size_t)i * N + n;
            us2[n] += cc2real(h->us[q], m);
            up2[n] += cc2real(h->up[q], m);
            en[n] += cc2real(h->om[q], m);
            if (h->l_heat) {
                buo[n] += cc22real(h->us[q], h->temp[q], m);
                nuv[n] += cc2real(dT[q], m) + (double)m * (double)m * h->or2[n] * cc2real(h->temp[q], m);
                flux[n] += cc22real(h->us[q], h->temp[q], m);
            }
            if (h->l_chem) {
                chp[n] += cc22real(h->us[q], h->xi[q], m);
                shv[n] += cc2real(dX[q], m) + (double)m * (double)m * h->or2[n] * cc2real(h->xi[q], m);
            }
            if (m == 0) {
                upa[n] += cc2real(h->up[q], m);
                if (h->l_heat) flux[n] += -h->TdiffFac * creal(dT[q]) - h->TdiffFac * h->dtcond[n];
            }
        }
        if (h->l_heat) nuv[n] *= h->r[n] * h->height[n];
        if (h->l_chem) shv[n] *= h->r[n] * h->height[n];
    }
    double *tmp = dalloc(N);
#define INTEG(expr) (({ for (int k = 0; k < N; k++) tmp[k] = (expr); }), rint_r(h, tmp))
    const double *r = h->r, *hg = h->height;
    double us2_2D = round_off(pi * INTEG(us2[k] * r[k])), up2_2D = round_off(pi * INTEG(up2[k] * r[k]));
    double upa_2D = round_off(pi * INTEG(upa[k] * r[k]));
    double us2_3D = round_off(pi * INTEG(us2[k] * r[k] * hg[k])), up2_3D = round_off(pi * INTEG(up2[k] * r[k] * hg[k]));
    double uz2_3D = round_off(pi * INTEG(us2[k] * h->beta[k] * h->beta[k] * r[k] * pow(hg[k], 3) / 12.0));
    double upa_3D = round_off(pi * INTEG(upa[k] * r[k] * hg[k]));
    double surf = pi * (h->r_cmb * h->r_cmb - h->r_icb * h->r_icb);
    double vol_otc = 4.0 / 3.0 * pi * pow(h->r_cmb * h->r_cmb - h->r_icb * h->r_icb, 1.5);
    double visc_2D = round_off(2.0 * pi * INTEG(en[k] * r[k]));
    if (h->c.kbotv == 1) visc_2D += 4.0 * pi * up2[N - 1];
    if (h->c.ktopv == 1) visc_2D -= 4.0 * pi * up2[0];
    double pow_2D = h->l_heat ? round_off(2.0 * pi * INTEG(h->BuoFac * buo[k] * h->rgrav[k] * r[k])) : 0.0;
    double chem_2D = h->l_chem ? round_off(2.0 * pi * INTEG(h->ChemFac * chp[k] * h->rgrav[k] * r[k])) : 0.0;
    double visc_3D = round_off(2.0 * pi * INTEG(en[k] * r[k] * hg[k]));
    if (h->c.kbotv == 1) visc_3D -= 4.0 * h->or1[N - 1] * up2[N - 1];
    if (h->c.ktopv == 1) visc_3D -= 4.0 * h->or1[0] * up2[0];
    double pow_3D = h->l_heat ? round_off(2.0 * pi * INTEG(h->BuoFac * buo[k] * h->rgrav[k] * r[k] * hg[k])) : 0.0;
    double chem_3D = h->l_chem ? round_off(2.0 * pi * INTEG(h->ChemFac * chp[k] * h->rgrav[k] * r[k] * hg[k])) : 0.0;
    double ptot_2D = round_off(2.0 * pi * INTEG(h->CorFac * (us2[k] + up2[k]) * h->ekpump[k] * r[k]));
    double pzon_2D = round_off(2.0 * pi * INTEG(h->CorFac * upa[k] * h->ekpump[k] * r[k]));
    double ptot_3D = round_off(2.0 * pi * INTEG(h->CorFac * (us2[k] + up2[k]) * h->ekpump[k] * hg[k] * r[k]));
    double pzon_3D = round_off(2.0 * pi * INTEG(h->CorFac * upa[k] * h->ekpump[k] * hg[k] * r[k]));
    /* m-spectra */
    double *S = calloc(5 * (size_t)nm, sizeof(double));
    for (int i = 0; i < nm; i++) {
        int m = h->idx2m[i];
        for (int j = 0; j < 5; j++) {
            for (int k = 0; k < N; k++) {
                size_t q = (size_t)i * N + k;
                cplx v = j == 0 ? h->us[q] : j == 1 ? h->up[q] : j == 2 ? h->om[q] : j == 3 ? (h->l_heat ? h->temp[q] : 0) : (h->l_chem ? h->xi[q] : 0);
                tmp[k] = pi * cc2real(v, m) * r[k] * hg[k];
            }
            S[(size_t)j * nm + i] = rint_r(h, tmp);
        }
        if (h->l_heat) S[4 * (size_t)nm + i] = S[3 * (size_t)nm + i]; /* spectra.f90:188 */
    }
    double L[5] = {0};
    {
        int bu = -1, be = -1, bv = -1;
        for (int i = 0; i < nm; i++) {
            if (h->idx2m[i] == 0) continue;
            if (bu < 0 || S[i] > S[bu]) bu = i;
            if (be < 0 || S[i] + S[nm + i] > S[be] + S[nm + be]) be = i;
            if (bv < 0 || S[2 * nm + i] > S[2 * nm + bv]) bv = i;
        }
        int lu = bu < 0 ? 0 : h->idx2m[bu], le = be < 0 ? 0 : h->idx2m[be], lv = bv < 0 ? 0 : h->idx2m[bv];
        L[0] = lu > 0 ? pi / lu : 0.0;
        L[1] = le > 0 ? pi / le : 0.0;
        L[2] = lv > 0 ? pi / lv : 0.0;
        double E = 0, Em = 0;
        for (int i = 0; i < nm; i++) { E += S[i] + S[nm + i]; Em += (double)h->idx2m[i] * (S[i] + S[nm + i]); }
        L[3] = fabs(E) > tiny ? pi * E / Em : 0.0;
        /* outputs.f90:697-709: the 2D dissipation length of a non-rotating run */
        if (h->c.l_non_rot)
            L[4] = fabs(visc_2D) > tiny ? sqrt(2.0 * (us2_2D + up2_2D) * h->ViscFac / visc_2D) : 0.0;
        else
            L[4] = fabs(visc_3D) > tiny ? sqrt(2.0 * (us2_3D + up2_3D) * h->ViscFac / visc_3D) : 0.0;
    }
    double heat[7] = {0}, chem[6] = {0};
    int mid = N / 2 - 1;
    if (h->l_heat) {
        heat[4] = round_off(creal(h->temp[0]) + h->tcond[0]);
        heat[5] = round_off(creal(h->temp[N - 1]) + h->tcond[N - 1]);
        heat[0] = round_off(fabs(h->dtcond[0]) > tiny ? 1.0 + creal(dT[0]) / h->dtcond[0] : 1.0);
        heat[1] = round_off(fabs(h->dtcond[N - 1]) > tiny ? 1.0 + creal(dT[N - 1]) / h->dtcond[N - 1] : 1.0);
        double nv = rint_r(h, nuv);
        double nc = INTEG(h->dtcond[k] * h->dtcond[k] * hg[k] * r[k]);
        heat[2] = round_off(1.0 + nv / nc);
        double *th = dalloc(N), *f1 = dalloc(N), *f2 = dalloc(N);
        for (int k = 0; k < N; k++) {
            double sr = r[k] / h->r_cmb;
            f1[k] = -h->TdiffFac * h->dtcond[k] * sr * sr;
            th[k] = asin(sr);
            f2[k] = flux[k] * sr * sr;
        }
        if (h->c.l_non_rot) {
            /* outputs.f90:763-767 */
            double fac = INTEG(-h->TdiffFac * h->dtcond[k] * r[k]);
            heat[3] = round_off(INTEG(flux[k] * r[k] / fac));
        } else {
            heat[3] = round_off(simps_x(N, f2, th) / simps_x(N, f1, th));
        }
        heat[6] = round_off(h->dtcond[mid] + creal(dT[mid]));
        free(th); free(f1); free(f2);
    }
    if (h->l_chem) {
        chem[3] = round_off(creal(h->xi[0]) + h->xicond[0]);
        chem[4] = round_off(creal(h->xi[N - 1]) + h->xicond[N - 1]);
        chem[0] = round_off(1.0 + creal(dX[0]) / h->dxicond[0]);
        chem[1] = round_off(1.0 + creal(dX[N - 1]) / h->dxicond[N - 1]);
        double sv = rint_r(h, shv);
        double sc = INTEG(h->dxicond[k] * h->dxicond[k] * hg[k] * r[k]);
        chem[2] = round_off(1.0 + sv / sc);
        chem[5] = round_off(h->dxicond[mid] + creal(dX[mid]));
    }
#undef INTEG
    double vals[42] = {h->time, us2_2D, up2_2D, upa_2D, us2_3D, up2_3D, uz2_3D, upa_3D,
                       pow_2D, chem_2D, visc_2D, pzon_2D, ptot_2D, pow_3D, chem_3D, visc_3D, pzon_3D, ptot_3D,
                       sqrt(2.0 * (us2_2D + up2_2D) / surf), sqrt(2.0 * upa_2D / surf),
                       sqrt(2.0 * (us2_2D + up2_2D - upa_2D) / surf),
                       sqrt(2.0 * (us2_3D + up2_3D + uz2_3D) / vol_otc), sqrt(2.0 * upa_3D / vol_otc),
                       sqrt(2.0 * (us2_3D + up2_3D + uz2_3D - upa_3D) / vol_otc),
                       L[0], L[1], L[2], L[3], L[4],
                       heat[0], heat[1], heat[2], heat[3], heat[4], heat[5], heat[6],
                       chem[0], chem[1], chem[2], chem[3], chem[4], chem[5]};
    if (h->c.l_non_rot) {
        /* a non-rotating run writes no 3D series (outputs.f90:102-121, 552, 580, 620, 685) */
        static const int z3[] = {4, 5, 6, 7, 13, 14, 15, 16, 17, 21, 22, 23};
        for (size_t q = 0; q < sizeof(z3) / sizeof(z3[0]); q++) vals[z3[q]] = 0.0;
    }
    memcpy(out, vals, sizeof(vals));
    if (spec) memcpy(spec, S, sizeof(double) * 5 * nm);
    free(S); free(P); free(tmp); free(dT); free(dX);
    return 0;
}

/* ---- unit operators ---- */
static void host_to_vecs(or_handle *h, const double *host, int n_m, cplx *v)
{
    for (int i = 0; i < n_m; i++)
        for (int k = 0; k < h->N; k++) {
            size_t s = 2 * ((size_t)k * n_m + i);
            v[(size_t)i * h->N + k] = host[s] + I * host[s + 1];
        }
}
static void vecs_to_host(or_handle *h, const cplx *v, int n_m, double *host)
{
    for (int i = 0; i < n_m; i++)
        for (int k = 0; k < h->N; k++) {
            size_t s = 2 * ((size_t)k * n_m + i);
            host[s] = creal(v[(size_t)i * h->N + k]);
            host[s + 1] = cimag(v[(size_t)i * h->N + k]);
        }
}

int or_costf(or_handle *h, double *f, int n_m)
{
    cplx *v = malloc(sizeof(cplx) * (size_t)n_m * h->N);
    host_to_vecs(h, f, n_m, v);
    for (int i = 0; i < n_m; i++) costf_vec(h, v + (size_t)i * h->N);
    vecs_to_host(h, v, n_m, f);
    free(v);
    return 0;
}

int or_get_dr(or_handle *h, const double *f, double *df, int n_m, int l_dct_in)
{
    size_t S = (size_t)n_m * h->N;
    cplx *v = malloc(sizeof(cplx) * S), *d = malloc(sizeof(cplx) * S);
    host_to_vecs(h, f, n_m, v);
    for (int i = 0; i < n_m; i++) get_dr_vec(h, v + (size_t)i * h->N, d + (size_t)i * h->N, l_dct_in);
    vecs_to_host(h, d, n_m, df);
    free(v);
    free(d);
    return 0;
}

int or_get_ddr(or_handle *h, const double *f, double *df, double *ddf, int n_m, int l_dct_in)
{
    size_t S = (size_t)n_m * h->N;
    cplx *v = malloc(sizeof(cplx) * S), *d = malloc(sizeof(cplx) * S), *dd = malloc(sizeof(cplx) * S);
    host_to_vecs(h, f, n_m, v);
    for (int i = 0; i < n_m; i++)
        get_ddr_vec(h, v + (size_t)i * h->N, d + (size_t)i * h->N, dd + (size_t)i * h->N, l_dct_in);
    vecs_to_host(h, d, n_m, df);
    vecs_to_host(h, dd, n_m, ddf);
    free(v);
    free(d);
    free(dd);
    return 0;
}

/* grid (n_phi, n_rad) real, phi fastest <-> spec (n_m, n_rad) complex, m fastest */
int or_fft(or_handle *h, const double *grid, double *spec, int n_rad)
{
    cplx *s = malloc(sizeof(cplx) * h->nm);
    for (int r = 0; r < n_rad; r++) {
        phi_fft(h, grid + (size_t)r * h->nphi, s, h->mmax);
        for (int i = 0; i < h->nm; i++) {
            spec[2 * ((size_t)r * h->nm + i)] = creal(s[i]);
            spec[2 * ((size_t)r * h->nm + i) + 1] = cimag(s[i]);
        }
    }
    free(s);
    return 0;
}

int or_ifft(or_handle *h, const double *spec, double *grid, int n_rad)
{
    cplx *s = malloc(sizeof(cplx) * h->nm);
    for (int r = 0; r < n_rad; r++) {
        for (int i = 0; i < h->nm; i++)
            s[i] = spec[2 * ((size_t)r * h->nm + i)] + I * spec[2 * ((size_t)r * h->nm + i) + 1];
        phi_ifft(h, s, grid + (size_t)r * h->nphi, h->mmax);
    }
    free(s);
    return 0;
}

int or_grid(or_handle *h, double *r, double *x)
{
    memcpy(r, h->r, sizeof(double) * h->N);
    memcpy(x, h->x, sizeof(double) * h->N);
    return 0;
}

int or_stencil(or_handle *h, int which, int m, int idx, int ku, double *out)
{
    if (!h->tabs) h->tabs = build_optabs(h);
    optabs *T = h->tabs;
    const exttab *tabs[4];
    ld coef[4];
    int nt = 0;
    ld M2 = (ld)m * m;
    switch (which) {
    case 0: tabs[0] = &T->I2R2; coef[0] = 1; nt = 1; break;
    case 1: tabs[0] = &T->I2R1; coef[0] = 1; nt = 1; break;
    case 2: tabs[0] = &T->L2_0; tabs[1] = &T->L2_2; coef[0] = 1; coef[1] = M2; nt = 2; break;
    case 3: tabs[0] = &T->R2; coef[0] = 1; nt = 1; break;
    case 4: tabs[0] = &T->I1R1; coef[0] = 1; nt = 1; break;
    case 5: tabs[0] = &T->I4R4; coef[0] = 1; nt = 1; break;
    case 6: tabs[0] = &T->I4R3; coef[0] = 1; nt = 1; break;
    case 7: tabs[0] = &T->LR_0; tabs[1] = &T->LR_2; coef[0] = 1; coef[1] = M2; nt = 2; break;
    case 8: tabs[0] = &T->LR2_0; tabs[1] = &T->LR2_2; tabs[2] = &T->LR2_4; coef[0] = 1; coef[1] = M2; coef[2] = M2 * M2; nt = 3; break;
    default: return -1;
    }
    for (int t = 0; t < nt; t++)
        if (tabs[t]->K > ku) return -2;
    ld *st = malloc(sizeof(ld) * (2 * ku + 1));
    comb_stencil(idx, ku, nt, tabs, coef, st);
    for (int i = 0; i < 2 * ku + 1; i++) out[i] = (double)st[i];
    free(st);
    return 0;
}

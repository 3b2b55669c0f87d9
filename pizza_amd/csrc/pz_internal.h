// Internal device-side data structures of libpizza_amd (not part of the ABI).
//
// Layout of every spectral field on the device: complex FP64 stored as
// double2 (x = re, y = im), "M-layout" f[m][r]: one contiguous radial column
// of n_r values per azimuthal wavenumber.  Time-scheme slots of a field are
// consecutive [slot][m][r] blocks.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pizza_amd.h"

typedef double2 cd;

#define PZ_MAX_RADICES 24
#define PZ_MAX_STAGES 8

// Time-scheme state that changes during the run.  It lives in device memory
// and is updated by k_courant, so a captured step replays without the host.
struct PzScheme {
    double dt[8];
    double wimp[8];
    double wimp_lin[8];
    double wexp[8];
    double bimp[PZ_MAX_STAGES][PZ_MAX_STAGES]; // dt-scaled Butcher tables
    double bexp[PZ_MAX_STAGES][PZ_MAX_STAGES];
    double bass_imp[PZ_MAX_STAGES], bass_exp[PZ_MAX_STAGES]; // dt-scaled assembly weights
    double time;
    double dt_r, dt_h; // last Courant limits
    int lMat;          // matrices must be rebuilt in this step
    int lMatNext;
    int l_new_dt;
    int error;         // 1: dt fell below dtMin
    int n_steps;
    int ab1;  // the first step after a restart uses first-order Adams-Bashforth
    int pad[2];
};

struct PzDev;
// A batch of independent runs advanced by one launch per kernel: block b of
// a batched launch does the work of block map[b].y of run map[b].x's own
// (flattened) grid.
struct PzBatch {
    const PzDev *const *runs;
    const int2 *map;
    int nruns;
};

// Everything a kernel needs: sizes, constants and device pointers.
struct PzDev {
    int N, nm, nphi, ncheb, L; // L = 2N-2 (Chebyshev even extension)
    int family, nstages, nexp, nimp, nold;
    int scheme, bridge; // enum pz_time_scheme; 1: multistep start-up steps use CNAB2
    int l_heat, l_chem, l_coriolis_imp, l_buo_imp, l_ek_pump;
    int ktopt, kbott, ktopxi, kbotxi, ktopv, kbotv;
    int nrad, radices[PZ_MAX_RADICES]; // azimuthal FFT plan
    int nradL, radicesL[PZ_MAX_RADICES]; // plan of length L (Chebyshev); 0: direct sums
    int l_exp_calc[8], l_imp_calc_rhs[8];
    double rnorm, cheb_fac, icheb_fac, boundary_fac, thr;
    double CorFac, BuoFac, TdiffFac, ChemFac, XidiffFac, ViscFac, tadvz, r_cmb, ek;
    double alpha, dtMax, dtMin, dt_fac, courfac, wimp_lin0;
    double bimp0[PZ_MAX_STAGES][PZ_MAX_STAGES], bexp0[PZ_MAX_STAGES][PZ_MAX_STAGES];
    // DIRK schemes with an assembly stage: the assembly weights, the
    // collocation boundary-derivative rows (robin_bc) and the inverses of
    // the psi assembly matrices [m][col][ldi]
    int l_assembly;
    int gj_b; // pivot columns per panel of the collocation rebuilds (8 or 16)
    double bass_imp0[PZ_MAX_STAGES], bass_exp0[PZ_MAX_STAGES];
    const double *dr_top, *dr_bot;
    double *A_inv;

    // radial functions (n_r) and transform tables
    const double *r, *or1, *or2, *rgrav, *beta, *ekpump, *ekp_up, *ekp_us, *ekp_dusdp;
    const double *tcond, *dtcond, *xicond, *dxicond, *delxr2, *delxh2, *drx, *ddrx, *x;
    const double *cosT, *sinT; // cos/sin(pi q/(N-1)), q in [0, L)
    const double *drb;         // [3][N] bulk factors of the radial derivatives
    const double *dbeta;       // d beta / dr (collocation psi operator)
    const double *radc;        // block holding all radial rows above (PZ_NRADC x n_r)
    // finite differences (radial_scheme = 'FD'): stencils
    // [dr 3N | ddr 3N | dddr 5N | ddddr 5N | dr_top 3 | dr_bot 3], and the
    // factorised systems: T and xi [m][d, low, up][N] (real), psi
    // [m][d, low1, low2, up1, up2][N] (complex), u_phi(m = 0) [d, low, up][N]
    int fd;
    const double *fdc;
    double *fT, *fX, *fU;
    cd *fP;
    const double *tau;         // tau lines: [rMat(2,N) | drMat(2,N) | d2rMat(2,N)]
    const cd *twid;            // exp(-2 pi i k / n_phi)
    const int *idx2m;
    const double *hdifT, *hdifXi, *hdifV;
    const cd *topt, *bott, *topxi, *botxi;

    // fields [m][r]
    cd *us, *up, *om, *psi, *temp, *xi, *temp_hat, *xi_hat, *psi_hat, *dtemp;
    cd *dVsT, *dVsXi, *dVsOm;
    cd *T_old, *T_impl, *T_expl, *X_old, *X_impl, *X_expl, *P_old, *P_impl, *P_expl;
    double *dtr, *dth; // per-radius Courant limits

    // m-independent RHS bands, row-major [n_r][2K+1] in column-offset order
    // (entry K+d is column i+d), already multiplied by rnorm
    const double *T_RHSE1, *T_RHSE2;           // K = 3, 4
    const double *P_RHSE1, *P_RHSE2, *P_RHSE3; // K = 4, 8, 7
    // per-m RHS bands
    const double *T_RHSI; // [nm][n_r][5]  K = 2
    const double *P_RHSI; // [nm][n_r][17] K = 8 (m = 0 uses the central 9)
    const cd *P_RHSIL;    // [nm][n_r][17]
    // per-m raw LHS stencils in the reference stencil order (index nb <->
    // column idx+K-nb), rows idx in [0, n_r)
    const double *T_S0, *T_S1, *X_S1;   // [nm][n_r][9]
    const double *P_S0, *P_S1, *P_S2;   // [nm][n_r][17]
    // factorised bordered LHS (written by k_build_lhs)
    // dinv: reciprocals of the diagonal of U of the banded block
    double *T_A1, *T_A2, *T_A3, *T_A4, *T_fac, *T_dinv;
    int *T_p1, *T_p4;
    double *X_A1, *X_A2, *X_A3, *X_A4, *X_fac, *X_dinv;
    int *X_p1, *X_p4;
    cd *P_A1, *P_A2, *P_A3, *P_A4, *P_dinv;
    double *P_fac;
    int *P_p1, *P_p4;
    int *info; // [3 * nm] factorisation info per (field, m)
    // explicit inverses of the preconditioned bordered matrices with the row
    // preconditioner folded in: x = INV * rhs.  Column-major per m,
    // [nm][n_r columns][ldi rows], ldi = n_r rounded up to even (zero pad)
    double *T_inv, *X_inv;
    cd *P_inv;
    int ldi;

    // collocation method (l_cheb_coll): T_n and its r-derivatives at the
    // grid points, [k][n] (chebyshev.f90:296-321); T/xi inverses reuse
    // T_inv/X_inv; psi: the first n_r columns of the inverse of the coupled
    // omega-psi system, [m][col][2 n_r]; u_phi(m = 0): [col][ldi]; scratch of
    // the Gauss-Jordan inversions, coll_ws bytes per m
    int coll;
    const double *cM, *cdM, *cd2M;
    cd *PC_inv;
    double *U_inv;
    char *coll_work;
    size_t coll_ws;

    // influence-matrix psi solve of the collocation method (matrix_solve =
    // 'influence', update_psi_coll_dmat.f90): per m the scaled inverses of
    // the omega and psi operators [m][col][ldi], the homogeneous solutions
    // [m][4][n_r] (omega_2, omega_3, psi_2, psi_3; Chebyshev space), the
    // 2 x 2 influence matrices [m][4], and the wall rows of the velocity
    // condition [2][n_r] (outer, inner)
    int dmat;
    double *OmI, *PsI, *dhom, *dinfl;
    const double *bcrow;
    // finite differences: 1 when the T and xi walls take the t_top/t_bot
    // patterns (l_fd_bc_pattern; the reference holds them at zero)
    int fd_pattern;
    // long azimuthal grids: u_s, u_phi, omega and T by radius, [f][n_r][n_m]
    // (k_gather_r), for k_radial_loop_inplace
    cd *rrow;

    PzScheme *ss;
};

// Carves the step state out of one block at base (base = nullptr: only
// sizes it) and returns its size in bytes: the fields, the time-scheme
// slots and the scheme state, each at a 256-byte boundary.  Every block of
// this size (snapshots, adjoint variables) shares the layout.
inline size_t arena_layout(PzDev &d, char *base, size_t F, int nold, int nimp, int nexp)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *q = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return q;
    };
    const size_t fb = F * sizeof(cd);
    cd **one[] = {&d.us, &d.up, &d.om, &d.psi, &d.temp, &d.xi, &d.temp_hat, &d.xi_hat, &d.psi_hat, &d.dtemp};
    for (cd **q : one) *q = (cd *)take(fb);
    cd **slot[] = {&d.T_old, &d.T_impl, &d.T_expl, &d.X_old, &d.X_impl, &d.X_expl, &d.P_old, &d.P_impl, &d.P_expl};
    const int cnt[] = {nold, nimp, nexp, nold, nimp, nexp, nold, nimp, nexp};
    for (int i = 0; i < 9; i++) *slot[i] = (cd *)take(fb * cnt[i]);
    d.ss = (PzScheme *)take(sizeof(PzScheme));
    return off;
}

// rows of the radial block: x, r, drx, ddrx, or1, or2, rgrav, beta, ekpump,
// ekp_up, ekp_us, ekp_dusdp, delxr2, delxh2, tcond, dtcond, xicond, dxicond,
// drb (3 rows), dbeta
#define PZ_NRADC 25
template <class D>
__host__ __device__ inline void set_radial_pointers(D &d, const double *rc, size_t N)
{
    d.x = rc;
    d.r = rc + N;
    d.drx = rc + 2 * N;
    d.ddrx = rc + 3 * N;
    d.or1 = rc + 4 * N;
    d.or2 = rc + 5 * N;
    d.rgrav = rc + 6 * N;
    d.beta = rc + 7 * N;
    d.ekpump = rc + 8 * N;
    d.ekp_up = rc + 9 * N;
    d.ekp_us = rc + 10 * N;
    d.ekp_dusdp = rc + 11 * N;
    d.delxr2 = rc + 12 * N;
    d.delxh2 = rc + 13 * N;
    d.tcond = rc + 14 * N;
    d.dtcond = rc + 15 * N;
    d.xicond = rc + 16 * N;
    d.dxicond = rc + 17 * N;
    d.drb = rc + 18 * N;
    d.dbeta = rc + 21 * N;
}

// The radial rows of one block (global memory or a kernel's LDS copy):
// one base pointer, rows addressed in the order of set_radial_pointers
// (a struct of nineteen pointers would hold 38 scalar registers).
struct RadView {
    const double *b;
    int N;
    __host__ __device__ RadView(const double *rc, int n) : b(rc), N(n) {}
    __host__ __device__ double row(int i, int k) const { return b[(size_t)i * N + k]; }
    __host__ __device__ double x(int k) const { return row(0, k); }
    __host__ __device__ double r(int k) const { return row(1, k); }
    __host__ __device__ double drx(int k) const { return row(2, k); }
    __host__ __device__ double ddrx(int k) const { return row(3, k); }
    __host__ __device__ double or1(int k) const { return row(4, k); }
    __host__ __device__ double or2(int k) const { return row(5, k); }
    __host__ __device__ double rgrav(int k) const { return row(6, k); }
    __host__ __device__ double beta(int k) const { return row(7, k); }
    __host__ __device__ double ekpump(int k) const { return row(8, k); }
    __host__ __device__ double ekp_up(int k) const { return row(9, k); }
    __host__ __device__ double ekp_us(int k) const { return row(10, k); }
    __host__ __device__ double ekp_dusdp(int k) const { return row(11, k); }
    __host__ __device__ double delxr2(int k) const { return row(12, k); }
    __host__ __device__ double delxh2(int k) const { return row(13, k); }
    __host__ __device__ double tcond(int k) const { return row(14, k); }
    __host__ __device__ double dtcond(int k) const { return row(15, k); }
    __host__ __device__ double xicond(int k) const { return row(16, k); }
    __host__ __device__ double dxicond(int k) const { return row(17, k); }
    __host__ __device__ double drb(int k) const { return row(18, k); } // k in [0, 3N)
    __host__ __device__ double dbeta(int k) const { return row(21, k); }
    // finite differences only: 1/r^3 and the higher derivatives of beta
    __host__ __device__ double or3(int k) const { return row(22, k); }
    __host__ __device__ double d2beta(int k) const { return row(23, k); }
    __host__ __device__ double d3beta(int k) const { return row(24, k); }
};

// sizes of the bordered storage per m
#define PZ_T_KL 4
#define PZ_T_KU 4
#define PZ_T_NT 2
#define PZ_T_LD (2 * PZ_T_KL + PZ_T_KU + 1) // 13
#define PZ_P_KL 8
#define PZ_P_KU 8
#define PZ_P_NT 4
#define PZ_P_LD (2 * PZ_P_KL + PZ_P_KU + 1) // 25

/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C CPU restatement of pizza's IMEX time step (integration method,
 * multistep CNAB2 and DIRK schemes), used as the checker of the HIP path.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * it; the product library never links or calls it.
 */
#ifndef PZ_ORACLE_H
#define PZ_ORACLE_H

#include "../include/pizza_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_handle or_handle;

void or_config_default(pz_config *cfg);
int or_create(const pz_config *cfg, or_handle **h);
void or_destroy(or_handle *h);
int or_dims(const or_handle *h, int *n_m_max, int *n_r_max, int *n_phi_max);
int or_set_field(or_handle *h, int field, const double *host);
int or_get_field(or_handle *h, int field, double *host);
int or_init_fields(or_handle *h, int init_t, double amp_t, int init_xi,
                   double amp_xi, int init_u, double amp_u);
int or_start(or_handle *h);
int or_step(or_handle *h, int nsteps);
int or_time(or_handle *h, double *time, double *dt);
int or_ekin_3d(or_handle *h, double out[4]);
int or_nusselt(or_handle *h, double out[2]);
int or_robin(or_handle *h, double at, double bt, const double *rtop, double ab, double bb, const double *rbot,
             double *f);
int or_ass_solve(or_handle *h, int mi, const double *omega, double *psi);
/* -L psi = omega of the finite-difference assembly stage, psi = 0 at the walls */
int or_fd_ell_solve(or_handle *h, int mi, const double *omega, double *psi);
int or_radial(or_handle *h, double *out);
/* checkpoint files of checkpoints.f90: the version-7 writer, and the
 * reader (versions 1-7) that maps the stored grid, truncation and scheme
 * onto this run; or_start afterwards performs the restart branch of
 * get_start_fields */
int or_write_checkpoint(or_handle *h, const char *path);
int or_read_checkpoint(or_handle *h, const char *path);
/* conducting profiles [tcond | dtcond | xicond | dxicond] (4 n_r values) */
int or_cond(or_handle *h, double *out);
/* the log-step time series of get_time_series (outputs.f90:418-830): 42
 * values [time, e_kin_2D(3), e_kin_3D(4), power_2D(5), power_3D(5),
 * reynolds_2D(3), reynolds_3D(3), length_scales(5), heat(7), chem(6)], and
 * the m-spectra of calculate_spectra (5 n_m values, may be NULL) */
int or_time_series(or_handle *h, double *out, double *spec);
/* OpenMP threads of the oracle: n > 0 sets them; returns the count in use */
int or_threads(int n);
/* planted mistake of the failure-demonstrating tests (0 = none) */
extern int or_plant;
int or_dirk_tables(int scheme, int *nstages, int *l_assembly, double *wimp_lin0, double *bimp, double *bexp,
                   double *bass_imp, double *bass_exp);

/* unit operators on (n_m, n_r) complex arrays, m fastest */
int or_costf(or_handle *h, double *f, int n_m);
int or_get_dr(or_handle *h, const double *f, double *df, int n_m, int l_dct_in);
int or_get_ddr(or_handle *h, const double *f, double *df, double *ddf, int n_m,
               int l_dct_in);
int or_fft(or_handle *h, const double *grid, double *spec, int n_rad);
int or_ifft(or_handle *h, const double *spec, double *grid, int n_rad);

/* time-scheme weights for the dt history dt[0..3] (dt[0] newest): family
 * (0 multistep, 1 DIRK), stage count, multistep weights (8 each) and the
 * dt-scaled Butcher tables (8 x 8, row-major); returns -1 for an unknown
 * scheme */
int or_scheme_weights(int scheme, double alpha, const double *dt, int *family, int *nstages,
                      double *wimp, double *wimp_lin, double *wexp, double *bimp, double *bexp);

/* finite-difference stencils [dr 3N | ddr 3N | dddr 5N | ddddr 5N | dr_top 3 |
 * dr_bot 3] (16 N + 6 values); -1 for a Chebyshev run */
int or_fd_stencils(or_handle *h, double *out);

/* radial grid and the Chebyshev x grid (n_r_max values each) */
int or_grid(or_handle *h, double *r, double *x);

/*
 * Folded band stencil of a sparse-Chebyshev operator for row idx (0-based)
 * with half-bandwidth ku: out[0..2ku] in the reference's stencil order
 * (out[0] = column idx+ku ... out[2ku] = column idx-ku).
 * which: 0 I2R2, 1 I2R1, 2 I2R2LAPL(m), 3 R2, 4 I1R1, 5 I4R4, 6 I4R3,
 *        7 I4R4LAPLROT(m), 8 I4R4LAPLROT2(m)
 */
int or_stencil(or_handle *h, int which, int m, int idx, int ku, double *out);

#ifdef __cplusplus
}
#endif

#endif

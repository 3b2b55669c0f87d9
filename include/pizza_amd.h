/*
 * pizza_amd — C ABI of the MI355X-native IMEX time step of pizza.
 *
 * Every entry point takes plain pointers and sizes; no C++ or torch types
 * cross this boundary.  Host arrays that hold spectral fields use the
 * reference's Fortran layout: complex(8) f(n_m_max, n_r_max), m fastest,
 * (re, im) interleaved — so a Fortran caller passes its arrays directly
 * through ISO_C_BINDING.
 *
 * Return codes: 0 = ok; > 0 = LAPACK-style "info" of a singular pivot
 * (the reference aborts with 'Singular matrix' in that case);
 * < 0 = invalid argument, unsupported configuration or HIP error.
 * pz_last_error() returns the message of the last failure of a handle.
 */
#ifndef PIZZA_AMD_H
#define PIZZA_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define PZ_N_T_BOUNDS 32 /* namelists.f90:23 (n_t_bounds) */

/* Time schemes (select_time_scheme.f90:15-49). */
enum pz_time_scheme {
    /* multistep family: initialize multistep_schemes.f90:56-104, weights :137-305 */
    PZ_CNAB2 = 1,
    PZ_CNLF = 2,
    PZ_MODCNAB = 3,
    PZ_SBDF2 = 4,  /* also BDF2AB2 */
    PZ_SBDF3 = 5,  /* also BDF3AB3 */
    PZ_SBDF4 = 6,  /* also BDF4AB4 */
    PZ_TVB33 = 7,
    /* DIRK family without an assembly stage: initialize dirk_schemes.f90:55-182,
       Butcher tables :286-525 */
    PZ_ARS222 = 101,
    PZ_ARS443 = 102,
    PZ_BPR353 = 103,
    PZ_DP1_242 = 104,
    PZ_DP2_242 = 105,
    PZ_LZ232 = 106,
    PZ_CK232 = 107,
    PZ_LZ453 = 108,
    PZ_PC2 = 109,
    /* DIRK family with an assembly stage (dirk_schemes.f90:63-118, 183-230;
       tables :299-392, 427-441, 526-711): the last stage only evaluates the
       explicit terms and the step ends with assemble_stage (mloop.f90:105-160) */
    PZ_SSP222 = 110,
    PZ_LSDIRK222 = 111,
    PZ_ARS232 = 112,
    PZ_ARS233 = 113,
    PZ_SSP322 = 114,
    PZ_SSP332 = 115,
    PZ_SSP433 = 116,
    PZ_ARS343 = 117,
    PZ_MARS343 = 118,
    PZ_KC343 = 119,
    PZ_KC564 = 120,
    PZ_KC785 = 121,
    PZ_DBM453 = 122
};

/* Field identifiers for pz_set_field / pz_get_field (fields.f90). */
enum pz_field {
    PZ_US = 0,      /* u_s        (m, r) physical r */
    PZ_UP = 1,      /* u_phi      (m, r) */
    PZ_OM = 2,      /* omega_z    (m, r) */
    PZ_PSI = 3,     /* psi        (m, r)  (m /= 0) */
    PZ_TEMP = 4,    /* T          (m, r) */
    PZ_XI = 5,      /* xi         (m, r) */
    PZ_TEMP_HAT = 6,/* Chebyshev coefficients of T (integration method) */
    PZ_XI_HAT = 7,
    PZ_PSI_HAT = 8,
    PZ_DTEMP = 9,   /* dT/dr (refreshed on log steps) */
    PZ_NFIELDS = 10
};

/*
 * Run configuration: the subset of the reference namelists
 * (namelists.f90:162-189, defaults :530-672) that the IMEX step reads.
 * pz_config_default() fills the reference defaults.
 */
typedef struct pz_config {
    /* &grid */
    int n_r_max, n_cheb_max, m_max, minc;
    /* &control */
    int time_scheme;       /* enum pz_time_scheme */
    int l_cheb_coll;       /* 1: 'colloc', 0: 'integ' (sparse Chebyshev) */
    int l_direct_solve;    /* matrix_solve='direct' */
    int l_coriolis_imp;    /* corio_term='implicit' */
    int l_buo_imp;         /* buo_term='implicit' */
    int l_non_rot;
    int time_scale_rot;    /* time_scale='rot' */
    int l_rerror_fix;
    double rerror_fac;
    double alpha;          /* Crank-Nicolson weight */
    double dtMax;
    double courfac;        /* >= 1e3 selects the scheme's own value */
    double dt_fac;
    /* &phys_param */
    double ra, ek, pr, raxi, sc, radratio, g0, g1, g2;
    int ktopt, kbott, ktopv, kbotv, ktopxi, kbotxi;
    int l_ek_pump, l_temp_advz;
    double tcond_fac, xicond_fac;
    double t_top[3 * PZ_N_T_BOUNDS], t_bot[3 * PZ_N_T_BOUNDS];
    double xi_top[3 * PZ_N_T_BOUNDS], xi_bot[3 * PZ_N_T_BOUNDS];
    /* &grid / &control: radial_scheme='FD' (second-order finite differences,
       radial.f90:86-99) on the stretched grid of finite_differences.f90:87-188 */
    double fd_stretch, fd_ratio;
    int l_finite_diff;
    /* 1: the finite-difference T and xi equations take the t_top/t_bot
       patterns as wall values (fixed value) or wall gradients (fixed flux,
       through the ghost points), as the spectral methods do.  No reference
       counterpart: the reference's finite differences hold the perturbation
       at zero on the walls (update_temp_fd.f90:85-110, its flux terms
       commented out at :142,162).  Default 0, the reference's behaviour. */
    int l_fd_bc_pattern;
} pz_config;

typedef struct pz_handle pz_handle;

/* Fills the reference namelist defaults (namelists.f90:530-672). */
void pz_config_default(pz_config *cfg);

/*
 * Creates a run on the current HIP device: truncation (truncation.f90:24-50),
 * radial functions (radial.f90:115-282), horizontal functions
 * (horizontal.f90:69-202), operator matrices and device state.
 * Replaces the init sequence of pizza.f90:63-199 for the step's state.
 */
int pz_create(const pz_config *cfg, pz_handle **h);
int pz_destroy(pz_handle *h);
const char *pz_last_error(const pz_handle *h);

/* Sizes of the spectral arrays: n_m_max, n_r_max, n_phi_max. */
int pz_dims(const pz_handle *h, int *n_m_max, int *n_r_max, int *n_phi_max);

/*
 * Host <-> device copies of one field; host = complex(8) (n_m_max, n_r_max),
 * m fastest.  Stream-ordered; pz_get_field synchronises.
 */
int pz_set_field(pz_handle *h, int field, const double *host);
int pz_get_field(pz_handle *h, int field, double *host);

/*
 * Reference deterministic initial conditions (init_fields.f90:260-360,
 * :455-509): init_t > 0 single mode, init_t == -1 Gaussian bubble; the same
 * for xi; init_u > 0 single mode, init_u < 0 axisymmetric u_phi.  Adds to the
 * fields already set (zero after pz_create).
 */
int pz_init_fields(pz_handle *h, int init_t, double amp_t, int init_xi,
                   double amp_xi, int init_u, double amp_u);

/*
 * Fresh-start post-processing of get_start_fields (init_fields.f90:96-112,
 * :186-255): dt(:) = dtMax, scheme weights, omega and dT/dr from the set
 * fields, the Chebyshev "hat" fields and the first implicit state.
 */
int pz_start(pz_handle *h);

/*
 * Advances nsteps full IMEX steps: the stage loop of time_loop
 * (step_time.f90:215-370) — radial loop, explicit assembly, Courant
 * (stage 1), m-loop solves.  Asynchronous on the handle's stream.
 */
int pz_step(pz_handle *h, int nsteps);

/*
 * The step one procedure of time_loop at a time (step_time.f90:215-370), for
 * a driver that keeps the reference's stage loop:
 *   for stage = 1..nstages (pz_nstages):
 *       pz_stage_explicit(h, stage)        radial_loop with its transposes
 *                                          (radial_loop.f90:231-394,
 *                                          step_time.f90:222-251)
 *       if stage == 1: pz_courant(h, ...)  dt_courant, dt roll, set_weights,
 *                                          rebuild flag (courant.f90:78-167,
 *                                          step_time.f90:256-304)
 *       pz_stage_implicit(h, stage, lMat)  finish_explicit_assembly and mloop
 *                                          (mloop.f90:36-103, 169-229)
 * is pz_step(h, 1) launch for launch (same kernels, same order), so the
 * results are bit for bit those of pz_step.
 * pz_stage_explicit runs the radial loop of the stage when the scheme
 * evaluates explicit terms there (l_exp_calc), else nothing.
 * pz_courant synchronises only when dtr or dth is not NULL; they receive
 * the smallest radial and horizontal Courant limits of the step.
 * pz_stage_implicit: on stage 1 it first rebuilds the matrices when lMat is
 * nonzero or dt changed (the rebuild of step_time.f90:277-304); the explicit
 * finish of the stage is fused into the per-m kernel of mloop; on the last
 * stage of a DIRK scheme with an assembly stage it runs assemble_stage
 * (mloop.f90:105-160) instead of mloop.  All three are asynchronous on the
 * handle's stream; stage must lie in 1..nstages.
 */
int pz_nstages(const pz_handle *h);
int pz_stage_explicit(pz_handle *h, int stage);
int pz_courant(pz_handle *h, double *dtr, double *dth);
int pz_stage_implicit(pz_handle *h, int stage, int lMat);

/*
 * Checkpoint files of checkpoints.f90.
 * pz_write_checkpoint writes a version-7 file of the run's state
 * (write_checkpoint_mloc, checkpoints.f90:33-183): the header (time, scheme
 * family and sizes, dt array, physical parameters, grid, radial scheme and
 * radii), then u_s, u_phi and the psi history, T and its history, xi and its
 * history (expl(2:nexp), impl(2:nimp), old(2:nold) of a multistep scheme).
 * Synchronises; needs a started run.
 * pz_read_checkpoint reads a version 1-7 file into the run
 * (read_checkpoint, checkpoints.f90:185-707): the fields and the history
 * slots the run's scheme keeps, mapped onto the run's grid (same grid: a
 * copy; Chebyshev to Chebyshev: through the coefficients; to or from finite
 * differences: cubic interpolation between the nearest stored radii), modes
 * above the stored m_max start from zero, the time and dt array of the file,
 * and the restart rules of the reference (first-order Adams-Bashforth first
 * step when ek or the radial method changed, CNAB2 start-up steps when the
 * stored history is shorter than the scheme needs).  pz_start must follow:
 * it then performs the restart branch of get_start_fields
 * (init_fields.f90:67-95), keeping the stored history.  pz_init_fields may
 * be called in between to add the namelist's perturbation, as the
 * reference does after reading.
 * pz_read_checkpoint_header reads only the header of a file (host code, no
 * device): r (r_len values, may be NULL) receives the stored radii,
 * n_arrays the number of (n_m_max, n_r_max) complex arrays that follow the
 * header_bytes of the header (phase-field arrays included).
 */
typedef struct pz_checkpoint_info {
    int version;
    double time;
    int l_cheb_coll;
    char family[11];   /* "MULTISTEP" or "DIRK" (versions >= 5) */
    int nexp, nimp, nold, n_dt;
    double dt[16];
    double ra, pr, raxi, sc, ek, stef, radratio;
    int n_r_max, m_max, minc;
    char rscheme[73];  /* "cheb" or "fd" */
    int n_max, order_boundary;
    double alph1, alph2;
    int l_heat, l_chem, l_phase_field;
    long long header_bytes;
    int n_arrays;
} pz_checkpoint_info;
int pz_write_checkpoint(pz_handle *h, const char *path);
int pz_read_checkpoint(pz_handle *h, const char *path);
int pz_read_checkpoint_header(const char *path, pz_checkpoint_info *info, double *r, int r_len);

/*
 * A frame file of one field (write_snapshot_mloc, output_frames.f90:96-167,
 * version 2): time, ra, ek, pr, radratio, sc, raxi, n_r_max, n_m_max, m_max,
 * minc, n_phi_max, r, tcond, xicond, idx2m, then the field complex(8)
 * (n_m_max, n_r_max).  Synchronises.
 */
int pz_write_frame(pz_handle *h, int field, const char *path);

/* Synchronises and returns the simulation time and current dt(1). */
int pz_time(pz_handle *h, double *time, double *dt);

/*
 * 3-D kinetic energy time-series entry of outputs.f90:552-571, 686-688:
 * out[0..3] = us2_3D, up2_3D, uz2_3D, up2_axi_3D.  Synchronises.
 */
int pz_ekin_3d(pz_handle *h, double out[4]);

/*
 * The log-step time series of get_time_series (outputs.f90:418-830), one
 * row of each file of the reference, from the current state.  Synchronises.
 *   ekin_2d     e_kin_2D: us2_2D, up2_2D, up2_axi_2D
 *   ekin_3d     e_kin_3D: us2_3D, up2_3D, uz2_3D, up2_axi_3D
 *   power_2d/3d power_2D/3D: buoyancy power, chemical power, viscous
 *               dissipation, zonal and total Ekman-pumping dissipation
 *   reynolds_2d/3d reynolds_2D/3D: Re, zonal Re, fluctuating Re (the
 *               volume outside the tangent cylinder of the sphere container)
 *   lengths     length_scales: dlus_peak, dlekin_peak, dlvort_peak,
 *               dlekin_int (get_lengthscales, outputs.f90:897-962), dl_diss
 *   heat        heat: NuTop, NuBot, Nu_vol, Nu_int, tTop, tBot, beta_t
 *   chem        chem: ShTop, ShBot, Sh_vol, xiTop, xiBot, beta_xi
 * The radial derivatives of T and xi are those of the run's method (get_dr
 * on Chebyshev grids, the ghost-zone stencil of the finite-difference step).
 */
typedef struct pz_series {
    double time;
    double ekin_2d[3], ekin_3d[4];
    double power_2d[5], power_3d[5];
    double reynolds_2d[3], reynolds_3d[3];
    double lengths[5];
    double heat[7];
    double chem[6];
} pz_series;
int pz_time_series(pz_handle *h, pz_series *out);

/*
 * The m-spectra of calculate_spectra (spectra.f90:104-197) at the current
 * state: per m (n_m_max values each, any may be NULL) the radial integrals
 * of pi r h times |u_s|^2, |u_phi|^2, |omega|^2, |T|^2 and the xi entry,
 * which the reference takes from the temperature profile
 * (spectra.f90:188; from |xi|^2 when there is no temperature).
 * Synchronises.
 */
int pz_spectra(pz_handle *h, double *us2_m, double *up2_m, double *enst_m, double *temp2_m, double *xi2_m);

/*
 * Classical Nusselt numbers of the heat time series (outputs.f90:729-755):
 * out[0] = NuTop, out[1] = NuBot, 1 + dT/dr of the m = 0 temperature
 * perturbation over dT_cond/dr at each wall.  Synchronises.
 */
int pz_nusselt(pz_handle *h, double out[2]);

/* The HIP stream (hipStream_t) the handle launches on. */
void *pz_stream(pz_handle *h);

/*
 * Unit-level operators on caller buffers (host memory), the reference
 * module procedures of the path:
 *   pz_costf  — costf1 (radial_scheme.f90:47-49; dct_fftw.f90:175-197),
 *               in place on f(n_m, n_r) complex, m fastest;
 *   pz_get_dr / pz_get_ddr — get_dr/get_ddr complex 2-D
 *               (radial_derivatives.f90:313-418, 487-591; dct_fftw.f90:258-397);
 *   pz_fft / pz_ifft — fourier::fft/ifft (fft_fftw.f90:44-90) over n_rad
 *               radii: spectral (n_m, n_rad) <-> grid (n_phi, n_rad).
 * l_dct_in = 0 means f already holds Chebyshev coefficients.
 */
int pz_costf(pz_handle *h, double *f, int n_m);
int pz_get_dr(pz_handle *h, const double *f, double *df, int n_m, int l_dct_in);
int pz_get_ddr(pz_handle *h, const double *f, double *df, double *ddf, int n_m,
               int l_dct_in);
int pz_fft(pz_handle *h, const double *grid, double *spec, int n_rad);
int pz_ifft(pz_handle *h, const double *spec, double *grid, int n_rad);

/*
 * pz_get_dr / pz_get_ddr on DEVICE buffers, without allocating: no
 * counterpart in the reference; for callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  d_f, d_df and d_ddf are device arrays complex(8)
 * (n_m, n_r_max) in the layout of pz_get_ddr (m fastest); d_ddf = NULL
 * computes the first derivative only (pz_get_dr).  The layout changes are
 * staged through a static device array of 3 x 262144 complex values that
 * each GPU holds until the process ends, so n_m * n_r_max <= 262144
 * (larger sizes return -2).  Asynchronous on the handle's stream
 * (pz_stream); the caller synchronises before reading d_df.  Every call on
 * one GPU shares that GPU's staging array: calls that overlap on one GPU
 * (handles on different streams) race.
 */
int pz_get_ddr_noalloc(pz_handle *h, const double *d_f, double *d_df, double *d_ddf, int n_m,
                       int l_dct_in);

/*
 * Kernel timing for the roofline report: launches one stage kernel reps
 * times between HIP events on the handle's stream and returns the mean
 * duration in ms.  which = 0: the m-loop kernel of the run's radial method
 * for stage 1 with its explicit part (it advances the fields and the
 * time-scheme slots); which = 1: the radial loop.  The step state (fields,
 * time-scheme slots and scheme state) is copied before the launches and
 * restored after them, so the run continues from where it was; the
 * per-stage advection terms that the next radial loop rewrites are left as
 * the launches wrote them.  Allocates a copy of the step state.  No
 * reference counterpart (measurement only).
 */
int pz_time_kernel(pz_handle *h, int which, int reps, double *avg_ms);

/*
 * Ensembles of independent runs.  The reference runs one case per process
 * and has no counterpart; the reference interface each member replaces is
 * pz_step's.  The members are ordinary handles, created and started with
 * pz_create / pz_start (their sizes and physical parameters may all differ);
 * they must share the time scheme and the radial method.  pz_ensemble_step
 * advances every member by nsteps steps with one batched launch per kernel
 * and stage for all of them (graph-captured), so that small runs share the
 * GPU instead of queueing; each member's result equals its own pz_step.
 * pz_get_field, pz_time and the other per-handle calls work on the members
 * afterwards (they are ordered after the ensemble's work).  The members must
 * outlive the ensemble.
 */
typedef struct pz_ensemble pz_ensemble;
int pz_ensemble_create(pz_handle *const *runs, int nruns, pz_ensemble **e);
int pz_ensemble_step(pz_ensemble *e, int nsteps);
int pz_ensemble_destroy(pz_ensemble *e);
const char *pz_ensemble_last_error(const pz_ensemble *e);
/*
 * The collocation m-loop shape of the classes with n_r > 65, picked by the
 * ensemble's autotuner on its first pz_ensemble_step: it times a few steps
 * of the whole ensemble under each policy from the same member states
 * (restored after every trial) and keeps the fastest.  Policies: 0 the
 * single-run shape (512 threads, radial rows in LDS), 1 256 threads
 * (n_r <= 128), 2 radial rows and transform tables read through L2, 3
 * both.  Policies 0 and 2
 * give each member's pz_step results bit for bit; 1 and 3 sum the dense
 * products in another order (rounding-level differences).  The environment
 * variable PZ_ENSEMBLE_POLICY=k fixes the policy.  trial_ms (4 values, may
 * be null) receives the timed step of each policy tried.  Returns 1 before
 * the first step.
 */
int pz_ensemble_policy(const pz_ensemble *e, int *policy, double *trial_ms);

/*
 * Discrete adjoint of the step.  The reference has no adjoint; the forward
 * map differentiated is pz_step's (collocation or finite-difference method,
 * multistep schemes, dt fixed over the window), its objectives are the
 * diagnostics of outputs.f90:
 *   PZ_OBJ_NUTOP  J = (1/K) sum_{n=1..K} NuTop(n), the outer-wall Nusselt
 *                 number (outputs.f90:729-755) after each of the K steps;
 *   PZ_OBJ_EKIN   J = (1/K) sum_{n=1..K} E(n), the kinetic energy
 *                 us2_3D + up2_3D + uz2_3D of the e_kin_3D series
 *                 (outputs.f90:552-570, without round_off) after each step;
 *   PZ_OBJ_TEMP   J = sum Re(conj(W) T(K)) over the spectral temperature
 *                 after the last step, W = weights (layout of pz_set_field).
 * pz_adjoint_forward runs K steps from the state pz_start left, keeping a
 * snapshot of the step state every ckpt_every steps, and returns J; the
 * handle ends at the state after K steps.  pz_adjoint_backward replays each
 * segment from its snapshot and runs the transposed step kernels backwards
 * through it; it returns dJ times the gradient of J with respect to the
 * temperature field given to pz_start (grad_temp, layout of pz_set_field)
 * and to the boundary values of T in the implicit solves (grad_bc,
 * complex(8) (n_m_max, 2): outer wall, inner wall; the fixed temperature or
 * flux pattern of ktopt/kbott, with the conducting state held fixed; with
 * finite differences the wall patterns of l_fd_bc_pattern, zero without
 * them), and restores the state after K steps.  Either output may be null.  Returns -7
 * if dt (and so the matrices) changed inside the window.
 */
enum pz_objective { PZ_OBJ_NUTOP = 0, PZ_OBJ_EKIN = 1, PZ_OBJ_TEMP = 2 };
int pz_adjoint_forward(pz_handle *h, int objective, const double *weights, int nsteps, int ckpt_every,
                       double *J);
int pz_adjoint_backward(pz_handle *h, double dJ, double *grad_temp, double *grad_bc);

#ifdef __cplusplus
}
#endif

#endif /* PIZZA_AMD_H */

! ISO_C_BINDING interface of libpizza_amd (include/pizza_amd.h).
!
! One interface per C entry point, same names and argument meaning.  Host
! arrays keep the reference layout: complex(cp) f(n_m_max, n_r_max), m
! fastest (fields.f90), so module arrays are passed as they are.
module pizza_amd_gpu
   use iso_c_binding
   implicit none

   integer, parameter :: PZ_N_T_BOUNDS = 32          ! namelists.f90:23

   ! time schemes (select_time_scheme.f90:15-49)
   integer(c_int), parameter :: PZ_CNAB2 = 1, PZ_CNLF = 2, PZ_MODCNAB = 3, PZ_SBDF2 = 4, PZ_SBDF3 = 5, &
                                PZ_SBDF4 = 6, PZ_TVB33 = 7
   integer(c_int), parameter :: PZ_ARS222 = 101, PZ_ARS443 = 102, PZ_BPR353 = 103, PZ_DP1_242 = 104, &
                                PZ_DP2_242 = 105, PZ_LZ232 = 106, PZ_CK232 = 107, PZ_LZ453 = 108, PZ_PC2 = 109
   integer(c_int), parameter :: PZ_SSP222 = 110, PZ_LSDIRK222 = 111, PZ_ARS232 = 112, PZ_ARS233 = 113, &
                                PZ_SSP322 = 114, PZ_SSP332 = 115, PZ_SSP433 = 116, PZ_ARS343 = 117, &
                                PZ_MARS343 = 118, PZ_KC343 = 119, PZ_KC564 = 120, PZ_KC785 = 121, PZ_DBM453 = 122
   ! fields (fields.f90)
   integer(c_int), parameter :: PZ_OBJ_NUTOP = 0, PZ_OBJ_EKIN = 1, PZ_OBJ_TEMP = 2
   integer(c_int), parameter :: PZ_US = 0, PZ_UP = 1, PZ_OM = 2, PZ_PSI = 3, PZ_TEMP = 4, PZ_XI = 5, &
                                PZ_TEMP_HAT = 6, PZ_XI_HAT = 7, PZ_PSI_HAT = 8, PZ_DTEMP = 9

   ! the namelist-derived run configuration, layout of struct pz_config
   type, bind(c) :: pz_config
      integer(c_int) :: n_r_max, n_cheb_max, m_max, minc
      integer(c_int) :: time_scheme, l_cheb_coll, l_direct_solve, l_coriolis_imp, l_buo_imp
      integer(c_int) :: l_non_rot, time_scale_rot, l_rerror_fix
      real(c_double) :: rerror_fac, alpha, dtMax, courfac, dt_fac
      real(c_double) :: ra, ek, pr, raxi, sc, radratio, g0, g1, g2
      integer(c_int) :: ktopt, kbott, ktopv, kbotv, ktopxi, kbotxi
      integer(c_int) :: l_ek_pump, l_temp_advz
      real(c_double) :: tcond_fac, xicond_fac
      real(c_double) :: t_top(3*PZ_N_T_BOUNDS), t_bot(3*PZ_N_T_BOUNDS)
      real(c_double) :: xi_top(3*PZ_N_T_BOUNDS), xi_bot(3*PZ_N_T_BOUNDS)
      real(c_double) :: fd_stretch, fd_ratio
      integer(c_int) :: l_finite_diff, l_fd_bc_pattern
   end type pz_config

   ! one log-step row of every time-series file, layout of struct pz_series
   type, bind(c) :: pz_series
      real(c_double) :: time
      real(c_double) :: ekin_2d(3), ekin_3d(4), power_2d(5), power_3d(5)
      real(c_double) :: reynolds_2d(3), reynolds_3d(3), lengths(5), heat(7), chem(6)
   end type pz_series

   interface
      integer(c_int) function pz_time_series(h, ts) bind(c, name='pz_time_series')
         import :: c_int, c_ptr, pz_series
         type(c_ptr), value :: h
         type(pz_series), intent(out) :: ts
      end function pz_time_series

      subroutine pz_config_default(cfg) bind(c, name='pz_config_default')
         import :: pz_config
         type(pz_config), intent(out) :: cfg
      end subroutine pz_config_default

      integer(c_int) function pz_create(cfg, h) bind(c, name='pz_create')
         import :: pz_config, c_int, c_ptr
         type(pz_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: h
      end function pz_create

      integer(c_int) function pz_destroy(h) bind(c, name='pz_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function pz_destroy

      type(c_ptr) function pz_last_error_c(h) bind(c, name='pz_last_error')
         import :: c_ptr
         type(c_ptr), value :: h
      end function pz_last_error_c

      integer(c_int) function pz_dims(h, n_m_max, n_r_max, n_phi_max) bind(c, name='pz_dims')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), intent(out) :: n_m_max, n_r_max, n_phi_max
      end function pz_dims

      integer(c_int) function pz_set_field(h, field, host) bind(c, name='pz_set_field')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: h
         integer(c_int), value :: field
         complex(c_double_complex), intent(in) :: host(*)
      end function pz_set_field

      integer(c_int) function pz_get_field(h, field, host) bind(c, name='pz_get_field')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: h
         integer(c_int), value :: field
         complex(c_double_complex), intent(out) :: host(*)
      end function pz_get_field

      integer(c_int) function pz_init_fields(h, init_t, amp_t, init_xi, amp_xi, init_u, amp_u) &
            bind(c, name='pz_init_fields')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: init_t, init_xi, init_u
         real(c_double), value :: amp_t, amp_xi, amp_u
      end function pz_init_fields

      integer(c_int) function pz_start(h) bind(c, name='pz_start')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function pz_start

      integer(c_int) function pz_step(h, nsteps) bind(c, name='pz_step')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: nsteps
      end function pz_step

      integer(c_int) function pz_nstages(h) bind(c, name='pz_nstages')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function pz_nstages

      integer(c_int) function pz_stage_explicit(h, stage) bind(c, name='pz_stage_explicit')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: stage
      end function pz_stage_explicit

      integer(c_int) function pz_courant(h, dtr, dth) bind(c, name='pz_courant')
         import :: c_int, c_ptr
         type(c_ptr), value :: h, dtr, dth
      end function pz_courant

      integer(c_int) function pz_stage_implicit(h, stage, lMat) bind(c, name='pz_stage_implicit')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: stage, lMat
      end function pz_stage_implicit

      integer(c_int) function pz_write_checkpoint(h, path) bind(c, name='pz_write_checkpoint')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: h
         character(kind=c_char), intent(in) :: path(*)
      end function pz_write_checkpoint

      integer(c_int) function pz_read_checkpoint(h, path) bind(c, name='pz_read_checkpoint')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: h
         character(kind=c_char), intent(in) :: path(*)
      end function pz_read_checkpoint

      integer(c_int) function pz_write_frame(h, field, path) bind(c, name='pz_write_frame')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: h
         integer(c_int), value :: field
         character(kind=c_char), intent(in) :: path(*)
      end function pz_write_frame

      integer(c_int) function pz_time(h, time, dt) bind(c, name='pz_time')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: time, dt
      end function pz_time

      integer(c_int) function pz_ekin_3d(h, e) bind(c, name='pz_ekin_3d')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: e(4)
      end function pz_ekin_3d

      integer(c_int) function pz_costf(h, f, n_m) bind(c, name='pz_costf')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: h
         complex(c_double_complex), intent(inout) :: f(*)
         integer(c_int), value :: n_m
      end function pz_costf

      integer(c_int) function pz_get_dr(h, f, df, n_m, l_dct_in) bind(c, name='pz_get_dr')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: h
         complex(c_double_complex), intent(in) :: f(*)
         complex(c_double_complex), intent(out) :: df(*)
         integer(c_int), value :: n_m, l_dct_in
      end function pz_get_dr

      integer(c_int) function pz_get_ddr(h, f, df, ddf, n_m, l_dct_in) bind(c, name='pz_get_ddr')
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: h
         complex(c_double_complex), intent(in) :: f(*)
         complex(c_double_complex), intent(out) :: df(*), ddf(*)
         integer(c_int), value :: n_m, l_dct_in
      end function pz_get_ddr

      integer(c_int) function pz_fft(h, grid, spec, n_rad) bind(c, name='pz_fft')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: h
         real(c_double), intent(in) :: grid(*)
         complex(c_double_complex), intent(out) :: spec(*)
         integer(c_int), value :: n_rad
      end function pz_fft

      integer(c_int) function pz_ifft(h, spec, grid, n_rad) bind(c, name='pz_ifft')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: h
         complex(c_double_complex), intent(in) :: spec(*)
         real(c_double), intent(out) :: grid(*)
         integer(c_int), value :: n_rad
      end function pz_ifft
   end interface

contains

   ! pz_last_error as a Fortran string
   function pz_last_error(h) result(msg)
      type(c_ptr), intent(in) :: h
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: s(:)
      type(c_ptr) :: p
      integer :: n
      interface
         integer(c_size_t) function c_strlen(str) bind(c, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: str
         end function c_strlen
         integer(c_int) function pz_ensemble_create(runs, nruns, e) bind(c, name='pz_ensemble_create')
         import :: c_int, c_ptr
         type(c_ptr), intent(in) :: runs(*)
         integer(c_int), value :: nruns
         type(c_ptr), intent(out) :: e
      end function
      integer(c_int) function pz_ensemble_step(e, nsteps) bind(c, name='pz_ensemble_step')
         import :: c_int, c_ptr
         type(c_ptr), value :: e
         integer(c_int), value :: nsteps
      end function
      integer(c_int) function pz_ensemble_destroy(e) bind(c, name='pz_ensemble_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: e
      end function
      integer(c_int) function pz_nusselt(h, nu) bind(c, name='pz_nusselt')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: nu(2)
      end function
      integer(c_int) function pz_adjoint_forward(h, objective, weights, nsteps, ckpt_every, J) &
            bind(c, name='pz_adjoint_forward')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, weights
         integer(c_int), value :: objective, nsteps, ckpt_every
         real(c_double), intent(out) :: J
      end function
      integer(c_int) function pz_adjoint_backward(h, dJ, grad_temp, grad_bc) bind(c, name='pz_adjoint_backward')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, grad_temp, grad_bc
         real(c_double), value :: dJ
      end function
   end interface
      p = pz_last_error_c(h)
      n = int(c_strlen(p))
      call c_f_pointer(p, s, [n])
      allocate(character(len=n) :: msg)
      msg = transfer(s(1:n), msg)
   end function pz_last_error

end module pizza_amd_gpu

! pizza_gpu — driver of the GPU IMEX step, the counterpart of pizza.f90's
! main program (pizza.f90:63-199) and time_loop (step_time.f90:96-405) for
! the path the library implements.  It reads the reference's input
! namelist, creates the run, applies the namelist's initial conditions
! (init_fields.f90:260-509), steps through libpizza_amd and writes the
! e_kin_3D.TAG time series in the reference's format (outputs.f90:686-687)
! on every log step (l_correct_step, step_time.f90:145).
!
! usage: pizza_gpu input.nml [-n n_time_steps] [--print-config]
program pizza_gpu
   use iso_c_binding
   use pizza_amd_gpu
   use namelists
   implicit none

   character(len=512) :: fname, arg
   type(pz_config) :: cfg
   type(c_ptr) :: h
   integer :: nargs, i, n_steps, n_done, chunk, u
   logical :: print_only
   real(c_double) :: time, dt, e(4)

   nargs = command_argument_count()
   if (nargs < 1) then
      write(*, '(a)') 'usage: pizza_gpu input.nml [-n n_time_steps] [--print-config]'
      error stop 1
   end if
   call get_command_argument(1, fname)
   n_steps = -1
   print_only = .false.
   i = 2
   do while (i <= nargs)
      call get_command_argument(i, arg)
      if (trim(arg) == '-n') then
         i = i + 1
         call get_command_argument(i, arg)
         read(arg, *) n_steps
      else if (trim(arg) == '--print-config') then
         print_only = .true.
      else
         write(*, '(a)') 'pizza_gpu: unknown argument '//trim(arg)
         error stop 1
      end if
      i = i + 1
   end do

   call read_namelists(trim(fname))
   call fill_config(cfg)
   if (n_steps < 0) n_steps = n_time_steps
   if (print_only) then
      call print_config(cfg)
   else
      call check_supported()
      call run()
   end if

contains

   subroutine run()

      call check(pz_create(cfg, h), 'pz_create')
      call check(pz_init_fields(h, int(init_t, c_int), real(amp_t, c_double), int(init_xi, c_int), &
                                real(amp_xi, c_double), int(init_u, c_int), real(amp_u, c_double)), 'pz_init_fields')
      call check(pz_start(h), 'pz_start')

      open(newunit=u, file='e_kin_3D.'//trim(tag), status='replace', action='write')
      n_done = 0
      do
         call check(pz_time(h, time, dt), 'pz_time')
         call check(pz_ekin_3d(h, e), 'pz_ekin_3d')
         write(u, '(1P, es20.12, 4es16.8)') time, e(1), e(2), e(3), e(4)
         if (n_done >= n_steps) exit
         chunk = min(max(n_log_step, 1), n_steps - n_done)
         call check(pz_step(h, int(chunk, c_int)), 'pz_step')
         n_done = n_done + chunk
      end do
      close(u)
      call check(pz_destroy(h), 'pz_destroy')

   end subroutine run

   subroutine check(ierr, what)
      integer(c_int), intent(in) :: ierr
      character(len=*), intent(in) :: what
      if (ierr /= 0) then
         write(*, '(a, i0, a)') 'pizza_gpu: '//what//' failed (', ierr, '): '//pz_last_error(h)
         error stop 3
      end if
   end subroutine check

   subroutine print_config(c)
      type(pz_config), intent(in) :: c
      write(*, '(a, i0)') 'n_r_max ', c%n_r_max
      write(*, '(a, i0)') 'n_cheb_max ', c%n_cheb_max
      write(*, '(a, i0)') 'm_max ', c%m_max
      write(*, '(a, i0)') 'minc ', c%minc
      write(*, '(a, i0)') 'time_scheme ', c%time_scheme
      write(*, '(a, i0)') 'l_cheb_coll ', c%l_cheb_coll
      write(*, '(a, i0)') 'l_direct_solve ', c%l_direct_solve
      write(*, '(a, i0)') 'l_coriolis_imp ', c%l_coriolis_imp
      write(*, '(a, i0)') 'l_buo_imp ', c%l_buo_imp
      write(*, '(a, i0)') 'l_non_rot ', c%l_non_rot
      write(*, '(a, i0)') 'time_scale_rot ', c%time_scale_rot
      write(*, '(a, i0)') 'l_rerror_fix ', c%l_rerror_fix
      write(*, '(a, es24.16)') 'rerror_fac ', c%rerror_fac
      write(*, '(a, es24.16)') 'alpha ', c%alpha
      write(*, '(a, es24.16)') 'dtMax ', c%dtMax
      write(*, '(a, es24.16)') 'courfac ', c%courfac
      write(*, '(a, es24.16)') 'dt_fac ', c%dt_fac
      write(*, '(a, es24.16)') 'ra ', c%ra
      write(*, '(a, es24.16)') 'ek ', c%ek
      write(*, '(a, es24.16)') 'pr ', c%pr
      write(*, '(a, es24.16)') 'raxi ', c%raxi
      write(*, '(a, es24.16)') 'sc ', c%sc
      write(*, '(a, es24.16)') 'radratio ', c%radratio
      write(*, '(a, es24.16)') 'g0 ', c%g0
      write(*, '(a, es24.16)') 'g1 ', c%g1
      write(*, '(a, es24.16)') 'g2 ', c%g2
      write(*, '(a, i0)') 'ktopt ', c%ktopt
      write(*, '(a, i0)') 'kbott ', c%kbott
      write(*, '(a, i0)') 'ktopv ', c%ktopv
      write(*, '(a, i0)') 'kbotv ', c%kbotv
      write(*, '(a, i0)') 'ktopxi ', c%ktopxi
      write(*, '(a, i0)') 'kbotxi ', c%kbotxi
      write(*, '(a, i0)') 'l_ek_pump ', c%l_ek_pump
      write(*, '(a, i0)') 'l_temp_advz ', c%l_temp_advz
      write(*, '(a, es24.16)') 'tcond_fac ', c%tcond_fac
      write(*, '(a, es24.16)') 'xicond_fac ', c%xicond_fac
      write(*, '(a, es24.16)') 'fd_stretch ', c%fd_stretch
      write(*, '(a, es24.16)') 'fd_ratio ', c%fd_ratio
      write(*, '(a, i0)') 'l_finite_diff ', c%l_finite_diff
      write(*, '(a, i0)') 'fd_pad ', c%fd_pad
      write(*, '(a, i0)') 'init_t ', init_t
      write(*, '(a, es24.16)') 'amp_t ', amp_t
      write(*, '(a, i0)') 'init_xi ', init_xi
      write(*, '(a, es24.16)') 'amp_xi ', amp_xi
      write(*, '(a, i0)') 'init_u ', init_u
      write(*, '(a, es24.16)') 'amp_u ', amp_u
      write(*, '(a, i0)') 'n_log_step ', n_log_step
      write(*, '(a, i0)') 'n_time_steps ', n_steps
      write(*, '(a, a)') 'tag ', trim(tag)
   end subroutine print_config

end program pizza_gpu

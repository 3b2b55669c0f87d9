! The procedures of the reference's stage loop, under the reference's
! names, each calling the stage-granular entry points of libpizza_amd
! (pz_stage_explicit, pz_courant, pz_stage_implicit).  A driver that keeps
! time_loop's structure (step_time.f90:215-370) calls them in the
! reference's order:
!
!    do n_stage=1,nstages
!       call radial_loop(h, n_stage)                ! step_time.f90:222-251
!       if ( n_stage == 1 ) call dt_courant(h)      ! step_time.f90:256-304
!       call finish_explicit_assembly(h, n_stage)   ! step_time.f90:254
!       call mloop(h, n_stage, lMat)                ! step_time.f90:306-345
!    end do
!
! which launches exactly what pz_step(h, 1) launches.  The spectral state
! lives on the GPU, so the procedures take the run handle instead of the
! module arrays; pz_get_field / pz_set_field move fields to and from the
! reference's (n_m_max, n_r_max) arrays.
module step_stages
   use iso_c_binding
   use pizza_amd_gpu
   implicit none

   private
   public :: radial_loop, dt_courant, finish_explicit_assembly, mloop, time_loop_step

   ! stage whose explicit finish was requested and not yet run
   integer :: n_finish_stage = 0

contains

   ! rloop::radial_loop (radial_loop.f90:231-394) with the M->R and R->M
   ! transposes around it (step_time.f90:222, 251): the m->phi FFTs, the
   ! nonlinear products, the phi->m FFTs and, on stage 1, the per-radius
   ! Courant limits
   subroutine radial_loop(h, n_stage)
      type(c_ptr), intent(in) :: h
      integer, intent(in) :: n_stage
      call check(h, pz_stage_explicit(h, int(n_stage, c_int)), 'radial_loop')
   end subroutine radial_loop

   ! courant_mod::dt_courant (courant.f90:78-167), the dt roll and the
   ! weights of the new step (step_time.f90:256-304); optionally returns
   ! the smallest radial and horizontal Courant limits
   subroutine dt_courant(h, dtr, dth)
      type(c_ptr), intent(in) :: h
      real(c_double), intent(out), optional, target :: dtr, dth
      type(c_ptr) :: pr, ph
      pr = c_null_ptr
      ph = c_null_ptr
      if (present(dtr)) pr = c_loc(dtr)
      if (present(dth)) ph = c_loc(dth)
      call check(h, pz_courant(h, pr, ph), 'dt_courant')
   end subroutine dt_courant

   ! finish_explicit_assembly (mloop.f90:169-229): the radial derivatives
   ! of the advective fluxes and the explicit Coriolis, buoyancy and
   ! pumping terms.  They are per-m work, so the library runs them at the
   ! start of the per-m kernel of the following mloop; this records the
   ! request, and mloop checks it was made for its own stage.
   subroutine finish_explicit_assembly(h, n_stage)
      type(c_ptr), intent(in) :: h
      integer, intent(in) :: n_stage
      if (.not. c_associated(h)) error stop 'finish_explicit_assembly: null run'
      n_finish_stage = n_stage
   end subroutine finish_explicit_assembly

   ! mloop (mloop.f90:36-103): the implicit solves of T, xi and psi per m,
   ! after the matrix rebuild on stage 1 when lMat is set or dt changed;
   ! on the last stage of a DIRK scheme with an assembly stage, the
   ! assemble_stage of mloop.f90:105-160 instead
   subroutine mloop(h, n_stage, lMat)
      type(c_ptr), intent(in) :: h
      integer, intent(in) :: n_stage
      logical, intent(in) :: lMat
      if (n_finish_stage /= 0 .and. n_finish_stage /= n_stage) &
         error stop 'mloop: finish_explicit_assembly was called for another stage'
      n_finish_stage = 0
      call check(h, pz_stage_implicit(h, int(n_stage, c_int), merge(1_c_int, 0_c_int, lMat)), 'mloop')
   end subroutine mloop

   ! one iteration of time_loop's stage loop (step_time.f90:215-370)
   subroutine time_loop_step(h)
      type(c_ptr), intent(in) :: h
      integer :: n_stage, nstages
      nstages = pz_nstages(h)
      do n_stage = 1, nstages
         call radial_loop(h, n_stage)
         if (n_stage == 1) call dt_courant(h)
         call finish_explicit_assembly(h, n_stage)
         call mloop(h, n_stage, .false.)
      end do
   end subroutine time_loop_step

   subroutine check(h, ierr, what)
      type(c_ptr), intent(in) :: h
      integer(c_int), intent(in) :: ierr
      character(len=*), intent(in) :: what
      if (ierr /= 0) then
         write(*, '(a, i0, a)') 'step_stages: '//what//' failed (', ierr, '): '//pz_last_error(h)
         error stop 3
      end if
   end subroutine check

end module step_stages

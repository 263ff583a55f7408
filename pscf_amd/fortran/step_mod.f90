!-----------------------------------------------------------------------
! step_mod -- drop-in replacement of src/step_mod.f90: init_step,
! make_propg and step_exp keep their interfaces; the propagator factors and
! the Richardson-extrapolated pseudo-spectral step run on the GPU.
!-----------------------------------------------------------------------
module step_mod
   use, intrinsic :: iso_c_binding
   use const_mod
   use fft_mod
   use pscf_hip_c
   implicit none

   private
   public :: init_step      ! allocate the device state
   public :: make_propg     ! exp(-ds*b^2*k^2/6), exp(-ds*omega/2) and halves
   public :: step_exp       ! one contour step

   integer(c_int64_t), save :: hstep = 0

contains

   subroutine init_step(n)
   integer, intent(IN) :: n(3)
   integer(c_int) :: ng(3)
   if (hstep /= 0) call hip_check(c_pscf_step_destroy(hstep), 'init_step')
   ng = n
   call hip_check(c_pscf_step_create(ng, hstep), 'init_step')
   end subroutine init_step

   subroutine make_propg(ds, b, ksq, omega)
   real(long), intent(IN) :: ds
   real(long), intent(IN) :: b
   real(long), intent(IN) :: ksq(0:,0:,0:)
   real(long), intent(IN) :: omega(0:,0:,0:)
   call hip_check(c_pscf_step_make_propg(hstep, ds, b, ksq, omega), 'make_propg')
   end subroutine make_propg

   subroutine step_exp(q_in, q_out, plan)
   real(long), intent(IN)     :: q_in(0:,0:,0:)
   real(long), intent(OUT)    :: q_out(0:,0:,0:)
   type(fft_plan), intent(IN) :: plan
   call hip_check(c_pscf_step_exp(hstep, q_in, q_out), 'step_exp')
   end subroutine step_exp

end module step_mod

!-----------------------------------------------------------------------
! fft_mod -- drop-in replacement of the reference's src/fft3_mod.f90
! (same module name, fft_plan type and public procedures) running the
! transforms on the GPU through libpscf_hip.  Unnormalised, sign -1 forward
! (r2c, x halved), sign +1 inverse (c2r), fftc with 1/M on direction +1.
! Device plans are cached per grid: the reference calls create_fft_plan
! freely (src/fft3_mod.f90:61-67 only records the grid).
!-----------------------------------------------------------------------
module fft_mod
   use, intrinsic :: iso_c_binding
   use const_mod
   use pscf_hip_c
   implicit none

   PRIVATE
   PUBLIC :: fft_plan
   PUBLIC :: create_fft_plan
   PUBLIC :: fftc
   PUBLIC :: fft
   PUBLIC :: ifft

   type fft_plan
      integer    ::  n(3)   ! grid dimensions, 1 for unused dimensions
      integer*8  ::  f      ! libpscf_hip plan handle
      integer*8  ::  r      ! same handle (kept for layout compatibility)
   end type fft_plan

   integer, parameter       :: max_cached = 16
   integer, save            :: n_cached = 0
   integer, save            :: cached_n(3, max_cached)
   integer(c_int64_t), save :: cached_h(max_cached)

contains

   subroutine create_fft_plan(ngrid, plan, fft_c2c)
   integer, intent(IN)           :: ngrid(3)
   type(fft_plan), intent(OUT)   :: plan
   logical, optional, intent(IN) :: fft_c2c
   integer :: i
   integer(c_int) :: ng(3)
   plan%n = ngrid
   do i = 1, n_cached
      if (all(cached_n(:, i) == ngrid)) then
         plan%f = cached_h(i)
         plan%r = cached_h(i)
         return
      end if
   end do
   if (n_cached == max_cached) stop 'fft_mod: too many distinct FFT grids'
   ng = ngrid
   n_cached = n_cached + 1
   call hip_check(c_pscf_fft_plan_create(ng, cached_h(n_cached)), 'create_fft_plan')
   cached_n(:, n_cached) = ngrid
   plan%f = cached_h(n_cached)
   plan%r = cached_h(n_cached)
   end subroutine create_fft_plan

   subroutine fft(plan, in, out)
   type(fft_plan), intent(IN)  :: plan
   real(long), intent(IN)      :: in(0:,0:,0:)
   complex(long), intent(OUT)  :: out(0:,0:,0:)
   call hip_check(c_pscf_fft_r2c(int(plan%f, c_int64_t), in, out), 'fft')
   end subroutine fft

   subroutine ifft(plan, in, out)
   type(fft_plan), intent(IN)  :: plan
   complex(long), intent(IN)   :: in(0:,0:,0:)
   real(long), intent(OUT)     :: out(0:,0:,0:)
   call hip_check(c_pscf_fft_c2r(int(plan%r, c_int64_t), in, out), 'ifft')
   end subroutine ifft

   subroutine fftc(direction, plan, in, out)
   integer, intent(IN)         :: direction
   type(fft_plan), intent(IN)  :: plan
   complex(long), intent(IN)   :: in(0:,0:,0:)
   complex(long), intent(OUT)  :: out(0:,0:,0:)
   call hip_check(c_pscf_fft_c2c(int(plan%f, c_int64_t), int(direction, c_int), in, out), 'fftc')
   end subroutine fftc

end module fft_mod

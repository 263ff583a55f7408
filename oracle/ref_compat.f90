!-----------------------------------------------------------------------
! ref_compat -- test infrastructure for the oracle build (oracle/Makefile).
!
! flang does not short-circuit `.and.`, so the reference's idiom
! `present(x) .and. x` dereferences an absent optional argument.  The
! oracle recipe rewrites each such expression (on copies of the reference
! sources under oracle/_ref/src) into a call to one of these helpers, which
! only read the argument when it is present.  Nothing else changes.
!-----------------------------------------------------------------------
module ref_compat
   implicit none
   private
   public :: opt_true, opt_gt1
contains
   logical function opt_true(x)
      logical, intent(IN), optional :: x
      opt_true = .false.
      if (present(x)) opt_true = x
   end function opt_true

   logical function opt_gt1(n)
      integer, intent(IN), optional :: n
      opt_gt1 = .false.
      if (present(n)) opt_gt1 = (n > 1)
   end function opt_gt1
end module ref_compat

!-----------------------------------------------------------------------
! grid_basis_mod -- drop-in replacement of src/grid_basis_mod.f90.
! basis_to_kgrid / kgrid_to_basis run on the GPU from the star tables and
! the device k -> wave map (G_to_bz with the reference's tie rule).  The
! karray_full and no_parity variants are used only by the RESPONSE command
! (src/response_mod.f90:460,1297), which this drop-in does not provide: they
! stop with a message.  check_symmetry is the reference's host diagnostic.
!-----------------------------------------------------------------------
module grid_basis_mod
   use, intrinsic :: iso_c_binding
   use const_mod
   use grid_mod
   use basis_mod
   use pscf_hip_c
   use pscf_hip_state
   implicit none

   PRIVATE
   PUBLIC :: basis_to_kgrid
   PUBLIC :: kgrid_to_basis
   PUBLIC :: check_symmetry

contains

   subroutine basis_to_kgrid(basis, kgrid, karray_full, no_parity)
   real(long),intent(IN)     :: basis(:)
   complex(long),intent(OUT) :: kgrid(0:,0:,0:)
   logical,optional          :: karray_full
   logical,optional          :: no_parity
   if (present(karray_full)) then
      if (karray_full) stop 'basis_to_kgrid: karray_full (RESPONSE) is not provided on the GPU path'
   end if
   if (present(no_parity)) then
      if (no_parity) stop 'basis_to_kgrid: no_parity (RESPONSE) is not provided on the GPU path'
   end if
   call ensure_system()
   call sync_cell()
   call hip_check(c_pscf_basis_to_kgrid(hsys, basis, kgrid, 0_c_int), 'basis_to_kgrid')
   end subroutine basis_to_kgrid

   subroutine kgrid_to_basis(kgrid, basis, no_parity)
   complex(long),intent(IN)  :: kgrid(0:,0:,0:)
   real(long),intent(OUT)    :: basis(:)
   logical,optional          :: no_parity
   if (present(no_parity)) then
      if (no_parity) stop 'kgrid_to_basis: no_parity (RESPONSE) is not provided on the GPU path'
   end if
   call ensure_system()
   call hip_check(c_pscf_kgrid_to_basis(hsys, kgrid, basis, 0_c_int), 'kgrid_to_basis')
   end subroutine kgrid_to_basis

   ! Diagnostic files <prefix>.dead / <prefix>.live (src/grid_basis_mod.f90:211-276)
   subroutine check_symmetry(kgrid, basis, prefix)
   complex(long), intent(IN)         :: kgrid(0:,0:,0:)
   real(long), intent(IN), optional  :: basis(:)
   character(*)                      :: prefix
   integer  :: i, j, k, n_dead, n_live, G(3), Gbz(3), i_wave
   n_dead = 0
   n_live = 0
   open(11, file=trim(prefix)//'.dead', status='unknown')
   do k = 0, ngrid(3)-1
      G(3) = k
      do j = 0, ngrid(2)-1
         G(2) = j
         do i = 0, ngrid(1)-1
            G(1) = i
            Gbz = G_to_bz(G)
            i_wave = which_wave(Gbz(1), Gbz(2), Gbz(3))
            if (i_wave < 1 .or. i_wave > N_wave) then
               n_dead = n_dead + 1
               if (i <= ngrid(1)/2) write(11,*) Gbz(1:dim), kgrid(i,j,k)
            else
               n_live = n_live + 1
            end if
         end do
      end do
   end do
   close(11)
   write(6,*)
   if (n_live /= N_wave) write(6,*) "********", prefix, " SERIOUS ********"
   write(6,*) 'n_live=', n_live
   write(6,*) 'N_wave=', N_wave
   write(6,*)
   open(12, file=trim(prefix)//'.live', status='unknown')
   G = 0
   Gbz = 0
   do i = 1, N_star
      write(12,*) 'star=', i, 'star_count=', star_count(i)
      do j = star_begin(i), star_end(i)
         Gbz(1:dim) = wave(1:dim, j)
         G = G_to_fft(Gbz)
         if (G(1) > ngrid(1)/2) then
            G = -G
            G = G_to_fft(G)
         end if
         if (present(basis)) then
            write(12,*) Gbz(1:dim), dble(kgrid(G(1),G(2),G(3))/coeff(j)) - basis(i)
         else
            write(12,*) Gbz(1:dim), dble(kgrid(G(1),G(2),G(3))/coeff(j))
         end if
      end do
      write(12,*)
   end do
   close(12)
   end subroutine check_symmetry

end module grid_basis_mod

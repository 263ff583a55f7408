!-----------------------------------------------------------------------
! chain_mod -- drop-in replacement of src/chain_mod.f90.  chain_grid_type
! keeps its components; the contour discretisation (block_bgn, block_ds)
! comes from libpscf_hip (the same rule drives the device slabs); qf/qr
! host slabs are allocated only when a caller asks (allocate_q), since the
! GPU scf_mod keeps its slabs on the device.
!-----------------------------------------------------------------------
module chain_mod
   use, intrinsic :: iso_c_binding
   use const_mod
   use fft_mod, only : fft_plan
   use pscf_hip_c
   implicit none

   private
   public :: chain_grid_type
   public :: null_chain_grid
   public :: make_chain_grid
   public :: destroy_chain_grid

   type chain_grid_type
      integer,    pointer     :: block_bgn(:)     ! 1st element of block
      real(long), pointer     :: block_ds(:)      ! step size for block
      real(long), pointer     :: qf(:,:,:,:)      ! qf(x,y,z,s)
      real(long), pointer     :: qr(:,:,:,:)      ! qr(x,y,z,s)
      real(long), pointer     :: rho(:,:,:,:)     ! rho(x,y,z,block)
      type(fft_plan)          :: plan
      real(long)              :: bigQ
      complex(long), pointer  :: del_qf(:,:,:,:)
      complex(long), pointer  :: del_qr(:,:,:,:)
      complex(long)           :: delQ
   end type

contains

   subroutine null_chain_grid(chain)
   type(chain_grid_type), intent(INOUT) :: chain
   nullify(chain%block_bgn, chain%block_ds, chain%qf, chain%qr, chain%rho, &
           chain%del_qf, chain%del_qr)
   end subroutine null_chain_grid

   subroutine make_chain_grid(chain, plan, nblk, blk_length, ds0, allocate_q, perturb, order)
   type(chain_grid_type), intent(INOUT) :: chain
   type(fft_plan),        intent(IN)    :: plan
   integer,               intent(IN)    :: nblk
   real(long),            intent(IN)    :: blk_length(:)
   real(long),            intent(IN)    :: ds0
   logical,               intent(IN)    :: allocate_q
   logical, optional,     intent(IN)    :: perturb
   integer, optional,     intent(IN)    :: order
   integer :: nx, ny, nz, bgns, error
   integer(c_int) :: ord
   integer(c_int) :: bgn(nblk+1)
   real(c_double) :: bds(nblk)

   nx = plan%n(1)-1
   ny = plan%n(2)-1
   nz = plan%n(3)-1
   chain%plan = plan
   if (.not. associated(chain%block_bgn)) then
      allocate(chain%block_bgn(nblk+1), STAT=error)
      if (error /= 0) stop 'chain%block_bgn allocation error'
   end if
   if (.not. associated(chain%block_ds)) then
      allocate(chain%block_ds(nblk), STAT=error)
      if (error /= 0) stop 'chain%block_ds allocation error'
   end if
   if (.not. associated(chain%rho)) then
      allocate(chain%rho(0:nx,0:ny,0:nz,nblk), STAT=error)
      if (error /= 0) stop 'chain%rho allocation error!'
   end if
   ord = -1
   if (present(order)) ord = order
   call hip_check(c_pscf_chain_discretize(int(nblk, c_int), blk_length(1:nblk), ds0, ord, &
                  bgn, bds), 'make_chain_grid')
   chain%block_bgn = bgn
   chain%block_ds = bds
   bgns = bgn(nblk+1)
   if (allocate_q) then
      allocate(chain%qf(0:nx,0:ny,0:nz,bgns), STAT=error)
      if (error /= 0) stop "chain%qf allocation error!"
      allocate(chain%qr(0:nx,0:ny,0:nz,bgns), STAT=error)
      if (error /= 0) stop "chain%qr allocation error!"
   end if
   if (present(perturb)) then
      if (perturb) then
         allocate(chain%del_qf(0:nx,0:ny,0:nz,bgns), STAT=error)
         if (error /= 0) stop "chain%qf allocation error!"
         allocate(chain%del_qr(0:nx,0:ny,0:nz,bgns), STAT=error)
         if (error /= 0) stop "chain%qr allocation error!"
      end if
   end if
   end subroutine make_chain_grid

   subroutine destroy_chain_grid(chain)
   type(chain_grid_type) :: chain
   if (associated(chain%block_bgn)) deallocate(chain%block_bgn)
   if (associated(chain%block_ds))  deallocate(chain%block_ds)
   if (associated(chain%rho))       deallocate(chain%rho)
   if (associated(chain%qf))        deallocate(chain%qf)
   if (associated(chain%qr))        deallocate(chain%qr)
   if (associated(chain%del_qf))    deallocate(chain%del_qf)
   if (associated(chain%del_qr))    deallocate(chain%del_qr)
   call null_chain_grid(chain)
   end subroutine destroy_chain_grid

end module chain_mod

!===============================================================================
!> ISO_C_BINDING interface of libm3drt for Fortran hosts (dispatch.x).
!>
!> The bind(C) types and interfaces mirror include/m3drt.h one to one.  Arrays
!> cross the boundary as type(c_ptr) device addresses; m3drt_device_alloc and the
!> copy helpers let a host without HIP bindings own device buffers.  Status
!> conventions are those of HOP's C ABI (HOP/standalone/hop_c_api.f90): every
!> call returns an integer status, 0 for success, and m3drt_last_error returns
!> the message.
!>
!> short_char2_gpu below is the drop-in for rt_patch_t%short_char2
!> (SRC/rt_solvers_mod.f90:224-507) followed by integrate_intensity
!> (SRC/rt_patch_mod.f90:3399-3431) over one direction: it takes the device copies
!> of the slab column's opacity and source function in the reference layout
!> work(f,x,y,z) and accumulates J.
!===============================================================================
module m3drt
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: M3DRT_OK = 0, M3DRT_F32 = 0, M3DRT_F64 = 1
  integer(c_int), parameter, public :: M3DRT_LINEAR2 = 0
  integer(c_int), parameter, public :: M3DRT_SPLIT_RATES = 1, M3DRT_GET_RATES = 2

  type, bind(C), public :: m3drt_sweep_desc
    integer(c_int32_t) :: nf, nx, ny, nz
    integer(c_int32_t) :: dir, prec, solver, flags
    real(c_double) :: ds(3), vec(3)
    real(c_double) :: maser_dtau_max, maser_Inu_max, wj
    integer(c_int64_t) :: cs_f, cs_x, cs_y, cs_z
    integer(c_int64_t) :: gs_f, gs_x, gs_y
    integer(c_int64_t) :: eta_ld, eta_nrow
  end type m3drt_sweep_desc

  type, bind(C), public :: m3drt_sweep_bufs
    type(c_ptr) :: chi = c_null_ptr, S = c_null_ptr
    type(c_ptr) :: ghost_chi = c_null_ptr, ghost_S = c_null_ptr, ghost_I = c_null_ptr
    type(c_ptr) :: chi500 = c_null_ptr, chi500_up = c_null_ptr, g500 = c_null_ptr
    type(c_ptr) :: eta_row = c_null_ptr, eta_mask = c_null_ptr
    type(c_ptr) :: J = c_null_ptr, PsiJ = c_null_ptr
    type(c_ptr) :: I_out = c_null_ptr, Psi_out = c_null_ptr
    type(c_ptr) :: Irem_out = c_null_ptr, escape_out = c_null_ptr
    type(c_ptr) :: exit_I = c_null_ptr
  end type m3drt_sweep_bufs

  public :: m3drt_create, m3drt_destroy, m3drt_synchronize, m3drt_last_error
  public :: m3drt_device_alloc, m3drt_device_free, m3drt_copy_to_device, m3drt_copy_to_host
  public :: m3drt_sc_sweep, m3drt_reference_strides, m3drt_native_strides
  public :: short_char2_gpu, m3drt_error_message

  interface
    integer(c_int) function m3drt_create(ctx, device, flags) bind(C, name='m3drt_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device, flags
    end function m3drt_create
    integer(c_int) function m3drt_destroy(ctx) bind(C, name='m3drt_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function m3drt_destroy
    integer(c_int) function m3drt_synchronize(ctx) bind(C, name='m3drt_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function m3drt_synchronize
    integer(c_int) function m3drt_last_error(buf, buflen) bind(C, name='m3drt_last_error')
      import :: c_int, c_char
      character(kind=c_char), dimension(*) :: buf
      integer(c_int), value :: buflen
    end function m3drt_last_error
    integer(c_int) function m3drt_device_alloc(ctx, bytes, ptr) bind(C, name='m3drt_device_alloc')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: bytes
      type(c_ptr), intent(out) :: ptr
    end function m3drt_device_alloc
    integer(c_int) function m3drt_device_free(ctx, ptr) bind(C, name='m3drt_device_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ptr
    end function m3drt_device_free
    integer(c_int) function m3drt_copy_to_device(ctx, dst, src, bytes) bind(C, name='m3drt_copy_to_device')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function m3drt_copy_to_device
    integer(c_int) function m3drt_copy_to_host(ctx, dst, src, bytes) bind(C, name='m3drt_copy_to_host')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function m3drt_copy_to_host
    integer(c_int) function m3drt_sc_sweep(ctx, desc, bufs) bind(C, name='m3drt_sc_sweep')
      import :: c_int, c_ptr, m3drt_sweep_desc, m3drt_sweep_bufs
      type(c_ptr), value :: ctx
      type(m3drt_sweep_desc), intent(in) :: desc
      type(m3drt_sweep_bufs), intent(in) :: bufs
    end function m3drt_sc_sweep
    integer(c_int) function m3drt_reference_strides(desc) bind(C, name='m3drt_reference_strides')
      import :: c_int, m3drt_sweep_desc
      type(m3drt_sweep_desc), intent(inout) :: desc
    end function m3drt_reference_strides
    integer(c_int) function m3drt_native_strides(desc) bind(C, name='m3drt_native_strides')
      import :: c_int, m3drt_sweep_desc
      type(m3drt_sweep_desc), intent(inout) :: desc
    end function m3drt_native_strides
  end interface

contains

  !-----------------------------------------------------------------------------
  !> The last error message as a Fortran string.
  !-----------------------------------------------------------------------------
  function m3drt_error_message() result(msg)
    character(len=256) :: msg
    character(kind=c_char) :: buf(256)
    integer :: i, st
    st = m3drt_last_error(buf, 256_c_int)
    msg = ' '
    do i = 1, 256
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
  end function m3drt_error_message

  !-----------------------------------------------------------------------------
  !> One direction of rt_patch_t%short_char2 + integrate_intensity on the device,
  !> in the reference layout: chi, S, J are device copies of work(f,x,y,1:nz,1:2)
  !> and Jnu(f,x,y,1:nz); ghost holds the upstream plane work(:,:,:,0|nz+1,1:3) as
  !> three consecutive (nf,nx,ny) planes; chi500/chi500_up (nx,ny,nz), g500 (nx,ny).
  !> wmu2 is the 2*wmu of integrate_intensity.  Returns the status.
  !-----------------------------------------------------------------------------
  integer function short_char2_gpu(ctx, nf, nx, ny, nz, dir, ds, vec, chi, S, ghost, chi500, &
                                   chi500_up, g500, wmu2, J, exit_I, maser_dtau_max, maser_Inu_max) &
                                   result(st)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nf, nx, ny, nz, dir
    real(8), intent(in) :: ds(3), vec(3), wmu2, maser_dtau_max, maser_Inu_max
    type(c_ptr), intent(in) :: chi, S, ghost, chi500, chi500_up, g500, J, exit_I
    type(m3drt_sweep_desc) :: d
    type(m3drt_sweep_bufs) :: b
    integer(c_size_t) :: plane
    d%nf = nf; d%nx = nx; d%ny = ny; d%nz = nz
    d%dir = dir
    d%prec = M3DRT_F32
    d%solver = M3DRT_LINEAR2
    d%flags = 0
    d%ds = ds
    d%vec = vec
    d%maser_dtau_max = maser_dtau_max
    d%maser_Inu_max = maser_Inu_max
    d%wj = wmu2
    d%eta_ld = 0
    d%eta_nrow = 0
    st = m3drt_reference_strides(d)
    if (st /= M3DRT_OK) return
    plane = int(nf, c_size_t) * nx * ny * 4
    b%chi = chi
    b%S = S
    b%ghost_chi = ghost
    b%ghost_S = transfer(transfer(ghost, 0_c_intptr_t) + plane, ghost)
    b%ghost_I = transfer(transfer(ghost, 0_c_intptr_t) + 2 * plane, ghost)
    b%chi500 = chi500
    b%chi500_up = chi500_up
    b%g500 = g500
    b%J = J
    b%exit_I = exit_I
    st = m3drt_sc_sweep(ctx, d, b)
  end function short_char2_gpu
end module m3drt

!===============================================================================
! A Fortran host calling the GPU sweep through the ISO_C_BINDING module, as
! dispatch.x's rt_patch_t would: reference-layout work arrays on the host, device
! copies through the library's helpers, short_char2_gpu per direction.
!
! Case 1: a 3D column with random opacity and a constant source function S0, the
! upstream ghost plane at I = S0.  Every linear step keeps I = S0 exactly (its
! weights sum to one, the upwind interpolation reproduces constants), so J must
! come back as wmu2*S0 everywhere and the exit plane as S0.  Prints PASS or FAIL.
!
! Case 2: the same column with a random source function; the inputs and the J and
! exit plane the GPU returns are written, as raw Fortran-order streams, to the file
! named by the first argument (when given), for the oracle to check:
!   nf nx ny nz (int32); ds(3) vec(3) wmu2 (real(8)); chi S (nf,nx,ny,nz);
!   ghost (nf,nx,ny,3: chi, S, I); c500 cup (nx,ny,nz); g500 (nx,ny); J; exit (nf,nx,ny)
!===============================================================================
program sweep_driver
  use iso_c_binding
  use m3drt
  implicit none
  integer, parameter :: nf = 16, nx = 10, ny = 9, nz = 12
  real(4), allocatable, target :: chi(:,:,:,:), S(:,:,:,:), ghost(:,:,:,:), J(:,:,:,:), ex(:,:,:)
  real(4), allocatable, target :: c500(:,:,:), cup(:,:,:), g500(:,:)
  type(c_ptr) :: ctx, dchi, dsrc, dgh, dJ, dex, dc5, dcu, dg5
  real(8) :: ds(3), vec(3), wmu2
  real(4) :: S0, err
  integer :: st, z, uo
  character(len=512) :: fout
  S0 = 0.0372
  allocate(chi(nf,nx,ny,nz), S(nf,nx,ny,nz), ghost(nf,nx,ny,3), J(nf,nx,ny,nz), ex(nf,nx,ny))
  allocate(c500(nx,ny,nz), cup(nx,ny,nz), g500(nx,ny))
  call random_seed()
  call random_number(chi)
  do z = 1, nz
    chi(:,:,:,z) = 1e-7 * exp(-0.35*(z-1)) * (1.0 + 10.0*chi(:,:,:,z))
    c500(:,:,z) = 2e-7 * exp(-0.35*(z-1))
  end do
  cup(:,:,2:nz) = c500(:,:,1:nz-1)
  g500 = 2e-7 * exp(0.35)
  cup(:,:,1) = g500
  S = S0
  ghost(:,:,:,1) = chi(:,:,:,1) * 1.4
  ghost(:,:,:,2) = S0
  ghost(:,:,:,3) = S0
  J = 0
  ds = [50d5, 50d5, 10d5]
  vec = [0.3d0, -0.4d0, sqrt(1d0 - 0.25d0)]
  wmu2 = 0.125d0
  st = m3drt_create(ctx, 0_c_int, 0_c_int)
  call check('create')
  call put(dchi, c_loc(chi), size(chi))
  call put(dsrc, c_loc(S), size(S))
  call put(dgh, c_loc(ghost), size(ghost))
  call put(dJ, c_loc(J), size(J))
  call put(dex, c_loc(ex), size(ex))
  call put(dc5, c_loc(c500), size(c500))
  call put(dcu, c_loc(cup), size(cup))
  call put(dg5, c_loc(g500), size(g500))
  st = short_char2_gpu(ctx, nf, nx, ny, nz, 1, ds, vec, dchi, dsrc, dgh, dc5, dcu, dg5, wmu2, dJ, dex, 2d0, 1d30)
  call check('short_char2_gpu')
  st = m3drt_synchronize(ctx)
  st = m3drt_copy_to_host(ctx, c_loc(J), dJ, int(size(J), c_size_t) * 4)
  call check('copy J')
  st = m3drt_copy_to_host(ctx, c_loc(ex), dex, int(size(ex), c_size_t) * 4)
  call check('copy exit')
  err = max(maxval(abs(J / real(wmu2 * S0, 4) - 1.0)), maxval(abs(ex / S0 - 1.0)))
  print '(a,es10.3)', 'max rel err ', err
  if (err < 1e-5) then
    print '(a)', 'PASS'
  else
    print '(a)', 'FAIL'
  end if
  if (command_argument_count() >= 1) then
    call get_command_argument(1, fout)
    call random_number(S)
    S = S0 * (0.5 + S)
    ghost(:,:,:,2) = S(:,:,:,1) * 1.1
    ghost(:,:,:,3) = ghost(:,:,:,2)
    J = 0
    st = m3drt_copy_to_device(ctx, dsrc, c_loc(S), int(size(S), c_size_t) * 4)
    call check('copy S')
    st = m3drt_copy_to_device(ctx, dgh, c_loc(ghost), int(size(ghost), c_size_t) * 4)
    call check('copy ghost')
    st = m3drt_copy_to_device(ctx, dJ, c_loc(J), int(size(J), c_size_t) * 4)
    call check('copy J')
    st = short_char2_gpu(ctx, nf, nx, ny, nz, 1, ds, vec, dchi, dsrc, dgh, dc5, dcu, dg5, wmu2, dJ, dex, 2d0, 1d30)
    call check('short_char2_gpu, random source')
    st = m3drt_synchronize(ctx)
    st = m3drt_copy_to_host(ctx, c_loc(J), dJ, int(size(J), c_size_t) * 4)
    call check('copy J')
    st = m3drt_copy_to_host(ctx, c_loc(ex), dex, int(size(ex), c_size_t) * 4)
    call check('copy exit')
    open(newunit=uo, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(uo) nf, nx, ny, nz
    write(uo) ds, vec, wmu2
    write(uo) chi, S, ghost, c500, cup, g500, J, ex
    close(uo)
    print '(a)', 'WROTE'
  end if
  st = m3drt_destroy(ctx)
contains
  subroutine put(d, h, n)
    type(c_ptr), intent(out) :: d
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: n
    st = m3drt_device_alloc(ctx, int(n, c_size_t) * 4, d)
    call check('alloc')
    st = m3drt_copy_to_device(ctx, d, h, int(n, c_size_t) * 4)
    call check('copy')
  end subroutine put
  subroutine check(what)
    character(len=*), intent(in) :: what
    if (st /= M3DRT_OK) then
      print '(a,a,a,a)', 'FAIL ', what, ': ', trim(m3drt_error_message())
      stop 1
    end if
  end subroutine check
end program sweep_driver

!===============================================================================
! TEST INFRASTRUCTURE.  Drives routines of the reference that compile on their
! own (SRC/ray_mod.f90 with solvers/rt/radau_mod.f90, HOP/interpol_mod.f90,
! HOP/optab_mod.f90), built from their sources by oracle/build_ref.sh, on inputs
! written by tests/golden/make_golden.py, and writes their outputs as golden
! vectors.
!
!   rt_golden.x <kind> <infile> <outfile>
!
! kinds:
!   quad   the angular quadratures of setup_quadrature (SRC/rt_params_mod.f90:
!          170-220): per request a scheme name (16 chars), ntheta, nphi,
!          Jnu_weights, one_dimensional, ds(3), nmu, mu(nmu) (custom) ->
!          nray, vec(3,nray), wts(nray)
!   pchip  the bottom ghost plane of set_upstream_boundary without the LTE ghost
!          (SRC/rt_patch_mod.f90:3287-3300): n, ds3, y(3,n) -> ynew(n) of
!          pchip_mono at one plane below the first, before the clamp at 0
!   optab  sample_interpolate of a sampled background table (HOP/optab_mod.f90:
!          127-287) -> node slopes and (back, scat) at the query cells
!   save   the statements of save_intensity (SRC/rt_patch_mod.f90:4151-4195) for
!          one patch and one set, with the constants of HOP/const_mod.f90:
!          nf, nx, ny, nmu, nu(nf), wmu(nmu), I(nf,nx,ny,nmu) -> i3, ie, flux
! All files are raw little-endian streams.
!===============================================================================
program rt_golden
  use ray_mod
  use interpol_mod, only: pchip_mono
  use optab_mod, only: sample_optab_type
  use const_mod, only: HH, CC, CM_TO_AA, pi
  implicit none
  character(len=256) :: kind, fin, fout
  integer :: ui, uo
  call get_command_argument(1, kind)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open(newunit=ui, file=trim(fin), access='stream', form='unformatted', status='old')
  open(newunit=uo, file=trim(fout), access='stream', form='unformatted', status='replace')
  select case (trim(kind))
  case ('quad')
    call do_quad()
  case ('pchip')
    call do_pchip()
  case ('optab')
    call do_optab()
  case ('save')
    call do_save()
  case default
    error stop 'rt_golden: unknown kind'
  end select
  close(ui)
  close(uo)

contains

  subroutine do_quad()
    integer :: nreq, r, ntheta, nphi, jnu, one_d, nmu, n
    character(len=16) :: scheme
    real(8) :: ds(3)
    real(8), allocatable :: mu(:), vec(:,:), wts(:)
    read(ui) nreq
    do r = 1, nreq
      read(ui) scheme, ntheta, nphi, jnu, one_d, ds, nmu
      if (allocated(mu)) deallocate(mu)
      allocate(mu(max(nmu, 1)))
      if (nmu > 0) read(ui) mu(1:nmu)
      one_dimensional = one_d /= 0
      select case (trim(scheme))
      case ('disk_center', 'set_a2', 'set_a4', 'set_a6', 'set_a8', 'set_b4', 'set_b6', 'set_b8')
        call set_carlson_quadratures(trim(scheme), vec, wts, Jnu_weights=jnu /= 0)
      case ('radau')
        call set_radau_quadratures(ntheta, nphi, vec, wts, Jnu_weights=jnu /= 0)
      case ('lobatto')
        call set_lobatto_quadratures(ntheta, nphi, vec, wts, Jnu_weights=jnu /= 0)
      case ('gauss')
        call set_gauss_quadratures(ntheta, nphi, vec, wts, Jnu_weights=jnu /= 0)
      case ('custom')
        call set_custom_quadratures(mu(1:nmu), nphi, vec, wts)
      case ('grid+', 'gridx', 'grid*')
        call set_grid_quadratures(trim(scheme), ds, vec, wts)
      case default
        error stop 'rt_golden: unknown quadrature'
      end select
      n = size(wts)
      write(uo) n
      write(uo) vec, wts
      deallocate(vec, wts)
    end do
    one_dimensional = .false.
  end subroutine do_quad

  subroutine do_pchip()
    integer :: n, j, i
    real(8) :: ds3
    real(4) :: xxold(3), xxnew(1), yn(1)
    real(4), allocatable :: y(:,:)
    read(ui) n, ds3
    allocate(y(3, n))
    read(ui) y
    xxold = [(i*ds3, i=1, 3)]
    xxnew = [(i*ds3, i=0, 0)]
    do j = 1, n
      call pchip_mono(xxold, y(:, j), xxnew, yn, sorted=.true.)
      write(uo) yn(1)
    end do
  end subroutine do_pchip

  subroutine do_optab()
    type(sample_optab_type) :: t
    integer :: ncell, nf
    real(8), allocatable :: temp(:,:,:), ne(:,:,:), lam_q(:)
    real(4), allocatable :: chi500(:,:,:), back(:,:,:,:), scat(:,:,:,:)
    read(ui) t%nnu, t%ntemp, t%nne
    read(ui) t%pca%mu, t%pca%rot, t%pca%c0, t%pca%dc
    allocate(t%lam(t%nnu), t%val(t%nnu, t%ntemp, t%nne), t%scat(t%nnu, t%ntemp, t%nne))
    read(ui) t%lam, t%val, t%scat
    read(ui) ncell, nf
    allocate(temp(ncell, 1, 1), ne(ncell, 1, 1), chi500(ncell, 1, 1), lam_q(nf))
    allocate(back(nf, ncell, 1, 1), scat(nf, ncell, 1, 1))
    read(ui) temp, ne, chi500, lam_q
    call t%build_derivatives()
    call t%interpolate(temp, ne, chi500, back, scat, lam_q, 1, nf)
    write(uo) t%dl, t%dt, t%de, t%sl, t%st, t%se, back, scat
  end subroutine do_optab

  subroutine do_save()
    real(8), parameter :: conversion = 2 * HH / (CC**3 * CM_TO_AA)
    integer :: nf, nx, ny, nmu, f, x, y, mu
    real(8), allocatable :: nu(:), wmu(:)
    real(4), allocatable :: mem(:,:,:,:), i3(:,:,:,:), ie(:,:), flux(:)
    read(ui) nf, nx, ny, nmu
    allocate(nu(nf), wmu(nmu), mem(nf, nx, ny, nmu), i3(nf, nx, ny, nmu), ie(nf, nmu), flux(nf))
    read(ui) nu, wmu, mem
    do f = 1, nf
      i3(f,:,:,:) = conversion * nu(f)**5 * mem(f,:,:,:)
    end do
    ie = 0.
    do y = 1, ny
      do x = 1, nx
        ie = ie + i3(:,x,y,:)
      end do
    end do
    ie = ie / real(nx*ny)
    flux = 0.
    do mu = 1, nmu
      flux = flux + real(2*pi*wmu(mu), 4) * ie(:,mu)
    end do
    write(uo) i3, ie, flux
  end subroutine do_save

end program rt_golden

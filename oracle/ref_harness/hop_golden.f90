!===============================================================================
! TEST INFRASTRUCTURE.  Drives the reference's own compiled routines (HOP of
! OpenMULTI3D, built from their sources by oracle/build_ref.sh) on inputs written
! by tests/golden/make_golden.py, and writes their outputs as golden vectors.
!
!   hop_golden.x <kind> <infile> <outfile>
!
! kinds: expneg (HOP/exp_neg.inc), qlgbot / qlgref / qlgline (HOP/qlg_mod.f90),
! lineq (HOP/op_math_mod.f90 lineq_no_pivot), opac_equi / opac_gen (HOP/linelist_mod.f90
! add_opacity_equi / add_opacity on a grid of any spacing).  All files are raw little-endian streams.
!===============================================================================
module expneg_h
  implicit none
contains
#include "exp_neg.inc"
end module expneg_h

program hop_golden
  use expneg_h
  use qlg
  use op_math_mod, only: lineq_no_pivot
  use line_types_mod, only: linelist_type
  use linelist_mod, only: add_opacity_equi, add_opacity
  implicit none
  character(len=256) :: kind, fin, fout
  integer :: ui, uo
  call get_command_argument(1, kind)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open(newunit=ui, file=trim(fin), access='stream', form='unformatted', status='old')
  open(newunit=uo, file=trim(fout), access='stream', form='unformatted', status='replace')
  call qlg_init
  select case (trim(kind))
  case ('expneg')
    call do_expneg()
  case ('qlgbot')
    call do_qlgbot()
  case ('qlgref')
    call do_qlgref()
  case ('qlgline')
    call do_qlgline()
  case ('lineq')
    call do_lineq()
  case ('opac_equi')
    call do_opac_equi(.false.)
  case ('opac_gen')
    call do_opac_equi(.true.)
  case default
    error stop 'hop_golden: unknown kind'
  end select
  close(ui)
  close(uo)
contains

  subroutine do_expneg()
    integer :: n, i
    real(4), allocatable :: t(:), E(:), u0(:)
    read(ui) n
    allocate(t(n), E(n), u0(n))
    read(ui) t
    do i = 1, n
      call exp_neg(t(i), E(i), u0(i))
    end do
    write(uo) E, u0
  end subroutine do_expneg

  subroutine do_qlgbot()
    integer :: n, i
    real(4), allocatable :: a(:), v(:), K(:)
    real(4) :: K1(1)
    read(ui) n
    allocate(a(n), v(n), K(n))
    read(ui) a, v
    do i = 1, n
      call qlg_bot_sweep(a(i), v(i), 0.0, 1, K1)
      K(i) = K1(1)
    end do
    write(uo) K
  end subroutine do_qlgbot

  subroutine do_qlgref()
    integer :: n, i
    real(8), allocatable :: a(:), v(:), K(:)
    real(8) :: K1(1)
    read(ui) n
    allocate(a(n), v(n), K(n))
    read(ui) a, v
    do i = 1, n
      call qlg_ref_f64(a(i), v(i), 0.0d0, 1, K1)
      K(i) = K1(1)
    end do
    write(uo) K
  end subroutine do_qlgref

  subroutine do_qlgline()
    integer :: nc, nf, c, fa, fb
    real(4) :: D, gL, S, ref, lnq, cap, hw
    real(8) :: llc, pos, idl
    real(8), allocatable :: llam(:)
    real(4), allocatable :: K(:)
    type(qlg_bot_cell_t) :: w
    read(ui) nc, nf
    allocate(llam(nf), K(nf))
    read(ui) llam
    do c = 1, nc
      read(ui) D, gL, S, ref, lnq, cap, llc, pos, idl
      w = qlg_bot_cell(D)
      K = 0
      call qlg_bot_line_sweep(w, gL, S, ref, lnq, nf, llam, llc, pos, idl, cap, fa, fb, hw, K)
      write(uo) fa, fb, hw, K
    end do
  end subroutine do_qlgline

  subroutine do_lineq()
    integer :: n, ns, s
    real(8), allocatable :: a(:,:), b(:)
    read(ui) n, ns
    allocate(a(n,n), b(n))
    do s = 1, ns
      read(ui) a, b
      call lineq_no_pivot(a, b, n)
      write(uo) b
    end do
  end subroutine do_lineq

  !-----------------------------------------------------------------------------
  ! One isotope's line list over a block of cells, as add_particle_opacity hands
  ! it to the kernel (HOP/linelist_mod.f90:1040-1046).
  !-----------------------------------------------------------------------------
  subroutine do_opac_equi(general)
    logical, intent(in) :: general
    type(linelist_type) :: ll
    integer :: nl, nf, m(3), np, nlev, flags, l
    logical :: do_nlte, has_ipart, do_scat, has_reach
    integer, allocatable :: active_i(:)
    real(8), allocatable :: llam(:)
    real(4), allocatable :: chi(:,:,:,:), eta(:,:,:,:), cont(:,:,:,:), logref(:,:,:,:)
    real(4), allocatable :: vproj(:,:,:), idf(:,:,:), vb2(:,:,:), ekt(:,:,:), ne(:,:,:)
    real(4), allocatable :: tothi(:,:,:), spe(:,:,:), lnbase(:,:,:,:), dep(:,:,:,:)
    real(4), allocatable :: ome(:,:,:,:), csc(:,:,:,:)
    integer, allocatable :: unsold_i(:), ipart(:)
    read(ui) nl, nf, m, np, nlev, flags
    do_nlte = iand(flags, 1) /= 0
    has_ipart = iand(flags, 2) /= 0
    do_scat = iand(flags, 4) /= 0
    has_reach = iand(flags, 8) /= 0
    ll%nline = nl
    allocate(ll%llam0(nl), ll%lower_ev(nl), ll%gpref(nl), ll%grad(nl), ll%gw1(nl), ll%gw2(nl), &
             ll%gstk(nl), ll%unsold(nl), ll%lev_i(nl), ll%lev_j(nl), ll%active(nl), ll%used(nl), &
             ll%strength(nl), ll%reach(nl), ll%reach_in(nl), unsold_i(nl), ipart(nl))
    read(ui) ll%llam0, ll%lower_ev, ll%gpref, ll%grad, ll%gw1, ll%gw2, ll%gstk, unsold_i, &
             ll%lev_i, ll%lev_j, ipart
    ll%unsold = unsold_i /= 0
    ll%active = .true.
    ll%used = .false.
    ll%strength = 0
    ll%reach = 0
    ll%reach_in = -1
    ll%has_reach = .false.
    if (has_reach) then
      ! a list whose every line carries a reach, as finalize_linelist marks it
      ! (HOP/linelist_mod.f90:375-378), and its active flags
      allocate(active_i(nl))
      read(ui) ll%reach_in, active_i
      ll%active = active_i /= 0
      ll%has_reach = all(ll%reach_in(1:nl) >= 0)
      if (ll%has_reach) ll%reach_in_max = maxval(ll%reach_in(1:nl))
    end if
    if (has_ipart) then
      allocate(ll%ipart(nl))
      ll%ipart = ipart
    end if
    allocate(llam(nf))
    read(ui) llam
    allocate(chi(nf,m(1),m(2),m(3)), eta(nf,m(1),m(2),m(3)), cont(nf,m(1),m(2),m(3)), &
             logref(nf,m(1),m(2),m(3)))
    allocate(vproj(m(1),m(2),m(3)), idf(m(1),m(2),m(3)), vb2(m(1),m(2),m(3)), ekt(m(1),m(2),m(3)), &
             ne(m(1),m(2),m(3)), tothi(m(1),m(2),m(3)), spe(m(1),m(2),m(3)))
    allocate(lnbase(m(1),m(2),m(3),np), dep(m(1),m(2),m(3),nlev))
    read(ui) chi, eta, cont, logref, vproj, idf, vb2, ekt, ne, tothi, spe, lnbase, dep
    if (do_scat) then
      allocate(ome(nf,m(1),m(2),m(3)), csc(nf,m(1),m(2),m(3)))
      read(ui) ome, csc
    else
      allocate(ome(0,0,0,0), csc(0,0,0,0))
    end if
    if (general) then
      call add_opacity(ll, llam, chi, eta, cont, logref, vproj, idf, vb2, ekt, lnbase, ne, &
                       tothi, spe, dep, do_nlte, ome, csc, do_scat)
    else
      call add_opacity_equi(ll, llam, chi, eta, cont, logref, vproj, idf, vb2, ekt, lnbase, ne, &
                            tothi, spe, dep, do_nlte, ome, csc, do_scat)
    end if
    write(uo) chi, eta
    write(uo) (merge(1, 0, ll%used(l)), l = 1, nl)
    write(uo) ll%strength, ll%reach
    if (do_scat) write(uo) csc
  end subroutine do_opac_equi
end program hop_golden

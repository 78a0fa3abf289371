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
  integer(c_int), parameter, public :: M3DRT_LINEAR2 = 0, M3DRT_QUADRATIC = 1, M3DRT_LINEAR = 2, &
                                        M3DRT_LONG_LINEAR = 3
  integer(c_int), parameter, public :: M3DRT_SPLIT_RATES = 1, M3DRT_GET_RATES = 2
  integer(c_int), parameter, public :: M3DRT_OPAC_NLTE = 1, M3DRT_OPAC_SCAT = 2
  integer(c_int), parameter, public :: M3DRT_RATES_PRECOND = 1, M3DRT_RATES_OWN_NORM = 2, &
                                       M3DRT_RATES_LARGEST = 4, M3DRT_RATES_SCAT = 8, M3DRT_RATES_FLAGS = 16
  integer(c_int), parameter, public :: M3DRT_BBOP_LTE = 1, M3DRT_BBOP_SPLIT = 2
  integer(c_int), parameter, public :: M3DRT_UNITS_SPLIT = 1, M3DRT_UNITS_SCAT = 2, &
                                       M3DRT_UNITS_SUPPRESS_MASERS = 4
  integer(c_int), parameter, public :: M3DRT_BG_SCAT_ALI = 1, M3DRT_BG_J_TO_B = 2
  integer(c_int), parameter, public :: M3DRT_ETA_GHOST = 1
  integer(c_int), parameter, public :: M3DRT_LCH_TOP = 0, M3DRT_LCH_BOTTOM = 1, M3DRT_LCH_UPSTREAM = 2
  integer(c_int), parameter, public :: M3DRT_GHOST_TOP = 0, M3DRT_GHOST_LTE = 1, M3DRT_GHOST_PCHIP = 2, &
                                       M3DRT_GHOST_UPSTREAM = 3

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

  type, bind(C), public :: m3drt_adjoint_bufs
    type(c_ptr) :: chi = c_null_ptr
    type(c_ptr) :: S = c_null_ptr
    type(c_ptr) :: ghost_chi = c_null_ptr
    type(c_ptr) :: ghost_S = c_null_ptr
    type(c_ptr) :: ghost_I = c_null_ptr
    type(c_ptr) :: chi500 = c_null_ptr
    type(c_ptr) :: chi500_up = c_null_ptr
    type(c_ptr) :: g500 = c_null_ptr
    type(c_ptr) :: I = c_null_ptr
    type(c_ptr) :: Jbar = c_null_ptr
    type(c_ptr) :: Ibar = c_null_ptr
    type(c_ptr) :: exit_bar = c_null_ptr
    type(c_ptr) :: chibar = c_null_ptr
    type(c_ptr) :: Sbar = c_null_ptr
    type(c_ptr) :: ghost_chibar = c_null_ptr
    type(c_ptr) :: ghost_Sbar = c_null_ptr
    type(c_ptr) :: ghost_Ibar = c_null_ptr
  end type m3drt_adjoint_bufs

  type, bind(C), public :: m3drt_linelist
    integer(c_int32_t) :: nline
    type(c_ptr) :: llam0 = c_null_ptr
    type(c_ptr) :: lower_ev = c_null_ptr
    type(c_ptr) :: gpref = c_null_ptr
    type(c_ptr) :: grad = c_null_ptr
    type(c_ptr) :: gw1 = c_null_ptr
    type(c_ptr) :: gw2 = c_null_ptr
    type(c_ptr) :: gstk = c_null_ptr
    type(c_ptr) :: unsold = c_null_ptr
    type(c_ptr) :: lev_i = c_null_ptr
    type(c_ptr) :: lev_j = c_null_ptr
    type(c_ptr) :: ipart = c_null_ptr
    type(c_ptr) :: used = c_null_ptr
    type(c_ptr) :: strength = c_null_ptr
    type(c_ptr) :: reach = c_null_ptr
    type(c_ptr) :: reach_in = c_null_ptr
    type(c_ptr) :: active = c_null_ptr
    type(c_ptr) :: sel_work = c_null_ptr
  end type m3drt_linelist

  type, bind(C), public :: m3drt_opacity_args
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: ncell
    integer(c_int32_t) :: npart
    integer(c_int32_t) :: nlev
    integer(c_int32_t) :: flags
    real(c_float) :: tolerance
    real(c_float) :: line_reject_frac
    integer(c_int64_t) :: sf
    integer(c_int64_t) :: sc
    type(c_ptr) :: llam = c_null_ptr
    type(c_ptr) :: vproj = c_null_ptr
    type(c_ptr) :: inv_df = c_null_ptr
    type(c_ptr) :: vbar2 = c_null_ptr
    type(c_ptr) :: e_kT = c_null_ptr
    type(c_ptr) :: ne = c_null_ptr
    type(c_ptr) :: tothi = c_null_ptr
    type(c_ptr) :: stark_pe = c_null_ptr
    type(c_ptr) :: lnbase = c_null_ptr
    type(c_ptr) :: dep = c_null_ptr
    type(c_ptr) :: chi = c_null_ptr
    type(c_ptr) :: eta = c_null_ptr
    type(c_ptr) :: cont = c_null_ptr
    type(c_ptr) :: logref = c_null_ptr
    type(c_ptr) :: one_m_eps = c_null_ptr
    type(c_ptr) :: chi_scat = c_null_ptr
  end type m3drt_opacity_args

  type, bind(C), public :: m3drt_rate_solve_args
    integer(c_int32_t) :: nk
    integer(c_int32_t) :: nrad
    integer(c_int32_t) :: nline
    integer(c_int32_t) :: npair
    integer(c_int32_t) :: ncell
    type(c_ptr) :: trans_i = c_null_ptr
    type(c_ptr) :: trans_j = c_null_ptr
    type(c_ptr) :: trans_active = c_null_ptr
    type(c_ptr) :: g = c_null_ptr
    type(c_ptr) :: ev = c_null_ptr
    type(c_ptr) :: Aji = c_null_ptr
    type(c_ptr) :: pair_i = c_null_ptr
    type(c_ptr) :: pair_j = c_null_ptr
    type(c_ptr) :: e_kT = c_null_ptr
    type(c_ptr) :: temp = c_null_ptr
    type(c_ptr) :: totn = c_null_ptr
    type(c_ptr) :: nstar = c_null_ptr
    type(c_ptr) :: b = c_null_ptr
    type(c_ptr) :: gamma_ij = c_null_ptr
    type(c_ptr) :: gamma_ji = c_null_ptr
    type(c_ptr) :: cup = c_null_ptr
    real(c_double) :: max_update
    real(c_double) :: min_pop_frac
    real(c_double) :: conv_pop_floor
    real(c_double) :: sor_w
    real(c_double) :: cswitch
    real(c_double) :: cut_x
    real(c_double) :: bf_cut_rate
    real(c_double) :: tmax_update
  end type m3drt_rate_solve_args

  type, bind(C), public :: m3drt_rate_solve_result
    real(c_double) :: maxerr
    integer(c_int32_t) :: epos
    integer(c_int32_t) :: elvl
    integer(c_int64_t) :: nclamp
    type(c_ptr) :: lvlerr = c_null_ptr
  end type m3drt_rate_solve_result

  type, bind(C), public :: m3drt_hermite_fields
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: pe4 = c_null_ptr
    type(c_ptr) :: temp = c_null_ptr
    type(c_ptr) :: dpe_dx = c_null_ptr
    type(c_ptr) :: dpe_dy = c_null_ptr
    type(c_ptr) :: dpe_dz = c_null_ptr
    type(c_ptr) :: dlnt_dx = c_null_ptr
    type(c_ptr) :: dlnt_dy = c_null_ptr
    type(c_ptr) :: dlnt_dz = c_null_ptr
    type(c_ptr) :: hknu = c_null_ptr
  end type m3drt_hermite_fields

  type, bind(C), public :: m3drt_lc_ghost
    integer(c_int32_t) :: kind = 0
    type(c_ptr) :: pe = c_null_ptr
    type(c_ptr) :: dpe_dz = c_null_ptr
    type(c_ptr) :: temp = c_null_ptr
    type(c_ptr) :: dlnt_dz = c_null_ptr
  end type m3drt_lc_ghost

  type, bind(C), public :: m3drt_bb_rates_args
    integer(c_int32_t) :: nline
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: ncell
    integer(c_int32_t) :: nk
    integer(c_int32_t) :: flags
    real(c_float) :: wmu
    real(c_float) :: wtot
    integer(c_int64_t) :: sf
    integer(c_int64_t) :: sc
    type(c_ptr) :: iblue = c_null_ptr
    type(c_ptr) :: ired = c_null_ptr
    type(c_ptr) :: lev_i = c_null_ptr
    type(c_ptr) :: lev_j = c_null_ptr
    type(c_ptr) :: Aij = c_null_ptr
    type(c_ptr) :: Bij = c_null_ptr
    type(c_ptr) :: lower_off = c_null_ptr
    type(c_ptr) :: q = c_null_ptr
    type(c_ptr) :: wnu = c_null_ptr
    type(c_ptr) :: wq = c_null_ptr
    type(c_ptr) :: g = c_null_ptr
    type(c_ptr) :: b = c_null_ptr
    type(c_ptr) :: nstar = c_null_ptr
    type(c_ptr) :: pop_ratio = c_null_ptr
    type(c_ptr) :: inv_df = c_null_ptr
    type(c_ptr) :: vproj = c_null_ptr
    type(c_ptr) :: e_kT = c_null_ptr
    type(c_ptr) :: adamp = c_null_ptr
    type(c_ptr) :: Inu = c_null_ptr
    type(c_ptr) :: Psi = c_null_ptr
    type(c_ptr) :: Bnu = c_null_ptr
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: gamma_ij = c_null_ptr
    type(c_ptr) :: gamma_ji = c_null_ptr
    integer(c_int32_t) :: split_offset
    integer(c_int32_t) :: npw
    type(c_ptr) :: Irem = c_null_ptr
    type(c_ptr) :: escape = c_null_ptr
    type(c_ptr) :: chi = c_null_ptr
    type(c_ptr) :: split_opac1 = c_null_ptr
    type(c_ptr) :: split_opac2 = c_null_ptr
    type(c_ptr) :: split_chi1 = c_null_ptr
    type(c_ptr) :: split_chi2 = c_null_ptr
    type(c_ptr) :: split_id = c_null_ptr
    type(c_ptr) :: split_chi_id = c_null_ptr
    type(c_ptr) :: chi_scat = c_null_ptr
    type(c_ptr) :: dJnu = c_null_ptr
    type(c_ptr) :: ntrans = c_null_ptr
    type(c_ptr) :: psi_flags_in = c_null_ptr
    type(c_ptr) :: psi_flags_out = c_null_ptr
  end type m3drt_bb_rates_args

  type, bind(C), public :: m3drt_bb_opacity_args
    integer(c_int32_t) :: nline
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: ncell
    integer(c_int32_t) :: nk
    integer(c_int32_t) :: flags
    integer(c_int32_t) :: split_offset
    real(c_float) :: wmu
    integer(c_int64_t) :: sf
    integer(c_int64_t) :: sc
    type(c_ptr) :: iblue = c_null_ptr
    type(c_ptr) :: ired = c_null_ptr
    type(c_ptr) :: lev_i = c_null_ptr
    type(c_ptr) :: lev_j = c_null_ptr
    type(c_ptr) :: Bij = c_null_ptr
    type(c_ptr) :: lower_off = c_null_ptr
    type(c_ptr) :: q = c_null_ptr
    type(c_ptr) :: wq = c_null_ptr
    type(c_ptr) :: wphi_line = c_null_ptr
    type(c_ptr) :: b = c_null_ptr
    type(c_ptr) :: nstar = c_null_ptr
    type(c_ptr) :: pop_ratio = c_null_ptr
    type(c_ptr) :: inv_df = c_null_ptr
    type(c_ptr) :: vproj = c_null_ptr
    type(c_ptr) :: e_kT = c_null_ptr
    type(c_ptr) :: adamp = c_null_ptr
    type(c_ptr) :: xhat = c_null_ptr
    type(c_ptr) :: ehat = c_null_ptr
    type(c_ptr) :: split_opac1 = c_null_ptr
    type(c_ptr) :: split_opac2 = c_null_ptr
    type(c_ptr) :: split_chi1 = c_null_ptr
    type(c_ptr) :: split_chi2 = c_null_ptr
    type(c_ptr) :: split_id = c_null_ptr
    type(c_ptr) :: split_chi_id = c_null_ptr
    type(c_ptr) :: wphi = c_null_ptr
  end type m3drt_bb_opacity_args

  type, bind(C), public :: m3drt_bf_rates_args
    integer(c_int32_t) :: ncont
    integer(c_int32_t) :: nline
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: ncell
    integer(c_int32_t) :: nk
    integer(c_int32_t) :: precond
    real(c_double) :: cut_x
    integer(c_int64_t) :: sf
    integer(c_int64_t) :: sc
    type(c_ptr) :: iblue = c_null_ptr
    type(c_ptr) :: ired = c_null_ptr
    type(c_ptr) :: lev_i = c_null_ptr
    type(c_ptr) :: lev_j = c_null_ptr
    type(c_ptr) :: active = c_null_ptr
    type(c_ptr) :: ev = c_null_ptr
    type(c_ptr) :: wnu = c_null_ptr
    type(c_ptr) :: alpha = c_null_ptr
    type(c_ptr) :: twohnu3_c2 = c_null_ptr
    type(c_ptr) :: e_kT = c_null_ptr
    type(c_ptr) :: nstar = c_null_ptr
    type(c_ptr) :: b = c_null_ptr
    type(c_ptr) :: Jnu = c_null_ptr
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: PsiJ = c_null_ptr
    type(c_ptr) :: gamma_ij = c_null_ptr
    type(c_ptr) :: gamma_ji = c_null_ptr
  end type m3drt_bf_rates_args

  type, bind(C), public :: m3drt_units_args
    integer(c_int64_t) :: n
    integer(c_int32_t) :: flags
    real(c_double) :: maser_chi_floor
    type(c_ptr) :: work1 = c_null_ptr
    type(c_ptr) :: work2 = c_null_ptr
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: split_opac1 = c_null_ptr
    type(c_ptr) :: split_opac2 = c_null_ptr
    type(c_ptr) :: chi_scat = c_null_ptr
    type(c_ptr) :: dJnu = c_null_ptr
    type(c_ptr) :: n_floored = c_null_ptr
  end type m3drt_units_args

  type, bind(C), public :: m3drt_background_args
    integer(c_int64_t) :: n
    integer(c_int32_t) :: flags
    real(c_float) :: wscat
    real(c_double) :: scat_op_max
    type(c_ptr) :: J = c_null_ptr
    type(c_ptr) :: B = c_null_ptr
    type(c_ptr) :: chi_cont = c_null_ptr
    type(c_ptr) :: sigma = c_null_ptr
    type(c_ptr) :: PsiJ = c_null_ptr
    type(c_ptr) :: Jali_old = c_null_ptr
    type(c_ptr) :: Jali = c_null_ptr
    type(c_ptr) :: eta = c_null_ptr
    type(c_ptr) :: dJnu = c_null_ptr
    type(c_ptr) :: n_neg = c_null_ptr
    type(c_ptr) :: diag = c_null_ptr
  end type m3drt_background_args

  type, bind(C), public :: m3drt_eta_args
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: nx
    integer(c_int32_t) :: ny
    integer(c_int32_t) :: nz
    integer(c_int32_t) :: dir
    integer(c_int32_t) :: flags
    integer(c_int32_t) :: rx
    integer(c_int32_t) :: ry
    type(c_ptr) :: chi = c_null_ptr
    type(c_ptr) :: S = c_null_ptr
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: ghost_chi = c_null_ptr
    type(c_ptr) :: ghost_S = c_null_ptr
    type(c_ptr) :: src = c_null_ptr
    type(c_ptr) :: mask = c_null_ptr
    type(c_ptr) :: ghost_src = c_null_ptr
  end type m3drt_eta_args

  type, bind(C), public :: m3drt_window_args
    integer(c_int32_t) :: flags
    real(c_double) :: maser_chi_floor
    type(c_ptr) :: chi_cont = c_null_ptr
    type(c_ptr) :: Bnu = c_null_ptr
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: n_floored = c_null_ptr
  end type m3drt_window_args

  type, bind(C), public :: m3drt_rotate_args
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: nx
    integer(c_int32_t) :: ny
    integer(c_int32_t) :: nz
    integer(c_int32_t) :: nk
    integer(c_int32_t) :: reserved
    integer(c_int64_t) :: cs_f
    integer(c_int64_t) :: cs_x
    integer(c_int64_t) :: cs_y
    integer(c_int64_t) :: cs_z
    integer(c_int64_t) :: cs_k
    real(c_double) :: vec(3)
    real(c_double) :: ds(3)
    real(c_double) :: bot2pivot
    real(c_double) :: ray_offset(2)
    type(c_ptr) :: in = c_null_ptr
    type(c_ptr) :: out = c_null_ptr
  end type m3drt_rotate_args

  type, bind(C), public :: m3drt_b1d_batch
    integer(c_int32_t) :: natm
    integer(c_int32_t) :: reserved
    type(c_ptr) :: nd = c_null_ptr
    type(c_ptr) :: nf = c_null_ptr
    type(c_ptr) :: nmu = c_null_ptr
    type(c_ptr) :: nline = c_null_ptr
    type(c_ptr) :: d_off = c_null_ptr
    type(c_ptr) :: l_off = c_null_ptr
    type(c_ptr) :: ld_off = c_null_ptr
    type(c_ptr) :: mu_off = c_null_ptr
    type(c_ptr) :: f_off = c_null_ptr
    type(c_ptr) :: i_off = c_null_ptr
    type(c_ptr) :: lam0 = c_null_ptr
    type(c_ptr) :: dlam = c_null_ptr
    type(c_ptr) :: dtau = c_null_ptr
    type(c_ptr) :: hckT = c_null_ptr
    type(c_ptr) :: ivth = c_null_ptr
    type(c_ptr) :: cont_a = c_null_ptr
    type(c_ptr) :: cont_b = c_null_ptr
    type(c_ptr) :: line_off = c_null_ptr
    type(c_ptr) :: line_cl = c_null_ptr
    type(c_ptr) :: line_rec = c_null_ptr
    type(c_ptr) :: mu = c_null_ptr
    type(c_ptr) :: wflux = c_null_ptr
    real(c_double) :: lam_ref
    type(c_ptr) :: I_top = c_null_ptr
    type(c_ptr) :: flux = c_null_ptr
  end type m3drt_b1d_batch

  type, bind(C), public :: m3drt_b1d_tile
    integer(c_int32_t) :: fpl
    integer(c_int32_t) :: threads
  end type m3drt_b1d_tile

  type, bind(C), public :: m3drt_background_lte_args
    integer(c_int32_t) :: nf
    integer(c_int64_t) :: ncell
    type(c_ptr) :: nu = c_null_ptr
    type(c_ptr) :: temp = c_null_ptr
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: Bnu = c_null_ptr
    integer(c_int64_t) :: sf = 0
    integer(c_int64_t) :: sc = 0
  end type m3drt_background_lte_args

  type, bind(C), public :: m3drt_boundary_args
    integer(c_int32_t) :: kind
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: nx
    integer(c_int32_t) :: ny
    integer(c_int64_t) :: ps_f = 0
    integer(c_int64_t) :: ps_x = 0
    integer(c_int64_t) :: ps_y = 0
    type(c_ptr) :: chi1 = c_null_ptr
    type(c_ptr) :: S1 = c_null_ptr
    type(c_ptr) :: chi2 = c_null_ptr
    type(c_ptr) :: S2 = c_null_ptr
    type(c_ptr) :: g500 = c_null_ptr
    type(c_ptr) :: chi500_1 = c_null_ptr
    type(c_ptr) :: T_ghost = c_null_ptr
    type(c_ptr) :: nu = c_null_ptr
    real(c_double) :: ds3 = 0
    type(c_ptr) :: exit_I = c_null_ptr
    type(c_ptr) :: mem_chi = c_null_ptr
    type(c_ptr) :: mem_S = c_null_ptr
    type(c_ptr) :: mem_I = c_null_ptr
    type(c_ptr) :: ghost_chi = c_null_ptr
    type(c_ptr) :: ghost_S = c_null_ptr
    type(c_ptr) :: ghost_I = c_null_ptr
  end type m3drt_boundary_args

  type, bind(C), public :: m3drt_quadrature_args
    type(c_ptr) :: scheme = c_null_ptr
    integer(c_int32_t) :: ntheta
    integer(c_int32_t) :: nphi
    integer(c_int32_t) :: jnu_weights
    integer(c_int32_t) :: one_dimensional
    type(c_ptr) :: mu = c_null_ptr
    real(c_double) :: ds(3)
  end type m3drt_quadrature_args

  type, bind(C), public :: m3drt_optab_table
    integer(c_int32_t) :: nnu
    integer(c_int32_t) :: ntemp
    integer(c_int32_t) :: nne
    real(c_double) :: mu(2)
    real(c_double) :: rot(4)
    real(c_double) :: c0(2)
    real(c_double) :: dc(2)
    type(c_ptr) :: lam = c_null_ptr
    type(c_ptr) :: val = c_null_ptr
    type(c_ptr) :: scat = c_null_ptr
  end type m3drt_optab_table

  type, bind(C), public :: m3drt_optab_query
    integer(c_int64_t) :: ncell
    integer(c_int32_t) :: nf
    type(c_ptr) :: temp = c_null_ptr
    type(c_ptr) :: ne = c_null_ptr
    type(c_ptr) :: chi500 = c_null_ptr
    type(c_ptr) :: lam_q = c_null_ptr
    type(c_ptr) :: back = c_null_ptr
    type(c_ptr) :: scat = c_null_ptr
    integer(c_int64_t) :: sf = 0
    integer(c_int64_t) :: sc = 0
  end type m3drt_optab_query

  type, bind(C), public :: m3drt_isotope
    type(m3drt_linelist) :: lines
    real(c_double) :: awgt
    real(c_double) :: frac
  end type m3drt_isotope

  type, bind(C), public :: m3drt_species
    integer(c_int32_t) :: niso
    type(c_ptr) :: iso = c_null_ptr
    type(c_ptr) :: n = c_null_ptr
    real(c_double) :: scale = 1
    type(c_ptr) :: part = c_null_ptr
    type(c_ptr) :: dep = c_null_ptr
    integer(c_int32_t) :: nlev = 0
    integer(c_int32_t) :: scat_group = 0
  end type m3drt_species

  type, bind(C), public :: m3drt_linelist_opacity_args
    integer(c_int32_t) :: nf
    integer(c_int32_t) :: nsp
    integer(c_int32_t) :: equidistant
    integer(c_int64_t) :: ncell
    type(c_ptr) :: llam = c_null_ptr
    type(c_ptr) :: sp = c_null_ptr
    type(c_ptr) :: temp = c_null_ptr
    type(c_ptr) :: vmic = c_null_ptr
    type(c_ptr) :: vproj = c_null_ptr
    type(c_ptr) :: e_kT = c_null_ptr
    type(c_ptr) :: ne = c_null_ptr
    type(c_ptr) :: tothi = c_null_ptr
    type(c_ptr) :: stark_pe = c_null_ptr
    type(c_ptr) :: all_chi = c_null_ptr
    type(c_ptr) :: all_eta = c_null_ptr
    type(c_ptr) :: one_m_eps = c_null_ptr
    type(c_ptr) :: chi_scat = c_null_ptr
    type(c_ptr) :: beta = c_null_ptr
    type(c_ptr) :: chi_lines = c_null_ptr
    type(c_ptr) :: eta_lines = c_null_ptr
    real(c_float) :: tolerance = 1e-4
    real(c_float) :: line_reject_frac = 0
  end type m3drt_linelist_opacity_args

  public :: m3drt_create, m3drt_destroy, m3drt_synchronize, m3drt_last_error
  public :: m3drt_device_alloc, m3drt_device_free, m3drt_copy_to_device, m3drt_copy_to_host
  public :: m3drt_sc_sweep, m3drt_reference_strides, m3drt_native_strides
  public :: m3drt_lc_sweep_hermite, m3drt_lc_sweep_hermite2
  public :: m3drt_sc_sweep_hermite, m3drt_sc_adjoint, m3drt_line_opacity_equi, m3drt_line_opacity
  public :: m3drt_window_opacity, m3drt_rotate_array
  public :: m3drt_bb_opacity, m3drt_bb_rates, m3drt_bf_rates, m3drt_work_to_transfer
  public :: m3drt_convert_to_work_units, m3drt_background_emission, m3drt_update_eta_mask
  public :: m3drt_normalize_rates, m3drt_rate_solve_f64
  public :: m3drt_b1d_synth_plain, m3drt_b1d_plan_create, m3drt_b1d_plan_run, m3drt_b1d_plan_info
  public :: m3drt_b1d_plan_destroy
  public :: m3drt_lte_background, m3drt_set_upstream_boundary, m3drt_set_downstream_boundary
  public :: m3drt_quadrature, m3drt_linear_weights, m3drt_parab_moments
  public :: m3drt_optab_plan_create, m3drt_optab_slopes, m3drt_optab_interpolate, m3drt_optab_plan_destroy
  public :: m3drt_add_linelist_opacities, m3drt_sc_sweep_noalloc
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
    integer(c_int) function m3drt_sc_sweep_hermite(ctx, desc, bufs, fields) bind(C, name='m3drt_sc_sweep_hermite')
      import :: c_int, c_ptr, m3drt_sweep_desc, m3drt_sweep_bufs, m3drt_hermite_fields
      type(c_ptr), value :: ctx
      type(m3drt_sweep_desc), intent(in) :: desc
      type(m3drt_sweep_bufs), intent(in) :: bufs
      type(m3drt_hermite_fields), intent(in) :: fields
    end function m3drt_sc_sweep_hermite
    integer(c_int) function m3drt_lc_sweep_hermite(ctx, desc, bufs, fields, ghost) &
        bind(C, name='m3drt_lc_sweep_hermite')
      import :: c_int, c_ptr, m3drt_sweep_desc, m3drt_sweep_bufs, m3drt_hermite_fields, m3drt_lc_ghost
      type(c_ptr), value :: ctx
      type(m3drt_sweep_desc), intent(in) :: desc
      type(m3drt_sweep_bufs), intent(in) :: bufs
      type(m3drt_hermite_fields), intent(in) :: fields
      type(m3drt_lc_ghost), intent(in) :: ghost
    end function m3drt_lc_sweep_hermite
    integer(c_int) function m3drt_lc_sweep_hermite2(ctx, desc, bufs, fields, ghost) &
        bind(C, name='m3drt_lc_sweep_hermite2')
      import :: c_int, c_ptr, m3drt_sweep_desc, m3drt_sweep_bufs, m3drt_hermite_fields, m3drt_lc_ghost
      type(c_ptr), value :: ctx
      type(m3drt_sweep_desc), intent(in) :: desc
      type(m3drt_sweep_bufs), intent(in) :: bufs
      type(m3drt_hermite_fields), intent(in) :: fields
      type(m3drt_lc_ghost), intent(in) :: ghost
    end function m3drt_lc_sweep_hermite2
    integer(c_int) function m3drt_sc_adjoint(ctx, desc, bufs) bind(C, name='m3drt_sc_adjoint')
      import :: c_int, c_ptr, m3drt_sweep_desc, m3drt_adjoint_bufs
      type(c_ptr), value :: ctx
      type(m3drt_sweep_desc), intent(in) :: desc
      type(m3drt_adjoint_bufs), intent(in) :: bufs
    end function m3drt_sc_adjoint
    integer(c_int) function m3drt_line_opacity_equi(ctx, ll, args) bind(C, name='m3drt_line_opacity_equi')
      import :: c_int, c_ptr, m3drt_linelist, m3drt_opacity_args
      type(c_ptr), value :: ctx
      type(m3drt_linelist), intent(in) :: ll
      type(m3drt_opacity_args), intent(in) :: args
    end function m3drt_line_opacity_equi
    integer(c_int) function m3drt_line_opacity(ctx, ll, args) bind(C, name='m3drt_line_opacity')
      import :: c_int, c_ptr, m3drt_linelist, m3drt_opacity_args
      type(c_ptr), value :: ctx
      type(m3drt_linelist), intent(in) :: ll
      type(m3drt_opacity_args), intent(in) :: args
    end function m3drt_line_opacity
    integer(c_int) function m3drt_bb_opacity(ctx, args) bind(C, name='m3drt_bb_opacity')
      import :: c_int, c_ptr, m3drt_bb_opacity_args
      type(c_ptr), value :: ctx
      type(m3drt_bb_opacity_args), intent(in) :: args
    end function m3drt_bb_opacity
    integer(c_int) function m3drt_bb_rates(ctx, args) bind(C, name='m3drt_bb_rates')
      import :: c_int, c_ptr, m3drt_bb_rates_args
      type(c_ptr), value :: ctx
      type(m3drt_bb_rates_args), intent(in) :: args
    end function m3drt_bb_rates
    integer(c_int) function m3drt_bf_rates(ctx, args) bind(C, name='m3drt_bf_rates')
      import :: c_int, c_ptr, m3drt_bf_rates_args
      type(c_ptr), value :: ctx
      type(m3drt_bf_rates_args), intent(in) :: args
    end function m3drt_bf_rates
    integer(c_int) function m3drt_work_to_transfer(ctx, args) bind(C, name='m3drt_work_to_transfer')
      import :: c_int, c_ptr, m3drt_units_args
      type(c_ptr), value :: ctx
      type(m3drt_units_args), intent(in) :: args
    end function m3drt_work_to_transfer
    integer(c_int) function m3drt_convert_to_work_units(ctx, n, chi_cont, xhat, ehat, beta) &
        bind(C, name='m3drt_convert_to_work_units')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr), value :: chi_cont, xhat, ehat, beta
    end function m3drt_convert_to_work_units
    integer(c_int) function m3drt_background_emission(ctx, args) bind(C, name='m3drt_background_emission')
      import :: c_int, c_ptr, m3drt_background_args
      type(c_ptr), value :: ctx
      type(m3drt_background_args), intent(in) :: args
    end function m3drt_background_emission
    integer(c_int) function m3drt_update_eta_mask(ctx, args) bind(C, name='m3drt_update_eta_mask')
      import :: c_int, c_ptr, m3drt_eta_args
      type(c_ptr), value :: ctx
      type(m3drt_eta_args), intent(in) :: args
    end function m3drt_update_eta_mask
    integer(c_int) function m3drt_normalize_rates(ctx, args, wphi) bind(C, name='m3drt_normalize_rates')
      import :: c_int, c_ptr, m3drt_rate_solve_args
      type(c_ptr), value :: ctx
      type(m3drt_rate_solve_args), intent(in) :: args
      type(c_ptr), value :: wphi
    end function m3drt_normalize_rates
    integer(c_int) function m3drt_rate_solve_f64(ctx, args, res) bind(C, name='m3drt_rate_solve_f64')
      import :: c_int, c_ptr, m3drt_rate_solve_args, m3drt_rate_solve_result
      type(c_ptr), value :: ctx
      type(m3drt_rate_solve_args), intent(in) :: args
      type(m3drt_rate_solve_result), intent(inout) :: res
    end function m3drt_rate_solve_f64
    integer(c_int) function m3drt_lte_background(ctx, args) bind(C, name='m3drt_lte_background')
      import :: c_int, c_ptr, m3drt_background_lte_args
      type(c_ptr), value :: ctx
      type(m3drt_background_lte_args), intent(in) :: args
    end function m3drt_lte_background
    integer(c_int) function m3drt_set_upstream_boundary(ctx, args) bind(C, name='m3drt_set_upstream_boundary')
      import :: c_int, c_ptr, m3drt_boundary_args
      type(c_ptr), value :: ctx
      type(m3drt_boundary_args), intent(in) :: args
    end function m3drt_set_upstream_boundary
    integer(c_int) function m3drt_set_downstream_boundary(ctx, args) bind(C, name='m3drt_set_downstream_boundary')
      import :: c_int, c_ptr, m3drt_boundary_args
      type(c_ptr), value :: ctx
      type(m3drt_boundary_args), intent(in) :: args
    end function m3drt_set_downstream_boundary
    ! vec(3, cap) and wts(cap) are host arrays; scheme a NUL-terminated string
    integer(c_int) function m3drt_quadrature(q, cap, vec, wts, nray) bind(C, name='m3drt_quadrature')
      import :: c_int, c_int32_t, c_double, m3drt_quadrature_args
      type(m3drt_quadrature_args), intent(in) :: q
      integer(c_int32_t), value :: cap
      real(c_double) :: vec(*), wts(*)
      integer(c_int32_t), intent(out) :: nray
    end function m3drt_quadrature
    integer(c_int) function m3drt_linear_weights(ctx, n, t, E, W, Psi, Esc) bind(C, name='m3drt_linear_weights')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr), value :: t, E, W, Psi, Esc
    end function m3drt_linear_weights
    integer(c_int) function m3drt_parab_moments(ctx, n, t, E, U0, U1, q) bind(C, name='m3drt_parab_moments')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      type(c_ptr), value :: t, E, U0, U1, q
    end function m3drt_parab_moments
    integer(c_int) function m3drt_optab_plan_create(ctx, table, plan) bind(C, name='m3drt_optab_plan_create')
      import :: c_int, c_ptr, m3drt_optab_table
      type(c_ptr), value :: ctx
      type(m3drt_optab_table), intent(in) :: table
      type(c_ptr), intent(out) :: plan
    end function m3drt_optab_plan_create
    integer(c_int) function m3drt_optab_slopes(plan, dl, dt, de, sl, st, se) bind(C, name='m3drt_optab_slopes')
      import :: c_int, c_ptr
      type(c_ptr), value :: plan, dl, dt, de, sl, st, se
    end function m3drt_optab_slopes
    integer(c_int) function m3drt_optab_interpolate(plan, query) bind(C, name='m3drt_optab_interpolate')
      import :: c_int, c_ptr, m3drt_optab_query
      type(c_ptr), value :: plan
      type(m3drt_optab_query), intent(in) :: query
    end function m3drt_optab_interpolate
    integer(c_int) function m3drt_optab_plan_destroy(plan) bind(C, name='m3drt_optab_plan_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function m3drt_optab_plan_destroy
    ! args%sp points to a host array of m3drt_species, each %iso to a host array of m3drt_isotope
    integer(c_int) function m3drt_add_linelist_opacities(ctx, args) bind(C, name='m3drt_add_linelist_opacities')
      import :: c_int, c_ptr, m3drt_linelist_opacity_args
      type(c_ptr), value :: ctx
      type(m3drt_linelist_opacity_args), intent(in) :: args
    end function m3drt_add_linelist_opacities
    integer(c_int) function m3drt_sc_sweep_noalloc(ctx, desc, bufs) bind(C, name='m3drt_sc_sweep_noalloc')
      import :: c_int, c_ptr, m3drt_sweep_desc, m3drt_sweep_bufs
      type(c_ptr), value :: ctx
      type(m3drt_sweep_desc), intent(in) :: desc
      type(m3drt_sweep_bufs), intent(in) :: bufs
    end function m3drt_sc_sweep_noalloc
    integer(c_int) function m3drt_rotate_array(ctx, args) bind(C, name='m3drt_rotate_array')
      import :: c_int, c_ptr, m3drt_rotate_args
      type(c_ptr), value :: ctx
      type(m3drt_rotate_args), intent(in) :: args
    end function m3drt_rotate_array
    integer(c_int) function m3drt_window_opacity(ctx, args, win) bind(C, name='m3drt_window_opacity')
      import :: c_int, c_ptr, m3drt_bb_opacity_args, m3drt_window_args
      type(c_ptr), value :: ctx
      type(m3drt_bb_opacity_args), intent(in) :: args
      type(m3drt_window_args), intent(in) :: win
    end function m3drt_window_opacity
    integer(c_int) function m3drt_b1d_synth_plain(ctx, batch, tile) bind(C, name='m3drt_b1d_synth_plain')
      import :: c_int, c_ptr, m3drt_b1d_batch
      type(c_ptr), value :: ctx
      type(m3drt_b1d_batch), intent(in) :: batch
      type(c_ptr), value :: tile
    end function m3drt_b1d_synth_plain
    integer(c_int) function m3drt_b1d_plan_create(ctx, batch, tile, plan) bind(C, name='m3drt_b1d_plan_create')
      import :: c_int, c_ptr, m3drt_b1d_batch
      type(c_ptr), value :: ctx
      type(m3drt_b1d_batch), intent(in) :: batch
      type(c_ptr), value :: tile
      type(c_ptr), intent(out) :: plan
    end function m3drt_b1d_plan_create
    integer(c_int) function m3drt_b1d_plan_run(plan) bind(C, name='m3drt_b1d_plan_run')
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function m3drt_b1d_plan_run
    integer(c_int) function m3drt_b1d_plan_info(plan, tile, times, ntimes) bind(C, name='m3drt_b1d_plan_info')
      import :: c_int, c_ptr, c_double, c_int32_t, m3drt_b1d_tile
      type(c_ptr), value :: plan
      type(m3drt_b1d_tile), intent(out) :: tile
      real(c_double), intent(out) :: times(*)
      integer(c_int32_t), intent(out) :: ntimes
    end function m3drt_b1d_plan_info
    integer(c_int) function m3drt_b1d_plan_destroy(plan) bind(C, name='m3drt_b1d_plan_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function m3drt_b1d_plan_destroy
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

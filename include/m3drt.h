/*
 * m3drt.h -- C ABI of the MI355X radiative-transfer kernels (libm3drt.so).
 *
 * This is the drop-in boundary for OpenMULTI3D's radiative-transfer hot path
 * (experiments/OpenMULTI3D/ of Bergemann-Group-development/dispatch).  Each entry
 * point names the reference procedure it replaces.  Conventions follow the
 * reference's own C ABI, HOP/standalone/hop_c_api.f90:1-18: every call returns an
 * int status (0 ok, >0 error), the message of the last error is read with
 * m3drt_last_error, arrays are allocated by the caller, and nothing is retained
 * by reference after a call returns.
 *
 * Array arguments are DEVICE pointers (memory of the context's device), unless a
 * helper below says otherwise.  Cubes are addressed through element strides
 * (f, x, y, z), so the reference's Fortran layout work(f,x,y,z) (f fastest) and
 * the native frequency-sliced layout (x fastest, then y, then f, then z) are both
 * accepted; the native one is the fast path on MI355X (see m3drt_native_strides).
 * chi500-like per-cell arrays are (x, y, z) with x fastest, as in the reference.
 */
#ifndef M3DRT_H
#define M3DRT_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define M3DRT_OK 0
#define M3DRT_EINVAL 1    /* invalid argument or configuration */
#define M3DRT_EDEVICE 2   /* HIP runtime error */
#define M3DRT_ENOMEM 3    /* device allocation failed */
#define M3DRT_ELIMIT 4    /* size beyond a documented kernel limit */

/* element precision of the floating-point cubes of a call */
#define M3DRT_F32 0
#define M3DRT_F64 1

/* formal solvers (rt_iteration_params solver=..., SRC/rt_params_mod.f90) */
#define M3DRT_LINEAR2 0   /* short_char2 in 3D, long_char2 in 1D (the default) */
#define M3DRT_QUADRATIC 1 /* long_char_quadratic along every column (the output pass, after
                             rotation for inclined rays; SRC/rt_solvers_mod.f90:891-1020), FP32 */
#define M3DRT_LINEAR 2    /* short_char, the 'linear' Multi3D-comparison preset: chi, S, I all
                             from the 4x4 stencil (SRC/rt_solvers_mod.f90:19-189), FP32 */
#define M3DRT_LONG_LINEAR 3 /* long_char_linear along every column: the trapezoid step with the
                               linear weights (SRC/rt_solvers_mod.f90:666-757; the 'linear'
                               output solver and 1D iteration solver), FP32, split rates */

/* sweep flags */
#define M3DRT_SPLIT_RATES 1  /* fill Irem and escape ('largest' rate formulation) */
#define M3DRT_GET_RATES 2    /* form Psi (Lambda* diagonal / chi) for Psi_out / PsiJ */

typedef struct m3drt_ctx m3drt_ctx;

/* Context: one per device and HIP stream.  flags reserved (0).  The reference has
 * no counterpart (its kernels are Fortran procedures of rt_patch_t); HOP's
 * hop_init (HOP/standalone/libhop.h) is the analogue. */
int m3drt_create(m3drt_ctx **ctx, int device, int flags);
int m3drt_destroy(m3drt_ctx *ctx);
/* Runs the context's work on an existing hipStream_t (NULL: the null stream). */
int m3drt_set_stream(m3drt_ctx *ctx, void *hip_stream);
int m3drt_synchronize(m3drt_ctx *ctx);
/* Copies the last error message of the calling thread; as hop_last_error,
 * HOP/standalone/libhop.h. */
int m3drt_last_error(char *buf, int buflen);

/* Device memory helpers for hosts without HIP bindings (the Fortran side). */
int m3drt_device_alloc(m3drt_ctx *ctx, size_t bytes, void **ptr);
int m3drt_device_free(m3drt_ctx *ctx, void *ptr);
int m3drt_copy_to_device(m3drt_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int m3drt_copy_to_host(m3drt_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);

/* Geometry and parameters of one directional sweep through a column of slabs. */
typedef struct {
  int32_t nf, nx, ny, nz;  /* frequencies, plane size, planes swept */
  int32_t dir;             /* +1: z increasing (up), -1: down (rt_patch_t%dir) */
  int32_t prec;            /* M3DRT_F32 or M3DRT_F64 for every float cube */
  int32_t solver;          /* M3DRT_LINEAR2, M3DRT_QUADRATIC, M3DRT_LINEAR or M3DRT_LONG_LINEAR */
  int32_t flags;           /* M3DRT_SPLIT_RATES | M3DRT_GET_RATES */
  double ds[3];            /* cell size in cm (self%ds * scaling%l) */
  double vec[3];           /* ray direction (rt_iter%vec(:,i_mu)) */
  double maser_dtau_max;   /* &nlte_params, default 2 (SRC/m3d_patch_mod.f90:319) */
  double maser_Inu_max;    /* default 1e30 (SRC/m3d_patch_mod.f90:320) */
  double wj;               /* weight of J += wj*I and PsiJ += wj*Psi (2*wmu) */
  int64_t cs_f, cs_x, cs_y, cs_z;  /* element strides of every cube */
  int64_t gs_f, gs_x, gs_y;        /* element strides of the ghost and exit planes */
  int64_t eta_ld;                  /* leading dimension of eta_mask (>= nf) */
  int64_t eta_nrow;                /* rows of eta_mask; eta_row values lie in 0..eta_nrow */
} m3drt_sweep_desc;

/* Buffers of one sweep.  Optional pointers may be NULL. */
typedef struct {
  const void *chi, *S;          /* (nf,nx,ny,nz) opacity and source function */
  const void *ghost_chi;        /* upstream ghost plane (nf,nx,ny): work(:,:,:,0|nz+1,1:3) */
  const void *ghost_S;
  const void *ghost_I;
  const float *chi500;          /* (nx,ny,nz) composition%chi500 */
  const float *chi500_up;       /* (nx,ny,nz) chi500 at the upwind point of this direction */
  const float *g500;            /* (nx,ny) chi500 of the ghost plane */
  const int32_t *eta_row;       /* (nx,ny,nz) row of the cell's eta flags, 0 none; or NULL */
  const uint8_t *eta_mask;      /* (eta_ld, eta_nrow) logical(1) flags of the eta step */
  void *J, *PsiJ;               /* accumulated in place (integrate_intensity) */
  void *I_out, *Psi_out;        /* full cubes of Inu and Psi/chi */
  void *Irem_out, *escape_out;  /* with M3DRT_SPLIT_RATES */
  void *exit_I;                 /* (nf,nx,ny) Inu on the last plane (exit plane / emergent) */
} m3drt_sweep_bufs;

/*
 * One formal solution along direction vec through nz planes, fused with the
 * angle integration of J and PsiJ.  Replaces, per direction:
 *   rt_patch_t%short_char2   SRC/rt_solvers_mod.f90:224-507 (nx*ny > 1)
 *   rt_patch_t%long_char2    SRC/rt_solvers_mod.f90:527-663 (nx = ny = 1)
 *   and with desc->solver: short_char (M3DRT_LINEAR), long_char_quadratic
 *   (M3DRT_QUADRATIC) and long_char_linear (M3DRT_LONG_LINEAR, :666-757) along columns
 *   rt_patch_t%integrate_intensity  SRC/rt_patch_mod.f90:3399-3431 (J, PsiJ)
 * over a whole column of slabs (rt_column_t%execute, SRC/rt_column_mod.f90:252-321
 * hands the exit plane downstream; here the sweep continues through it).
 * No amr grading (amr='').
 * Limits (M3DRT_ELIMIT): nx <= 1024; the plane must fit the 160 KB LDS with the
 * intensity at least (nx*ny*sizeof(elem) <= 160 KB).
 */
int m3drt_sc_sweep(m3drt_ctx *ctx, const m3drt_sweep_desc *desc, const m3drt_sweep_bufs *bufs);

/* m3drt_sc_sweep without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same checks and kernels as m3drt_sc_sweep,
 * with the per-cell scratch of the sweep (8 bytes per cell and plane of the
 * straight-line kernel, 2 x 4 or 2 x 8 of the general one) in a fixed 32 MiB device
 * array the library holds on every device it runs on until the process ends; a
 * call that needs more is refused with M3DRT_ELIMIT before any launch (the
 * 120 x 120 x 230 FP32 box takes 26.6 MB).  Every call on one device shares the
 * array: calls that overlap on one device (two streams or threads) race. */
int m3drt_sc_sweep_noalloc(m3drt_ctx *ctx, const m3drt_sweep_desc *desc, const m3drt_sweep_bufs *bufs);

/* ------------------------------------------------------------------ adjoint sweep */

/* Buffers of one direction's adjoint (reverse) sweep.  Layouts and strides as the
 * forward's (m3drt_sweep_desc).  Optional pointers may be NULL. */
typedef struct {
  const float *chi, *S;                  /* the forward's inputs */
  const float *ghost_chi, *ghost_S, *ghost_I;
  const float *chi500, *chi500_up, *g500;
  const float *I;                        /* the forward's intensity cube (its I_out) */
  const float *Jbar;                     /* cotangent of J (the forward added wj*I to J) */
  const float *Ibar;                     /* cotangent of the intensity cube itself */
  const float *exit_bar;                 /* (nf,nx,ny) cotangent of exit_I */
  float *chibar, *Sbar;                  /* accumulated (+=): cotangents of chi and S */
  float *ghost_chibar, *ghost_Sbar, *ghost_Ibar;  /* accumulated: of the upstream ghost plane */
} m3drt_adjoint_bufs;

/*
 * The reverse-mode derivative of m3drt_sc_sweep's short_char2 step (solver
 * M3DRT_LINEAR2, FP32, no eta flags, no split rates): given the cotangents of J,
 * of the intensity cube and of the exit plane, accumulates those of chi and S (and
 * of the upstream ghost plane), chi500 held fixed.  The reference has no adjoint;
 * this serves the abundance and temperature fits of the Python front end.  Limits
 * (M3DRT_ELIMIT): 13 bytes per plane cell must fit the LDS (nx*ny <= 12600).
 */
int m3drt_sc_adjoint(m3drt_ctx *ctx, const m3drt_sweep_desc *desc, const m3drt_adjoint_bufs *bufs);

/* ------------------------------------------------------------------ line opacity */

/* flags of m3drt_opacity_args */
#define M3DRT_OPAC_NLTE 1   /* lines with both levels in dep take b_i, b_j (do_nlte) */
#define M3DRT_OPAC_SCAT 2   /* LTE lines also add (1-eps)*K to chi_scat (do_scat) */

/* One isotope's line list, structure of arrays, ascending in llam0
 * (linelist_type, HOP/line_types_mod.f90:22-74).  Device pointers. */
typedef struct {
  int32_t nline;
  const double *llam0;                /* ln(lambda0) of the line centres */
  const float *lower_ev, *gpref;      /* E_low [eV]; loggf*ln10 + log(const/nu0) */
  const float *grad, *gw1, *gw2, *gstk;  /* damping recipes, scaled by 1/(4 pi nu0) */
  const int32_t *unsold;              /* nonzero: Unsold van der Waals, else Barklem */
  const int32_t *lev_i, *lev_j;       /* 1-based levels in dep, 0: none */
  const int32_t *ipart;               /* 1-based population group of each line; NULL: 1 */
  int32_t *used;                      /* out (may be NULL): set where the line deposits */
  float *strength, *reach;            /* out (may be NULL): maxima over the cells */
  /* select_lines (HOP/linelist_mod.f90:1110-1142).  reach_in: every line's reach in
   * ln(lambda) (the list's has_reach: pass it only when every line carries one, as
   * finalize_linelist sets it, :375-378); then the lines are narrowed by bisection on
   * the centres with the largest reach as margin and compared one by one with their
   * own, each margin widened by max |vproj| of the call's cells and half a grid step,
   * and a line centred beyond the grid takes its tolerance reference at the nearer
   * grid end with its window capped by its reach (:594-602, 660-667, 844-858,
   * 914-921).  active: nonzero for the lines to visit (NULL: all).  sel_work: device
   * scratch of nline + 1 int32 the selection is written to, required with reach_in
   * or active. */
  const float *reach_in;
  const int32_t *active;
  int32_t *sel_work;
} m3drt_linelist;

/* The cells and spectra of one call.  (f, cell) arrays are addressed f*sf + cell*sc
 * (elements); per-cell arrays are (ncell); lnbase (ncell, npart) and dep (ncell,
 * nlev) cell-fastest, as the reference's (m1,m2,m3,n) arrays. */
typedef struct {
  int32_t nf, ncell, npart, nlev, flags;
  float tolerance;                    /* bbtrans_mod tolerance, default 1e-4 */
  float line_reject_frac;             /* default 0 */
  int64_t sf, sc;
  const double *llam;                 /* (nf) ascending ln(lambda) grid, equidistant */
  const float *vproj, *inv_df, *vbar2, *e_kT, *ne, *tothi, *stark_pe;
  const float *lnbase;                /* ln n - ln U + ln inv_df per population group */
  const float *dep;                   /* departure coefficients (NLTE) */
  float *chi, *eta;                   /* accumulated line opacity and emissivity */
  const float *cont, *logref;         /* continuum and ln(tol cont / voigt_peak) */
  const float *one_m_eps;             /* scattering fraction (M3DRT_OPAC_SCAT) */
  float *chi_scat;                    /* accumulated scattering opacity (M3DRT_OPAC_SCAT) */
} m3drt_opacity_args;

/*
 * Adds the line opacity of one isotope's line list to chi and eta with the
 * QLG-bot Voigt kernel.  Replaces add_opacity_equi, HOP/linelist_mod.f90:729-960
 * (without reach-based chunking), as add_particle_opacity (HOP/linelist_mod.f90:
 * 970-1070) calls it per isotope.
 */
int m3drt_line_opacity_equi(m3drt_ctx *ctx, const m3drt_linelist *ll, const m3drt_opacity_args *o);

/*
 * add_opacity (HOP/linelist_mod.f90:469-719): the same summation on an ascending
 * ln(lambda) grid of any spacing (spectrum%equidistant false): per cell a forward
 * cursor finds each line centre's nearest node, the profile is set up on the
 * ln(lambda) abscissa, and its window is found by bisection (bisect_nearest,
 * :1075-1095).  Every line of the list is visited (select_lines without reaches).
 * Arguments as m3drt_line_opacity_equi; nf >= 1.
 */
int m3drt_line_opacity(m3drt_ctx *ctx, const m3drt_linelist *ll, const m3drt_opacity_args *o);

/* ------------------------------------------------------------------ statistical equilibrium */

/* A model atom and its cells, FP64 except the radiative rate sums.  Per-cell
 * arrays are cell-fastest: nstar[i*ncell + c] (the reference's (x,y,z,level)),
 * gamma_ij[kr*ncell + c], cup[q*ncell + c].  Levels and transitions 1-based. */
typedef struct {
  int32_t nk, nrad, nline, npair, ncell;   /* levels, transitions (lines first), pairs, cells */
  const int32_t *trans_i, *trans_j;        /* (nrad) lower, upper level */
  const int32_t *trans_active;             /* (nrad) */
  const double *g, *ev;                    /* (nk) statistical weight, energy [eV] */
  const double *Aji;                       /* (nline) Einstein A, the pair of a cut line */
  const int32_t *pair_i, *pair_j;          /* (npair) collisional pairs i < j */
  const double *e_kT, *temp, *totn;        /* (ncell) 1 eV/kT, T [K], atoms per cm^3 */
  const double *nstar;                     /* (nk, ncell) LTE populations */
  double *b;                               /* (nk, ncell) departure coefficients, updated */
  float *gamma_ij, *gamma_ji;              /* (nrad, ncell) rate sums (thermal, deviation for lines) */
  const double *cup;                       /* (npair, ncell) upward collision rates [1/s]; or NULL */
  double max_update;                       /* &nlte_params defaults: 1e2 */
  double min_pop_frac;                     /* 1e-30 */
  double conv_pop_floor;                   /* 0 */
  double sor_w;                            /* 1 */
  double cswitch;                          /* 1 */
  double cut_x;                            /* 70 */
  double bf_cut_rate;                      /* 1 [1/s] */
  double tmax_update;                      /* 1e7 K */
} m3drt_rate_solve_args;

typedef struct {
  double maxerr;      /* signed largest (weighted) relative change of any level */
  int32_t epos;       /* its cell (0-based, x fastest), -1 if none */
  int32_t elvl;       /* its level (0-based) */
  int64_t nclamp;     /* corrections cut by the limiter */
  double *lvlerr;     /* host array (nk) of each level's signed largest change, or NULL */
} m3drt_rate_solve_result;

/*
 * One statistical-equilibrium update of every cell: the rate matrix of
 * setup_transition_matrix (SRC/atom_mod.f90:740-851) with line_rate_coefficients
 * (862-884), solved by lineq_no_pivot (HOP/op_math_mod.f90:420-449), then
 * update_populations' limiter, under-relaxation, conservation and floor
 * (SRC/atom_mod.f90:439-732).  Returns the convergence measure the NLTE task
 * compares with convlim (SRC/nlte_patch_mod.f90:711-770).  A cell without a
 * finite solution fails the call with its index, as the reference stops there.
 * Synchronizes the context's stream to return the result.
 */
int m3drt_rate_solve_f64(m3drt_ctx *ctx, const m3drt_rate_solve_args *args, m3drt_rate_solve_result *res);

/* normalize_profiles (SRC/atom_mod.f90:377-431): line rate sums times the inverted
 * profile integrals wphi (nline, ncell; NULL: none), continuum downward sums times
 * n*_i/n*_j.  Uses trans_*, nline, nrad, ncell, nstar, gamma_ij/ji of args. */
int m3drt_normalize_rates(m3drt_ctx *ctx, const m3drt_rate_solve_args *args, const float *wphi);

/* ------------------------------------------------------------------ quadratic 3D sweep */

/* The composition fields short_char_hermite reads besides chi, S and I. */
typedef struct {
  const float *beta;                        /* cube (the sweep's layout): stimulated-emission factor */
  const float *pe4, *temp;                  /* (nz, ny, nx) electron pressure, temperature */
  const float *dpe_dx, *dpe_dy, *dpe_dz;    /* (nz, ny, nx) gradients of Pe [per cm] */
  const float *dlnt_dx, *dlnt_dy, *dlnt_dz; /* (nz, ny, nx) gradients of ln T [per cm] */
  const float *hknu;                        /* (nf) h nu / k */
} m3drt_hermite_fields;

/* short_char_hermite (SRC/rt_solvers_mod.f90:1917-2248) + integrate_intensity: the 3D
 * short-characteristics sweep with quadratic (parabolic-moment) weights and the
 * product-rule slopes of chi (through Pe) and S (through B).  FP32 (FP64 step length
 * and weights), as the reference.  desc->solver is ignored. */
int m3drt_sc_sweep_hermite(m3drt_ctx *ctx, const m3drt_sweep_desc *desc, const m3drt_sweep_bufs *bufs,
                           const m3drt_hermite_fields *fields);

/* ------------------------------------------------------------------ cubic-Bezier column sweep */

/* The boundary below the first plane of long_char_hermite (SRC/rt_solvers_mod.f90:1100-1130). */
#define M3DRT_LCH_TOP 0      /* the zeroed top ghost (chi = 0): no thermodynamic state there; the
                                first interval takes secant slopes at its upwind end (herm_first) */
#define M3DRT_LCH_BOTTOM 1   /* the LTE ghost below the domain with the model's Pe and T there
                                (bottom_ghost_at, ghost_thermo) */
#define M3DRT_LCH_UPSTREAM 2 /* the last plane of an upstream neighbour (comp_up) */

/* Thermodynamics of the ghost plane, (ny, nx) x fastest; unused for M3DRT_LCH_TOP. */
typedef struct {
  int32_t kind;                             /* M3DRT_LCH_TOP, _BOTTOM or _UPSTREAM */
  const float *pe, *dpe_dz, *temp, *dlnt_dz;
} m3drt_lc_ghost;

/* long_char_hermite (SRC/rt_solvers_mod.f90:1052-1504) + integrate_intensity
 * (SRC/rt_patch_mod.f90:3399-3431): the cubic-Bezier long-characteristics solver along
 * every column (the output pass; inclined rays after m3drt_rotate_array), chi's
 * slopes by the product rule through Pe, the source's through B_nu and S/B_nu, the
 * control points clamped to [min, max] of the end values.  FP32, as the reference.
 * Of fields, beta, pe4, temp, dpe_dz, dlnt_dz and hknu are read (the x, y gradients
 * may be NULL).  Needs nz >= 2; desc->solver is ignored; no split rates or eta flags. */
int m3drt_lc_sweep_hermite(m3drt_ctx *ctx, const m3drt_sweep_desc *desc, const m3drt_sweep_bufs *bufs,
                           const m3drt_hermite_fields *fields, const m3drt_lc_ghost *ghost);

/* long_char_hermite2 (SRC/rt_solvers_mod.f90:1536-1895), the 'hermite2' output solver:
 * long_char_hermite's product-rule optical depth (slopes clamped to 3x the secant)
 * with long_char_quadratic's MC-limited source slope and parab_moments weights (A, B, C
 * in FP64, as the reference).  Of fields only pe4 and dpe_dz are read, of ghost only
 * pe and dpe_dz.  Otherwise as m3drt_lc_sweep_hermite. */
int m3drt_lc_sweep_hermite2(m3drt_ctx *ctx, const m3drt_sweep_desc *desc, const m3drt_sweep_bufs *bufs,
                            const m3drt_hermite_fields *fields, const m3drt_lc_ghost *ghost);

/* ------------------------------------------------------------------ line rates */

#define M3DRT_RATES_PRECOND 1   /* &nlte_params preconditioning=T */
#define M3DRT_RATES_OWN_NORM 2  /* the line lies inside the window: own profile normalisation */
#define M3DRT_RATES_LARGEST 4   /* rate_formulation='largest' (split rates; with PRECOND the
                                   tracked-contribution integrands) */
#define M3DRT_RATES_SCAT 8      /* line scattering: other += chi_scat * dJnu */
#define M3DRT_RATES_FLAGS 16    /* psi_flags: read in-force bits, raise new ones */

/* One direction's line-rate integrals.  (nline, nf) line arrays; (f, cell) fields at
 * f*sf + cell*sc; (nk, ncell) and (nline, ncell) per-cell arrays cell-fastest. */
typedef struct {
  int32_t nline, nf, ncell, nk, flags;
  float wmu, wtot;                          /* direction weight, sum of the weights */
  int64_t sf, sc;
  const int32_t *iblue, *ired;              /* (nline) 1-based window of each line */
  const int32_t *lev_i, *lev_j;             /* (nline) 1-based levels */
  const double *Aij, *Bij;                  /* (nline) */
  const float *lower_off;                   /* (nline) E_low of the line minus level E_i [eV] */
  const float *q;                           /* (nline, nf) profile abscissa (line%q) */
  const double *wnu, *wq;                   /* (nline, nf) rate and profile quadrature weights */
  const double *g, *b, *nstar, *pop_ratio;  /* (nk), (nk, ncell) x2, (ncell) */
  const float *inv_df, *vproj, *e_kT;       /* (ncell) */
  const float *adamp;                       /* (nline, ncell) damping in Doppler units */
  const float *Inu, *Psi, *Bnu, *beta;      /* this direction's intensity and Psi; B_nu, beta */
  float *gamma_ij, *gamma_ji;               /* (nline, ncell) th and dev, accumulated */
  /* LARGEST | PRECOND: the solver's split parts and the tracked contributions */
  int32_t split_offset, npw;                /* index base of this atom's lines; psi_flags words */
  const float *Irem, *escape, *chi;         /* (f, cell) Irem, escape, total chi (work slot 1) */
  const float *split_opac1, *split_opac2;   /* (f, cell) largest eta term, the rest */
  const float *split_chi1, *split_chi2;     /* (f, cell) largest chi term, the rest */
  const int32_t *split_id, *split_chi_id;   /* (f, cell) index of the largest terms */
  const float *chi_scat, *dJnu;             /* (f, cell), M3DRT_RATES_SCAT */
  const int32_t *ntrans;                    /* (nline) transition numbers, M3DRT_RATES_FLAGS */
  const int32_t *psi_flags_in;              /* (npw, ncell) words: bits in force */
  int32_t *psi_flags_out;                   /* (npw, ncell) words: bits raised (OR-ed in) */
} m3drt_bb_rates_args;

/* bound_bound_rates for one direction (SRC/rt_patch_mod.f90:2852-3142), FP32 with
 * Kahan-compensated frequency sums: 'original' (with or without preconditioning),
 * 'largest' without preconditioning, and 'largest' with it (Rybicki & Hummer
 * integrands from Irem, escape and the tracked contributions, psi_flags).  A line's
 * window is clipped to [1, nf]; a line whose level index lies outside [1, nk] is
 * skipped. */
int m3drt_bb_rates(m3drt_ctx *ctx, const m3drt_bb_rates_args *args);

/* ------------------------------------------------------------------ line opacity of a model atom */

#define M3DRT_BBOP_LTE 1    /* b = 1 (the LTE companion set) */
#define M3DRT_BBOP_SPLIT 2  /* track the largest contribution per point (add_largest) */

typedef struct {
  int32_t nline, nf, ncell, nk, flags, split_offset;
  float wmu;                                /* direction weight for wphi */
  int64_t sf, sc;
  const int32_t *iblue, *ired, *lev_i, *lev_j;
  const double *Bij;                        /* (nline) */
  const float *lower_off, *q;               /* (nline), (nline, nf) */
  const double *wq;                         /* (nline, nf) */
  const int32_t *wphi_line;                 /* (nline) 1: add this direction's wphi share; or NULL */
  const double *b, *nstar, *pop_ratio;      /* (nk, ncell) x2, (ncell) */
  const float *inv_df, *vproj, *e_kT;       /* (ncell) */
  const float *adamp;                       /* (nline, ncell) */
  float *xhat, *ehat;                       /* (f, cell) accumulated */
  float *split_opac1, *split_opac2, *split_chi1, *split_chi2;  /* (f, cell), SPLIT */
  int32_t *split_id, *split_chi_id;         /* (f, cell), SPLIT */
  float *wphi;                              /* (nline, ncell) += wmu * sum(phi wq) */
} m3drt_bb_opacity_args;

/* add_bound_bound_opacities (SRC/rt_patch_mod.f90:2689-2831) for one direction: the
 * model atom's (non-hydrogen) lines into xhat/ehat, add_largest tracking (3150-3169)
 * and the direction's wphi share.  Lines visit each point in list order. */
int m3drt_bb_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *args);

/* ------------------------------------------------------------------ bound-free rates */

typedef struct {
  int32_t ncont, nline, nf, ncell, nk, precond;
  double cut_x;                             /* radiative_cut threshold (dev * e/kT) */
  int64_t sf, sc;
  const int32_t *iblue, *ired, *lev_i, *lev_j, *active;  /* (ncont), 1-based */
  const double *ev;                         /* (nk) level energies [eV] */
  const double *wnu;                        /* (ncont, nf) frequency weights */
  const float *alpha;                       /* (ncont, nf) cross sections */
  const float *twohnu3_c2;                  /* (nf) */
  const double *e_kT;                       /* (ncell) */
  const double *nstar, *b;                  /* (nk, ncell) */
  const float *Jnu, *beta, *PsiJ;           /* (f, cell) */
  float *gamma_ij, *gamma_ji;               /* (nline + ncont, ncell); continua from row nline */
} m3drt_bf_rates_args;

/* bound_free_rates (SRC/rt_patch_mod.f90:3193-3259): photoionisation and
 * recombination sums from J, with the PsiJ preconditioning. */
int m3drt_bf_rates(m3drt_ctx *ctx, const m3drt_bf_rates_args *args);

/* ------------------------------------------------------------------ working units */

#define M3DRT_UNITS_SPLIT 1            /* eta from the tracked parts split_opac1/2 */
#define M3DRT_UNITS_SCAT 2             /* eta += chi_scat * dJnu (line scattering) */
#define M3DRT_UNITS_SUPPRESS_MASERS 4  /* chi = max(chi, floor) instead of the signed floor */

/* n (f, cell) points of one window, every array in the same layout. */
typedef struct {
  int64_t n;
  int32_t flags;
  double maser_chi_floor;              /* |chi| >= maser_chi_floor * |xhat| (as real(4)) */
  float *work1, *work2;                /* in: xhat, ehat; out: chi, S = eta/chi */
  const float *beta;                   /* stimulated-emission factor */
  const float *split_opac1, *split_opac2;  /* M3DRT_UNITS_SPLIT */
  const float *chi_scat, *dJnu;        /* M3DRT_UNITS_SCAT */
  int64_t *n_floored;                  /* device counter or NULL: += points whose chi was floored */
} m3drt_units_args;

/* work_to_transfer_units (SRC/rt_patch_mod.f90:3571-3651), in place. */
int m3drt_work_to_transfer(m3drt_ctx *ctx, const m3drt_units_args *args);

#define M3DRT_BG_SCAT_ALI 1  /* the continuum-scattering ALI step (scattering_preconditioning) */
#define M3DRT_BG_J_TO_B 2    /* initial_Jnu = Planck on the first iteration: J = B */

typedef struct {
  int64_t n;
  int32_t flags;
  float wscat;                         /* 1, or 0 to drop the scattering term */
  double scat_op_max;                  /* the clamp of op = sigma * PsiJ */
  float *J;                            /* in: J_FS; out: J - B (J_ali - B with the ALI step) */
  const float *B, *chi_cont, *sigma;
  const float *PsiJ, *Jali_old;        /* M3DRT_BG_SCAT_ALI */
  float *Jali;                         /* M3DRT_BG_SCAT_ALI: the new J_ali, or NULL */
  float *eta;                          /* out: chi_cont B + wscat sigma (J - B) */
  float *dJnu;                         /* out: wscat (J - B), or NULL */
  int64_t *n_neg;                      /* device counter or NULL: ALI values that went negative */
  double *diag;                        /* device [3] or NULL: max op_raw, op, J_ali/J_FS (max-ed in) */
} m3drt_background_args;

/* lte_background_emission (SRC/rt_patch_mod.f90:2127-2328), elementwise over a window. */
int m3drt_background_emission(m3drt_ctx *ctx, const m3drt_background_args *args);

/* convert_to_work_units (SRC/rt_patch_mod.f90:3518-3536): xhat = chi_cont + ehat,
 * then ehat = ehat / beta, over n points. */
int m3drt_convert_to_work_units(m3drt_ctx *ctx, int64_t n, const float *chi_cont, float *xhat, float *ehat,
                                const float *beta);

/* ------------------------------------------------------------------ one window, fused */

/* The continuum and conversion arrays of a fused window (every (f, cell) array in the
 * layout of the m3drt_bb_opacity_args it goes with). */
typedef struct {
  int32_t flags;                       /* M3DRT_UNITS_SUPPRESS_MASERS, or 0 */
  double maser_chi_floor;              /* as m3drt_units_args */
  const float *chi_cont, *Bnu, *beta;  /* (f, cell): the continuum opacity, B_nu, beta */
  int64_t *n_floored;                  /* device counter or NULL */
} m3drt_window_args;

/* One direction's opacity pass over a window in one kernel: foreground's continuum
 * (eta = chi_cont B_nu), convert_to_work_units (SRC/rt_patch_mod.f90:3518-3536), the
 * split tracking seeded with the continuum as contribution 0 (SRC/rt_patch_mod.f90:
 * 1467-1488), add_bound_bound_opacities with add_largest (:2689-2831, 3150-3169) and
 * work_to_transfer_units (:3571-3651, with M3DRT_BBOP_SPLIT its split form).  Writes
 * chi into args->xhat, S into args->ehat and, with M3DRT_BBOP_SPLIT, the six tracking
 * arrays; the same values as that sequence of m3drt_convert_to_work_units,
 * m3drt_bb_opacity and m3drt_work_to_transfer.  No wphi sums (wphi_line must be NULL)
 * and no line scattering (M3DRT_UNITS_SCAT). */
int m3drt_window_opacity(m3drt_ctx *ctx, const m3drt_bb_opacity_args *args, const m3drt_window_args *win);

/* ------------------------------------------------------------------ eta-step flags */

#define M3DRT_ETA_GHOST 1  /* the upstream ghost plane fires too (not the top of the domain) */

/* Dense flags, native layout: cubes (z, f, y, x) x fastest, the ghost plane (f, y, x). */
typedef struct {
  int32_t nf, nx, ny, nz, dir, flags;
  int32_t rx, ry;                       /* eta_rx, eta_ry: the spread of a flag (0, 0 for columns) */
  const float *chi, *S, *beta;          /* the transfer-unit chi, S and beta of the window */
  const float *ghost_chi, *ghost_S;     /* upstream ghost plane (M3DRT_ETA_GHOST) */
  uint8_t *src;                         /* (z, f, y, x) pairs that fired (grows) */
  uint8_t *mask;                        /* (z, f, y, x) pairs that take the eta step (grows) */
  uint8_t *ghost_src;                   /* (f, y, x) ghost pairs that fired (grows) */
} m3drt_eta_args;

/* update_eta_mask (SRC/rt_patch_mod.f90:3671-3802) with dense flags in place of the
 * reference's rows: the same flag sets.  The sweep takes them as eta_row[cell] =
 * cell + 1 (one row per cell) and eta_mask = mask viewed per z plane, or any
 * compaction of them.  Clearing the three arrays is the reference's eta reset. */
int m3drt_update_eta_mask(m3drt_ctx *ctx, const m3drt_eta_args *args);

/* ------------------------------------------------------------------ rotation for the output pass */

/* Cubes (f, x, y, z, k) addressed by element strides, k counting the cubes (chi and S
 * of the work array, or J), every z plane shifted periodically with bilinear
 * interpolation so that the rays of direction vec become vertical, the plane at the
 * pivot left in place: rotate_array (SRC/rt_patch_mod.f90:3809-3863), out of place. */
typedef struct {
  int32_t nf, nx, ny, nz, nk;
  int32_t reserved;
  int64_t cs_f, cs_x, cs_y, cs_z, cs_k;  /* element strides, the same for in and out */
  double vec[3];                         /* the ray's direction (vec[2] != 0) */
  double ds[3];                          /* cell size */
  double bot2pivot;                      /* pivot_z - bottom of the patch, as the caller forms it */
  double ray_offset[2];                  /* rt_output_params ray_offset [cells] */
  const float *in;
  float *out;                            /* must not overlap in */
} m3drt_rotate_args;

int m3drt_rotate_array(m3drt_ctx *ctx, const m3drt_rotate_args *args);

/* ------------------------------------------------------------------ batched 1D synthesis */

/* A batch of plane-parallel LTE atmospheres of uneven depth, frequency and angle
 * counts, each on its own equidistant wavelength grid.  Per (atmosphere, frequency)
 * one lane walks the depth points from the bottom up: it sums the atmosphere's line
 * opacities (QLG-bot Voigt profiles, HOP/qlg_mod.f90:538-627, each line only inside
 * its window, as add_opacity_equi windows them, HOP/linelist_mod.f90:853-906) onto
 * the continuum, forms S = B_nu(T), and integrates every upward mu with the linear
 * step of long_char2 in tau500 (SRC/rt_solvers_mod.f90:527-663); the deepest point
 * is thermalised (I = S).  In the reference each such atmosphere is one dispatch.x
 * run (1D: long_char2 per frequency window and direction, add_opacity_equi per
 * species); the batch has no single counterpart.
 *
 * Every array is a device array.  Per-depth, per-line, per-mu and per-(line, depth)
 * arrays are flat over the batch, each atmosphere's entries starting at its offset.
 * Wavelengths in cm, opacities relative to chi500, intensities in occupation units
 * (I / (2 h nu^3 / c^2)). */
#define M3DRT_B1D_MAXMU 8

typedef struct {
  int32_t natm;
  int32_t reserved;
  const int32_t *nd, *nf, *nmu, *nline;  /* (natm): depth points (>= 2), frequencies, upward mu, lines */
  const int64_t *d_off, *l_off, *ld_off, *mu_off;  /* (natm) first depth / line / (line, depth) record / mu */
  const int64_t *f_off;                 /* (natm) first element of flux */
  const int64_t *i_off;                 /* (natm) first element of I_top ((mu, f), f fastest) */
  const double *lam0, *dlam;            /* (natm) lam_f = lam0 + f dlam */
  const float *dtau;                    /* (depth) vertical tau500 step from point d+1 up to d (last unused) */
  const float *hckT;                    /* (depth) h c / (k T) [cm]: B = 1 / (exp(hckT / lam) - 1) */
  const float *ivth;                    /* (depth) 1 / Doppler speed [s/cm] */
  const float *cont_a, *cont_b;         /* (depth) chi_cont / chi500 = cont_a + cont_b (lam / lam_ref)^3 */
  const float *line_off;                /* (line) lam_l - lam0 of the line's atmosphere [cm] */
  const float *line_cl;                 /* (line) c / lam_l [1/s]: v = |lam - lam_l| line_cl ivth */
  const float *line_rec;                /* (line, depth), the atmosphere's records depth-major ([d][l]),
                                           4 floats each: damping a, centre strength s (the line's
                                           chi / chi500 is s K(a, v)), window half-width [cm] (< 0:
                                           the line is skipped at that depth), reserved */
  const float *mu, *wflux;              /* (mu) cosine (0 < mu <= 1) and flux weight */
  double lam_ref;                       /* [cm] */
  float *I_top;                         /* out: emergent intensity of every (atmosphere, mu, f) */
  float *flux;                          /* out: sum_mu wflux I_top, every (atmosphere, f) */
} m3drt_b1d_batch;

/* Tile shape of the batched kernel: frequencies per lane (a wavefront covers 64 fpl
 * consecutive frequencies of one atmosphere) and threads per workgroup (the waves of
 * a workgroup share one LDS copy of the QLG-bot tables). */
typedef struct {
  int32_t fpl;        /* 1 or 2 */
  int32_t threads;    /* 256, 512 or 1024 */
} m3drt_b1d_tile;

typedef struct m3drt_b1d_plan m3drt_b1d_plan;

/* The plain path: one launch per atmosphere (a static grid of 64-frequency tiles), the
 * same per-lane arithmetic.  tile NULL: fpl 1, 256 threads. */
int m3drt_b1d_synth_plain(m3drt_ctx *ctx, const m3drt_b1d_batch *batch, const m3drt_b1d_tile *tile);

/* A plan for repeated synthesis of one batch (the arrays stay where they are; their
 * contents may change between runs): the work list of (atmosphere, frequency tile)
 * items in decreasing cost, the per-(line, depth) profile coefficients, and the
 * persistent launch (a counter reset and one kernel whose waves take items until the
 * list is exhausted) captured in a HIP graph.  tile NULL: the autotuner times every
 * tile shape on a strided sample of the work list and keeps the fastest. */
int m3drt_b1d_plan_create(m3drt_ctx *ctx, const m3drt_b1d_batch *batch, const m3drt_b1d_tile *tile,
                          m3drt_b1d_plan **plan);
/* Re-forms the profile coefficients from line_rec (after its contents changed) and
 * replays the graph, ordered after the context stream's earlier work. */
int m3drt_b1d_plan_run(m3drt_b1d_plan *plan);
/* The tile in use and, when the autotuner chose it, every candidate's time on the
 * sample in ms (times[i] for candidate i of the order fpl 1/2 x threads 256/512/1024;
 * < 0 where not timed); ntimes receives the count (6). */
int m3drt_b1d_plan_info(const m3drt_b1d_plan *plan, m3drt_b1d_tile *tile, double *times, int32_t *ntimes);
int m3drt_b1d_plan_destroy(m3drt_b1d_plan *plan);

/* Strides of the native frequency-sliced layout (x fastest, then y, f, z) for a
 * cube and for a plane; fills cs_* and gs_* of desc from its nf, nx, ny. */
int m3drt_native_strides(m3drt_sweep_desc *desc);
/* Strides of the reference layout work(f,x,y,z), f fastest. */
int m3drt_reference_strides(m3drt_sweep_desc *desc);

/* The line-list opacity of a composition: replaces add_linelist_opacities
 * (HOP/composition_mod.f90:1663-1954) for its metals and molecules (the hydrogen
 * Stark lines are a separate profile) together with add_particle_opacity
 * (HOP/linelist_mod.f90:970-1070) for each of their isotopes.
 *   logref = ln(tolerance/voigt_peak) + ln all_chi (-huge where all_chi <= 0), once,
 *     from the working opacity before any line (:1766-1772);
 *   per species and isotope: inv_df = c/sqrt(2kT/(m_u awgt) + vmic^2),
 *     vbar2 = 8kT/(pi m_u) (m_u/m_H + 1/awgt) (HOP/ion_mod.f90:198-221),
 *     lnbase = ln max(n, tiny) [+ ln scale] - part + ln inv_df + ln frac, and the
 *     Voigt summation of m3drt_line_opacity_equi (equidistant) or m3drt_line_opacity
 *     into the increments chi_l, eta_l, with all_chi as the tolerance reference;
 *     species with departures (dep) in NLTE, the others in LTE with the line
 *     scattering fraction of their stage group (one_m_eps group 1 for neutrals,
 *     2 for ions, :1928-1934) when chi_scat is given;
 *   all_chi += chi_l, all_eta += eta_l (and chi_lines, eta_lines = the increments);
 *   chi_scat *= 1 - beta (working to physical units, :1879-1890) when beta is given.
 * A composition of one list of at most 4 lines (no other list, no scattering split,
 * chi_lines and eta_lines NULL) is summed without the increments: each point's line
 * sum, formed from zero in line order, is added to all_chi / all_eta where the lines
 * are deposited, with the same result bit for bit and no (f, cell) scratch.
 * All (f, cell) arrays have cells fastest (f*ncell + cell): all_chi, all_eta,
 * chi_scat, beta, chi_lines, eta_lines, and one_m_eps (f, cell, 2).  Species and
 * isotope descriptors are HOST arrays, their members DEVICE arrays. */
typedef struct {
  m3drt_linelist lines;       /* the isotope's line list (nline 0: skipped) */
  double awgt;                /* atomic weight [u] */
  double frac;                /* isotope fraction (1 for the main list) */
} m3drt_isotope;
typedef struct {
  int32_t niso;
  const m3drt_isotope *iso;   /* host array of niso */
  const double *n;            /* (ncell) number density [cm^-3] */
  double scale;               /* population_scale (1: none) */
  const double *part;         /* (ncell) ln of the partition function */
  const float *dep;           /* (ncell, nlev) departure coefficients, NULL: LTE */
  int32_t nlev;
  int32_t scat_group;         /* one_m_eps group of an LTE species: 1 neutral, 2 ion, 0 none */
} m3drt_species;
typedef struct {
  int32_t nf, nsp, equidistant;
  int64_t ncell;
  const double *llam;                              /* (nf) ln(lambda) */
  const m3drt_species *sp;                         /* host array of nsp */
  const float *temp, *vmic, *vproj, *e_kT, *ne, *tothi, *stark_pe;  /* (ncell); vmic [km/s] */
  float *all_chi, *all_eta;                        /* working xhat, ehat: added to */
  const float *one_m_eps;                          /* (f, cell, 2) or NULL */
  float *chi_scat;                                 /* or NULL (with one_m_eps) */
  const float *beta;                               /* or NULL */
  float *chi_lines, *eta_lines;                    /* or NULL: the increments */
  float tolerance, line_reject_frac;               /* 1e-4 and 0 in the reference */
} m3drt_linelist_opacity_args;
int m3drt_add_linelist_opacities(m3drt_ctx *ctx, const m3drt_linelist_opacity_args *args);

/* The LTE background factors per (frequency, cell): replaces the beta/Bnu loop of
 * lte_background_opacities (SRC/rt_patch_mod.f90:2004-2023): h_kT = h/k / T and
 * hv_kT = nu h_kT in FP32, beta = max(exp(-hv_kT), the smallest normal FP32 number),
 * Bnu = beta/(1 - beta) (occupation units), with the reference's h and k
 * (HOP/const_mod.f90:39-43).  nu (nf) and temp (ncell) FP64; beta and Bnu FP32 at
 * f*sf + cell*sc (sf = sc = 0: sf = ncell, sc = 1). */
typedef struct {
  int32_t nf;
  int64_t ncell;
  const double *nu, *temp;
  float *beta, *Bnu;
  int64_t sf, sc;
} m3drt_background_lte_args;
int m3drt_lte_background(m3drt_ctx *ctx, const m3drt_background_lte_args *args);

/* The sampled background-opacity table of HOP (HOP/optab_mod.f90): the LTE
 * continuum extinction and its scattering part per chi500, sampled on (wavelength,
 * PC axis 1, PC axis 2) nodes, the PC axes a rotation of (ln T, ln ne).
 * m3drt_optab_plan_create replaces sample_build_derivatives (:127-166): it copies
 * the table to the device with its monotone PCHIP node slopes along every axis.
 * m3drt_optab_interpolate replaces sample_interpolate (:181-287), the separable
 * tensor-product cubic Hermite of lte_background_opacities' sampled path
 * (SRC/rt_patch_mod.f90:2115-2117): back = (kappa + kappa_scat) chi500 and
 * scat = kappa_scat chi500 per (frequency, cell), at f*sf + cell*sc (zeros: sf =
 * ncell, sc = 1).  m3drt_optab_slopes copies out the node slopes (dl, dt, de of
 * val, sl, st, se of scat), each shaped like val.  Table and query arrays are
 * DEVICE arrays; the plan keeps its own copies. */
typedef struct {
  int32_t nnu, ntemp, nne;
  double mu[2];              /* mean of (ln T, ln ne) */
  double rot[4];             /* PC directions as columns, column-major: rot(1,1), rot(2,1), rot(1,2), rot(2,2) */
  double c0[2], dc[2];       /* first node and spacing of each PC axis */
  const double *lam;         /* (nnu) wavelength nodes, ascending */
  const float *val, *scat;   /* (nnu, ntemp, nne), nnu fastest */
} m3drt_optab_table;
typedef struct m3drt_optab_plan m3drt_optab_plan;
typedef struct {
  int64_t ncell;
  int32_t nf;
  const double *temp, *ne;   /* (ncell) [K], [cm^-3] */
  const float *chi500;       /* (ncell) */
  const double *lam_q;       /* (nf) wavelengths of the frequencies, the table's unit */
  float *back, *scat;
  int64_t sf, sc;
} m3drt_optab_query;
int m3drt_optab_plan_create(m3drt_ctx *ctx, const m3drt_optab_table *table, m3drt_optab_plan **plan);
int m3drt_optab_slopes(m3drt_optab_plan *plan, float *dl, float *dt, float *de, float *sl, float *st, float *se);
int m3drt_optab_interpolate(m3drt_optab_plan *plan, const m3drt_optab_query *query);
int m3drt_optab_plan_destroy(m3drt_optab_plan *plan);

/* Boundary planes of a directional sweep.  m3drt_set_upstream_boundary replaces
 * set_upstream_boundary (SRC/rt_patch_mod.f90:3270-3374): it fills the ghost plane
 * (chi, S, I) the sweep starts from, of one of four kinds:
 *   M3DRT_GHOST_TOP       zeros (the top of the domain, downward rays);
 *   M3DRT_GHOST_LTE       lte_ghost (:3352-3373): chi = chi1 * g500/chi500_1,
 *                         S = I = Bnu(T_ghost) as m3drt_lte_background forms it;
 *   M3DRT_GHOST_PCHIP     the other solvers' bottom ghost (:3287-3300): chi and S
 *                         extrapolated by pchip_mono (HOP/interpol_mod.f90:289-365)
 *                         from the first three planes, i*ds3 apart, to one plane below,
 *                         clamped at 0; I = S (only the first two planes enter);
 *   M3DRT_GHOST_UPSTREAM  the upstream slab's exit plane mem (chi, S, I), copied.
 * m3drt_set_downstream_boundary replaces set_downstream_boundary (:3471-3511): it
 * copies the exit plane (chi1, S1: the slab's last plane in the sweep direction;
 * exit_I: its intensity) into mem.  Planes are (nf, ny, nx) addressed
 * f*ps_f + x*ps_x + y*ps_y (zeros: native strides nx*ny, 1, nx); every plane of one
 * call shares the strides. */
#define M3DRT_GHOST_TOP 0
#define M3DRT_GHOST_LTE 1
#define M3DRT_GHOST_PCHIP 2
#define M3DRT_GHOST_UPSTREAM 3
typedef struct {
  int32_t kind, nf, nx, ny;
  int64_t ps_f, ps_x, ps_y;
  const float *chi1, *S1, *chi2, *S2;   /* the first (PCHIP: and second) plane of the slab */
  const float *g500, *chi500_1, *T_ghost; /* LTE: per cell (nx*ny, x fastest) */
  const double *nu;                     /* LTE: frequencies [Hz] (nf) */
  double ds3;                           /* PCHIP: cell size in z */
  const float *exit_I;                  /* downstream: the exit plane's intensity */
  float *mem_chi, *mem_S, *mem_I;       /* upstream: read; downstream: written */
  float *ghost_chi, *ghost_S, *ghost_I; /* upstream: written */
} m3drt_boundary_args;
int m3drt_set_upstream_boundary(m3drt_ctx *ctx, const m3drt_boundary_args *args);
int m3drt_set_downstream_boundary(m3drt_ctx *ctx, const m3drt_boundary_args *args);

/* Angular quadrature of a transfer pass: replaces setup_quadrature
 * (SRC/rt_params_mod.f90:170-220) and the sets it calls, set_radau_quadratures,
 * set_lobatto_quadratures, set_gauss_quadratures, set_carlson_quadratures,
 * set_custom_quadratures, set_grid_quadratures and nphi_copies
 * (SRC/ray_mod.f90:36-467; Gauss-Radau nodes of solvers/rt/radau_mod.f90:15-115).
 * The defaults of the reference are "radau", ntheta 2, nphi 4 for the iteration
 * pass (jnu_weights 1) and "lobatto", ntheta 4, nphi 4 for the output pass
 * (jnu_weights 0; SRC/rt_params_mod.f90:122-130).  HOST arrays: vec (cap, 3) holds
 * the directions as the reference's vec(3, nray), wts (cap) the weights; *nray is
 * set to the number of rays also when cap is too small (M3DRT_ELIMIT).  Ranges of
 * the reference that stop its run (ntheta beyond a table) are M3DRT_EINVAL. */
typedef struct {
  const char *scheme;       /* "radau", "lobatto", "gauss", "custom", "grid+", "gridx", "grid*",
                               "disk_center", "set_a2" ... "set_a8", "set_b4" ... "set_b8", "none" */
  int32_t ntheta, nphi;     /* inclinations (custom: the number of mu values), azimuths */
  int32_t jnu_weights;      /* 1: the set mirrored downward, weights summing to 1 (mean intensity);
                               0: the upward rays with flux weights (w * mu) */
  int32_t one_dimensional;  /* plane-parallel atmosphere: one azimuth per inclination */
  const double *mu;         /* custom: ntheta inclinations (host) */
  double ds[3];             /* cell size, for the grid schemes */
} m3drt_quadrature_args;
int m3drt_quadrature(const m3drt_quadrature_args *q, int32_t cap, double *vec, double *wts, int32_t *nray);

/* The FP32 step weights the sweep kernels use, elementwise over n optical depths t
 * (device arrays of n floats each).
 * m3drt_linear_weights replaces linear_weights / linear_weights_array
 * (SRC/rt_solvers_mod.f90:785-822): E = exp(-t), W = u0 - Psi, Psi = 1 - u0/t,
 * Esc = 1 - Psi with u0 = 1 - E.
 * m3drt_parab_moments replaces parab_moments (SRC/rt_solvers_mod.f90:862-877, the
 * pure procedure of rt_patch_mod's interface): E = exp(-t), U0 = 1 - E,
 * U1 = t - U0, q = (t^2 - 2 U1)/t.
 * Both take E and 1 - E from exp_neg's reduction (HOP/exp_neg.inc:20-37), t
 * clamped at 87 for the exponential. */
int m3drt_linear_weights(m3drt_ctx *ctx, int64_t n, const float *t, float *E, float *W, float *Psi,
                         float *Esc);
int m3drt_parab_moments(m3drt_ctx *ctx, int64_t n, const float *t, float *E, float *U0, float *U1, float *q);

#ifdef __cplusplus
}
#endif
#endif

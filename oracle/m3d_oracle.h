/*
 * m3d_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of OpenMULTI3D's radiative-transfer hot path, used as the
 * checker for the HIP kernels of dispatch_amd.  Only tests/, the smoke() entry
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Array convention (the reference's, Fortran column-major, frequency fastest):
 *   work-like cubes  a[f + nf*(x + nx*(y + ny*z))]
 * with z running over the ghost planes too where stated (0 .. nz+1).
 *
 * Each function cites the reference file:line it restates; paths are relative to
 * experiments/OpenMULTI3D/ of Bergemann-Group-development/dispatch.
 */
#ifndef M3D_ORACLE_H
#define M3D_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* HOP/exp_neg.inc:20-37 */
void orc_exp_neg(float t, float *E, float *u0);
/* HOP/exp_neg.inc:43-49 */
void orc_exp_neg0(float t, float *E, float *u0);
/* SRC/rt_solvers_mod.f90:785-797 (linear_weights), 803-809 (lw_series) */
void orc_linear_weights(float t, float *E, float *W, float *Psi, float *Esc);
float orc_lw_series(float t);
/* SRC/rt_solvers_mod.f90:862-877 */
void orc_parab_moments(float t, float *E, float *U0, float *U1, float *q);
/* vectorised helpers for tests */
void orc_linear_weights_n(int n, const float *t, float *E, float *W, float *Psi, float *Esc);
void orc_parab_moments_n(int n, const float *t, float *E, float *U0, float *U1, float *q);
void orc_exp_neg_n(int n, const float *t, float *E, float *u0);

/* SRC/ray_mod.f90:581-596 */
void orc_cc_kernel(double s, double k[4]);
/* SRC/ray_mod.f90:493-508; ds in cm */
typedef struct { double xkern[4], ykern[4], dsup; int upo1, upo2; } orc_ray_t;
void orc_set_stencil(const double ds[3], const double vec[3], orc_ray_t *ray);

/* SRC/ray_mod.f90:269-311 + 392-467 (set_gauss_quadratures, nphi_copies).
 * Returns the number of rays; vec is (3, nray), wts (nray).  Arrays must hold
 * 2*ntheta*nphi entries (mirror) or ntheta*nphi. */
int orc_gauss_quadrature(int ntheta, int nphi, int jnu_weights, int one_dimensional,
                         double *vec, double *wts);

/* Parameters of one 3D short-characteristics call (SRC/rt_solvers_mod.f90:224-507). */
typedef struct {
  int nf, nx, ny, nz;
  int dir;                 /* +1: z increasing (up), -1 */
  double ds[3];            /* cell size in cm (self%ds * scaling%l) */
  double vec[3];           /* ray direction */
  double maser_dtau_max;   /* SRC/m3d_patch_mod.f90:319 default 2 */
  double maser_Inu_max;    /* SRC/m3d_patch_mod.f90:320 default 1e30 */
  int split_rates;         /* fill Irem and escape */
  int get_rates;           /* fill Psi */
  int nf_stride;           /* leading dimension of the cubes (0: nf); a frequency
                              window fs:fe is passed as pointers offset by fs */
} orc_sc_params;

/*
 * short_char2: work arrays chi, S, I each hold (nf, nx, ny, nz+2): planes 0 and
 * nz+1 are the ghost planes (the upstream one filled by the caller).  chi500 is
 * (nx, ny, nz), chi500_up (nx, ny, nz) for this direction, g500 (nx, ny) the
 * ghost plane's chi500.  eta_row (nx, ny, nz) or NULL; eta_mask (nf, nrow) bytes.
 * Outputs: I (slot 3, planes 1..nz), Psi/Irem/escape (nf, nx, ny, nz) or NULL.
 */
void orc_short_char2_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                         const float *chi500, const float *chi500_up, const float *g500,
                         const int32_t *eta_row, const uint8_t *eta_mask,
                         float *Psi, float *Irem, float *escape);
/* the same algorithm with every quantity in FP64 and exact weights: the "plain
 * FP64" reference against which FP32 errors are stated */
void orc_short_char2_f64(const orc_sc_params *p, const double *chi, const double *S, double *I,
                         const double *chi500, const double *chi500_up, const double *g500,
                         const int32_t *eta_row, const uint8_t *eta_mask,
                         double *Psi, double *Irem, double *escape);

/* short_char, SRC/rt_solvers_mod.f90:19-189 (4x4 cubic for chi, S, I) */
void orc_short_char_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                        float *Psi, float *Irem, float *escape);

/* long_char2, SRC/rt_solvers_mod.f90:527-663; nx, ny columns; g500 (nx,ny) */
void orc_long_char2_f32(const orc_sc_params *p, const float *chi, const float *S, float *I,
                        const float *chi500, const float *g500,
                        const int32_t *eta_row, const uint8_t *eta_mask,
                        float *Psi, float *Irem, float *escape);
void orc_long_char2_f64(const orc_sc_params *p, const double *chi, const double *S, double *I,
                        const double *chi500, const double *g500,
                        const int32_t *eta_row, const uint8_t *eta_mask,
                        double *Psi, double *Irem, double *escape);

/* long_char_quadratic + quad_step, SRC/rt_solvers_mod.f90:891-1020 */
void orc_long_char_quadratic_f32(const orc_sc_params *p, const float *chi, const float *S,
                                 float *I, float *Psi);

/* integrate_intensity, SRC/rt_patch_mod.f90:3399-3431: J += 2*wmu*I over planes 1..nz */
void orc_integrate_intensity(int nf, int nx, int ny, int nz, float wmu, const float *I,
                             const float *Psi, float *J, float *PsiJ);

/* CPU baseline: short_char2 + integrate_intensity over nwin frequency windows on
 * OpenMP threads (the reference's n_windows columns) */
void orc_sweep_windows_f32(const orc_sc_params *p, int nwin, const float *chi, const float *S,
                           float *I, const float *chi500, const float *chi500_up, const float *g500,
                           float wmu2, float *J);

#ifdef __cplusplus
}
#endif
#endif

/* m3d_b1d.h -- TEST INFRASTRUCTURE ONLY: oracle of the batched 1D synthesis
 * (one plane-parallel atmosphere). */
#ifndef M3D_B1D_H
#define M3D_B1D_H
#ifdef __cplusplus
extern "C" {
#endif

/* One atmosphere: nd depth points (0 = top), nf frequencies lam0 + f dlam, nmu upward
 * mu, nline lines; line_rec (nd, nline, 4) = (a, s, window half-width, -).  Writes
 * I_top (nmu, nf) and flux (nf) in FP64; chi_out (nd, nf) if not NULL: chi / chi500. */
void orc_b1d_atmosphere(int nd, int nf, int nmu, int nline, double lam0, double dlam, const float *dtau,
                        const float *hckT, const float *ivth, const float *cont_a, const float *cont_b,
                        const float *line_off, const float *line_cl, const float *line_rec, const float *mu,
                        const float *wflux, double lam_ref, double *I_top, double *flux, double *chi_out);

#ifdef __cplusplus
}
#endif
#endif

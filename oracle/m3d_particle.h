/* TEST INFRASTRUCTURE ONLY: per-isotope population base and tolerance reference
 * restatements (m3d_particle.c). */
#ifndef M3D_PARTICLE_H
#define M3D_PARTICLE_H
void orc_particle_base(long ncell, const float *temp, const float *vmic, const double *n, double scale,
                       const double *part, double awgt, double frac, float *idf, float *vb2, float *lnbase);
void orc_logref(long n, const float *chi, float tolerance, float *logref);
#endif

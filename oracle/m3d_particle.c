/*
 * m3d_particle.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the per-isotope
 * preamble of add_particle_opacity (HOP/linelist_mod.f90:1012-1030 of
 * experiments/OpenMULTI3D/ in Bergemann-Group-development/dispatch), with
 * inverse_doppler_factor and compute_vbar2 (HOP/ion_mod.f90:198-221), and the
 * tolerance reference of add_linelist_opacities (HOP/composition_mod.f90:1766-1772);
 * the checker for dispatch_amd/csrc/linelist.hip.
 *
 * In the reference's mixed precision: the real(4) parameters 2k/m_u, 8k/(pi m_u),
 * m_u/m_H; dop = (const T)/awgt in FP64 stored FP32, + (vmic 1e5)^2 in FP32;
 * inv_df = c/sqrt(dop) in FP64 stored FP32; vbar2 = (const T)(m_u/m_H + 1/awgt)
 * in FP64 stored FP32; lnbase = ln max(n, tiny) in FP64 stored FP32, + ln(scale)
 * as FP32, then (- part + ln inv_df + ln frac) in FP64 stored FP32.  ln inv_df is
 * the FP32 logarithm (glibc logf here).
 *
 * Pinning: inv_df, vbar2 and lnbase against their FP64 closed forms
 * (tests/test_oracle_linelist.py).
 */
#include <float.h>
#include <math.h>
#include "m3d_particle.h"

#define CC 2.99792458e10
#define UU 1.6605655e-24
#define BK 1.380662e-16
#define MH 1.6726e-24
#define PI_REF 3.14159265359

void orc_particle_base(long ncell, const float *temp, const float *vmic, const double *n, double scale,
                       const double *part, double awgt, double frac, float *idf, float *vb2, float *lnbase)
{
  const float cdop = (float)(2.0 * BK / UU);
  const float cvb = (float)(8.0 * BK / (PI_REF * UU)), inv_mh = (float)(UU / MH);
  for (long c = 0; c < ncell; c++) {
    const float T = temp[c];
    float dop = (float)((double)(cdop * T) / awgt);
    if (vmic[c] > 0) {
      const float v = vmic[c] * 1e5f;
      dop = dop + v * v;
    }
    idf[c] = (float)(CC / (double)sqrtf(dop));
    vb2[c] = (float)((double)(cvb * T) * ((double)inv_mh + 1.0 / awgt));
    float lb = (float)log(n[c] > DBL_MIN ? n[c] : DBL_MIN);
    if (scale != 1.0) lb = lb + (float)log(scale);
    lnbase[c] = (float)((((double)lb - part[c]) + (double)logf(idf[c])) + (double)(float)log(frac));
  }
}

void orc_logref(long n, const float *chi, float tolerance, float *logref)
{
  const float lt = (float)log((double)(tolerance / 0.5641895835f));
  for (long i = 0; i < n; i++) logref[i] = chi[i] > 0.0f ? lt + logf(chi[i]) : -FLT_MAX;
}

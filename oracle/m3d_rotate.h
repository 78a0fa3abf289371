/* m3d_rotate.h -- TEST INFRASTRUCTURE ONLY: oracle of rotate_array. */
#ifndef M3D_ROTATE_H
#define M3D_ROTATE_H
#ifdef __cplusplus
extern "C" {
#endif
/* cubes in the reference layout array(f, x, y, z, k), f fastest, out of place */
void orc_rotate_array(const double vec[3], const double ds[3], double bot2pivot, const double ray_offset[2],
                      int nf, int nx, int ny, int nz, int nk, const float *in, float *out);
#ifdef __cplusplus
}
#endif
#endif

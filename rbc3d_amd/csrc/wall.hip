// Wall (vessel) surfaces: placeholder entry points until the wall kernels land.
#include "internal.h"

int rbc_add_int_on_walls_dev(rbc3d_ctx *c, double, int, const double *, const int *, const int *,
                             const double *, double *) {
  (void)c;
  rbc_set_error("wall integrals are not available in this build");
  return RBC3D_EINVAL;
}

extern "C" int rbc3d_set_wall(rbc3d_ctx *c, int, int, const double *, const int *) {
  (void)c;
  rbc_set_error("walls are not available in this build");
  return RBC3D_EINVAL;
}
extern "C" int rbc3d_set_wall_density(rbc3d_ctx *c, const double *) {
  (void)c;
  return RBC3D_EINVAL;
}
extern "C" int rbc3d_prepare_sing_int_on_wall(rbc3d_ctx *c) {
  (void)c;
  return RBC3D_EINVAL;
}
extern "C" int rbc3d_wall_self_nnz(rbc3d_ctx *c, long long *nnz) {
  (void)c;
  if (nnz) *nnz = 0;
  return RBC3D_EINVAL;
}
extern "C" int rbc3d_add_int_on_walls(rbc3d_ctx *c, double, int, const double *, const int *,
                                      const int *, const double *, double *) {
  (void)c;
  return RBC3D_EINVAL;
}
extern "C" int rbc3d_wall_dof(rbc3d_ctx *c, int *dof) {
  (void)c;
  if (dof) *dof = 0;
  return RBC3D_EINVAL;
}
extern "C" int rbc3d_wall_matmult(rbc3d_ctx *c, const double *, double *) {
  (void)c;
  return RBC3D_EINVAL;
}
extern "C" int rbc3d_wall_residual_vel(rbc3d_ctx *c, double *) {
  (void)c;
  return RBC3D_EINVAL;
}

// Device helpers shared by the ray kernels (render and gradient).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace at3d {

// LOCATE_GRID_CELL (shdomsub2.f:4043-4206) for a base grid without a tree.
__device__ inline int locate_cell_ray(const RayArgs& a, double& x0, double& y0, double z0) {
  const float* XG = a.xgrid;
  const float* YG = a.ygrid;
  const float* ZG = a.zgrid;
  if (!a.open_x) {
    const double xd = (double)(XG[a.nx] - XG[0]);
    if (x0 < XG[0])
      x0 = x0 - xd * (double)((int)((x0 - XG[0]) / xd) - 1);
    else if (x0 > XG[a.nx])
      x0 = x0 - xd * (double)(int)((x0 - XG[0]) / xd);
  }
  if (!a.open_y) {
    const double yd = (double)(YG[a.ny] - YG[0]);
    if (y0 < YG[0])
      y0 = y0 - yd * (double)((int)((y0 - YG[0]) / yd) - 1);
    else if (y0 > YG[a.ny])
      y0 = y0 - yd * (double)(int)((y0 - YG[0]) / yd);
  }
  int il = 0, iu = a.nx + 1;
  while (iu - il > 1) {
    const int im = (iu + il) / 2;
    if (x0 >= XG[im - 1]) il = im; else iu = im;
  }
  int ix = max(il, 1);
  il = 0;
  iu = a.ny + 1;
  while (iu - il > 1) {
    const int im = (iu + il) / 2;
    if (y0 >= YG[im - 1]) il = im; else iu = im;
  }
  int iy = max(il, 1);
  il = 0;
  iu = a.nz;
  while (iu - il > 1) {
    const int im = (iu + il) / 2;
    if (z0 >= ZG[im - 1]) il = im; else iu = im;
  }
  const int iz = max(il, 1);
  if (a.open_x) {
    if (x0 < XG[0]) ix = 1;
    else if (x0 > XG[a.nx - 1]) ix = a.nx + 1;
    else ix = ix + 1;
  }
  if (a.open_y) {
    if (y0 < YG[0]) iy = 1;
    else if (y0 > YG[a.ny - 1]) iy = a.ny + 1;
    else iy = iy + 1;
  }
  return (iz + (a.nz - 1) * (iy - 1) + (a.nz - 1) * a.nyc * (ix - 1)) - 1;
}

}  // namespace at3d

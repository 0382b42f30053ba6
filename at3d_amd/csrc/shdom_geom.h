// Base-grid geometry shared by the host planner and the GPU sweep kernel.
//
// Everything here is compiled twice (host and gfx950) from the same source with
// floating-point contraction disabled, so the host dependency planner and the
// device sweep take bit-identical decisions when they trace the same
// characteristic.  Indices are 0-based; the reference's 1-based conventions are
// noted where they matter.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace at3d {

// Face corner table: the four cell corners on each face -X,+X,-Y,+Y,-Z,+Z.
// Corner c = bitx + 2*bity + 4*bitz (the reference's GRIDPTR corner index - 1,
// INIT_CELL_STRUCTURE shdomsub2.f:84-164; GRIDFACE data shdomsub1.f:3733-3734).
__host__ __device__ inline int face_corner(int face, int k) {
  // faces: 0:-X 1:+X 2:-Y 3:+Y 4:-Z 5:+Z
  const int8_t tab[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                            {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
  return tab[face][k];
}

// Static grid topology (device or host pointers).
struct GridTopo {
  const float* gridpos;     // [npts][3]
  const int32_t* gridptr;   // [ncells][8]
  const int32_t* neighptr;  // [ncells][6], -1 for a domain boundary
  const int8_t* cellflags;  // [ncells] bit0: independent pixel in X, bit1: in Y
  int npts, ncells;
};

// Per-ordinate trace-back direction (BACK_INT_GRID3D_UNPOL shdomsub1.f:3736-3785).
struct OrdGeom {
  double cx, cy, cz, cxinv, cyinv, czinv;
  double eps;
  int ioct;  // 0..7 = (reference IOCT) - 1 = bitx + 2*bity + 4*bitz
  int bitx, bity, bitz;
};

// One cell step of the short-characteristic trace (the loop body of
// BACK_INT_GRID3D_UNPOL, shdomsub1.f:3822-3920): from the current position in
// cell `icell`, find the exit plane, the exit face corners, the bilinear face
// weights and the next cell.  Returns false on a negative path length (the
// reference stops there).
struct TraceStep {
  double so;        // path length through the cell
  double u, v;      // face interpolation factors
  int i[4];         // face corner points
  int next;         // next cell or -1 at a domain boundary
  int face;         // exit face (0..5: -x, +x, -y, +y, -z, +z)
};

__host__ __device__ inline bool trace_step(const GridTopo& g, const OrdGeom& o, int icell,
                                           double& xe, double& ye, double& ze, TraceStep& t) {
#pragma clang fp contract(off)
  const int8_t flags = g.cellflags[icell];
  const bool ipinx = (flags & 1) != 0;
  const bool ipiny = (flags & 2) != 0;
  const int iopp = g.gridptr[icell * 8 + (7 - o.ioct)];
  const float* P = g.gridpos;
  double sox = ipinx ? 1.0e20 : ((double)P[iopp * 3 + 0] - xe) * o.cxinv;
  double soy = ipiny ? 1.0e20 : ((double)P[iopp * 3 + 1] - ye) * o.cyinv;
  double soz = ((double)P[iopp * 3 + 2] - ze) * o.czinv;
  double so = sox < soy ? sox : soy;
  so = so < soz ? so : soz;
  if (so < -o.eps) return false;
  xe = xe + so * o.cx;
  ye = ye + so * o.cy;
  ze = ze + so * o.cz;
  int iface, jface;
  if (sox <= soz && sox <= soy) {
    iface = 1 - o.bitx;  // 2-BITX (1-based) -> 0-based
    jface = 1;
  } else if (soy <= soz) {
    iface = 3 - o.bity;
    jface = 2;
  } else {
    iface = 5 - o.bitz;
    jface = 3;
  }
  const int inext = g.neighptr[icell * 6 + iface];
  const int* gp = g.gridptr + icell * 8;
  const int i1 = gp[face_corner(iface, 0)], i2 = gp[face_corner(iface, 1)];
  const int i3 = gp[face_corner(iface, 2)], i4 = gp[face_corner(iface, 3)];
  double u, v;
  if (jface == 1) {
    u = (ze - (double)P[i1 * 3 + 2]) / ((double)P[i3 * 3 + 2] - (double)P[i1 * 3 + 2]);
    v = ipiny ? 0.5 : (ye - (double)P[i1 * 3 + 1]) / ((double)P[i2 * 3 + 1] - (double)P[i1 * 3 + 1]);
  } else if (jface == 2) {
    u = (ze - (double)P[i1 * 3 + 2]) / ((double)P[i3 * 3 + 2] - (double)P[i1 * 3 + 2]);
    v = ipinx ? 0.5 : (xe - (double)P[i1 * 3 + 0]) / ((double)P[i2 * 3 + 0] - (double)P[i1 * 3 + 0]);
  } else {
    u = ipiny ? 0.5 : (ye - (double)P[i1 * 3 + 1]) / ((double)P[i3 * 3 + 1] - (double)P[i1 * 3 + 1]);
    v = ipinx ? 0.5 : (xe - (double)P[i1 * 3 + 0]) / ((double)P[i2 * 3 + 0] - (double)P[i1 * 3 + 0]);
  }
  // Boundary wrapping: snap the crossing coordinate to the next cell's entry corner.
  // (selects, not branches: a branch per coordinate makes the compiler
  // address the three as an array in scratch memory)
  if (inext >= 0) {
    const int q = g.gridptr[inext * 8 + o.ioct];
    const double snap = (double)P[q * 3 + (jface - 1)];
    xe = jface == 1 ? snap : xe;
    ye = jface == 2 ? snap : ye;
    ze = jface == 3 ? snap : ze;
  }
  t.so = so;
  t.u = u;
  t.v = v;
  t.i[0] = i1;
  t.i[1] = i2;
  t.i[2] = i3;
  t.i[3] = i4;
  t.next = inext;
  t.face = iface;
  return true;
}

// Interpolation factors of the point (xe, ye, ze) on face `face` of cell
// `icell` (the formulas of trace_step for that face) and the face's corner
// points.  Used to express a step that leaves the cell through a shared edge
// (a tie between two faces) on the face its class uses.
__host__ __device__ inline void face_uv(const GridTopo& g, int icell, int face, double xe, double ye,
                                        double ze, double& u, double& v, int corners[4]) {
#pragma clang fp contract(off)
  const int8_t flags = g.cellflags[icell];
  const bool ipinx = (flags & 1) != 0;
  const bool ipiny = (flags & 2) != 0;
  const int* gp = g.gridptr + icell * 8;
  const float* P = g.gridpos;
  const int i1 = gp[face_corner(face, 0)], i2 = gp[face_corner(face, 1)];
  const int i3 = gp[face_corner(face, 2)], i4 = gp[face_corner(face, 3)];
  const int jface = face / 2 + 1;
  if (jface == 1) {
    u = (ze - (double)P[i1 * 3 + 2]) / ((double)P[i3 * 3 + 2] - (double)P[i1 * 3 + 2]);
    v = ipiny ? 0.5 : (ye - (double)P[i1 * 3 + 1]) / ((double)P[i2 * 3 + 1] - (double)P[i1 * 3 + 1]);
  } else if (jface == 2) {
    u = (ze - (double)P[i1 * 3 + 2]) / ((double)P[i3 * 3 + 2] - (double)P[i1 * 3 + 2]);
    v = ipinx ? 0.5 : (xe - (double)P[i1 * 3 + 0]) / ((double)P[i2 * 3 + 0] - (double)P[i1 * 3 + 0]);
  } else {
    u = ipiny ? 0.5 : (ye - (double)P[i1 * 3 + 1]) / ((double)P[i3 * 3 + 1] - (double)P[i1 * 3 + 1]);
    v = ipinx ? 0.5 : (xe - (double)P[i1 * 3 + 0]) / ((double)P[i2 * 3 + 0] - (double)P[i1 * 3 + 0]);
  }
  corners[0] = i1;
  corners[1] = i2;
  corners[2] = i3;
  corners[3] = i4;
}

// Cell integral of the short characteristic (shdomsub1.f:3948-3987): linear
// extinction and linear source*extinction across the cell, with the reference's
// tau branches.  The literal constants are the reference's single-precision
// literals promoted to double.
__host__ __device__ inline void cell_integral(double ext0, double ext1, double srcext0,
                                              double srcext1, double so, double& src,
                                              double& abscell, double& transcell) {
#pragma clang fp contract(off)
  const double ext = 0.5 * (ext0 + ext1);
  const double tau = ext * so;
  if (tau >= 0.5) {
    transcell = exp(-tau);
    abscell = 1.0 - transcell;
  } else {
    abscell = tau * (1.0 - 0.5 * tau * (1.0 - (double)0.33333333333f * tau * (1 - 0.25 * tau)));
    transcell = 1.0 - abscell;
  }
  if (tau <= 2.0) {
    if (ext == 0.0) {
      src = 0.0;
    } else {
      src = (0.5 * (srcext0 + srcext1) +
             (double)0.08333333333f * (ext0 * srcext1 - ext1 * srcext0) * so) / ext;
    }
  } else {
    double ext0p = ext0, srcext0p = srcext0;
    if (tau > 4.0) {
      ext0p = ext1 + (ext0 - ext1) * 4.0 / tau;
      if (ext0 > 0.0) srcext0p = srcext0 * ext0p / ext0;
    }
    src = 1.0 / (ext0p + ext1) *
          (srcext0p + srcext1 +
           (ext0p * srcext1 - ext1 * srcext0p) * 2.0 / (ext0p + ext1) *
               (1 - 2 / tau + 2 * transcell / abscell));
  }
  if (src < 0.0) src = 0.0;
}

}  // namespace at3d

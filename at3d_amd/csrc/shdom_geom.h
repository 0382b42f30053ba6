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
//
// Cells refined by adaptive splitting form a binary tree per base cell
// (DIVIDE_CELL, shdomsub1.f:5286-5368): a split cell keeps its index and
// points to its two children, which are consecutive cells (the negative side
// first).  Neighbour entries follow the reference's signed convention in
// 0-based form: c >= 0 is an end cell whose face contains this cell's face
// (same size or larger), -1 a domain boundary, and -(c+2) a cell c whose
// face is larger than the cell's own only before it was split: the
// neighbours across that face are smaller cells found down c's tree.
struct GridTopo {
  const float* gridpos;     // [npts][3]
  const int32_t* gridptr;   // [ncells][8]
  const int32_t* neighptr;  // [ncells][6], see above
  const int8_t* cellflags;  // [ncells] bit0: independent pixel in X, bit1: in Y,
                            // bits 2-3: split direction (1 X, 2 Y, 3 Z) of a split cell
  int npts, ncells;
  const int32_t* treeptr = nullptr;  // [ncells][2] parent, first child (-1: none); null without a tree
};

// Neighbour entry helpers (see GridTopo).
__host__ __device__ inline bool nb_smaller(int v) { return v <= -2; }
__host__ __device__ inline int nb_cell(int v) { return v >= 0 ? v : (v == -1 ? -1 : -v - 2); }
__host__ __device__ inline int nb_encode(int cell, bool smaller) { return smaller ? -cell - 2 : cell; }

// NEXT_CELL (shdomsub1.f:4470-4518): the end cell across face `iface`
// (0..5) of a cell whose neighbour entry there is -(c+2), found by walking
// down c's tree with the crossing point (xe, ye, ze); jface = 1, 2, 3 for an
// x, y, z face.
__host__ __device__ inline int next_cell_down(const GridTopo& g, int c, int iface, int jface, double xe,
                                              double ye, double ze) {
  int ic = c;
  while (g.treeptr[ic * 2 + 1] >= 0) {
    const int dir = (g.cellflags[ic] >> 2) & 3;
    const int ic1 = g.treeptr[ic * 2 + 1];
    if (dir == jface) {
      ic = ic1 + 1 - (iface & 1);
    } else {
      const float* split = g.gridpos + (size_t)g.gridptr[ic1 * 8 + 7] * 3;
      const double coord = dir == 1 ? xe : (dir == 2 ? ye : ze);
      ic = coord > (double)split[dir - 1] ? ic1 + 1 : ic1;
    }
  }
  return ic;
}

// The tree descent of LOCATE_GRID_CELL (shdomsub2.f:4181-4199): the end cell
// of base cell `c` that holds (x, y, z).
__host__ __device__ inline int locate_in_tree(const GridTopo& g, int c, double x, double y, double z) {
  if (!g.treeptr) return c;
  int ic = c;
  while (g.treeptr[ic * 2 + 1] >= 0) {
    const int dir = (g.cellflags[ic] >> 2) & 3;
    int child = g.treeptr[ic * 2 + 1] + 1;
    const float* p1 = g.gridpos + (size_t)g.gridptr[child * 8] * 3;
    const double coord = dir == 1 ? x : (dir == 2 ? y : z);
    if (dir >= 1 && dir <= 3 && coord < (double)p1[dir - 1]) child -= 1;
    ic = child;
  }
  return ic;
}

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
  const int nb = g.neighptr[icell * 6 + iface];
  int inext = nb;
  // a face onto smaller cells: the next cell down the neighbour's tree, and
  // its (smaller, more accurate) face holds the interpolation corners
  int kcell = icell, kface = iface;
  if (nb_smaller(nb)) {
    inext = next_cell_down(g, nb_cell(nb), iface, jface, xe, ye, ze);
    kcell = inext;
    kface = iface ^ 1;
  }
  const int* gp = g.gridptr + kcell * 8;
  const int i1 = gp[face_corner(kface, 0)], i2 = gp[face_corner(kface, 1)];
  const int i3 = gp[face_corner(kface, 2)], i4 = gp[face_corner(kface, 3)];
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
// literals promoted to double.  The source is clamped at zero for I, not for
// the Q and U of a polarized sweep (clamp false; BACK_INT_GRID3D :3640).
__host__ __device__ inline void cell_integral(double ext0, double ext1, double srcext0,
                                              double srcext1, double so, double& src,
                                              double& abscell, double& transcell, bool clamp = true) {
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
  if (clamp && src < 0.0) src = 0.0;
}

}  // namespace at3d

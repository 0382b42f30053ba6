// Adaptive cell splitting (SPLIT_GRID and the routines it calls,
// shdomsub1.f:4703-5794) on the host.
//
// The reference refines its grid serially: cells are ranked by a criterion
// computed from the source function, divided one at a time, and every
// division updates the cell tree and the neighbour pointers of the cells
// around it.  The decisions depend on the order of the divisions, so this
// restatement keeps the reference's order of operations and its
// single-precision arithmetic; the GPU works on the refined grid afterwards.
// Indices are 0-based; neighbour entries use the encoding of GridTopo
// (shdom_geom.h).  The splitting runs a few times per solve, on host copies
// of the source and radiance fields.
#include "split_grid.h"

#include <math.h>

#include <algorithm>
#include <array>
#include <stdexcept>

#pragma clang fp contract(off)

namespace at3d {
namespace {

struct SplitError : std::runtime_error {
  explicit SplitError(const std::string& m) : std::runtime_error(m) {}
};

// Corner pairs of the four cell edges along direction idir (1 x, 2 y, 3 z):
// corner c = bitx + 2 bity + 4 bitz, the pair differing in the idir bit.
const int8_t kEdgeCorner[3][4][2] = {{{0, 1}, {2, 3}, {4, 5}, {6, 7}},
                                     {{0, 2}, {1, 3}, {4, 6}, {5, 7}},
                                     {{0, 4}, {1, 5}, {2, 6}, {3, 7}}};
// The two faces (0..5: -x +x -y +y -z +z) that hold each of those edges.
const int8_t kEdgeFaces[3][4][2] = {{{2, 4}, {3, 4}, {2, 5}, {3, 5}},
                                    {{0, 4}, {1, 4}, {0, 5}, {1, 5}},
                                    {{0, 2}, {1, 2}, {0, 3}, {1, 3}}};

inline float pos(const Plan& p, int pt, int dim) { return p.gridpos[(size_t)pt * 3 + dim - 1]; }
inline int& nbr(Plan& p, int c, int f) { return p.neighptr[(size_t)c * 6 + f]; }
inline int& child(Plan& p, int c) { return p.treeptr[(size_t)c * 2 + 1]; }
inline int split_dir(const Plan& p, int c) { return (p.cellflags[c] >> 2) & 3; }

GridTopo topo(const Plan& p) {
  GridTopo g{p.gridpos.data(), p.gridptr.data(), p.neighptr.data(), p.cellflags.data(), p.npts, p.ncells};
  g.treeptr = p.treeptr.data();
  return g;
}

// The end cell across face `iface` of `icell` at the point (xe, ye, ze)
// (NEXT_CELL, shdomsub1.f:4470-4518); -1 at a domain boundary.
int next_cell(const Plan& p, int icell, int iface, double xe, double ye, double ze) {
  const int v = p.neighptr[(size_t)icell * 6 + iface];
  if (!nb_smaller(v)) return nb_cell(v);
  return next_cell_down(topo(p), nb_cell(v), iface, iface / 2 + 1, xe, ye, ze);
}

// Sets the neighbour across face `iface` to end cell `in` for every cell of
// `icell`'s tree that borders the face (INHERIT_NEIGHBOR, shdomsub1.f:5454-5506).
void inherit_neighbor(Plan& p, int icell, int iface, int in) {
  const int jface = iface / 2 + 1;
  std::vector<int> stack;
  int ic = icell;
  for (;;) {
    nbr(p, ic, iface) = in;
    if (child(p, ic) < 0) {
      if (stack.empty()) return;
      ic = stack.back();
      stack.pop_back();
    } else if (split_dir(p, ic) == jface) {
      ic = child(p, ic) + (iface & 1);
    } else {
      stack.push_back(child(p, ic) + 1);
      ic = child(p, ic);
    }
  }
}

// Resolves face `iface` of the new cell `ic` against the cells down its
// neighbour's tree, and updates those cells' pointers back to it
// (MATCH_NEIGHBOR_FACE, shdomsub1.f:5371-5451).
void match_neighbor_face(Plan& p, int iface, int ic) {
  int in = nb_cell(nbr(p, ic, iface));
  if (in < 0) return;
  const int jface = iface / 2 + 1;
  const int opp = iface ^ 1;
  const int ic1 = p.gridptr[(size_t)ic * 8], ic8 = p.gridptr[(size_t)ic * 8 + 7];
  const int dir1 = jface % 3 + 1, dir2 = (jface + 1) % 3 + 1;
  float cpos[4];
  cpos[dir1] = (pos(p, ic1, dir1) + pos(p, ic8, dir1)) / 2;
  cpos[dir2] = (pos(p, ic1, dir2) + pos(p, ic8, dir2)) / 2;
  bool done = false;
  while (!done && ic != in) {
    const int in1 = p.gridptr[(size_t)in * 8], in8 = p.gridptr[(size_t)in * 8 + 7];
    // the new face lies within the neighbour's face
    if (pos(p, ic1, dir1) >= pos(p, in1, dir1) && pos(p, ic8, dir1) <= pos(p, in8, dir1) &&
        pos(p, ic1, dir2) >= pos(p, in1, dir2) && pos(p, ic8, dir2) <= pos(p, in8, dir2))
      nbr(p, ic, iface) = nb_encode(in, child(p, in) >= 0);
    // the neighbour's face lies within the new face: it (and its tree) now
    // border the new cell; otherwise its pointer no longer ends at an end cell
    if (pos(p, in1, dir1) >= pos(p, ic1, dir1) && pos(p, in8, dir1) <= pos(p, ic8, dir1) &&
        pos(p, in1, dir2) >= pos(p, ic1, dir2) && pos(p, in8, dir2) <= pos(p, ic8, dir2)) {
      inherit_neighbor(p, in, opp, ic);
    } else {
      const int c = nb_cell(nbr(p, in, opp));
      if (c >= 0) nbr(p, in, opp) = nb_encode(c, true);
    }
    if (child(p, in) < 0) {
      done = true;
    } else {
      const int dir = split_dir(p, in);
      const int inn = child(p, in);
      if (dir == jface)
        in = inn + 1 - (iface & 1);
      else
        in = cpos[dir] > pos(p, p.gridptr[(size_t)inn * 8 + 7], dir) ? inn + 1 : inn;
    }
  }
}

// An existing grid point at (xp, yp, zp) on face `iface` of `icell`, looked
// up among the cells across that face (MATCH_GRID_POINT, shdomsub1.f:5599-5699);
// -1 if there is none.
int match_grid_point(const Plan& p, float xp, float yp, float zp, int icell, int iface) {
  const int idir = iface / 2 + 1, kface = iface ^ 1;
  int ic = nb_cell(p.neighptr[(size_t)icell * 6 + iface]);
  if (ic < 0) return -1;
  std::vector<int> stack;
  for (;;) {
    while (p.treeptr[(size_t)ic * 2 + 1] < 0) {
      for (int i = 0; i < 4; i++) {
        const int ipt = p.gridptr[(size_t)ic * 8 + face_corner(kface, i)];
        if (xp == pos(p, ipt, 1) && yp == pos(p, ipt, 2) && zp == pos(p, ipt, 3)) return ipt;
      }
      if (stack.empty()) return -1;
      ic = stack.back();
      stack.pop_back();
    }
    const int dir = split_dir(p, ic);
    if (dir == 0) throw SplitError("MATCH_GRID_POINT: No split direction");
    const int ic1 = p.treeptr[(size_t)ic * 2 + 1];
    if (dir == idir) {
      ic = ic1 + 1 - (iface & 1);
    } else {
      ic = ic1;
      const float c = dir == 1 ? xp : (dir == 2 ? yp : zp);
      const float split = pos(p, p.gridptr[(size_t)ic1 * 8 + 7], dir);
      if (c == split)
        stack.push_back(ic + 1);   // the point may lie on either child's face
      else if (c > split)
        ic = ic + 1;
    }
  }
}

struct NewPoint {
  int p1, p2, ip;   // parents and the new point (ip < 0: an existing point was matched)
};

// Grid pointers of the two new cells and the up to four new points on the
// edges crossing the split (NEW_GRID_POINTS, shdomsub1.f:5509-5596).
void new_grid_points(Plan& p, int idir, int icell, int newcell, NewPoint np[4]) {
  for (int k = 0; k < 8; k++) {
    p.gridptr[(size_t)newcell * 8 + k] = p.gridptr[(size_t)icell * 8 + k];
    p.gridptr[(size_t)(newcell + 1) * 8 + k] = p.gridptr[(size_t)icell * 8 + k];
  }
  for (int i = 0; i < 4; i++) {
    const int i1 = kEdgeCorner[idir - 1][i][0], i2 = kEdgeCorner[idir - 1][i][1];
    const int ip1 = p.gridptr[(size_t)icell * 8 + i1], ip2 = p.gridptr[(size_t)icell * 8 + i2];
    const float xp = (pos(p, ip1, 1) + pos(p, ip2, 1)) / 2;
    const float yp = (pos(p, ip1, 2) + pos(p, ip2, 2)) / 2;
    const float zp = (pos(p, ip1, 3) + pos(p, ip2, 3)) / 2;
    // the point may exist in one of the three cells around the edge
    const int f1 = kEdgeFaces[idir - 1][i][0], f2 = kEdgeFaces[idir - 1][i][1];
    int m = match_grid_point(p, xp, yp, zp, icell, f1);
    if (m < 0) m = match_grid_point(p, xp, yp, zp, icell, f2);
    if (m < 0) {
      const int c2 = nb_cell(p.neighptr[(size_t)icell * 6 + f1]);
      if (c2 >= 0) m = match_grid_point(p, xp, yp, zp, c2, f2);
    }
    if (m < 0) {
      m = p.npts++;
      p.gridpos.push_back(xp);
      p.gridpos.push_back(yp);
      p.gridpos.push_back(zp);
      np[i] = NewPoint{ip1, ip2, m};
    } else {
      np[i] = NewPoint{ip1, ip2, -1};
    }
    p.gridptr[(size_t)newcell * 8 + i2] = m;
    p.gridptr[(size_t)(newcell + 1) * 8 + i1] = m;
  }
}

// Divides end cell `icell` in direction idir (1 x, 2 y, 3 z) into two new
// cells (DIVIDE_CELL, shdomsub1.f:5286-5368).
void divide_cell(Plan& p, int icell, int idir, NewPoint np[4]) {
  if (p.treeptr[(size_t)icell * 2 + 1] >= 0) throw SplitError("DIVIDE_CELL: Cannot divide already split cell.");
  const int newcell = p.ncells;
  p.ncells += 2;
  p.gridptr.resize((size_t)p.ncells * 8, 0);
  p.neighptr.resize((size_t)p.ncells * 6, -1);
  p.treeptr.resize((size_t)p.ncells * 2, -1);
  p.cellflags.resize(p.ncells, 0);
  child(p, icell) = newcell;
  p.treeptr[(size_t)newcell * 2] = icell;
  p.treeptr[(size_t)(newcell + 1) * 2] = icell;
  p.cellflags[icell] = (int8_t)(p.cellflags[icell] | (idir << 2));
  p.cellflags[newcell] = (int8_t)(p.cellflags[icell] & 3);
  p.cellflags[newcell + 1] = (int8_t)(p.cellflags[icell] & 3);
  new_grid_points(p, idir, icell, newcell, np);
  const int fplus = 2 * idir - 1, fminus = 2 * idir - 2;
  for (int iface = 0; iface < 6; iface++) {
    const int v = nbr(p, icell, iface);
    if (v == icell) {
      nbr(p, newcell, iface) = newcell;
      nbr(p, newcell + 1, iface) = newcell + 1;
    } else if (iface == fplus) {
      nbr(p, newcell, iface) = newcell + 1;
      nbr(p, newcell + 1, iface) = v;
    } else if (iface == fminus) {
      nbr(p, newcell + 1, iface) = newcell;
      nbr(p, newcell, iface) = v;
    } else {
      nbr(p, newcell, iface) = v;
      nbr(p, newcell + 1, iface) = v;
    }
    match_neighbor_face(p, iface, newcell);
    match_neighbor_face(p, iface, newcell + 1);
  }
}

// The adaptive criterion of end cell `icell` in x, y and z, its largest value
// and direction (CELL_SPLIT_TEST, shdomsub1.f:5702-5791): per edge crossing
// the direction, the root-sum-square of the extinction x source difference
// over the SH terms, divided by the mean extinction, times the edge's
// absorptance; averaged over the four edges.
void cell_split_test(const Plan& p, const SplitFields& f, int icell, float& maxadapt, int& idir) {
  const float C0 = 0.282095f;
  const float* E = p.total_ext.data();
  maxadapt = -1.0f;
  idir = 0;
  for (int id = 1; id <= 3; id++) {
    float sum1 = 0.0f;
    int num = 0;
    for (int ie = 0; ie < 4; ie++) {
      const int ip1 = p.gridptr[(size_t)icell * 8 + kEdgeCorner[id - 1][ie][0]];
      const int ip2 = p.gridptr[(size_t)icell * 8 + kEdgeCorner[id - 1][ie][1]];
      if (ip1 == ip2) continue;
      num++;
      const float* s1 = &f.S[(size_t)ip1 * f.ld];
      const float* s2 = &f.S[(size_t)ip2 * f.ld];
      const int ns1 = f.ns[ip1], ns2 = f.ns[ip2], ns = std::min(ns1, ns2);
      float jay = 0.0f;
      for (int j = 0; j < ns; j++) {
        const float d = E[ip2] * s2[j] - E[ip1] * s1[j];
        jay = jay + d * d;
      }
      for (int j = ns; j < ns1; j++) {
        const float d = E[ip1] * s1[j];
        jay = jay + d * d;
      }
      for (int j = ns; j < ns2; j++) {
        const float d = E[ip2] * s2[j];
        jay = jay + d * d;
      }
      const float ext = 0.5f * (E[ip1] + E[ip2]);
      jay = ext > 0.0f ? C0 * sqrtf(jay) / ext : 0.0f;
      const float tau = fabsf(ext * (pos(p, ip2, id) - pos(p, ip1, id)));
      sum1 = sum1 + fabsf(jay) * (1 - expf(-tau));
    }
    const float a = num > 0 ? sum1 / (float)num : 0.0f;
    if (a > maxadapt) {
      maxadapt = a;
      idir = id;
    }
  }
}

// Whether end cell `icell` should be divided to keep the adaptive grid
// smooth, and in which direction (0: no) (GRID_SMOOTH_TEST, shdomsub1.f:5794-5890):
// finer cells on both sides, neighbours split the same other way on both
// sides, or a neighbour more than 2.5 times finer beside a cell with source.
int grid_smooth_test(const Plan& p, const SplitFields& f, int icell) {
  const int8_t edge[3][2] = {{0, 1}, {0, 2}, {0, 4}};
  int idir = 0;
  float sizeratio = 1.0f;
  for (int id = 1; id <= 3; id++) {
    const int ip1 = p.gridptr[(size_t)icell * 8 + edge[id - 1][0]];
    const int ip2 = p.gridptr[(size_t)icell * 8 + edge[id - 1][1]];
    if (ip1 == ip2) continue;
    const float cur = fabsf(pos(p, ip1, id) - pos(p, ip2, id));
    const float inv = 1.0f / cur;
    float gridsize[2];
    int dirsplit[2];
    for (int j = 0; j < 2; j++) {
      dirsplit[j] = 0;
      const int iface = 2 * (id - 1) + j;
      double xe = 0.0, ye = 0.0, ze = 0.0;
      for (int i = 0; i < 4; i++) {
        const int ip = p.gridptr[(size_t)icell * 8 + face_corner(iface, i)];
        xe = xe + (double)(pos(p, ip, 1) * 0.25f);
        ye = ye + (double)(pos(p, ip, 2) * 0.25f);
        ze = ze + (double)(pos(p, ip, 3) * 0.25f);
      }
      const int inc = next_cell(p, icell, iface, xe, ye, ze);
      if (inc < 0) {
        gridsize[j] = cur;
      } else {
        gridsize[j] = fabsf(pos(p, p.gridptr[(size_t)inc * 8], id) - pos(p, p.gridptr[(size_t)inc * 8 + 7], id));
        if (p.treeptr[(size_t)inc * 2] < 0) gridsize[j] = cur;   // a base cell
        if (id <= 2 && ((p.cellflags[inc] >> (id - 1)) & 1)) gridsize[j] = cur;
      }
      const int v = p.neighptr[(size_t)icell * 6 + iface];
      if (nb_smaller(v)) dirsplit[j] = split_dir(p, nb_cell(v));
    }
    if (gridsize[0] * inv < 0.75f && gridsize[1] * inv < 0.75f) idir = id;
    if (dirsplit[0] > 0 && dirsplit[0] != id && dirsplit[0] == dirsplit[1]) idir = dirsplit[0];
    const float ratio = std::min(gridsize[0] * inv, gridsize[1] * inv);
    if (ratio < 0.4f && ratio < sizeratio) {
      int isum = 0;
      for (int i = 0; i < 8; i++) isum += f.ns[p.gridptr[(size_t)icell * 8 + i]];
      if (isum > 0) {
        idir = id;
        sizeratio = ratio;
      }
    }
  }
  return idir;
}

// Sorts x into decreasing order carrying y along, with the partition and
// insertion steps of Singleton's quicksort (R. C. Singleton, CACM Algorithm
// 347, 1969), the method of the reference's SSORT (shdomsub2.f:4961-5217),
// so that equal criteria keep the reference's relative order.
void sort_decreasing(std::vector<float>& x, std::vector<int64_t>& y, int n) {
  if (n < 1) return;
  // 1-based views; the decreasing sort is an increasing sort of -x
  float* X = x.data() - 1;
  int64_t* Y = y.data() - 1;
  for (int i = 1; i <= n; i++) X[i] = -X[i];
  int il[64], iu[64];
  int m = 1, i = 1, j = n, k, l, ij;
  float r = 0.375f, t;
  int64_t ty;
  enum { SELECT, SPLIT, POP, SMALL } step = SELECT;
  bool adjust = true;   // enter SELECT through the r update (label 110)
  for (;;) {
    if (step == SELECT) {
      if (adjust) {
        if (i == j) {
          step = POP;
          continue;
        }
        r = r <= 0.5898437f ? r + 3.90625e-2f : r - 0.21875f;
      }
      adjust = true;
      k = i;
      ij = i + (int)((float)(j - i) * r);
      t = X[ij];
      ty = Y[ij];
      if (X[i] > t) {
        X[ij] = X[i]; X[i] = t; t = X[ij];
        Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij];
      }
      l = j;
      if (X[j] < t) {
        X[ij] = X[j]; X[j] = t; t = X[ij];
        Y[ij] = Y[j]; Y[j] = ty; ty = Y[ij];
        if (X[i] > t) {
          X[ij] = X[i]; X[i] = t; t = X[ij];
          Y[ij] = Y[i]; Y[i] = ty; ty = Y[ij];
        }
      }
      for (;;) {
        do l--; while (X[l] > t);
        do k++; while (X[k] < t);
        if (k > l) break;
        std::swap(X[l], X[k]);
        std::swap(Y[l], Y[k]);
      }
      if (m > 63) throw SplitError("sort stack exceeded");
      if (l - i > j - k) {
        il[m] = i; iu[m] = l; i = k;
      } else {
        il[m] = k; iu[m] = j; j = l;
      }
      m++;
      step = SPLIT;
    } else if (step == POP) {
      m--;
      if (m == 0) break;
      i = il[m];
      j = iu[m];
      step = SPLIT;
    } else if (step == SPLIT) {
      if (j - i >= 1) {
        step = SELECT;
        adjust = false;   // label 120: partition again without the r update
        continue;
      }
      if (i == 1) {
        step = SELECT;
        continue;
      }
      i--;
      step = SMALL;
    } else {   // SMALL: straight insertion over a short segment (labels 170-180)
      for (;;) {
        i++;
        if (i == j) break;
        t = X[i + 1];
        ty = Y[i + 1];
        if (X[i] <= t) continue;
        k = i;
        do {
          X[k + 1] = X[k];
          Y[k + 1] = Y[k];
          k--;
        } while (t < X[k]);
        X[k + 1] = t;
        Y[k + 1] = ty;
      }
      step = POP;
    }
  }
  for (int q = 1; q <= n; q++) X[q] = -X[q];
}

// Optics of one new point from the property grid with delta-M scaling
// (INTERPOLATE_POINT, shdomsub1.f:5020-5072).
struct PointOptics {
  std::vector<float> ext, alb, wt;   // [npart], [npart], [npart][8]
  std::vector<int32_t> iph;          // [npart][8]
  float total_ext = 0.f;
};

void new_point_optics(const Plan& p, float x, float y, float z, PointOptics& o) {
  const int npart = p.npart, nl = p.nleg + 1, ML = p.ml;
  o.ext.assign(npart, 0.f);
  o.alb.assign(npart, 0.f);
  o.wt.assign((size_t)npart * 8, 0.f);
  o.iph.assign((size_t)npart * 8, 1);
  o.total_ext = 0.0f;
  for (int ipa = 0; ipa < npart; ipa++) {
    float e, a;
    point_optics(p, ipa, x, y, z, e, a, &o.iph[(size_t)ipa * 8], &o.wt[(size_t)ipa * 8]);
    if (p.d.deltam) {
      const int32_t* ip = &o.iph[(size_t)ipa * 8];
      const float* w = &o.wt[(size_t)ipa * 8];
      float fd;
      if (w[0] >= 0.999f) {
        fd = p.legen[(size_t)(ip[0] - 1) * nl + ML + 1];
      } else {
        fd = 0.0f;
        for (int q = 0; q < 8; q++) fd = fd + p.legen[(size_t)(ip[q] - 1) * nl + ML + 1] * w[q];
      }
      const float a0 = a;
      e = (1.0f - a0 * fd) * e;
      a = (1.0f - fd) * a0 / (1.0f - a0 * fd);
    }
    // (the reference adds the gas absorption first, which is zero here)
    o.total_ext = o.total_ext + e;
    o.ext[ipa] = e;
    o.alb[ipa] = a;
  }
}

// The scattering albedo and albedo-weighted, delta-M renormalised phase
// function of a new point (INTERPOLATE_POINT, shdomsub1.f:5121-5172).
void new_point_phase(const Plan& p, const PointOptics& o, std::vector<float>& legent, float& albedo) {
  const int npart = p.npart, nl = p.nleg + 1, ML = p.ml;
  legent.assign(nl, 0.f);
  std::vector<float> l1(nl);
  double alb = 0.0;
  for (int ipa = 0; ipa < npart; ipa++) {
    const double scat = (double)(o.ext[ipa] * o.alb[ipa]);
    alb = alb + scat;
    const int32_t* ip = &o.iph[(size_t)ipa * 8];
    const float* w = &o.wt[(size_t)ipa * 8];
    if (w[0] >= 0.999f) {
      for (int l = 0; l < nl; l++) l1[l] = p.legen[(size_t)(ip[0] - 1) * nl + l];
    } else {
      std::fill(l1.begin(), l1.end(), 0.f);
      for (int q = 0; q < 8; q++) {
        if (w[q] <= 1e-5f) continue;
        for (int l = 0; l < nl; l++) l1[l] = l1[l] + p.legen[(size_t)(ip[q] - 1) * nl + l] * w[q];
      }
    }
    if (p.d.deltam) {
      const float fd = l1[ML + 1];
      for (int l = 0; l <= ML; l++) l1[l] = l1[l] / (1 - fd);
    }
    for (int l = 0; l < nl; l++) legent[l] = (float)((double)legent[l] + scat * (double)l1[l]);
  }
  if (alb > 1e-10) {
    for (int l = 0; l < nl; l++) legent[l] = (float)((double)legent[l] / alb);
  } else {
    for (int l = 0; l < nl; l++) legent[l] = legent[l] / (float)npart;
  }
  alb = o.total_ext > 1e-10f ? alb / (double)o.total_ext : 0.0;
  legent[0] = 1.0f;
  albedo = (float)alb;
}

// CALC_SOURCE_PNT (shdomsub1.f:823-910) at a new point with NSTOKES 3: the
// I, Q and U source terms 0..ns-1 from the radiance terms 0..nr-1 and the
// solar direct beam, with the point's mixed, delta-M scaled elements a1
// (legent, from new_point_phase), a2, a3 and b1 (made as COMPUTE_SOURCE
// mixes them).  Q and U start at the fifth term, as the reference's loops do.
void new_point_source_pol(const Plan& p, const PointOptics& o, const std::vector<float>& a1, float albedo,
                          float flux0, int ns, int nr, const float* RI, const float* RQ, const float* RU, float* SI,
                          float* SQ, float* SU) {
#pragma clang fp contract(off)
  const int ML = p.ml, nl = p.nleg + 1;
  LegTable t1 = plan_leg_table(p);
  t1.npw = 8;
  const LegTable te[3] = {LegTable{p.legen2.data(), nl, ML, p.legen2.size(), 8},
                          LegTable{p.legen3.data(), nl, ML, p.legen3.size(), 8},
                          LegTable{p.legen5.data(), nl, ML, p.legen5.size(), 8}};
  std::vector<float> a2(ML + 1), a3(ML + 1), b1(ML + 1);
  float* el[3] = {a2.data(), a3.data(), b1.data()};
  for (int e = 0; e < 3; e++)
    point_element_constants(t1, te[e], p.npart, p.d.deltam != 0, o.ext.data(), o.alb.data(), o.iph.data(),
                            o.wt.data(), 1, el[e]);
  for (int j = 0; j < ns; j++) {
    const int l = p.lofj[j];
    float sI = 0.f, sQ = 0.f, sU = 0.f;
    if (j < nr) {
      sI = a1[l] * RI[j];
      sI = sI + b1[l] * RQ[j];
      if (j >= 4) {
        sQ = b1[l] * RI[j] + a2[l] * RQ[j];
        sU = a3[l] * RU[j];
      }
    }
    sI = albedo * sI;
    sQ = albedo * sQ;
    sU = albedo * sU;
    sI = sI + flux0 * albedo * a1[l] * p.ylmsun[j];
    if (j >= 4) sQ = sQ + flux0 * albedo * b1[l] * p.ylmsun[j];
    SI[j] = sI;
    SQ[j] = sQ;
    SU[j] = sU;
  }
}

struct Splitter {
  Plan& p;
  SplitFields& f;
  DirectBeamGeom beam;
  std::vector<float> extdirp;
  bool beam_ready = false;
  int64_t sum_ns = 0, sum_nr = 0;
  int old_npts = 0;
  std::vector<PointOptics> added;   // optics of the points appended in this call

  Splitter(Plan& p_, SplitFields& f_) : p(p_), f(f_) {
    for (int i = 0; i < p.npts; i++) {
      sum_ns += f.ns[i];
      sum_nr += f.nr[i];
    }
    old_npts = p.npts;
  }

  // New points of the batch whose medium, direct beam and source are still
  // to be computed: (point, first parent, second parent).
  std::vector<std::array<int, 3>> pending;

  // INTERPOLATE_POINT (shdomsub1.f:4934-5283) for the new points of one
  // division: the radiance (mean of the parents' expansions) and the term
  // counts at once, since the smoothing test and the capacity bounds read
  // them; the medium, direct beam and source wait for finish_points (they
  // depend only on the point and its parents).
  void interpolate_points(const NewPoint np[4]) {
    for (int i = 0; i < 4; i++) {
      if (np[i].ip < 0) continue;
      const int ip1 = np[i].p1, ip2 = np[i].p2, ip = np[i].ip;
      const int ld = f.ld;
      const int nr1 = f.nr[ip1], nr2 = f.nr[ip2], nr = std::max(nr1, nr2);
      f.R.resize((size_t)(ip + 1) * ld, 0.f);
      float* R = &f.R[(size_t)ip * ld];
      for (int j = 0; j < nr; j++) {
        const float r1 = j < nr1 ? f.R[(size_t)ip1 * ld + j] : 0.0f;
        const float r2 = j < nr2 ? f.R[(size_t)ip2 * ld + j] : 0.0f;
        R[j] = 0.5f * (r1 + r2);
      }
      f.nr.push_back(nr);
      if (f.nst == 3)
        for (int k = 0; k < 2; k++) {
          std::vector<float>& Rk = f.Rqu[k];
          Rk.resize((size_t)(ip + 1) * ld, 0.f);
          for (int j = 0; j < nr; j++) {
            const float r1 = j < nr1 ? Rk[(size_t)ip1 * ld + j] : 0.0f;
            const float r2 = j < nr2 ? Rk[(size_t)ip2 * ld + j] : 0.0f;
            Rk[(size_t)ip * ld + j] = 0.5f * (r1 + r2);
          }
          f.Squ[k].resize((size_t)(ip + 1) * ld, 0.f);
          f.DSqu[k].resize((size_t)(ip + 1) * ld, 0.f);
        }
      const int ns = std::max(f.ns[ip1], f.ns[ip2]);
      f.S.resize((size_t)(ip + 1) * ld, 0.f);
      f.DS.resize((size_t)(ip + 1) * ld, 0.f);
      f.ns.push_back(ns);
      f.nos.push_back(0);
      p.total_ext.push_back(0.0f);
      added.emplace_back();
      sum_ns += ns;
      sum_nr += nr;
      pending.push_back({ip, ip1, ip2});
    }
  }

  // The medium from the property grid, the direct beam and the source
  // (CALC_SOURCE_PNT_UNPOL, shdomsub1.f:913-958, kept to the parents' larger
  // truncation) of the pending points, on the host threads.
  void finish_points() {
    if (pending.empty()) return;
    if (!beam_ready) {
      direct_beam_setup(p, extdirp, beam);
      beam_ready = true;
    }
    const int n = (int)pending.size();
    std::vector<char> bad(n, 0);
    parallel_for(n, [&](int k) {
      const int ip = pending[k][0];
      const float x = pos(p, ip, 1), y = pos(p, ip, 2), z = pos(p, ip, 3);
      PointOptics& o = added[ip - old_npts];
      new_point_optics(p, x, y, z, o);
      p.total_ext[ip] = o.total_ext;
      float dirflux = 0.f;
      if (!direct_beam_point(beam, x, y, z, dirflux, nullptr)) {
        bad[k] = 1;
        return;
      }
      std::vector<float> legent;
      float albedo;
      new_point_phase(p, o, legent, albedo);
      const float flux0 = dirflux * (1.0f / fabsf(p.d.solarmu));
      const int ld = f.ld, ns = f.ns[ip], nr = f.nr[ip];
      const float* R = &f.R[(size_t)ip * ld];
      float* S = &f.S[(size_t)ip * ld];
      if (f.nst == 3) {
        new_point_source_pol(p, o, legent, albedo, flux0, ns, nr, R, &f.Rqu[0][(size_t)ip * ld],
                             &f.Rqu[1][(size_t)ip * ld], S, &f.Squ[0][(size_t)ip * ld], &f.Squ[1][(size_t)ip * ld]);
        return;
      }
      for (int j = 0; j < ns; j++) {
        const int l = p.lofj[j];
        float s = flux0 * albedo * legent[l] * p.ylmsun[j];
        if (j < nr) s = s + albedo * legent[l] * R[j];
        S[j] = s;
      }
    });
    pending.clear();
    for (char b : bad)
      if (b) throw SplitError("DIRECT_BEAM_PROP: ray left the property grid");
  }

  bool over(int maxcells, int maxpts, int maxwork, int maxsh, int maxrad) const {
    return p.ncells > maxcells || p.npts > maxpts || (int64_t)p.npts * p.nphi0max > maxwork ||
           sum_ns > maxsh || sum_nr > maxrad;
  }

  // The species-major optics arrays of the plan with the appended points.
  void merge_optics() {
    const int n0 = old_npts, n1 = p.npts, npart = p.npart;
    if (n1 == n0) return;
    std::vector<float> ext((size_t)npart * n1), alb((size_t)npart * n1), wt((size_t)npart * n1 * 8);
    std::vector<int32_t> iph((size_t)npart * n1 * 8);
    for (int ipa = 0; ipa < npart; ipa++) {
      std::copy_n(&p.extinct[(size_t)ipa * n0], n0, &ext[(size_t)ipa * n1]);
      std::copy_n(&p.albedo[(size_t)ipa * n0], n0, &alb[(size_t)ipa * n1]);
      std::copy_n(&p.iphase[(size_t)ipa * n0 * 8], (size_t)n0 * 8, &iph[(size_t)ipa * n1 * 8]);
      std::copy_n(&p.phaseinterpwt[(size_t)ipa * n0 * 8], (size_t)n0 * 8, &wt[(size_t)ipa * n1 * 8]);
      for (int k = 0; k < n1 - n0; k++) {
        const PointOptics& o = added[k];
        const size_t ip = (size_t)n0 + k;
        ext[(size_t)ipa * n1 + ip] = o.ext[ipa];
        alb[(size_t)ipa * n1 + ip] = o.alb[ipa];
        std::copy_n(&o.iph[(size_t)ipa * 8], 8, &iph[((size_t)ipa * n1 + ip) * 8]);
        std::copy_n(&o.wt[(size_t)ipa * 8], 8, &wt[((size_t)ipa * n1 + ip) * 8]);
      }
    }
    p.extinct.swap(ext);
    p.albedo.swap(alb);
    p.iphase.swap(iph);
    p.phaseinterpwt.swap(wt);
  }
};

}  // namespace

int split_grid(Plan& p, SplitFields& f, bool dosplit, bool& outofmem, float cursplitacc, float& splitcrit,
               std::string& err) {
  if (p.treeptr.empty()) p.treeptr.assign((size_t)p.ncells * 2, -1);
  Splitter sp(p, f);
  try {
    if (dosplit) {
      bool outofmem0 = outofmem;
      int icell1 = 0;
      std::vector<float> crit;
      std::vector<int64_t> ind;
      // batches: the cells made by one batch are tested in the next
      while (icell1 < p.ncells) {
        crit.clear();
        ind.clear();
        {
          // the end cells' criteria on the host threads, kept in cell order
          const int nc = p.ncells - icell1;
          std::vector<float> ca(nc);
          std::vector<int> cd(nc, -1);
          parallel_for((nc + 255) / 256, [&](int b) {
            for (int k = b * 256; k < std::min(nc, b * 256 + 256); k++) {
              const int ic = icell1 + k;
              if (p.treeptr[(size_t)ic * 2 + 1] >= 0) continue;
              cell_split_test(p, f, ic, ca[k], cd[k]);
            }
          });
          for (int k = 0; k < nc; k++)
            if (cd[k] >= 0) {
              crit.push_back(ca[k]);
              ind.push_back(4 * (int64_t)(icell1 + k) + cd[k]);
            }
        }
        icell1 = p.ncells;
        const int n = (int)crit.size();
        sort_decreasing(crit, ind, n);
        // room left for the cells split to smooth the grid (the reference's
        // single-precision bounds)
        const float frac = 0.03f;
        const int maxcells = (int)((float)p.maxic - frac * (float)(p.maxic - p.ncells) - 2.0f);
        const int maxpts = (int)((float)p.maxig - frac * (float)(p.maxig - p.npts) - 4.0f);
        const int maxwork = (int)((float)p.maxido - frac * (float)(p.maxido - p.nphi0max * p.npts) -
                                  (float)(4 * p.nphi0max));
        const int maxsh = (int)((float)p.maxiv - frac * (float)(p.maxiv - sp.sum_ns) - (float)(4 * p.nlm));
        const int maxrad = (int)((float)(p.maxiv + p.maxig) - frac * (float)(p.maxiv + p.maxig - sp.sum_nr) -
                                 (float)(4 * p.nlm));
        int i = 0;
        while (i < n && crit[i] > cursplitacc && !outofmem0) {
          if (sp.over(maxcells, maxpts, maxwork, maxsh, maxrad)) {
            outofmem0 = true;
          } else {
            NewPoint np[4];
            divide_cell(p, (int)(ind[i] >> 2), (int)(ind[i] & 3), np);
            sp.interpolate_points(np);
          }
          i++;
        }
        // split more cells to make the adaptive grid smoother
        while (!outofmem && i < n) {
          if (sp.over(p.maxic - 2, p.maxig - 4, p.maxido - 4 * p.nphi0max, p.maxiv - 4 * p.nlm,
                      p.maxiv + p.maxig - 4 * p.nlm)) {
            outofmem = true;
          } else {
            const int icell = (int)(ind[i] >> 2);
            const int idir = grid_smooth_test(p, f, icell);
            if (idir > 0) {
              NewPoint np[4];
              divide_cell(p, icell, idir, np);
              sp.interpolate_points(np);
            }
          }
          i++;
        }
        sp.finish_points();   // the next batch's criteria read the new points' source
      }
      if (outofmem0) outofmem = true;
    }
    // the largest criterion after the divisions
    splitcrit = split_criterion(p, f);
  } catch (const SplitError& e) {
    err = e.what();
    return AT3D_ERR_NUMERIC;
  }
  sp.merge_optics();
  return AT3D_OK;
}

float split_criterion(const Plan& p, const SplitFields& f) {
  const bool tree = !p.treeptr.empty();
  const int nchunk = 64;
  std::vector<float> part(nchunk, 0.0f);
  // the maximum does not depend on the order the cells are visited in
  parallel_for(nchunk, [&](int c) {
    const int c0 = (int)((int64_t)p.ncells * c / nchunk), c1 = (int)((int64_t)p.ncells * (c + 1) / nchunk);
    float m = 0.0f;
    for (int ic = c0; ic < c1; ic++) {
      if (tree && p.treeptr[(size_t)ic * 2 + 1] >= 0) continue;
      float a;
      int idir;
      cell_split_test(p, f, ic, a, idir);
      m = std::max(m, a);
    }
    part[c] = m;
  });
  float crit = 0.0f;
  for (float m : part) crit = std::max(crit, m);
  return crit;
}

}  // namespace at3d

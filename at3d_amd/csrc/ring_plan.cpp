// Host-side layout of the level-ordered sweep.
//
// The reference sweeps one ordinate at a time in SWEEPING_ORDER
// (shdomsub1.f:3261-3351) with the boundary radiances set beforehand
// (PATH_INTEGRATION :2042-2086).  The GPU sweep walks dependency levels (the
// plan's levels, shifted up by one here) and keeps the last levels' values in
// LDS; boundary points become entries of the level before their first
// reader, so they travel through the same ring.  Sectors of eight (or four)
// consecutive points (32 or 16 bytes of a point-ordered radiance row) whose
// points span at most as many levels are written whole from the ring after
// their last point.
#include "ring_plan.h"

#include <algorithm>
#include <climits>
#include <utility>

namespace at3d {

// One octant's levels and entry order (see build_ring_plan).
static bool build_octant_levels(const Plan& p, int o, bool used, const std::vector<int32_t>& cls,
                                RingOctant& ro, std::string& err) {
  const int npts = p.npts;
  ro = RingOctant();
  if (!used) return true;
  const OctantPlan& op = p.oct[o];
  const bool up = (o >> 2) & 1;
  const std::vector<int32_t>& bc = up ? p.bcbot : p.bctop;
  const int bitx = o & 1, bity = (o >> 1) & 1, bitz = (o >> 2) & 1;
  std::vector<int32_t> bindex(npts, -1);
  for (size_t b = 0; b < bc.size(); b++) bindex[bc[b]] = (int32_t)b;
  ro.level.assign(npts, 0);
  for (int i = 0; i < npts; i++)
    if (op.level[i] > 0) ro.level[i] = op.level[i] + 1;
  // boundary points: one level before the first point whose start cell has
  // them on a far face (the faces a one-cell step can end on)
  std::vector<int32_t> first(npts, INT_MAX);
  const int nsw = op.lvl_off[op.nlevels];
  for (int j = 0; j < nsw; j++) {
    const int c = op.lvl_pts[j], pt = c >= 0 ? c : ~c, cell = op.lvl_cell[j];
    const int8_t fl = p.cellflags[cell];
    int faces[3], nf = 0;
    if (!(fl & 1)) faces[nf++] = 1 - bitx;
    if (!(fl & 2)) faces[nf++] = 3 - bity;
    faces[nf++] = 5 - bitz;
    for (int f = 0; f < nf; f++)
      for (int k = 0; k < 4; k++) {
        const int q = p.gridptr[(size_t)cell * 8 + face_corner(faces[f], k)];
        if (bindex[q] >= 0) first[q] = std::min(first[q], ro.level[pt]);
      }
  }
  int nl = 1;
  for (int i = 0; i < npts; i++) nl = std::max(nl, ro.level[i]);
  for (size_t b = 0; b < bc.size(); b++) {
    const int q = bc[b];
    ro.level[q] = first[q] == INT_MAX ? 1 : std::max(1, first[q] - 1);
  }
  ro.nlev = nl;
  // entries by level: the plan's order (level, rank) for points, then the
  // boundary points in boundary order (regrouped by source class below)
  std::vector<int32_t> cnt(nl + 2, 0);
  for (int i = 0; i < npts; i++) cnt[ro.level[i]]++;
  if (cnt[0] != 0) {
    err = "ring layout: a grid point is neither swept nor a boundary point";
    return false;
  }
  ro.off.assign(nl + 1, 0);
  for (int L = 1; L <= nl; L++) ro.off[L] = ro.off[L - 1] + cnt[L];
  ro.code.assign(npts, 0);
  ro.cell.assign(npts, -1);
  ro.pos.assign(npts, 0);
  std::vector<int32_t> fill(ro.off.begin(), ro.off.end());
  for (int j = 0; j < nsw; j++) {
    const int c = op.lvl_pts[j], pt = c >= 0 ? c : ~c;
    const int q = fill[ro.level[pt] - 1]++;
    ro.code[q] = pt;
    ro.cell[q] = op.lvl_cell[j];
    ro.pos[pt] = q;
  }
  for (size_t b = 0; b < bc.size(); b++) {
    const int pt = bc[b];
    const int q = fill[ro.level[pt] - 1]++;
    ro.code[q] = -1 - (int32_t)b;
    ro.pos[pt] = q;
  }
  for (int L = 1; L <= nl; L++)
    if (fill[L - 1] != ro.off[L]) {
      err = "ring layout: level lists do not cover the grid";
      return false;
    }
  // a level's entries are independent of each other, so their order is
  // free: entries of like source class are grouped (ascending on odd levels,
  // descending on even ones, so neighbouring levels meet in one class), and
  // a run of positions of the level-ordered SH->DO transform mostly needs
  // only the few source terms of its class
  std::vector<std::pair<int32_t, int32_t>> ent;
  std::vector<int32_t> code2, cell2;
  for (int L = 1; L <= nl; L++) {
    const int a = ro.off[L - 1], b = ro.off[L];
    ent.clear();
    for (int q = a; q < b; q++) {
      const int c = ro.code[q];
      const int pt = c >= 0 ? c : bc[-1 - c];
      ent.emplace_back((L & 1) ? cls[pt] : -cls[pt], q);
    }
    std::stable_sort(ent.begin(), ent.end(),
                     [](const std::pair<int32_t, int32_t>& x, const std::pair<int32_t, int32_t>& y) {
                       return x.first < y.first;
                     });
    code2.resize(b - a);
    cell2.resize(b - a);
    for (int k = 0; k < b - a; k++) {
      code2[k] = ro.code[ent[k].second];
      cell2[k] = ro.cell[ent[k].second];
    }
    for (int k = 0; k < b - a; k++) {
      const int q = a + k, c = code2[k];
      ro.code[q] = c;
      ro.cell[q] = cell2[k];
      ro.pos[c >= 0 ? c : bc[-1 - c]] = q;
    }
  }
  return true;
}

bool build_ring_plan(const Plan& p, const bool used[8], size_t lds_budget, RingPlan& r,
                     std::string& err) {
  const int npts = p.npts;
  r.nlev_max = 0;
  std::vector<int32_t> cls;
  plan_source_class(p, cls);
  // octants are independent: one host thread each
  std::string errs[8];
  bool oks[8];
  parallel_for(8, [&](int o) { oks[o] = build_octant_levels(p, o, used[o], cls, r.oct[o], errs[o]); });
  for (int o = 0; o < 8; o++)
    if (!oks[o]) {
      err = errs[o];
      return false;
    }
  for (int o = 0; o < 8; o++)
    if (used[o]) r.nlev_max = std::max(r.nlev_max, r.oct[o].nlev);
  // a level reads RING_WINDOW levels back and writes a sector whose points
  // span sector_points levels after its last one: the ring holds window
  // consecutive levels
  auto ring_slots = [&](int spts) {
    const int window = std::max(RING_WINDOW, spts) + 1;
    int m = 1;
    for (int o = 0; o < 8; o++) {
      if (!used[o]) continue;
      const RingOctant& ro = r.oct[o];
      for (int L = 1; L <= ro.nlev; L++) m = std::max(m, ro.off[L] - ro.off[std::max(L - window, 0)]);
    }
    return m;
  };
  r.sector_points = 8;
  r.slots = ring_slots(8);
  if (ring_plan_lds_bytes(r.slots, r.nlev_max, p.nz * 4) > lds_budget) {
    r.sector_points = 4;
    r.slots = ring_slots(4);
  }
  const int SP = r.sector_points;
  // sectors (need the ring size), octants on their own host threads
  parallel_for(8, [&](int o) {
    if (!used[o]) return;
    RingOctant& ro = r.oct[o];
    const int nl = ro.nlev;
    ro.staged.assign(npts, 0);
    std::vector<std::vector<int32_t>> by_level(nl + 1);
    for (int s0 = 0; s0 + SP <= npts; s0 += SP) {
      int lo = INT_MAX, hi = 0;
      for (int k = 0; k < SP; k++) {
        lo = std::min(lo, ro.level[s0 + k]);
        hi = std::max(hi, ro.level[s0 + k]);
      }
      if (lo < hi - (SP - 1)) continue;
      by_level[hi].push_back(s0);
      for (int k = 0; k < SP; k++) ro.staged[s0 + k] = 1;
    }
    ro.sec_off.assign(nl + 2, 0);
    for (int L = 1; L <= nl; L++) ro.sec_off[L + 1] = ro.sec_off[L] + (int32_t)by_level[L].size();
    ro.sec_p0.clear();
    ro.sec_slot.clear();
    for (int L = 1; L <= nl; L++)
      for (int s0 : by_level[L]) {
        ro.sec_p0.push_back(s0);
        for (int k = 0; k < 8; k++)
          ro.sec_slot.push_back(k < SP ? (uint16_t)(ro.pos[s0 + k] % r.slots) : (uint16_t)0);
      }
  });
  return true;
}

}  // namespace at3d

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

// One octant's ring levels and entry order before grouping by source class:
// the plan's levels shifted up by one, each boundary point inserted one level
// before the first point whose start cell has it on a far face (the faces a
// one-cell step can end on), points in the plan's (level, rank) order and
// then the boundary points in boundary order.
bool ring_octant_base(const Plan& p, int o, OctantPlan& op, std::string& err) {
  const int npts = p.npts;
  const bool up = (o >> 2) & 1;
  const std::vector<int32_t>& bc = up ? p.bcbot : p.bctop;
  const int bitx = o & 1, bity = (o >> 1) & 1, bitz = (o >> 2) & 1;
  std::vector<int32_t> bindex(npts, -1);
  for (size_t b = 0; b < bc.size(); b++) bindex[bc[b]] = (int32_t)b;
  std::vector<int32_t>& level = op.ring_level;
  level.assign(npts, 0);
  for (int i = 0; i < npts; i++)
    if (op.level[i] > 0) level[i] = op.level[i] + 1;
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
        if (bindex[q] >= 0) first[q] = std::min(first[q], level[pt]);
      }
  }
  int nl = 1;
  for (int i = 0; i < npts; i++) nl = std::max(nl, level[i]);
  for (size_t b = 0; b < bc.size(); b++) {
    const int q = bc[b];
    level[q] = first[q] == INT_MAX ? 1 : std::max(1, first[q] - 1);
  }
  op.ring_nlev = nl;
  std::vector<int32_t> cnt(nl + 2, 0);
  for (int i = 0; i < npts; i++) cnt[level[i]]++;
  if (cnt[0] != 0) {
    err = "ring layout: a grid point is neither swept nor a boundary point";
    return false;
  }
  std::vector<int32_t>& off = op.ring_off;
  off.assign(nl + 1, 0);
  for (int L = 1; L <= nl; L++) off[L] = off[L - 1] + cnt[L];
  op.ring_code.assign(npts, 0);
  op.ring_cell.assign(npts, -1);
  op.ring_pos.assign(npts, 0);
  std::vector<int32_t> fill(off.begin(), off.end());
  for (int j = 0; j < nsw; j++) {
    const int c = op.lvl_pts[j], pt = c >= 0 ? c : ~c;
    const int q = fill[level[pt] - 1]++;
    op.ring_code[q] = pt;
    op.ring_cell[q] = op.lvl_cell[j];
    op.ring_pos[pt] = q;
  }
  for (size_t b = 0; b < bc.size(); b++) {
    const int pt = bc[b];
    const int q = fill[level[pt] - 1]++;
    op.ring_code[q] = -1 - (int32_t)b;
    op.ring_pos[pt] = q;
  }
  for (int L = 1; L <= nl; L++)
    if (fill[L - 1] != off[L]) {
      err = "ring layout: level lists do not cover the grid";
      return false;
    }
  return true;
}

// One octant's levels and entry order (see build_ring_plan): the plan's
// ring_* order with each level's entries grouped by source class.
static bool build_octant_levels(const Plan& p, int o, bool used, const std::vector<int32_t>& cls, int ncls,
                                RingOctant& ro, std::string& err) {
  ro = RingOctant();
  if (!used) return true;
  const OctantPlan& op = p.oct[o];
  if (op.ring_off.empty()) {
    err = "ring layout: the plan's levels carry no ring order for a used octant";
    return false;
  }
  const std::vector<int32_t>& bc = ((o >> 2) & 1) ? p.bcbot : p.bctop;
  const int nl = op.ring_nlev;
  ro.nlev = nl;
  ro.level = op.ring_level;
  ro.off = op.ring_off;
  ro.code = op.ring_code;
  ro.cell = op.ring_cell;
  ro.pos = op.ring_pos;
  // a level's entries are independent of each other, so their order is
  // free: entries of like source class are grouped (ascending on odd levels,
  // descending on even ones, so neighbouring levels meet in one class; a
  // stable counting sort), and a run of positions of the level-ordered
  // SH->DO transform mostly needs only the few source terms of its class
  std::vector<int32_t> cnt(ncls + 1), code2, cell2;
  for (int L = 1; L <= nl; L++) {
    const int a = ro.off[L - 1], b = ro.off[L];
    auto key = [&](int q) {
      const int c = op.ring_code[q];
      const int k = cls[c >= 0 ? c : bc[-1 - c]];
      return (L & 1) ? k : ncls - 1 - k;
    };
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int q = a; q < b; q++) cnt[key(q) + 1]++;
    for (int k = 0; k < ncls; k++) cnt[k + 1] += cnt[k];
    code2.resize(b - a);
    cell2.resize(b - a);
    for (int q = a; q < b; q++) {
      const int d = cnt[key(q)]++;
      code2[d] = op.ring_code[q];
      cell2[d] = op.ring_cell[q];
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

bool build_ring_plan(const Plan& p, const bool used[8], size_t lds_budget, RingPlan& r, std::string& err,
                     const std::vector<int32_t>* cls_in, int slot_bytes) {
  const int npts = p.npts;
  r.nlev_max = 0;
  std::vector<int32_t> cls;
  if (cls_in)
    cls = *cls_in;
  else
    plan_source_class(p, cls);
  int ncls = 1;
  for (int c : cls) ncls = std::max(ncls, c + 1);
  // octants are independent: one host thread each
  std::string errs[8];
  bool oks[8];
  parallel_for(8, [&](int o) { oks[o] = build_octant_levels(p, o, used[o], cls, ncls, r.oct[o], errs[o]); });
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
  if (ring_plan_lds_bytes(r.slots, r.nlev_max, p.nz * 4, slot_bytes) > lds_budget) {
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
    // the level each sector completes on (0: not staged), then the sectors
    // by completion level (a counting sort, in point order within a level)
    const int nsec = npts / SP;
    std::vector<int32_t> done(nsec, 0);
    ro.sec_off.assign(nl + 2, 0);
    for (int sc = 0; sc < nsec; sc++) {
      const int s0 = sc * SP;
      int lo = INT_MAX, hi = 0;
      for (int k = 0; k < SP; k++) {
        lo = std::min(lo, ro.level[s0 + k]);
        hi = std::max(hi, ro.level[s0 + k]);
      }
      if (lo < hi - (SP - 1)) continue;
      done[sc] = hi;
      ro.sec_off[hi + 1]++;
      for (int k = 0; k < SP; k++) ro.staged[s0 + k] = 1;
    }
    for (int L = 1; L <= nl; L++) ro.sec_off[L + 1] += ro.sec_off[L];
    const int nst = ro.sec_off[nl + 1];
    ro.sec_p0.assign(nst, 0);
    ro.sec_slot.assign((size_t)nst * 8, 0);
    std::vector<int32_t> fill(ro.sec_off.begin(), ro.sec_off.end());
    for (int sc = 0; sc < nsec; sc++) {
      if (!done[sc]) continue;
      const int d = fill[done[sc]]++, s0 = sc * SP;
      ro.sec_p0[d] = s0;
      for (int k = 0; k < SP; k++) ro.sec_slot[(size_t)d * 8 + k] = (uint16_t)(ro.pos[s0 + k] % r.slots);
    }
  });
  return true;
}

}  // namespace at3d

// Host-side layout of the level-ordered sweep (CPU only).  See ring_plan.cpp.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "host_plan.h"

namespace at3d {

// Per octant: the sweep levels with the hemisphere's boundary points inserted
// as entries of their own (each one level before the first one-cell step that
// reads it), so a sweep workgroup can keep every value it gathers in a ring of
// recent levels; and the 32-byte sectors of the point-ordered radiance row
// whose eight points all lie within eight consecutive levels, written whole
// once their last point is done.
struct RingOctant {
  int nlev = 0;
  std::vector<int32_t> off;       // [nlev+1] level L spans positions off[L-1] .. off[L]
  std::vector<int32_t> code;      // [npts] per position: point >= 0, or -1-b for boundary point b
  std::vector<int32_t> cell;      // [npts] start cell of a point entry (-1 for boundary entries)
  std::vector<int32_t> level;     // [npts] ring level of each point
  std::vector<int32_t> pos;       // [npts] position of each point
  std::vector<int8_t> staged;     // [npts] 1: written by its sector
  std::vector<int32_t> sec_off;   // [nlev+2] sectors completing at level L: sec_off[L] .. sec_off[L+1]
  std::vector<int32_t> sec_p0;    // first point of each sector
  std::vector<uint16_t> sec_slot; // [nsec][8] radiance ring slot of each of its points
};

constexpr int RING_WINDOW = 3;    // earlier levels a one-cell step may read through the ring
constexpr int SECTOR_POINTS = 8;  // 32-byte sectors of the FP32 radiance rows

struct RingPlan {
  RingOctant oct[8];
  int c4 = 0;   // source ring slots: >= the entries of RING_WINDOW + 1 consecutive levels
  int c9 = 0;   // radiance ring slots: a multiple of c4, >= the entries of SECTOR_POINTS + 1 levels
  int nlev_max = 0;
};

// Builds the ring layout for the octants with ordinates (used[o] != 0).
// Returns false (err set) if the plan's level lists are incomplete.
bool build_ring_plan(const Plan& p, const bool used[8], RingPlan& r, std::string& err);

}  // namespace at3d

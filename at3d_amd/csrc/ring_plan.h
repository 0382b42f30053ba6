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
// recent levels; and the sectors of the point-ordered radiance row (eight
// points, 32 bytes, or four, 16 bytes, whichever ring fits in LDS) whose
// points all lie within as many consecutive levels, written whole once their
// last point is done.
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
  std::vector<uint16_t> sec_slot; // [nsec][8] ring slot of each of its points (4 used with 4-point sectors)
};

constexpr int RING_WINDOW = 3;    // earlier levels a one-cell step may read through the ring
// LDS a gfx950 workgroup may allocate (hipDeviceAttributeMaxSharedMemoryPerBlock
// there): the budget of host-only layout queries made without a device; the
// solver state takes the device's own attribute
constexpr size_t GFX950_LDS_BYTES = 160 * 1024;

struct RingPlan {
  RingOctant oct[8];
  int sector_points = 8;  // 8 (32-byte sectors) or 4 (16-byte sectors)
  int slots = 0;  // ring slots: >= the entries of max(RING_WINDOW, sector_points) + 1 consecutive levels
  int nlev_max = 0;
};

// Bytes of LDS the level-ordered sweep needs for a ring of `slots` entries
// of `slot_bytes` each, `nlev` levels and `ncls` step classes.
inline size_t ring_plan_lds_bytes(int slots, int nlev, int ncls, int slot_bytes = 8) {
  return (size_t)ncls * 16 + (size_t)slots * slot_bytes + (size_t)(nlev + 2) * 9 + 4;
}

// Builds the ring layout for the octants with ordinates (used[o] != 0), with
// 32-byte sectors if that ring fits in lds_budget bytes, else 16-byte ones.
// Returns false (err set) if the plan's level lists are incomplete.
// cls: each point's source class (the plan's, from its source constants,
// when null).
// slot_bytes: 8 for a ring of (radiance, source) pairs, 4 for radiances alone.
bool build_ring_plan(const Plan& p, const bool used[8], size_t lds_budget, RingPlan& r, std::string& err,
                     const std::vector<int32_t>* cls = nullptr, int slot_bytes = 8);

}  // namespace at3d

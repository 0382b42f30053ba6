// Adaptive cell splitting on the host (SPLIT_GRID, shdomsub1.f:4703-4931).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "host_plan.h"

namespace at3d {

// Host copies of the per-point solution fields the splitting reads and
// extends with the new points (dense rows of `ld` terms; the counts are the
// reference's SHPTR / RSHPTR / OSHPTR extents).
struct SplitFields {
  int ld = 0;
  std::vector<float> S, R, DS;           // [npts][ld] source, radiance, delta-source (I)
  std::vector<int32_t> ns, nr, nos;      // [npts]
  // polarized solves (NSTOKES 3): the Q and U components, each [npts][ld]
  // with the counts above
  int nst = 1;
  std::vector<float> Squ[2], Rqu[2], DSqu[2];
};

// The splitting schedule of SOLUTION_ITERATIONS (shdomsub1.f:636-720) across
// the iterations of one solve call.
struct SplitSchedule {
  bool testing = true, outofmem = false;
  float cursplitacc = 0.f, startsplitacc = 0.f, avgsolcrit = 0.f, splitcrit = 0.f;
};

// One SPLIT_GRID call on plan `p` (its grid, tree and per-point optics are
// extended in place; the caller refreshes the derived plan data with
// plan_refresh_points when points were added): with `dosplit`, cells whose
// criterion exceeds `cursplitacc` are divided in batches, largest criterion
// first, then cells that make the grid smoother, within the capacities
// MAXIG / MAXIC / MAXIV / MAXIDO; new points get their optics from the
// property grid, the direct beam, the mean of their parents' radiances and
// the source computed from those.  `splitcrit` returns the largest criterion
// of the (new) end cells.
int split_grid(Plan& p, SplitFields& f, bool dosplit, bool& outofmem, float cursplitacc, float& splitcrit,
               std::string& err);

// SPLITCRIT alone (the largest criterion of the end cells): the splitting
// test of an iteration that divides nothing.  Reads f.S and f.ns only.
float split_criterion(const Plan& p, const SplitFields& f);

}  // namespace at3d

// plan.h -- per-structure linked-cell plans, computed on the host with the
// reference's FP64 arithmetic (src/ext/lclist.f90, src/geometry.f90).
#pragma once

#include <string>
#include <vector>

#include "aenet_amd_internal.h"

namespace aenet_amd {

struct PlanHost {
  std::vector<StructDev> st;
  std::vector<int> atom_struct;   // [natoms]
  std::vector<int> cell_struct;   // [ncells]
  std::vector<int> cell_ijk;      // [3*ncells]
  std::vector<double> cell_ctr;   // [3*ncells] Cartesian centres
  std::vector<int> Tvec, Cvec;    // 3 ints per vector
  std::vector<int> self_t;        // Tvec rank (1-based) of each self image
  std::vector<double> self_d;
  int natoms = 0, ncells = 0;
};

// latt: 9 doubles per structure (vector k = latt[9*s+3k .. 3k+2]);
// coords: Cartesian, 3 per atom (needed for the bounding box of isolated
// structures).  Returns 0 or an AENET_STATUS_* code with a message.
int build_plan(double rc_max, int nstruct, const int *natoms, const double *latt,
               const int *pbc, const double *coords, PlanHost &plan, std::string &err);

// Plan for one structure whose input coordinates are already lattice
// coordinates of the given lattice (aenet_nbl_init, src/aenet.f90:571-607):
// no bounding box, identity reciprocal lattice.
int build_plan_lattice(double rc_max, int natoms, const double *latt, int pbc, PlanHost &plan,
                       std::string &err);

}  // namespace aenet_amd

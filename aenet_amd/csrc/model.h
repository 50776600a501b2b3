// model.h -- host-side potential set: .nn loading and packing of the
// per-species parameters the gfx950 kernels read.
#pragma once

#include <string>
#include <vector>

#include "aenet_amd_internal.h"

namespace aenet_amd {

struct NetHost {
  int nlayers = 0, nnodes_max = 0, Wsize = 0, nvalues = 0;
  std::vector<int> nnodes, fa, iw, iv;
  std::vector<double> W;
};

struct SetupHost {
  std::string description, atomtype, sftype;
  std::vector<std::string> envtypes;
  double Rc_min = 0, Rc_max = 0;
  int nsf = 0, nsfparam = 0, neval = 0;
  std::vector<int> sf, sfenv;          // sfenv: 2 per function (1-based local types)
  std::vector<double> sfparam;         // nsfparam per function
  std::vector<double> vmin, vmax, avg, cov;
  std::vector<int> ltype;              // [ntypes+1]: global -> local (1-based), 0 if absent
};

struct PotHost {
  bool init = false;
  std::string file;
  NetHost net;
  SetupHost stp;
  double E_scale = 1, E_shift = 0, E_atom = 0;
  std::string ts_file;
  double ts_scale = 1, ts_shift = 0;
  std::vector<std::string> ts_types;
  std::vector<double> ts_E_atom;
};

// Chebyshev term counts (order + 1) the atom kernel has a register layout
// for: radial up to 24 with angular up to 8, or both up to 16 (the
// per-lane accumulators of a descriptor hold 2 (radial + angular) <= 64
// terms, typespin included)
inline bool cheb_orders_ok(int radial_terms, int angular_terms) {
  return (radial_terms <= 24 && angular_terms <= 8) || (radial_terms <= 16 && angular_terms <= 16);
}

// Reads one .nn file (Fortran unformatted sequential; 36 records).
// Returns 0 on success, AENET_ERR_IO / a format error code otherwise.
int read_nn(const std::string &path, PotHost &pot, std::string &err);

struct PotSet {
  int ntypes = 0;
  std::vector<std::string> names;
  std::vector<PotHost> pots;
  bool loaded = false;
  double Rc_min = 0, Rc_max = 0;
  int nnb_max = 0, nsf_max = 0;
  int desc = DESC_CHEBYSHEV;
  // packed for the device
  std::vector<SpeciesDev> species;
  std::vector<float> W;
  std::vector<double> norm_avg, norm_scale;
  BPDev bp{};
  std::vector<BPFunc> bpf;
  std::vector<BPAngK> bpk;  // parallel to bpf
  std::vector<BPAngGrid> bpg;  // grid blocks
  int max_rN = 0, max_aN = 0;

  // called when the last potential is loaded (aenet_load_potential,
  // src/aenet.f90:282-289): derives cutoffs, capacities and the packing.
  int finalize(std::string &err);
};

int nmax_nbdist(double rmin, double rmax);
bool name_eq(const std::string &a, const std::string &b);

}  // namespace aenet_amd

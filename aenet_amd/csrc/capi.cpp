// capi.cpp -- the libaenet C ABI (include/aenet.h) and the batched
// extensions (include/aenet_amd.h), backed by the gfx950 kernels.
//
// Like the reference library (src/aenet.f90:122-132) the potential set is
// process-global; a mutex serializes calls that use it.
#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "aenet_capi_status.h"
#include "kernels.h"
#include "mc.h"
#include "engine.h"
#include "model.h"

using namespace aenet_amd;

extern "C" {
int AENET_OK = AENET_STATUS_OK;
int AENET_ERR_INIT = AENET_STATUS_ERR_INIT;
int AENET_ERR_MALLOC = AENET_STATUS_ERR_MALLOC;
int AENET_ERR_IO = AENET_STATUS_ERR_IO;
int AENET_ERR_TYPE = AENET_STATUS_ERR_TYPE;
int AENET_TYPELEN = 2;
int AENET_PATHLEN = 1024;
bool AENET_TRUE = true;
bool AENET_FALSE = false;
int aenet_nsf_max = 0;
int aenet_nnb_max = 0;
double aenet_Rc_min = 0.0;
double aenet_Rc_max = 0.0;
}

namespace {

std::mutex g_mu;
thread_local std::string g_err;

struct Global {
  bool init = false;
  PotSet set;
  std::unique_ptr<DeviceModel> dev;
  hipStream_t stream = nullptr;
  EnvScratch env;
  // standalone fingerprint basis (aenet_sfb_*)
  bool sfb_init = false;
  PotSet sfb_set;
  std::unique_ptr<DeviceModel> sfb_dev;
  int sfb_N = 0, sfb_rN = 0;
  double sfb_rRc = 0.0;
  // neighbor list state (aenet_nbl_*)
  bool nbl_init = false;
  int nbl_natoms = 0, nbl_cap = 0;
  std::vector<int> nbl_count, nbl_list, nbl_type;
  std::vector<double> nbl_coo, nbl_dist;
};

Global &G() {
  static Global g;
  return g;
}

int set_err(int st, const std::string &msg) {
  g_err = msg;
  return st;
}

hipStream_t lib_stream() {
  Global &g = G();
  if (!g.stream) {
    if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess) g.stream = nullptr;
  }
  return g.stream;
}

int ensure_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return set_err(AENET_STATUS_ERR_DEVICE, "no HIP device available: libaenet_amd runs on AMD GPUs only");
  return AENET_STATUS_OK;
}

std::string cname(const char *s) {
  // TYPELEN+1 characters up to the terminating NUL (src/aenet.f90:184-189)
  std::string out;
  for (int i = 0; i < 3 && s && s[i]; ++i) out.push_back(s[i]);
  if (out.size() > 2) out.resize(2);
  return out;
}

void clear_set(Global &g) {
  g.set = PotSet();
  g.dev.reset();
  g.init = false;
  aenet_nsf_max = 0;
  aenet_nnb_max = 0;
  aenet_Rc_min = 0.0;
  aenet_Rc_max = 0.0;
}

// one environment through the device kernel; adds forces into F (FP64)
int eval_env(Global &g, const double *coo_i, int type_i, int n_j, const double *coo_j, const int *type_j,
             const int *index_j, int index_i, double *F, double *E_i) {
  if (type_i < 1 || type_i > g.set.ntypes) return set_err(AENET_STATUS_ERR_TYPE, "central type out of range");
  for (int j = 0; j < n_j; ++j)
    if (type_j[j] < 1 || type_j[j] > g.set.ntypes)
      return set_err(AENET_STATUS_ERR_TYPE, "neighbor type out of range");
  int start[2] = {0, n_j};
  bool forces = F != nullptr;
  std::vector<double> dFj(forces ? 3 * (size_t)n_j : 0);
  double dFi[3] = {0, 0, 0};
  std::string err;
  int st = g.env.eval(*g.dev, 1, start, &type_i, coo_i, coo_j, type_j, forces, E_i, dFi, dFj.data(), nullptr,
                      0, lib_stream(), err);
  if (st) return set_err(st, err);
  if (forces) {
    for (int x = 0; x < 3; ++x) F[3 * (size_t)(index_i - 1) + x] += dFi[x];
    for (int j = 0; j < n_j; ++j)
      for (int x = 0; x < 3; ++x) F[3 * (size_t)(index_j[j] - 1) + x] += dFj[3 * (size_t)j + x];
  }
  return AENET_STATUS_OK;
}

}  // namespace

struct aenet_amd_batch {
  Batch batch;
};

struct aenet_amd_model {
  PotSet set;
  DeviceModel dev;
  DBuf<int> flags;  // error word of the training kernels
  EnvScratch env;   // explicit environments (descriptor derivatives)
};

struct aenet_amd_mc {
  MonteCarlo mc;
};

namespace {
// the model an entry point works with: the given handle, else the set
// loaded through aenet_load_potential
const DeviceModel *pick_model(const aenet_amd_model *m) {
  if (m) return &m->dev;
  Global &g = G();
  return g.dev ? g.dev.get() : nullptr;
}
}  // namespace

extern "C" {

const char *aenet_amd_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------
// aenet_init / aenet_final / aenet_load_potential (src/aenet.f90:140-331)
// ---------------------------------------------------------------------

void aenet_init(int ntypes, char *atom_types[], int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (g.init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_init: already initialized");
    return;
  }
  if (ntypes < 1 || ntypes > AENET_AMD_MAX_TYPES) {
    *stat = set_err(AENET_STATUS_ERR_MALLOC, "aenet_init: 1 to 8 species are supported");
    return;
  }
  g.set = PotSet();
  g.set.ntypes = ntypes;
  for (int i = 0; i < ntypes; ++i) g.set.names.push_back(cname(atom_types[i]));
  g.set.pots.assign(ntypes, PotHost());
  g.init = true;
}

void aenet_final(int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (g.init) clear_set(g);
}

void aenet_print_info(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!g.init) {
    std::printf(" Nothing to print. AenetLib is not initialized.\n");
    return;
  }
  for (int t = 0; t < g.set.ntypes; ++t) {
    const PotHost &p = g.set.pots[t];
    if (!p.init) continue;
    std::printf(" Atomic species : %s\n File name      : %s\n\n", p.stp.atomtype.c_str(), p.file.c_str());
    std::printf(" Network: %d layers (", p.net.nlayers);
    for (int l = 0; l < p.net.nlayers; ++l) std::printf(l ? "-%d" : "%d", p.net.nnodes[l]);
    std::printf("), %d weights\n", p.net.Wsize);
    std::printf(" Descriptor: %s, %d functions, Rc_min %.3f, Rc_max %.3f\n", p.stp.sftype.c_str(), p.stp.nsf,
                p.stp.Rc_min, p.stp.Rc_max);
    std::printf(" E_scale %.10g  E_shift %.10g  E_atom %.10g\n\n", p.E_scale, p.E_shift, p.E_atom);
  }
  std::fflush(stdout);
}

void aenet_load_potential(int type_id, char *filename, int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (!filename || std::strlen(filename) > 1024) {
    *stat = set_err(AENET_STATUS_ERR_MALLOC, "aenet_load_potential: path longer than 1024 characters");
    return;
  }
  if (!g.init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_load_potential: call aenet_init first");
    return;
  }
  if (type_id < 1 || type_id > g.set.ntypes) {
    *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_load_potential: type id out of range");
    return;
  }
  struct stat sb;
  if (::stat(filename, &sb) != 0) {
    *stat = set_err(AENET_STATUS_ERR_IO, std::string("aenet_load_potential: file not found: ") + filename);
    return;
  }
  std::string err;
  PotHost pot;
  int st = read_nn(filename, pot, err);
  if (st) {
    *stat = set_err(st, err);
    return;
  }
  g.set.pots[type_id - 1] = pot;
  g.set.loaded = false;
  bool all = true;
  for (auto &p : g.set.pots) all = all && p.init;
  if (all) {
    st = g.set.finalize(err);
    if (st) {
      *stat = set_err(st, err);
      return;
    }
    if ((st = ensure_device())) {
      *stat = st;
      return;
    }
    g.dev.reset(new DeviceModel());
    if (g.dev->upload(g.set, lib_stream())) {
      *stat = set_err(AENET_STATUS_ERR_DEVICE, "aenet_load_potential: device upload failed");
      return;
    }
    aenet_Rc_min = g.set.Rc_min;
    aenet_Rc_max = g.set.Rc_max;
    aenet_nnb_max = g.set.nnb_max;
    aenet_nsf_max = g.set.nsf_max;
  }
}

bool aenet_all_loaded(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!g.init) return true;  // the reference loops over zero types
  for (auto &p : g.set.pots)
    if (!p.init) return false;
  return true;
}

double aenet_free_atom_energy(int type_id) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!g.init || type_id < 1 || type_id > g.set.ntypes || !g.set.pots[type_id - 1].init) return NAN;
  return g.set.pots[type_id - 1].E_atom;
}

// ---------------------------------------------------------------------
// per-atom evaluation (src/aenet.f90:359-486)
// ---------------------------------------------------------------------

void aenet_atomic_energy(double coo_i[3], int type_i, int n_j, double coo_j[], int type_j[], double *E_i,
                         int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!g.init || !g.set.loaded) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_atomic_energy: potentials not loaded");
    return;
  }
  *stat = eval_env(g, coo_i, type_i, n_j, coo_j, type_j, nullptr, 0, nullptr, E_i);
}

void aenet_atomic_energy_and_forces(double coo_i[3], int type_i, int index_i, int n_j, double coo_j[],
                                    int type_j[], int index_j[], int natoms, double *E_i, double F[],
                                    int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!g.init || !g.set.loaded) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_atomic_energy_and_forces: potentials not loaded");
    return;
  }
  if (index_i < 1 || index_i > natoms) {
    *stat = set_err(AENET_STATUS_ERR_TYPE, "atom index out of range");
    return;
  }
  for (int j = 0; j < n_j; ++j)
    if (index_j[j] < 1 || index_j[j] > natoms) {
      *stat = set_err(AENET_STATUS_ERR_TYPE, "neighbor index out of range");
      return;
    }
  *stat = eval_env(g, coo_i, type_i, n_j, coo_j, type_j, index_j, index_i, F, E_i);
}

// geo_type_conv (src/geometry.f90:440-461) over the loaded species
void aenet_convert_atom_types(int ntypes_in, char *atom_types[], int natoms_in, int type_id_in[],
                              int type_id_out[], int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (!g.init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_convert_atom_types: not initialized");
    return;
  }
  for (int i = 0; i < natoms_in; ++i) {
    int it = type_id_in[i];
    int out = 0;
    if (it >= 1 && it <= ntypes_in) {
      std::string nm = cname(atom_types[it - 1]);
      for (int t = 0; t < g.set.ntypes; ++t)
        if (name_eq(g.set.names[t], nm)) {
          out = t + 1;
          break;
        }
    }
    if (out == 0) {
      *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_convert_atom_types: unknown species");
      return;
    }
    type_id_out[i] = out;
  }
}

// ---------------------------------------------------------------------
// standalone Chebyshev basis (src/aenet.f90:652-799)
// ---------------------------------------------------------------------

void aenet_sfb_init(int ntypes, char *atom_types[], int radial_order, int angular_order, double radial_Rc,
                    double angular_Rc, int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (g.sfb_init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_sfb_init: already initialized");
    return;
  }
  if (ntypes < 1 || ntypes > AENET_AMD_MAX_TYPES || radial_order < 0 || radial_order > 31 ||
      angular_order < 0 || angular_order > 15 || radial_Rc <= 0.0) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_sfb_init: unsupported basis parameters");
    return;
  }
  int st = ensure_device();
  if (st) {
    *stat = st;
    return;
  }
  // a potential set with one network-free species per type and the
  // descriptor of new_SFBasis (src/ext/sfbasis.f90:83-151)
  PotSet ps;
  ps.ntypes = ntypes;
  std::vector<std::string> names;
  for (int i = 0; i < ntypes; ++i) names.push_back(cname(atom_types[i]));
  ps.names = names;
  int nsf = (radial_order + 1 + angular_order + 1) * (ntypes > 1 ? 2 : 1);
  for (int t = 0; t < ntypes; ++t) {
    PotHost p;
    p.init = true;
    p.net.nlayers = 2;
    p.net.nnodes = {nsf, 1};
    p.net.fa = {0, 0};
    p.net.iw = {0, nsf + 1};
    p.net.iv = {0, nsf + 1};
    p.net.Wsize = nsf + 1;
    p.net.W.assign(nsf + 1, 0.0);
    p.stp.atomtype = names[t];
    p.stp.envtypes = names;
    p.stp.sftype = "Chebyshev";
    p.stp.nsf = nsf;
    p.stp.nsfparam = 4;
    p.stp.sfparam.assign(4 * (size_t)nsf, 0.0);
    p.stp.sfparam[0] = radial_Rc;
    p.stp.sfparam[1] = radial_order;
    p.stp.sfparam[2] = angular_Rc;
    p.stp.sfparam[3] = angular_order;
    p.stp.Rc_min = 1.0;
    p.stp.Rc_max = std::max(radial_Rc, angular_Rc);
    p.stp.avg.assign(nsf, 0.0);
    p.stp.cov.assign(nsf, 1.0);
    p.stp.sf.assign(nsf, 0);
    p.stp.sfenv.assign(2 * (size_t)nsf, 0);
    ps.pots.push_back(p);
  }
  std::string err;
  st = ps.finalize(err);
  if (st) {
    *stat = set_err(st, err);
    return;
  }
  g.sfb_set = ps;
  g.sfb_dev.reset(new DeviceModel());
  if (g.sfb_dev->upload(g.sfb_set, lib_stream())) {
    *stat = set_err(AENET_STATUS_ERR_DEVICE, "aenet_sfb_init: device upload failed");
    return;
  }
  g.sfb_N = nsf;
  g.sfb_rN = radial_order + 1;
  g.sfb_rRc = radial_Rc;
  g.sfb_init = true;
}

void aenet_sfb_final(int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (!g.sfb_init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_sfb_final: not initialized");
    return;
  }
  g.sfb_dev.reset();
  g.sfb_set = PotSet();
  g.sfb_init = false;
}

int aenet_sfb_nvalues(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return G().sfb_init ? G().sfb_N : 0;
}

void aenet_sfb_eval(int itype0, double coo0[3], int nat, int itype1[], double coo1[], int nv, double values[],
                    int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (!g.sfb_init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_sfb_eval: not initialized");
    return;
  }
  if (nv != g.sfb_N) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_sfb_eval: wrong number of basis functions");
    return;
  }
  int nt = g.sfb_set.ntypes;
  if (itype0 < 1 || itype0 > nt) {
    *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_sfb_eval: type out of range");
    return;
  }
  for (int j = 0; j < nat; ++j)
    if (itype1[j] < 1 || itype1[j] > nt) {
      *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_sfb_eval: type out of range");
      return;
    }
  int start[2] = {0, nat};
  std::string err;
  int st = g.env.eval(*g.sfb_dev, 1, start, &itype0, coo0, coo1, itype1, false, nullptr, nullptr, nullptr,
                      values, nv, lib_stream(), err);
  if (st) *stat = set_err(st, err);
}

// sfb_eval with its derivatives (deriv0, deriv1 of src/ext/sfbasis.f90:
// 440-571, which aenet_sfb_eval does not return): the environment is
// evaluated as nv copies, copy f with the unit vector of feature f in
// place of the network's dE/dG, so the force kernel returns -dG_f/dR_j and
// -dG_f/dR_0.  The derivatives are FP32 sums (relative error ~1e-6).
void aenet_sfb_eval_deriv(int itype0, double coo0[3], int nat, int itype1[], double coo1[], int nv, double values[],
                          double deriv0[], double deriv1[], int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (!g.sfb_init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_sfb_eval_deriv: not initialized");
    return;
  }
  if (nv != g.sfb_N || nat < 0) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_sfb_eval_deriv: wrong number of basis functions");
    return;
  }
  const int nt = g.sfb_set.ntypes;
  if (itype0 < 1 || itype0 > nt) {
    *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_sfb_eval_deriv: type out of range");
    return;
  }
  for (int j = 0; j < nat; ++j)
    if (itype1[j] < 1 || itype1[j] > nt) {
      *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_sfb_eval_deriv: type out of range");
      return;
    }
  std::vector<int> start(nv + 1), ct(nv, itype0), tj((size_t)nv * nat);
  std::vector<double> ci(3 * (size_t)nv), cj(3 * (size_t)nv * nat), E(nv), dFi(3 * (size_t)nv),
      dFj(3 * (size_t)nv * nat), G((size_t)nv * nv);
  for (int f = 0; f <= nv; ++f) start[f] = f * nat;
  for (int f = 0; f < nv; ++f) {
    std::memcpy(&ci[3 * (size_t)f], coo0, 3 * sizeof(double));
    std::memcpy(&cj[3 * (size_t)f * nat], coo1, 3 * (size_t)nat * sizeof(double));
    std::memcpy(&tj[(size_t)f * nat], itype1, (size_t)nat * sizeof(int));
  }
  std::string err;
  int st = g.env.eval(*g.sfb_dev, nv, start.data(), ct.data(), ci.data(), cj.data(), tj.data(), true, E.data(),
                      dFi.data(), dFj.data(), G.data(), nv, lib_stream(), err, nv);
  if (st) {
    *stat = set_err(st, err);
    return;
  }
  std::memcpy(values, G.data(), (size_t)nv * sizeof(double));
  // deriv0(3, nv), deriv1(3, nv, nat), column-major
  for (int f = 0; f < nv; ++f)
    for (int x = 0; x < 3; ++x) {
      if (deriv0) deriv0[3 * (size_t)f + x] = -dFi[3 * (size_t)f + x];
      if (deriv1)
        for (int j = 0; j < nat; ++j)
          deriv1[x + 3 * ((size_t)f + (size_t)nv * j)] = -dFj[3 * ((size_t)f * nat + j) + x];
    }
}

// sfb_reconstruct_radial (src/ext/sfbasis.f90:581-637): the radial
// distribution function from the first r_N expansion coefficients
void aenet_sfb_reconstruct_radial(int nv, double values[], int nx, double x[], double y[], int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *stat = AENET_STATUS_OK;
  if (!g.sfb_init) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_sfb_reconstruct_radial: not initialized");
    return;
  }
  (void)nv;
  const double PI = 3.14159265358979, PI_INV = 1.0 / PI;
  const double Rc = g.sfb_rRc;
  const int rN = g.sfb_rN;
  double dx = Rc / (double)(nx - 1);
  for (int i = 0; i < nx; ++i) x[i] = (double)i * dx;
  std::vector<double> f(rN);
  for (int i = 0; i < nx - 1; ++i) {
    double xx = (2.0 * x[i] - 0.0 - Rc) / (Rc - 0.0);
    f[0] = 1.0;
    if (rN > 1) f[1] = xx;
    for (int k = 2; k < rN; ++k) f[k] = 2.0 * xx * f[k - 1] - f[k - 2];
    double r = x[i] / Rc;
    double w = 0.5 / std::sqrt(r - r * r);
    for (int k = 0; k < rN; ++k) f[k] = f[k] * w * PI_INV;
    f[0] = 0.5 * f[0];
    double acc = 0.0;
    for (int k = rN - 1; k >= 0; --k) acc = acc + values[k] * f[k];
    y[i] = acc;
  }
  if (nx > 0) y[nx - 1] = 0.0;
}

// ---------------------------------------------------------------------
// neighbor lists (src/aenet.f90:571-639)
// ---------------------------------------------------------------------

void aenet_nbl_init(double latt_vec[9], int natoms, int atom_types[], double coords[], bool cartesian, bool pbc) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  g.nbl_init = false;
  if (!g.init || !g.set.loaded || natoms < 1) {
    set_err(AENET_STATUS_ERR_INIT, "aenet_nbl_init: potentials not loaded");
    return;
  }
  // lattice coordinates: crystallographic reciprocal lattice (no 2*pi)
  std::vector<double> frac(3 * (size_t)natoms);
  const double *A = latt_vec;
  if (cartesian) {
    double r[3][3], B[9];
    auto vp = [](const double *a, const double *b, double *c) {
      c[0] = a[1] * b[2] - a[2] * b[1];
      c[1] = a[2] * b[0] - a[0] * b[2];
      c[2] = a[0] * b[1] - a[1] * b[0];
    };
    vp(A + 3, A + 6, r[0]);
    vp(A + 0, A + 6, r[1]);
    vp(A + 0, A + 3, r[2]);
    for (int x = 0; x < 3; ++x) r[1][x] = -r[1][x];
    auto a = [A](int i, int j) { return A[(j - 1) * 3 + (i - 1)]; };
    double V = a(1, 1) * a(2, 2) * a(3, 3) + a(2, 1) * a(3, 2) * a(1, 3) + a(3, 1) * a(1, 2) * a(2, 3) -
               a(3, 1) * a(2, 2) * a(1, 3) - a(1, 1) * a(3, 2) * a(2, 3) - a(2, 1) * a(1, 2) * a(3, 3);
    for (int i = 0; i < 3; ++i)
      for (int x = 0; x < 3; ++x) B[x * 3 + i] = r[i][x] / V;
    for (int i = 0; i < natoms; ++i)
      for (int rr = 0; rr < 3; ++rr) {
        double acc = B[rr] * coords[3 * i] + B[3 + rr] * coords[3 * i + 1];
        frac[3 * i + rr] = acc + B[6 + rr] * coords[3 * i + 2];
      }
  } else {
    std::memcpy(frac.data(), coords, frac.size() * sizeof(double));
  }
  // the batch machinery takes Cartesian input; give it L*frac and a
  // lattice whose reciprocal maps it back exactly is not guaranteed, so
  // the lattice coordinates are handed over directly
  Batch b;
  std::string err;
  int pbcv = pbc ? 1 : 0;
  int st = b.create_lattice_coords(g.set, natoms, latt_vec, pbcv, err);
  if (st) {
    set_err(st, err);
    return;
  }
  hipStream_t s = lib_stream();
  DBuf<double> dfrac;
  DBuf<int> dtypes;
  if (dfrac.upload(frac.data(), frac.size(), s) || dtypes.upload(atom_types, natoms, s)) {
    set_err(AENET_STATUS_ERR_MALLOC, "aenet_nbl_init: device allocation failed");
    return;
  }
  int cap = std::min(g.set.nnb_max, 512);
  for (int attempt = 0; attempt < 2; ++attempt) {
    DBuf<int> cnt, lst, typ;
    DBuf<double> coo, dst;
    if (cnt.alloc(natoms) || lst.alloc((size_t)natoms * cap) || typ.alloc((size_t)natoms * cap) ||
        coo.alloc(3 * (size_t)natoms * cap) || dst.alloc((size_t)natoms * cap)) {
      set_err(AENET_STATUS_ERR_MALLOC, "aenet_nbl_init: device allocation failed");
      return;
    }
    st = b.neighbor_lists(dfrac.p, dtypes.p, cap, cnt.p, coo.p, dst.p, lst.p, typ.p, s);
    std::string ferr;
    int fst = st ? st : b.check_flags(s, ferr);
    if (fst == AENET_STATUS_ERR_CAPACITY && attempt == 0 && cap < g.set.nnb_max) {
      cap = g.set.nnb_max;
      continue;
    }
    if (fst) {
      set_err(fst, ferr);
      return;
    }
    g.nbl_count.resize(natoms);
    g.nbl_list.resize((size_t)natoms * cap);
    g.nbl_type.resize((size_t)natoms * cap);
    g.nbl_coo.resize(3 * (size_t)natoms * cap);
    g.nbl_dist.resize((size_t)natoms * cap);
    (void)hipMemcpyAsync(g.nbl_count.data(), cnt.p, natoms * sizeof(int), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(g.nbl_list.data(), lst.p, g.nbl_list.size() * sizeof(int), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(g.nbl_type.data(), typ.p, g.nbl_type.size() * sizeof(int), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(g.nbl_coo.data(), coo.p, g.nbl_coo.size() * sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(g.nbl_dist.data(), dst.p, g.nbl_dist.size() * sizeof(double), hipMemcpyDeviceToHost, s);
    // wrapped coordinates back, Cartesian: matmul(latt_vec, coo_latt)
    std::vector<double> wrapped(3 * (size_t)natoms);
    (void)hipMemcpyAsync(wrapped.data(), b.wrapped_coords(), wrapped.size() * sizeof(double), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
      set_err(AENET_STATUS_ERR_DEVICE, "aenet_nbl_init: device error");
      return;
    }
    for (int i = 0; i < natoms; ++i)
      for (int rr = 0; rr < 3; ++rr) {
        double acc = A[rr] * wrapped[3 * i] + A[3 + rr] * wrapped[3 * i + 1];
        coords[3 * i + rr] = acc + A[6 + rr] * wrapped[3 * i + 2];
      }
    g.nbl_natoms = natoms;
    g.nbl_cap = cap;
    g.nbl_init = true;
    return;
  }
}

void aenet_nbl_final(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  g.nbl_init = false;
  g.nbl_count.clear();
  g.nbl_list.clear();
  g.nbl_type.clear();
  g.nbl_coo.clear();
  g.nbl_dist.clear();
}

void aenet_nbl_neighbors(int iatom, int *nnb, double nbcoo[], double nbdist[], int nblist[], int nbtype[]) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!g.nbl_init || iatom < 1 || iatom > g.nbl_natoms) {
    set_err(AENET_STATUS_ERR_INIT, "aenet_nbl_neighbors: no neighbor list for this atom");
    *nnb = 0;
    return;
  }
  int n = g.nbl_count[iatom - 1];
  int m = std::min(n, std::max(*nnb, 0));
  size_t base = (size_t)(iatom - 1) * g.nbl_cap;
  std::memcpy(nbcoo, &g.nbl_coo[3 * base], 3 * (size_t)m * sizeof(double));
  std::memcpy(nbdist, &g.nbl_dist[base], (size_t)m * sizeof(double));
  std::memcpy(nblist, &g.nbl_list[base], (size_t)m * sizeof(int));
  std::memcpy(nbtype, &g.nbl_type[base], (size_t)m * sizeof(int));
  // the reference stops the program when the caller's arrays are too
  // small (src/ext/lclist.f90:552-556, :653-660); here *nnb is the number
  // of entries written and the error is recorded for aenet_amd_last_error
  if (m < n)
    set_err(AENET_STATUS_ERR_CAPACITY, "aenet_nbl_neighbors: atom " + std::to_string(iatom) + " has " +
                                           std::to_string(n) + " neighbors, room for " + std::to_string(*nnb) +
                                           "; the list is truncated");
  *nnb = m;
}

// ---------------------------------------------------------------------
// batched entry points (include/aenet_amd.h)
// ---------------------------------------------------------------------

void aenet_structures_energy_and_forces(int nstruct, const int natoms[], const double latt[], const int pbc[],
                                        const int types[], const double coords[], double E_tot[],
                                        double E_coh[], double E_atoms[], double F[], int *stat) {
  aenet_amd::structures_energy_and_forces(nstruct, natoms, latt, pbc, types, coords, E_tot, E_coh, E_atoms, F,
                                          nullptr, "aenet_structures_energy_and_forces", stat);
}

}  // extern "C"

namespace aenet_amd {

int set_last_error(int st, const std::string &msg) { return set_err(st, msg); }

void structures_energy_and_forces(int nstruct, const int natoms[], const double latt[], const int pbc[],
                                  const int types[], const double coords[], double E_tot[], double E_coh[],
                                  double E_atoms[], double F[], ScratchArena *arena, const char *who, int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!g.init || !g.set.loaded) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "potentials not loaded");
    return;
  }
  if (nstruct < 1 || !natoms || !latt || !pbc || !types || !coords) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, std::string(who) + ": no structures or missing arrays");
    return;
  }
  hipStream_t s = lib_stream();
  // with an arena, the previous call's work on it must be over before its
  // memory is rewritten (an error path may have left kernels queued)
  if (arena && hipStreamSynchronize(s) != hipSuccess) {
    *stat = set_err(AENET_STATUS_ERR_DEVICE, "device error");
    return;
  }
  ArenaScope scope(arena);
  const auto fail = [&](int st, const std::string &msg) {
    if (arena && arena->overflow)
      return set_err(AENET_STATUS_ERR_CAPACITY, std::string(who) + ": the structures need more device scratch than its " +
                                                    std::to_string(arena->cap >> 20) + " MiB");
    return set_err(st, msg);
  };
  std::string err;
  Batch b;
  int st = b.create(g.set, nstruct, natoms, latt, pbc, coords, err);
  if (st) {
    *stat = fail(st, err);
    return;
  }
  int N = b.natoms();
  for (int i = 0; i < N; ++i)
    if (types[i] < 1 || types[i] > g.set.ntypes) {
      *stat = set_err(AENET_STATUS_ERR_TYPE, "atom type out of range");
      return;
    }
  DBuf<double> dc, dE, dES, dF;
  DBuf<int> dt;
  if (dc.upload(coords, 3 * (size_t)N, s) || dt.upload(types, N, s) || dE.alloc(N) || dES.alloc(2 * (size_t)nstruct) ||
      (F && dF.alloc(3 * (size_t)N))) {
    *stat = fail(AENET_STATUS_ERR_MALLOC, "device allocation failed");
    return;
  }
  st = b.eval(*g.dev, dc.p, dt.p, dE.p, dES.p, F ? dF.p : nullptr, s);
  if (!st) st = b.check_flags(s, err);
  if (st == AENET_STATUS_ERR_CAPACITY && b.nb_cap() < AENET_AMD_NB_CAP) {
    // denser than estimated: once more with the largest list capacity
    err.clear();
    st = b.set_nb_cap(AENET_AMD_NB_CAP);
    if (!st) st = b.eval(*g.dev, dc.p, dt.p, dE.p, dES.p, F ? dF.p : nullptr, s);
    if (!st) st = b.check_flags(s, err);
  }
  if (st) {
    *stat = fail(st, err.empty() ? "kernel launch failed" : err);
    return;
  }
  std::vector<double> es(2 * (size_t)nstruct);
  (void)hipMemcpyAsync(es.data(), dES.p, es.size() * sizeof(double), hipMemcpyDeviceToHost, s);
  if (E_atoms) (void)hipMemcpyAsync(E_atoms, dE.p, N * sizeof(double), hipMemcpyDeviceToHost, s);
  if (F) (void)hipMemcpyAsync(F, dF.p, 3 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    *stat = set_err(AENET_STATUS_ERR_DEVICE, "device error");
    return;
  }
  for (int i = 0; i < nstruct; ++i) {
    if (E_tot) E_tot[i] = es[i];
    if (E_coh) E_coh[i] = es[nstruct + i];
  }
  *stat = AENET_STATUS_OK;
}

}  // namespace aenet_amd

extern "C" {

void aenet_amd_batch_create(int nstruct, const int natoms[], const double latt[], const int pbc[],
                            const double coords_host[], aenet_amd_batch **batch, int *stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  *batch = nullptr;
  if (!g.init || !g.set.loaded) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "potentials not loaded");
    return;
  }
  std::unique_ptr<aenet_amd_batch> b(new aenet_amd_batch());
  std::string err;
  int st = b->batch.create(g.set, nstruct, natoms, latt, pbc, coords_host, err);
  if (st) {
    *stat = set_err(st, err);
    return;
  }
  *batch = b.release();
  *stat = AENET_STATUS_OK;
}

void aenet_amd_batch_destroy(aenet_amd_batch *batch) { delete batch; }

void aenet_amd_batch_replan(aenet_amd_batch *batch, const double coords_host[], int *stat) {
  if (!batch || !coords_host) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_batch_replan: no batch or no coordinates");
    return;
  }
  std::string err;
  int st = batch->batch.replan(coords_host, err);
  *stat = st ? set_err(st, err) : AENET_STATUS_OK;
}

int aenet_amd_batch_has_isolated(const aenet_amd_batch *batch) {
  return batch && batch->batch.has_isolated() ? 1 : 0;
}

int aenet_amd_batch_natoms(const aenet_amd_batch *batch) { return batch ? batch->batch.natoms() : 0; }

void aenet_amd_batch_eval(aenet_amd_batch *batch, const double *d_coords, const int *d_types, double *d_E_atoms,
                          double *d_E_struct, double *d_F, void *stream, int *stat) {
  // the lock keeps aenet_final from releasing the potentials during the launch
  std::lock_guard<std::mutex> lk(g_mu);
  Global &g = G();
  if (!batch || !g.dev) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or potentials not loaded");
    return;
  }
  if (!batch->batch.fits(*g.dev)) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "potential cutoff exceeds the cutoff the batch was created for");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.eval(*g.dev, d_coords, d_types, d_E_atoms, d_E_struct, d_F, s);
  if (*stat) set_err(*stat, "kernel launch failed");
}

void aenet_amd_batch_set_timing(aenet_amd_batch *batch, int enable) {
  if (batch) batch->batch.set_timing(enable != 0);
}

void aenet_amd_batch_kernel_time(aenet_amd_batch *batch, double *ms, int *count, int *stat) {
  if (!batch) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch");
    return;
  }
  *stat = batch->batch.kernel_time(*ms, *count);
  if (*stat) set_err(*stat, "event timing failed");
}

void aenet_amd_batch_search_time(aenet_amd_batch *batch, double *ms, int *count, int *stat) {
  if (!batch) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch");
    return;
  }
  *stat = batch->batch.search_time(*ms, *count);
  if (*stat) set_err(*stat, "event timing failed");
}

void aenet_amd_batch_check(aenet_amd_batch *batch, void *stream, int *stat) {
  if (!batch) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch");
    return;
  }
  std::string err;
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.check_flags(s, err);
  if (*stat) set_err(*stat, err);
}

void aenet_amd_batch_neighbors(aenet_amd_batch *batch, const double *d_coords, const int *d_types, int cap,
                               int *d_count, double *d_coo, double *d_dist, int *d_list, int *d_type, void *stream,
                               int *stat) {
  if (!batch) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.neighbor_lists(d_coords, d_types, cap, d_count, d_coo, d_dist, d_list, d_type, s);
  if (*stat) set_err(*stat, "kernel launch failed");
}

void aenet_amd_batch_descriptors(aenet_amd_batch *batch, const double *d_coords, const int *d_types, int scaled,
                                 double *d_G, int stride, void *stream, int *stat) {
  Global &g = G();
  if (!batch || !g.dev) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or potentials not loaded");
    return;
  }
  if (stride < g.set.nsf_max) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "descriptor stride smaller than nsf_max");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.descriptors(*g.dev, d_coords, d_types, scaled, d_G, stride, s);
  if (*stat) set_err(*stat, "kernel launch failed");
}


// ---------------------------------------------------------------------
// independent potential sets (include/aenet_amd.h)
// ---------------------------------------------------------------------

void aenet_amd_model_load(int ntypes, char *atom_types[], char *files[], aenet_amd_model **model, int *stat) {
  *model = nullptr;
  if (ntypes < 1 || ntypes > AENET_AMD_MAX_TYPES) {
    *stat = set_err(AENET_STATUS_ERR_MALLOC, "aenet_amd_model_load: 1 to 8 species are supported");
    return;
  }
  if ((*stat = ensure_device())) return;
  std::unique_ptr<aenet_amd_model> m(new aenet_amd_model());
  m->set.ntypes = ntypes;
  for (int i = 0; i < ntypes; ++i) m->set.names.push_back(cname(atom_types[i]));
  m->set.pots.assign(ntypes, PotHost());
  std::string err;
  for (int t = 0; t < ntypes; ++t) {
    struct stat sb;
    if (!files[t] || ::stat(files[t], &sb) != 0) {
      *stat = set_err(AENET_STATUS_ERR_IO, std::string("aenet_amd_model_load: file not found: ") +
                                               (files[t] ? files[t] : "(null)"));
      return;
    }
    int st = read_nn(files[t], m->set.pots[t], err);
    if (st) {
      *stat = set_err(st, err);
      return;
    }
  }
  int st = m->set.finalize(err);
  if (st) {
    *stat = set_err(st, err);
    return;
  }
  hipStream_t s = lib_stream();
  if (m->dev.upload(m->set, s) || m->flags.alloc(1)) {
    *stat = set_err(AENET_STATUS_ERR_DEVICE, "aenet_amd_model_load: device upload failed");
    return;
  }
  launch_fill_i32(m->flags.p, 0, 1, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    *stat = set_err(AENET_STATUS_ERR_DEVICE, "aenet_amd_model_load: device error");
    return;
  }
  *model = m.release();
  *stat = AENET_STATUS_OK;
}

void aenet_amd_model_destroy(aenet_amd_model *model) { delete model; }

void aenet_amd_model_info(const aenet_amd_model *model, double *Rc_max, int *nsf_max, int *wsize, int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no model");
    return;
  }
  if (Rc_max) *Rc_max = dm->rc_max;
  if (nsf_max) *nsf_max = dm->nsf_max;
  if (wsize) *wsize = dm->wsize;
  *stat = AENET_STATUS_OK;
}

void aenet_amd_model_species_weights(const aenet_amd_model *model, int type_id, int *offset, int *size,
                                     int *stat) {
  const PotSet *ps = model ? &model->set : (G().dev ? &G().set : nullptr);
  if (!ps || type_id < 1 || type_id > ps->ntypes) {
    *stat = set_err(AENET_STATUS_ERR_TYPE, "no model or type id out of range");
    return;
  }
  const SpeciesDev &sp = ps->species[type_id - 1];
  *offset = sp.w_off[0];
  *size = ps->pots[type_id - 1].net.Wsize;
  *stat = AENET_STATUS_OK;
}

void aenet_amd_model_weights(aenet_amd_model *model, float *d_W, int to_model, void *stream, int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!dm || !d_W) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no model or no buffer");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  float *W = const_cast<float *>(dm->W.p);
  hipError_t e = to_model ? hipMemcpyAsync(W, d_W, dm->wsize * sizeof(float), hipMemcpyDeviceToDevice, s)
                          : hipMemcpyAsync(d_W, W, dm->wsize * sizeof(float), hipMemcpyDeviceToDevice, s);
  *stat = e == hipSuccess ? AENET_STATUS_OK : set_err(AENET_STATUS_ERR_DEVICE, "weight copy failed");
}

void aenet_amd_model_descriptor_derivatives(aenet_amd_model *model, int nenv, const int env_start[],
                                            const int type_i[], const double coo_i[], const double coo_j[],
                                            const int type_j[], int stride, double values[], double deriv_i[],
                                            double deriv_j[], int *stat) {
  if (!model || nenv < 0 || !env_start || stride < model->dev.nsf_max) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_model_descriptor_derivatives: bad arguments");
    return;
  }
  const int nt = model->set.ntypes;
  for (int e = 0; e < nenv; ++e) {
    if (type_i[e] < 1 || type_i[e] > nt || env_start[e + 1] < env_start[e]) {
      *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_amd_model_descriptor_derivatives: bad environment");
      return;
    }
    for (int q = env_start[e]; q < env_start[e + 1]; ++q)
      if (type_j[q] < 1 || type_j[q] > nt) {
        *stat = set_err(AENET_STATUS_ERR_TYPE, "aenet_amd_model_descriptor_derivatives: type out of range");
        return;
      }
  }
  // environment e becomes nsf(e) copies, copy f differentiating feature f;
  // chunks of environments keep the copies' neighbor data bounded
  const hipStream_t s = lib_stream();
  int e0 = 0;
  while (e0 < nenv) {
    int e1 = e0;
    size_t copies = 0, nb = 0;
    while (e1 < nenv) {
      const int nsf = model->set.pots[type_i[e1] - 1].stp.nsf;
      const size_t add = (size_t)nsf * (env_start[e1 + 1] - env_start[e1]);
      if (e1 > e0 && (nb + add > (size_t)(4 << 20) || copies + nsf > 65536)) break;
      copies += nsf;
      nb += add;
      ++e1;
    }
    // copies are laid out feature-major within an environment and padded
    // to a multiple of coef_unit = stride, so copy c differentiates c % stride
    const size_t nc = (size_t)(e1 - e0) * stride;
    std::vector<int> start(nc + 1), ct(nc), tj;
    std::vector<double> ci(3 * nc), cj;
    size_t at = 0;
    for (int e = e0; e < e1; ++e) {
      const int n = env_start[e + 1] - env_start[e];
      for (int f = 0; f < stride; ++f) {
        const size_t c = (size_t)(e - e0) * stride + f;
        start[c] = (int)at;
        ct[c] = type_i[e];
        std::memcpy(&ci[3 * c], &coo_i[3 * (size_t)e], 3 * sizeof(double));
        // features beyond the species' nsf: an empty environment (nothing to differentiate)
        if (f < model->set.pots[type_i[e] - 1].stp.nsf) {
          cj.insert(cj.end(), coo_j + 3 * (size_t)env_start[e], coo_j + 3 * (size_t)env_start[e + 1]);
          tj.insert(tj.end(), type_j + env_start[e], type_j + env_start[e + 1]);
          at += n;
        }
      }
    }
    start[nc] = (int)at;
    std::vector<double> E(nc), dFi(3 * nc), dFj(3 * at + 3), G(nc * stride);
    std::string err;
    int st = model->env.eval(model->dev, (int)nc, start.data(), ct.data(), ci.data(), cj.data(), tj.data(), true,
                             E.data(), dFi.data(), dFj.data(), G.data(), stride, s, err, stride);
    if (st) {
      *stat = set_err(st, err);
      return;
    }
    for (int e = e0; e < e1; ++e) {
      const int n = env_start[e + 1] - env_start[e];
      const size_t c0 = (size_t)(e - e0) * stride;
      if (values) std::memcpy(&values[(size_t)e * stride], &G[c0 * stride], stride * sizeof(double));
      for (int f = 0; f < stride; ++f) {
        const size_t c = c0 + f;
        const bool has = f < model->set.pots[type_i[e] - 1].stp.nsf;
        for (int x = 0; x < 3; ++x) {
          if (deriv_i) deriv_i[((size_t)e * stride + f) * 3 + x] = has ? -dFi[3 * c + x] : 0.0;
          if (deriv_j)
            for (int j = 0; j < n; ++j)
              deriv_j[((size_t)(env_start[e] + j) * stride + f) * 3 + x] =
                  has ? -dFj[3 * ((size_t)start[c] + j) + x] : 0.0;
        }
      }
    }
    e0 = e1;
  }
  *stat = AENET_STATUS_OK;
}

void aenet_amd_batch_create_model(const aenet_amd_model *model, int nstruct, const int natoms[], const double latt[],
                                  const int pbc[], const double coords_host[], aenet_amd_batch **batch, int *stat) {
  if (!model) {
    aenet_amd_batch_create(nstruct, natoms, latt, pbc, coords_host, batch, stat);
    return;
  }
  *batch = nullptr;
  std::unique_ptr<aenet_amd_batch> b(new aenet_amd_batch());
  std::string err;
  int st = b->batch.create(model->set, nstruct, natoms, latt, pbc, coords_host, err);
  if (st) {
    *stat = set_err(st, err);
    return;
  }
  *batch = b.release();
  *stat = AENET_STATUS_OK;
}

void aenet_amd_batch_eval_model(aenet_amd_batch *batch, const aenet_amd_model *model, const double *d_coords,
                                const int *d_types, double *d_E_atoms, double *d_E_struct, double *d_F, void *stream,
                                int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  if (!batch->batch.fits(*dm)) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "model cutoff exceeds the cutoff the batch was created for");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.eval(*dm, d_coords, d_types, d_E_atoms, d_E_struct, d_F, s);
  if (*stat) set_err(*stat, "kernel launch failed");
}

void aenet_amd_batch_autotune(aenet_amd_batch *batch, const aenet_amd_model *model, const double *d_coords,
                              const int *d_types, int forces, int repeats, double results[], int *ncand,
                              int chosen[3], void *stream, int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  if (!d_coords || !d_types || repeats < 1) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_batch_autotune: bad arguments");
    return;
  }
  if (!batch->batch.fits(*dm)) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "model cutoff exceeds the cutoff the batch was created for");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  std::vector<std::array<double, 4>> out;
  std::string err;
  *stat = batch->batch.autotune(*dm, d_coords, d_types, forces != 0, repeats, s, out, err);
  if (*stat) {
    set_err(*stat, err);
    return;
  }
  if (ncand) *ncand = (int)out.size();
  double best = 1e300;
  int bw = 0, bs = -1, bt = 0;
  for (size_t k = 0; k < out.size(); ++k) {
    if (results && k < 32)
      for (int c = 0; c < 4; ++c) results[4 * k + c] = out[k][c];
    if (out[k][3] < best) {
      best = out[k][3];
      bw = (int)out[k][0];
      bs = (int)out[k][1];
      bt = (int)out[k][2];
    }
  }
  if (chosen) {
    chosen[0] = bw;
    chosen[1] = bs;
    chosen[2] = bt;
  }
}

void aenet_amd_batch_set_launch(aenet_amd_batch *batch, const aenet_amd_model *model, int waves, int staged,
                                int tile, int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  std::string err;
  *stat = batch->batch.set_launch(*dm, waves, staged, tile, err);
  if (*stat) set_err(*stat, "aenet_amd_batch_set_launch: " + err);
}

void aenet_amd_batch_set_buckets(aenet_amd_batch *batch, const aenet_amd_model *model, int n, const int caps[],
                                 const int shapes[], int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  std::string err;
  *stat = batch->batch.set_buckets(*dm, n, caps, shapes, err);
  if (*stat) set_err(*stat, "aenet_amd_batch_set_buckets: " + err);
}

void aenet_amd_batch_buckets(aenet_amd_batch *batch, const aenet_amd_model *model, int *n, int caps[],
                             int shapes[], int atoms[], double ms[], double totals[2], int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm || !n) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  Batch::BucketPlan pl;
  batch->batch.buckets(*dm, pl);
  *n = pl.n;
  for (int k = 0; k < pl.n; ++k) {
    if (caps) caps[k] = pl.cap[k];
    if (shapes)
      for (int c = 0; c < 3; ++c) shapes[3 * k + c] = pl.shape[k][c];
    if (atoms) atoms[k] = pl.atoms[k];
    if (ms) ms[k] = pl.ms[k];
  }
  if (totals) {
    totals[0] = pl.ms_single;
    totals[1] = pl.ms_buckets;
  }
  *stat = AENET_STATUS_OK;
}

void aenet_amd_bucket_caps(const int counts[], int natoms, int nb_cap, int caps[], int *ncaps) {
  std::vector<int> c(counts, counts + (natoms > 0 ? natoms : 0));
  const std::vector<int> r = bucket_caps_for(c, nb_cap);
  *ncaps = (int)r.size();
  for (size_t k = 0; k < r.size(); ++k) caps[k] = r[k];
}

void aenet_amd_batch_eval_active(aenet_amd_batch *batch, const aenet_amd_model *model, const double *d_coords,
                                 const int *d_types, const int *d_active, const int *d_nactive, int max_active,
                                 int reuse_lists, double *d_E_atoms, void *stream, int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  if (!d_active || !d_nactive || !d_E_atoms || max_active < 0 || max_active > batch->batch.natoms()) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_batch_eval_active: bad arguments");
    return;
  }
  if (!batch->batch.fits(*dm)) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "model cutoff exceeds the cutoff the batch was created for");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.eval_active(*dm, d_coords, d_types, d_active, d_nactive, max_active, reuse_lists != 0,
                                   d_E_atoms, s);
  if (*stat) set_err(*stat, "kernel launch failed");
}

void aenet_amd_batch_vjp(aenet_amd_batch *batch, const aenet_amd_model *model, const double *d_coords,
                         const int *d_types, const double *d_g, double *d_grad_coords, float *d_grad_W, void *stream,
                         int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  if (!d_g || !d_grad_coords || !batch->batch.fits(*dm)) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_batch_vjp: bad arguments or model cutoff too large");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.vjp(*dm, d_coords, d_types, d_g, d_grad_coords, d_grad_W, s);
  if (*stat) set_err(*stat, "kernel launch failed");
}

void aenet_amd_batch_descriptors_model(aenet_amd_batch *batch, const aenet_amd_model *model, const double *d_coords,
                                       const int *d_types, int scaled, double *d_G, int stride, void *stream,
                                       int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!batch || !dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no batch or no model");
    return;
  }
  if (stride < dm->nsf_max || !batch->batch.fits(*dm)) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "descriptor stride smaller than nsf_max, or model cutoff too large");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  *stat = batch->batch.descriptors(*dm, d_coords, d_types, scaled, d_G, stride, s);
  if (*stat) set_err(*stat, "kernel launch failed");
}

void aenet_amd_model_net_grad(aenet_amd_model *model, int natoms, const double *d_G, int stride, const int *d_types,
                              const double *d_E_ref, const double *d_weight, const double *d_upstream, double coef,
                              double *d_E_out, float *d_grad, double *d_loss, void *stream, int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!dm) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no model");
    return;
  }
  if (natoms < 0 || !d_G || !d_types || stride < dm->nsf_max || (!d_E_ref && !d_upstream && d_grad)) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_model_net_grad: bad arguments");
    return;
  }
  NetGradArgs a;
  a.G = d_G;
  a.stride = stride;
  a.types = d_types;
  a.E_ref = d_E_ref;
  a.weight = d_weight;
  a.upstream = d_upstream;
  a.coef = coef;
  a.E_out = d_E_out;
  a.grad = d_grad;
  a.loss = d_loss;
  a.flags = model ? model->flags.p : nullptr;
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  launch_net_grad(dm->view(), a, natoms, dm->nsf_max, dm->nl_max, dm->wsize, d_grad != nullptr, s, dm->hmax);
  *stat = hipGetLastError() == hipSuccess ? AENET_STATUS_OK : set_err(AENET_STATUS_ERR_DEVICE, "launch failed");
}

void aenet_amd_model_adam(aenet_amd_model *model, const float *d_grad, float *d_m, float *d_v, int t, double lr,
                          double b1, double b2, double eps, void *stream, int *stat) {
  const DeviceModel *dm = pick_model(model);
  if (!dm || !d_grad || !d_m || !d_v || t < 1) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_model_adam: bad arguments");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  launch_adam(const_cast<float *>(dm->W.p), d_grad, d_m, d_v, dm->wsize, t, (float)lr, (float)b1, (float)b2,
              (float)eps, s);
  *stat = hipGetLastError() == hipSuccess ? AENET_STATUS_OK : set_err(AENET_STATUS_ERR_DEVICE, "launch failed");
}

void aenet_amd_model_check(aenet_amd_model *model, void *stream, int *stat) {
  if (!model) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no model");
    return;
  }
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream
  int h = 0;
  if (hipMemcpyAsync(&h, model->flags.p, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    *stat = set_err(AENET_STATUS_ERR_DEVICE, "device error");
    return;
  }
  if (h) {
    launch_fill_i32(model->flags.p, 0, 1, s);
    *stat = set_err(AENET_STATUS_ERR_TYPE, "training kernel reported an atom type outside the model");
    return;
  }
  *stat = AENET_STATUS_OK;
}


// ---------------------------------------------------------------------
// Monte Carlo (include/aenet_amd.h)
// ---------------------------------------------------------------------

void aenet_amd_mc_create(const aenet_amd_model *full, const aenet_amd_model *first, int nchains, int natoms,
                         const double latt[9], const int types[], const double coords[], unsigned long long seed,
                         aenet_amd_mc **mc, int *stat) {
  *mc = nullptr;
  if (!full) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "aenet_amd_mc_create: no full potential");
    return;
  }
  if (first) {
    const PotSet &s = first->set;
    bool ok = s.desc == DESC_CHEBYSHEV && s.ntypes == full->set.ntypes && s.nsf_max <= 32;
    for (const SpeciesDev &sp : s.species) {
      ok = ok && sp.a_Rc == 0.0f && sp.nlayers == 3 && sp.nnodes[1] <= 16 && sp.nnodes[2] == 1 && sp.r_N <= 16;
    }
    if (!ok) {
      *stat = set_err(AENET_AMD_ERR_ARG_VALUE,
                      "first-stage potential must be radial-only Chebyshev (angular Rc 0) with at most 32 "
                      "features and one hidden layer of at most 16 nodes, for the same species as the full "
                      "potential");
      return;
    }
  }
  std::unique_ptr<aenet_amd_mc> m(new aenet_amd_mc());
  std::string err;
  int st = m->mc.create(full->set, &full->dev, first ? &first->dev : nullptr, nchains, natoms, latt, types, coords,
                        seed, err);
  if (st) {
    *stat = set_err(st, err);
    return;
  }
  *mc = m.release();
  *stat = AENET_STATUS_OK;
}

void aenet_amd_mc_destroy(aenet_amd_mc *mc) { delete mc; }

void aenet_amd_mc_set(aenet_amd_mc *mc, double beta, double p_swap, double max_disp, int max_screen, int *stat) {
  if (!mc || !(beta >= 0.0) || !(p_swap >= 0.0 && p_swap <= 1.0) || !(max_disp >= 0.0) || max_screen < 1 ||
      max_screen > 1024) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_mc_set: bad arguments");
    return;
  }
  mc->mc.set_params(beta, p_swap, max_disp, max_screen);
  *stat = AENET_STATUS_OK;
}

void aenet_amd_mc_refresh(aenet_amd_mc *mc, int *stat) {
  if (!mc) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
    return;
  }
  std::string err;
  *stat = mc->mc.refresh(err);
  if (*stat) set_err(*stat, err);
}

void aenet_amd_mc_run(aenet_amd_mc *mc, int nsteps, int delayed, int sample_every, int use_graph, int *stat) {
  if (!mc) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
    return;
  }
  std::string err;
  *stat = mc->mc.run(nsteps, delayed != 0, sample_every, use_graph != 0, err);
  if (*stat) set_err(*stat, err);
}

void aenet_amd_mc_stats(aenet_amd_mc *mc, long long out[4], int reset, int *stat) {
  if (!mc) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
    return;
  }
  *stat = mc->mc.stats(out);
  if (!*stat && reset) mc->mc.reset_stats();
  if (*stat) set_err(*stat, "device error");
}

void aenet_amd_mc_energies(aenet_amd_mc *mc, double E_chain[], int *stat) {
  *stat = mc ? mc->mc.energies(E_chain) : set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
}

void aenet_amd_mc_samples(aenet_amd_mc *mc, double out[], long long *nsamples, int *stat) {
  if (!mc) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
    return;
  }
  long long n = mc->mc.samples(out);
  *nsamples = n < 0 ? 0 : n;
  *stat = n < 0 ? set_err(AENET_STATUS_ERR_DEVICE, "device error") : AENET_STATUS_OK;
}

void aenet_amd_mc_trace(aenet_amd_mc *mc, long long max_props, double out[], long long *nprops, int *stat) {
  if (!mc) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
    return;
  }
  if (max_props >= 0) {
    *stat = mc->mc.set_trace(max_props);
    if (*stat) set_err(*stat, "device allocation failed");
    if (nprops) *nprops = 0;
    return;
  }
  long long n = mc->mc.trace(out);
  if (nprops) *nprops = n < 0 ? 0 : n;
  *stat = n < 0 ? set_err(AENET_STATUS_ERR_DEVICE, "device error") : AENET_STATUS_OK;
}

void aenet_amd_mc_state(aenet_amd_mc *mc, double coords[], int types[], int *stat) {
  *stat = mc ? mc->mc.state(coords, types) : set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
}

void aenet_amd_mc_first_stage_state(aenet_amd_mc *mc, double E_atoms[], double G[], int *nsf, int *stat) {
  if (!mc) {
    *stat = set_err(AENET_STATUS_ERR_INIT, "no Monte Carlo state");
    return;
  }
  if (nsf) *nsf = mc->mc.nsf_student();
  *stat = mc->mc.student_state(E_atoms, G);
}

void aenet_amd_model_fit_cells(aenet_amd_model *model, int ncells, int natoms, const double latt[9],
                               const double *d_coords, const int *d_types, const int *d_sel, int nsel,
                               const double *d_E_ref, const double *d_F_ref, double coefE, double coefF,
                               float *d_grad, double *d_loss, void *stream, int *stat) {
  if (!model || !d_coords || !d_types || !d_sel || !d_E_ref || !d_F_ref || !d_grad || !d_loss || ncells < 1 ||
      natoms < 2 || natoms > 256 || nsel < 0) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "aenet_amd_model_fit_cells: bad arguments");
    return;
  }
  const PotSet &s = model->set;
  bool ok = s.desc == DESC_CHEBYSHEV;
  for (const SpeciesDev &sp : s.species)
    ok = ok && sp.a_Rc == 0.0f && sp.r_N <= 16 && sp.nlayers == 3 && sp.nnodes[1] <= 16 && sp.nnodes[2] == 1 &&
         sp.fa[2] == 0;
  if (!ok) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE,
                    "fit: radial-only Chebyshev basis (at most 16 terms), one hidden layer of at most 16 nodes, "
                    "linear output");
    return;
  }
  MCDev m;
  std::memset(&m, 0, sizeof(m));
  m.nchains = ncells;
  m.natoms = natoms;
  std::memcpy(m.L, latt, sizeof(m.L));
  if (!mc_invert_cell(m.L, m.Binv) || mc_min_width(m.L) <= 2.0 * s.Rc_max) {
    *stat = set_err(AENET_AMD_ERR_ARG_VALUE, "fit: singular cell or cell narrower than twice the cutoff");
    return;
  }
  m.coords = const_cast<double *>(d_coords);
  m.types = const_cast<int *>(d_types);
  hipStream_t st = (hipStream_t)stream;  // NULL: the default stream
  launch_first_stage_fit(m, model->dev.view(), d_sel, nsel, d_E_ref, d_F_ref, coefE, coefF, d_grad, model->dev.wsize,
                         d_loss, st);
  *stat = hipGetLastError() == hipSuccess ? AENET_STATUS_OK : set_err(AENET_STATUS_ERR_DEVICE, "launch failed");
}

}  // extern "C"

// engine.h -- device-resident potential sets and structure batches.
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <map>
#include <string>
#include <vector>

#include "aenet_amd_internal.h"
#include "kernels.h"
#include "model.h"
#include "plan.h"

namespace aenet_amd {

// A fixed region of device memory handed out front to back.  While a
// thread has one installed (ArenaScope), every DBuf it allocates takes its
// memory from the region instead of hipMalloc, and releasing it frees
// nothing.  The region is the caller's; its contents are not cleared.
struct ScratchArena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  bool overflow = false;  // a request did not fit
  void *take(size_t bytes) {
    const size_t b = (bytes + 255) & ~(size_t)255;
    if (overflow || b > cap - used) {
      overflow = true;
      return nullptr;
    }
    void *q = base + used;
    used += b;
    return q;
  }
};
ScratchArena *&current_arena();  // of the calling thread (null: hipMalloc)
struct ArenaScope {
  ScratchArena *prev;
  explicit ArenaScope(ScratchArena *a) : prev(current_arena()) { current_arena() = a; }
  ~ArenaScope() { current_arena() = prev; }
  ArenaScope(const ArenaScope &) = delete;
  ArenaScope &operator=(const ArenaScope &) = delete;
};

// RAII device buffer
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  bool owned = false;  // from hipMalloc (not from an arena)
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  int alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    if (ScratchArena *a = current_arena()) {
      p = static_cast<T *>(a->take(count * sizeof(T)));
      if (!p) return -1;
      n = count;
      return 0;
    }
    if (hipMalloc((void **)&p, count * sizeof(T)) != hipSuccess) {
      p = nullptr;
      return -1;
    }
    owned = true;
    n = count;
    return 0;
  }
  int upload(const T *src, size_t count, hipStream_t s) {
    if (alloc(count)) return -1;
    if (count && hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    return 0;
  }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = false;
  }
};

struct DeviceModel {
  DBuf<SpeciesDev> species;
  DBuf<float> W;
  DBuf<double> avg, scale;
  DBuf<BPDev> bp;
  DBuf<BPFunc> bpf;
  DBuf<BPAngK> bpk;
  DBuf<BPAngGrid> bpg;
  int ntypes = 0, desc = 0, max_rN = 0, max_aN = 0, nsf_max = 0, wsize = 0, nl_max = 0, hmax = 1, nnorm = 0, net_uniform = 0;
  double rc_max = 0.0;
  AtomSizes sizes(int natoms, int cap, const int *active = nullptr, const int *nactive = nullptr) const {
    return AtomSizes{natoms, cap, nsf_max, wsize, nl_max, hmax, active, nactive, 0, -1, nnorm, net_uniform, -1};
  }
  int upload(const PotSet &ps, hipStream_t s);
  ModelDev view() const;
};

// neighbor-count bucket caps for the neighbor counts of a batch's atoms
// (the last cap is nb_cap; one cap: no buckets)
std::vector<int> bucket_caps_for(const std::vector<int> &counts, int nb_cap);

// a fixed set of structures (shapes, lattices, atom counts); coordinates
// and types may change between evaluations
class Batch {
 public:
  int create(const PotSet &ps, int nstruct, const int *natoms, const double *latt, const int *pbc,
             const double *coords, std::string &err);
  // one structure given in lattice coordinates of latt (aenet_nbl_init)
  int create_lattice_coords(const PotSet &ps, int natoms, const double *latt, int pbc, std::string &err);
  // new bounding boxes of the isolated structures from coordinates (host,
  // Cartesian), as geo_get_bounds takes them; periodic structures keep
  // their plans.  The measured launch shapes are kept.
  int replan(const double *coords, std::string &err);
  bool has_isolated() const;
  const double *wrapped_coords() const { return s_.p; }
  // all pointers are device pointers; E_struct receives 2 per structure
  // (total, cohesive); F may be null when forces are not wanted
  int eval(const DeviceModel &dm, const double *d_coords, const int *d_types, double *d_E_atom,
           double *d_E_struct, double *d_F, hipStream_t stream);
  // atomic energies of the atoms d_active[0 .. *d_nactive-1] only (no
  // forces, no structure sums); max_active bounds *d_nactive.  With
  // reuse_lists the neighbor lists of the previous evaluation of these
  // coordinates are used as they are (the atoms must have been searched).
  int eval_active(const DeviceModel &dm, const double *d_coords, const int *d_types, const int *d_active,
                  const int *d_nactive, int max_active, bool reuse_lists, double *d_E_atom, hipStream_t stream);
  // vector-Jacobian product of the atomic energies: d_gR (3 per atom) =
  // sum_i g_i dE_i/dR (overwritten); d_gW (wsize floats, may be null) +=
  // sum_i g_i dE_i/dW
  int vjp(const DeviceModel &dm, const double *d_coords, const int *d_types, const double *d_g, double *d_gR,
          float *d_gW, hipStream_t stream);
  // the model's cutoff must not exceed the one the batch was planned for
  bool fits(const DeviceModel &dm) const { return dm.rc_max <= rc_max_ * (1.0 + 1e-12); }
  // bit-exact neighbor lists of every atom
  int neighbor_lists(const double *d_coords, const int *d_types, int cap, int *d_count, double *d_coo,
                     double *d_dist, int *d_list, int *d_type, hipStream_t stream);
  // descriptor values of every atom (FP64, stride nsf_max)
  int descriptors(const DeviceModel &dm, const double *d_coords, const int *d_types, int scaled,
                  double *d_G, int stride, hipStream_t stream);
  int check_flags(hipStream_t stream, std::string &err);
  // measured launch shape of the atom kernel for model dm: every feasible
  // (wavefronts per workgroup, weight staging) is timed over `repeats`
  // full evaluations (HIP events around the atom kernel) and the fastest
  // kept for the batch's later evaluations with dm; out receives
  // (waves, staged, average ms) per candidate
  int autotune(const DeviceModel &dm, const double *d_coords, const int *d_types, bool forces, int repeats,
               hipStream_t stream, std::vector<std::array<double, 4>> &out, std::string &err);
  // a launch shape given by the caller (as one the tuner keeps); waves = 0
  // returns to the occupancy model.  Drops a bucket plan of dm.
  int set_launch(const DeviceModel &dm, int waves, int staged, int tile, std::string &err);
  // Neighbor-count buckets of the atom kernel for dm: n caps (ascending,
  // the last one at least the batch's neighbor capacity, which it becomes)
  // and a launch shape (waves, staged, tile) per bucket; n = 0 drops the
  // plan.  The tuner makes one where it measures faster than one launch.
  int set_buckets(const DeviceModel &dm, int n, const int *caps, const int *shapes, std::string &err);
  // the plan in use for dm: n (0: none), caps, shapes (3 per bucket) and,
  // when the tuner made it, each bucket's measured atom-kernel time and
  // the totals with and without buckets (ms; 0: not measured)
  struct BucketPlan {
    int n = 0;
    int cap[MAX_BUCKETS] = {};
    std::array<int, 3> shape[MAX_BUCKETS] = {};
    double ms[MAX_BUCKETS] = {};
    int atoms[MAX_BUCKETS] = {};  // atoms per bucket when measured (-1: unknown)
    double ms_single = 0.0, ms_buckets = 0.0;
  };
  void buckets(const DeviceModel &dm, BucketPlan &out) const;
  // HIP events around the atom kernel of every evaluation while enabled
  void set_timing(bool on);
  // summed duration (ms) and count of the timed atom-kernel launches since
  // the last call; synchronizes on the recorded events
  int kernel_time(double &ms, int &count);
  // the same for the neighbor-search launches of full evaluations
  int search_time(double &ms, int &count);
  ~Batch();
  int natoms() const { return plan_.natoms; }
  int nb_cap() const { return nb_cap_; }
  // the batch's compact neighbor lists and error word, for callers that
  // fill the lists themselves (eval_active with reuse_lists)
  NbDev nb_lists() const { return nb_view(); }
  int *flags_ptr() const { return flags_.p; }
  int set_nb_cap(int cap);
  int nstruct() const { return (int)plan_.st.size(); }
  const PlanHost &plan() const { return plan_; }

 private:
  BatchDev view(const double *d_coords, const int *d_types) const;
  // atom-kernel sizes of dm with the batch's measured launch shape
  AtomSizes sizes_for(const DeviceModel &dm, int natoms, const int *active = nullptr,
                      const int *nactive = nullptr) const;
  std::map<const DeviceModel *, std::array<int, 3>> tune_;  // model -> (waves, staged, tile network)
  std::map<const DeviceModel *, BucketPlan> buckets_;     // model -> neighbor-count buckets
  DBuf<int> bucket_lists_, bucket_counts_;                // [MAX_BUCKETS * natoms], [MAX_BUCKETS]
  int timed_bucket_ = -1;                                 // events around one bucket's launch only (tuner)
  // the atom kernel(s) of a full evaluation: one launch, or the bucket
  // launches of dm's plan
  void launch_atoms(const DeviceModel &dm, bool forces, const ModelDev &md, const BatchDev &b, const NbDev &nb,
                    const OutDev &o, hipStream_t stream);
  int tune_buckets(const DeviceModel &dm, const double *d_coords, const int *d_types, bool forces, int repeats,
                   double ms_single, hipStream_t stream, std::string &err);
  int allocate(std::string &err);
  PlanHost plan_;
  double rc_max_ = 0.0;
  bool lattice_coords_ = false;
  std::vector<double> occ_;  // occupied volume fraction of each structure (density estimate)
  DBuf<StructDev> st_;
  DBuf<int> atom_struct_, cell_struct_, cell_ijk_, Tvec_, Cvec_, self_t_;
  DBuf<double> self_d_;
  DBuf<double> s_;
  DBuf<int> cellof_, cell_count_, cell_fill_, cell_start_, cell_atoms_, cell_tmp_, combo_count_, combos_, flags_;
  DBuf<int> cell_work_, fmark_, fcnt_;
  DBuf<unsigned long long> fslot_;
  int fslots_ = 0;
  // triplet lists of the BP kernel (batches made for BP potentials)
  bool bp_ = false;
  DBuf<int> trip_;
  int trip_slots_ = 0;
  DBuf<double> cell_ctr_, sorted_s_;
  int combo_cap_ = 0;
  int kept_cap_ = AENET_AMD_KEPT_CAP;
  int combo_max_ = 1;
  int work_cap_ = 0;
  // compact neighbor lists
  int nb_cap_ = 0;
  DBuf<int> nb_count_, nb_j_;
  DBuf<float> nb_R_;
  NbDev nb_view() const;
  bool timing_ = false;
  std::vector<hipEvent_t> ev_, sev_;
  DBuf<double> scratch_E_;
  DBuf<double> scratch_G_;
};

// explicit environments (per-atom entry points).  Inputs and outputs of a
// call travel in one packed block each way between pinned host memory and
// one device buffer: [start | ctype | tj | ci | cj | flags | E | dFi | dFj]
// (the call's latency is mostly copies, so one of each).
struct EnvScratch {
  DBuf<char> dev;
  char *pinned = nullptr;
  size_t bytes = 0;
  ~EnvScratch();
  // coef_unit > 0: environment g differentiates descriptor feature
  // g mod coef_unit instead of the energy (h_dFi, h_dFj = -dG/dR)
  int eval(const DeviceModel &dm, int nenv, const int *h_start, const int *h_ctype, const double *h_ci,
           const double *h_cj, const int *h_tj, bool forces, double *h_E, double *h_dFi, double *h_dFj,
           double *h_G, int G_stride, hipStream_t stream, std::string &err, int coef_unit = 0);
};

// Body of aenet_structures_energy_and_forces (capi.cpp).  With an arena,
// every device buffer of the call (inputs, outputs, plan, cell lists,
// neighbor lists) comes from it; a request beyond it fails with
// AENET_STATUS_ERR_CAPACITY.
void structures_energy_and_forces(int nstruct, const int natoms[], const double latt[], const int pbc[],
                                  const int types[], const double coords[], double E_tot[], double E_coh[],
                                  double E_atoms[], double F[], ScratchArena *arena, const char *who, int *stat);
// records the calling thread's last error (aenet_amd_last_error); returns st
int set_last_error(int st, const std::string &msg);

}  // namespace aenet_amd

// engine.cpp -- device-resident potential sets and structure batches.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "aenet_capi_status.h"
#include "kernels.h"

namespace aenet_amd {

ScratchArena *&current_arena() {
  static thread_local ScratchArena *a = nullptr;
  return a;
}

int DeviceModel::upload(const PotSet &ps, hipStream_t s) {
  ntypes = ps.ntypes;
  desc = ps.desc;
  max_rN = ps.max_rN;
  max_aN = ps.max_aN;
  nsf_max = ps.nsf_max;
  wsize = (int)ps.W.size();
  nnorm = (int)ps.norm_avg.size();
  rc_max = ps.Rc_max;
  nl_max = 0;
  for (const SpeciesDev &sd : ps.species) {
    nl_max = std::max(nl_max, sd.nlayers);
    for (int l = 1; l < sd.nlayers; ++l) hmax = std::max(hmax, sd.nnodes[l]);
  }
  // the matrix-tile network: one architecture, hidden layers up to 64 nodes
  net_uniform = !ps.species.empty() && hmax <= 64;
  for (const SpeciesDev &sd : ps.species) {
    const SpeciesDev &s0 = ps.species[0];
    bool same = sd.nlayers == s0.nlayers && sd.nlayers >= 2 && sd.nnodes[sd.nlayers - 1] == 1;
    for (int l = 0; same && l < sd.nlayers; ++l) same = sd.nnodes[l] == s0.nnodes[l];
    if (!same) net_uniform = 0;
  }
  if (species.upload(ps.species.data(), ps.species.size(), s)) return AENET_STATUS_ERR_DEVICE;
  // 32 words of padding: the first layer's products read up to 31 words
  // past a row (first_layer_from_lanes); the copy lives until the stream
  // synchronization below
  std::vector<float> wpad(ps.W);
  wpad.resize(ps.W.size() + 32, 0.0f);
  if (W.upload(wpad.data(), wpad.size(), s)) return AENET_STATUS_ERR_DEVICE;
  if (avg.upload(ps.norm_avg.data(), ps.norm_avg.size(), s)) return AENET_STATUS_ERR_DEVICE;
  if (scale.upload(ps.norm_scale.data(), ps.norm_scale.size(), s)) return AENET_STATUS_ERR_DEVICE;
  if (bp.upload(&ps.bp, 1, s)) return AENET_STATUS_ERR_DEVICE;
  if (bpf.upload(ps.bpf.data(), ps.bpf.size(), s)) return AENET_STATUS_ERR_DEVICE;
  if (bpk.upload(ps.bpk.data(), ps.bpk.size(), s)) return AENET_STATUS_ERR_DEVICE;
  if (bpg.upload(ps.bpg.data(), ps.bpg.size(), s)) return AENET_STATUS_ERR_DEVICE;
  if (hipStreamSynchronize(s) != hipSuccess) return AENET_STATUS_ERR_DEVICE;
  return AENET_STATUS_OK;
}

ModelDev DeviceModel::view() const {
  ModelDev m;
  m.species = species.p;
  m.W = W.p;
  m.norm_avg = avg.p;
  m.norm_scale = scale.p;
  m.bp = bp.p;
  m.bpf = bpf.p;
  m.bpk = bpk.p;
  m.bpg = bpg.p;
  m.ntypes = ntypes;
  m.desc = desc;
  return m;
}

// fraction of a periodic cell's volume the atoms occupy: per lattice
// direction one minus the widest empty gap between the atoms' fractional
// coordinates (a slab's vacuum, wherever it sits in the cell), counted
// only when the gap is wider than the cutoff (the spacing of atomic
// planes is not vacuum)
static double occupied_fraction(const double *L, int n, const double *cart, double rc) {
  // L: column-major lattice (column k = vector k); fractional = L^-1 x
  const double det = L[0] * (L[4] * L[8] - L[7] * L[5]) - L[3] * (L[1] * L[8] - L[7] * L[2]) +
                     L[6] * (L[1] * L[5] - L[4] * L[2]);
  if (n < 2 || !(std::fabs(det) > 1e-300)) return 1.0;
  double inv[9];
  inv[0] = (L[4] * L[8] - L[7] * L[5]) / det;
  inv[3] = (L[6] * L[5] - L[3] * L[8]) / det;
  inv[6] = (L[3] * L[7] - L[6] * L[4]) / det;
  inv[1] = (L[7] * L[2] - L[1] * L[8]) / det;
  inv[4] = (L[0] * L[8] - L[6] * L[2]) / det;
  inv[7] = (L[6] * L[1] - L[0] * L[7]) / det;
  inv[2] = (L[1] * L[5] - L[4] * L[2]) / det;
  inv[5] = (L[3] * L[2] - L[0] * L[5]) / det;
  inv[8] = (L[0] * L[4] - L[3] * L[1]) / det;
  double occ = 1.0;
  std::vector<double> f(n);
  for (int k = 0; k < 3; ++k) {
    for (int i = 0; i < n; ++i) {
      double v = inv[k] * cart[3 * i] + inv[3 + k] * cart[3 * i + 1] + inv[6 + k] * cart[3 * i + 2];
      f[i] = v - std::floor(v);
    }
    std::sort(f.begin(), f.end());
    double gap = 1.0 - (f[n - 1] - f[0]);
    for (int i = 1; i < n; ++i) gap = std::max(gap, f[i] - f[i - 1]);
    // width of the cell along the normal of the other two vectors: |det| / |a_i x a_j|
    const double *a = L + 3 * ((k + 1) % 3), *b = L + 3 * ((k + 2) % 3);
    const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    const double width = std::fabs(det) / std::sqrt(cx * cx + cy * cy + cz * cz);
    if (gap * width > rc) occ *= std::max(0.05, 1.0 - gap);
  }
  return occ;
}

int Batch::create(const PotSet &ps, int nstruct, const int *natoms, const double *latt, const int *pbc,
                  const double *coords, std::string &err) {
  rc_max_ = ps.Rc_max;
  bp_ = ps.desc != DESC_CHEBYSHEV;
  int st = build_plan(ps.Rc_max, nstruct, natoms, latt, pbc, coords, plan_, err);
  if (st) return st;
  occ_.assign(nstruct, 1.0);
  for (int s = 0, a0 = 0; s < nstruct; a0 += natoms[s], ++s)
    if (pbc[s] && coords) occ_[s] = occupied_fraction(plan_.st[s].L, natoms[s], coords + 3 * (size_t)a0, ps.Rc_max);
  return allocate(err);
}

bool Batch::has_isolated() const {
  if (lattice_coords_) return false;
  for (const StructDev &d : plan_.st)
    if (!d.pbc) return true;
  return false;
}

int Batch::replan(const double *coords, std::string &err) {
  if (!has_isolated()) return AENET_STATUS_OK;
  const int ns = (int)plan_.st.size();
  std::vector<int> natoms(ns), pbc(ns);
  std::vector<double> latt(9 * (size_t)ns);
  for (int s = 0; s < ns; ++s) {
    natoms[s] = plan_.st[s].natoms;
    pbc[s] = plan_.st[s].pbc;
    std::memcpy(&latt[9 * (size_t)s], plan_.st[s].L, 9 * sizeof(double));
  }
  for (size_t i = 0; i < 3 * (size_t)plan_.natoms; ++i)
    if (!std::isfinite(coords[i])) {
      err = "non-finite coordinates";
      return AENET_STATUS_ERR_COORD;
    }
  PlanHost np;
  int st = build_plan(rc_max_, ns, natoms.data(), latt.data(), pbc.data(), coords, np, err);
  if (st) return st;
  const int cap = nb_cap_;
  plan_ = std::move(np);
  nb_cap_ = 0;
  st = allocate(err);
  if (st) return st;
  // keep a list capacity that an earlier evaluation needed
  return cap > nb_cap_ ? set_nb_cap(cap) : AENET_STATUS_OK;
}

int Batch::create_lattice_coords(const PotSet &ps, int natoms, const double *latt, int pbc, std::string &err) {
  rc_max_ = ps.Rc_max;
  bp_ = ps.desc != DESC_CHEBYSHEV;
  lattice_coords_ = true;
  int st = build_plan_lattice(ps.Rc_max, natoms, latt, pbc, plan_, err);
  if (st) return st;
  return allocate(err);
}

int Batch::allocate(std::string &err) {
  const PlanHost &p = plan_;
  combo_cap_ = 1;
  for (auto &d : p.st) combo_cap_ = std::max(combo_cap_, std::min((d.nC + 1) * (d.nT + 1), AENET_AMD_COMBO_CAP));
  hipStream_t s = nullptr;
  bool bad = false;
  bad |= st_.upload(p.st.data(), p.st.size(), s) != 0;
  bad |= atom_struct_.upload(p.atom_struct.data(), p.atom_struct.size(), s) != 0;
  bad |= cell_struct_.upload(p.cell_struct.data(), p.cell_struct.size(), s) != 0;
  bad |= cell_ijk_.upload(p.cell_ijk.data(), p.cell_ijk.size(), s) != 0;
  bad |= Tvec_.upload(p.Tvec.data(), p.Tvec.size(), s) != 0;
  bad |= Cvec_.upload(p.Cvec.data(), p.Cvec.size(), s) != 0;
  bad |= self_t_.upload(p.self_t.data(), p.self_t.size(), s) != 0;
  bad |= self_d_.upload(p.self_d.data(), p.self_d.size(), s) != 0;
  bad |= cell_ctr_.upload(p.cell_ctr.data(), p.cell_ctr.size(), s) != 0;
  bad |= s_.alloc(3 * (size_t)p.natoms) != 0;
  bad |= sorted_s_.alloc(3 * (size_t)p.natoms) != 0;
  bad |= cellof_.alloc(p.natoms) != 0;
  bad |= cell_count_.alloc(p.ncells) != 0;
  bad |= cell_fill_.alloc(p.ncells) != 0;
  bad |= cell_start_.alloc(p.ncells + 1) != 0;
  bad |= cell_atoms_.alloc(p.natoms) != 0;
  bad |= cell_tmp_.alloc(p.natoms) != 0;
  bad |= combo_count_.alloc(p.ncells) != 0;
  bad |= combos_.alloc(2 * (size_t)p.ncells * combo_cap_) != 0;
  bad |= flags_.alloc(1) != 0;
  work_cap_ = p.natoms / CS_WAVES + p.ncells + 1;
  bad |= cell_work_.alloc(4 * (size_t)work_cap_) != 0;
  bad |= scratch_E_.alloc(p.natoms) != 0;
  bad |= fmark_.alloc(p.natoms) != 0;
  // slots for the atom kernels' workgroup force tables: used only when one
  // round of wavefronts covers the batch (a few thousand atoms)
  fslots_ = p.natoms <= FORCE_SLOT_MAX_ATOMS ? FORCE_SLOTS : 0;
  bad |= fcnt_.alloc(fslots_ ? p.natoms : 1) != 0;
  bad |= fslot_.alloc(fslots_ ? 3 * (size_t)p.natoms * fslots_ : 1) != 0;
  if (bp_ && !current_arena()) {
    // one list per wavefront the BP kernel can hold resident at once (a
    // persistent grid has at most a wavefront per atom, rounded up to a
    // workgroup); none in a fixed scratch arena, where the force pass takes
    // the candidate pass again
    trip_slots_ = std::min(p.natoms + 64, device_limits().cus * BP_TRIP_WAVES_PER_CU);
    bad |= trip_.alloc((size_t)trip_slots_ * BP_TRIP_CAP) != 0;
  }
  if (bad) {
    err = "device allocation failed";
    return AENET_STATUS_ERR_MALLOC;
  }
  // neighbor capacity from the densest structure: atoms within the list
  // radius at that density, with headroom
  double est = 0.0;
  for (size_t si = 0; si < p.st.size(); ++si) {
    const StructDev &d = p.st[si];
    double V = std::fabs(d.L[0] * (d.L[4] * d.L[8] - d.L[5] * d.L[7]) - d.L[3] * (d.L[1] * d.L[8] - d.L[2] * d.L[7]) +
                         d.L[6] * (d.L[1] * d.L[5] - d.L[2] * d.L[4]));
    if (si < occ_.size()) V *= occ_[si];
    double rho = V > 0.0 ? d.natoms / V : 0.0;
    double e = rho * 4.18879020478639 * d.rc_list * d.rc_list * d.rc_list + d.nself;
    // an isolated structure has no images: at most natoms - 1 neighbors
    // (its bounding box overstates the density of a small cluster)
    if (!d.pbc) e = std::min(e, (double)(d.natoms - 1));
    est = std::max(est, e);
  }
  int cap = (int)(1.2 * est) + 16;
  cap = std::max(64, std::min(AENET_AMD_NB_CAP, (cap + 15) / 16 * 16));
  int st = set_nb_cap(cap);
  if (st) {
    err = "device allocation failed";
    return st;
  }
  // hipMalloc does not clear memory: the error word and the cell counters
  // must start at zero (each step leaves the counters zeroed again)
  launch_fill_i32(flags_.p, 0, 1, s);
  launch_fill_i32(fmark_.p, 0, p.natoms, s);
  if (fslots_) launch_fill_i32(fcnt_.p, 0, p.natoms, s);
  launch_fill_i32(cell_count_.p, 0, p.ncells, s);
  launch_fill_i32(cell_fill_.p, 0, p.ncells, s);
  // the (cell, image) combinations depend on the lattice only
  launch_combos(view(nullptr, nullptr), combo_count_.p, combos_.p, s);
  {
    // the search's LDS holds the largest combination count of a home cell
    std::vector<int> cc(p.ncells);
    if (hipMemcpy(cc.data(), combo_count_.p, p.ncells * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      err = "device error";
      return AENET_STATUS_ERR_DEVICE;
    }
    int mx = 1;
    for (int v : cc) mx = std::max(mx, v);
    kept_cap_ = std::min(AENET_AMD_KEPT_CAP, (mx + 15) / 16 * 16);
    combo_max_ = mx;
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    err = "device error";
    return AENET_STATUS_ERR_DEVICE;
  }
  return check_flags(s, err);
}

int Batch::set_nb_cap(int cap) {
  if (cap < 1) cap = 1;
  if (cap > AENET_AMD_NB_CAP) cap = AENET_AMD_NB_CAP;
  if (cap == nb_cap_ && nb_R_.p) return AENET_STATUS_OK;
  size_t n = (size_t)plan_.natoms;
  if (nb_count_.alloc(n) || nb_j_.alloc(n * cap) || nb_R_.alloc(4 * n * cap)) return AENET_STATUS_ERR_MALLOC;
  nb_cap_ = cap;
  return AENET_STATUS_OK;
}

NbDev Batch::nb_view() const {
  NbDev nb;
  nb.cap = nb_cap_;
  nb.count = nb_count_.p;
  nb.R = nb_R_.p;
  nb.j = nb_j_.p;
  return nb;
}

BatchDev Batch::view(const double *d_coords, const int *d_types) const {
  BatchDev b;
  std::memset(&b, 0, sizeof(b));
  b.st = st_.p;
  b.nstruct = (int)plan_.st.size();
  b.natoms = plan_.natoms;
  b.atom_struct = atom_struct_.p;
  b.cart = d_coords;
  b.types = d_types;
  b.s = s_.p;
  b.cellof = cellof_.p;
  b.ncells = plan_.ncells;
  b.cell_struct = cell_struct_.p;
  b.cell_ijk = cell_ijk_.p;
  b.cell_count = cell_count_.p;
  b.cell_fill = cell_fill_.p;
  b.cell_start = cell_start_.p;
  b.cell_atoms = cell_atoms_.p;
  b.cell_ctr = cell_ctr_.p;
  b.sorted_s = sorted_s_.p;
  b.combo_count = combo_count_.p;
  b.combos = combos_.p;
  b.combo_cap = combo_cap_;
  b.kept_cap = kept_cap_;
  b.combo_max = combo_max_;
  b.cell_work = cell_work_.p;
  b.work_cap = work_cap_;
  b.Tvec = Tvec_.p;
  b.Cvec = Cvec_.p;
  b.self_t = self_t_.p;
  b.self_d = self_d_.p;
  b.flags = flags_.p;
  b.fmark = fmark_.p;
  b.fcnt = fcnt_.p;
  b.fslot = fslot_.p;
  b.fslots = fslots_;
  b.trip = bp_ ? trip_.p : nullptr;
  b.trip_cap = BP_TRIP_CAP;
  b.trip_slots = bp_ ? trip_slots_ : 0;
  return b;
}

int Batch::eval(const DeviceModel &dm, const double *d_coords, const int *d_types, double *d_E_atom,
                double *d_E_struct, double *d_F, hipStream_t stream) {
  BatchDev b = view(d_coords, d_types);
  ModelDev md = dm.view();
  NbDev nb = nb_view();
  launch_binning(b, cell_tmp_.p, d_F, stream);
  {
    hipEvent_t s0 = nullptr, s1 = nullptr;
    if (timing_ && hipEventCreate(&s0) == hipSuccess && hipEventCreate(&s1) == hipSuccess)
      (void)hipEventRecord(s0, stream);
    launch_search(b, nb, stream);
    if (s0 && s1) {
      (void)hipEventRecord(s1, stream);
      sev_.push_back(s0);
      sev_.push_back(s1);
    }
  }
  OutDev o;
  std::memset(&o, 0, sizeof(o));
  o.E_atom = d_E_atom ? d_E_atom : scratch_E_.p;
  o.F = d_F;
  launch_atoms(dm, d_F != nullptr, md, b, nb, o, stream);
  // E_struct = [E_tot(0..ns-1), E_coh(0..ns-1)]; the forces into FP64
  launch_struct_energy(b, md, o.E_atom, d_E_struct, d_E_struct ? d_E_struct + b.nstruct : nullptr, d_F, 1.0,
                       stream);
  return hipGetLastError() == hipSuccess ? AENET_STATUS_OK : AENET_STATUS_ERR_DEVICE;
}

int Batch::vjp(const DeviceModel &dm, const double *d_coords, const int *d_types, const double *d_g, double *d_gR,
               float *d_gW, hipStream_t stream) {
  BatchDev b = view(d_coords, d_types);
  ModelDev md = dm.view();
  NbDev nb = nb_view();
  launch_binning(b, cell_tmp_.p, d_gR, stream);
  launch_search(b, nb, stream);
  OutDev o;
  std::memset(&o, 0, sizeof(o));
  o.E_atom = scratch_E_.p;
  o.F = d_gR;
  o.g_atom = d_g;
  EnvDev e;
  std::memset(&e, 0, sizeof(e));
  const AtomSizes z = sizes_for(dm, b.natoms);
  if (dm.desc == DESC_CHEBYSHEV)
    launch_atom_cheb(dm.max_rN, dm.max_aN, 0, true, z, md, b, nb, o, e, stream);
  else
    launch_atom_bp(0, true, z, md, b, nb, o, e, stream);
  // the kernels accumulate forces, -gradient
  launch_struct_energy(b, md, nullptr, nullptr, nullptr, d_gR, -1.0, stream);
  if (d_gW) {
    const size_t need = (size_t)b.natoms * dm.nsf_max;
    if (scratch_G_.n < need && scratch_G_.alloc(need)) return AENET_STATUS_ERR_MALLOC;
    OutDev od;
    std::memset(&od, 0, sizeof(od));
    od.E_atom = scratch_E_.p;
    od.G = scratch_G_.p;
    od.G_scaled = 1;
    od.G_stride = dm.nsf_max;
    od.desc_only = 1;
    if (dm.desc == DESC_CHEBYSHEV)
      launch_atom_cheb(dm.max_rN, dm.max_aN, 0, false, z, md, b, nb, od, e, stream);
    else
      launch_atom_bp(0, false, z, md, b, nb, od, e, stream);
    NetGradArgs a;
    std::memset(&a, 0, sizeof(a));
    a.G = scratch_G_.p;
    a.stride = dm.nsf_max;
    a.types = d_types;
    a.upstream = d_g;
    a.grad = d_gW;
    a.flags = flags_.p;
    launch_net_grad(md, a, b.natoms, dm.nsf_max, dm.nl_max, dm.wsize, true, stream, dm.hmax);
  }
  return hipGetLastError() == hipSuccess ? AENET_STATUS_OK : AENET_STATUS_ERR_DEVICE;
}

int Batch::eval_active(const DeviceModel &dm, const double *d_coords, const int *d_types, const int *d_active,
                       const int *d_nactive, int max_active, bool reuse_lists, double *d_E_atom,
                       hipStream_t stream) {
  if (max_active <= 0) return AENET_STATUS_OK;
  BatchDev b = view(d_coords, d_types);
  ModelDev md = dm.view();
  NbDev nb = nb_view();
  if (!reuse_lists) {
    launch_binning(b, cell_tmp_.p, nullptr, stream);
    launch_search(b, nb, stream, d_active, d_nactive, max_active);
  }
  OutDev o;
  std::memset(&o, 0, sizeof(o));
  o.E_atom = d_E_atom;
  EnvDev e;
  std::memset(&e, 0, sizeof(e));
  const AtomSizes z = sizes_for(dm, max_active, d_active, d_nactive);
  if (dm.desc == DESC_CHEBYSHEV)
    launch_atom_cheb(dm.max_rN, dm.max_aN, 0, false, z, md, b, nb, o, e, stream);
  else
    launch_atom_bp(0, false, z, md, b, nb, o, e, stream);
  return hipGetLastError() == hipSuccess ? AENET_STATUS_OK : AENET_STATUS_ERR_DEVICE;
}

int Batch::neighbor_lists(const double *d_coords, const int *d_types, int cap, int *d_count, double *d_coo,
                          double *d_dist, int *d_list, int *d_type, hipStream_t stream) {
  BatchDev b = view(d_coords, d_types);
  launch_binning(b, cell_tmp_.p, nullptr, stream);
  OutDev o;
  std::memset(&o, 0, sizeof(o));
  o.nb_cap = cap;
  o.nb_count = d_count;
  o.nb_coo = d_coo;
  o.nb_dist = d_dist;
  o.nb_list = d_list;
  o.nb_type = d_type;
  launch_neighbor_lists(b, o, stream);
  return hipGetLastError() == hipSuccess ? AENET_STATUS_OK : AENET_STATUS_ERR_DEVICE;
}

int Batch::descriptors(const DeviceModel &dm, const double *d_coords, const int *d_types, int scaled,
                       double *d_G, int stride, hipStream_t stream) {
  BatchDev b = view(d_coords, d_types);
  ModelDev md = dm.view();
  NbDev nb = nb_view();
  launch_binning(b, cell_tmp_.p, nullptr, stream);
  launch_search(b, nb, stream);
  OutDev o;
  std::memset(&o, 0, sizeof(o));
  o.E_atom = scratch_E_.p;
  o.G = d_G;
  o.G_scaled = scaled;
  o.G_stride = stride;
  o.desc_only = 1;
  EnvDev e;
  std::memset(&e, 0, sizeof(e));
  if (dm.desc == DESC_CHEBYSHEV)
    launch_atom_cheb(dm.max_rN, dm.max_aN, 0, false, sizes_for(dm, b.natoms), md, b, nb, o, e, stream);
  else
    launch_atom_bp(0, false, sizes_for(dm, b.natoms), md, b, nb, o, e, stream);
  return hipGetLastError() == hipSuccess ? AENET_STATUS_OK : AENET_STATUS_ERR_DEVICE;
}

// events around [begin, end) of a launch sequence while timing
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  void begin(bool on, hipStream_t s) {
    if (on && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess)
      (void)hipEventRecord(e0, s);
  }
  void end(hipStream_t s, std::vector<hipEvent_t> &ev) {
    if (e0 && e1) {
      (void)hipEventRecord(e1, s);
      ev.push_back(e0);
      ev.push_back(e1);
    }
    e0 = e1 = nullptr;
  }
};

void Batch::launch_atoms(const DeviceModel &dm, bool forces, const ModelDev &md, const BatchDev &b, const NbDev &nb,
                         const OutDev &o, hipStream_t stream) {
  EnvDev e;
  std::memset(&e, 0, sizeof(e));
  const auto it = buckets_.find(&dm);
  if (it == buckets_.end() || it->second.n < 2) {
    EventPair t;
    t.begin(timing_, stream);
    if (dm.desc == DESC_CHEBYSHEV)
      launch_atom_cheb(dm.max_rN, dm.max_aN, 0, forces, sizes_for(dm, b.natoms), md, b, nb, o, e, stream);
    else
      launch_atom_bp(0, forces, sizes_for(dm, b.natoms), md, b, nb, o, e, stream);
    t.end(stream, ev_);
    return;
  }
  // one launch per neighbor-count bucket, each over the device list of its
  // atoms, with LDS regions of the bucket's capacity and its own shape
  const BucketPlan &pl = it->second;
  BucketCaps bc;
  bc.n = pl.n;
  for (int k = 0; k < pl.n; ++k) bc.cap[k] = std::min(pl.cap[k], nb_cap_);
  EventPair all;
  all.begin(timing_ && timed_bucket_ < 0, stream);
  launch_bucket_atoms(nb.count, b.natoms, bc, bucket_lists_.p, bucket_counts_.p, stream);
  for (int k = 0; k < pl.n; ++k) {
    // the persistent grid sized for the atoms the bucket held when measured,
    // with room (any count is evaluated, in more rounds)
    const int expect = pl.atoms[k] < 0 ? b.natoms : std::min(b.natoms, pl.atoms[k] + pl.atoms[k] / 4 + 256);
    AtomSizes z = dm.sizes(expect, bc.cap[k], bucket_lists_.p + (size_t)k * b.natoms, bucket_counts_.p + k);
    z.tune_waves = pl.shape[k][0];
    z.tune_staged = pl.shape[k][1];
    z.tune_tile = pl.shape[k][2];
    EventPair one;
    one.begin(timing_ && timed_bucket_ == k, stream);
    if (dm.desc == DESC_CHEBYSHEV)
      launch_atom_cheb(dm.max_rN, dm.max_aN, 0, forces, z, md, b, nb, o, e, stream);
    else
      launch_atom_bp(0, forces, z, md, b, nb, o, e, stream);
    one.end(stream, ev_);
  }
  all.end(stream, ev_);
}

// Bucket caps for a histogram of neighbor counts: powers of two from 16
// below the capacity, then the capacity; a bucket with fewer than
// max(natoms / 32, 2048) atoms joins the next larger one (the last bucket
// always stays).  One cap: no buckets.
std::vector<int> bucket_caps_for(const std::vector<int> &counts, int nb_cap) {
  if (nb_cap < 1) return {};
  std::vector<int> cand;
  for (int c = 16; c < nb_cap && (int)cand.size() < MAX_BUCKETS - 1; c *= 2) cand.push_back(c);
  cand.push_back(nb_cap);
  std::vector<long> in(cand.size(), 0);
  for (int n : counts) {
    size_t k = 0;
    while (k + 1 < cand.size() && n > cand[k]) ++k;
    ++in[k];
  }
  const long minimum = std::max<long>((long)counts.size() / 32, 2048);
  std::vector<int> caps;
  long carry = 0;
  for (size_t k = 0; k < cand.size(); ++k) {
    carry += in[k];
    if (k + 1 == cand.size() || carry >= minimum) {
      caps.push_back(cand[k]);
      carry = 0;
    }
  }
  return caps;
}

int Batch::set_buckets(const DeviceModel &dm, int n, const int *caps, const int *shapes, std::string &err) {
  if (n == 0) {
    buckets_.erase(&dm);
    return AENET_STATUS_OK;
  }
  if (n < 0 || n > MAX_BUCKETS || !caps || !shapes) {
    err = "bucket count out of range";
    return AENET_AMD_ERR_ARG_VALUE;
  }
  const bool cheb = dm.desc == DESC_CHEBYSHEV;
  const int max_waves = cheb ? CHEB_MAX_WAVES : ATOM_MAX_WAVES;
  BucketPlan pl;
  pl.n = n;
  for (int k = 0; k < n; ++k) pl.atoms[k] = -1;
  for (int k = 0; k < n; ++k) {
    if (caps[k] < 1 || (k > 0 && caps[k] <= caps[k - 1]) || (k + 1 < n && caps[k] >= nb_cap_)) {
      err = "bucket caps must ascend and stay below the batch's neighbor capacity but for the last";
      return AENET_AMD_ERR_ARG_VALUE;
    }
    pl.cap[k] = k + 1 < n ? caps[k] : nb_cap_;
    const int waves = shapes[3 * k], staged = shapes[3 * k + 1], tile = shapes[3 * k + 2];
    if (waves < 1 || waves > max_waves || (waves & (waves - 1)) || staged < 0 || staged > 1 || tile < 0 ||
        tile > 1 || (tile && !(cheb && dm.net_uniform))) {
      err = "bucket launch shape out of range";
      return AENET_AMD_ERR_ARG_VALUE;
    }
    AtomSizes z = dm.sizes(plan_.natoms, pl.cap[k]);
    z.tune_waves = waves;
    z.tune_staged = staged;
    z.tune_tile = tile;
    if (!(cheb ? atom_cheb_shape_ok(z) : atom_bp_shape_ok(z))) {
      err = "bucket launch shape does not fit the device";
      return AENET_AMD_ERR_ARG_VALUE;
    }
    pl.shape[k] = {waves, staged, tile};
  }
  if (bucket_lists_.n < (size_t)MAX_BUCKETS * plan_.natoms &&
      bucket_lists_.alloc((size_t)MAX_BUCKETS * std::max(plan_.natoms, 1)))
    return AENET_STATUS_ERR_MALLOC;
  if (bucket_counts_.n < (size_t)MAX_BUCKETS && bucket_counts_.alloc(MAX_BUCKETS)) return AENET_STATUS_ERR_MALLOC;
  buckets_[&dm] = pl;
  return AENET_STATUS_OK;
}

void Batch::buckets(const DeviceModel &dm, BucketPlan &out) const {
  const auto it = buckets_.find(&dm);
  out = it == buckets_.end() ? BucketPlan() : it->second;
}

int Batch::tune_buckets(const DeviceModel &dm, const double *d_coords, const int *d_types, bool forces, int repeats,
                        double ms_single, hipStream_t stream, std::string &err) {
  buckets_.erase(&dm);
  const int nat = plan_.natoms;
  if (nat < 4096 || nb_cap_ <= 16) return AENET_STATUS_OK;
  DBuf<double> E_struct, F;
  if (E_struct.alloc(2 * (size_t)plan_.st.size()) || (forces && F.alloc(3 * (size_t)nat))) {
    err = "device allocation failed";
    return AENET_STATUS_ERR_MALLOC;
  }
  double* const Fp = forces ? F.p : nullptr;
  // the neighbor counts of these coordinates
  int st = eval(dm, d_coords, d_types, scratch_E_.p, E_struct.p, Fp, stream);
  std::vector<int> cnt(nat);
  if (!st && (hipStreamSynchronize(stream) != hipSuccess ||
              hipMemcpy(cnt.data(), nb_count_.p, nat * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess))
    st = AENET_STATUS_ERR_DEVICE;
  if (st) {
    err = "evaluation failed while tuning";
    return st;
  }
  const std::vector<int> caps = bucket_caps_for(cnt, nb_cap_);
  if (caps.size() < 2) return AENET_STATUS_OK;
  const auto single = tune_.find(&dm);
  std::vector<int> shapes;
  for (size_t k = 0; k < caps.size(); ++k) {
    shapes.push_back(single != tune_.end() ? single->second[0] : 1);
    shapes.push_back(single != tune_.end() ? std::max(single->second[1], 0) : 0);
    shapes.push_back(0);
  }
  // the single launch's shape may not fit a bucket (it does: a smaller
  // capacity needs less LDS); any failure drops the plan
  st = set_buckets(dm, (int)caps.size(), caps.data(), shapes.data(), err);
  if (st) {
    buckets_.erase(&dm);
    return st == AENET_STATUS_ERR_MALLOC ? st : AENET_STATUS_OK;
  }
  BucketPlan &pl = buckets_[&dm];
  for (int k = 0; k < pl.n; ++k) {
    long a = 0;
    for (int n : cnt) {
      int b = pl.n - 1;
      for (int j = pl.n - 2; j >= 0; --j)
        if (n <= pl.cap[j]) b = j;
      a += b == k;
    }
    pl.atoms[k] = (int)a;
  }
  const bool was_timing = timing_;
  const bool cheb = dm.desc == DESC_CHEBYSHEV;
  const int max_waves = cheb ? CHEB_MAX_WAVES : ATOM_MAX_WAVES;
  const int max_tile = cheb && dm.net_uniform ? 1 : 0;
  auto timed = [&](double &avg) -> int {
    double t0 = 0.0;
    int c0 = 0;
    (void)kernel_time(t0, c0);
    timing_ = false;
    int s = eval(dm, d_coords, d_types, scratch_E_.p, E_struct.p, Fp, stream);  // warm-up
    timing_ = true;
    for (int r = 0; r < repeats && !s; ++r) s = eval(dm, d_coords, d_types, scratch_E_.p, E_struct.p, Fp, stream);
    timing_ = false;
    double ms = 0.0;
    int count = 0;
    if (!s) s = kernel_time(ms, count);
    avg = count > 0 ? ms / count : 1e300;
    return s;
  };
  for (int k = 0; k < pl.n && !st; ++k) {
    timed_bucket_ = k;
    double best = 1e300;
    std::array<int, 3> keep = pl.shape[k];
    for (int tile = 0; tile <= max_tile && !st; ++tile)
      for (int waves = 1; waves <= max_waves && !st; waves *= 2)
        for (int staged = 0; staged <= 1 && !st; ++staged) {
          AtomSizes z = dm.sizes(nat, pl.cap[k]);
          z.tune_waves = waves;
          z.tune_staged = staged;
          z.tune_tile = tile;
          if (!(cheb ? atom_cheb_shape_ok(z) : atom_bp_shape_ok(z))) continue;
          pl.shape[k] = {waves, staged, tile};
          double avg = 0.0;
          st = timed(avg);
          if (!st && avg < best) {
            best = avg;
            keep = pl.shape[k];
          }
        }
    pl.shape[k] = keep;
    pl.ms[k] = best < 1e300 ? best : 0.0;
  }
  timed_bucket_ = -1;
  double total = 1e300;
  if (!st) st = timed(total);
  timing_ = was_timing;
  {
    double t1 = 0.0;
    int c1 = 0;
    (void)search_time(t1, c1);  // the tuner's own search launches are not reported
  }
  if (st) {
    buckets_.erase(&dm);
    err = "evaluation failed while tuning";
    return st;
  }
  pl.ms_single = ms_single;
  pl.ms_buckets = total;
  // buckets only where they measure faster than the single launch
  if (!(total < 0.98 * ms_single)) buckets_.erase(&dm);
  return AENET_STATUS_OK;
}

AtomSizes Batch::sizes_for(const DeviceModel &dm, int natoms, const int *active, const int *nactive) const {
  AtomSizes z = dm.sizes(natoms, nb_cap_, active, nactive);
  const auto it = tune_.find(&dm);
  if (it != tune_.end()) {
    z.tune_waves = it->second[0];
    z.tune_staged = it->second[1];
    z.tune_tile = it->second[2];
  }
  return z;
}

int Batch::autotune(const DeviceModel &dm, const double *d_coords, const int *d_types, bool forces, int repeats,
                    hipStream_t stream, std::vector<std::array<double, 4>> &out, std::string &err) {
  out.clear();
  repeats = std::max(1, std::min(repeats, 100));
  const int nat = plan_.natoms;
  DBuf<double> E_struct, F;
  if (E_struct.alloc(2 * (size_t)plan_.st.size()) || (forces && F.alloc(3 * (size_t)nat))) {
    err = "device allocation failed";
    return AENET_STATUS_ERR_MALLOC;
  }
  const bool was_timing = timing_;
  double t0 = 0.0;
  int c0 = 0;
  (void)kernel_time(t0, c0);  // drop events of earlier evaluations
  tune_.erase(&dm);
  buckets_.erase(&dm);
  const int max_waves = dm.desc == DESC_CHEBYSHEV ? CHEB_MAX_WAVES : ATOM_MAX_WAVES;
  // the matrix-tile network is a candidate where the Chebyshev kernel has it
  const int max_tile = dm.desc == DESC_CHEBYSHEV && dm.net_uniform ? 1 : 0;
  double best = 1e300;
  std::array<int, 3> choice = {0, -1, -1};
  int st = AENET_STATUS_OK;
  for (int tile = 0; tile <= max_tile && !st; ++tile)
  for (int waves = 1; waves <= max_waves && !st; waves *= 2) {
    for (int staged = 0; staged <= 1 && !st; ++staged) {
      AtomSizes z = dm.sizes(nat, nb_cap_);
      z.tune_waves = waves;
      z.tune_staged = staged;
      z.tune_tile = tile;
      const bool ok = dm.desc == DESC_CHEBYSHEV ? atom_cheb_shape_ok(z) : atom_bp_shape_ok(z);
      if (!ok) continue;
      tune_[&dm] = {waves, staged, tile};
      timing_ = false;
      st = eval(dm, d_coords, d_types, scratch_E_.p, E_struct.p, forces ? F.p : nullptr, stream);  // warm-up
      timing_ = true;
      for (int r = 0; r < repeats && !st; ++r)
        st = eval(dm, d_coords, d_types, scratch_E_.p, E_struct.p, forces ? F.p : nullptr, stream);
      double ms = 0.0;
      int count = 0;
      if (!st) st = kernel_time(ms, count);
      if (!st && count > 0) {
        const double avg = ms / count;
        out.push_back({(double)waves, (double)staged, (double)tile, avg});
        if (avg < best) {
          best = avg;
          choice = {waves, staged, tile};
        }
      }
    }
  }
  timing_ = was_timing;
  {
    double t1 = 0.0;
    int c1 = 0;
    (void)search_time(t1, c1);  // the tuner's own search launches are not reported
  }
  if (st) {
    tune_.erase(&dm);
    err = "evaluation failed while tuning";
    return st;
  }
  st = check_flags(stream, err);
  if (st) {
    tune_.erase(&dm);
    return st;
  }
  if (out.empty()) {
    tune_.erase(&dm);
    err = "no launch shape fits";
    return AENET_STATUS_ERR_DEVICE;
  }
  tune_[&dm] = choice;
  // then neighbor-count buckets, kept where they beat the single launch
  return tune_buckets(dm, d_coords, d_types, forces, repeats, best, stream, err);
}

int Batch::set_launch(const DeviceModel &dm, int waves, int staged, int tile, std::string &err) {
  buckets_.erase(&dm);
  if (waves == 0) {
    tune_.erase(&dm);
    return AENET_STATUS_OK;
  }
  const bool cheb = dm.desc == DESC_CHEBYSHEV;
  const int max_waves = cheb ? CHEB_MAX_WAVES : ATOM_MAX_WAVES;
  if (waves < 0 || waves > max_waves || (waves & (waves - 1)) || staged < 0 || staged > 1 || tile < 0 || tile > 1) {
    err = "launch shape out of range";
    return AENET_AMD_ERR_ARG_VALUE;
  }
  if (tile && !(cheb && dm.net_uniform)) {
    err = "the tile network needs Chebyshev descriptors and one network architecture for every species";
    return AENET_AMD_ERR_ARG_VALUE;
  }
  AtomSizes z = dm.sizes(plan_.natoms, nb_cap_);
  z.tune_waves = waves;
  z.tune_staged = staged;
  z.tune_tile = tile;
  if (!(cheb ? atom_cheb_shape_ok(z) : atom_bp_shape_ok(z))) {
    err = "launch shape does not fit the device";
    return AENET_AMD_ERR_ARG_VALUE;
  }
  tune_[&dm] = {waves, staged, tile};
  return AENET_STATUS_OK;
}

void Batch::set_timing(bool on) { timing_ = on; }

static int drain_events(std::vector<hipEvent_t> &ev, double &ms, int &count) {
  ms = 0.0;
  count = 0;
  int st = AENET_STATUS_OK;
  for (size_t i = 0; i + 1 < ev.size(); i += 2) {
    float t = 0.0f;
    if (hipEventSynchronize(ev[i + 1]) != hipSuccess || hipEventElapsedTime(&t, ev[i], ev[i + 1]) != hipSuccess)
      st = AENET_STATUS_ERR_DEVICE;
    ms += t;
    ++count;
  }
  for (auto e : ev) (void)hipEventDestroy(e);
  ev.clear();
  return st;
}

int Batch::kernel_time(double &ms, int &count) { return drain_events(ev_, ms, count); }

int Batch::search_time(double &ms, int &count) { return drain_events(sev_, ms, count); }

Batch::~Batch() {
  for (auto e : ev_) (void)hipEventDestroy(e);
  for (auto e : sev_) (void)hipEventDestroy(e);
}

static std::string flag_text(int f) {
  std::string s;
  if (f & AENET_AMD_FLAG_NB_OVERFLOW) s += " neighbor capacity exceeded;";
  if (f & AENET_AMD_FLAG_COMBO_OVERFLOW) s += " cell-image combination capacity exceeded;";
  if (f & AENET_AMD_FLAG_KEPT_OVERFLOW) s += " per-atom cell-image capacity exceeded;";
  if (f & AENET_AMD_FLAG_BAD_TYPE) s += " atom type out of range;";
  if (f & AENET_AMD_FLAG_TRIPLET_OVERFLOW) s += " triplet capacity exceeded;";
  if (f & AENET_AMD_FLAG_BAD_COORD) s += " non-finite or out-of-range coordinates;";
  return s;
}

int Batch::check_flags(hipStream_t stream, std::string &err) {
  int f = 0;
  if (hipMemcpyAsync(&f, flags_.p, sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess) {
    err = "device error";
    return AENET_STATUS_ERR_DEVICE;
  }
  if (f) {
    err = "kernel reported:" + flag_text(f);
    int zero = 0;
    (void)hipMemcpy(flags_.p, &zero, sizeof(int), hipMemcpyHostToDevice);
    if (f & AENET_AMD_FLAG_BAD_TYPE) return AENET_STATUS_ERR_TYPE;
    return (f & ~AENET_AMD_FLAG_BAD_COORD) ? AENET_STATUS_ERR_CAPACITY : AENET_STATUS_ERR_COORD;
  }
  return AENET_STATUS_OK;
}

EnvScratch::~EnvScratch() {
  if (pinned) (void)hipHostFree(pinned);
}

int EnvScratch::eval(const DeviceModel &dm, int nenv, const int *h_start, const int *h_ctype, const double *h_ci,
                     const double *h_cj, const int *h_tj, bool forces, double *h_E, double *h_dFi,
                     double *h_dFj, double *h_G, int G_stride, hipStream_t stream, std::string &err,
                     int coef_unit) {
  const size_t nnb = (size_t)h_start[nenv];
  int max_n = 1;
  for (int i = 0; i < nenv; ++i) max_n = std::max(max_n, h_start[i + 1] - h_start[i]);
  if (max_n > AENET_AMD_NB_CAP) {
    err = "more than 512 neighbors in one environment";
    return AENET_STATUS_ERR_CAPACITY;
  }
  max_n = (max_n + 31) / 32 * 32;
  // packed layout (byte offsets, 8-aligned double sections)
  auto up8 = [](size_t v) { return (v + 7) / 8 * 8; };
  const size_t o_start = 0, o_ctype = o_start + 4 * (size_t)(nenv + 1), o_tj = o_ctype + 4 * (size_t)nenv;
  const size_t o_ci = up8(o_tj + 4 * nnb), o_cj = o_ci + 24 * (size_t)nenv, o_flags = o_cj + 24 * nnb;
  const size_t o_E = o_flags + 8, o_dFi = o_E + 8 * (size_t)nenv, o_dFj = o_dFi + 24 * (size_t)nenv;
  const size_t o_end = o_dFj + 24 * nnb;
  if (o_end > bytes || !dev.p || !pinned) {
    const size_t want = std::max(o_end + o_end / 2, (size_t)(64 << 10));
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    bytes = 0;
    if (dev.alloc(want) || hipHostMalloc((void **)&pinned, want, hipHostMallocDefault) != hipSuccess) {
      pinned = nullptr;
      err = "device or pinned host allocation failed";
      return AENET_STATUS_ERR_MALLOC;
    }
    bytes = want;
  }
  DBuf<double> G;
  if (h_G && G.alloc((size_t)nenv * G_stride)) {
    err = "device allocation failed";
    return AENET_STATUS_ERR_MALLOC;
  }
  std::memcpy(pinned + o_start, h_start, 4 * (size_t)(nenv + 1));
  std::memcpy(pinned + o_ctype, h_ctype, 4 * (size_t)nenv);
  if (nnb) std::memcpy(pinned + o_tj, h_tj, 4 * nnb);
  std::memcpy(pinned + o_ci, h_ci, 24 * (size_t)nenv);
  if (nnb) std::memcpy(pinned + o_cj, h_cj, 24 * nnb);
  std::memset(pinned + o_flags, 0, 8);
  if (hipMemcpyAsync(dev.p, pinned, o_E, hipMemcpyHostToDevice, stream) != hipSuccess) {
    err = "device error";
    return AENET_STATUS_ERR_DEVICE;
  }
  char *d = dev.p;
  BatchDev b;
  std::memset(&b, 0, sizeof(b));
  b.flags = reinterpret_cast<int *>(d + o_flags);
  OutDev o;
  std::memset(&o, 0, sizeof(o));
  o.G = h_G ? G.p : nullptr;
  o.G_scaled = 0;
  o.G_stride = G_stride;
  o.desc_only = h_G && !h_E;
  o.coef_unit = coef_unit;
  EnvDev e;
  e.nenv = nenv;
  e.env_start = reinterpret_cast<int *>(d + o_start);
  e.env_type = reinterpret_cast<int *>(d + o_ctype);
  e.coo_i = reinterpret_cast<double *>(d + o_ci);
  e.coo_j = reinterpret_cast<double *>(d + o_cj);
  e.type_j = reinterpret_cast<int *>(d + o_tj);
  e.E = reinterpret_cast<double *>(d + o_E);
  e.dF_i = reinterpret_cast<double *>(d + o_dFi);
  e.dF_j = reinterpret_cast<double *>(d + o_dFj);
  ModelDev md = dm.view();
  NbDev nb;
  std::memset(&nb, 0, sizeof(nb));
  if (dm.desc == DESC_CHEBYSHEV)
    launch_atom_cheb(dm.max_rN, dm.max_aN, 1, forces, dm.sizes(nenv, max_n), md, b, nb, o, e, stream);
  else
    launch_atom_bp(1, forces, dm.sizes(nenv, max_n), md, b, nb, o, e, stream);
  if (hipGetLastError() != hipSuccess) {
    err = "kernel launch failed";
    return AENET_STATUS_ERR_DEVICE;
  }
  const size_t back = forces ? o_end : o_dFi;
  (void)hipMemcpyAsync(pinned + o_flags, d + o_flags, back - o_flags, hipMemcpyDeviceToHost, stream);
  if (h_G) (void)hipMemcpyAsync(h_G, G.p, (size_t)nenv * G_stride * sizeof(double), hipMemcpyDeviceToHost, stream);
  if (hipStreamSynchronize(stream) != hipSuccess) {
    err = "device error";
    return AENET_STATUS_ERR_DEVICE;
  }
  int f = 0;
  std::memcpy(&f, pinned + o_flags, sizeof(int));
  if (f) {
    err = "kernel reported:" + flag_text(f);
    return (f & AENET_AMD_FLAG_BAD_TYPE) ? AENET_STATUS_ERR_TYPE : AENET_STATUS_ERR_CAPACITY;
  }
  if (h_E) std::memcpy(h_E, pinned + o_E, 8 * (size_t)nenv);
  if (forces) {
    std::memcpy(h_dFi, pinned + o_dFi, 24 * (size_t)nenv);
    if (nnb) std::memcpy(h_dFj, pinned + o_dFj, 24 * nnb);
  }
  return AENET_STATUS_OK;
}

}  // namespace aenet_amd

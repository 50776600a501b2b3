// mc.cpp -- host side of the GPU Monte Carlo (kernels in mc.hip).
#include "mc.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "aenet_capi_status.h"
#include "kernels.h"

namespace aenet_amd {

namespace {

bool invert3_impl(const double *A, double *inv) {
  // A row-major 3x3
  const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                     A[2] * (A[3] * A[7] - A[4] * A[6]);
  if (!(std::fabs(det) > 1e-300)) return false;
  const double d = 1.0 / det;
  inv[0] = (A[4] * A[8] - A[5] * A[7]) * d;
  inv[1] = (A[2] * A[7] - A[1] * A[8]) * d;
  inv[2] = (A[1] * A[5] - A[2] * A[4]) * d;
  inv[3] = (A[5] * A[6] - A[3] * A[8]) * d;
  inv[4] = (A[0] * A[8] - A[2] * A[6]) * d;
  inv[5] = (A[2] * A[3] - A[0] * A[5]) * d;
  inv[6] = (A[3] * A[7] - A[4] * A[6]) * d;
  inv[7] = (A[1] * A[6] - A[0] * A[7]) * d;
  inv[8] = (A[0] * A[4] - A[1] * A[3]) * d;
  return true;
}

// smallest distance between opposite faces of the cell (rows = vectors)
double min_width_impl(const double *L) {
  auto cross = [](const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
  };
  double c[3];
  cross(L + 3, L + 6, c);
  const double V = std::fabs(L[0] * c[0] + L[1] * c[1] + L[2] * c[2]);
  double w = 1e300;
  for (int k = 0; k < 3; ++k) {
    cross(L + 3 * ((k + 1) % 3), L + 3 * ((k + 2) % 3), c);
    w = std::min(w, V / std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
  }
  return w;
}

}  // namespace

bool mc_invert_cell(const double *L, double *inv) { return invert3_impl(L, inv); }
double mc_min_width(const double *L) { return min_width_impl(L); }

MonteCarlo::~MonteCarlo() {
  drop_graphs();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void MonteCarlo::drop_graphs() {
  for (int k = 0; k < 2; ++k) {
    if (exec_[k]) (void)hipGraphExecDestroy(exec_[k]);
    if (graph_[k]) (void)hipGraphDestroy(graph_[k]);
    exec_[k] = nullptr;
    graph_[k] = nullptr;
  }
}

int MonteCarlo::create(const PotSet &tset, const DeviceModel *teacher, const DeviceModel *student, int C, int N,
                       const double *latt, const int *types, const double *coords, unsigned long long seed,
                       std::string &err) {
  if (C < 1 || N < 2 || N > 1024) {
    err = "1 or more chains of 2 to 1024 atoms";
    return AENET_STATUS_ERR_ARG;
  }
  teacher_ = teacher;
  student_ = student;
  C_ = C;
  N_ = N;
  std::memcpy(L_, latt, sizeof(L_));
  if (!mc_invert_cell(L_, Binv_)) {
    err = "singular cell";
    return AENET_STATUS_ERR_ARG;
  }
  rc_t_ = teacher->rc_max + 1.0e-3;
  double rc_s = 0.0;
  if (student) {
    rc_s = student->rc_max;
    nsf_s_ = student->nsf_max;
    rn_s_ = student->max_rN;
    nh_s_ = student->hmax;
    if (!mc_screen_fits(N, student->ntypes, rn_s_)) {
      err = "cell too large for the first-stage kernel's LDS";
      return AENET_STATUS_ERR_ARG;
    }
  }
  // the first stage and the active-atom test use minimum images
  if (mc_min_width(L_) <= 2.0 * std::max(rc_t_, rc_s)) {
    err = "cell narrower than twice the cutoff: minimum-image Monte Carlo needs a larger cell";
    return AENET_STATUS_ERR_ARG;
  }
  for (long i = 0; i < (long)C * N; ++i)
    if (types[i] < 1 || types[i] > tset.ntypes) {
      err = "atom type out of range";
      return AENET_STATUS_ERR_TYPE;
    }
  std::vector<int> nat(C, N), pbc(C, 1);
  std::vector<double> lat(9 * (size_t)C);
  for (int c = 0; c < C; ++c) std::memcpy(&lat[9 * (size_t)c], latt, 9 * sizeof(double));
  int st = batch_.create(tset, C, nat.data(), lat.data(), pbc.data(), coords, err);
  if (st) return st;
  if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) {
    err = "stream creation failed";
    return AENET_STATUS_ERR_DEVICE;
  }
  const size_t n = (size_t)C * N;
  bool bad = false;
  bad |= coords_.upload(coords, 3 * n, stream_) != 0;
  bad |= types_.upload(types, n, stream_) != 0;
  bad |= old_pos_.alloc(3 * (size_t)C) || u1_.alloc(C) || u2_.alloc(C) || dEs_.alloc(C);
  bad |= E_t_.alloc(n) || E_trial_.alloc(n) || E_chain_.alloc(C) || E_struct_.alloc(2 * (size_t)C);
  bad |= G_s_.alloc(n * nsf_s_) || G_trial_.alloc(n * nsf_s_) || E_s_.alloc(n) || E_s_trial_.alloc(n);
  bad |= kind_.alloc(C) || a_.alloc(C) || b_.alloc(C) || need_.alloc(C) || nactive_.alloc(1);
  bad |= active_.alloc(n) || act_start_.alloc(C) || act_count_.alloc(C) || affected_.alloc(n);
  bad |= stats_.alloc(4 * (size_t)C) || step_.alloc(1) || nprop_.alloc(C);
  if (student) bad |= fs_tab_.alloc(first_stage_table_bytes(student->ntypes) / 8) != 0;
  std::vector<unsigned long long> r(C);
  for (int c = 0; c < C; ++c) r[c] = seed ^ (0x9E3779B97F4A7C15ull * (unsigned long long)(c + 1));
  bad |= rng_.upload(r.data(), C, stream_) != 0;
  if (bad) {
    err = "device allocation failed";
    return AENET_STATUS_ERR_MALLOC;
  }
  launch_fill_i32(affected_.p, 0, (long)n, stream_);
  launch_fill_i32(need_.p, 0, C, stream_);
  launch_fill_i32(act_count_.p, 0, C, stream_);
  launch_fill_i32(reinterpret_cast<int *>(stats_.p), 0, 8L * C, stream_);
  launch_fill_i32(reinterpret_cast<int *>(step_.p), 0, 2, stream_);
  launch_fill_i32(reinterpret_cast<int *>(nprop_.p), 0, 2L * C, stream_);
  if (hipStreamSynchronize(stream_) != hipSuccess) {
    err = "device error";
    return AENET_STATUS_ERR_DEVICE;
  }
  return refresh(err);
}

void MonteCarlo::set_params(double beta, double p_swap, double max_disp, int max_screen) {
  beta_ = beta;
  p_swap_ = p_swap;
  max_disp_ = max_disp;
  max_screen_ = max_screen;
  drop_graphs();
}

MCDev MonteCarlo::view() const {
  MCDev m;
  std::memset(&m, 0, sizeof(m));
  m.nchains = C_;
  m.natoms = N_;
  std::memcpy(m.L, L_, sizeof(L_));
  std::memcpy(m.Binv, Binv_, sizeof(Binv_));
  m.coords = coords_.p;
  m.types = types_.p;
  m.rng = rng_.p;
  m.beta = beta_;
  m.p_swap = p_swap_;
  m.max_disp = max_disp_;
  m.max_screen = max_screen_;
  m.rc_t = rc_t_;
  m.kind = kind_.p;
  m.site_a = a_.p;
  m.site_b = b_.p;
  m.old_pos = old_pos_.p;
  m.u_stage1 = u1_.p;
  m.u_stage2 = u2_.p;
  m.need = need_.p;
  m.dE_s = dEs_.p;
  m.nactive = nactive_.p;
  m.active = active_.p;
  m.act_start = act_start_.p;
  m.act_count = act_count_.p;
  m.E_t = E_t_.p;
  m.E_trial = E_trial_.p;
  m.E_chain = E_chain_.p;
  m.nsf_s = nsf_s_;
  m.rn_s = rn_s_;
  m.nh_s = nh_s_;
  m.fs_tab = fs_tab_.p;
  m.G_s = G_s_.p;
  m.G_trial = G_trial_.p;
  m.E_s = E_s_.p;
  m.E_s_trial = E_s_trial_.p;
  m.affected = affected_.p;
  m.stats = stats_.p;
  m.step = step_.p;
  m.nprop = nprop_.p;
  m.samples = sample_every_ > 0 ? samples_.p : nullptr;
  m.trace = max_trace_ > 0 ? trace_.p : nullptr;
  m.max_trace = max_trace_;
  m.sample_every = sample_every_;
  m.max_samples = max_samples_;
  return m;
}

int MonteCarlo::refresh(std::string &err) {
  // every atom through the same minimum-image lists the steps use
  const MCDev m = view();
  launch_mc_all_active(m, stream_);
  launch_mc_neighbors(m, batch_.nb_lists(), rc_t_, batch_.flags_ptr(), stream_);
  int st = batch_.eval_active(*teacher_, coords_.p, types_.p, active_.p, nactive_.p, C_ * N_, true, E_t_.p, stream_);
  if (st) return st;
  launch_mc_chain_sum(m, stream_);
  st = batch_.check_flags(stream_, err);
  if (st) return st;
  if (student_) {
    launch_first_stage_table(student_->view(), fs_tab_.p, stream_);
    launch_mc_student_init(m, student_->view(), stream_);
  }
  if (hipStreamSynchronize(stream_) != hipSuccess) {
    err = "device error";
    return AENET_STATUS_ERR_DEVICE;
  }
  return AENET_STATUS_OK;
}

int MonteCarlo::step_once(const MCDev &m, bool delayed) {
  ModelDev sm;
  if (delayed) sm = student_->view();
  launch_mc_step(m, delayed ? &sm : nullptr, !delayed, 0, stream_);
  launch_mc_neighbors(m, batch_.nb_lists(), rc_t_, batch_.flags_ptr(), stream_);
  int st = batch_.eval_active(*teacher_, coords_.p, types_.p, active_.p, nactive_.p, C_ * N_, true, E_trial_.p,
                              stream_);
  if (st) return st;
  launch_mc_step(m, delayed ? &sm : nullptr, !delayed, 1, stream_);
  return hipGetLastError() == hipSuccess ? AENET_STATUS_OK : AENET_STATUS_ERR_DEVICE;
}

int MonteCarlo::run(int nsteps, bool delayed, int sample_every, bool graph, std::string &err) {
  if (delayed && !student_) {
    err = "delayed acceptance needs a first-stage potential";
    return AENET_STATUS_ERR_ARG;
  }
  if (nsteps <= 0) return AENET_STATUS_OK;
  // samples of this run; the step counter restarts
  drop_graphs();
  sample_every_ = sample_every > 0 ? sample_every : 0;
  const long long max_props = (long long)nsteps * (delayed ? max_screen_ : 1);
  max_samples_ = sample_every_ ? max_props / sample_every_ : 0;
  if (max_samples_ > 0 && samples_.alloc((size_t)max_samples_ * C_)) {
    err = "device allocation failed";
    return AENET_STATUS_ERR_MALLOC;
  }
  launch_fill_i32(reinterpret_cast<int *>(step_.p), 0, 2, stream_);
  launch_fill_i32(reinterpret_cast<int *>(nprop_.p), 0, 2L * C_, stream_);
  if (max_trace_ > 0) launch_mc_trace_init(trace_.p, max_trace_ * C_ * 4, stream_);
  const MCDev m = view();
  const int k = delayed ? 1 : 0;
  // the first stage from scratch: its weights may have changed since the
  // last run, and plain steps move atoms without updating its state
  if (delayed) {
    launch_first_stage_table(student_->view(), fs_tab_.p, stream_);
    launch_mc_student_init(m, student_->view(), stream_);
  }
  int st = AENET_STATUS_OK;
  if (graph) {
    if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      err = "graph capture failed to start";
      return AENET_STATUS_ERR_DEVICE;
    }
    st = step_once(m, delayed);
    hipError_t e = hipStreamEndCapture(stream_, &graph_[k]);
    if (st || e != hipSuccess || hipGraphInstantiate(&exec_[k], graph_[k], nullptr, nullptr, 0) != hipSuccess) {
      drop_graphs();
      err = "graph capture failed";
      return st ? st : AENET_STATUS_ERR_DEVICE;
    }
    for (int i = 0; i < nsteps; ++i)
      if (hipGraphLaunch(exec_[k], stream_) != hipSuccess) {
        err = "graph launch failed";
        return AENET_STATUS_ERR_DEVICE;
      }
  } else {
    for (int i = 0; i < nsteps && !st; ++i) st = step_once(m, delayed);
    if (st) {
      err = "kernel launch failed";
      return st;
    }
  }
  if (hipStreamSynchronize(stream_) != hipSuccess) {
    err = "device error";
    return AENET_STATUS_ERR_DEVICE;
  }
  return batch_.check_flags(stream_, err);
}

int MonteCarlo::set_trace(long long max_props) {
  drop_graphs();
  max_trace_ = 0;
  if (max_props > 0) {
    if (trace_.alloc((size_t)max_props * C_ * 4)) return AENET_STATUS_ERR_MALLOC;
    max_trace_ = max_props;
  }
  return AENET_STATUS_OK;
}

long long MonteCarlo::trace(double *out) {
  if (max_trace_ <= 0) return 0;
  if (out && (hipMemcpyAsync(out, trace_.p, (size_t)max_trace_ * C_ * 4 * sizeof(double), hipMemcpyDeviceToHost,
                             stream_) != hipSuccess ||
              hipStreamSynchronize(stream_) != hipSuccess))
    return -1;
  return max_trace_;
}

int MonteCarlo::stats(long long out[4]) {
  std::vector<long long> h(4 * (size_t)C_);
  if (hipMemcpyAsync(h.data(), stats_.p, h.size() * sizeof(long long), hipMemcpyDeviceToHost, stream_) != hipSuccess ||
      hipStreamSynchronize(stream_) != hipSuccess)
    return AENET_STATUS_ERR_DEVICE;
  for (int k = 0; k < 4; ++k) out[k] = 0;
  for (int c = 0; c < C_; ++c)
    for (int k = 0; k < 4; ++k) out[k] += h[4 * (size_t)c + k];
  return AENET_STATUS_OK;
}

void MonteCarlo::reset_stats() {
  launch_fill_i32(reinterpret_cast<int *>(stats_.p), 0, 8L * C_, stream_);
  (void)hipStreamSynchronize(stream_);
}

int MonteCarlo::energies(double *E_chain) {
  if (hipMemcpyAsync(E_chain, E_chain_.p, C_ * sizeof(double), hipMemcpyDeviceToHost, stream_) != hipSuccess ||
      hipStreamSynchronize(stream_) != hipSuccess)
    return AENET_STATUS_ERR_DEVICE;
  return AENET_STATUS_OK;
}

long long MonteCarlo::samples(double *out) {
  if (max_samples_ <= 0) return 0;
  std::vector<long long> np(C_);
  if (hipMemcpyAsync(np.data(), nprop_.p, C_ * sizeof(long long), hipMemcpyDeviceToHost, stream_) != hipSuccess ||
      hipStreamSynchronize(stream_) != hipSuccess)
    return -1;
  long long n = max_samples_;
  for (int c = 0; c < C_; ++c) n = std::min(n, np[c] / sample_every_);
  if (out && n > 0 &&
      (hipMemcpyAsync(out, samples_.p, (size_t)n * C_ * sizeof(double), hipMemcpyDeviceToHost, stream_) != hipSuccess ||
       hipStreamSynchronize(stream_) != hipSuccess))
    return -1;
  return n;
}

int MonteCarlo::state(double *coords, int *types) {
  const size_t n = (size_t)C_ * N_;
  if ((coords && hipMemcpyAsync(coords, coords_.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, stream_)) ||
      (types && hipMemcpyAsync(types, types_.p, n * sizeof(int), hipMemcpyDeviceToHost, stream_)) ||
      hipStreamSynchronize(stream_) != hipSuccess)
    return AENET_STATUS_ERR_DEVICE;
  return AENET_STATUS_OK;
}

int MonteCarlo::student_state(double *E_s, double *G_s) {
  const size_t n = (size_t)C_ * N_;
  if ((E_s && hipMemcpyAsync(E_s, E_s_.p, n * sizeof(double), hipMemcpyDeviceToHost, stream_)) ||
      (G_s && hipMemcpyAsync(G_s, G_s_.p, n * nsf_s_ * sizeof(double), hipMemcpyDeviceToHost, stream_)) ||
      hipStreamSynchronize(stream_) != hipSuccess)
    return AENET_STATUS_ERR_DEVICE;
  return AENET_STATUS_OK;
}

}  // namespace aenet_amd

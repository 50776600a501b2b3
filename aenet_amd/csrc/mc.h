// mc.h -- Monte Carlo of independent periodic chains on the GPU (host side
// of mc.hip): plain Metropolis with the full potential, or delayed
// acceptance with a radial-only first-stage potential.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>

#include "engine.h"

namespace aenet_amd {

// inverse of a cell (rows = vectors) and its smallest face-to-face width
bool mc_invert_cell(const double *L, double *inv);
double mc_min_width(const double *L);

class MonteCarlo {
 public:
  // C chains of N atoms in the same periodic cell latt (rows = vectors);
  // types (C*N, 1-based) and coords (3*C*N) on the host.  student may be
  // null (plain runs only).
  int create(const PotSet &teacher_set, const DeviceModel *teacher, const DeviceModel *student, int C, int N,
             const double *latt, const int *types, const double *coords, unsigned long long seed,
             std::string &err);
  void set_params(double beta, double p_swap, double max_disp, int max_screen);
  // energies (and first-stage descriptors) from scratch
  int refresh(std::string &err);
  // nsteps steps; a plain step makes one proposal per chain, a
  // delayed-acceptance step up to max_screen (the first-stage tries up to
  // the first that passes); samples of the chain energies every
  // sample_every proposals of each chain (0: none); graph: capture one
  // step, replay it
  int run(int nsteps, bool delayed, int sample_every, bool graph, std::string &err);
  int stats(long long out[4]);
  void reset_stats();
  int energies(double *E_chain);
  // the samples every chain completed (returns their count per chain)
  long long samples(double *out);
  // per-proposal, per-chain record of the decisions of the next runs (0: off)
  int set_trace(long long max_props);
  long long trace(double *out);
  int state(double *coords, int *types);
  int student_state(double *E_s, double *G_s);
  int nsf_student() const { return nsf_s_; }
  ~MonteCarlo();

 private:
  MCDev view() const;
  int step_once(const MCDev &m, bool delayed);
  void drop_graphs();
  const DeviceModel *teacher_ = nullptr, *student_ = nullptr;
  Batch batch_;
  hipStream_t stream_ = nullptr;
  DBuf<double> coords_, old_pos_, u1_, u2_, dEs_, E_t_, E_trial_, E_chain_, G_s_, G_trial_, E_s_, E_s_trial_;
  DBuf<double> samples_, E_struct_, trace_, fs_tab_;  // fs_tab_: first-stage table (bytes / 8)
  long long max_trace_ = 0;
  DBuf<int> types_, kind_, a_, b_, need_, nactive_, active_, act_start_, act_count_, affected_;
  DBuf<unsigned long long> rng_;
  DBuf<long long> stats_, step_, nprop_;
  int C_ = 0, N_ = 0, nsf_s_ = 1, rn_s_ = 0, nh_s_ = 0;
  double L_[9] = {0}, Binv_[9] = {0};
  double rc_t_ = 0.0;
  double beta_ = 1.0 / (8.617333262e-5 * 600.0), p_swap_ = 0.5, max_disp_ = 0.075;
  int max_screen_ = 4;
  int sample_every_ = 0;
  long long max_samples_ = 0;
  hipGraphExec_t exec_[2] = {nullptr, nullptr};
  hipGraph_t graph_[2] = {nullptr, nullptr};
};

}  // namespace aenet_amd

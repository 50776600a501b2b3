// kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include "aenet_amd_internal.h"

namespace aenet_amd {

// wavefronts per workgroup at most: the Chebyshev atom kernel takes up to
// a CU's worth, so that one staged copy of the network weights serves
// every wavefront of the CU; the other atom kernels take up to 4 (fewer
// when LDS is short)
constexpr int CHEB_MAX_WAVES = 16;
constexpr int ATOM_MAX_WAVES = 4;

// per-atom slots for the workgroup force tables' sums: a batch of up to
// FORCE_SLOT_MAX_ATOMS atoms (one round of wavefronts covers it) gets
// FORCE_SLOTS per atom (the workgroups that share a target); a sum that
// finds them taken goes to the atom's fixed-point F directly
constexpr int FORCE_SLOTS = 32;
constexpr int FORCE_SLOT_MAX_ATOMS = 16384;

// the BP kernel's triplet lists: the angular value pass records the
// triplets that contribute, the force pass replays them.  One list of
// BP_TRIP_CAP entries per resident wavefront (a block that does not fit
// takes the candidate pass again); BP_TRIP_WAVES_PER_CU bounds the
// resident wavefronts of a CU
constexpr int BP_TRIP_CAP = 4096;
constexpr int BP_TRIP_WAVES_PER_CU = 16;

// neighbor-count buckets of a batch's atoms: bucket k holds the atoms whose
// neighbor count n has cap[k-1] < n <= cap[k] (the last cap is the batch's
// capacity), each bucket evaluated by its own atom-kernel launch with LDS
// regions of its cap and its own measured launch shape
constexpr int MAX_BUCKETS = 6;
struct BucketCaps {
  int n;
  int cap[MAX_BUCKETS];
};

// kernels.hip
void launch_fill_i32(int *p, int v, long n, hipStream_t s);
// the atoms into their buckets: lists[k * natoms ..] and counts[k] (cleared
// here first); within a bucket, runs of 256 consecutive atoms keep their
// order
void launch_bucket_atoms(const int *nb_count, int natoms, const BucketCaps &bc, int *lists, int *counts,
                         hipStream_t s);
void launch_fill_f64(double *p, double v, long n, hipStream_t s);
void launch_scale_f64(double *p, double a, long n, hipStream_t s);
// (cell, image) combinations of every home cell; once per batch
void launch_combos(const BatchDev &b, int *combo_count, int *combos, hipStream_t s);
// cell binning of a batch (cell_count/cell_fill are cleared here); tmp: natoms ints
// F (3 per atom, may be null) is cleared on the way
void launch_binning(const BatchDev &b, int *tmp, double *F, hipStream_t s);
// compact neighbor lists for the atom kernels
// active (optional, device): search only atoms active[0 .. *nactive-1];
// max_active bounds *nactive
void launch_search(const BatchDev &b, const NbDev &nb, hipStream_t s, const int *active = nullptr,
                   const int *nactive = nullptr, int max_active = 0);
// bit-exact FP64 neighbor lists
void launch_neighbor_lists(const BatchDev &b, const OutDev &o, hipStream_t s);
// per-structure energies (E_tot null: none) and, with F, the fixed-point
// forces of the atom kernels turned into FP64 times fscale
void launch_struct_energy(const BatchDev &b, const ModelDev &md, const double *E_atom, double *E_tot,
                          double *E_coh, double *F, double fscale, hipStream_t s);

// atom_cheb.hip: mode 0 = structures, 1 = explicit environments
// properties of the current device the launchers size their grids by
struct DeviceLimits {
  int cus;               // compute units
  size_t lds_per_block;  // bytes of LDS one workgroup may allocate
  size_t lds_per_cu;     // bytes of LDS per compute unit
};
const DeviceLimits &device_limits();

// set bits of the batch's error word from the host side of a launch
void raise_flag(int *flags, int bits, hipStream_t s);

// the launch shape that z asks for (measured choice or the occupancy
// model's) fits the device
bool atom_cheb_shape_ok(const AtomSizes &z);
bool atom_bp_shape_ok(const AtomSizes &z);
void launch_atom_cheb(int rN, int aN, int mode, bool forces, const AtomSizes &z, const ModelDev &md,
                      const BatchDev &b, const NbDev &nb, const OutDev &o, const EnvDev &e,
                      hipStream_t stream);

// train.hip
// hmax: the widest non-input layer of the model
void launch_net_grad(const ModelDev &md, const NetGradArgs &a, int natoms, int nsf_max, int nl_max, int wsize,
                     bool grad, hipStream_t stream, int hmax = AENET_AMD_MAX_HIDDEN);
void launch_adam(float *W, const float *g, float *m, float *v, int n, int t, float lr, float b1, float b2,
                 float eps, hipStream_t stream);

// mc.hip: phase 0 = propose (+ first stage) + active lists, phase 1 = finish
void launch_mc_step(const MCDev &m, const ModelDev *student, bool plain, int phase, hipStream_t s);
void launch_mc_trace_init(double *trace, long long n, hipStream_t s);
// the first-stage screening kernel's LDS holds a chain of natoms atoms
bool mc_screen_fits(int natoms, int ntypes, int rn);
size_t first_stage_table_bytes(int ntypes);
void launch_first_stage_table(const ModelDev &student, void *tab, hipStream_t s);
void launch_mc_student_init(const MCDev &m, const ModelDev &student, hipStream_t s);
// neighbor lists (list radius rc_list) of the active atoms by minimum image
void launch_mc_neighbors(const MCDev &m, const NbDev &nb, double rc_list, int *flags, hipStream_t s);
void launch_mc_all_active(const MCDev &m, hipStream_t s);
void launch_mc_chain_sum(const MCDev &m, hipStream_t s);
// fit of the first-stage potential on cells (m: coords/types of the cells,
// cells[ncells] their indices); grad += gradient, loss[0..1] += squared
// energy and force residuals
void launch_first_stage_fit(const MCDev &m, const ModelDev &md, const int *cells, int ncells, const double *E_ref,
                            const double *F_ref, double coefE, double coefF, float *grad, int wsize, double *loss,
                            hipStream_t s);

// atom_bp.hip
void launch_atom_bp(int mode, bool forces, const AtomSizes &z, const ModelDev &md, const BatchDev &b,
                    const NbDev &nb, const OutDev &o, const EnvDev &e, hipStream_t stream);

}  // namespace aenet_amd

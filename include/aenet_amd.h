/*
 * aenet_amd.h -- batched entry points of libaenet_amd.so (MI355X, gfx950).
 *
 * These have no single counterpart in the reference: they batch what
 * predict.x's get_energy (src/predict.F90:574-706) does per atom -- neighbor
 * list, descriptors, network and forces for whole structures -- into one
 * device pass.  They use the potentials loaded with aenet_init /
 * aenet_load_potential (include/aenet.h).  Status codes are the AENET_*
 * values; codes 5-8 are this library's own (see AENET_AMD_ERR_*).
 *
 * Conventions: lattices are 9 doubles per structure, vector k at
 * [3k, 3k+1, 3k+2] (the column-major latticeVec(3,3) of the reference, or
 * equivalently the rows of an ASE cell); coordinates are Cartesian,
 * 3 doubles per atom, structures concatenated; types are 1-based ids of
 * the loaded potentials; isolated structures (pbc = 0) use the bounding box
 * of geo_get_bounds (src/geometry.f90:538-580) as the reference does.
 */
#ifndef AENET_AMD_H_INCLUDED
#define AENET_AMD_H_INCLUDED

#ifdef __cplusplus
extern "C" {
#endif

#define AENET_AMD_ERR_CAPACITY 5
#define AENET_AMD_ERR_DEVICE 6
#define AENET_AMD_ERR_ARG 7
#define AENET_AMD_ERR_COORD 8 /* non-finite coordinates, or an isolated
                                 structure left its bounding box */

/* Energies and forces of nstruct structures given in host memory.
 * E_tot/E_coh: per structure (may be NULL); E_atoms: per atom (may be
 * NULL); F: 3 per atom (set, not accumulated; NULL for energies only).
 * Equivalent of calling predict.x's get_energy on each structure. */
void aenet_structures_energy_and_forces(int nstruct, const int natoms[],
                                        const double latt[], const int pbc[],
                                        const int types[], const double coords[],
                                        double E_tot[], double E_coh[],
                                        double E_atoms[], double F[], int *stat);

/* aenet_structures_energy_and_forces without device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results; every device buffer of the call (inputs,
 * outputs, plan, cell and neighbor lists) is taken from one fixed 64 MiB
 * device array that each GPU the library runs on holds until the process
 * ends.  Structures that need more than that (about 20,000 atoms of dense
 * bulk at a 6.5 A cutoff) are refused with AENET_AMD_ERR_CAPACITY and a
 * message.  Calls are serialized by the library's lock; each call
 * overwrites the array, so calls that overlapped on one GPU (from
 * different threads of the library without that lock) would race.  The
 * host-side planning still uses ordinary host memory. */
void aenet_structures_energy_and_forces_noalloc(int nstruct, const int natoms[],
                                                const double latt[], const int pbc[],
                                                const int types[], const double coords[],
                                                double E_tot[], double E_coh[],
                                                double E_atoms[], double F[], int *stat);

/* Device-resident batches: the structures' shapes (lattices, atom counts,
 * pbc) are fixed at creation; coordinates and types are device pointers
 * passed to each evaluation, so a caller that keeps its structures in HBM
 * (Monte Carlo, molecular dynamics, training) never copies them. */
typedef struct aenet_amd_batch aenet_amd_batch;
typedef struct aenet_amd_model aenet_amd_model;  /* potential-set handles, below */

void aenet_amd_batch_create(int nstruct, const int natoms[], const double latt[],
                            const int pbc[], const double coords_host[],
                            aenet_amd_batch **batch, int *stat);
void aenet_amd_batch_destroy(aenet_amd_batch *batch);
int aenet_amd_batch_natoms(const aenet_amd_batch *batch);

/* Isolated structures (pbc = 0) are binned in the bounding box of the
 * coordinates given at creation (geo_get_bounds).  An evaluation whose
 * device coordinates leave that box raises AENET_AMD_ERR_COORD at the next
 * aenet_amd_batch_check; until then (unchecked calls, graph replays) atoms
 * outside the box are binned into its edge cells and neighbors beyond
 * them can be missed.  aenet_amd_batch_replan makes new bounding boxes
 * from host coordinates (periodic structures are unchanged); the
 * evaluation is then repeated.  has_isolated: 1 if the batch has any
 * isolated structure. */
void aenet_amd_batch_replan(aenet_amd_batch *batch, const double coords_host[], int *stat);
int aenet_amd_batch_has_isolated(const aenet_amd_batch *batch);

/* d_* are device pointers; stream is the hipStream_t every entry point
 * with device pointers launches on (NULL: the default stream, as in HIP),
 * so work is ordered with the caller's own.  d_E_struct holds 2*nstruct doubles: E_tot then E_coh.  d_F
 * (3 per atom) is overwritten; NULL skips forces.  The call is
 * asynchronous; aenet_amd_batch_check synchronizes and reports capacity
 * or type errors raised by the kernels. */
void aenet_amd_batch_eval(aenet_amd_batch *batch, const double *d_coords,
                          const int *d_types, double *d_E_atoms,
                          double *d_E_struct, double *d_F, void *stream,
                          int *stat);
void aenet_amd_batch_check(aenet_amd_batch *batch, void *stream, int *stat);

/* Timing of the fused atom kernel: while enabled, every evaluation records
 * HIP events around the atom-kernel launch on its stream;
 * aenet_amd_batch_kernel_time returns the summed duration (ms) and number
 * of timed launches since the previous call (it waits for them). */
void aenet_amd_batch_set_timing(aenet_amd_batch *batch, int enable);
void aenet_amd_batch_kernel_time(aenet_amd_batch *batch, double *ms, int *count,
                                 int *stat);
/* the same for the neighbor-search launches (k_search_cells) of the timed
 * evaluations */
void aenet_amd_batch_search_time(aenet_amd_batch *batch, double *ms, int *count,
                                 int *stat);

/* Autotuner of the atom kernel's launch shape for this batch and model
 * (NULL: the potentials of aenet_load_potential): every feasible triple
 * (wavefronts per workgroup: 1, 2, 4, .. 16 for Chebyshev, .. 4 for
 * Behler-Parrinello; network weights staged in LDS or read through the
 * caches; network per wavefront or, for Chebyshev potentials whose species
 * share one network architecture, the networks of a workgroup's atoms as
 * one matrix-core tile) is timed with HIP events over `repeats` full
 * evaluations of the given device coordinates and types (with forces if
 * forces != 0), and the fastest is used by the batch's later evaluations
 * with that model.  Without it the launch shape comes from an occupancy
 * model, with the network per wavefront.  results (may be NULL, room for 32
 * rows): per candidate (waves, staged, tile, average ms); *ncand = their
 * number; chosen[0..2] = the kept (waves, staged, tile).
 * Synchronous.  No counterpart in the reference. */
void aenet_amd_batch_autotune(aenet_amd_batch *batch, const aenet_amd_model *model,
                              const double *d_coords, const int *d_types, int forces,
                              int repeats, double results[], int *ncand, int chosen[3],
                              void *stream, int *stat);

/* Sets the atom kernel's launch shape for this batch and model as the
 * autotuner would keep it: waves per workgroup (a power of two up to 16 for
 * Chebyshev, 4 for Behler-Parrinello), weights staged in LDS (0/1), tile
 * network (0/1; Chebyshev, one architecture for every species).  waves = 0
 * returns to the occupancy model.  AENET_AMD_ERR_ARG_VALUE for a shape out
 * of range or one that does not fit the device.  No counterpart in the
 * reference. */
void aenet_amd_batch_set_launch(aenet_amd_batch *batch, const aenet_amd_model *model,
                                int waves, int staged, int tile, int *stat);

/* Neighbor-count buckets of the atom kernel for this batch and model: the
 * atoms whose neighbor count n has caps[k-1] < n <= caps[k] are evaluated
 * by launch k, with LDS regions sized for caps[k] and the launch shape
 * shapes[3k .. 3k+2] = (waves, staged, tile) as in aenet_amd_batch_set_launch.
 * caps ascend; all but the last stay below the batch's neighbor capacity,
 * and the last becomes that capacity.  Each evaluation puts the atoms into
 * their buckets on the device (one small launch), so the counts may change
 * between evaluations.  n = 0 drops the plan; n <= 6.  The results equal
 * those of a single launch (per-atom energies and the 64-bit fixed-point
 * force sums do not depend on the launch shape without the tile network).
 * aenet_amd_batch_autotune makes such a plan itself, after timing the
 * single launch, where the batch has at least 4096 atoms, its neighbor
 * counts fill more than one bucket, and the buckets measure at least 2 %
 * faster.  aenet_amd_batch_set_launch drops the plan.  No counterpart in
 * the reference. */
void aenet_amd_batch_set_buckets(aenet_amd_batch *batch, const aenet_amd_model *model,
                                 int n, const int caps[], const int shapes[], int *stat);

/* The bucket plan in use: *n (0: none), caps[n], shapes[3n], atoms[n]
 * (atoms per bucket when the tuner measured), ms[n] (each
 * bucket's measured atom-kernel time) and totals = (single launch, bucket
 * launches with the bucketing kernel) in ms as the tuner measured them
 * (atoms -1 and times 0: set by the caller).  Arrays may be NULL; room for 6 buckets.  No
 * counterpart in the reference. */
void aenet_amd_batch_buckets(aenet_amd_batch *batch, const aenet_amd_model *model, int *n,
                             int caps[], int shapes[], int atoms[], double ms[], double totals[2],
                             int *stat);

/* The bucket caps the autotuner would take for these neighbor counts
 * (host arrays, no device needed): powers of two from 16 below nb_cap,
 * then nb_cap; a bucket with fewer than max(natoms/32, 2048) atoms joins
 * the next larger one.  *ncaps = 1: no buckets.  No counterpart in the
 * reference. */
void aenet_amd_bucket_caps(const int counts[], int natoms, int nb_cap, int caps[], int *ncaps);

/* Bit-exact neighbor lists (lcl_nbdist_cart, src/ext/lclist.f90:520-677)
 * of every atom of the batch, device pointers, capacity cap per atom:
 * d_count[natoms], d_coo[natoms*cap*3], d_dist/d_list/d_type[natoms*cap];
 * d_list holds 1-based atom indices within each structure. */
void aenet_amd_batch_neighbors(aenet_amd_batch *batch, const double *d_coords,
                               const int *d_types, int cap, int *d_count,
                               double *d_coo, double *d_dist, int *d_list,
                               int *d_type, void *stream, int *stat);

/* Descriptor values of every atom (d_G: natoms*stride doubles), unscaled
 * (stp_eval without normalization) or normalized (scaled != 0). */
void aenet_amd_batch_descriptors(aenet_amd_batch *batch, const double *d_coords,
                                 const int *d_types, int scaled, double *d_G,
                                 int stride, void *stream, int *stat);

/* ------------------------------------------------------------------
 * Independent potential sets.  The reference keeps one set of potentials
 * in module globals (src/aenet.f90:122-132); a caller that needs two at
 * once -- a cheap model and the full one in delayed-acceptance Monte
 * Carlo, or a network under training beside its teacher -- loads them as
 * handles.  Passing NULL as the model of an entry point below means the
 * set loaded with aenet_load_potential.
 * ------------------------------------------------------------------ */

/* files[t] is the .nn potential of species atom_types[t] (type id t+1). */
void aenet_amd_model_load(int ntypes, char *atom_types[], char *files[],
                          aenet_amd_model **model, int *stat);
void aenet_amd_model_destroy(aenet_amd_model *model);
void aenet_amd_model_info(const aenet_amd_model *model, double *Rc_max,
                          int *nsf_max, int *wsize, int *stat);
/* offset and length of species type_id's network in the weight array
 * (the .nn layout: per layer a column-major (n_out, n_in+1) block, bias
 * last) */
void aenet_amd_model_species_weights(const aenet_amd_model *model, int type_id,
                                     int *offset, int *size, int *stat);
/* d_W: wsize floats on the device; to_model = 0 copies the model's
 * weights out, 1 replaces them */
void aenet_amd_model_weights(aenet_amd_model *model, float *d_W, int to_model,
                             void *stream, int *stat);

/* Descriptor values and their derivatives for explicit environments
 * (what generate.x writes for force training, src/pytorchoutput.f90:105-139,
 * from sfderiv_i / sfderiv_j of stp_eval): environment e is the central
 * atom of species type_i[e] at coo_i[3e] with neighbors env_start[e] ..
 * env_start[e+1]-1 of coo_j / type_j.  Out (host, may be NULL): values
 * [nenv][stride] (unscaled), deriv_i [nenv][stride][3] (d/dR of the central
 * atom), deriv_j [neighbor][stride][3] in the neighbors' order.  stride >=
 * the model's largest feature count; features beyond a species' count are
 * zero.  FP32 derivative sums (relative error ~1e-6).  Synchronous. */
void aenet_amd_model_descriptor_derivatives(aenet_amd_model *model, int nenv,
                                            const int env_start[], const int type_i[],
                                            const double coo_i[], const double coo_j[],
                                            const int type_j[], int stride, double values[],
                                            double deriv_i[], double deriv_j[], int *stat);

/* a batch planned for the model's cutoff */
void aenet_amd_batch_create_model(const aenet_amd_model *model, int nstruct,
                                  const int natoms[], const double latt[],
                                  const int pbc[], const double coords_host[],
                                  aenet_amd_batch **batch, int *stat);
/* aenet_amd_batch_eval with a given model (its cutoff must not exceed the
 * batch's) */
void aenet_amd_batch_eval_model(aenet_amd_batch *batch, const aenet_amd_model *model,
                                const double *d_coords, const int *d_types,
                                double *d_E_atoms, double *d_E_struct, double *d_F,
                                void *stream, int *stat);
/* Atomic energies of a subset of atoms: d_active[0 .. *d_nactive-1]
 * (device), max_active >= *d_nactive.  Only those entries of d_E_atoms
 * are written.  With reuse_lists != 0 the neighbor lists built by the
 * previous evaluation of the same coordinates are used (the subset must
 * have been part of that evaluation).  For local energy changes in Monte
 * Carlo. */
void aenet_amd_batch_eval_active(aenet_amd_batch *batch, const aenet_amd_model *model,
                                 const double *d_coords, const int *d_types,
                                 const int *d_active, const int *d_nactive,
                                 int max_active, int reuse_lists, double *d_E_atoms,
                                 void *stream, int *stat);
/* Vector-Jacobian product of the atomic energies E_i of a batch, for
 * reverse-mode differentiation: d_grad_coords (3 per atom, overwritten)
 * = sum_i g_i dE_i/dR and, if d_grad_W is not NULL, d_grad_W (wsize
 * floats) += sum_i g_i dE_i/dW.  d_g: natoms doubles (device). */
void aenet_amd_batch_vjp(aenet_amd_batch *batch, const aenet_amd_model *model,
                         const double *d_coords, const int *d_types, const double *d_g,
                         double *d_grad_coords, float *d_grad_W, void *stream, int *stat);
void aenet_amd_batch_descriptors_model(aenet_amd_batch *batch, const aenet_amd_model *model,
                                       const double *d_coords, const int *d_types,
                                       int scaled, double *d_G, int stride,
                                       void *stream, int *stat);

/* Network training on normalized descriptors (d_G: natoms*stride, as
 * aenet_amd_batch_descriptors_model(scaled=1) writes them).  Computes the
 * atomic energies (d_E_out, optional) and adds to d_grad (wsize floats)
 * the gradient of sum_i g_i E_i with respect to the weights -- the
 * reference's ff_wderiv (src/ext/feedforward.f90:763-860) times E_scale,
 * contracted with g.  g_i = coef * weight_i * (E_i - E_ref_i) when d_E_ref
 * is given (then d_loss, optional, receives += sum weight_i (E_i -
 * E_ref_i)^2), else g_i = d_upstream[i].  d_grad NULL: energies only. */
void aenet_amd_model_net_grad(aenet_amd_model *model, int natoms, const double *d_G,
                              int stride, const int *d_types, const double *d_E_ref,
                              const double *d_weight, const double *d_upstream,
                              double coef, double *d_E_out, float *d_grad,
                              double *d_loss, void *stream, int *stat);
/* One Adam step on the model's weights (src/optimize.f90:1012-1046);
 * d_m, d_v: wsize floats, zero before the first step; t >= 1. */
void aenet_amd_model_adam(aenet_amd_model *model, const float *d_grad, float *d_m,
                          float *d_v, int t, double lr, double b1, double b2,
                          double eps, void *stream, int *stat);
/* synchronizes and reports errors raised by the training kernels */
void aenet_amd_model_check(aenet_amd_model *model, void *stream, int *stat);

/* ------------------------------------------------------------------
 * Metropolis Monte Carlo of independent periodic chains on the GPU: nchains
 * cells of natoms atoms (latt: 9 doubles, rows = lattice vectors, the
 * same cell for every chain; types and coords on the host, chains
 * concatenated).  Moves: swap of two atoms of different species
 * (probability p_swap) or displacement of one atom by U(-max_disp,
 * max_disp)^3; acceptance at inverse temperature beta (1/eV).  With a
 * first-stage potential (radial-only Chebyshev, small network) the run
 * can use delayed acceptance: proposals are screened with the first
 * stage's energy change (evaluated incrementally, no neighbor search), and
 * only those that pass are decided by the full potential with
 * min(1, exp(-beta (dE_full - dE_first))), which keeps the chain exact for
 * the full potential.  In a delayed-acceptance step each chain screens up
 * to max_screen proposals and sends the first that passes to the full
 * potential; those rejected by the first stage are complete steps of the
 * chain, so every chain follows sequential delayed acceptance while the
 * full potential gets a proposal from nearly every chain.  The full potential re-evaluates only the atoms
 * within its cutoff of the moved sites.  Each step's kernels can be
 * captured in a HIP graph and replayed.  No counterpart in the
 * reference. */
typedef struct aenet_amd_mc aenet_amd_mc;

void aenet_amd_mc_create(const aenet_amd_model *full, const aenet_amd_model *first,
                         int nchains, int natoms, const double latt[9], const int types[],
                         const double coords[], unsigned long long seed,
                         aenet_amd_mc **mc, int *stat);
void aenet_amd_mc_destroy(aenet_amd_mc *mc);
/* max_screen: first-stage tries per chain and delayed-acceptance step
 * (1 .. 1024; 1 screens one proposal per step) */
void aenet_amd_mc_set(aenet_amd_mc *mc, double beta, double p_swap, double max_disp,
                      int max_screen, int *stat);
/* energies (and first-stage descriptors) recomputed from scratch */
void aenet_amd_mc_refresh(aenet_amd_mc *mc, int *stat);
/* nsteps steps (synchronous): plain, one proposal per chain and step;
 * delayed != 0, two-stage, up to max_screen proposals per chain and step.
 * sample_every > 0 records every chain's energy after each sample_every-th
 * proposal of that chain (chains of a delayed run advance at different
 * rates) */
void aenet_amd_mc_run(aenet_amd_mc *mc, int nsteps, int delayed, int sample_every,
                      int use_graph, int *stat);
/* sums over chains: proposed, accepted, full-potential evaluations, swaps
 * proposed; reset != 0 clears them */
void aenet_amd_mc_stats(aenet_amd_mc *mc, long long out[4], int reset, int *stat);
void aenet_amd_mc_energies(aenet_amd_mc *mc, double E_chain[], int *stat);
/* out (may be NULL): nsamples*nchains energies of the last run, nsamples =
 * the samples every chain completed */
void aenet_amd_mc_samples(aenet_amd_mc *mc, double out[], long long *nsamples, int *stat);
void aenet_amd_mc_state(aenet_amd_mc *mc, double coords[], int types[], int *stat);
/* Decision record: max_props >= 0 arms it for the next runs (each chain's
 * proposals counted from the start of each run; 0 disarms); max_props < 0
 * copies it into out (may be NULL): per proposal index and chain 4
 * doubles -- first-stage energy change, full-potential change (0 if not
 * evaluated), stage-1 outcome (1 passed, 0 rejected, -1 no record),
 * accepted. */
void aenet_amd_mc_trace(aenet_amd_mc *mc, long long max_props, double out[], long long *nprops,
                        int *stat);
/* first-stage atomic energies and unscaled descriptors (nsf per atom) */
void aenet_amd_mc_first_stage_state(aenet_amd_mc *mc, double E_atoms[], double G[], int *nsf,
                                    int *stat);

/* Fit of a first-stage potential (radial-only Chebyshev, one hidden layer
 * of at most 16 nodes, linear output) to reference energies and forces of
 * ncells cells of natoms atoms (<= 256) with one lattice: d_coords
 * (3*ncells*natoms), d_types, d_E_ref (atomic energies), d_F_ref
 * (forces).  The cells d_sel[0 .. nsel-1] contribute; d_grad (wsize
 * floats) += gradient of coefE/2 sum_i (E_i - E_ref_i)^2 + coefF/2 sum_i
 * |F_i - F_ref_i|^2; d_loss[0], d_loss[1] += the two sums of squares.
 * The force term is differentiated through the network's input gradient
 * (forward tangent along the residual's pull-back, then reverse). */
void aenet_amd_model_fit_cells(aenet_amd_model *model, int ncells, int natoms, const double latt[9],
                               const double *d_coords, const int *d_types, const int *d_sel,
                               int nsel, const double *d_E_ref, const double *d_F_ref,
                               double coefE, double coefF, float *d_grad, double *d_loss,
                               void *stream, int *stat);

/* Text of the last error of the calling thread. */
const char *aenet_amd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif

// aenet_amd_internal.h -- data layouts shared by the host code and the
// gfx950 kernels.  Plain structs only; no torch types.
#pragma once

#include <stdint.h>

#define AENET_AMD_MAX_TYPES 8
#define AENET_AMD_MAX_LAYERS 8
#define AENET_AMD_MAX_HIDDEN 512     // hidden layer width (wider than a wavefront: in chunks of 64 nodes)
#define AENET_AMD_MAX_NSF 256        // descriptor length per species
#define AENET_AMD_NB_CAP 512         // neighbors held in LDS per atom
#define AENET_AMD_KEPT_CAP 512       // (cell, image) combos kept per atom
#define AENET_AMD_COMBO_CAP 2048     // (cell, image) combos per home cell

// error flags raised by kernels (bit mask in a device int)
#define AENET_AMD_FLAG_NB_OVERFLOW 1
#define AENET_AMD_FLAG_COMBO_OVERFLOW 2
#define AENET_AMD_FLAG_KEPT_OVERFLOW 4
#define AENET_AMD_FLAG_BAD_TYPE 8
#define AENET_AMD_FLAG_TRIPLET_OVERFLOW 16
#define AENET_AMD_FLAG_BAD_COORD 32

enum { DESC_CHEBYSHEV = 0, DESC_BEHLER = 1 };

// Per-species parameters (one per global atom type).  Offsets index the
// model's float weight array and double normalization arrays.
struct SpeciesDev {
  int nsf;
  int nlayers;
  int nnodes[AENET_AMD_MAX_LAYERS];
  int fa[AENET_AMD_MAX_LAYERS];    // activation of layer l (l >= 1)
  int w_off[AENET_AMD_MAX_LAYERS]; // start of the weight block of layer l -> l+1
  int norm_off;
  // Chebyshev basis (new_SFBasis layout; 0-based feature offsets)
  int r_N, a_N, multi;
  int r_i1, a_i1, r_i2, a_i2;
  float r_Rc, a_Rc;
  float spin[AENET_AMD_MAX_TYPES]; // typespin of neighbor global type t (index t-1)
  // Behler-Parrinello: row of the symmetry-function tables used for this
  // central species and the local type of each global neighbor type
  int bp_row;
  int bp_ltype[AENET_AMD_MAX_TYPES];
  double E_scale, E_shift, E_atom;
};

// Behler-Parrinello tables (sf_init layout, 0-based): per (row, local type2)
// a radial block, per (row, local type2 <= type3) an angular block.
struct BPDev {
  int ntypes;
  int rad_off[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];   // feature offset of radial block
  int rad_n[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];     // number of G2 (G1/G3 folded in kind)
  int rad_p0[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];    // first parameter record
  int ang_off[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];
  int ang_n[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];
  int ang_p0[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];
  float ang_rcmax[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];  // largest cutoff of the block
  int ang_g5[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];       // block holds G5 (no r_jk cutoff)
  int ang_one_seg[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];  // all functions: one cutoff, one kind
  // one segment and every lambda +-1: the block has BPAngK constants (at
  // ang_p0, sign-sorted) with ang_np of lambda = +1 first
  int ang_fast[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];
  int ang_np[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];
  // fast-path block whose functions are every combination of a set of
  // (lambda, zeta) pairs with a set of etas: 1 + its BPAngGrid record (0: none)
  int ang_grid[AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES * AENET_AMD_MAX_TYPES];
};

// one radial (G1/G2/G3) or angular (G4/G5) function record
struct BPFunc {
  int kind;      // 1..5
  float Rc, p1, p2, p3;   // G2: Rs, eta; G3: kappa; G4/G5: lambda, zeta, eta
};

// Constants of one G4/G5 function of a block the angular fast path takes,
// with arg = (1 + lambda cos) / 2:
//   G = 2 arg^zeta exp(-eta S) C = 2 arg 2^(A log2(arg) + B S) C
struct BPAngK {
  float A;    // zeta - 1
  float B;    // -eta log2(e)
  float K1;   // zeta lambda / 2: d/dcos of the function is K1 2^(..) C times 2
  float K2;   // -2 eta: d/d(r^2) of the exponent's eta S term
  int feat;   // the function's place in the block (feature ang_off + feat)
  int pad[3];
};

// Per-structure plan: lattice-derived data computed on the host with the
// reference's arithmetic, plus offsets into the batch's flat arrays.
struct StructDev {
  double L[9];       // lattice, column k = vector k
  double B[9];       // 2*pi-scaled reciprocal lattice (geo_recip_lattice)
  double shift[3];   // isolated structures: bounding-box origin
  double frac_div;   // 2*pi (Cartesian input) or 1 (lattice coordinates given)
  double Rc2;        // (Rc_max + 1e-3)^2
  double rc_list;    // Rc_max + 1e-3
  double cell_rad;   // circumradius of one grid cell (all cells of a structure alike)
  double cell_gap[3];  // spacing of the grid's planes along each lattice direction
  int pbc;
  int atom0, natoms;
  int nc[3];
  int cell0, ncells;
  int nT, T0;        // lattice translations (3 ints each) at T0
  int nC, C0;        // cell stencil vectors at C0
  int nself, self0;  // self images: Tvec ranks (1-based) and distances at self0
};

// sizes an atom kernel launch depends on
struct AtomSizes {
  int natoms;   // central atoms (structures) or environments
  int cap;      // neighbor capacity per atom
  int nsf_max;  // longest descriptor
  int wsize;    // total network weights
  int nl_max;   // most network layers (input layer included)
  int hmax;     // widest non-input layer
  // optional subset of central atoms (structures mode): atom indices
  // active[0 .. *nactive-1]; natoms is then the largest possible count
  const int *active;
  const int *nactive;
  // launch shape chosen by measurement (0 / -1: the occupancy model's)
  int tune_waves;    // wavefronts per workgroup
  int tune_staged;   // network weights staged in LDS (0 / 1)
  int nnorm;         // normalization entries of all species (staged with the weights)
  // every species has the same layer count and widths, one output node:
  // the networks of a workgroup's atoms can run as one matrix tile
  int net_uniform;
  int tune_tile;     // that tile network chosen by measurement (-1: not measured, off)
};

// network training (train.hip): per-atom normalized descriptors in, atomic
// energies and the weight gradient of sum_i g_i E_i out
struct NetGradArgs {
  const double *G;         // [natoms*stride] normalized descriptors G'
  int stride;
  const int *types;        // [natoms] 1-based species
  const double *E_ref;     // optional targets: g_i = coef * weight_i * (E_i - E_ref_i)
  const double *weight;    // optional per-atom weights (with E_ref)
  const double *upstream;  // g_i given directly (when E_ref is null)
  double coef;
  double *E_out;           // optional [natoms] predicted atomic energies
  float *grad;             // [wsize] gradient, accumulated
  double *loss;            // optional: += sum_i weight_i (E_i - E_ref_i)^2
  int *flags;              // optional error word
};

// An angular block whose G4 (or G5) functions are the nz x ne combinations
// of nz (lambda = +-1, zeta) pairs and ne etas, at one cutoff.  With arg =
// (1 + lambda cos) / 2 the function of pair iz and eta ie factorizes,
//   G = 2 arg^zeta exp(-eta S) C = Z_iz E_ie,
//   Z_iz = 2^(zeta log2(arg) + 1),  E_ie = 2^(-eta log2(e) S) C,
// so a triplet takes nz + ne exponentials instead of one per function.
// Slot s = iz BP_GRID_NE + ie holds the function's place in the block.
constexpr int BP_GRID_NZ = 8;
constexpr int BP_GRID_NE = 4;
struct BPAngGrid {
  int nz, ne;
  int nzp;                 // pairs with lambda = +1, which come first
  float zeta[BP_GRID_NZ];
  float K1[BP_GRID_NZ];    // zeta lambda / 2 (lambda = +1 for iz < nzp, else -1)
  float B[BP_GRID_NE];     // -eta log2(e)
  float K2[BP_GRID_NE];    // -2 eta
  int feat[BP_GRID_NZ * BP_GRID_NE];  // place in the block (feature ang_off + feat), -1: none
};

struct ModelDev {
  const SpeciesDev *species;  // [ntypes]
  const float *W;             // network weights
  const double *norm_avg;     // [sum nsf]
  const double *norm_scale;
  const BPDev *bp;
  const BPFunc *bpf;
  const BPAngK *bpk;  // parallel to bpf (angular fast-path blocks only)
  const BPAngGrid *bpg;  // grid blocks (BPDev::ang_grid)
  int ntypes;
  int desc;
};

// wavefronts (central atoms of one home cell) per workgroup of the
// cell-major neighbor search
constexpr int CS_WAVES = 8;

struct BatchDev {
  const StructDev *st;
  int nstruct;
  int natoms;
  const int *atom_struct;     // [natoms]
  const double *cart;         // [3*natoms]
  const int *types;           // [natoms], 1-based
  double *s;                  // [3*natoms] wrapped lattice coordinates
  int *cellof;                // [natoms] global cell id
  int ncells;                 // total cells
  const int *cell_struct;     // [ncells]
  const int *cell_ijk;        // [3*ncells] 1-based cell indices
  const double *cell_ctr;     // [3*ncells] Cartesian cell centres (fixed per plan)
  int *cell_count;            // [ncells]
  int *cell_fill;             // [ncells]
  int *cell_start;            // [ncells+1]
  int *cell_atoms;            // [natoms] atom ids, cells in order, descending id within a cell
  double *sorted_s;           // [3*natoms] lattice coordinates in cell_atoms order
  const int *combo_count;     // [ncells] (cell, image) combinations of each home cell
  const int *combos;          // [2*ncells*combo_cap] pairs (cell, image rank), visiting order
  int combo_cap;
  int kept_cap;               // largest combination count of a home cell (search LDS size)
  int combo_max;              // largest combination count of a home cell, unclamped (cell search LDS)
  int *cell_work;             // [4*work_cap] cell-major search items: (home cell, first atom, atoms, structure) or -1
  int work_cap;               // natoms/CS_WAVES + ncells + 1
  const int *Tvec;            // all lattice translations
  const int *Cvec;            // all stencil vectors
  const int *self_t;          // self-image ranks
  const double *self_d;       // self-image distances
  int *flags;
  int *fmark;                 // [natoms] a force contribution outside the fixed-point range (see scatter_forces)
  // per-atom slots for the workgroup force tables' sums (see the atom
  // kernels' table flush): fcnt[natoms] slots taken, fslot[natoms *
  // fslots * 3] 64-bit sums; fslots = 0: none (direct atomics)
  int *fcnt;
  unsigned long long *fslot;
  int fslots;
  // the BP kernel's per-wavefront triplet lists: trip[slot * trip_cap ..]
  // for wavefront slots below trip_slots (null: none)
  int *trip;
  int trip_cap;
  int trip_slots;
};

// compact neighbor lists written by the search kernel for the atom kernels
struct NbDev {
  int cap;
  int *count;          // [natoms]
  float *R;            // [4*natoms*cap]: R_ij (A) and |R_ij|
  int *j;              // [natoms*cap] global atom index of the neighbor
};

// outputs of the atom kernel
struct OutDev {
  double *E_atom;      // [natoms] atomic energies (incl. E_atom offset)
  double *F;           // [3*natoms] forces, accumulated with FP64 atomics
  double *G;           // optional [natoms*G_stride] descriptors (unscaled or scaled)
  int G_scaled;
  int G_stride;
  int desc_only;       // stop after the descriptors
  const double *g_atom;  // optional [natoms]: forces of sum_i g_i E_i instead of sum_i E_i
  int coef_unit;       // explicit environments: > 0 replaces the network's dE/dG of environment g
                       // by the unit vector of feature g mod coef_unit (descriptor derivatives)
  // neighbor-list mode (bit-exact lists)
  int nb_cap;
  int *nb_count;
  double *nb_coo;      // [natoms*nb_cap*3]
  double *nb_dist;
  int *nb_list;        // 1-based atom index within the structure
  int *nb_type;
};

// explicit environments (per-atom C-ABI entry points): neighbors given
struct EnvDev {
  int nenv;
  const int *env_start;   // [nenv+1]
  const int *env_type;    // [nenv] 1-based central type
  const double *coo_i;    // [3*nenv]
  const double *coo_j;    // [3*total]
  const int *type_j;      // [total]
  double *E;              // [nenv]
  double *dF_i;           // [3*nenv] force on the central atom
  double *dF_j;           // [3*total] force on each neighbor entry
};

// Monte Carlo of independent periodic chains (mc.hip): nchains cells of
// natoms atoms each, all with the same lattice
enum { MC_NONE = 0, MC_SWAP = 1, MC_DISP = 2 };

struct MCDev {
  int nchains, natoms;
  double L[9];      // rows: lattice vectors (the same cell for every chain)
  double Binv[9];   // inverse of L (fractional = v * Binv)
  double *coords;   // [3*nchains*natoms] Cartesian
  int *types;       // [nchains*natoms]
  unsigned long long *rng;  // [nchains] generator states
  double beta, p_swap, max_disp, rc_t;
  int max_screen;   // first-stage tries per chain and step (delayed acceptance)
  // the step's move per chain
  int *kind, *site_a, *site_b;
  double *old_pos;  // [3*nchains]
  double *u_stage1, *u_stage2;
  int *need;        // chain goes to the full potential
  double *dE_s;     // first-stage energy change
  // atoms re-evaluated by the full potential, one block per chain
  int *nactive, *active, *act_start, *act_count;
  // full potential: atomic energies (current, trial), chain energies
  double *E_t, *E_trial, *E_chain;
  // first-stage potential: unscaled descriptors and atomic energies
  int nsf_s, rn_s, nh_s;  // features per atom; largest radial term count and hidden width
  const void *fs_tab;     // first-stage table (mc.hip), built per run
  double *G_s, *G_trial, *E_s, *E_s_trial;
  int *affected;
  long long *stats;  // [4*nchains]: proposed, accepted, full evaluations, swaps
  long long *step;
  long long *nprop;  // [nchains] proposals resolved since the start of the run
  double *trace;     // optional [max_trace*nchains*4] per proposal: dE_first, dE_full, stage-1 outcome, accepted
  long long max_trace;
  double *samples;   // [max_samples*nchains] chain energies every sample_every proposals
  int sample_every;
  long long max_samples;
};

/*
 * aenet_amd.h -- batched entry points of libaenet_amd.so (MI355X, gfx950).
 *
 * These have no single counterpart in the reference: they batch what
 * predict.x's get_energy (src/predict.F90:574-706) does per atom -- neighbor
 * list, descriptors, network and forces for whole structures -- into one
 * device pass.  They use the potentials loaded with aenet_init /
 * aenet_load_potential (include/aenet.h).  Status codes are the AENET_*
 * values; codes 5-7 are this library's own (see AENET_AMD_ERR_*).
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

/* Energies and forces of nstruct structures given in host memory.
 * E_tot/E_coh: per structure (may be NULL); E_atoms: per atom (may be
 * NULL); F: 3 per atom (set, not accumulated; NULL for energies only).
 * Equivalent of calling predict.x's get_energy on each structure. */
void aenet_structures_energy_and_forces(int nstruct, const int natoms[],
                                        const double latt[], const int pbc[],
                                        const int types[], const double coords[],
                                        double E_tot[], double E_coh[],
                                        double E_atoms[], double F[], int *stat);

/* Device-resident batches: the structures' shapes (lattices, atom counts,
 * pbc) are fixed at creation; coordinates and types are device pointers
 * passed to each evaluation, so a caller that keeps its structures in HBM
 * (Monte Carlo, molecular dynamics, training) never copies them. */
typedef struct aenet_amd_batch aenet_amd_batch;

void aenet_amd_batch_create(int nstruct, const int natoms[], const double latt[],
                            const int pbc[], const double coords_host[],
                            aenet_amd_batch **batch, int *stat);
void aenet_amd_batch_destroy(aenet_amd_batch *batch);
int aenet_amd_batch_natoms(const aenet_amd_batch *batch);

/* d_* are device pointers; stream is a hipStream_t (NULL: the library's
 * stream).  d_E_struct holds 2*nstruct doubles: E_tot then E_coh.  d_F
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

/* Text of the last error of the calling thread. */
const char *aenet_amd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif

/*
 * aenet_oracle.h -- FP64 CPU restatement of aenet's energy/force path.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests hold the HIP implementation in
 * aenet_amd/ to this library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function follows the reference Fortran sources (ænet 2.0.4); the
 * file:line of the routine it restates is given next to each declaration.
 * Type ids and atom indices are 1-based as in the reference; arrays are
 * column-major (x,y,z interleaved per atom / neighbor).
 */
#ifndef AENET_ORACLE_H
#define AENET_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

enum { ORC_OK = 0, ORC_ERR_INIT = 1, ORC_ERR_MALLOC = 2, ORC_ERR_IO = 3,
       ORC_ERR_TYPE = 4, ORC_ERR_CAPACITY = 5, ORC_ERR_FORMAT = 6 };

/* aenet_init (src/aenet.f90:140-169); names are up to 2 characters */
orc_ctx *orc_create(int ntypes, const char *const *names);
void orc_destroy(orc_ctx *c);
/* aenet_load_potential (src/aenet.f90:250-291) + load_NNPot (src/potential.f90:111-157) */
int orc_load(orc_ctx *c, int type_id, const char *path);
int orc_all_loaded(const orc_ctx *c);
double orc_rc_min(const orc_ctx *c);
double orc_rc_max(const orc_ctx *c);
int orc_nnb_max(const orc_ctx *c);
int orc_nsf(const orc_ctx *c, int type_id);
int orc_wsize(const orc_ctx *c, int type_id);
double orc_e_atom(const orc_ctx *c, int type_id);

/* stp_eval (src/sfsetup.f90:887-979) without the running statistics; with
 * scaled != 0 the values and derivatives are normalized (stp_normalize,
 * src/sfsetup.f90:985-1035).  dGi: 3*nsf, dGj: 3*nsf*n_j (may be NULL). */
int orc_descriptors(orc_ctx *c, int type_i, const double coo_i[3], int n_j,
                    const double *coo_j, const int *type_j, int scaled,
                    double *G, double *dGi, double *dGj);

/* ff_eval + ff_deriv + ff_wderiv (src/ext/feedforward.f90:586-860) of the
 * network of type_id; y[1]; dy_dx[nx] and dy_dw[Wsize] may be NULL. */
int orc_network(orc_ctx *c, int type_id, const double *x, double *y,
                double *dy_dx, double *dy_dw);

/* aenet_atomic_energy (src/aenet.f90:359-408) */
int orc_atomic_energy(orc_ctx *c, const double coo_i[3], int type_i, int n_j,
                      const double *coo_j, const int *type_j, double *E_i);
/* aenet_atomic_energy_and_forces (src/aenet.f90:410-486); F is accumulated */
int orc_atomic_energy_and_forces(orc_ctx *c, const double coo_i[3], int type_i,
                                 int index_i, int n_j, const double *coo_j,
                                 const int *type_j, const int *index_j,
                                 int natoms, double *E_i, double *F);

/* geo_init_xsf Cartesian -> fractional (src/geometry.f90:661-736), with
 * geo_get_bounds for isolated structures (src/geometry.f90:538-580).
 * latt_out receives the lattice the neighbor list must use. */
void orc_cart_to_frac(const double latt[9], int pbc, int natoms,
                      const double *cart, double *frac, double latt_out[9]);

/* lcl_init + lcl_nbdist_cart for every atom (src/ext/lclist.f90:99-159,
 * 520-677).  frac is wrapped in place like cell_assign_atoms does.  Lists
 * are written with stride cap: nbcoo[3*cap*i], nbdist/nblist/nbtype[cap*i]. */
int orc_neighbors(double rc_min, double rc_max, const double latt[9], int pbc,
                  int natoms, const int *types, double *frac, int cap,
                  int *nnb, double *nbcoo, double *nbdist, int *nblist,
                  int *nbtype);

/* predict.x get_energy (src/predict.F90:574-706), serial; atoms with
 * (i-1) mod nranks == rank only (its round-robin split, :632).  E_at and F
 * may be NULL. */
int orc_structure(orc_ctx *c, const double latt[9], int pbc, int natoms,
                  const int *types, const double *cart, int rank, int nranks,
                  double *Ecoh, double *Etot, double *E_at, double *F);

/* lcl_nmax_nbdist (src/ext/lclist.f90:229-245) */
int orc_nmax_nbdist(double rmin, double rmax);

/* Adam step of src/optimize.f90:1012-1046 on n weights (t = step count,
 * 1-based after increment as in the reference). */
void orc_adam_step(int n, double *W, const double *g, double *m, double *v,
                   int t, double lr, double b1, double b2, double eps);

#ifdef __cplusplus
}
#endif
#endif

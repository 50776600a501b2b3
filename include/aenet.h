/*
 * aenet.h -- drop-in C interface of libaenet, implemented on AMD MI355X
 * (gfx950) by libaenet_amd.so.
 *
 * Every declaration below replaces the reference routine of the same name
 * (ænet 2.0.4); the reference interface it stands in for is cited next to
 * it.  Semantics follow the reference: type ids and atom indices are
 * 1-based, arrays are column-major (x,y,z interleaved per atom/neighbor),
 * forces are ACCUMULATED into F (never zeroed), energies include the free
 * atom energy, and status codes are the AENET_* values below.  Differences:
 * errors that abort the process in the reference return a status instead.
 */
#ifndef AENET_H_INCLUDED
#define AENET_H_INCLUDED

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/aenet.h:4, src/aenet.f90:171-193 */
void aenet_init(int ntypes, char *atom_types[], int *stat);
/* src/aenet.h:5, src/aenet.f90:197-222 */
void aenet_final(int *stat);
/* src/aenet.h:6, src/aenet.f90:228-244 */
void aenet_print_info(void);
/* src/aenet.h:7, src/aenet.f90:293-314 */
void aenet_load_potential(int type_id, char *filename, int *stat);
/* src/aenet.h:8, src/aenet.f90:316-331 */
_Bool aenet_all_loaded(void);

/* src/aenet.h:10, src/aenet.f90:337-346 (argument by value as the
 * reference header declares it) */
double aenet_free_atom_energy(int type_id);

/* src/aenet.h:12-14, src/aenet.f90:359-408 */
void aenet_atomic_energy(double coo_i[3], int type_i, int n_j,
                         double coo_j[], int type_j[], double *E_i,
                         int *stat);

/* src/aenet.h:16-19, src/aenet.f90:410-486 */
void aenet_atomic_energy_and_forces(double coo_i[3], int type_i, int index_i,
                                    int n_j, double coo_j[], int type_j[],
                                    int index_j[], int natoms, double *E_i,
                                    double F[], int *stat);

/* src/aenet.h:21-23, src/aenet.f90:525-553 */
void aenet_convert_atom_types(int ntypes_in, char *atom_types[],
                              int natoms_in, int type_id_in[],
                              int type_id_out[], int *stat);

/* src/aenet.h:25-27, src/aenet.f90:681-711 */
void aenet_sfb_init(int ntypes, char *atom_types[], int radial_order,
                    int angular_order, double radial_Rc, double angular_Rc,
                    int *stat);
/* src/aenet.h:29, src/aenet.f90:715-729 */
void aenet_sfb_final(int *stat);
/* src/aenet.h:31, src/aenet.f90:733-745 */
int aenet_sfb_nvalues(void);
/* src/aenet.h:33-34, src/aenet.f90:749-771 */
void aenet_sfb_eval(int itype0, double coo0[3], int nat, int itype1[],
                    double coo1[], int nv, double values[], int *stat);
/* src/aenet.h:36-37, src/aenet.f90:775-799 */
/* Not in the reference's header: sfb_eval with the derivatives it computes
 * internally (src/ext/sfbasis.f90:440-571) -- deriv0(3,nv) with respect to
 * the central atom and deriv1(3,nv,nat) with respect to each neighbor,
 * column-major as there (either may be NULL).  Values are unscaled, as in
 * aenet_sfb_eval. */
void aenet_sfb_eval_deriv(int itype0, double coo0[3], int nat, int itype1[],
                          double coo1[], int nv, double values[], double deriv0[],
                          double deriv1[], int *stat);
void aenet_sfb_reconstruct_radial(int nv, double values[], int nx,
                                  double x[], double y[], int *stat);

/* Exported by the reference library but absent from its header
 * (src/aenet.f90:571-639).  latt_vec: 9 doubles, column k = lattice vector
 * k; coords are replaced by the wrapped Cartesian coordinates; nnb is the
 * capacity on entry and the neighbor count on return. */
void aenet_nbl_init(double latt_vec[9], int natoms, int atom_types[],
                    double coords[], bool cartesian, bool pbc);
void aenet_nbl_final(void);
void aenet_nbl_neighbors(int iatom, int *nnb, double nbcoo[],
                         double nbdist[], int nblist[], int nbtype[]);

/* src/aenet.h:39-50, src/aenet.f90:101-118 */
extern int AENET_OK;
extern int AENET_ERR_INIT;
extern int AENET_ERR_MALLOC;
extern int AENET_ERR_IO;
extern int AENET_ERR_TYPE;
extern int AENET_TYPELEN;
extern int AENET_PATHLEN;
extern bool AENET_TRUE;
extern bool AENET_FALSE;

extern int aenet_nsf_max;
extern int aenet_nnb_max;
extern double aenet_Rc_min;
extern double aenet_Rc_max;

#ifdef __cplusplus
}
#endif

#endif

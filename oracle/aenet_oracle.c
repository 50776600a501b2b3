/*
 * aenet_oracle.c -- FP64 CPU restatement of aenet's energy/force path.
 *
 * TEST INFRASTRUCTURE ONLY: the tests hold the HIP path in aenet_amd/ to it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product never links it.  Compiled with -ffp-contract=off so
 * that the FP64 neighbor arithmetic is evaluated term by term, in the order
 * the reference Fortran writes it.
 *
 * Parity pins (tests/test_oracle.py): the reference library built from
 * its own sources (oracle/_ref), the golden energy of
 * python3/aenet/tests/test_ase.py:58, FCC neighbor counts of
 * src/tests/test_lclist.f90, closed forms and finite differences.
 */
#include "aenet_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* PI as written in src/ext/sfbasis.f90:72 and src/ext/symmfunc.f90:38 */
#define PI_SF 3.14159265358979
/* PI of src/constants.f90:30 (geometry) */
#define PI_GEO 3.141592653589793

/* ------------------------------------------------------------------ */
/* data types (src/ext/feedforward.f90:55-89, src/sfsetup.f90:81-154,  */
/* src/potential.f90:68-103, src/trainset.f90:86-127)                  */
/* ------------------------------------------------------------------ */

typedef struct {
  int nlayers, nnodes_max, Wsize, nvalues;
  int *nnodes; /* [nlayers] */
  int *f_a;    /* [nlayers]; f_a[l] is the activation of layer l (l>=1) */
  int *iw;     /* [nlayers] */
  int *iv;     /* [nlayers] */
  double *W;   /* [Wsize] */
} orc_net;

typedef struct {
  char atomtype[3];
  int nenv;
  char (*envtypes)[3];
  double Rc_min, Rc_max;
  char sftype[101];
  int nsf, nsfparam;
  int *sf;         /* [nsf] */
  double *sfparam; /* [nsfparam*nsf] column-major */
  int *sfenv;      /* [2*nsf] column-major */
  double *sfval_min, *sfval_max, *sfval_avg, *sfval_cov;
  int *ltype; /* [ntypes_global+1], 1-based global -> 1-based local, 0 if absent */
  int *gtype; /* [nenv+1] */
  /* Chebyshev basis (new_SFBasis, src/ext/sfbasis.f90:83-151) */
  int r_order, a_order, r_N, a_N, N;
  double r_Rc, a_Rc;
  int r_i1, r_f1, a_i1, a_f1, r_i2, r_f2, a_i2, a_f2;
  int multi;
  double *typespin; /* [nenv+1] indexed by local type */
} orc_setup;

typedef struct {
  int init;
  double E_scale, E_shift, E_atom;
  orc_net net;
  orc_setup stp;
  double ts_scale, ts_shift;
  int ts_ntypes;
} orc_pot;

/* Behler2011 tables (src/ext/symmfunc.f90:59-77, sf_init :91-140) */
typedef struct {
  int ntypes, npairs, ntriples, nGmax;
  int *idx2;   /* [ntypes*ntypes] */
  int *idx3;   /* [ntypes^3] */
  int *nG2;    /* [5*npairs] (funct-1 + 5*(pij-1)) */
  int *nG3;    /* [5*ntriples] */
  int *iG02;   /* [npairs] */
  int *iG03;   /* [ntriples] */
  double *pG1; /* [nGmax*npairs] */
  double *pG2; /* [3*nGmax*npairs] */
  double *pG3; /* [2*nGmax*npairs] */
  double *pG4; /* [4*nGmax*ntriples] */
  double *pG5; /* [4*nGmax*ntriples] */
  int *nG_type;
} orc_bp;

struct orc_ctx {
  int ntypes;
  char (*names)[3];
  orc_pot *pot;
  int loaded;
  double Rc_min, Rc_max;
  int nnb_max, nsf_max;
  int is_bp;
  orc_bp bp;
};

/* ------------------------------------------------------------------ */
/* Fortran unformatted sequential records                             */
/* ------------------------------------------------------------------ */

typedef struct {
  unsigned char *buf;
  size_t len, pos;
  int bad;
} recfile;

static const unsigned char *rec_next(recfile *f, size_t *n) {
  int32_t a, b;
  if (f->bad || f->pos + 4 > f->len) { f->bad = 1; return NULL; }
  memcpy(&a, f->buf + f->pos, 4);
  if (a < 0 || f->pos + 8 + (size_t)a > f->len) { f->bad = 1; return NULL; }
  memcpy(&b, f->buf + f->pos + 4 + a, 4);
  if (a != b) { f->bad = 1; return NULL; }
  const unsigned char *p = f->buf + f->pos + 4;
  f->pos += 8 + (size_t)a;
  *n = (size_t)a;
  return p;
}

static int rec_ints(recfile *f, int *out, int n) {
  size_t len;
  const unsigned char *p = rec_next(f, &len);
  if (!p || len != (size_t)n * 4) { f->bad = 1; return -1; }
  memcpy(out, p, len);
  return 0;
}

static int rec_dbls(recfile *f, double *out, int n) {
  size_t len;
  const unsigned char *p = rec_next(f, &len);
  if (!p || len != (size_t)n * 8) { f->bad = 1; return -1; }
  memcpy(out, p, len);
  return 0;
}

/* character(len=k) record: copied, trailing blanks stripped */
static int rec_chars(recfile *f, char *out, int k, int count) {
  size_t len;
  const unsigned char *p = rec_next(f, &len);
  if (!p || len != (size_t)k * count) { f->bad = 1; return -1; }
  for (int c = 0; c < count; ++c) {
    char *o = out + (size_t)c * (k + 1);
    memcpy(o, p + (size_t)c * k, k);
    o[k] = 0;
    for (int i = k - 1; i >= 0 && (o[i] == ' ' || o[i] == 0); --i) o[i] = 0;
  }
  return 0;
}

static int rec_skip(recfile *f) {
  size_t len;
  return rec_next(f, &len) ? 0 : -1;
}

static char *trim_dup(const char *s) {
  while (*s == ' ') ++s;
  return (char *)s;
}

static int name_eq(const char *a, const char *b) {
  a = trim_dup(a);
  b = trim_dup(b);
  size_t la = strlen(a), lb = strlen(b);
  while (la && a[la - 1] == ' ') --la;
  while (lb && b[lb - 1] == ' ') --lb;
  return la == lb && strncmp(a, b, la) == 0;
}

/* ------------------------------------------------------------------ */
/* potential loading                                                  */
/* ------------------------------------------------------------------ */

static void free_pot(orc_pot *p) {
  free(p->net.nnodes); free(p->net.f_a); free(p->net.iw); free(p->net.iv);
  free(p->net.W);
  free(p->stp.envtypes); free(p->stp.sf); free(p->stp.sfparam);
  free(p->stp.sfenv); free(p->stp.sfval_min); free(p->stp.sfval_max);
  free(p->stp.sfval_avg); free(p->stp.sfval_cov); free(p->stp.ltype);
  free(p->stp.gtype); free(p->stp.typespin);
  memset(p, 0, sizeof(*p));
}

/* new_SFBasis (src/ext/sfbasis.f90:83-151) for the Chebyshev setup
 * (setup_basis_chebyshev, src/sfsetup.f90:1162-1179) */
static void setup_chebyshev(orc_setup *s) {
  s->r_Rc = s->sfparam[0];
  s->r_order = (int)lround(s->sfparam[1]);
  s->a_Rc = s->sfparam[2];
  s->a_order = (int)lround(s->sfparam[3]);
  s->r_N = s->r_order + 1;
  s->a_N = s->a_order + 1;
  s->N = s->r_N + s->a_N;
  s->r_i1 = 1;
  s->r_f1 = s->r_i1 + s->r_N - 1;
  s->a_i1 = s->r_f1 + 1;
  s->a_f1 = s->a_i1 + s->a_N - 1;
  s->r_i2 = s->a_f1 + 1;
  s->r_f2 = s->r_i2 + s->r_N - 1;
  s->a_i2 = s->r_f2 + 1;
  s->a_f2 = s->a_i2 + s->a_N - 1;
  s->multi = s->nenv > 1;
  if (s->multi) s->N *= 2;
  s->typespin = (double *)calloc((size_t)s->nenv + 1, sizeof(double));
  int sp = -s->nenv / 2;
  for (int i = 1; i <= s->nenv; ++i) {
    if (sp == 0 && s->nenv % 2 == 0) sp += 1;
    s->typespin[i] = (double)sp;
    sp += 1;
  }
}

static int load_pot(orc_ctx *c, orc_pot *p, const char *path) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return ORC_ERR_IO;
  fseek(fp, 0, SEEK_END);
  long sz = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  recfile f = {0};
  f.buf = (unsigned char *)malloc((size_t)sz);
  f.len = (size_t)sz;
  if (fread(f.buf, 1, f.len, fp) != f.len) { fclose(fp); free(f.buf); return ORC_ERR_IO; }
  fclose(fp);
  memset(p, 0, sizeof(*p));
  orc_net *n = &p->net;
  /* load_Network, src/ext/feedforward.f90:291-306 */
  rec_ints(&f, &n->nlayers, 1);
  rec_ints(&f, &n->nnodes_max, 1);
  rec_ints(&f, &n->Wsize, 1);
  rec_ints(&f, &n->nvalues, 1);
  if (f.bad || n->nlayers < 2 || n->nlayers > 64 || n->Wsize <= 0) { free(f.buf); return ORC_ERR_FORMAT; }
  n->nnodes = (int *)calloc(n->nlayers, sizeof(int));
  n->f_a = (int *)calloc(n->nlayers, sizeof(int));
  n->iw = (int *)calloc(n->nlayers, sizeof(int));
  n->iv = (int *)calloc(n->nlayers, sizeof(int));
  n->W = (double *)calloc(n->Wsize, sizeof(double));
  rec_ints(&f, n->nnodes, n->nlayers);
  rec_ints(&f, n->f_a + 1, n->nlayers - 1);
  rec_ints(&f, n->iw, n->nlayers);
  rec_ints(&f, n->iv, n->nlayers);
  rec_dbls(&f, n->W, n->Wsize);
  /* load_Setup, src/sfsetup.f90:585-617 */
  orc_setup *s = &p->stp;
  rec_skip(&f); /* description */
  rec_chars(&f, s->atomtype, 2, 1);
  rec_ints(&f, &s->nenv, 1);
  if (f.bad || s->nenv < 1 || s->nenv > 64) { free(f.buf); free_pot(p); return ORC_ERR_FORMAT; }
  s->envtypes = (char(*)[3])calloc(s->nenv, 3);
  rec_chars(&f, (char *)s->envtypes, 2, s->nenv);
  rec_dbls(&f, &s->Rc_min, 1);
  rec_dbls(&f, &s->Rc_max, 1);
  rec_chars(&f, s->sftype, 100, 1);
  rec_ints(&f, &s->nsf, 1);
  rec_ints(&f, &s->nsfparam, 1);
  if (f.bad || s->nsf < 1 || s->nsfparam < 1) { free(f.buf); free_pot(p); return ORC_ERR_FORMAT; }
  s->sf = (int *)calloc(s->nsf, sizeof(int));
  s->sfparam = (double *)calloc((size_t)s->nsfparam * s->nsf, sizeof(double));
  s->sfenv = (int *)calloc(2 * (size_t)s->nsf, sizeof(int));
  s->sfval_min = (double *)calloc(s->nsf, sizeof(double));
  s->sfval_max = (double *)calloc(s->nsf, sizeof(double));
  s->sfval_avg = (double *)calloc(s->nsf, sizeof(double));
  s->sfval_cov = (double *)calloc(s->nsf, sizeof(double));
  rec_ints(&f, s->sf, s->nsf);
  rec_dbls(&f, s->sfparam, s->nsfparam * s->nsf);
  rec_ints(&f, s->sfenv, 2 * s->nsf);
  int neval;
  rec_ints(&f, &neval, 1);
  rec_dbls(&f, s->sfval_min, s->nsf);
  rec_dbls(&f, s->sfval_max, s->nsf);
  rec_dbls(&f, s->sfval_avg, s->nsf);
  rec_dbls(&f, s->sfval_cov, s->nsf);
  /* load_TrnSet_info, src/trainset.f90:430-440 */
  rec_skip(&f); /* file */
  rec_skip(&f); /* normalized */
  rec_dbls(&f, &p->ts_scale, 1);
  rec_dbls(&f, &p->ts_shift, 1);
  rec_ints(&f, &p->ts_ntypes, 1);
  if (f.bad || p->ts_ntypes < 1 || p->ts_ntypes > 1024) { free(f.buf); free_pot(p); return ORC_ERR_FORMAT; }
  char(*tsnames)[3] = (char(*)[3])calloc(p->ts_ntypes, 3);
  double *E_atom = (double *)calloc(p->ts_ntypes, sizeof(double));
  rec_chars(&f, (char *)tsnames, 2, p->ts_ntypes);
  rec_dbls(&f, E_atom, p->ts_ntypes);
  rec_skip(&f); rec_skip(&f); rec_skip(&f);
  int bad = f.bad;
  free(f.buf);
  if (bad) { free(tsnames); free(E_atom); free_pot(p); return ORC_ERR_FORMAT; }
  /* load_NNPot, src/potential.f90:145-151 */
  p->E_scale = 1.0 / p->ts_scale;
  p->E_shift = p->ts_shift;
  int it = 0;
  for (int i = 0; i < p->ts_ntypes; ++i)
    if (name_eq(tsnames[i], s->atomtype)) { it = i + 1; break; }
  p->E_atom = it ? E_atom[it - 1] : 0.0;
  free(tsnames);
  free(E_atom);
  /* stp_set_global_types, src/sfsetup.f90:721-755 */
  s->ltype = (int *)calloc((size_t)c->ntypes + 1, sizeof(int));
  s->gtype = (int *)calloc((size_t)s->nenv + 1, sizeof(int));
  for (int i = 1; i <= c->ntypes; ++i)
    for (int j = 1; j <= s->nenv; ++j)
      if (name_eq(c->names[i - 1], s->envtypes[j - 1])) { s->ltype[i] = j; break; }
  for (int i = 1; i <= s->nenv; ++i)
    for (int j = 1; j <= c->ntypes; ++j)
      if (name_eq(c->names[j - 1], s->envtypes[i - 1])) { s->gtype[i] = j; break; }
  p->init = 1;
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Behler2011 set-up (sf_init, sf_add_rad/ang, sf_build_index)         */
/* src/ext/symmfunc.f90:91-305, src/sfsetup.f90:1287-1347              */
/* ------------------------------------------------------------------ */

static void bp_free(orc_bp *b) {
  free(b->idx2); free(b->idx3); free(b->nG2); free(b->nG3); free(b->iG02);
  free(b->iG03); free(b->pG1); free(b->pG2); free(b->pG3); free(b->pG4);
  free(b->pG5); free(b->nG_type);
  memset(b, 0, sizeof(*b));
}

#define IDX2(b, i, j) (b)->idx2[((i)-1) * (b)->ntypes + ((j)-1)]
#define IDX3(b, i, j, k) (b)->idx3[(((i)-1) * (b)->ntypes + ((j)-1)) * (b)->ntypes + ((k)-1)]

static void bp_build_index(orc_bp *b) {
  for (int t1 = 1; t1 <= b->ntypes; ++t1) {
    int i = 0;
    for (int t2 = 1; t2 <= b->ntypes; ++t2) {
      int pij = IDX2(b, t1, t2);
      b->iG02[pij - 1] = i;
      for (int f = 0; f < 5; ++f) i += b->nG2[f + 5 * (pij - 1)];
      for (int t3 = t2; t3 <= b->ntypes; ++t3) {
        int tijk = IDX3(b, t1, t2, t3);
        b->iG03[tijk - 1] = i;
        for (int f = 0; f < 5; ++f) i += b->nG3[f + 5 * (tijk - 1)];
      }
    }
  }
}

static int bp_init(orc_ctx *c) {
  orc_bp *b = &c->bp;
  int nt = c->ntypes, nGmax = 0;
  for (int t = 0; t < nt; ++t)
    if (c->pot[t].stp.nsf > nGmax) nGmax = c->pot[t].stp.nsf;
  b->ntypes = nt;
  b->nGmax = nGmax;
  b->idx2 = (int *)calloc((size_t)nt * nt, sizeof(int));
  b->idx3 = (int *)calloc((size_t)nt * nt * nt, sizeof(int));
  int i2 = 0, i3 = 0;
  for (int i = 1; i <= nt; ++i)
    for (int j = 1; j <= nt; ++j) {
      IDX2(b, i, j) = ++i2;
      IDX3(b, i, j, j) = ++i3;
      for (int k = j + 1; k <= nt; ++k) {
        ++i3;
        IDX3(b, i, j, k) = i3;
        IDX3(b, i, k, j) = i3;
      }
    }
  b->npairs = i2;
  b->ntriples = i3;
  b->nG_type = (int *)calloc(nt + 1, sizeof(int));
  b->nG2 = (int *)calloc(5 * (size_t)i2, sizeof(int));
  b->nG3 = (int *)calloc(5 * (size_t)i3, sizeof(int));
  b->iG02 = (int *)calloc(i2, sizeof(int));
  b->iG03 = (int *)calloc(i3, sizeof(int));
  b->pG1 = (double *)calloc((size_t)nGmax * i2, sizeof(double));
  b->pG2 = (double *)calloc(3 * (size_t)nGmax * i2, sizeof(double));
  b->pG3 = (double *)calloc(2 * (size_t)nGmax * i2, sizeof(double));
  b->pG4 = (double *)calloc(4 * (size_t)nGmax * i3, sizeof(double));
  b->pG5 = (double *)calloc(4 * (size_t)nGmax * i3, sizeof(double));
  /* setup_symmfunc_Behler2011(itype, stp): type1 = global potential index */
  for (int t1 = 1; t1 <= nt; ++t1) {
    orc_setup *s = &c->pot[t1 - 1].stp;
    for (int isf = 0; isf < s->nsf; ++isf) {
      int funct = s->sf[isf];
      const double *pp = s->sfparam + (size_t)s->nsfparam * isf;
      int t2 = s->sfenv[2 * isf], t3 = s->sfenv[2 * isf + 1];
      if (funct < 1 || funct > 5 || t2 < 1 || t2 > nt) return ORC_ERR_FORMAT;
      if (funct <= 3) {
        int pij = IDX2(b, t1, t2);
        int iG = b->nG2[funct - 1 + 5 * (pij - 1)];
        if (iG >= nGmax) return ORC_ERR_FORMAT;
        b->nG2[funct - 1 + 5 * (pij - 1)] = iG + 1;
        b->nG_type[t1] += 1;
        if (funct == 1) b->pG1[iG + nGmax * (pij - 1)] = pp[0];
        if (funct == 2) {
          double *q = b->pG2 + 3 * ((size_t)iG + nGmax * (pij - 1));
          q[0] = pp[0]; q[1] = pp[1]; q[2] = pp[2];
        }
        if (funct == 3) {
          double *q = b->pG3 + 2 * ((size_t)iG + nGmax * (pij - 1));
          q[0] = pp[0]; q[1] = pp[1];
        }
      } else {
        if (t3 < 1 || t3 > nt) return ORC_ERR_FORMAT;
        int tijk = IDX3(b, t1, t2, t3);
        int iG = b->nG3[funct - 1 + 5 * (tijk - 1)];
        if (iG >= nGmax) return ORC_ERR_FORMAT;
        b->nG3[funct - 1 + 5 * (tijk - 1)] = iG + 1;
        b->nG_type[t1] += 1;
        double *q = (funct == 4 ? b->pG4 : b->pG5) + 4 * ((size_t)iG + nGmax * (tijk - 1));
        q[0] = pp[0]; q[1] = pp[1]; q[2] = pp[2]; q[3] = pp[3];
      }
      bp_build_index(b);
    }
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* context                                                            */
/* ------------------------------------------------------------------ */

orc_ctx *orc_create(int ntypes, const char *const *names) {
  if (ntypes < 1) return NULL;
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->ntypes = ntypes;
  c->names = (char(*)[3])calloc(ntypes, 3);
  for (int i = 0; i < ntypes; ++i) {
    strncpy(c->names[i], names[i], 2);
    c->names[i][2] = 0;
  }
  c->pot = (orc_pot *)calloc(ntypes, sizeof(orc_pot));
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  for (int i = 0; i < c->ntypes; ++i) free_pot(&c->pot[i]);
  bp_free(&c->bp);
  free(c->pot);
  free(c->names);
  free(c);
}

int orc_nmax_nbdist(double rmin, double rmax) {
  double V_atom = pow(0.5 * rmin, 3);
  double V_cut = pow(rmax + 0.5 * rmin, 3);
  return (int)ceil(V_cut / V_atom * 0.7405);
}

int orc_load(orc_ctx *c, int type_id, const char *path) {
  if (!c) return ORC_ERR_INIT;
  if (type_id < 1 || type_id > c->ntypes) return ORC_ERR_TYPE;
  FILE *fp = fopen(path, "rb");
  if (!fp) return ORC_ERR_IO;
  fclose(fp);
  if (c->pot[type_id - 1].init) free_pot(&c->pot[type_id - 1]);
  int st = load_pot(c, &c->pot[type_id - 1], path);
  if (st) return st;
  if (orc_all_loaded(c)) {
    /* pot_get_range (src/potential.f90:244-266) */
    c->Rc_min = c->pot[0].stp.Rc_min;
    c->Rc_max = c->pot[0].stp.Rc_max;
    for (int t = 0; t < c->ntypes; ++t) {
      if (c->pot[t].stp.Rc_min < c->Rc_min) c->Rc_min = c->pot[t].stp.Rc_min;
      if (c->pot[t].stp.Rc_max > c->Rc_max) c->Rc_max = c->pot[t].stp.Rc_max;
    }
    c->nnb_max = orc_nmax_nbdist(c->Rc_min, c->Rc_max);
    /* stp_init (src/sfsetup.f90:771-826) */
    c->nsf_max = 0;
    for (int t = 0; t < c->ntypes; ++t)
      if (c->pot[t].stp.nsf > c->nsf_max) c->nsf_max = c->pot[t].stp.nsf;
    c->is_bp = 0;
    int ncheb = 0;
    for (int t = 0; t < c->ntypes; ++t) {
      orc_setup *s = &c->pot[t].stp;
      char low[101];
      for (int i = 0; i < 101; ++i) low[i] = (char)((s->sftype[i] >= 'A' && s->sftype[i] <= 'Z') ? s->sftype[i] + 32 : s->sftype[i]);
      if (strcmp(low, "chebyshev") == 0) {
        if (!s->typespin) setup_chebyshev(s);
        ++ncheb;
      } else if (strcmp(low, "behler2011") == 0) {
        c->is_bp = 1;
      } else {
        return ORC_ERR_FORMAT;
      }
    }
    if (ncheb && c->is_bp) return ORC_ERR_FORMAT;
    if (c->is_bp) {
      bp_free(&c->bp);
      st = bp_init(c);
      if (st) return st;
    }
    c->loaded = 1;
  }
  return ORC_OK;
}

int orc_all_loaded(const orc_ctx *c) {
  for (int t = 0; t < c->ntypes; ++t)
    if (!c->pot[t].init) return 0;
  return 1;
}
double orc_rc_min(const orc_ctx *c) { return c->Rc_min; }
double orc_rc_max(const orc_ctx *c) { return c->Rc_max; }
int orc_nnb_max(const orc_ctx *c) { return c->nnb_max; }
int orc_nsf(const orc_ctx *c, int t) { return c->pot[t - 1].stp.nsf; }
int orc_wsize(const orc_ctx *c, int t) { return c->pot[t - 1].net.Wsize; }
double orc_e_atom(const orc_ctx *c, int t) { return c->pot[t - 1].E_atom; }

/* ------------------------------------------------------------------ */
/* Chebyshev basis                                                    */
/* ------------------------------------------------------------------ */

/* chebyshev_polynomial (src/ext/chebyshev.f90:41-75) */
static void cheb_T(double r, double r0, double r1, int n, double *T) {
  double x = (2.0 * r - r0 - r1) / (r1 - r0);
  T[0] = 1.0;
  if (n > 0) {
    T[1] = x;
    for (int i = 2; i <= n; ++i) T[i] = 2.0 * x * T[i - 1] - T[i - 2];
  }
}

/* chebyshev_polynomial_d1 (src/ext/chebyshev.f90:81-128) */
static void cheb_dT(double r, double r0, double r1, int n, double *dT) {
  double x = (2.0 * r - r0 - r1) / (r1 - r0);
  dT[0] = 0.0;
  if (n > 0) {
    double U1 = 1.0, U2 = 2.0 * x, U3;
    dT[1] = U1;
    for (int i = 2; i <= n; ++i) {
      dT[i] = U2 * (double)i;
      U3 = 2.0 * x * U2 - U1;
      U1 = U2;
      U2 = U3;
    }
  }
  for (int i = 0; i <= n; ++i) dT[i] = dT[i] * 2.0 / (r1 - r0);
}

/* sfb_fc, sfb_fc_d1 (src/ext/sfbasis.f90:719-752) */
static double sfb_fc(double R, double Rc) {
  return R >= Rc ? 0.0 : 0.5 * (cos(PI_SF / Rc * R) + 1.0);
}
static double sfb_fc_d1(double R, double Rc) {
  if (R >= Rc) return 0.0;
  double a = PI_SF / Rc;
  return -0.5 * a * sin(a * R);
}

/* sfb_eval (src/ext/sfbasis.f90:440-571), values and optional derivatives.
 * itype1 are setup-local type ids (1-based). */
static void sfb_eval(const orc_setup *s, const double *coo0, int nat,
                     const int *itype1, const double *coo1, double *values,
                     double *d0, double *d1) {
  const double EPS = 1.0e-12;
  int nv = s->N, nmax = s->r_N > s->a_N ? s->r_N : s->a_N;
  double *v = (double *)malloc(sizeof(double) * nmax);
  double *f = (double *)malloc(sizeof(double) * nmax);
  double *df = (double *)malloc(sizeof(double) * nmax);
  double *di = (double *)malloc(sizeof(double) * 3 * nmax);
  double *dj = (double *)malloc(sizeof(double) * 3 * nmax);
  double *dk = (double *)malloc(sizeof(double) * 3 * nmax);
  int deriv = d0 && d1;
  if (deriv) {
    memset(d0, 0, sizeof(double) * 3 * nv);
    memset(d1, 0, sizeof(double) * 3 * nv * (size_t)nat);
  }
  memset(values, 0, sizeof(double) * nv);
  double s_j = 1.0, s_k;
  for (int j = 0; j < nat; ++j) {
    double Rij[3] = {coo1[3 * j] - coo0[0], coo1[3 * j + 1] - coo0[1], coo1[3 * j + 2] - coo0[2]};
    double dij = sqrt(Rij[0] * Rij[0] + Rij[1] * Rij[1] + Rij[2] * Rij[2]);
    if (dij <= s->r_Rc && dij > EPS) {
      /* sfb_radial (src/ext/sfbasis.f90:762-793) */
      int N = s->r_N;
      double w = sfb_fc(dij, s->r_Rc);
      cheb_T(dij, 0.0, s->r_Rc, s->r_order, f);
      for (int n = 0; n < N; ++n) v[n] = w * f[n];
      if (deriv) {
        double dw = sfb_fc_d1(dij, s->r_Rc);
        cheb_dT(dij, 0.0, s->r_Rc, s->r_order, df);
        for (int n = 0; n < N; ++n)
          for (int x = 0; x < 3; ++x) {
            di[3 * n + x] = -Rij[x] / dij * (dw * f[n] + w * df[n]);
            dj[3 * n + x] = -di[3 * n + x];
          }
      }
      for (int n = 0; n < N; ++n) values[s->r_i1 - 1 + n] += v[n];
      if (deriv)
        for (int n = 0; n < N; ++n)
          for (int x = 0; x < 3; ++x) {
            d0[3 * (s->r_i1 - 1 + n) + x] += di[3 * n + x];
            d1[3 * (s->r_i1 - 1 + n) + x + 3 * (size_t)nv * j] += dj[3 * n + x];
          }
      if (s->multi) {
        s_j = s->typespin[itype1[j]];
        for (int n = 0; n < N; ++n) values[s->r_i2 - 1 + n] += s_j * v[n];
        if (deriv)
          for (int n = 0; n < N; ++n)
            for (int x = 0; x < 3; ++x) {
              d0[3 * (s->r_i2 - 1 + n) + x] += s_j * di[3 * n + x];
              d1[3 * (s->r_i2 - 1 + n) + x + 3 * (size_t)nv * j] += s_j * dj[3 * n + x];
            }
      }
    }
    if (dij > s->a_Rc) continue;
    for (int k = j + 1; k < nat; ++k) {
      double Rik[3] = {coo1[3 * k] - coo0[0], coo1[3 * k + 1] - coo0[1], coo1[3 * k + 2] - coo0[2]};
      double dik = sqrt(Rik[0] * Rik[0] + Rik[1] * Rik[1] + Rik[2] * Rik[2]);
      if (dik > s->a_Rc || dik < EPS) continue;
      double cosv = (Rij[0] * Rik[0] + Rij[1] * Rik[1] + Rij[2] * Rik[2]) / (dij * dik);
      /* sfb_angular (src/ext/sfbasis.f90:797-858) */
      int N = s->a_N;
      double fcj = sfb_fc(dij, s->a_Rc), fck = sfb_fc(dik, s->a_Rc);
      double w = fcj * fck;
      cheb_T(cosv, -1.0, 1.0, s->a_order, f);
      for (int n = 0; n < N; ++n) v[n] = w * f[n];
      if (deriv) {
        double dfcj = sfb_fc_d1(dij, s->a_Rc), dfck = sfb_fc_d1(dik, s->a_Rc);
        cheb_dT(cosv, -1.0, 1.0, s->a_order, df);
        double id_ij2 = 1.0 / (dij * dij), id_ik2 = 1.0 / (dik * dik), id_ijik = 1.0 / (dij * dik);
        double dic[3], djc[3], dkc[3], diw[3], djw[3], dkw[3];
        for (int x = 0; x < 3; ++x) {
          dic[x] = cosv * (Rij[x] * id_ij2 + Rik[x] * id_ik2) - (Rij[x] + Rik[x]) * id_ijik;
          djc[x] = -cosv * Rij[x] * id_ij2 + Rik[x] * id_ijik;
          dkc[x] = -cosv * Rik[x] * id_ik2 + Rij[x] * id_ijik;
          diw[x] = -(dfcj * fck * Rij[x] / dij + fcj * dfck * Rik[x] / dik);
          djw[x] = dfcj * fck * Rij[x] / dij;
          dkw[x] = fcj * dfck * Rik[x] / dik;
        }
        for (int n = 0; n < N; ++n)
          for (int x = 0; x < 3; ++x) {
            di[3 * n + x] = diw[x] * f[n] + w * df[n] * dic[x];
            dj[3 * n + x] = djw[x] * f[n] + w * df[n] * djc[x];
            dk[3 * n + x] = dkw[x] * f[n] + w * df[n] * dkc[x];
          }
      }
      for (int n = 0; n < N; ++n) values[s->a_i1 - 1 + n] += v[n];
      if (deriv)
        for (int n = 0; n < N; ++n)
          for (int x = 0; x < 3; ++x) {
            d0[3 * (s->a_i1 - 1 + n) + x] += di[3 * n + x];
            d1[3 * (s->a_i1 - 1 + n) + x + 3 * (size_t)nv * j] += dj[3 * n + x];
            d1[3 * (s->a_i1 - 1 + n) + x + 3 * (size_t)nv * k] += dk[3 * n + x];
          }
      if (s->multi) {
        s_k = s->typespin[itype1[k]];
        for (int n = 0; n < N; ++n) values[s->a_i2 - 1 + n] += s_j * s_k * v[n];
        if (deriv)
          for (int n = 0; n < N; ++n)
            for (int x = 0; x < 3; ++x) {
              d0[3 * (s->a_i2 - 1 + n) + x] += s_j * s_k * di[3 * n + x];
              d1[3 * (s->a_i2 - 1 + n) + x + 3 * (size_t)nv * j] += s_j * s_k * dj[3 * n + x];
              d1[3 * (s->a_i2 - 1 + n) + x + 3 * (size_t)nv * k] += s_j * s_k * dk[3 * n + x];
            }
      }
    }
  }
  free(v); free(f); free(df); free(di); free(dj); free(dk);
}

/* ------------------------------------------------------------------ */
/* Behler-Parrinello symmetry functions                               */
/* ------------------------------------------------------------------ */

/* sf_cut (src/ext/symmfunc.f90:823-838) */
static void sf_cut(double R, double Rc, double *fc, double *dfc) {
  if (R >= Rc) { *fc = 0.0; *dfc = 0.0; }
  else {
    *fc = 0.5 * (cos(PI_SF * R / Rc) + 1.0);
    *dfc = -0.5 * PI_SF / Rc * sin(PI_SF * R / Rc);
  }
}

/* sf_F2_ij (src/ext/symmfunc.f90:662-682) */
static void sf_F2(double R, double Rc, double eta, double *F, double *dF) {
  double fc, dfc;
  sf_cut(R, Rc, &fc, &dfc);
  double fexp = exp(-eta * R * R);
  *F = fexp * fc;
  *dF = fexp * (dfc - 2.0 * eta * R * fc);
}

/* sf_F1_ijk (src/ext/symmfunc.f90:625-654) */
static void sf_F1(const double *vij, const double *vik, double Rij, double Rik,
                  double cost, double lambda, double zeta, double *F,
                  double *dFj, double *dFk) {
  double arg = 0.5 * (1.0 + lambda * cost);
  double pref = 0.5 * zeta * lambda * pow(arg, zeta - 1.0);
  *F = pow(arg, zeta);
  for (int x = 0; x < 3; ++x) {
    dFj[x] = pref * (-cost * (vij[x] / Rij) + vik[x] / Rij);
    dFk[x] = pref * (-cost * (vik[x] / Rik) + vij[x] / Rik);
  }
}

/* sf_fingerprint (src/ext/symmfunc.f90:311-431) with G1-G5 updates
 * (:464-811).  type1 and typej are setup-local type ids. */
static void sf_fingerprint(const orc_bp *b, int type1, const double *Xi, int nx,
                           const double *X, const int *typej, int n, double *G,
                           double *dGi, double *dGj) {
  int deriv = dGi && dGj;
  memset(G, 0, sizeof(double) * n);
  if (deriv) {
    memset(dGi, 0, sizeof(double) * 3 * n);
    memset(dGj, 0, sizeof(double) * 3 * n * (size_t)nx);
  }
  for (int j = 0; j < nx; ++j) {
    int pij = IDX2(b, type1, typej[j]);
    double R1[3] = {X[3 * j] - Xi[0], X[3 * j + 1] - Xi[1], X[3 * j + 2] - Xi[2]};
    double Rij = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2]);
    for (int x = 0; x < 3; ++x) R1[x] = R1[x] / Rij;
    int iG = b->iG02[pij - 1];
    double *dGjj = deriv ? dGj + 3 * (size_t)n * j : NULL;
    /* G1 */
    for (int g = 0; g < b->nG2[0 + 5 * (pij - 1)]; ++g) {
      double fc, dfc;
      sf_cut(Rij, b->pG1[g + b->nGmax * (pij - 1)], &fc, &dfc);
      G[iG] += fc;
      if (deriv)
        for (int x = 0; x < 3; ++x) { dGi[3 * iG + x] -= dfc * R1[x]; dGjj[3 * iG + x] += dfc * R1[x]; }
      ++iG;
    }
    /* G2 (sf_G2_ij, :497-518) */
    for (int g = 0; g < b->nG2[1 + 5 * (pij - 1)]; ++g) {
      const double *q = b->pG2 + 3 * ((size_t)g + b->nGmax * (pij - 1));
      double fc, dfc;
      sf_cut(Rij, q[0], &fc, &dfc);
      double arg = Rij - q[1];
      double fexp = exp(-q[2] * arg * arg);
      double Gv = fexp * fc, dG = fexp * (dfc - 2.0 * q[2] * arg * fc);
      G[iG] += Gv;
      if (deriv)
        for (int x = 0; x < 3; ++x) { dGi[3 * iG + x] -= dG * R1[x]; dGjj[3 * iG + x] += dG * R1[x]; }
      ++iG;
    }
    /* G3 (sf_G3_ij, :557-580) */
    for (int g = 0; g < b->nG2[2 + 5 * (pij - 1)]; ++g) {
      const double *q = b->pG3 + 2 * ((size_t)g + b->nGmax * (pij - 1));
      double fc, dfc;
      sf_cut(Rij, q[0], &fc, &dfc);
      double arg = q[1] * Rij;
      double Gv = cos(arg) * fc, dG = cos(arg) * dfc - q[1] * sin(arg) * fc;
      G[iG] += Gv;
      if (deriv)
        for (int x = 0; x < 3; ++x) { dGi[3 * iG + x] -= dG * R1[x]; dGjj[3 * iG + x] += dG * R1[x]; }
      ++iG;
    }
    for (int k = j + 1; k < nx; ++k) {
      int tijk = IDX3(b, type1, typej[j], typej[k]);
      double R2[3] = {X[3 * k] - Xi[0], X[3 * k + 1] - Xi[1], X[3 * k + 2] - Xi[2]};
      double Rik = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
      if (!(Rik <= 1.0e-8))
        for (int x = 0; x < 3; ++x) R2[x] = R2[x] / Rik;
      double R3[3] = {X[3 * k] - X[3 * j], X[3 * k + 1] - X[3 * j + 1], X[3 * k + 2] - X[3 * j + 2]};
      double Rjk = sqrt(R3[0] * R3[0] + R3[1] * R3[1] + R3[2] * R3[2]);
      if (!(Rjk <= 1.0e-8))
        for (int x = 0; x < 3; ++x) R3[x] = R3[x] / Rjk;
      double cost = R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2];
      if (cost < -1.0) cost = -1.0;
      if (cost > 1.0) cost = 1.0;
      int iGa = b->iG03[tijk - 1];
      double *dGkk = deriv ? dGj + 3 * (size_t)n * k : NULL;
      /* G4 (sf_G4_update, :686-747) */
      for (int g = 0; g < b->nG3[3 + 5 * (tijk - 1)]; ++g) {
        const double *q = b->pG4 + 4 * ((size_t)g + b->nGmax * (tijk - 1));
        double F1, dF1j[3], dF1k[3], F2ij, dF2ij, F2ik, dF2ik, F2jk, dF2jk;
        sf_F1(R1, R2, Rij, Rik, cost, q[1], q[2], &F1, dF1j, dF1k);
        sf_F2(Rij, q[0], q[3], &F2ij, &dF2ij);
        sf_F2(Rik, q[0], q[3], &F2ik, &dF2ik);
        sf_F2(Rjk, q[0], q[3], &F2jk, &dF2jk);
        double G4 = F1 * F2ij * F2ik * F2jk;
        G[iGa] += 2.0 * G4;
        if (deriv)
          for (int x = 0; x < 3; ++x) {
            double di = F2jk * (-(dF1j[x] + dF1k[x]) * F2ij * F2ik - R1[x] * F1 * dF2ij * F2ik - R2[x] * F1 * F2ij * dF2ik);
            double dj = dF1j[x] * F2ij * F2ik * F2jk + R1[x] * F1 * dF2ij * F2ik * F2jk - R3[x] * F1 * F2ij * F2ik * dF2jk;
            double dk = dF1k[x] * F2ij * F2ik * F2jk + R2[x] * F1 * F2ij * dF2ik * F2jk + R3[x] * F1 * F2ij * F2ik * dF2jk;
            dGi[3 * iGa + x] += 2.0 * di;
            dGjj[3 * iGa + x] += 2.0 * dj;
            dGkk[3 * iGa + x] += 2.0 * dk;
          }
        ++iGa;
      }
      /* G5 (sf_G5_update, :753-811) */
      for (int g = 0; g < b->nG3[4 + 5 * (tijk - 1)]; ++g) {
        const double *q = b->pG5 + 4 * ((size_t)g + b->nGmax * (tijk - 1));
        double F1, dF1j[3], dF1k[3], F2ij, dF2ij, F2ik, dF2ik;
        sf_F1(R1, R2, Rij, Rik, cost, q[1], q[2], &F1, dF1j, dF1k);
        sf_F2(Rij, q[0], q[3], &F2ij, &dF2ij);
        sf_F2(Rik, q[0], q[3], &F2ik, &dF2ik);
        G[iGa] += 2.0 * F1 * F2ij * F2ik;
        if (deriv)
          for (int x = 0; x < 3; ++x) {
            double di = -(dF1j[x] + dF1k[x]) * F2ij * F2ik - R1[x] * F1 * dF2ij * F2ik - R2[x] * F1 * F2ij * dF2ik;
            double dj = dF1j[x] * F2ij * F2ik + R1[x] * F1 * dF2ij * F2ik;
            double dk = dF1k[x] * F2ij * F2ik + R2[x] * F1 * F2ij * dF2ik;
            dGi[3 * iGa + x] += 2.0 * di;
            dGjj[3 * iGa + x] += 2.0 * dj;
            dGkk[3 * iGa + x] += 2.0 * dk;
          }
        ++iGa;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* stp_eval + stp_normalize                                           */
/* ------------------------------------------------------------------ */

static int stp_eval(orc_ctx *c, int itype0, const double *coo0, int n,
                    const double *coo1, const int *type1, int scaled,
                    double *G, double *dGi, double *dGj) {
  orc_setup *s = &c->pot[itype0 - 1].stp;
  int t0 = s->ltype[itype0];
  int *t1 = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    if (type1[i] < 1 || type1[i] > c->ntypes) { free(t1); return ORC_ERR_TYPE; }
    t1[i] = s->ltype[type1[i]];
  }
  if (c->is_bp) {
    if (t0 < 1) { free(t1); return ORC_ERR_TYPE; }
    sf_fingerprint(&c->bp, t0, coo0, n, coo1, t1, s->nsf, G, dGi, dGj);
  } else {
    sfb_eval(s, coo0, n, t1, coo1, G, dGi, dGj);
  }
  free(t1);
  if (scaled) {
    for (int isf = 0; isf < s->nsf; ++isf) {
      double shift = s->sfval_avg[isf];
      double sd = sqrt(s->sfval_cov[isf] - shift * shift);
      double scale = sd <= 1.0e-10 ? 0.0 : 1.0 / sd;
      G[isf] = scale * (G[isf] - shift);
      if (dGi && dGj) {
        for (int x = 0; x < 3; ++x) dGi[3 * isf + x] *= scale;
        for (int j = 0; j < n; ++j)
          for (int x = 0; x < 3; ++x) dGj[3 * isf + x + 3 * (size_t)s->nsf * j] *= scale;
      }
    }
  }
  return ORC_OK;
}

int orc_descriptors(orc_ctx *c, int type_i, const double coo_i[3], int n_j,
                    const double *coo_j, const int *type_j, int scaled,
                    double *G, double *dGi, double *dGj) {
  if (!c || !c->loaded) return ORC_ERR_INIT;
  if (type_i < 1 || type_i > c->ntypes) return ORC_ERR_TYPE;
  return stp_eval(c, type_i, coo_i, n_j, coo_j, type_j, scaled, G, dGi, dGj);
}

/* ------------------------------------------------------------------ */
/* feed-forward network                                               */
/* ------------------------------------------------------------------ */

/* ff_activate (src/ext/feedforward.f90:932-969) */
static void ff_activate(int t, double x, double *y, double *dy) {
  const double a = 1.7159, b = 0.666666666666667, cc = 0.1;
  double tb;
  switch (t) {
  case 0: *y = x; *dy = 1.0; break;
  case 1: *y = tanh(x); *dy = 1.0 - (*y) * (*y); break;
  case 2: *y = 1.0 / (1.0 + exp(-x)); *dy = (*y) * (1.0 - *y); break;
  case 3: tb = tanh(b * x); *y = a * tb; *dy = a * (1.0 - tb * tb) * b; break;
  case 4: tb = tanh(b * x); *y = a * tb + cc * x; *dy = a * (1.0 - tb * tb) * b + cc; break;
  default: *y = 0.0; *dy = 0.0;
  }
}

/* ff_eval (src/ext/feedforward.f90:586-655): values/derivs have nvalues. */
static double ff_eval(const orc_net *n, const double *x, double *values, double *derivs) {
  int nn1 = n->nnodes[0];
  for (int i = 0; i < nn1; ++i) values[i] = x[i];
  int iw1 = 1, iv1 = 1, nn2 = 1;
  double *work = (double *)malloc(sizeof(double) * (n->nnodes_max + 1));
  for (int l = 0; l < n->nlayers - 1; ++l) {
    int iw2 = n->iw[l + 1], iv2 = n->iv[l + 1];
    values[iv2 - 1] = 1.0;
    nn2 = n->nnodes[l + 1];
    for (int h = 0; h < nn2; ++h) {
      double acc = 0.0;
      for (int cidx = 0; cidx < nn1 + 1; ++cidx)
        acc += n->W[iw1 - 1 + cidx * nn2 + h] * values[iv1 - 1 + cidx];
      work[h] = acc;
    }
    iv1 = iv2 + 1;
    iw1 = iw2 + 1;
    for (int h = 0; h < nn2; ++h)
      ff_activate(n->f_a[l + 1], work[h], &values[iv1 - 1 + h], &derivs[iv1 - 1 + h]);
    nn1 = nn2;
  }
  free(work);
  return values[iv1 - 1];
}

/* ff_deriv (src/ext/feedforward.f90:661-757) for a single output: returns
 * dy/dx in dy[nx] and the layer Jacobians F'W in jac[Wsize]. */
static void ff_deriv(const orc_net *n, const double *derivs, double *jac, double *dy) {
  int M = n->nnodes_max + 1;
  double *w4 = (double *)calloc((size_t)M * M * M, sizeof(double));
  double *w3 = (double *)calloc((size_t)M * M, sizeof(double));
  double *tmp = (double *)calloc((size_t)M * M * M, sizeof(double));
  memset(jac, 0, sizeof(double) * n->Wsize);
  int nn1 = n->nnodes[0], nn0 = nn1, nn2 = 1;
  for (int i = 0; i < M; ++i) w4[i + (size_t)M * i] = 1.0;
  int iw1 = 1;
  for (int l = 0; l < n->nlayers - 1; ++l) {
    int iw2 = n->iw[l + 1], iv2 = n->iv[l + 1];
    nn2 = n->nnodes[l + 1];
    for (int h = 0; h < nn2; ++h)
      for (int cidx = 0; cidx < nn1; ++cidx)
        w3[h + (size_t)M * cidx] = derivs[iv2 + h] * n->W[iw1 - 1 + cidx * nn2 + h];
    for (int cidx = 0; cidx < nn1; ++cidx)
      for (int h = 0; h < nn2; ++h) jac[iw1 - 1 + cidx * nn2 + h] = w3[h + (size_t)M * cidx];
    for (int h = 0; h < nn2; ++h)
      for (int q = 0; q < nn0; ++q) {
        double acc = 0.0;
        for (int m = 0; m < nn1; ++m) acc += w3[h + (size_t)M * m] * w4[m + (size_t)M * q];
        tmp[h + (size_t)M * q] = acc;
      }
    for (int h = 0; h < nn2; ++h)
      for (int q = 0; q < nn0; ++q) w4[h + (size_t)M * q] = tmp[h + (size_t)M * q];
    iw1 = iw2 + 1;
    nn1 = nn2;
  }
  for (int q = 0; q < nn0; ++q) dy[q] = w4[0 + (size_t)M * q];
  free(w4); free(w3); free(tmp);
}

/* ff_wderiv (src/ext/feedforward.f90:763-860) for a single output. */
static void ff_wderiv(const orc_net *n, const double *values, const double *derivs,
                      const double *jac, double *dydw) {
  int M = n->nnodes_max + 1;
  double *w3 = (double *)calloc((size_t)M * M, sizeof(double));
  double *w2 = (double *)calloc((size_t)M * M, sizeof(double));
  double *tmp = (double *)calloc((size_t)M * M, sizeof(double));
  for (int i = 0; i < M; ++i) w3[i + (size_t)M * i] = 1.0;
  int iw2 = n->iw[n->nlayers - 1], iv2 = n->iv[n->nlayers - 1];
  int nn2 = n->nnodes[n->nlayers - 1], nn3 = nn2;
  for (int l = n->nlayers - 2; l >= 0; --l) {
    int iw1 = n->iw[l] + 1, iv1 = n->iv[l] + 1, nn1 = n->nnodes[l];
    for (int a = 0; a < nn2; ++a)
      for (int b = 0; b < nn1 + 1; ++b) w2[a + (size_t)M * b] = derivs[iv2 + a] * values[iv1 - 1 + b];
    /* dy_dw(1:nn3, iw1:iw2) = w3(1:nn3,1:nn2) * work4, work4(in2, in0) */
    for (int b = 0; b < nn1 + 1; ++b)
      for (int a = 0; a < nn2; ++a) {
        /* column in0 = b*nn2 + a has the single nonzero work2(a,b) at row a */
        for (int y = 0; y < nn3; ++y) dydw[(iw1 - 1) + b * nn2 + a] = w3[y + (size_t)M * a] * w2[a + (size_t)M * b];
      }
    for (int y = 0; y < nn3; ++y)
      for (int q = 0; q < nn1; ++q) {
        double acc = 0.0;
        for (int a = 0; a < nn2; ++a) acc += w3[y + (size_t)M * a] * jac[iw1 - 1 + q * nn2 + a];
        tmp[y + (size_t)M * q] = acc;
      }
    for (int y = 0; y < nn3; ++y)
      for (int q = 0; q < nn1; ++q) w3[y + (size_t)M * q] = tmp[y + (size_t)M * q];
    iv2 = iv1 - 1;
    iw2 = iw1 - 1;
    nn2 = nn1;
  }
  free(w3); free(w2); free(tmp);
}

int orc_network(orc_ctx *c, int t, const double *x, double *y, double *dydx, double *dydw) {
  if (!c || t < 1 || t > c->ntypes || !c->pot[t - 1].init) return ORC_ERR_INIT;
  const orc_net *n = &c->pot[t - 1].net;
  if (n->nnodes[n->nlayers - 1] != 1) return ORC_ERR_FORMAT;
  double *values = (double *)calloc(n->nvalues, sizeof(double));
  double *derivs = (double *)calloc(n->nvalues, sizeof(double));
  double *jac = (double *)calloc(n->Wsize, sizeof(double));
  double *dy = (double *)calloc(n->nnodes[0], sizeof(double));
  *y = ff_eval(n, x, values, derivs);
  ff_deriv(n, derivs, jac, dy);
  if (dydx) memcpy(dydx, dy, sizeof(double) * n->nnodes[0]);
  if (dydw) ff_wderiv(n, values, derivs, jac, dydw);
  free(values); free(derivs); free(jac); free(dy);
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* per-atom energies and forces (src/aenet.f90:359-486)                */
/* ------------------------------------------------------------------ */

int orc_atomic_energy(orc_ctx *c, const double coo_i[3], int type_i, int n_j,
                      const double *coo_j, const int *type_j, double *E_i) {
  if (!c || !c->loaded) return ORC_ERR_INIT;
  if (type_i < 1 || type_i > c->ntypes) return ORC_ERR_TYPE;
  orc_pot *p = &c->pot[type_i - 1];
  double *G = (double *)calloc(c->nsf_max, sizeof(double));
  double *values = (double *)calloc(p->net.nvalues, sizeof(double));
  double *derivs = (double *)calloc(p->net.nvalues, sizeof(double));
  int st = stp_eval(c, type_i, coo_i, n_j, coo_j, type_j, 1, G, NULL, NULL);
  if (!st) {
    double y = ff_eval(&p->net, G, values, derivs);
    *E_i = p->E_scale * y + p->E_shift;
    *E_i = *E_i + p->E_atom;
  }
  free(G); free(values); free(derivs);
  return st;
}

int orc_atomic_energy_and_forces(orc_ctx *c, const double coo_i[3], int type_i,
                                 int index_i, int n_j, const double *coo_j,
                                 const int *type_j, const int *index_j,
                                 int natoms, double *E_i, double *F) {
  if (!c || !c->loaded) return ORC_ERR_INIT;
  if (type_i < 1 || type_i > c->ntypes) return ORC_ERR_TYPE;
  orc_pot *p = &c->pot[type_i - 1];
  int nsf = p->stp.nsf;
  double *G = (double *)calloc(c->nsf_max, sizeof(double));
  double *dGi = (double *)calloc(3 * (size_t)nsf, sizeof(double));
  double *dGj = (double *)calloc(3 * (size_t)nsf * (n_j > 0 ? n_j : 1), sizeof(double));
  double *values = (double *)calloc(p->net.nvalues, sizeof(double));
  double *derivs = (double *)calloc(p->net.nvalues, sizeof(double));
  double *jac = (double *)calloc(p->net.Wsize, sizeof(double));
  double *dEdG = (double *)calloc(nsf, sizeof(double));
  int st = stp_eval(c, type_i, coo_i, n_j, coo_j, type_j, 1, G, dGi, dGj);
  if (!st) {
    double y = ff_eval(&p->net, G, values, derivs);
    ff_deriv(&p->net, derivs, jac, dEdG);
    *E_i = p->E_scale * y + p->E_shift;
    *E_i = *E_i + p->E_atom;
    double m[3];
    for (int x = 0; x < 3; ++x) {
      double acc = 0.0;
      for (int g = 0; g < nsf; ++g) acc += dGi[3 * g + x] * dEdG[g];
      m[x] = acc;
    }
    for (int x = 0; x < 3; ++x) F[3 * (size_t)(index_i - 1) + x] = F[3 * (size_t)(index_i - 1) + x] - p->E_scale * m[x];
    for (int j = 0; j < n_j; ++j) {
      for (int x = 0; x < 3; ++x) {
        double acc = 0.0;
        for (int g = 0; g < nsf; ++g) acc += dGj[3 * g + x + 3 * (size_t)nsf * j] * dEdG[g];
        m[x] = acc;
      }
      for (int x = 0; x < 3; ++x)
        F[3 * (size_t)(index_j[j] - 1) + x] = F[3 * (size_t)(index_j[j] - 1) + x] - p->E_scale * m[x];
    }
  }
  (void)natoms;
  free(G); free(dGi); free(dGj); free(values); free(derivs); free(jac); free(dEdG);
  return st;
}

/* ------------------------------------------------------------------ */
/* geometry (src/geometry.f90)                                        */
/* ------------------------------------------------------------------ */

static void vprod(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* geo_cell_volume (:467-483) -- signed */
static double geo_cell_volume(const double *A) {
#define A_(i, j) A[((j)-1) * 3 + ((i)-1)]
  return A_(1, 1) * A_(2, 2) * A_(3, 3) + A_(2, 1) * A_(3, 2) * A_(1, 3) + A_(3, 1) * A_(1, 2) * A_(2, 3)
       - A_(3, 1) * A_(2, 2) * A_(1, 3) - A_(1, 1) * A_(3, 2) * A_(2, 3) - A_(2, 1) * A_(1, 2) * A_(3, 3);
#undef A_
}

/* geo_recip_lattice (:491-514) with 2*PI scaling; B column-major */
static void geo_recip(const double *A, double *B) {
  double r[3][3];
  vprod(A + 3, A + 6, r[0]);
  vprod(A + 0, A + 6, r[1]);
  vprod(A + 0, A + 3, r[2]);
  for (int x = 0; x < 3; ++x) r[1][x] = -r[1][x];
  double V = geo_cell_volume(A);
  for (int i = 0; i < 3; ++i)
    for (int x = 0; x < 3; ++x) B[x * 3 + i] = r[i][x] / V;
  for (int k = 0; k < 9; ++k) B[k] = B[k] * 2.0 * PI_GEO;
}

void orc_cart_to_frac(const double latt[9], int pbc, int natoms,
                      const double *cart, double *frac, double latt_out[9]) {
  double *work = (double *)malloc(sizeof(double) * 3 * (size_t)(natoms > 0 ? natoms : 1));
  memcpy(work, cart, sizeof(double) * 3 * (size_t)natoms);
  if (pbc) {
    memcpy(latt_out, latt, sizeof(double) * 9);
  } else {
    /* geo_get_bounds (:538-580) */
    const double EPS = 2.0e-6;
    double mn[3], mx[3];
    for (int x = 0; x < 3; ++x) {
      mn[x] = work[x];
      mx[x] = work[x];
      for (int i = 1; i < natoms; ++i) {
        if (work[3 * i + x] < mn[x]) mn[x] = work[3 * i + x];
        if (work[3 * i + x] > mx[x]) mx[x] = work[3 * i + x];
      }
    }
    memset(latt_out, 0, sizeof(double) * 9);
    for (int x = 0; x < 3; ++x) latt_out[4 * x] = mx[x] - mn[x] + EPS;
    double shift[3];
    for (int x = 0; x < 3; ++x) shift[x] = mn[x] - 0.5 * EPS;
    for (int i = 0; i < natoms; ++i)
      for (int x = 0; x < 3; ++x) work[3 * i + x] = work[3 * i + x] - shift[x];
  }
  double B[9];
  geo_recip(latt_out, B);
  for (int i = 0; i < natoms; ++i)
    for (int r = 0; r < 3; ++r) {
      double acc = B[r] * work[3 * i] + B[3 + r] * work[3 * i + 1];
      acc = acc + B[6 + r] * work[3 * i + 2];
      frac[3 * i + r] = acc / (2.0 * PI_GEO);
    }
  free(work);
}

/* ------------------------------------------------------------------ */
/* linked-cell neighbor list (src/ext/lclist.f90)                      */
/* ------------------------------------------------------------------ */

typedef struct {
  int nat, pbc;
  double L[9]; /* column-major: L[3*k+x] = component x of vector k */
  double *s;   /* wrapped fractional coordinates, 3 per atom */
  const int *type;
  int nc[3];
  int *head, *next, *cellof; /* cellof: 3 ints per atom (1-based) */
  int nC, nT;
  int *Cvec, *Tvec;
  int nmax_nblist;
} lcl_t;

static int cidx(const lcl_t *l, const int *ic) {
  return ((ic[0] - 1) * l->nc[1] + (ic[1] - 1)) * l->nc[2] + (ic[2] - 1);
}

/* sphere_supercell (:1070-1098) */
static void sphere_supercell(const double *A, double r, int *n) {
#define A_(i, j) A[((j)-1) * 3 + ((i)-1)]
  double V = fabs(geo_cell_volume(A));
  double row1[3] = {A_(1, 1), A_(1, 2), A_(1, 3)};
  double row2[3] = {A_(2, 1), A_(2, 2), A_(2, 3)};
  double row3[3] = {A_(3, 1), A_(3, 2), A_(3, 3)};
#undef A_
  double b[3][3];
  vprod(row2, row3, b[0]);
  vprod(row1, row3, b[1]);
  vprod(row1, row2, b[2]);
  for (int x = 0; x < 3; ++x) {
    b[0][x] = b[0][x] / V;
    b[1][x] = -b[1][x] / V;
    b[2][x] = b[2][x] / V;
  }
  for (int k = 0; k < 3; ++k) {
    double bl = sqrt(b[k][0] * b[k][0] + b[k][1] * b[k][1] + b[k][2] * b[k][2]);
    n[k] = (int)ceil(r * bl + 0.001);
  }
}

static int fmod_pos(int a, int n) {
  int m = a % n;
  return m < 0 ? m + n : m;
}

/* translation_vectors (:1235-1386); returns count, fills T (3 per vector) */
static int translation_vectors(double Rc, const double *A, int **Tout, const int *nc, int pbc) {
  int n[3];
  sphere_supercell(A, Rc + 1.0e-3, n);
  for (int k = 0; k < 3; ++k) n[k] = (n[k] / 2 + 1) + 1;
  if (nc)
    for (int k = 0; k < 3; ++k) n[k] = n[k] < nc[k] - 1 ? n[k] : nc[k] - 1;
  int cap = (2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1);
  int *T = (int *)malloc(sizeof(int) * 3 * (size_t)(cap > 0 ? cap : 1));
  unsigned char *seen = NULL;
  if (nc && pbc) seen = (unsigned char *)calloc((size_t)nc[0] * nc[1] * nc[2], 1);
  int iT = 0;
  for (int i3 = -n[2]; i3 <= n[2]; ++i3)
    for (int i2 = -n[1]; i2 <= n[1]; ++i2)
      for (int i1 = -n[0]; i1 <= n[0]; ++i1) {
        if (i1 == 0 && i2 == 0 && i3 == 0) continue;
        int redundant = 0, key = 0;
        if (seen) {
          /* cell_from_Cvec([1,1,1], T, nc, .true.) and in_list over the cells stored so far */
          int ic0 = fmod_pos(i1, nc[0]), ic1 = fmod_pos(i2, nc[1]), ic2 = fmod_pos(i3, nc[2]);
          key = (ic0 * nc[1] + ic1) * nc[2] + ic2;
          redundant = seen[key];
        }
        if (!redundant) {
          T[3 * iT] = i1; T[3 * iT + 1] = i2; T[3 * iT + 2] = i3;
          ++iT;
          if (seen) seen[key] = 1;
        }
      }
  free(seen);
  *Tout = T;
  return iT;
}

/* inner_sphere (:1015-1047) */
static double inner_sphere(const double *A) {
  double v[3], r, h;
  vprod(A + 0, A + 3, v);
  double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  r = fabs(v[0] * A[6] + v[1] * A[7] + v[2] * A[8]);
  vprod(A + 0, A + 6, v);
  nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h = fabs(v[0] * A[3] + v[1] * A[4] + v[2] * A[5]);
  r = r < h ? r : h;
  vprod(A + 3, A + 6, v);
  nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h = fabs(v[0] * A[0] + v[1] * A[1] + v[2] * A[2]);
  r = r < h ? r : h;
  return 0.5 * r;
}

/* cell_multiples (:725-782) */
static void cell_multiples(double r_max, const double *L, double *grid, int *nc) {
  double v[3], h[3];
  vprod(L + 0, L + 3, v);
  double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h[2] = fabs(v[0] * L[6] + v[1] * L[7] + v[2] * L[8]);
  vprod(L + 0, L + 6, v);
  nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h[1] = fabs(v[0] * L[3] + v[1] * L[4] + v[2] * L[5]);
  vprod(L + 3, L + 6, v);
  nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h[0] = fabs(v[0] * L[0] + v[1] * L[1] + v[2] * L[2]);
  double hmin = h[0];
  if (h[1] < hmin) hmin = h[1];
  if (h[2] < hmin) hmin = h[2];
  if (hmin < 0.5) hmin = 0.5;
  for (int k = 0; k < 3; ++k) {
    nc[k] = (int)floor(h[k] / hmin);
    if (nc[k] < 1) nc[k] = 1;
  }
  for (int k = 0; k < 3; ++k)
    for (int x = 0; x < 3; ++x) grid[3 * k + x] = L[3 * k + x] / (double)nc[k];
  double r = inner_sphere(grid);
  if (r > 0.5 * r_max) {
    for (int k = 0; k < 3; ++k) nc[k] = (int)floor(nc[k] * (2.0 * r / r_max));
    for (int k = 0; k < 3; ++k)
      for (int x = 0; x < 3; ++x) grid[3 * k + x] = L[3 * k + x] / (double)nc[k];
  }
}

static void lcl_free(lcl_t *l) {
  free(l->head); free(l->next); free(l->cellof); free(l->Cvec); free(l->Tvec);
}

/* lcl_init (:99-159); s is wrapped in place */
static void lcl_init(lcl_t *l, double r_min, double r_max, const double *L, int nat,
                     const int *type, double *s, int pbc) {
  memset(l, 0, sizeof(*l));
  l->nat = nat;
  l->pbc = pbc;
  memcpy(l->L, L, sizeof(double) * 9);
  l->s = s;
  l->type = type;
  double grid[9];
  cell_multiples(r_max, L, grid, l->nc);
  int ncell = l->nc[0] * l->nc[1] * l->nc[2];
  l->head = (int *)calloc(ncell, sizeof(int));
  l->next = (int *)calloc((size_t)nat + 1, sizeof(int));
  l->cellof = (int *)calloc(3 * ((size_t)nat + 1), sizeof(int));
  /* cell_assign_atoms (:692-719) with wrap_cooLatt (:953-987), cell_get (:833-845) */
  for (int i = 1; i <= nat; ++i) {
    double *c = s + 3 * (i - 1);
    if (pbc) {
      for (int x = 0; x < 3; ++x) {
        while (c[x] < 0.0) c[x] = c[x] + 1.0;
        while (c[x] >= 1.0) c[x] = c[x] - 1.0;
      }
    }
    int ic[3];
    for (int x = 0; x < 3; ++x) {
      double t = c[x] * (double)l->nc[x] + 0.5;
      int v = (int)lround(t);
      ic[x] = v < l->nc[x] ? v : l->nc[x];
    }
    int id = cidx(l, ic);
    l->next[i] = l->head[id];
    l->head[id] = i;
    for (int x = 0; x < 3; ++x) l->cellof[3 * i + x] = ic[x];
  }
  l->nC = translation_vectors(r_max, grid, &l->Cvec, l->nc, pbc);
  l->nT = 0;
  l->Tvec = NULL;
  if (pbc) l->nT = translation_vectors(r_max, L, &l->Tvec, NULL, 1);
  /* lcl_nmax_cell (:251-271), lcl_nmax_nblist (:215-223) */
  double Vc = fabs(geo_cell_volume(grid));
  double Va = sqrt(32.0) * pow(0.5 * r_min, 3);
  int nmax_cell = (int)ceil(Vc / Va);
  long nb = (long)nmax_cell * (l->nC + 1);
  l->nmax_nblist = (int)(nb < nat - 1 ? nb : nat - 1);
}

/* lcl_nblist (:281-376); returns count or -1 when nmax_nblist is exceeded */
static int lcl_nblist(const lcl_t *l, int iatom, int *list, unsigned char *mark) {
  int nnb = l->nmax_nblist, inb = 0;
  const int *ic0 = l->cellof + 3 * iatom;
  int iat = l->head[cidx(l, ic0)];
  if (iat != 0) {
    if (iat != iatom) {
      if (++inb > nnb) return -1;
      list[inb - 1] = iat;
      mark[iat] = 1;
    }
    while (l->next[iat] != 0) {
      iat = l->next[iat];
      if (iat == iatom) continue;
      if (++inb > nnb) return -1;
      list[inb - 1] = iat;
      mark[iat] = 1;
    }
  }
  for (int iv = 0; iv < l->nC; ++iv) {
    int ic[3];
    const int *cv = l->Cvec + 3 * iv;
    if (l->pbc) {
      for (int x = 0; x < 3; ++x) ic[x] = fmod_pos(ic0[x] + cv[x] - 1, l->nc[x]) + 1;
    } else {
      int out = 0;
      for (int x = 0; x < 3; ++x) {
        ic[x] = ic0[x] + cv[x];
        if (ic[x] > l->nc[x] || ic[x] < 1) out = 1;
      }
      if (out) continue;
    }
    iat = l->head[cidx(l, ic)];
    if (iat != 0) {
      if (!mark[iat]) {
        if (++inb > nnb) return -1;
        list[inb - 1] = iat;
        mark[iat] = 1;
      }
      while (l->next[iat] != 0) {
        iat = l->next[iat];
        if (!mark[iat]) {
          if (++inb > nnb) return -1;
          list[inb - 1] = iat;
          mark[iat] = 1;
        }
      }
    }
  }
  return inb;
}

static void matvec(const double *L, const double *v, double *out) {
  for (int r = 0; r < 3; ++r) {
    double acc = L[r] * v[0] + L[3 + r] * v[1];
    out[r] = acc + L[6 + r] * v[2];
  }
}

/* lcl_nbdist_cart (:520-677) for atom iatom (1-based); returns count or
 * a negative status */
static int lcl_nbdist_cart(const lcl_t *l, int iatom, double Rc, int cap, double *nbcoo,
                           double *nbdist, int *nblist, int *nbtype, int *cand,
                           unsigned char *mark) {
  const double EPS = 1.0e-3;
  double Rc2 = (Rc + EPS) * (Rc + EPS);
  int nc = lcl_nblist(l, iatom, cand, mark);
  for (int k = 0; k < (nc > 0 ? nc : 0); ++k) mark[cand[k]] = 0;
  if (nc < 0) {
    memset(mark, 0, (size_t)l->nat + 1);
    return -4;
  }
  int n = 0;
  const double *si = l->s + 3 * (iatom - 1);
  for (int iT = 0; iT < l->nT; ++iT) {
    double Tv[3] = {(double)l->Tvec[3 * iT], (double)l->Tvec[3 * iT + 1], (double)l->Tvec[3 * iT + 2]};
    double cart[3];
    matvec(l->L, Tv, cart);
    double d2 = cart[0] * cart[0] + cart[1] * cart[1];
    d2 = d2 + cart[2] * cart[2];
    if (d2 <= Rc2) {
      if (++n > cap) return -1;
      nblist[n - 1] = iatom;
      nbtype[n - 1] = l->type[iatom - 1];
      double c2[3] = {si[0] + Tv[0], si[1] + Tv[1], si[2] + Tv[2]};
      matvec(l->L, c2, nbcoo + 3 * (n - 1));
      nbdist[n - 1] = sqrt(d2);
    }
  }
  double coo1[3];
  matvec(l->L, si, coo1);
  for (int inb = 0; inb < nc; ++inb) {
    int iat = cand[inb];
    const double *sj = l->s + 3 * (iat - 1);
    double coo2[3], cart[3];
    matvec(l->L, sj, coo2);
    for (int x = 0; x < 3; ++x) cart[x] = coo2[x] - coo1[x];
    double d2 = cart[0] * cart[0] + cart[1] * cart[1];
    d2 = d2 + cart[2] * cart[2];
    if (d2 <= Rc2) {
      if (++n > cap) return -2;
      nblist[n - 1] = iat;
      nbtype[n - 1] = l->type[iat - 1];
      memcpy(nbcoo + 3 * (n - 1), coo2, sizeof(coo2));
      nbdist[n - 1] = sqrt(d2);
    }
    for (int iT = 0; iT < l->nT; ++iT) {
      double c2[3] = {sj[0] + (double)l->Tvec[3 * iT], sj[1] + (double)l->Tvec[3 * iT + 1],
                      sj[2] + (double)l->Tvec[3 * iT + 2]};
      matvec(l->L, c2, coo2);
      for (int x = 0; x < 3; ++x) cart[x] = coo2[x] - coo1[x];
      d2 = cart[0] * cart[0] + cart[1] * cart[1];
      d2 = d2 + cart[2] * cart[2];
      if (d2 <= Rc2) {
        if (++n > cap) return -3;
        nblist[n - 1] = iat;
        nbtype[n - 1] = l->type[iat - 1];
        memcpy(nbcoo + 3 * (n - 1), coo2, sizeof(coo2));
        nbdist[n - 1] = sqrt(d2);
      }
    }
  }
  return n;
}

int orc_neighbors(double rc_min, double rc_max, const double latt[9], int pbc,
                  int natoms, const int *types, double *frac, int cap,
                  int *nnb, double *nbcoo, double *nbdist, int *nblist,
                  int *nbtype) {
  lcl_t l;
  lcl_init(&l, rc_min, rc_max, latt, natoms, types, frac, pbc);
  int *cand = (int *)malloc(sizeof(int) * ((size_t)natoms + 1));
  unsigned char *mark = (unsigned char *)calloc((size_t)natoms + 1, 1);
  int st = ORC_OK;
  for (int i = 1; i <= natoms; ++i) {
    int n = lcl_nbdist_cart(&l, i, rc_max, cap, nbcoo + 3 * (size_t)cap * (i - 1),
                            nbdist + (size_t)cap * (i - 1), nblist + (size_t)cap * (i - 1),
                            nbtype + (size_t)cap * (i - 1), cand, mark);
    if (n < 0) { st = ORC_ERR_CAPACITY; nnb[i - 1] = 0; break; }
    nnb[i - 1] = n;
  }
  free(cand);
  free(mark);
  lcl_free(&l);
  return st;
}

/* ------------------------------------------------------------------ */
/* predict.x get_energy (src/predict.F90:574-706)                      */
/* ------------------------------------------------------------------ */

int orc_structure(orc_ctx *c, const double latt[9], int pbc, int natoms,
                  const int *types, const double *cart, int rank, int nranks,
                  double *Ecoh, double *Etot, double *E_at, double *F) {
  if (!c || !c->loaded) return ORC_ERR_INIT;
  double L[9];
  double *s = (double *)malloc(sizeof(double) * 3 * (size_t)natoms);
  orc_cart_to_frac(latt, pbc, natoms, cart, s, L);
  lcl_t l;
  lcl_init(&l, c->Rc_min, c->Rc_max, L, natoms, types, s, pbc);
  int cap = c->nnb_max;
  double *nbcoo = (double *)malloc(sizeof(double) * 3 * (size_t)cap);
  double *nbdist = (double *)malloc(sizeof(double) * (size_t)cap);
  int *nblist = (int *)malloc(sizeof(int) * (size_t)cap);
  int *nbtype = (int *)malloc(sizeof(int) * (size_t)cap);
  int *cand = (int *)malloc(sizeof(int) * ((size_t)natoms + 1));
  unsigned char *mark = (unsigned char *)calloc((size_t)natoms + 1, 1);
  if (F) memset(F, 0, sizeof(double) * 3 * (size_t)natoms);
  if (E_at) memset(E_at, 0, sizeof(double) * (size_t)natoms);
  double ecoh = 0.0, etot = 0.0;
  int st = ORC_OK;
  for (int i = 1; i <= natoms; ++i) {
    if ((i - 1) % nranks != rank) continue;
    int ti = types[i - 1];
    double coo_i[3];
    matvec(L, s + 3 * (i - 1), coo_i);
    int n = lcl_nbdist_cart(&l, i, c->Rc_max, cap, nbcoo, nbdist, nblist, nbtype, cand, mark);
    if (n < 0) { st = ORC_ERR_CAPACITY; break; }
    double Ei;
    if (F)
      st = orc_atomic_energy_and_forces(c, coo_i, ti, i, n, nbcoo, nbtype, nblist, natoms, &Ei, F);
    else
      st = orc_atomic_energy(c, coo_i, ti, n, nbcoo, nbtype, &Ei);
    if (st) break;
    etot = etot + Ei;
    ecoh = ecoh + Ei - c->pot[ti - 1].E_atom;
    if (E_at) E_at[i - 1] = Ei;
  }
  *Ecoh = ecoh;
  *Etot = etot;
  free(s); free(nbcoo); free(nbdist); free(nblist); free(nbtype); free(cand); free(mark);
  lcl_free(&l);
  return st;
}

/* ------------------------------------------------------------------ */
/* Adam (src/optimize.f90:1012-1046)                                   */
/* ------------------------------------------------------------------ */

void orc_adam_step(int n, double *W, const double *g, double *m, double *v,
                   int t, double lr, double b1, double b2, double eps) {
  for (int i = 0; i < n; ++i) {
    m[i] = b1 * m[i] + (1.0 - b1) * g[i];
    v[i] = b2 * v[i] + (1.0 - b2) * (g[i] * g[i]);
    double m2 = 1.0 / (1.0 - pow(b1, t)) * m[i];
    double v2 = 1.0 / (1.0 - pow(b2, t)) * v[i];
    W[i] = W[i] + (-lr / (sqrt(v2) + eps) * m2);
  }
}

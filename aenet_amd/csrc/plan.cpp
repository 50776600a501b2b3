// plan.cpp -- linked-cell plans with the reference's FP64 arithmetic.
//
// Compiled with -ffp-contract=off: every expression is evaluated term by
// term in the order written, which is the order of the reference Fortran.
//   cell grid          cell_multiples       (src/ext/lclist.f90:725-782)
//   translations       translation_vectors  (src/ext/lclist.f90:1235-1386)
//   image range        sphere_supercell     (src/ext/lclist.f90:1070-1098)
//   self images        lcl_nbdist_cart (2)  (src/ext/lclist.f90:589-616)
//   reciprocal lattice geo_recip_lattice    (src/geometry.f90:491-514)
//   isolated bounds    geo_get_bounds       (src/geometry.f90:538-580)
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "aenet_capi_status.h"

namespace aenet_amd {

namespace {

const double PI_GEO = 3.141592653589793;  // src/constants.f90:30

void vprod(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// geo_cell_volume (src/geometry.f90:467-483), signed; A column-major
double cell_volume(const double *A) {
  auto a = [A](int i, int j) { return A[(j - 1) * 3 + (i - 1)]; };
  return a(1, 1) * a(2, 2) * a(3, 3) + a(2, 1) * a(3, 2) * a(1, 3) + a(3, 1) * a(1, 2) * a(2, 3) -
         a(3, 1) * a(2, 2) * a(1, 3) - a(1, 1) * a(3, 2) * a(2, 3) - a(2, 1) * a(1, 2) * a(3, 3);
}

void recip_lattice(const double *A, double *B) {
  double r[3][3];
  vprod(A + 3, A + 6, r[0]);
  vprod(A + 0, A + 6, r[1]);
  vprod(A + 0, A + 3, r[2]);
  for (int x = 0; x < 3; ++x) r[1][x] = -r[1][x];
  double V = cell_volume(A);
  for (int i = 0; i < 3; ++i)
    for (int x = 0; x < 3; ++x) B[x * 3 + i] = r[i][x] / V;
  for (int k = 0; k < 9; ++k) B[k] = B[k] * 2.0 * PI_GEO;
}

double inner_sphere(const double *A) {
  double v[3], r, h, nv;
  vprod(A + 0, A + 3, v);
  nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  r = std::fabs(v[0] * A[6] + v[1] * A[7] + v[2] * A[8]);
  vprod(A + 0, A + 6, v);
  nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h = std::fabs(v[0] * A[3] + v[1] * A[4] + v[2] * A[5]);
  r = r < h ? r : h;
  vprod(A + 3, A + 6, v);
  nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h = std::fabs(v[0] * A[0] + v[1] * A[1] + v[2] * A[2]);
  r = r < h ? r : h;
  return 0.5 * r;
}

void cell_multiples(double r_max, const double *L, double *grid, int *nc) {
  double v[3], h[3], nv;
  vprod(L + 0, L + 3, v);
  nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h[2] = std::fabs(v[0] * L[6] + v[1] * L[7] + v[2] * L[8]);
  vprod(L + 0, L + 6, v);
  nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h[1] = std::fabs(v[0] * L[3] + v[1] * L[4] + v[2] * L[5]);
  vprod(L + 3, L + 6, v);
  nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int x = 0; x < 3; ++x) v[x] = v[x] / nv;
  h[0] = std::fabs(v[0] * L[0] + v[1] * L[1] + v[2] * L[2]);
  double hmin = h[0];
  if (h[1] < hmin) hmin = h[1];
  if (h[2] < hmin) hmin = h[2];
  if (hmin < 0.5) hmin = 0.5;
  for (int k = 0; k < 3; ++k) {
    nc[k] = (int)std::floor(h[k] / hmin);
    if (nc[k] < 1) nc[k] = 1;
  }
  for (int k = 0; k < 3; ++k)
    for (int x = 0; x < 3; ++x) grid[3 * k + x] = L[3 * k + x] / (double)nc[k];
  double r = inner_sphere(grid);
  if (r > 0.5 * r_max) {
    for (int k = 0; k < 3; ++k) nc[k] = (int)std::floor(nc[k] * (2.0 * r / r_max));
    for (int k = 0; k < 3; ++k)
      for (int x = 0; x < 3; ++x) grid[3 * k + x] = L[3 * k + x] / (double)nc[k];
  }
}

void sphere_supercell(const double *A, double r, int *n) {
  auto a = [A](int i, int j) { return A[(j - 1) * 3 + (i - 1)]; };
  double V = std::fabs(cell_volume(A));
  // rows of A, as the reference takes them (avec(2,:), avec(3,:), ...)
  double row1[3] = {a(1, 1), a(1, 2), a(1, 3)};
  double row2[3] = {a(2, 1), a(2, 2), a(2, 3)};
  double row3[3] = {a(3, 1), a(3, 2), a(3, 3)};
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
    double bl = std::sqrt(b[k][0] * b[k][0] + b[k][1] * b[k][1] + b[k][2] * b[k][2]);
    n[k] = (int)std::ceil(r * bl + 0.001);
  }
}

int modp(int a, int n) {
  int m = a % n;
  return m < 0 ? m + n : m;
}

// translation vectors in the reference's loop order; with nc (cell
// stencil, pbc) vectors that point to an already listed cell are dropped
void translation_vectors(double Rc, const double *A, const int *nc, int pbc, std::vector<int> &T) {
  int n[3];
  sphere_supercell(A, Rc + 1.0e-3, n);
  for (int k = 0; k < 3; ++k) n[k] = (n[k] / 2 + 1) + 1;
  if (nc)
    for (int k = 0; k < 3; ++k) n[k] = n[k] < nc[k] - 1 ? n[k] : nc[k] - 1;
  std::vector<unsigned char> seen;
  if (nc && pbc) seen.assign((size_t)nc[0] * nc[1] * nc[2], 0);
  T.clear();
  for (int i3 = -n[2]; i3 <= n[2]; ++i3)
    for (int i2 = -n[1]; i2 <= n[1]; ++i2)
      for (int i1 = -n[0]; i1 <= n[0]; ++i1) {
        if (i1 == 0 && i2 == 0 && i3 == 0) continue;
        if (!seen.empty()) {
          size_t key = ((size_t)modp(i1, nc[0]) * nc[1] + modp(i2, nc[1])) * nc[2] + modp(i3, nc[2]);
          if (seen[key]) continue;
          seen[key] = 1;
        }
        T.push_back(i1);
        T.push_back(i2);
        T.push_back(i3);
      }
}

}  // namespace

static int build_plan_impl(double rc_max, int nstruct, const int *natoms, const double *latt, const int *pbc,
                           const double *coords, bool lattice_coords, PlanHost &plan, std::string &err);

int build_plan(double rc_max, int nstruct, const int *natoms, const double *latt, const int *pbc,
               const double *coords, PlanHost &plan, std::string &err) {
  return build_plan_impl(rc_max, nstruct, natoms, latt, pbc, coords, false, plan, err);
}

int build_plan_lattice(double rc_max, int natoms, const double *latt, int pbc, PlanHost &plan,
                       std::string &err) {
  return build_plan_impl(rc_max, 1, &natoms, latt, &pbc, nullptr, true, plan, err);
}

static int build_plan_impl(double rc_max, int nstruct, const int *natoms, const double *latt, const int *pbc,
                           const double *coords, bool lattice_coords, PlanHost &plan, std::string &err) {
  plan = PlanHost();
  plan.st.resize(nstruct);
  const double EPS_LIST = 1.0e-3;
  int atom0 = 0;
  for (int s = 0; s < nstruct; ++s) {
    StructDev &d = plan.st[s];
    std::memset(&d, 0, sizeof(d));
    int n = natoms[s];
    if (n < 1) {
      err = "structure with no atoms";
      return AENET_STATUS_ERR_ARG;
    }
    d.pbc = pbc[s] ? 1 : 0;
    d.atom0 = atom0;
    d.natoms = n;
    if (d.pbc || lattice_coords) {
      std::memcpy(d.L, latt + 9 * (size_t)s, 9 * sizeof(double));
      if (std::fabs(cell_volume(d.L)) <= 0.0) {
        err = "singular lattice";
        return AENET_STATUS_ERR_ARG;
      }
    } else {
      // geo_get_bounds: orthogonal bounding box, slightly enlarged
      const double EPS = 2.0e-6;
      const double *c = coords + 3 * (size_t)atom0;
      double mn[3], mx[3];
      for (int x = 0; x < 3; ++x) {
        mn[x] = c[x];
        mx[x] = c[x];
        for (int i = 1; i < n; ++i) {
          if (c[3 * i + x] < mn[x]) mn[x] = c[3 * i + x];
          if (c[3 * i + x] > mx[x]) mx[x] = c[3 * i + x];
        }
      }
      for (int x = 0; x < 3; ++x) {
        d.L[4 * x] = mx[x] - mn[x] + EPS;
        d.shift[x] = mn[x] - 0.5 * EPS;
      }
    }
    if (lattice_coords) {
      for (int k = 0; k < 9; ++k) d.B[k] = (k % 4 == 0) ? 1.0 : 0.0;
      d.frac_div = 1.0;
    } else {
      recip_lattice(d.L, d.B);
      d.frac_div = 2.0 * PI_GEO;
    }
    d.rc_list = rc_max + EPS_LIST;
    d.Rc2 = (rc_max + EPS_LIST) * (rc_max + EPS_LIST);
    double grid[9];
    cell_multiples(rc_max, d.L, grid, d.nc);
    long nc = (long)d.nc[0] * d.nc[1] * d.nc[2];
    if (nc > (1L << 24)) {
      err = "cell grid too large";
      return AENET_STATUS_ERR_CAPACITY;
    }
    d.cell0 = plan.ncells;
    d.ncells = (int)nc;
    for (int a = 1; a <= d.nc[0]; ++a)
      for (int b = 1; b <= d.nc[1]; ++b)
        for (int c = 1; c <= d.nc[2]; ++c) {
          plan.cell_struct.push_back(s);
          plan.cell_ijk.push_back(a);
          plan.cell_ijk.push_back(b);
          plan.cell_ijk.push_back(c);
          double f[3] = {((double)a - 0.5) / d.nc[0], ((double)b - 0.5) / d.nc[1], ((double)c - 0.5) / d.nc[2]};
          for (int r = 0; r < 3; ++r) plan.cell_ctr.push_back(d.L[r] * f[0] + d.L[3 + r] * f[1] + d.L[6 + r] * f[2]);
        }
    // circumradius of a grid cell: half its longest body diagonal
    {
      double rad = 0.0;
      const int sg[4][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}};
      for (int k = 0; k < 4; ++k) {
        double v[3] = {0, 0, 0};
        for (int q = 0; q < 3; ++q)
          for (int r = 0; r < 3; ++r) v[r] += sg[k][q] * grid[3 * q + r];
        rad = std::max(rad, 0.5 * std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      }
      d.cell_rad = rad * (1.0 + 1.0e-12) + 1.0e-9;
      // plane spacing of family k: cell volume / area of the face spanned
      // by the other two grid vectors
      const double *g0 = grid, *g1 = grid + 3, *g2 = grid + 6;
      const double *gv[3] = {g0, g1, g2};
      double x01[3] = {g0[1] * g1[2] - g0[2] * g1[1], g0[2] * g1[0] - g0[0] * g1[2], g0[0] * g1[1] - g0[1] * g1[0]};
      const double vol = std::fabs(x01[0] * g2[0] + x01[1] * g2[1] + x01[2] * g2[2]);
      for (int k = 0; k < 3; ++k) {
        const double *u = gv[(k + 1) % 3], *w = gv[(k + 2) % 3];
        const double c[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        const double area = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        d.cell_gap[k] = area > 0.0 ? vol / area : 0.0;
      }
    }
    plan.ncells += (int)nc;
    std::vector<int> C, T;
    translation_vectors(rc_max, grid, d.nc, d.pbc, C);
    d.nC = (int)C.size() / 3;
    d.C0 = (int)plan.Cvec.size() / 3;
    plan.Cvec.insert(plan.Cvec.end(), C.begin(), C.end());
    d.nT = 0;
    d.T0 = (int)plan.Tvec.size() / 3;
    d.nself = 0;
    d.self0 = (int)plan.self_t.size();
    if (d.pbc) {
      translation_vectors(rc_max, d.L, nullptr, 1, T);
      d.nT = (int)T.size() / 3;
      plan.Tvec.insert(plan.Tvec.end(), T.begin(), T.end());
      // periodic images of the central atom within the cutoff
      for (int iT = 0; iT < d.nT; ++iT) {
        double Tv[3] = {(double)T[3 * iT], (double)T[3 * iT + 1], (double)T[3 * iT + 2]};
        double cart[3];
        for (int r = 0; r < 3; ++r) {
          double acc = d.L[r] * Tv[0] + d.L[3 + r] * Tv[1];
          cart[r] = acc + d.L[6 + r] * Tv[2];
        }
        double d2 = cart[0] * cart[0] + cart[1] * cart[1];
        d2 = d2 + cart[2] * cart[2];
        if (d2 <= d.Rc2) {
          plan.self_t.push_back(iT + 1);
          plan.self_d.push_back(std::sqrt(d2));
          d.nself++;
        }
      }
    }
    for (int i = 0; i < n; ++i) plan.atom_struct.push_back(s);
    atom0 += n;
  }
  plan.natoms = atom0;
  return AENET_STATUS_OK;
}

}  // namespace aenet_amd

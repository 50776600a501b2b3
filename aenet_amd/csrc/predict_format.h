// predict_format.h -- the structure record and the printed blocks of
// predict.x (src/predict.F90:439-568, src/aeio.f90:90-181), shared by the
// driver (predict.cpp).  A line of Fortran list-directed output starts
// with one blank; F editing maps to printf's %w.df (a value that rounds
// to zero keeps its sign in both).
#pragma once

#include <cmath>
#include <cstdio>
#include <ctime>
#include <string>
#include <vector>

namespace predict_fmt {

const double PI_GEO = 3.141592653589793;  // src/constants.f90:30

inline std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

struct Structure {
  std::string file;
  bool pbc = false;
  double L[9] = {0};     // lattice vector k at L[3k..3k+2] (bounding box for isolated structures)
  double origin[3] = {0, 0, 0};
  std::vector<std::string> names;   // species in order of first appearance
  std::vector<int> xsf_type;        // 1-based into names
  std::vector<int> type;            // 1-based into the potentials' species
  std::vector<double> cart;         // input Cartesian coordinates (3 per atom)
  int natoms() const { return (int)xsf_type.size(); }
};

// lattice coordinates of Cartesian x (geo_recip_lattice and the conversion
// of geo_init_xsf), wrapped into [0,1) for periodic structures as the
// neighbor list does (wrap_cooLatt, src/ext/lclist.f90:953-987); returns
// the Cartesian position the reference prints, L s + origin
inline void printed_position(const Structure &s, const double *x, double out[3]) {
  const double *a = s.L, *b = s.L + 3, *c = s.L + 6;
  double B[3][3];
  B[0][0] = b[1] * c[2] - b[2] * c[1];
  B[0][1] = b[2] * c[0] - b[0] * c[2];
  B[0][2] = b[0] * c[1] - b[1] * c[0];
  B[1][0] = -(a[1] * c[2] - a[2] * c[1]);
  B[1][1] = -(a[2] * c[0] - a[0] * c[2]);
  B[1][2] = -(a[0] * c[1] - a[1] * c[0]);
  B[2][0] = a[1] * b[2] - a[2] * b[1];
  B[2][1] = a[2] * b[0] - a[0] * b[2];
  B[2][2] = a[0] * b[1] - a[1] * b[0];
  const double V = a[0] * B[0][0] + a[1] * B[0][1] + a[2] * B[0][2];
  double y[3] = {x[0] - s.origin[0], x[1] - s.origin[1], x[2] - s.origin[2]};
  double f[3];
  for (int r = 0; r < 3; ++r) {
    double acc = 0.0;
    for (int q = 0; q < 3; ++q) acc += (B[r][q] / V * 2.0 * PI_GEO) * y[q];
    f[r] = acc / (2.0 * PI_GEO);
    if (s.pbc) {
      while (f[r] < 0.0) f[r] += 1.0;
      while (f[r] >= 1.0) f[r] -= 1.0;
    }
  }
  for (int r = 0; r < 3; ++r) out[r] = s.L[r] * f[0] + s.L[3 + r] * f[1] + s.L[6 + r] * f[2] + s.origin[r];
}

// ---------------------------------------------------------------------
// output blocks (src/predict.F90:439-568, src/aeio.f90:90-181).  A line
// of Fortran list-directed output starts with one blank.
// ---------------------------------------------------------------------

inline void header(const std::string &str, char c) {
  std::string bar(70, c);
  int l = (int)str.size();
  std::string mid = l > 70 ? "" : std::string((70 - l) / 2, ' ') + str;
  mid.resize(70, ' ');
  std::printf(" %s\n %s\n %s\n", bar.c_str(), mid.c_str(), bar.c_str());
}

inline std::string timestamp() {
  std::time_t t = std::time(nullptr);
  std::tm tm = *std::localtime(&t);
  char buf[64];
  std::snprintf(buf, sizeof buf, "%04d-%02d-%02d  %02d:%02d:%02d", tm.tm_year + 1900, tm.tm_mon + 1, tm.tm_mday,
                tm.tm_hour, tm.tm_min, tm.tm_sec);
  return buf;
}

inline void print_fileinfo(int istruc, const Structure &s, int ntypes) {
  std::printf(" Structure number  : %d\n", istruc);
  std::printf(" File name         : %s\n", trim(s.file).c_str());
  std::printf(" Number of atoms   : %d\n", s.natoms());
  std::printf(" Number of species : %d\n", ntypes);
  std::printf(" \n");
  std::printf(" Lattice vectors (Angstrom):\n");
  std::printf(" \n");
  const char *lab = "abc";
  for (int k = 0; k < 3; ++k)
    std::printf("   %c = (   %15.8f  %15.8f  %15.8f )\n", lab[k], s.L[3 * k], s.L[3 * k + 1], s.L[3 * k + 2]);
  std::printf(" \n");
}

// Fortran F editing of a value that rounds to zero keeps its sign; so does
// printf, which matches.
inline void print_coordinates(int iter, const Structure &s, const double *x, const double *F, const double *E_atom) {
  std::string h = "Cartesian atomic coordinates";
  h += iter == 0 ? " (input)" : " (optimized)";
  if (F) h += " and corresponding atomic forces";
  h += ":";
  std::printf(" %s\n \n", h.c_str());
  std::printf("   %14s%14s%14s", "     x      ", "     y      ", "     z      ");
  if (F) std::printf("%14s%14s%14s", "     Fx     ", "    Fy      ", "    Fz      ");
  if (E_atom) std::printf("%14s", "   E_atom   ");
  std::printf("\n");
  std::printf("   %14s%14s%14s", "    (Ang)   ", "    (Ang)   ", "    (Ang)   ");
  if (F) std::printf("%14s%14s%14s", "  (eV/Ang)  ", "  (eV/Ang)  ", "  (eV/Ang)  ");
  if (E_atom) std::printf("%14s", "    (eV)    ");
  std::printf("\n");
  std::printf(" %s", std::string(44, '-').c_str());
  if (F) std::printf("%s", std::string(42, '-').c_str());
  if (E_atom) std::printf("%s", std::string(14, '-').c_str());
  std::printf("\n");
  for (int i = 0; i < s.natoms(); ++i) {
    double p[3];
    printed_position(s, x + 3 * i, p);
    std::string sym = s.names[s.xsf_type[i] - 1];
    sym.resize(2, ' ');
    std::printf(" %s  %12.6f  %12.6f  %12.6f", sym.c_str(), p[0], p[1], p[2]);
    if (F) std::printf("  %12.6f  %12.6f  %12.6f", F[3 * i], F[3 * i + 1], F[3 * i + 2]);
    if (E_atom) std::printf("  %12.6f", E_atom[i]);
    std::printf("\n");
  }
  std::printf(" \n");
}

struct ForceStats {
  double mav[3], max[3], avg[3], rms;
  int imax;
};

// calc_rms_force (src/predict.F90:709-749)
inline ForceStats force_stats(int n, const double *F) {
  ForceStats st;
  for (int x = 0; x < 3; ++x) st.mav[x] = st.max[x] = st.avg[x] = 0.0;
  st.rms = 0.0;
  st.imax = 0;
  double a2max = 0.0;
  for (int i = 0; i < n; ++i) {
    double a2 = 0.0;
    for (int x = 0; x < 3; ++x) {
      st.avg[x] += F[3 * i + x];
      st.mav[x] += std::fabs(F[3 * i + x]);
      a2 += F[3 * i + x] * F[3 * i + x];
    }
    st.rms += a2;
    if (a2 > a2max) {
      a2max = a2;
      for (int x = 0; x < 3; ++x) st.max[x] = F[3 * i + x];
      st.imax = i + 1;
    }
  }
  for (int x = 0; x < 3; ++x) {
    st.avg[x] /= n;
    st.mav[x] /= n;
  }
  st.rms = std::sqrt(st.rms / n);
  return st;
}

inline void print_energy(double Ecoh, double Etot, int n, const double *F) {
  std::printf(" Cohesive energy            :  %20.8f eV\n", Ecoh);
  std::printf(" Total energy               :  %20.8f eV\n", Etot);
  if (F) {
    ForceStats st = force_stats(n, F);
    std::printf(" Mean force (must be zero)  :  %12.6f  %12.6f  %12.6f\n", st.avg[0], st.avg[1], st.avg[2]);
    std::printf(" Mean absolute force        :  %12.6f  %12.6f  %12.6f\n", st.mav[0], st.mav[1], st.mav[2]);
    std::printf(" Maximum force              :  %12.6f  %12.6f  %12.6f\n", st.max[0], st.max[1], st.max[2]);
    std::printf(" RMS force                  :  %12.6f\n", st.rms);
    std::printf(" The maximum force is acting on atom %d.\n", st.imax);
    std::printf(" All forces are given in eV/Angstrom.\n");
  }
  std::printf(" \n");
}

}  // namespace predict_fmt

// predict.cpp -- predict.x on the GPU: energies, forces and relaxations of
// a list of XSF structures with the potentials of a predict.in file.
//
//   predict.x <predict.in> [<structure.xsf>]
//
// Input and output follow the reference program (src/predict.F90): the
// predict.in keywords of read_InpPredict (src/input.f90:358-400), the XSF
// reader of src/ext/xsflib.f90:109-208, and the printed blocks of
// print_fileinfo, print_coordinates and print_energy
// (src/predict.F90:439-568).  Unlike the reference, which evaluates one
// structure after another, every structure of the list (or every structure
// still relaxing) goes to the GPU in one batched call
// (aenet_structures_energy_and_forces); the blocks are then printed in the
// list's order.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "aenet.h"
#include "aenet_amd.h"
#include "predict_format.h"

namespace {

using namespace predict_fmt;

std::string lower(std::string s) {
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

std::vector<std::string> split(const std::string &s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string w;
  while (is >> w) out.push_back(w);
  return out;
}

// Fortran logical / real literals as the reference's list-directed reads
// take them
bool parse_logical(const std::string &w, bool &v) {
  std::string s = lower(w);
  if (!s.empty() && s[0] == '.') s = s.substr(1);
  if (s.empty()) return false;
  if (s[0] == 't') v = true;
  else if (s[0] == 'f') v = false;
  else return false;
  return true;
}

double parse_real(const std::string &w) {
  std::string s = w;
  for (char &c : s)
    if (c == 'd' || c == 'D') c = 'e';
  return std::strtod(s.c_str(), nullptr);
}

// ---------------------------------------------------------------------
// predict.in (src/input.f90:358-400; lines are read as aeio_readline
// does, src/aeio.f90:187-213: trimmed, and skipped when blank or starting
// with '!', '#' or '%'; keywords are the first word, case-insensitive)
// ---------------------------------------------------------------------

struct Input {
  std::vector<std::string> types, nets;
  bool forces = false, timing = false, atomic_energies = false, debug = false;
  bool relax = false;
  std::string relax_method = "bfgs";
  int relax_steps = 99;
  double F_conv = 0.01, E_conv = 0.001;
  double dmax[3] = {1.0, 1.0, 1.0};
  std::vector<std::string> files;
  int verbosity = 1;
};

std::vector<std::string> content_lines(const std::string &path) {
  std::ifstream f(path);
  std::vector<std::string> out;
  std::string l;
  while (std::getline(f, l)) {
    std::string t = trim(l);
    if (t.empty() || t[0] == '!' || t[0] == '#' || t[0] == '%') continue;
    out.push_back(t);
  }
  return out;
}

int find_keyword(const std::vector<std::string> &L, const std::string &kw) {
  for (size_t i = 0; i < L.size(); ++i) {
    std::vector<std::string> w = split(L[i]);
    if (!w.empty() && lower(w[0]) == kw) return (int)i;
  }
  return -1;
}

// io_readval: "name=value" anywhere in the line (case-insensitive name)
bool readval(const std::string &line, const std::string &name, std::string &val) {
  std::string l = lower(line), n = lower(name) + "=";
  size_t p = l.find(n);
  if (p == std::string::npos) return false;
  std::istringstream is(line.substr(p + n.size()));
  return (bool)(is >> val);
}

bool read_input(const std::string &path, Input &inp, std::string &err) {
  std::ifstream test(path);
  if (!test) {
    err = "Error: File not found: " + path;
    return false;
  }
  std::vector<std::string> L = content_lines(path);
  int k = find_keyword(L, "types");
  if (k < 0 || k + 1 >= (int)L.size()) {
    err = "Error: no atom TYPES specified in file `" + path + "'";
    return false;
  }
  int nt = std::atoi(L[k + 1].c_str());
  for (int i = 0; i < nt && k + 2 + i < (int)L.size(); ++i) inp.types.push_back(split(L[k + 2 + i])[0]);
  if ((int)inp.types.size() != nt || nt < 1) {
    err = "Error: incomplete TYPES block in `" + path + "'";
    return false;
  }
  k = find_keyword(L, "networks");
  if (k < 0) {
    err = "Error: no NETWORKS specified in file `" + path + "'";
    return false;
  }
  inp.nets.assign(nt, "");
  for (int i = 0; i < nt && k + 1 + i < (int)L.size(); ++i) {
    std::vector<std::string> w = split(L[k + 1 + i]);
    if (w.size() < 2) continue;
    for (int t = 0; t < nt; ++t)
      if (lower(w[0]) == lower(inp.types[t])) inp.nets[t] = w[1];
  }
  for (int t = 0; t < nt; ++t)
    if (inp.nets[t].empty()) {
      err = "Error: no network for species " + inp.types[t];
      return false;
    }
  auto flag = [&](const char *kw, bool &dst) {
    int i = find_keyword(L, kw);
    if (i < 0) return;
    std::vector<std::string> w = split(L[i]);
    bool v = true;
    if (w.size() > 1 && parse_logical(w[1], v)) dst = v;
    else dst = true;
  };
  flag("forces", inp.forces);
  flag("timing", inp.timing);
  flag("print_atomic_energies", inp.atomic_energies);
  flag("debug", inp.debug);
  k = find_keyword(L, "relax");
  if (k >= 0) {
    inp.relax = true;
    inp.forces = true;
    if (k + 1 < (int)L.size()) {
      const std::string &opt = L[k + 1];
      std::string v;
      if (readval(opt, "method", v)) inp.relax_method = lower(v);
      if (readval(opt, "F_conv", v)) inp.F_conv = parse_real(v);
      if (readval(opt, "E_conv", v)) inp.E_conv = parse_real(v);
      if (readval(opt, "steps", v)) inp.relax_steps = std::atoi(v.c_str());
      std::string l = lower(opt);
      size_t p = l.find("dmax=");
      if (p != std::string::npos) {
        std::istringstream is(opt.substr(p + 5));
        std::string a, b, c;
        if (is >> a >> b >> c) {
          inp.dmax[0] = parse_real(a);
          inp.dmax[1] = parse_real(b);
          inp.dmax[2] = parse_real(c);
        }
      }
    }
  }
  k = find_keyword(L, "files");
  if (k >= 0 && k + 1 < (int)L.size()) {
    int nf = std::atoi(L[k + 1].c_str());
    for (int i = 0; i < nf && k + 2 + i < (int)L.size(); ++i) inp.files.push_back(L[k + 2 + i]);
  }
  k = find_keyword(L, "verbosity");
  if (k >= 0) {
    std::vector<std::string> w = split(L[k]);
    std::string v = w.size() > 1 ? lower(w[1]) : "normal";
    if (v == "low") inp.verbosity = 0;
    else if (v == "normal") inp.verbosity = 1;
    else if (v == "high") inp.verbosity = 2;
  }
  return true;
}

// ---------------------------------------------------------------------
// XSF structures (src/ext/xsflib.f90:109-208 and geo_init_xsf,
// src/geometry.f90:661-736)
// ---------------------------------------------------------------------

bool read_atoms(std::ifstream &f, Structure &s, int expect) {
  std::string l;
  while ((expect < 0 || s.natoms() < expect) && std::getline(f, l)) {
    if (trim(l).empty()) {
      if (expect < 0) break;
      continue;
    }
    std::vector<std::string> w = split(l);
    if (w.size() < 4) return false;
    std::string el = w[0];
    int t = -1;
    for (size_t i = 0; i < s.names.size(); ++i)
      if (s.names[i] == el) t = (int)i + 1;
    if (t < 0) {
      s.names.push_back(el);
      t = (int)s.names.size();
    }
    s.xsf_type.push_back(t);
    for (int x = 0; x < 3; ++x) s.cart.push_back(parse_real(w[1 + x]));
  }
  return expect < 0 || s.natoms() == expect;
}

bool read_xsf(const std::string &path, Structure &s, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "Error: File not found: " + path;
    return false;
  }
  s.file = path;
  std::string l;
  while (std::getline(f, l)) {
    std::string t = trim(l);
    if (t.empty() || t[0] == '#') continue;
    if (t.find("PRIMVEC") != std::string::npos) {
      s.pbc = true;
      for (int k = 0; k < 3; ++k) {
        if (!std::getline(f, l)) break;
        std::vector<std::string> w = split(l);
        for (int x = 0; x < 3 && x < (int)w.size(); ++x) s.L[3 * k + x] = parse_real(w[x]);
      }
      continue;
    }
    if (t.find("PRIMCOORD") != std::string::npos) {
      if (!std::getline(f, l)) break;
      int n = std::atoi(trim(l).c_str());
      if (!read_atoms(f, s, n)) {
        err = "Error: invalid input file: " + path;
        return false;
      }
      continue;
    }
    if (t.find("ATOMS") != std::string::npos) {
      read_atoms(f, s, -1);
      continue;
    }
  }
  if (s.natoms() == 0) {
    err = "Error: invalid input file: " + path;
    return false;
  }
  if (!s.pbc) {
    // geo_get_bounds (src/geometry.f90:538-580)
    const double EPS = 2.0e-6;
    for (int x = 0; x < 3; ++x) {
      double mn = s.cart[x], mx = s.cart[x];
      for (int i = 1; i < s.natoms(); ++i) {
        mn = std::min(mn, s.cart[3 * i + x]);
        mx = std::max(mx, s.cart[3 * i + x]);
      }
      s.L[4 * x] = mx - mn + EPS;
      s.origin[x] = mn - 0.5 * EPS;
    }
  }
  return true;
}

// ---------------------------------------------------------------------
// batched evaluation
// ---------------------------------------------------------------------

struct Result {
  double Ecoh = 0.0, Etot = 0.0;
  std::vector<double> E, F;
};

// energies (and forces) of the structures sel[] at coordinates X[s], in
// batches of up to ~1M atoms per GPU call
bool evaluate(const std::vector<Structure> &S, const std::vector<int> &sel, const std::vector<std::vector<double>> &X,
              bool forces, std::vector<Result> &R, std::string &err) {
  const long MAX_ATOMS = 1 << 20;
  size_t b = 0;
  while (b < sel.size()) {
    size_t e = b;
    long na = 0;
    while (e < sel.size() && (e == b || na + S[sel[e]].natoms() <= MAX_ATOMS)) na += S[sel[e++]].natoms();
    const int ns = (int)(e - b);
    std::vector<int> nat(ns), pbc(ns), types;
    std::vector<double> latt(9 * (size_t)ns), coords;
    for (int k = 0; k < ns; ++k) {
      const Structure &s = S[sel[b + k]];
      nat[k] = s.natoms();
      pbc[k] = s.pbc ? 1 : 0;
      std::memcpy(&latt[9 * (size_t)k], s.L, sizeof s.L);
      types.insert(types.end(), s.type.begin(), s.type.end());
      coords.insert(coords.end(), X[sel[b + k]].begin(), X[sel[b + k]].end());
    }
    std::vector<double> Et(ns), Ec(ns), Ea(na), F(forces ? 3 * (size_t)na : 0);
    int stat = 0;
    aenet_structures_energy_and_forces(ns, nat.data(), latt.data(), pbc.data(), types.data(), coords.data(),
                                       Et.data(), Ec.data(), Ea.data(), forces ? F.data() : nullptr, &stat);
    if (stat != 0) {
      err = std::string("Error: evaluation failed: ") + aenet_amd_last_error();
      return false;
    }
    size_t a0 = 0;
    for (int k = 0; k < ns; ++k) {
      Result &r = R[sel[b + k]];
      r.Ecoh = Ec[k];
      r.Etot = Et[k];
      r.E.assign(Ea.begin() + a0, Ea.begin() + a0 + nat[k]);
      if (forces) r.F.assign(F.begin() + 3 * a0, F.begin() + 3 * (a0 + nat[k]));
      a0 += nat[k];
    }
    b = e;
  }
  return true;
}

// ---------------------------------------------------------------------
// relaxation: limited-memory BFGS per structure in Cartesian coordinates
// (gradient -F), with the step of each lattice coordinate capped at 0.1
// as predict.x caps it (src/predict.F90:260-261) and backtracking on the
// Armijo condition; every evaluation is one iteration of the RELAX loop
// (src/predict.F90:234-286).  Converged when every force component is
// within F_conv (the projected-gradient test of the reference's L-BFGS-B
// with pgtol = F_conv).
// ---------------------------------------------------------------------

struct Relax {
  std::vector<double> x, g;            // accepted point and its gradient
  double E = 0.0;
  std::vector<std::vector<double>> s_hist, y_hist;
  std::vector<double> d;               // search direction
  double alpha = 1.0, gd = 0.0;
  bool conv = false, done = false;
  int iter = 0;
  double E0 = 0.0, Frms_prev = 0.0;
  std::string log;
};

void lbfgs_direction(Relax &r) {
  const size_t n = r.g.size();
  std::vector<double> q(r.g);
  const int m = (int)r.s_hist.size();
  std::vector<double> a(m), rho(m);
  for (int k = m - 1; k >= 0; --k) {
    double sy = 0.0, sq = 0.0;
    for (size_t i = 0; i < n; ++i) {
      sy += r.s_hist[k][i] * r.y_hist[k][i];
      sq += r.s_hist[k][i] * q[i];
    }
    rho[k] = 1.0 / sy;
    a[k] = rho[k] * sq;
    for (size_t i = 0; i < n; ++i) q[i] -= a[k] * r.y_hist[k][i];
  }
  double gamma = 1.0;
  if (m > 0) {
    double sy = 0.0, yy = 0.0;
    for (size_t i = 0; i < n; ++i) {
      sy += r.s_hist[m - 1][i] * r.y_hist[m - 1][i];
      yy += r.y_hist[m - 1][i] * r.y_hist[m - 1][i];
    }
    gamma = sy / yy;
  } else {
    // first step: unit-free start, the largest force moves 0.1 A
    double gmax = 0.0;
    for (double v : r.g) gmax = std::max(gmax, std::fabs(v));
    gamma = gmax > 0.0 ? 0.1 / gmax : 1.0;
  }
  for (size_t i = 0; i < n; ++i) q[i] *= gamma;
  for (int k = 0; k < m; ++k) {
    double yq = 0.0;
    for (size_t i = 0; i < n; ++i) yq += r.y_hist[k][i] * q[i];
    const double bk = rho[k] * yq;
    for (size_t i = 0; i < n; ++i) q[i] += r.s_hist[k][i] * (a[k] - bk);
  }
  r.d.assign(n, 0.0);
  double gd = 0.0;
  for (size_t i = 0; i < n; ++i) {
    r.d[i] = -q[i];
    gd += r.g[i] * r.d[i];
  }
  if (!(gd < 0.0)) {  // not a descent direction: restart from steepest descent
    r.s_hist.clear();
    r.y_hist.clear();
    double gmax = 0.0;
    for (double v : r.g) gmax = std::max(gmax, std::fabs(v));
    const double sc = gmax > 0.0 ? 0.1 / gmax : 1.0;
    gd = 0.0;
    for (size_t i = 0; i < n; ++i) {
      r.d[i] = -sc * r.g[i];
      gd += r.g[i] * r.d[i];
    }
  }
  r.gd = gd;
  r.alpha = 1.0;
}

// largest step fraction that keeps every lattice-coordinate change <= cap
double step_cap(const Structure &s, const std::vector<double> &d, double cap) {
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
  double mx = 0.0;
  for (size_t i = 0; i < d.size() / 3; ++i)
    for (int r = 0; r < 3; ++r) {
      double f = (B[r][0] * d[3 * i] + B[r][1] * d[3 * i + 1] + B[r][2] * d[3 * i + 2]) / V;
      mx = std::max(mx, std::fabs(f));
    }
  return mx > cap ? cap / mx : 1.0;
}

void usage() {
  std::printf(" \n predict.x -- Predict/interpolate atomic energy.\n %s\n", std::string(70, '-').c_str());
  std::printf(" Usage: predict.x <input-file> [<structure files>]\n \n");
  std::printf(" See the documentation or the source code for a description of the\n");
  std::printf(" input file format.  Structure files can either be listed in the\n");
  std::printf(" input file, or specified on the command line.\n \n");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, " Error: No input file specified.\n");
    usage();
    return 1;
  }
  Input inp;
  std::string err;
  if (!read_input(argv[1], inp, err)) {
    std::fprintf(stderr, " %s\n", err.c_str());
    usage();
    return 1;
  }
  if (argc > 2) inp.files.assign(1, argv[2]);
  if (inp.relax && inp.relax_method != "bfgs") {
    std::fprintf(stderr, " Error: unknown relaxation method: %s\n", inp.relax_method.c_str());
    return 1;
  }

  // potentials (aenet_init / aenet_load_potential, src/predict.F90:367-385)
  const int nt = (int)inp.types.size();
  std::vector<char *> names(nt), files(nt);
  for (int t = 0; t < nt; ++t) {
    names[t] = const_cast<char *>(inp.types[t].c_str());
    files[t] = const_cast<char *>(inp.nets[t].c_str());
  }
  int stat = 0;
  aenet_init(nt, names.data(), &stat);
  if (stat != 0) {
    std::fprintf(stderr, " Error: aenet initialization failed\n");
    return 1;
  }
  for (int t = 0; t < nt; ++t) {
    aenet_load_potential(t + 1, files[t], &stat);
    if (stat != 0) {
      std::fprintf(stderr, " Error: could not load ANN potentials\n");
      aenet_final(&stat);
      return 1;
    }
  }
  if (inp.verbosity > 0) {
    header("Atomic Energy Network Interpolation", '=');
    header(timestamp(), ' ');
    std::printf(" \n");
    std::printf(" Copyright (C) 2015-2018 Nongnuch Artrith and Alexander Urban\n \n");
    std::printf(" This program is distributed in the hope that it will be useful, but\n");
    std::printf(" WITHOUT ANY WARRANTY; without even the implied warranty of\n");
    std::printf(" MERCHANTABILITY or FITNESS FOR A PARTICULAR PURPOSE.  See the\n");
    std::printf(" Mozilla Public License, v. 2.0, for more details.\n \n");
    std::fflush(stdout);
    aenet_print_info();
    std::fflush(stdout);
  }
  if (inp.files.empty()) {
    std::fprintf(stderr, " Error: no input structures specified in %s\n", argv[1]);
    aenet_final(&stat);
    return 1;
  }

  // read every structure; species names -> potential type ids
  std::vector<Structure> S(inp.files.size());
  for (size_t k = 0; k < S.size(); ++k) {
    if (!read_xsf(inp.files[k], S[k], err)) {
      std::fprintf(stderr, " %s\n", err.c_str());
      aenet_final(&stat);
      return 1;
    }
    Structure &s = S[k];
    s.type.resize(s.natoms());
    for (int i = 0; i < s.natoms(); ++i) {
      const std::string &el = s.names[s.xsf_type[i] - 1];
      int id = -1;
      for (int t = 0; t < nt; ++t)
        if (inp.types[t] == el) id = t + 1;
      if (id < 0) {
        std::fprintf(stderr, " Error: unknown species %s in %s\n", el.c_str(), s.file.c_str());
        aenet_final(&stat);
        return 1;
      }
      s.type[i] = id;
    }
  }
  const size_t ns = S.size();
  std::vector<std::vector<double>> X(ns), X0(ns);
  for (size_t k = 0; k < ns; ++k) X[k] = X0[k] = S[k].cart;
  std::vector<int> all(ns);
  for (size_t k = 0; k < ns; ++k) all[k] = (int)k;
  std::vector<Result> R(ns);
  const auto t0 = std::chrono::steady_clock::now();
  if (!evaluate(S, all, X, inp.forces, R, err)) {
    std::fprintf(stderr, " %s\n", err.c_str());
    aenet_final(&stat);
    return 1;
  }
  std::vector<Result> R0 = R;

  // relaxation of every structure at once: one batched call per iteration
  std::vector<Relax> RX;
  if (inp.relax) {
    RX.resize(ns);
    for (size_t k = 0; k < ns; ++k) {
      Relax &r = RX[k];
      r.x = X[k];
      r.E = r.E0 = R[k].Ecoh;
      r.g.resize(R[k].F.size());
      for (size_t i = 0; i < r.g.size(); ++i) r.g[i] = -R[k].F[i];
    }
    for (int it = 1; it <= inp.relax_steps; ++it) {
      std::vector<int> act;
      for (size_t k = 0; k < ns; ++k) {
        Relax &r = RX[k];
        if (r.done) continue;
        r.iter = it;
        const Result &cur = R[k];
        ForceStats st = force_stats(S[k].natoms(), cur.F.data());
        char line[160];
        if (it > 1)
          std::snprintf(line, sizeof line, " %5d  %16.8f  %16.8f  %16.8f\n", it, cur.Ecoh - r.E0, st.rms,
                        st.rms - r.Frms_prev);
        else
          std::snprintf(line, sizeof line, " %5d  %16.8f  %16.8f\n", it, cur.Ecoh - r.E0, st.rms);
        r.log += line;
        r.Frms_prev = st.rms;
        // the point just evaluated: accept it, or shorten the step
        bool first = r.d.empty();
        bool accept = first || cur.Ecoh <= r.E + 1e-4 * r.alpha * r.gd || r.alpha < 1e-6;
        if (!first && accept) {
          std::vector<double> sv(r.x.size()), yv(r.x.size());
          double sy = 0.0;
          for (size_t i = 0; i < r.x.size(); ++i) {
            sv[i] = X[k][i] - r.x[i];
            yv[i] = -cur.F[i] - r.g[i];
            sy += sv[i] * yv[i];
          }
          if (sy > 1e-12) {
            r.s_hist.push_back(sv);
            r.y_hist.push_back(yv);
            if (r.s_hist.size() > 20) {
              r.s_hist.erase(r.s_hist.begin());
              r.y_hist.erase(r.y_hist.begin());
            }
          }
          r.x = X[k];
          r.E = cur.Ecoh;
          for (size_t i = 0; i < r.g.size(); ++i) r.g[i] = -cur.F[i];
        }
        if (accept) {
          double gmax = 0.0;
          for (double v : r.g) gmax = std::max(gmax, std::fabs(v));
          if (gmax <= inp.F_conv) {
            r.conv = r.done = true;
            X[k] = r.x;
            continue;
          }
          lbfgs_direction(r);
          r.alpha = step_cap(S[k], r.d, 0.1);
        } else {
          r.alpha *= 0.5;
        }
        for (size_t i = 0; i < r.x.size(); ++i) X[k][i] = r.x[i] + r.alpha * r.d[i];
        act.push_back((int)k);
      }
      if (act.empty()) break;
      if (!evaluate(S, act, X, true, R, err)) {
        std::fprintf(stderr, " %s\n", err.c_str());
        aenet_final(&stat);
        return 1;
      }
    }
    // the last evaluated point of a structure that did not converge
    for (size_t k = 0; k < ns; ++k) RX[k].done = true;
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  // output, structure by structure, in the reference's order
  if (inp.verbosity > 0) {
    header("Energy evaluation", '-');
    std::printf(" \n");
  }
  for (size_t k = 0; k < ns; ++k) {
    const Structure &s = S[k];
    if (inp.verbosity > 0) print_fileinfo((int)k + 1, s, (int)s.names.size());
    const Result &r0 = R0[k];
    print_coordinates(0, s, X0[k].data(), inp.forces ? r0.F.data() : nullptr,
                      inp.atomic_energies ? r0.E.data() : nullptr);
    const Result *fin = &r0;
    if (inp.relax) {
      const Relax &rx = RX[k];
      if (inp.verbosity > 0) {
        std::printf(" Geometry optimization:\n \n");
        std::printf("        %s%s\n", "   energy change  ", "     rms force    ");
        std::printf("        %s%s\n", "        (eV)      ", "     (eV/Ang)     ");
        std::printf(" %s\n", std::string(60, '-').c_str());
        std::printf("%s", rx.log.c_str());
        if (rx.conv) std::printf("    converged after %d iterations.\n", rx.iter);
      }
      std::printf(" \n");
      fin = &R[k];
      print_coordinates(rx.iter, s, X[k].data(), R[k].F.data(), nullptr);
    }
    if (inp.forces)
      print_energy(fin->Ecoh, fin->Etot, s.natoms(), fin->F.data());
    else
      print_energy(fin->Ecoh, fin->Etot, s.natoms(), nullptr);
    if (k + 1 < ns) std::printf(" %s\n \n", std::string(70, '+').c_str());
  }
  if (inp.timing)
    std::fprintf(stderr, " predict.x: %zu structures evaluated on the GPU in %.3f s\n", ns, secs);
  if (inp.verbosity > 0) {
    header(timestamp(), ' ');
    header("Atomic Energy Network done.", '=');
  }
  aenet_final(&stat);
  return 0;
}

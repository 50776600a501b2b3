// model.cpp -- .nn reader and per-species parameter packing.
//
// File layout: load_Network (src/ext/feedforward.f90:291-306), load_Setup
// (src/sfsetup.f90:585-609), load_TrnSet_info (src/trainset.f90:430-440);
// derived energy constants as load_NNPot (src/potential.f90:145-151).
#include <algorithm>
#include "model.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

#include "aenet_capi_status.h"

namespace aenet_amd {

namespace {

struct RecReader {
  std::vector<unsigned char> buf;
  size_t pos = 0;
  bool bad = false;

  const unsigned char *next(size_t &n) {
    int32_t a, b;
    if (bad || pos + 4 > buf.size()) { bad = true; return nullptr; }
    std::memcpy(&a, buf.data() + pos, 4);
    if (a < 0 || pos + 8 + (size_t)a > buf.size()) { bad = true; return nullptr; }
    std::memcpy(&b, buf.data() + pos + 4 + a, 4);
    if (a != b) { bad = true; return nullptr; }
    const unsigned char *p = buf.data() + pos + 4;
    pos += 8 + (size_t)a;
    n = (size_t)a;
    return p;
  }
  template <class T>
  bool vec(std::vector<T> &out, size_t count) {
    size_t n;
    const unsigned char *p = next(n);
    if (!p || n != count * sizeof(T)) { bad = true; return false; }
    out.resize(count);
    std::memcpy(out.data(), p, n);
    return true;
  }
  template <class T>
  bool one(T &v) {
    std::vector<T> t;
    if (!vec(t, 1)) return false;
    v = t[0];
    return true;
  }
  bool strs(std::vector<std::string> &out, size_t k) {
    size_t n;
    const unsigned char *p = next(n);
    if (!p || n % k) { bad = true; return false; }
    out.clear();
    for (size_t i = 0; i < n; i += k) {
      std::string s((const char *)p + i, k);
      while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
      out.push_back(s);
    }
    return true;
  }
  bool str(std::string &out, size_t k) {
    std::vector<std::string> v;
    if (!strs(v, k) || v.size() != 1) { bad = true; return false; }
    out = v[0];
    return true;
  }
  bool skip() {
    size_t n;
    return next(n) != nullptr;
  }
};

std::string trimmed(const std::string &s) {
  size_t a = s.find_first_not_of(' ');
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(' ');
  return s.substr(a, b - a + 1);
}

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

}  // namespace

bool name_eq(const std::string &a, const std::string &b) { return trimmed(a) == trimmed(b); }

int nmax_nbdist(double rmin, double rmax) {
  // lcl_nmax_nbdist (src/ext/lclist.f90:229-245)
  double V_atom = std::pow(0.5 * rmin, 3);
  double V_cut = std::pow(rmax + 0.5 * rmin, 3);
  return (int)std::ceil(V_cut / V_atom * 0.7405);
}

int read_nn(const std::string &path, PotHost &pot, std::string &err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return AENET_STATUS_ERR_IO;
  }
  RecReader r;
  r.buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  NetHost &n = pot.net;
  r.one(n.nlayers);
  r.one(n.nnodes_max);
  r.one(n.Wsize);
  r.one(n.nvalues);
  if (r.bad || n.nlayers < 2 || n.nlayers > AENET_AMD_MAX_LAYERS || n.Wsize <= 0) {
    err = path + ": not a .nn potential file (network header)";
    return AENET_STATUS_ERR_IO;
  }
  r.vec(n.nnodes, n.nlayers);
  std::vector<int> fa;
  r.vec(fa, n.nlayers - 1);
  n.fa.assign(n.nlayers, 0);
  for (int l = 1; l < n.nlayers; ++l) n.fa[l] = fa[l - 1];
  r.vec(n.iw, n.nlayers);
  r.vec(n.iv, n.nlayers);
  r.vec(n.W, n.Wsize);
  SetupHost &s = pot.stp;
  r.str(s.description, 1024);
  r.str(s.atomtype, 2);
  int nenv = 0;
  r.one(nenv);
  if (r.bad || nenv < 1 || nenv > AENET_AMD_MAX_TYPES) {
    err = path + ": bad ENV record";
    return AENET_STATUS_ERR_IO;
  }
  r.strs(s.envtypes, 2);
  r.one(s.Rc_min);
  r.one(s.Rc_max);
  r.str(s.sftype, 100);
  r.one(s.nsf);
  r.one(s.nsfparam);
  if (r.bad || s.nsf < 1 || s.nsf > AENET_AMD_MAX_NSF || s.nsfparam < 1) {
    err = path + ": bad descriptor header";
    return AENET_STATUS_ERR_IO;
  }
  r.vec(s.sf, s.nsf);
  r.vec(s.sfparam, (size_t)s.nsfparam * s.nsf);
  r.vec(s.sfenv, 2 * (size_t)s.nsf);
  r.one(s.neval);
  r.vec(s.vmin, s.nsf);
  r.vec(s.vmax, s.nsf);
  r.vec(s.avg, s.nsf);
  r.vec(s.cov, s.nsf);
  r.str(pot.ts_file, 1024);
  r.skip();  // normalized
  r.one(pot.ts_scale);
  r.one(pot.ts_shift);
  int nt = 0;
  r.one(nt);
  if (r.bad || nt < 1 || nt > 1024) {
    err = path + ": bad training-set record (a legacy 35-record file?)";
    return AENET_STATUS_ERR_IO;
  }
  r.strs(pot.ts_types, 2);
  r.vec(pot.ts_E_atom, nt);
  r.skip();
  r.skip();
  r.skip();
  if (r.bad) {
    err = path + ": truncated .nn file";
    return AENET_STATUS_ERR_IO;
  }
  for (int l = 0; l + 1 < n.nlayers; ++l)
    if (n.nnodes[l + 1] > AENET_AMD_MAX_HIDDEN && l + 1 < n.nlayers - 1) {
      err = path + ": hidden layers wider than 512 nodes are not supported";
      return AENET_STATUS_ERR_IO;
    }
  if (n.nnodes[n.nlayers - 1] != 1 || n.nnodes[0] != s.nsf) {
    err = path + ": network input/output does not match the descriptor";
    return AENET_STATUS_ERR_IO;
  }
  pot.E_scale = 1.0 / pot.ts_scale;
  pot.E_shift = pot.ts_shift;
  pot.E_atom = 0.0;
  for (int i = 0; i < nt; ++i)
    if (name_eq(pot.ts_types[i], s.atomtype)) {
      pot.E_atom = pot.ts_E_atom[i];
      break;
    }
  pot.file = path;
  pot.init = true;
  return AENET_STATUS_OK;
}

int PotSet::finalize(std::string &err) {
  // pot_get_range (src/potential.f90:244-266), lcl_nmax_nbdist, stp_init
  Rc_min = pots[0].stp.Rc_min;
  Rc_max = pots[0].stp.Rc_max;
  for (auto &p : pots) {
    Rc_min = std::min(Rc_min, p.stp.Rc_min);
    Rc_max = std::max(Rc_max, p.stp.Rc_max);
  }
  nnb_max = nmax_nbdist(Rc_min, Rc_max);
  nsf_max = 0;
  for (auto &p : pots) nsf_max = std::max(nsf_max, p.stp.nsf);
  std::string kind = lower(trimmed(pots[0].stp.sftype));
  for (auto &p : pots)
    if (lower(trimmed(p.stp.sftype)) != kind) {
      err = "mixing of descriptor types is not supported";
      return AENET_STATUS_ERR_IO;
    }
  if (kind == "chebyshev")
    desc = DESC_CHEBYSHEV;
  else if (kind == "behler2011")
    desc = DESC_BEHLER;
  else {
    err = "unknown descriptor type " + kind;
    return AENET_STATUS_ERR_IO;
  }
  // stp_set_global_types (src/sfsetup.f90:721-755)
  for (auto &p : pots) {
    p.stp.ltype.assign(ntypes + 1, 0);
    for (int i = 1; i <= ntypes; ++i)
      for (size_t j = 0; j < p.stp.envtypes.size(); ++j)
        if (name_eq(names[i - 1], p.stp.envtypes[j])) {
          p.stp.ltype[i] = (int)j + 1;
          break;
        }
  }
  species.assign(ntypes, SpeciesDev{});
  W.clear();
  norm_avg.clear();
  norm_scale.clear();
  max_rN = max_aN = 0;
  for (int t = 0; t < ntypes; ++t) {
    PotHost &p = pots[t];
    SpeciesDev &sd = species[t];
    sd.nsf = p.stp.nsf;
    sd.nlayers = p.net.nlayers;
    for (int l = 0; l < p.net.nlayers; ++l) {
      sd.nnodes[l] = p.net.nnodes[l];
      sd.fa[l] = p.net.fa[l];
      sd.w_off[l] = (int)W.size() + p.net.iw[l];
    }
    for (double w : p.net.W) W.push_back((float)w);
    // stp_normalize scaling (src/sfsetup.f90:1005-1031)
    sd.norm_off = (int)norm_avg.size();
    for (int i = 0; i < p.stp.nsf; ++i) {
      double shift = p.stp.avg[i];
      double sdv = std::sqrt(p.stp.cov[i] - shift * shift);
      norm_avg.push_back(shift);
      norm_scale.push_back(sdv <= 1.0e-10 ? 0.0 : 1.0 / sdv);
    }
    sd.E_scale = p.E_scale;
    sd.E_shift = p.E_shift;
    sd.E_atom = p.E_atom;
    int nenv = (int)p.stp.envtypes.size();
    if (desc == DESC_CHEBYSHEV) {
      // setup_basis_chebyshev (src/sfsetup.f90:1162-1179) + new_SFBasis
      // (src/ext/sfbasis.f90:83-151)
      double r_Rc = p.stp.sfparam[0];
      int r_order = (int)std::lround(p.stp.sfparam[1]);
      double a_Rc = p.stp.sfparam[2];
      int a_order = (int)std::lround(p.stp.sfparam[3]);
      sd.r_N = r_order + 1;
      sd.a_N = a_order + 1;
      sd.r_Rc = (float)r_Rc;
      sd.a_Rc = (float)a_Rc;
      sd.multi = nenv > 1;
      sd.r_i1 = 0;
      sd.a_i1 = sd.r_N;
      sd.r_i2 = sd.r_N + sd.a_N;
      sd.a_i2 = 2 * sd.r_N + sd.a_N;
      int N = (sd.r_N + sd.a_N) * (sd.multi ? 2 : 1);
      if (N != p.stp.nsf) {
        err = p.file + ": Chebyshev feature count does not match nsf";
        return AENET_STATUS_ERR_IO;
      }
      std::vector<double> typespin(nenv + 1, 0.0);
      int s = -nenv / 2;
      for (int i = 1; i <= nenv; ++i) {
        if (s == 0 && nenv % 2 == 0) s += 1;
        typespin[i] = (double)s;
        s += 1;
      }
      for (int g = 1; g <= ntypes; ++g) sd.spin[g - 1] = (float)typespin[p.stp.ltype[g]];
      max_rN = std::max(max_rN, sd.r_N);
      max_aN = std::max(max_aN, sd.a_N);
      if (!cheb_orders_ok(sd.r_N, sd.a_N)) {
        err = p.file + ": Chebyshev orders beyond radial 23 with angular 7, or both 15, are not supported";
        return AENET_STATUS_ERR_IO;
      }
      if (!cheb_orders_ok(max_rN, max_aN)) {
        err = p.file + ": the set's Chebyshev orders together (radial " + std::to_string(max_rN - 1) + ", angular " +
              std::to_string(max_aN - 1) + ") exceed radial 23 with angular 7, or both 15";
        return AENET_STATUS_ERR_IO;
      }
    } else {
      sd.bp_row = p.stp.ltype[t + 1] - 1;
      for (int g = 1; g <= ntypes; ++g) sd.bp_ltype[g - 1] = p.stp.ltype[g];
    }
  }
  if (desc == DESC_BEHLER) {
    // sf_init / sf_add_rad / sf_add_ang / sf_build_index
    // (src/ext/symmfunc.f90:91-305): rows are the global potential
    // indices (setup_symmfunc_Behler2011(itype, stp), src/sfsetup.f90:1298)
    const int MT = AENET_AMD_MAX_TYPES;
    std::memset(&bp, 0, sizeof(bp));
    bp.ntypes = ntypes;
    bpf.clear();
    bpk.clear();
    bpg.clear();
    // registration lists per (row, funct, t2[, t3])
    for (int row = 0; row < ntypes; ++row) {
      const SetupHost &s = pots[row].stp;
      int off = 0;
      for (int t2 = 0; t2 < ntypes; ++t2) {
        int key2 = row * MT + t2;
        bp.rad_off[key2] = off;
        bp.rad_p0[key2] = (int)bpf.size();
        for (int funct = 1; funct <= 3; ++funct)
          for (int i = 0; i < s.nsf; ++i)
            if (s.sf[i] == funct && s.sfenv[2 * i] == t2 + 1) {
              const double *q = &s.sfparam[(size_t)s.nsfparam * i];
              BPFunc f{funct, (float)q[0], (float)q[1], (float)q[2], 0.0f};
              bpf.push_back(f);
            }
        bp.rad_n[key2] = (int)bpf.size() - bp.rad_p0[key2];
        off += bp.rad_n[key2];
        for (int t3 = t2; t3 < ntypes; ++t3) {
          int key3 = (row * MT + t2) * MT + t3;
          int key3b = (row * MT + t3) * MT + t2;
          bp.ang_off[key3] = off;
          bp.ang_p0[key3] = (int)bpf.size();
          for (int funct = 4; funct <= 5; ++funct)
            for (int i = 0; i < s.nsf; ++i) {
              int a = s.sfenv[2 * i] - 1, b = s.sfenv[2 * i + 1] - 1;
              if (s.sf[i] == funct && ((a == t2 && b == t3) || (a == t3 && b == t2))) {
                const double *q = &s.sfparam[(size_t)s.nsfparam * i];
                BPFunc f{funct, (float)q[0], (float)q[1], (float)q[2], (float)q[3]};
                bpf.push_back(f);
              }
            }
          bp.ang_n[key3] = (int)bpf.size() - bp.ang_p0[key3];
          bp.ang_rcmax[key3] = 0.0f;
          bp.ang_g5[key3] = 0;
          bp.ang_one_seg[key3] = 1;
          for (int k = bp.ang_p0[key3]; k < (int)bpf.size(); ++k) {
            bp.ang_rcmax[key3] = std::max(bp.ang_rcmax[key3], bpf[k].Rc);
            if (bpf[k].kind == 5) bp.ang_g5[key3] = 1;
            const BPFunc &f0 = bpf[bp.ang_p0[key3]];
            if (bpf[k].Rc != f0.Rc || bpf[k].kind != f0.kind) bp.ang_one_seg[key3] = 0;
          }
          // the fast path's constants: lambda = +1 first, then -1, each in
          // file order
          bp.ang_fast[key3] = 0;
          bp.ang_np[key3] = 0;
          if (bp.ang_one_seg[key3] && bp.ang_n[key3] > 0) {
            const int p0 = bp.ang_p0[key3], n = bp.ang_n[key3];
            bool lam_pm1 = true;
            for (int k = p0; k < p0 + n; ++k) lam_pm1 = lam_pm1 && (bpf[k].p1 == 1.0f || bpf[k].p1 == -1.0f);
            if (lam_pm1) {
              bpk.resize(bpf.size());
              int q = p0;
              for (int sign = 1; sign >= -1; sign -= 2)
                for (int k = p0; k < p0 + n; ++k) {
                  const BPFunc &f = bpf[k];
                  if (f.p1 != (float)sign) continue;
                  BPAngK c{};
                  c.A = f.p2 - 1.0f;
                  c.B = (float)(-(double)f.p3 * 1.4426950408889634);
                  c.K1 = 0.5f * f.p2 * f.p1;
                  c.K2 = -2.0f * f.p3;
                  c.feat = k - p0;
                  bpk[q++] = c;
                  if (sign == 1) ++bp.ang_np[key3];
                }
              bp.ang_fast[key3] = 1;
            }
          }
          // the grid path: a fast block whose functions are every
          // combination of its (lambda, zeta) pairs and its etas, once each
          bp.ang_grid[key3] = 0;
          if (bp.ang_fast[key3]) {
            const int p0 = bp.ang_p0[key3], n = bp.ang_n[key3];
            std::vector<std::pair<float, float>> lz;
            std::vector<float> et;
            for (int k = p0; k < p0 + n; ++k) {
              const std::pair<float, float> v(bpf[k].p1, bpf[k].p2);
              if (std::find(lz.begin(), lz.end(), v) == lz.end()) lz.push_back(v);
              if (std::find(et.begin(), et.end(), bpf[k].p3) == et.end()) et.push_back(bpf[k].p3);
            }
            // lambda = +1 pairs first (the kernel takes the sign from the index)
            std::stable_sort(lz.begin(), lz.end(), [](const std::pair<float, float> &a, const std::pair<float, float> &b) {
              return a.first > b.first;
            });
            if ((int)lz.size() <= BP_GRID_NZ && (int)et.size() <= BP_GRID_NE && (int)(lz.size() * et.size()) == n) {
              BPAngGrid g{};
              g.nz = (int)lz.size();
              g.ne = (int)et.size();
              for (int q = 0; q < BP_GRID_NZ * BP_GRID_NE; ++q) g.feat[q] = -1;
              bool once = true;
              for (int k = p0; k < p0 + n; ++k) {
                const int iz = (int)(std::find(lz.begin(), lz.end(), std::make_pair(bpf[k].p1, bpf[k].p2)) - lz.begin());
                const int ie = (int)(std::find(et.begin(), et.end(), bpf[k].p3) - et.begin());
                int &slot = g.feat[iz * BP_GRID_NE + ie];
                if (slot >= 0) once = false;
                slot = k - p0;
              }
              g.nzp = 0;
              for (int iz = 0; iz < g.nz; ++iz) {
                if (lz[iz].first > 0.0f) ++g.nzp;
                g.zeta[iz] = lz[iz].second;
                g.K1[iz] = 0.5f * lz[iz].second * lz[iz].first;
              }
              for (int ie = 0; ie < g.ne; ++ie) {
                g.B[ie] = (float)(-(double)et[ie] * 1.4426950408889634);
                g.K2[ie] = -2.0f * et[ie];
              }
              if (once) {
                bpg.push_back(g);
                bp.ang_grid[key3] = (int)bpg.size();
              }
            }
          }
          bp.ang_grid[key3b] = bp.ang_grid[key3];
          bp.ang_fast[key3b] = bp.ang_fast[key3];
          bp.ang_np[key3b] = bp.ang_np[key3];
          bp.ang_off[key3b] = bp.ang_off[key3];
          bp.ang_p0[key3b] = bp.ang_p0[key3];
          bp.ang_n[key3b] = bp.ang_n[key3];
          bp.ang_rcmax[key3b] = bp.ang_rcmax[key3];
          bp.ang_g5[key3b] = bp.ang_g5[key3];
          bp.ang_one_seg[key3b] = bp.ang_one_seg[key3];
          off += bp.ang_n[key3];
        }
      }
      for (int i = 0; i < s.nsf; ++i)
        if (s.sf[i] < 1 || s.sf[i] > 5 || s.sfenv[2 * i] < 1 || s.sfenv[2 * i] > ntypes ||
            (s.sf[i] >= 4 && (s.sfenv[2 * i + 1] < 1 || s.sfenv[2 * i + 1] > ntypes))) {
          err = pots[row].file + ": unsupported symmetry function entry";
          return AENET_STATUS_ERR_IO;
        }
    }
    bpk.resize(bpf.size());
    for (int t = 0; t < ntypes; ++t) {
      if (species[t].bp_row < 0) {
        err = "central species missing from its own ENV list";
        return AENET_STATUS_ERR_TYPE;
      }
    }
  }
  loaded = true;
  return AENET_STATUS_OK;
}

}  // namespace aenet_amd

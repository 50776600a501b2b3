// mc.hip -- Metropolis Monte Carlo of many independent periodic chains on
// the GPU, plain and delayed-acceptance (two-stage) with a cheap radial-only
// potential as the first stage.
//
// One MC step sends at most one proposal per chain to the full potential:
//   k_mc_propose : (plain) per chain, a swap of two atoms of different
//                  species or a displacement of one atom by U(-d, d)^3,
//                  applied in place (the old values are kept for a
//                  revert); SplitMix64 generator per chain, state in device
//                  memory
//   k_mc_screen  : (delayed acceptance) per chain, one thread per atom:
//                  proposals drawn from the chain's current state are
//                  screened with the first-stage energy change, exp(-beta
//                  dE_s), until one passes (at most max_screen per step);
//                  one rejected here leaves the chain where it is and is a
//                  complete MC step of that chain, so the chain's sequence
//                  of proposals and decisions is that of sequential delayed
//                  acceptance, while the full potential gets a proposal from
//                  nearly every chain in every step.  The radial Chebyshev
//                  descriptor of every atom is kept, and only the pair terms
//                  with the moved sites change it, so the first-stage change
//                  needs no neighbor search; the moved atoms are recomputed
//                  over the whole cell; per-thread network
//   k_mc_active  : per chain that goes to the full potential, the atoms
//                  within its cutoff of a moved site, as one contiguous
//                  block of a batch-wide list
//   (full potential on the listed atoms: the batch's active-atom path)
//   k_mc_finish  : per chain, dE of the full potential over its block; the
//                  decision (plain: exp(-beta dE); second stage:
//                  exp(-beta (dE - dE_s))); commit or revert
// Statistics, the decision record and the energy samples are kept per
// resolved proposal, in each chain's own count of proposals since the start
// of the run (the chains of a delayed-acceptance run advance by different
// numbers of proposals per step; sampling at step boundaries would weight
// states by their chance of passing the first stage).
// The second stage makes the chain exact for the full potential whatever
// the first-stage potential is (Christen & Fox, J. Comput. Graph. Stat.
// 14 (2005) 795).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "aenet_amd_internal.h"
#include "atom_common.h"
#include "kernels.h"
#include "kernels_common.h"

namespace aenet_amd {

namespace {

// SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): Weyl sequence + mixer
__device__ __forceinline__ uint64_t splitmix_next(uint64_t &s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform01(uint64_t &s) {
  return (double)(splitmix_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

// minimum-image displacement b - a in a periodic cell (rows of L are the
// lattice vectors, Binv its inverse); valid for distances below half the
// smallest cell width, which the host checks against the cutoffs
__device__ __forceinline__ void min_image(const MCDev &m, const double *a, const double *b, double *d) {
  double v[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double f[3];
  for (int k = 0; k < 3; ++k) {
    f[k] = m.Binv[k] * v[0] + m.Binv[3 + k] * v[1] + m.Binv[6 + k] * v[2];
    f[k] -= rint(f[k]);
  }
  for (int x = 0; x < 3; ++x) d[x] = f[0] * m.L[x] + f[1] * m.L[3 + x] + f[2] * m.L[6 + x];
}

__device__ __forceinline__ double dist(const MCDev &m, const double *a, const double *b) {
  double d[3];
  min_image(m, a, b, d);
  return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// the same in FP32 after the FP64 difference (the first stage's pair terms
// are FP32: their distances need no more)
__device__ __forceinline__ float dist_f(const MCDev &m, const double *a, const double *b) {
  const float v[3] = {(float)(b[0] - a[0]), (float)(b[1] - a[1]), (float)(b[2] - a[2])};
  float f[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f[k] = (float)m.Binv[k] * v[0] + (float)m.Binv[3 + k] * v[1] + (float)m.Binv[6 + k] * v[2];
    f[k] -= rintf(f[k]);
  }
  float d2 = 0.0f;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const float d = f[0] * (float)m.L[x] + f[1] * (float)m.L[3 + x] + f[2] * (float)m.L[6 + x];
    d2 += d * d;
  }
  return __builtin_amdgcn_sqrtf(d2);
}

constexpr int SR = 16;  // radial terms of the first-stage descriptor
constexpr int SW = 16;  // first-stage hidden width

// first-stage descriptor of one atom: plain and typespin-weighted radial
// sums (features r_i1+n and r_i2+n of the new_SFBasis layout); the angular
// features of a radial-only basis are zero
struct SDesc {
  double r[SR], s[SR];
};

__device__ __forceinline__ void sdesc_zero(SDesc &G) {
#pragma unroll
  for (int n = 0; n < SR; ++n) G.r[n] = G.s[n] = 0.0;
}

// radial Chebyshev terms of one pair (sfb_eval radial part,
// src/ext/sfbasis.f90:489-522): vr_n += w_plain fc T_n, vs_n += w_spin fc T_n
// for n < rN, cutoff Rc
__device__ __forceinline__ void radial_terms(double Rc, int rN, double r, double w_plain, double w_spin, double *vr,
                                             double *vs) {
  if (!(r <= Rc && r > 1.0e-12)) return;
  const double fc = 0.5 * (cos(3.14159265358979 * r / Rc) + 1.0);
  const double x = 2.0 * r / Rc - 1.0;
  double Tm = 1.0, T = x;
#pragma unroll
  for (int n = 0; n < SR; ++n) {
    double tn;
    if (n == 0) {
      tn = 1.0;
    } else if (n == 1) {
      tn = x;
    } else {
      const double Tn = 2.0 * x * T - Tm;
      Tm = T;
      T = Tn;
      tn = Tn;
    }
    if (n < rN) {
      vr[n] += w_plain * fc * tn;
      vs[n] += w_spin * fc * tn;
    }
  }
}

// the same terms in FP32 arithmetic for the Monte Carlo first stage (a
// screen only: the second stage keeps the chain exact), summed into the
// FP64 descriptors vr, vs (plain and typespin)
template <int RN, class T>
__device__ __forceinline__ void radial_terms_f(float Rc, int rN, double r, float w_plain, float w_spin, T *vr, T *vs) {
  const float rf = (float)r;
  if (!(rf <= Rc && rf > 1.0e-12f)) return;
  const float q = rf * __builtin_amdgcn_rcpf(Rc);
  const float fc = 0.5f * (__cosf(3.14159265f * q) + 1.0f);
  const float x = 2.0f * q - 1.0f;
  float Tm = 1.0f, Tc = x;
#pragma unroll
  for (int n = 0; n < RN; ++n) {
    float tn;
    if (n == 0) {
      tn = 1.0f;
    } else if (n == 1) {
      tn = x;
    } else {
      const float Tn = 2.0f * x * Tc - Tm;
      Tm = Tc;
      Tc = Tn;
      tn = Tn;
    }
    if (n < rN) {
      const float t = fc * tn;
      vr[n] += (T)(w_plain * t);
      vs[n] += (T)(w_spin * t);
    }
  }
}

__device__ __forceinline__ void student_pair(const SpeciesDev &sp, double r, double w_plain, double w_spin,
                                             SDesc &G) {
  radial_terms(sp.r_Rc, sp.r_N, r, w_plain, w_spin, G.r, G.s);
}

__device__ __forceinline__ float act_f(int f, float z) {
  switch (f) {
    case 0: return z;
    case 1: return tanhf(z);
    case 2: return 1.0f / (1.0f + expf(-z));
    case 3: return 1.7159f * tanhf(0.666666666666667f * z);
    case 4: return 1.7159f * tanhf(0.666666666666667f * z) + 0.1f * z;
    default: return 0.0f;
  }
}

// the activations for the Monte Carlo first stage (a screen: the second
// stage keeps the chain exact) with the hardware exponential:
// tanh(z) = sign(z) (1 - e^(-2|z|)) / (1 + e^(-2|z|)), absolute error about
// 1e-7; the sigmoid likewise
__device__ __forceinline__ float tanh_fast(float z) {
  const float t = __expf(-2.0f * fabsf(z));
  return copysignf(__fdividef(1.0f - t, 1.0f + t), z);
}

__device__ __forceinline__ float act_fast(int f, float z) {
  switch (f) {
    case 0: return z;
    case 1: return tanh_fast(z);
    case 2: return __fdividef(1.0f, 1.0f + __expf(-z));
    case 3: return 1.7159f * tanh_fast(0.666666666666667f * z);
    case 4: return 1.7159f * tanh_fast(0.666666666666667f * z) + 0.1f * z;
    default: return 0.0f;
  }
}

// first-stage potential of one species in compact form, held in LDS by the
// Monte Carlo kernels: the radial basis, the normalization of its plain
// and typespin features, the one-hidden-layer network (weights k-major;
// the constant input of the zero angular features folded into the hidden
// bias) and the energy epilogue
struct FirstStage {
  double avg[2 * SR], scl[2 * SR];  // plain features n, typespin features SR + n
  double E_scale, E_shift, E_atom;
  float W[SW][2 * SR];              // hidden node k: weights of the features
  float b[SW], Wo[SW];
  float bo, Rc;
  float spin[AENET_AMD_MAX_TYPES];
  int rN, multi, nh, fa_h, fa_o, r_i1, r_i2, pad;
};
static_assert(sizeof(FirstStage) % 8 == 0, "FirstStage keeps the doubles after it aligned");
typedef __attribute__((address_space(4))) const FirstStage ConstFirstStage;

// the workgroup fills tab[0 .. ntypes-1] from the model (one hidden layer)
__device__ void load_first_stage(const ModelDev &md, FirstStage *tab) {
  for (int t = 0; t < md.ntypes; ++t) {
    const SpeciesDev &sp = md.species[t];
    FirstStage &S = tab[t];
    const int nh = sp.nnodes[1];
    const float *W0 = md.W + sp.w_off[0];
    const float *W1 = md.W + sp.w_off[1];
    const double *avg = md.norm_avg + sp.norm_off;
    const double *scl = md.norm_scale + sp.norm_off;
    for (int i = threadIdx.x; i < SW * 2 * SR; i += blockDim.x) {
      const int k = i / (2 * SR), f = i - k * 2 * SR;
      const int q = f >= SR, n = f - q * SR;
      float w = 0.0f;
      if (n < sp.r_N && k < nh && (q == 0 || sp.multi)) w = W0[((q ? sp.r_i2 : sp.r_i1) + n) * nh + k];
      S.W[k][f] = w;
    }
    for (int f = threadIdx.x; f < 2 * SR; f += blockDim.x) {
      const int q = f >= SR, n = f - q * SR;
      const bool on = n < sp.r_N && (q == 0 || sp.multi);
      const int i = (q ? sp.r_i2 : sp.r_i1) + n;
      S.avg[f] = on ? avg[i] : 0.0;
      S.scl[f] = on ? scl[i] : 0.0;
    }
    if (threadIdx.x < SW) {
      const int k = threadIdx.x;
      float b = 0.0f, wo = 0.0f;
      if (k < nh) {
        b = W0[sp.nnodes[0] * nh + k];
        for (int n = 0; n < sp.a_N; ++n) {
          b += W0[(sp.a_i1 + n) * nh + k] * (float)(scl[sp.a_i1 + n] * (0.0 - avg[sp.a_i1 + n]));
          if (sp.multi) b += W0[(sp.a_i2 + n) * nh + k] * (float)(scl[sp.a_i2 + n] * (0.0 - avg[sp.a_i2 + n]));
        }
        wo = W1[k];
      }
      S.b[k] = b;
      S.Wo[k] = wo;
    }
    if (threadIdx.x == 0) {
      S.E_scale = sp.E_scale;
      S.E_shift = sp.E_shift;
      S.E_atom = sp.E_atom;
      S.bo = W1[nh];
      S.Rc = sp.r_Rc;
      for (int u = 0; u < AENET_AMD_MAX_TYPES; ++u) S.spin[u] = sp.spin[u];
      S.rN = sp.r_N;
      S.multi = sp.multi;
      S.nh = nh;
      S.fa_h = sp.fa[1];
      S.fa_o = sp.fa[2];
      S.r_i1 = sp.r_i1;
      S.r_i2 = sp.r_i2;
      S.pad = 0;
    }
  }
}

// the workgroup copies the table into LDS
__device__ __forceinline__ void copy_first_stage(const FirstStage *src, int ntypes, FirstStage *dst) {
  const double *a = reinterpret_cast<const double *>(src);
  double *b = reinterpret_cast<double *>(dst);
  const int n = ntypes * (int)(sizeof(FirstStage) / 8);
  for (int k = threadIdx.x; k < n; k += blockDim.x) b[k] = a[k];
}

// the first stage's atomic energy from its unscaled descriptor:
// normalization (stp_normalize), ff_eval in per-thread FP32 registers,
// energy epilogue in FP64; RN radial terms and NH hidden nodes at most
// (the table's entries beyond the species' own sizes are zero)
template <int RN, int NH, class Table>
__device__ __forceinline__ double fs_energy(const Table &S, const SDesc &G) {
  float x[2 * RN];
#pragma unroll
  for (int n = 0; n < RN; ++n) {
    x[n] = (float)(S.scl[n] * (G.r[n] - S.avg[n]));
    x[RN + n] = (float)(S.scl[SR + n] * (G.s[n] - S.avg[SR + n]));
  }
  float y = S.bo;
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    float z = S.b[k];
#pragma unroll
    for (int n = 0; n < RN; ++n) z += S.W[k][n] * x[n];
#pragma unroll
    for (int n = 0; n < RN; ++n) z += S.W[k][SR + n] * x[RN + n];
    y += S.Wo[k] * act_fast(S.fa_h, z);
  }
  y = act_fast(S.fa_o, y);
  return S.E_scale * (double)y + S.E_shift + S.E_atom;
}

// the first-stage descriptors of a chain's atoms: a copy in LDS (stride
// 2*rn: plain terms, then typespin terms) or the global array (new_SFBasis
// layout, nsf per atom)
struct FSRead {
  const double *G;
  int stride, rn;
  bool lds;
};

template <int RN>
__device__ __forceinline__ void fs_load(const FSRead &R, const FirstStage &S, int j, SDesc &G) {
  const double *p = R.G + (size_t)j * R.stride;
#pragma unroll
  for (int n = 0; n < RN; ++n) {
    if (R.lds) {
      G.r[n] = n < R.rn ? p[n] : 0.0;
      G.s[n] = n < R.rn ? p[R.rn + n] : 0.0;
    } else {
      G.r[n] = n < S.rN ? p[S.r_i1 + n] : 0.0;
      G.s[n] = (n < S.rN && S.multi) ? p[S.r_i2 + n] : 0.0;
    }
  }
}

// stores the features in the new_SFBasis layout (angular ones zero)
__device__ __forceinline__ void fs_store(const FirstStage &S, int nsf, double *dst, const SDesc &G) {
  for (int f = 0; f < nsf; ++f) dst[f] = 0.0;
#pragma unroll
  for (int n = 0; n < SR; ++n) {
    if (n < S.rN) {
      dst[S.r_i1 + n] = G.r[n];
      if (S.multi) dst[S.r_i2 + n] = G.s[n];
    }
  }
}

// descriptor of atom j over every other atom of its chain (positions in LDS)
__device__ __forceinline__ void student_full(const MCDev &m, const SpeciesDev &sp, const double *sx, int j, size_t base, SDesc &G) {
  sdesc_zero(G);
  for (int k = 0; k < m.natoms; ++k) {
    if (k == j) continue;
    const double r = dist(m, &sx[3 * j], &sx[3 * k]);
    student_pair(sp, r, 1.0, sp.spin[m.types[base + k] - 1], G);
  }
}

// one proposal: kind, sites, displacement, and the uniforms of the two
// acceptance tests
struct Move {
  int kind, a, b;
  double d[3];
  double u1, u2;
};

// a proposal from the chain's current types pt: a swap of two atoms of
// different species with probability p_swap (kind MC_NONE when 64 draws
// find no atom of another species), else a displacement of one atom by
// U(-max_disp, max_disp)^3
__device__ __forceinline__ Move draw_move(const MCDev &m, const int *pt, uint64_t &s) {
  Move mv;
  const int N = m.natoms;
  mv.kind = MC_NONE;
  mv.a = 0;
  mv.b = 0;
  mv.d[0] = mv.d[1] = mv.d[2] = 0.0;
  if (uniform01(s) < m.p_swap) {
    mv.a = min((int)(uniform01(s) * N), N - 1);
    const int ta = pt[mv.a];
    for (int k = 0; k < 64; ++k) {
      const int cand = min((int)(uniform01(s) * N), N - 1);
      if (pt[cand] != ta) {
        mv.b = cand;
        mv.kind = MC_SWAP;
        break;
      }
    }
  } else {
    mv.kind = MC_DISP;
    mv.a = min((int)(uniform01(s) * N), N - 1);
    for (int x = 0; x < 3; ++x) mv.d[x] = (2.0 * uniform01(s) - 1.0) * m.max_disp;
  }
  mv.u1 = uniform01(s);
  mv.u2 = uniform01(s);
  return mv;
}

// A proposal evaluated from the chain's current state, the move not
// applied: positions sx and types pt are the chain's, xnew the displaced
// atom's new position.  type_after: atom j's species after the move
// (0-based)
__device__ __forceinline__ int type_after(const Move &mv, const int *pt, int j) {
  if (mv.kind == MC_SWAP) {
    if (j == mv.a) return pt[mv.b] - 1;
    if (j == mv.b) return pt[mv.a] - 1;
  }
  return pt[j] - 1;
}

// the first-stage descriptor of atom j (species S after the move) under the
// proposal mv: Gm the proposal's moved-site descriptors, R the chain's
// current descriptors; false (G untouched) when the move leaves atom j's
// descriptor unchanged
template <int RN>
__device__ __forceinline__ bool trial_desc_view(const MCDev &m, const FirstStage &S, int j, const Move &mv,
                                                const double *sx, const int *pt, const double *xnew,
                                                const double (*Gm)[2 * SR], const FSRead &R, SDesc &G) {
  if (j == mv.a || (mv.kind == MC_SWAP && j == mv.b)) {
    const int q = j == mv.a ? 0 : 1;
#pragma unroll
    for (int n = 0; n < RN; ++n) {
      G.r[n] = Gm[q][n];
      G.s[n] = Gm[q][SR + n];
    }
    return true;
  }
  const float Rc = S.Rc;
  if (mv.kind == MC_DISP) {
    // atom a leaves its position sx[a] for xnew
    const float ro = dist_f(m, &sx[3 * j], &sx[3 * mv.a]);
    const float rn = dist_f(m, &sx[3 * j], xnew);
    if (!(ro <= Rc || rn <= Rc)) return false;
    fs_load<RN>(R, S, j, G);
    const float spa = S.spin[pt[mv.a] - 1];
    radial_terms_f<RN>(S.Rc, S.rN, ro, -1.0f, -spa, G.r, G.s);
    radial_terms_f<RN>(S.Rc, S.rN, rn, 1.0f, spa, G.r, G.s);
    return true;
  }
  // swap: a and b exchange species; only the typespin weights of the
  // pairs (j, a) and (j, b) change
  const float ra = dist_f(m, &sx[3 * j], &sx[3 * mv.a]);
  const float rb = dist_f(m, &sx[3 * j], &sx[3 * mv.b]);
  if (!(ra <= Rc || rb <= Rc)) return false;
  fs_load<RN>(R, S, j, G);
  const float ds = S.spin[pt[mv.b] - 1] - S.spin[pt[mv.a] - 1];  // new minus old weight of the pair (j, a)
  radial_terms_f<RN>(S.Rc, S.rN, ra, 0.0f, ds, G.r, G.s);
  radial_terms_f<RN>(S.Rc, S.rN, rb, 0.0f, -ds, G.r, G.s);
  return true;
}

// the first-stage descriptor of moved site q (0: a, 1: b) of proposal mv
// over the whole cell, by one wavefront, the move not applied: per-lane
// FP32 pair terms over the lane's atoms, a reduce-scatter over the
// wavefront (first level FP32, then FP64; lane l ends with feature l/2),
// written to G[2*SR] (plain terms, then typespin)
template <int RN>
__device__ __forceinline__ void moved_site_wave(const MCDev &m, const FirstStage *tab, const Move &mv, int q,
                                                const double *sx, const int *pt, const double *xnew, double *G) {
  const int N = m.natoms;
  const int lane = lane_id();
  const int st = q == 0 ? mv.a : mv.b;
  const FirstStage &S = tab[type_after(mv, pt, st)];
  const double *xs = (mv.kind == MC_DISP && q == 0) ? xnew : &sx[3 * st];
  float v[2 * SR];
#pragma unroll
  for (int f = 0; f < 2 * SR; ++f) v[f] = 0.0f;
  for (int j = lane; j < N; j += WAVE) {
    if (j == st) continue;
    const float r = dist_f(m, xs, &sx[3 * j]);
    radial_terms_f<RN>(S.Rc, S.rN, r, 1.0f, S.spin[type_after(mv, pt, j)], v, v + SR);
  }
  const double t = reduce_scatter<2 * SR>(v);
  if ((lane & 1) == 0) G[lane >> 1] = t;
}

// per-chain bookkeeping held in registers by one thread while a kernel
// resolves proposals: proposals since the start of the run, statistics
// (proposed, accepted, full-potential evaluations, swaps) and the chain
// energy after the latest decision
struct ChainLog {
  long long nprop, st[4];
  double E;
};

__device__ __forceinline__ ChainLog log_open(const MCDev &m, int c) {
  ChainLog L;
  L.nprop = m.nprop[c];
  for (int k = 0; k < 4; ++k) L.st[k] = 0;
  L.E = m.E_chain[c];
  return L;
}

__device__ __forceinline__ void log_close(const MCDev &m, int c, const ChainLog &L) {
  m.nprop[c] = L.nprop;
  for (int k = 0; k < 4; ++k)
    if (L.st[k]) m.stats[4 * (size_t)c + k] += L.st[k];
}

// one resolved proposal of chain c: statistics, the decision record and
// the energy samples, in the chain's own count of proposals since the
// start of the run; L.E is the energy after the decision
__device__ __forceinline__ void log_proposal(const MCDev &m, int c, ChainLog &L, int kind, double dE_first,
                                             double dE_full, int stage1, bool accept) {
  L.st[0] += 1;
  if (accept) L.st[1] += 1;
  if (stage1 == 1) L.st[2] += 1;
  if (kind == MC_SWAP) L.st[3] += 1;
  const long long k = L.nprop;
  if (m.trace && k < m.max_trace) {
    double *tr = m.trace + 4 * ((size_t)k * m.nchains + c);
    tr[0] = dE_first;
    tr[1] = dE_full;
    tr[2] = (double)stage1;
    tr[3] = accept ? 1.0 : 0.0;
  }
  L.nprop = k + 1;
  if (m.samples && m.sample_every > 0 && (k + 1) % m.sample_every == 0) {
    const long long q = (k + 1) / m.sample_every - 1;
    if (q < m.max_samples) m.samples[(size_t)q * m.nchains + c] = L.E;
  }
}

}  // namespace

__global__ void k_mc_propose(MCDev m) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) {
    *m.nactive = 0;
    *m.step += 1;
  }
  if (c >= m.nchains) return;
  uint64_t s = m.rng[c];
  const size_t base = (size_t)c * m.natoms;
  const Move mv = draw_move(m, m.types + base, s);
  if (mv.kind == MC_SWAP) {
    const int ta = m.types[base + mv.a];
    m.types[base + mv.a] = m.types[base + mv.b];
    m.types[base + mv.b] = ta;
  } else if (mv.kind == MC_DISP) {
    for (int x = 0; x < 3; ++x) {
      const double old = m.coords[3 * (base + mv.a) + x];
      m.old_pos[3 * c + x] = old;
      m.coords[3 * (base + mv.a) + x] = old + mv.d[x];
    }
  }
  m.kind[c] = mv.kind;
  m.site_a[c] = mv.a;
  m.site_b[c] = mv.b;
  m.u_stage1[c] = mv.u1;
  m.u_stage2[c] = mv.u2;
  m.need[c] = mv.kind != MC_NONE ? 1 : 0;
  m.rng[c] = s;
}

// proposals a screen round evaluates together (from the same chain state)
constexpr int SCREEN_ROUND = 4;

// LDS of k_mc_screen in doubles: positions [3N], trial energies of the
// round's proposals [SCREEN_ROUND][N], the first-stage tables, with the
// cache the current energies [N] and descriptors [N][2*rn], then types [N]
// and the proposals that change each atom's descriptor [N] (ints, bit p)
__host__ __device__ inline size_t screen_lds_bytes(int N, int ntypes, int rn, bool cache) {
  size_t d = 3 * (size_t)N + SCREEN_ROUND * (size_t)N;
  if (cache) d += (size_t)N * (1 + 2 * rn);
  return 8 * d + ntypes * sizeof(FirstStage) + 2 * 4 * (size_t)N;
}

// the first-stage table of every species, built once per run (one
// workgroup); the kernels read it with uniform addresses
__global__ __launch_bounds__(256) void k_first_stage_table(ModelDev md, FirstStage *tab) { load_first_stage(md, tab); }

// One workgroup per chain, threads over the chain's atoms; CACHE: the
// chain's first-stage state is copied into LDS once per step; RN, NH: the
// first stage's radial terms and hidden nodes at most.  A rejection by the
// first stage leaves the chain where it is, so every proposal of a step
// starts from the same state: the proposals are drawn SCREEN_ROUND at a
// time, evaluated together (each thread over its atoms for every proposal),
// then decided in the order drawn; the first that passes goes on to the
// full potential and the ones after it are dropped (their random numbers
// unused), which is sequential delayed acceptance.  The network runs once
// per species present among a wavefront's affected atoms, with the
// weights of tabg (uniform loads); the LDS copy of the table serves the
// per-lane lookups
template <bool CACHE, int RN, int NH>
__global__ __launch_bounds__(256) void k_mc_screen(MCDev m, ModelDev md, const FirstStage *__restrict__ tabg) {
  extern __shared__ double lds[];
  __shared__ double Gm[SCREEN_ROUND][2][2 * SR];  // descriptors of each proposal's moved sites
  __shared__ double red[4][SCREEN_ROUND];
  __shared__ Move smv[SCREEN_ROUND];
  __shared__ double xnew[SCREEN_ROUND][3];
  __shared__ int s_live[SCREEN_ROUND];  // proposal p is part of this step (k < max_screen)
  __shared__ int s_pick;
  const int c = blockIdx.x;
  const int N = m.natoms;
  const int rn = m.rn_s;
  const size_t base = (size_t)c * N;
  const int nwaves = (blockDim.x + WAVE - 1) / WAVE;
  const int wave = threadIdx.x / WAVE;
  double *sx = lds;
  double *sE = sx + 3 * N;  // [SCREEN_ROUND][N]
  FirstStage *tab = reinterpret_cast<FirstStage *>(sE + SCREEN_ROUND * N);
  double *Ec = reinterpret_cast<double *>(tab + md.ntypes);
  double *Gc = Ec + (CACHE ? N : 0);
  int *pt = reinterpret_cast<int *>(Gc + (CACHE ? (size_t)N * 2 * rn : 0));
  int *saff = pt + N;
  if (c == 0 && threadIdx.x == 0) {
    *m.nactive = 0;
    *m.step += 1;
  }
  copy_first_stage(tabg, md.ntypes, tab);
  for (int k = threadIdx.x; k < 3 * N; k += blockDim.x) sx[k] = m.coords[3 * base + k];
  for (int k = threadIdx.x; k < N; k += blockDim.x) pt[k] = m.types[base + k];
  __syncthreads();
  if (CACHE) {
    // k = j * 2rn + f stepped by the workgroup size: (j, f) advanced by
    // (bs / 2rn, bs mod 2rn) with a carry instead of a division per element
    const int w2 = 2 * rn, bs = blockDim.x;
    const int dj = bs / w2, df = bs - dj * w2;
    int j = threadIdx.x / w2, f = threadIdx.x - j * w2;
    for (int k = threadIdx.x; k < N * w2; k += bs) {
      const FirstStage &S = tab[pt[j] - 1];
      const int q = f >= rn, n = f - q * rn;
      Gc[k] = (n < S.rN && (q == 0 || S.multi)) ? m.G_s[(base + j) * m.nsf_s + (q ? S.r_i2 : S.r_i1) + n] : 0.0;
      j += dj;
      f += df;
      if (f >= w2) {
        f -= w2;
        ++j;
      }
    }
    for (int k = threadIdx.x; k < N; k += blockDim.x) Ec[k] = m.E_s[base + k];
  }
  const FSRead R = CACHE ? FSRead{Gc, 2 * rn, rn, true} : FSRead{m.G_s + base * m.nsf_s, m.nsf_s, rn, false};
  const double *Ecur = CACHE ? Ec : m.E_s + base;
  uint64_t s = 0;
  ChainLog L;  // thread 0
  if (threadIdx.x == 0) {
    s = m.rng[c];
    L = log_open(m, c);
  }
  double dE_pass = 0.0;  // thread 0: first-stage change of the proposal that passed
  int pick = -1;
  for (int k0 = 0; k0 < m.max_screen && pick < 0; k0 += SCREEN_ROUND) {
    __syncthreads();  // the previous round's shared values are no longer read
    if (threadIdx.x == 0) {
      for (int p = 0; p < SCREEN_ROUND; ++p) {
        const bool live = k0 + p < m.max_screen;
        Move mv;
        if (live) {
          mv = draw_move(m, pt, s);
        } else {
          mv.kind = MC_NONE;
          mv.a = mv.b = 0;
        }
        if (mv.kind == MC_DISP)
          for (int x = 0; x < 3; ++x) xnew[p][x] = sx[3 * mv.a + x] + mv.d[x];
        smv[p] = mv;
        s_live[p] = live ? 1 : 0;
      }
    }
    __syncthreads();
    // the moved sites' descriptors of every proposal of the round, a
    // wavefront each in turn
    {
      int i = 0;
      for (int p = 0; p < SCREEN_ROUND; ++p) {
        const Move mv = smv[p];
        if (mv.kind == MC_NONE) continue;  // not a proposal (uniform)
        const int nq = mv.kind == MC_SWAP ? 2 : 1;
        for (int q = 0; q < nq; ++q, ++i)
          if (i % nwaves == wave) moved_site_wave<RN>(m, tab, mv, q, sx, pt, xnew[p], Gm[p][q]);
      }
    }
    __syncthreads();
    // every proposal's change over the affected atoms; uniform loops, so
    // that the species index and the table address are wave-uniform
    // (scalar loads of the weights)
    for (int j0 = 0; j0 < N; j0 += blockDim.x) {
      const int j = j0 + threadIdx.x;
      const bool on = j < N;
      int bits = 0;
      for (int p = 0; p < SCREEN_ROUND; ++p) {
        const Move mv = smv[p];
        double dE = 0.0;
        if (mv.kind != MC_NONE) {
          const int tj = on ? type_after(mv, pt, j) : 0;
          SDesc G;
          const bool aff = on && trial_desc_view<RN>(m, tab[tj], j, mv, sx, pt, xnew[p], Gm[p], R, G);
          double E = 0.0;
          for (int t = 0; t < md.ntypes; ++t) {
            const bool mine = aff && tj == t;
            if (wave_ballot(mine)) {
              // the species' table through the constant address space: scalar loads
              const ConstFirstStage *St = (const ConstFirstStage *)(tabg + uniform(t));
              if (mine) E = fs_energy<RN, NH>(*St, G);
            }
          }
          if (aff) {
            sE[p * N + j] = E;
            dE = E - Ecur[j];
            bits |= 1 << p;
          }
        }
        dE = wave_sum(dE);
        if (lane_id() == 0) red[wave][p] = j0 == 0 ? dE : red[wave][p] + dE;
      }
      if (on) saff[j] = bits;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      // decisions in the order drawn: the first that passes goes on
      for (int p = 0; p < SCREEN_ROUND && pick < 0; ++p) {
        const Move mv = smv[p];
        if (!s_live[p] || mv.kind == MC_NONE) continue;  // no atom of another species: not a proposal
        double tot = 0.0;
        for (int w = 0; w < nwaves; ++w) tot += red[w][p];
        if (log(mv.u1) < -m.beta * tot) {
          pick = p;
          dE_pass = tot;
        } else {
          // rejected by the first stage: the chain stays where it is
          log_proposal(m, c, L, mv.kind, tot, 0.0, 0, false);
        }
      }
      s_pick = pick;
    }
    __syncthreads();
    pick = s_pick;
  }
  if (pick >= 0) {
    // the proposal that passed goes to the full potential, applied in place
    const Move mv = smv[pick];
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
      const size_t g = base + j;
      const bool aff = (saff[j] >> pick) & 1;
      if (aff) {
        const FirstStage &S = tab[type_after(mv, pt, j)];
        SDesc G;
        trial_desc_view<RN>(m, S, j, mv, sx, pt, xnew[pick], Gm[pick], R, G);
        fs_store(S, m.nsf_s, m.G_trial + g * m.nsf_s, G);
        m.E_s_trial[g] = sE[pick * N + j];
      }
      m.affected[g] = aff ? 1 : 0;
    }
    if (threadIdx.x == 0) {
      if (mv.kind == MC_SWAP) {
        m.types[base + mv.a] = pt[mv.b];
        m.types[base + mv.b] = pt[mv.a];
      } else {
        for (int x = 0; x < 3; ++x) {
          m.old_pos[3 * c + x] = sx[3 * mv.a + x];
          m.coords[3 * (base + mv.a) + x] = xnew[pick][x];
        }
      }
      m.kind[c] = mv.kind;
      m.site_a[c] = mv.a;
      m.site_b[c] = mv.b;
      m.u_stage1[c] = mv.u1;
      m.u_stage2[c] = mv.u2;
      m.dE_s[c] = dE_pass;
      m.need[c] = 1;
    }
  } else if (threadIdx.x == 0) {
    m.kind[c] = MC_NONE;
    m.need[c] = 0;
  }
  if (threadIdx.x == 0) {
    m.rng[c] = s;
    log_close(m, c, L);
  }
}

// one workgroup per chain: the atoms the full potential must re-evaluate
__global__ __launch_bounds__(1024) void k_mc_active(MCDev m) {
  __shared__ int wcount[16];
  __shared__ int base_out;
  const int c = blockIdx.x;
  const int j = threadIdx.x;
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  const int kind = m.kind[c];
  const bool need = m.need[c] != 0;
  bool in = false;
  if (need && j < N) {
    const int a = m.site_a[c], b = m.site_b[c];
    const double *xj = &m.coords[3 * (base + j)];
    const double rc = m.rc_t;
    if (j == a || (kind == MC_SWAP && j == b)) {
      in = true;
    } else if (kind == MC_DISP) {
      in = dist(m, xj, &m.old_pos[3 * c]) <= rc || dist(m, xj, &m.coords[3 * (base + a)]) <= rc;
    } else {
      in = dist(m, xj, &m.coords[3 * (base + a)]) <= rc || dist(m, xj, &m.coords[3 * (base + b)]) <= rc;
    }
  }
  const unsigned long long msk = wave_ballot(in);
  const int w = j / WAVE;
  if (lane_id() == 0) wcount[w] = popc64(msk);
  __syncthreads();
  if (j == 0) {
    int tot = 0;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    for (int k = 0; k < nw; ++k) {
      const int t = wcount[k];
      wcount[k] = tot;
      tot += t;
    }
    base_out = tot ? atomicAdd(m.nactive, tot) : 0;
    m.act_start[c] = base_out;
    m.act_count[c] = tot;
  }
  __syncthreads();
  if (in) m.active[base_out + wcount[w] + mask_prefix(msk)] = (int)(base + j);
}

// one wavefront per chain
// one workgroup of up to 4 wavefronts per chain: dE of the full potential
// over its block, the decision, then commit or revert; the affected atoms'
// first-stage state is copied by the wavefronts in parallel, a chunk of 64
// atoms each
__global__ __launch_bounds__(256) void k_mc_finish(MCDev m, int plain) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  const int kind = m.kind[c];
  const bool need = m.need[c] != 0;
  const int s0 = m.act_start[c], n = need ? m.act_count[c] : 0;
  double dE = 0.0;
  for (int k = tid; k < n; k += blockDim.x) {
    const int g = m.active[s0 + k];
    dE += m.E_trial[g] - m.E_t[g];
  }
  dE = wave_sum(dE);
  if (lane == 0) red[wave] = dE;
  __syncthreads();
  dE = 0.0;
  for (int w = 0; w < nw; ++w) dE += red[w];  // fixed order
  bool accept = false;
  if (need) {
    const double arg = plain ? -m.beta * dE : -m.beta * (dE - m.dE_s[c]);
    accept = log(m.u_stage2[c]) < arg;
  }
  if (accept) {
    for (int k = tid; k < n; k += blockDim.x) {
      const int g = m.active[s0 + k];
      m.E_t[g] = m.E_trial[g];
    }
    if (!plain) {
      // the affected atoms' first-stage state: their rows copied by groups
      // of nsf consecutive lanes (rows = 64 / nsf atoms a pass), the loads
      // of four passes issued before their stores
      const int nsf = m.nsf_s;
      const double *__restrict__ src = m.G_trial;
      double *__restrict__ dst = m.G_s;
      const int rows = nsf <= WAVE ? WAVE / nsf : 1;
      const int r = nsf <= WAVE ? lane / nsf : 0;
      for (int j0 = wave * WAVE; j0 < N; j0 += nw * WAVE) {
        const int j = j0 + lane;
        const bool aff = j < N && m.affected[base + j] != 0;
        if (aff) m.E_s[base + j] = m.E_s_trial[base + j];
        unsigned long long msk = wave_ballot(aff);
        while (msk) {
          size_t at[4];
          double val[4];
          bool ok[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            // this lane's row of the pass: the r-th set bit left in msk
            unsigned long long mm = msk;
            for (int k = 0; k < r && mm; ++k) mm &= mm - 1;
            ok[u] = mm != 0 && r < rows;
            at[u] = ok[u] ? (base + j0 + __builtin_ctzll(mm)) * nsf : 0;
            for (int k = 0; k < rows && msk; ++k) msk &= msk - 1;
          }
          if (nsf <= WAVE) {
            const int f = lane - r * nsf;
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (ok[u]) val[u] = src[at[u] + f];
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (ok[u]) dst[at[u] + f] = val[u];
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (ok[u])
                for (int f = lane; f < nsf; f += WAVE) dst[at[u] + f] = src[at[u] + f];
          }
        }
      }
    }
  } else if (tid == 0 && need) {
    const int a = m.site_a[c], b = m.site_b[c];
    if (kind == MC_SWAP) {
      const int ta = m.types[base + a];
      m.types[base + a] = m.types[base + b];
      m.types[base + b] = ta;
    } else {
      for (int x = 0; x < 3; ++x) m.coords[3 * (base + a) + x] = m.old_pos[3 * c + x];
    }
  }
  if (tid == 0) {
    if (accept) m.E_chain[c] += dE;
    if (need) {
      ChainLog L = log_open(m, c);
      log_proposal(m, c, L, kind, plain ? 0.0 : m.dE_s[c], dE, 1, accept);
      log_close(m, c, L);
    }
  }
}

// decision record: every entry (dE_first, dE_full, stage-1 outcome,
// accepted) = (0, 0, -1, 0) until a proposal fills it
__global__ void k_mc_trace_init(double *tr, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) tr[i] = (i & 3) == 2 ? -1.0 : 0.0;
}

// first-stage state from scratch: every atom's descriptor over the whole
// cell and its energy, with the arithmetic of k_mc_screen
__global__ __launch_bounds__(256) void k_mc_student_init(MCDev m, ModelDev md, const FirstStage *tabg) {
  extern __shared__ double lds[];
  const int c = blockIdx.x;
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  double *sx = lds;
  FirstStage *tab = reinterpret_cast<FirstStage *>(sx + 3 * N);
  copy_first_stage(tabg, md.ntypes, tab);
  for (int k = threadIdx.x; k < 3 * N; k += blockDim.x) sx[k] = m.coords[3 * base + k];
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    const size_t g = base + j;
    const FirstStage &S = tab[m.types[g] - 1];
    SDesc G;
    sdesc_zero(G);
    for (int k = 0; k < N; ++k) {
      if (k == j) continue;
      radial_terms_f<SR>(S.Rc, S.rN, dist(m, &sx[3 * j], &sx[3 * k]), 1.0f, S.spin[m.types[base + k] - 1], G.r, G.s);
    }
    fs_store(S, m.nsf_s, m.G_s + g * m.nsf_s, G);
    m.E_s[g] = fs_energy<SR, SW>(S, G);
    m.affected[g] = 0;
  }
}

// ---------------------------------------------------------------------
// Fitting the first-stage potential to energies and forces of a set of
// cells (all with one lattice, natoms <= 256, one hidden layer, linear
// output).  One workgroup per cell, one thread per atom:
//   A: descriptor, forward pass, c_j = dE_j/dG_j
//   B: forces F_i = sum_j u_ij sum_f phi'_f(r_ij) (c_j[f] w_f(i) + c_i[f]
//      w_f(j)), w_f = 1 or the typespin; residual r_i = F_i - F_ref_i
//   C: v_j = sum_i (dG_j/dR_i) . r_i (the residual pulled back into
//      descriptor space)
//   D: the gradient of coefE/2 (E_j - E_ref_j)^2 and of the force loss
//      coefF/2 sum_i |r_i|^2, whose weight gradient is -coefF d/dW (grad_G
//      E_j . v_j) (forward tangent of the network along v_j, then reverse)
// ---------------------------------------------------------------------

namespace {

__device__ __forceinline__ void student_pair_d(const SpeciesDev &sp, double r, double (&dphi)[SR]) {
#pragma unroll
  for (int n = 0; n < SR; ++n) dphi[n] = 0.0;
  const double Rc = sp.r_Rc;
  if (!(r < Rc && r > 1.0e-12)) return;
  const double a = 3.14159265358979 / Rc;
  const double fc = 0.5 * (cos(a * r) + 1.0);
  const double dfc = -0.5 * a * sin(a * r);
  const double x = 2.0 * r / Rc - 1.0, dx = 2.0 / Rc;
  double Tm = 1.0, T = x, Um = 0.0, U = 1.0;  // T_n and U_{n-1}
#pragma unroll
  for (int n = 0; n < SR; ++n) {
    double tn, dtn;
    if (n == 0) {
      tn = 1.0;
      dtn = 0.0;
    } else if (n == 1) {
      tn = x;
      dtn = 1.0;
    } else {
      const double Tn = 2.0 * x * T - Tm;
      const double Un = 2.0 * x * U - Um;  // U_{n-1}
      Tm = T;
      T = Tn;
      Um = U;
      U = Un;
      tn = Tn;
      dtn = n * Un;
    }
    if (n < sp.r_N) dphi[n] = dfc * tn + fc * dtn * dx;
  }
}

__device__ __forceinline__ void act_d(int f, float z, float h, float &d1, float &d2) {
  switch (f) {
    case 0: d1 = 1.0f; d2 = 0.0f; break;
    case 1: d1 = 1.0f - h * h; d2 = -2.0f * h * d1; break;
    case 2: d1 = h * (1.0f - h); d2 = d1 * (1.0f - 2.0f * h); break;
    case 3:
    case 4: {
      const float t = tanhf(0.666666666666667f * z);
      const float s = 1.0f - t * t;
      d1 = 1.7159f * 0.666666666666667f * s + (f == 4 ? 0.1f : 0.0f);
      d2 = 1.7159f * 0.444444444444444f * (-2.0f * t * s);
    } break;
    default: d1 = 0.0f; d2 = 0.0f;
  }
}

}  // namespace

__global__ __launch_bounds__(256) void k_first_stage_fit(MCDev m, ModelDev md, const int *cells, const double *E_ref,
                                                         const double *F_ref, double coefE, double coefF, float *grad,
                                                         int wsize, double *loss) {
  extern __shared__ double sh[];
  double *sx = sh;                                          // [3N]
  float *cr = reinterpret_cast<float *>(sx + 3 * m.natoms);  // [N*SR] dE/dG plain
  float *cs = cr + m.natoms * SR;                            // [N*SR] dE/dG typespin
  float *res = cs + m.natoms * SR;                           // [3N] force residual
  float *gacc = res + 3 * m.natoms;                          // [wsize]
  const int N = m.natoms;
  const int cell = cells[blockIdx.x];
  const size_t base = (size_t)cell * N;
  const int j = threadIdx.x;
  for (int k = j; k < 3 * N; k += blockDim.x) sx[k] = m.coords[3 * base + k];
  for (int k = j; k < wsize; k += blockDim.x) gacc[k] = 0.0f;
  __syncthreads();
  const bool on = j < N;
  const size_t g = base + (on ? j : 0);
  const int tj = on ? m.types[g] : 1;
  const SpeciesDev &sp = md.species[tj - 1];
  const double *avg = md.norm_avg + sp.norm_off;
  const double *scl = md.norm_scale + sp.norm_off;
  const int nin = sp.nnodes[0], nh = sp.nnodes[1];
  const float *W1 = md.W + sp.w_off[0];
  const float *W2 = md.W + sp.w_off[1];
  const float Es = (float)sp.E_scale;
  float xr[SR], xs[SR], z[SW], h[SW], d1[SW], d2[SW];
  double lossE = 0.0, lossF = 0.0;
  double Ej = 0.0;
  // ---- A
  {
    SDesc G;
    if (on) {
      student_full(m, sp, sx, j, base, G);
    } else {
      sdesc_zero(G);
    }
#pragma unroll
    for (int n = 0; n < SR; ++n) {
      xr[n] = n < sp.r_N ? (float)(scl[sp.r_i1 + n] * (G.r[n] - avg[sp.r_i1 + n])) : 0.0f;
      xs[n] = (n < sp.r_N && sp.multi) ? (float)(scl[sp.r_i2 + n] * (G.s[n] - avg[sp.r_i2 + n])) : 0.0f;
    }
    float y = W2[nh];
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      if (k < nh) {
        float zz = W1[nin * nh + k];
#pragma unroll
        for (int n = 0; n < SR; ++n) {
          if (n < sp.r_N) {
            zz += W1[(sp.r_i1 + n) * nh + k] * xr[n];
            if (sp.multi) zz += W1[(sp.r_i2 + n) * nh + k] * xs[n];
          }
        }
        for (int n = 0; n < sp.a_N; ++n) {
          zz += W1[(sp.a_i1 + n) * nh + k] * (float)(scl[sp.a_i1 + n] * (0.0 - avg[sp.a_i1 + n]));
          if (sp.multi) zz += W1[(sp.a_i2 + n) * nh + k] * (float)(scl[sp.a_i2 + n] * (0.0 - avg[sp.a_i2 + n]));
        }
        z[k] = zz;
        h[k] = act_f(sp.fa[1], zz);
        act_d(sp.fa[1], zz, h[k], d1[k], d2[k]);
        y += W2[k] * h[k];
      } else {
        z[k] = h[k] = d1[k] = d2[k] = 0.0f;
      }
    }
    Ej = sp.E_scale * (double)y + sp.E_shift + sp.E_atom;
    // c_j = dE_j/dG_j (unnormalized G)
#pragma unroll
    for (int n = 0; n < SR; ++n) {
      float ar = 0.0f, as = 0.0f;
#pragma unroll
      for (int k = 0; k < SW; ++k) {
        if (k < nh && n < sp.r_N) {
          const float q = W2[k] * d1[k];
          ar += q * W1[(sp.r_i1 + n) * nh + k];
          if (sp.multi) as += q * W1[(sp.r_i2 + n) * nh + k];
        }
      }
      if (on) {
        cr[j * SR + n] = n < sp.r_N ? Es * ar * (float)scl[sp.r_i1 + n] : 0.0f;
        cs[j * SR + n] = (n < sp.r_N && sp.multi) ? Es * as * (float)scl[sp.r_i2 + n] : 0.0f;
      }
    }
  }
  __syncthreads();
  // ---- B
  if (on) {
    double F[3] = {0.0, 0.0, 0.0};
    const double spin_j = sp.spin[tj - 1];
    for (int i = 0; i < N; ++i) {
      if (i == j) continue;
      double d[3];
      min_image(m, &sx[3 * j], &sx[3 * i], d);  // R_i - R_j
      const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (!(r < sp.r_Rc)) continue;
      double dphi[SR];
      student_pair_d(sp, r, dphi);
      const int ti = m.types[base + i];
      const double spin_i = sp.spin[ti - 1];
      double sgm = 0.0;
#pragma unroll
      for (int n = 0; n < SR; ++n)
        if (n < sp.r_N)
          sgm += dphi[n] * ((double)cr[i * SR + n] + (double)cs[i * SR + n] * spin_j + (double)cr[j * SR + n] +
                            (double)cs[j * SR + n] * spin_i);
      for (int x = 0; x < 3; ++x) F[x] += sgm * d[x] / r;
    }
    for (int x = 0; x < 3; ++x) {
      const double rr = F[x] - F_ref[3 * g + x];
      res[3 * j + x] = (float)rr;
      lossF += rr * rr;
    }
    const double re = Ej - E_ref[g];
    lossE = re * re;
  }
  __syncthreads();
  // ---- C, D
  if (on) {
    float vr[SR], vs[SR];
#pragma unroll
    for (int n = 0; n < SR; ++n) vr[n] = vs[n] = 0.0f;
    for (int i = 0; i < N; ++i) {
      if (i == j) continue;
      double d[3];
      min_image(m, &sx[3 * j], &sx[3 * i], d);  // R_i - R_j
      const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (!(r < sp.r_Rc)) continue;
      double dphi[SR];
      student_pair_d(sp, r, dphi);
      const float proj = (float)((d[0] * (res[3 * i] - res[3 * j]) + d[1] * (res[3 * i + 1] - res[3 * j + 1]) +
                                  d[2] * (res[3 * i + 2] - res[3 * j + 2])) / r);
      const float spin_i = sp.spin[m.types[base + i] - 1];
#pragma unroll
      for (int n = 0; n < SR; ++n) {
        vr[n] += (float)dphi[n] * proj;
        vs[n] += (float)dphi[n] * spin_i * proj;
      }
    }
    // tangent of the normalized input and of the hidden pre-activations
    float tr[SR], ts[SR], zt[SW];
#pragma unroll
    for (int n = 0; n < SR; ++n) {
      tr[n] = n < sp.r_N ? (float)scl[sp.r_i1 + n] * vr[n] : 0.0f;
      ts[n] = (n < sp.r_N && sp.multi) ? (float)scl[sp.r_i2 + n] * vs[n] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      float a = 0.0f;
      if (k < nh) {
#pragma unroll
        for (int n = 0; n < SR; ++n) {
          if (n < sp.r_N) {
            a += W1[(sp.r_i1 + n) * nh + k] * tr[n];
            if (sp.multi) a += W1[(sp.r_i2 + n) * nh + k] * ts[n];
          }
        }
      }
      zt[k] = a;
    }
    const float gE = (float)(coefE * (Ej - E_ref[g])) * Es;
    const float gF = -(float)coefF * Es;
    const int o1 = sp.w_off[0], o2 = sp.w_off[1];
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      if (k < nh) {
        const float w2 = W2[k];
        // output layer: weight from hidden k, then bias
        atomicAdd(&gacc[o2 + k], gE * h[k] + gF * d1[k] * zt[k]);
        const float bk = gE * w2 * d1[k] + gF * w2 * d2[k] * zt[k];
        atomicAdd(&gacc[o1 + nin * nh + k], bk);
        // input features: x'_c and its tangent
        for (int c = 0; c < nin; ++c) {
          float xc, tc;
          int n = c - sp.r_i1;
          if (n >= 0 && n < sp.r_N) {
            xc = 0.0f;
            tc = 0.0f;
#pragma unroll
            for (int q = 0; q < SR; ++q)
              if (q == n) {
                xc = xr[q];
                tc = tr[q];
              }
          } else if (sp.multi && (n = c - sp.r_i2) >= 0 && n < sp.r_N) {
            xc = 0.0f;
            tc = 0.0f;
#pragma unroll
            for (int q = 0; q < SR; ++q)
              if (q == n) {
                xc = xs[q];
                tc = ts[q];
              }
          } else {
            xc = (float)(scl[c] * (0.0 - avg[c]));  // angular feature of a radial-only basis
            tc = 0.0f;
          }
          atomicAdd(&gacc[o1 + c * nh + k], bk * xc + gF * w2 * d1[k] * tc);
        }
      }
    }
    atomicAdd(&gacc[o2 + nh], gE);
  }
  // loss sums and the gradient out
  double le = wave_sum(lossE), lf = wave_sum(lossF);
  if (lane_id() == 0) {
    unsafeAtomicAdd(&loss[0], le);
    unsafeAtomicAdd(&loss[1], lf);
  }
  __syncthreads();
  for (int k = j; k < wsize; k += blockDim.x) {
    const float v = gacc[k];
    if (v != 0.0f) atomicAdd(&grad[k], v);
  }
}

void launch_first_stage_fit(const MCDev &m, const ModelDev &md, const int *cells, int ncells, const double *E_ref,
                            const double *F_ref, double coefE, double coefF, float *grad, int wsize, double *loss,
                            hipStream_t s) {
  if (ncells <= 0) return;
  const size_t lds = 3 * (size_t)m.natoms * sizeof(double) + (2 * (size_t)m.natoms * SR + 3 * (size_t)m.natoms +
                                                             (size_t)wsize) * sizeof(float);
  hipLaunchKernelGGL(k_first_stage_fit, dim3(ncells), dim3(256), lds, s, m, md, cells, E_ref, F_ref, coefE, coefF,
                     grad, wsize, loss);
}

// neighbor lists of the active atoms by minimum image within their chain
// (the cell is wider than twice the list radius, so every neighbor has one
// image in range); one wavefront per atom, persistent grid
__global__ __launch_bounds__(256) void k_mc_neighbors(MCDev m, NbDev nb, double rc2, int *flags) {
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  const int N = m.natoms;
  const int total = *m.nactive;
  for (int k = blockIdx.x * nw + wave; k < total; k += gridDim.x * nw) {
    const int g = m.active[k];
    const int c = g / N, j = g - c * N;
    const size_t base = (size_t)c * N;
    const double xj[3] = {m.coords[3 * (size_t)g], m.coords[3 * (size_t)g + 1], m.coords[3 * (size_t)g + 2]};
    int count = 0;
    for (int b0 = 0; b0 < N; b0 += WAVE) {
      const int q = b0 + lane;
      bool in = false;
      double d[3] = {0.0, 0.0, 0.0}, r2 = 0.0;
      if (q < N && q != j) {
        min_image(m, xj, &m.coords[3 * (base + q)], d);
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        in = r2 <= rc2;
      }
      const unsigned long long msk = wave_ballot(in);
      const int slot = count + mask_prefix(msk);
      if (in && slot < nb.cap) {
        const size_t e = (size_t)g * nb.cap + slot;
        reinterpret_cast<float4 *>(nb.R)[e] = make_float4((float)d[0], (float)d[1], (float)d[2], (float)sqrt(r2));
        nb.j[e] = (int)(base + q);
      }
      count += popc64(msk);
    }
    if (lane == 0) {
      if (count > nb.cap) atomicOr(flags, AENET_AMD_FLAG_NB_OVERFLOW);
      nb.count[g] = min(count, nb.cap);
    }
  }
}

// every atom active (refresh)
__global__ void k_mc_all_active(MCDev m) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = m.nchains * m.natoms;
  if (g == 0) *m.nactive = n;
  if (g < n) m.active[g] = g;
}

// chain energies from the atomic energies
__global__ __launch_bounds__(256) void k_mc_chain_sum(MCDev m) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  double e = 0.0;
  for (int j = threadIdx.x; j < m.natoms; j += blockDim.x) e += m.E_t[(size_t)c * m.natoms + j];
  e = wave_sum(e);
  if (lane_id() == 0) red[threadIdx.x / WAVE] = e;
  __syncthreads();
  if (threadIdx.x == 0) m.E_chain[c] = red[0] + red[1] + red[2] + red[3];
}

void launch_mc_neighbors(const MCDev &m, const NbDev &nb, double rc_list, int *flags, hipStream_t s) {
  const int grid = std::min((m.nchains * m.natoms + 3) / 4, device_limits().cus * 8);
  hipLaunchKernelGGL(k_mc_neighbors, dim3(std::max(grid, 1)), dim3(256), 0, s, m, nb, rc_list * rc_list, flags);
}

void launch_mc_trace_init(double *trace, long long n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_mc_trace_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, trace, n);
}

void launch_mc_all_active(const MCDev &m, hipStream_t s) {
  const int n = m.nchains * m.natoms;
  hipLaunchKernelGGL(k_mc_all_active, dim3((n + 255) / 256), dim3(256), 0, s, m);
}

void launch_mc_chain_sum(const MCDev &m, hipStream_t s) {
  hipLaunchKernelGGL(k_mc_chain_sum, dim3(m.nchains), dim3(256), 0, s, m);
}

size_t first_stage_table_bytes(int ntypes) { return ntypes * sizeof(FirstStage); }

void launch_first_stage_table(const ModelDev &student, void *tab, hipStream_t s) {
  hipLaunchKernelGGL(k_first_stage_table, dim3(1), dim3(256), 0, s, student, static_cast<FirstStage *>(tab));
}

void launch_mc_student_init(const MCDev &m, const ModelDev &student, hipStream_t s) {
  const int bs = std::min(256, ((m.natoms + WAVE - 1) / WAVE) * WAVE);
  hipLaunchKernelGGL(k_mc_student_init, dim3(m.nchains), dim3(bs),
                     3 * (size_t)m.natoms * sizeof(double) + student.ntypes * sizeof(FirstStage), s, m, student,
                     static_cast<const FirstStage *>(m.fs_tab));
}

template <bool CACHE, int RN, int NH>
static void launch_screen(const MCDev &m, const ModelDev &student, size_t lds, int bs, hipStream_t s) {
  hipLaunchKernelGGL((k_mc_screen<CACHE, RN, NH>), dim3(m.nchains), dim3(bs), lds, s, m, student,
                     static_cast<const FirstStage *>(m.fs_tab));
}

bool mc_screen_fits(int natoms, int ntypes, int rn) {
  return screen_lds_bytes(natoms, ntypes, rn, false) <= device_limits().lds_per_block;
}

void launch_mc_step(const MCDev &m, const ModelDev *student, bool plain, int phase, hipStream_t s) {
  const int C = m.nchains;
  const int bs = ((m.natoms + WAVE - 1) / WAVE) * WAVE;
  if (phase == 0) {
    if (plain)
      hipLaunchKernelGGL(k_mc_propose, dim3((C + 63) / 64), dim3(64), 0, s, m);
    else {
      // the chain's first-stage state in LDS when two workgroups still fit on a CU
      const size_t cached = screen_lds_bytes(m.natoms, student->ntypes, m.rn_s, true);
      const bool cache = 2 * cached <= device_limits().lds_per_cu && cached <= device_limits().lds_per_block;
      const size_t lds = cache ? cached : screen_lds_bytes(m.natoms, student->ntypes, m.rn_s, false);
      const bool small = m.rn_s <= 9 && m.nh_s <= 8;  // radial order 8, 8 hidden nodes
      const int wg = std::min(256, bs);
      if (cache && small)
        launch_screen<true, 9, 8>(m, *student, lds, wg, s);
      else if (cache)
        launch_screen<true, SR, SW>(m, *student, lds, wg, s);
      else if (small)
        launch_screen<false, 9, 8>(m, *student, lds, wg, s);
      else
        launch_screen<false, SR, SW>(m, *student, lds, wg, s);
    }
    hipLaunchKernelGGL(k_mc_active, dim3(C), dim3(bs), 0, s, m);
  } else {
    hipLaunchKernelGGL(k_mc_finish, dim3(C), dim3(std::min(256, bs)), 0, s, m, plain ? 1 : 0);
  }
}

}  // namespace aenet_amd

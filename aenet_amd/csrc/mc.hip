// mc.hip -- Metropolis Monte Carlo of many independent periodic chains on
// the GPU, plain and delayed-acceptance (two-stage) with a cheap radial-only
// potential as the first stage.
//
// One MC step moves every chain once:
//   k_mc_propose : per chain, a swap of two atoms of different species or a
//                  displacement of one atom by U(-d, d)^3, applied in place
//                  (the old values are kept for a revert); SplitMix64
//                  generator per chain, state in device memory
//   k_mc_student : (delayed acceptance) per chain, one thread per atom: the
//                  radial Chebyshev descriptor of every atom is kept, and
//                  only the pair terms with the moved sites change it, so
//                  the first-stage energy change needs no neighbor search;
//                  the moved atoms are recomputed over the whole cell;
//                  per-thread network; stage-1 decision exp(-beta dE_s)
//   k_mc_active  : per chain that goes to the full potential, the atoms
//                  within its cutoff of a moved site, as one contiguous
//                  block of a batch-wide list
//   (full potential on the listed atoms: the batch's active-atom path)
//   k_mc_finish  : per chain, dE of the full potential over its block; the
//                  decision (plain: exp(-beta dE); second stage:
//                  exp(-beta (dE - dE_s))); commit or revert; statistics and
//                  energy samples
// The second stage makes the chain exact for the full potential whatever
// the first-stage potential is (Christen & Fox, J. Comput. Graph. Stat.
// 14 (2005) 795).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "aenet_amd_internal.h"
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
// src/ext/sfbasis.f90:489-522): r_n += w_plain fc T_n, s_n += w_spin fc T_n
__device__ __forceinline__ void student_pair(const SpeciesDev &sp, double r, double w_plain, double w_spin,
                                             SDesc &G) {
  const double Rc = sp.r_Rc;
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
    if (n < sp.r_N) {
      G.r[n] += w_plain * fc * tn;
      G.s[n] += w_spin * fc * tn;
    }
  }
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

// the first stage's atomic energy from its unscaled descriptor:
// normalization (stp_normalize), ff_eval in per-thread FP32 registers,
// energy epilogue in FP64
__device__ __forceinline__ double student_energy(const ModelDev &md, const SpeciesDev &sp, const SDesc &G) {
  const double *avg = md.norm_avg + sp.norm_off;
  const double *scl = md.norm_scale + sp.norm_off;
  float xr[SR], xs[SR];
#pragma unroll
  for (int n = 0; n < SR; ++n) {
    xr[n] = n < sp.r_N ? (float)(scl[sp.r_i1 + n] * (G.r[n] - avg[sp.r_i1 + n])) : 0.0f;
    xs[n] = (n < sp.r_N && sp.multi) ? (float)(scl[sp.r_i2 + n] * (G.s[n] - avg[sp.r_i2 + n])) : 0.0f;
  }
  float h[SW];
  {
    const int n1 = sp.nnodes[0], n2 = sp.nnodes[1];
    const float *W = md.W + sp.w_off[0];
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      if (k < n2) {
        float z = W[n1 * n2 + k];
#pragma unroll
        for (int n = 0; n < SR; ++n) {
          if (n < sp.r_N) {
            z += W[(sp.r_i1 + n) * n2 + k] * xr[n];
            if (sp.multi) z += W[(sp.r_i2 + n) * n2 + k] * xs[n];
          }
        }
        // angular features of a radial-only basis: G = 0
        for (int n = 0; n < sp.a_N; ++n) {
          z += W[(sp.a_i1 + n) * n2 + k] * (float)(scl[sp.a_i1 + n] * (0.0 - avg[sp.a_i1 + n]));
          if (sp.multi) z += W[(sp.a_i2 + n) * n2 + k] * (float)(scl[sp.a_i2 + n] * (0.0 - avg[sp.a_i2 + n]));
        }
        h[k] = act_f(sp.fa[1], z);
      } else {
        h[k] = 0.0f;
      }
    }
  }
  for (int l = 1; l + 1 < sp.nlayers; ++l) {
    const int n1 = sp.nnodes[l], n2 = sp.nnodes[l + 1];
    const float *W = md.W + sp.w_off[l];
    float o[SW];
#pragma unroll
    for (int k = 0; k < SW; ++k) {
      if (k < n2) {
        float z = W[n1 * n2 + k];
#pragma unroll
        for (int c = 0; c < SW; ++c)
          if (c < n1) z += W[c * n2 + k] * h[c];
        o[k] = act_f(sp.fa[l + 1], z);
      } else {
        o[k] = 0.0f;
      }
    }
#pragma unroll
    for (int k = 0; k < SW; ++k) h[k] = o[k];
  }
  double E = sp.E_scale * (double)h[0] + sp.E_shift;
  return E + sp.E_atom;
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

__device__ __forceinline__ void sdesc_load(const MCDev &m, const SpeciesDev &sp, size_t g, SDesc &G) {
  const double *p = m.G_s + g * m.nsf_s;
#pragma unroll
  for (int n = 0; n < SR; ++n) {
    G.r[n] = n < sp.r_N ? p[sp.r_i1 + n] : 0.0;
    G.s[n] = (n < sp.r_N && sp.multi) ? p[sp.r_i2 + n] : 0.0;
  }
}

// stores the features in the new_SFBasis layout (angular ones zero)
__device__ __forceinline__ void sdesc_store(const MCDev &m, const SpeciesDev &sp, double *dst, const SDesc &G) {
  for (int f = 0; f < m.nsf_s; ++f) dst[f] = 0.0;
#pragma unroll
  for (int n = 0; n < SR; ++n) {
    if (n < sp.r_N) {
      dst[sp.r_i1 + n] = G.r[n];
      if (sp.multi) dst[sp.r_i2 + n] = G.s[n];
    }
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
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  int kind = MC_NONE, a = 0, b = 0;
  if (uniform01(s) < m.p_swap) {
    a = min((int)(uniform01(s) * N), N - 1);
    const int ta = m.types[base + a];
    for (int k = 0; k < 64; ++k) {
      const int cand = min((int)(uniform01(s) * N), N - 1);
      if (m.types[base + cand] != ta) {
        b = cand;
        kind = MC_SWAP;
        break;
      }
    }
    if (kind == MC_SWAP) {
      const int tb = m.types[base + b];
      m.types[base + a] = tb;
      m.types[base + b] = ta;
    }
  } else {
    kind = MC_DISP;
    a = min((int)(uniform01(s) * N), N - 1);
    for (int x = 0; x < 3; ++x) {
      const double old = m.coords[3 * (base + a) + x];
      m.old_pos[3 * c + x] = old;
      m.coords[3 * (base + a) + x] = old + (2.0 * uniform01(s) - 1.0) * m.max_disp;
    }
  }
  m.kind[c] = kind;
  m.site_a[c] = a;
  m.site_b[c] = b;
  m.u_stage1[c] = uniform01(s);
  m.u_stage2[c] = uniform01(s);
  m.rng[c] = s;
}

// one workgroup per chain, threads over the chain's atoms
__global__ __launch_bounds__(256) void k_mc_student(MCDev m, ModelDev md) {
  extern __shared__ double sx[];  // [3*N] positions of the chain
  __shared__ double red[16];
  __shared__ double Gm[2][2 * SR];  // descriptors of the moved sites, summed over the workgroup
  const int c = blockIdx.x;
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  const int kind = m.kind[c];
  for (int k = threadIdx.x; k < 3 * N; k += blockDim.x) sx[k] = m.coords[3 * base + k];
  if (threadIdx.x < 4 * SR) (&Gm[0][0])[threadIdx.x] = 0.0;
  __syncthreads();
  // the moved sites' descriptors: every other atom adds its pair term
  if (kind != MC_NONE) {
    const int a = m.site_a[c], b = m.site_b[c];
    const int ta = m.types[base + a];
    const int tb = m.types[base + b];
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
      const int tj = m.types[base + j];
      if (j != a) {
        const SpeciesDev &spa = md.species[ta - 1];
        const double r = dist(m, &sx[3 * a], &sx[3 * j]);
        if (r <= spa.r_Rc) {
          SDesc t;
          sdesc_zero(t);
          student_pair(spa, r, 1.0, spa.spin[tj - 1], t);
#pragma unroll
          for (int n = 0; n < SR; ++n) {
            if (n < spa.r_N) {
              unsafeAtomicAdd(&Gm[0][n], t.r[n]);
              unsafeAtomicAdd(&Gm[0][SR + n], t.s[n]);
            }
          }
        }
      }
      if (kind == MC_SWAP && j != b) {
        const SpeciesDev &spb = md.species[tb - 1];
        const double r = dist(m, &sx[3 * b], &sx[3 * j]);
        if (r <= spb.r_Rc) {
          SDesc t;
          sdesc_zero(t);
          student_pair(spb, r, 1.0, spb.spin[tj - 1], t);
#pragma unroll
          for (int n = 0; n < SR; ++n) {
            if (n < spb.r_N) {
              unsafeAtomicAdd(&Gm[1][n], t.r[n]);
              unsafeAtomicAdd(&Gm[1][SR + n], t.s[n]);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  double dE = 0.0;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    const size_t g = base + j;
    if (kind == MC_NONE) {
      m.affected[g] = 0;
      continue;
    }
    const int a = m.site_a[c], b = m.site_b[c];
    const int tj = m.types[g];  // new type (swaps already applied)
    const SpeciesDev &sp = md.species[tj - 1];
    SDesc G;
    bool affected = false;
    const bool full = (j == a) || (kind == MC_SWAP && j == b);
    if (full) {
      const int q = j == a ? 0 : 1;
#pragma unroll
      for (int n = 0; n < SR; ++n) {
        G.r[n] = Gm[q][n];
        G.s[n] = Gm[q][SR + n];
      }
      affected = true;
    } else {
      sdesc_load(m, sp, g, G);
      if (kind == MC_DISP) {
        // atom a left old_pos for its new position
        const double ro = dist(m, &sx[3 * j], &m.old_pos[3 * c]);
        const double rn = dist(m, &sx[3 * j], &sx[3 * a]);
        if (ro <= sp.r_Rc || rn <= sp.r_Rc) {
          const double spa = sp.spin[m.types[base + a] - 1];
          student_pair(sp, ro, -1.0, -spa, G);
          student_pair(sp, rn, 1.0, spa, G);
          affected = true;
        }
      } else {
        // swap: a and b exchanged species in place; only the typespin
        // weights of the pairs (j, a) and (j, b) change
        const int sites[2] = {a, b};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int st = sites[q];
          const double r = dist(m, &sx[3 * j], &sx[3 * st]);
          if (r <= sp.r_Rc) {
            const double s_new = sp.spin[m.types[base + st] - 1];
            const double s_old = sp.spin[m.types[base + sites[1 - q]] - 1];
            student_pair(sp, r, 0.0, s_new - s_old, G);
            affected = true;
          }
        }
      }
    }
    if (affected) {
      const double E = student_energy(md, sp, G);
      dE += E - m.E_s[g];
      sdesc_store(m, sp, m.G_trial + g * m.nsf_s, G);
      m.E_s_trial[g] = E;
    }
    m.affected[g] = affected ? 1 : 0;
  }
  // chain sum of dE_s
  double v = wave_sum(dE);
  if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)((blockDim.x + WAVE - 1) / WAVE); ++w) t += red[w];
    m.dE_s[c] = t;
    bool pass = kind != MC_NONE && log(m.u_stage1[c]) < -m.beta * t;
    m.need[c] = pass ? 1 : 0;
  }
}

// one workgroup per chain: the atoms the full potential must re-evaluate
__global__ __launch_bounds__(1024) void k_mc_active(MCDev m, int plain) {
  __shared__ int wcount[16];
  __shared__ int base_out;
  const int c = blockIdx.x;
  const int j = threadIdx.x;
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  const int kind = m.kind[c];
  if (plain && j == 0) m.need[c] = kind != MC_NONE ? 1 : 0;
  __syncthreads();
  const bool need = plain ? (kind != MC_NONE) : (m.need[c] != 0);
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
__global__ __launch_bounds__(64) void k_mc_finish(MCDev m, int plain) {
  const int c = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  const int kind = m.kind[c];
  const bool need = m.need[c] != 0;
  double dE = 0.0;
  if (need) {
    const int s0 = m.act_start[c], n = m.act_count[c];
    for (int k = lane; k < n; k += WAVE) {
      const int g = m.active[s0 + k];
      dE += m.E_trial[g] - m.E_t[g];
    }
  }
  dE = wave_sum(dE);
  bool accept = false;
  if (need) {
    const double arg = plain ? -m.beta * dE : -m.beta * (dE - m.dE_s[c]);
    accept = log(m.u_stage2[c]) < arg;
  }
  if (accept) {
    const int s0 = m.act_start[c], n = m.act_count[c];
    for (int k = lane; k < n; k += WAVE) {
      const int g = m.active[s0 + k];
      m.E_t[g] = m.E_trial[g];
    }
    if (!plain) {
      for (int j = lane; j < N; j += WAVE) {
        const size_t g = base + j;
        if (m.affected[g]) {
          for (int f = 0; f < m.nsf_s; ++f) m.G_s[g * m.nsf_s + f] = m.G_trial[g * m.nsf_s + f];
          m.E_s[g] = m.E_s_trial[g];
        }
      }
    }
  } else if (lane == 0 && kind != MC_NONE) {
    const int a = m.site_a[c], b = m.site_b[c];
    if (kind == MC_SWAP) {
      const int ta = m.types[base + a];
      m.types[base + a] = m.types[base + b];
      m.types[base + b] = ta;
    } else {
      for (int x = 0; x < 3; ++x) m.coords[3 * (base + a) + x] = m.old_pos[3 * c + x];
    }
  }
  if (lane == 0) {
    if (m.trace) {
      const long long k = *m.step - 1;
      if (k >= 0 && k < m.max_trace) {
        double *tr = m.trace + 4 * ((size_t)k * m.nchains + c);
        tr[0] = plain ? 0.0 : m.dE_s[c];
        tr[1] = need ? dE : 0.0;
        tr[2] = kind == MC_NONE ? -1.0 : (need ? 1.0 : 0.0);
        tr[3] = accept ? 1.0 : 0.0;
      }
    }
    if (accept) m.E_chain[c] += dE;
    if (kind != MC_NONE) {
      m.stats[4 * c] += 1;                     // proposed
      if (accept) m.stats[4 * c + 1] += 1;     // accepted
      if (need) m.stats[4 * c + 2] += 1;       // full-potential evaluations
      if (kind == MC_SWAP) m.stats[4 * c + 3] += 1;
    }
    const long step = *m.step;
    if (m.samples && m.sample_every > 0 && step % m.sample_every == 0) {
      const long k = step / m.sample_every - 1;
      if (k >= 0 && k < m.max_samples) m.samples[(size_t)k * m.nchains + c] = m.E_chain[c];
    }
  }
}

// first-stage state from scratch: every atom's descriptor over the whole
// cell and its energy, with the arithmetic of k_mc_student
__global__ __launch_bounds__(256) void k_mc_student_init(MCDev m, ModelDev md) {
  extern __shared__ double sx[];
  const int c = blockIdx.x;
  const int N = m.natoms;
  const size_t base = (size_t)c * N;
  for (int k = threadIdx.x; k < 3 * N; k += blockDim.x) sx[k] = m.coords[3 * base + k];
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    const size_t g = base + j;
    const SpeciesDev &sp = md.species[m.types[g] - 1];
    SDesc G;
    student_full(m, sp, sx, j, base, G);
    sdesc_store(m, sp, m.G_s + g * m.nsf_s, G);
    m.E_s[g] = student_energy(md, sp, G);
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

void launch_mc_all_active(const MCDev &m, hipStream_t s) {
  const int n = m.nchains * m.natoms;
  hipLaunchKernelGGL(k_mc_all_active, dim3((n + 255) / 256), dim3(256), 0, s, m);
}

void launch_mc_chain_sum(const MCDev &m, hipStream_t s) {
  hipLaunchKernelGGL(k_mc_chain_sum, dim3(m.nchains), dim3(256), 0, s, m);
}

void launch_mc_student_init(const MCDev &m, const ModelDev &student, hipStream_t s) {
  const int bs = std::min(256, ((m.natoms + WAVE - 1) / WAVE) * WAVE);
  hipLaunchKernelGGL(k_mc_student_init, dim3(m.nchains), dim3(bs), 3 * (size_t)m.natoms * sizeof(double), s, m,
                     student);
}

void launch_mc_step(const MCDev &m, const ModelDev *student, bool plain, int phase, hipStream_t s) {
  const int C = m.nchains;
  const int bs = ((m.natoms + WAVE - 1) / WAVE) * WAVE;
  if (phase == 0) {
    hipLaunchKernelGGL(k_mc_propose, dim3((C + 63) / 64), dim3(64), 0, s, m);
    if (!plain)
      hipLaunchKernelGGL(k_mc_student, dim3(C), dim3(std::min(256, bs)), 3 * (size_t)m.natoms * sizeof(double), s,
                         m, *student);
    hipLaunchKernelGGL(k_mc_active, dim3(C), dim3(bs), 0, s, m, plain ? 1 : 0);
  } else {
    hipLaunchKernelGGL(k_mc_finish, dim3(C), dim3(64), 0, s, m, plain ? 1 : 0);
  }
}

}  // namespace aenet_amd

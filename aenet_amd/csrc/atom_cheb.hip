// atom_cheb.hip -- per-atom kernel for Chebyshev (AUC) descriptors.
//
// Persistent workgroups of 4 wavefronts; each wavefront evaluates one
// central atom at a time:
//   1. neighbors  : compact list of k_search (structures) or the caller's
//                   list (per-atom entry points)
//   2. descriptors: sfb_eval (src/ext/sfbasis.f90:440-571) -- radial terms
//                   over lanes = pairs, angular terms over lanes = triplets;
//                   Chebyshev recurrences in FP32 registers; the per-lane
//                   partial sums are combined by a reduce-scatter butterfly
//                   (first level FP32, then FP64) that leaves feature f in
//                   lane f
//   3. normalize  : stp_normalize (src/sfsetup.f90:985-1035) in FP64
//   4. network    : ff_eval / ff_deriv (src/ext/feedforward.f90:586-757).
//                   Structures with one network architecture for every
//                   species: the atoms of a workgroup round form the rows
//                   of 16-row matrix tiles (v_mfma_f32_16x16x4f32), each
//                   weight read once per round instead of once per atom.
//                   Otherwise per wavefront, lanes = nodes of a layer.
//                   Weights of every species held in LDS for the whole
//                   workgroup (or read through the caches)
//   5. forces     : F_j -= E_scale dy/dG dG/dR_j (src/aenet.f90:471-477):
//                   pair and triplet derivatives contracted with dE/dG in
//                   FP32, per-neighbor sums in LDS, fixed-point (64-bit integer)
//                   atomics into F, so that forces do not depend on the
//                   order of the sums
#include <hip/hip_runtime.h>

#include <algorithm>

#include "aenet_amd_internal.h"
#include "kernels.h"
#include "kernels_common.h"
#include "atom_common.h"

namespace aenet_amd {

namespace {
constexpr float EPS_R = 1.0e-12f;
constexpr int AUXN = 64;  // angular neighbors whose per-neighbor terms are tabulated in LDS
}  // namespace

// Per-wavefront LDS region, laid out after the weights in dynamic shared
// memory; cap = neighbor capacity.
struct WaveLds {
  SpeciesDev *spc;  // the central atom's species record
  int *row;    // [1] the atom's species in the round's network tile (0: none)
  float4 *aux; // [AUXN] per angular neighbor: 1/r, fc, dfc/dr, typespin
  float *R;    // [3*cap]
  float *r;    // [cap]
  int *t;      // [cap] neighbor types
  int *j;      // [cap] neighbor ids (global atom or environment entry)
  int *ang;    // [cap] angular neighbor list
  float *fs;   // [3*cap] per-neighbor forces
  NetLds net;  // network stage
};

struct ChebLaunch {
  int natoms;    // number of central atoms (structures) or environments
  int cap;       // neighbor capacity of the LDS region
  int nsf;       // max features
  int nl;        // max layers
  int wsize;     // total weights
  int w_in_lds;  // weights staged in LDS
  int waves;     // wavefronts per workgroup
  int hs;        // network layer stride
  const int *active;   // optional subset of central atoms (mode 0)
  const int *nactive;
  int tune_waves, tune_staged;  // host side: measured launch shape
  int nnorm;     // normalization entries (averages and scales staged after the weights)
  int tile_net;  // networks of a round as one matrix tile (mode 0, one architecture)
  int ftab;      // entries of the workgroup force table (0: scatter straight to global memory)
};

// words staged once per workgroup: the weights (rounded to 4 words), then
// the normalization averages and scales (FP64)
__host__ __device__ inline int staged_words(const ChebLaunch &P) { return (P.wsize + 3) / 4 * 4 + 4 * P.nnorm; }

__host__ __device__ inline int wave_lds_words(const ChebLaunch &P) {
  return (SPECIES_WORDS + 4 + 4 * AUXN + 10 * P.cap + net_lds_words(P.nsf, P.nl, P.hs) + 3) / 4 * 4;
}

__device__ inline WaveLds carve(float *base, const ChebLaunch &P) {
  const int cap = P.cap;
  WaveLds w;
  w.spc = reinterpret_cast<SpeciesDev *>(base);
  base += SPECIES_WORDS;
  w.row = reinterpret_cast<int *>(base);
  base += 4;
  w.aux = reinterpret_cast<float4 *>(base);
  w.R = base + 4 * AUXN;
  w.r = w.R + 3 * cap;
  w.t = reinterpret_cast<int *>(w.r + cap);
  w.j = w.t + cap;
  w.ang = w.j + cap;
  w.fs = reinterpret_cast<float *>(w.ang + cap);
  w.net = carve_net(w.fs + 3 * cap, P.nsf, P.nl, P.hs);
  return w;
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// The networks of one round of a workgroup (ff_eval / ff_deriv,
// src/ext/feedforward.f90:586-757, with the chain through the
// normalization and the energy scale as network() has it) on the matrix
// cores.  Row q of a 16-row tile is wavefront q's atom; its input (the
// normalized descriptor), activations and back-propagation buffers stay in
// wavefront q's LDS region, and *row holds its species (0: no atom).  A
// task is one 16-column tile of one layer for one species present in the
// round: rows of other species enter as zeros and their results are
// dropped.  Every species has the same layer widths (nodes[], layer count
// nl, one output node); offsets, activations and energy scale come from
// the species record of the species' first row.
//   forward : Z[q][j] = sum_k X[q][k] W_l[k][j] + b_l[j], K = n_l
//   backward: D_l[q][k] = sum_j D_l+1[q][j] W_l[k][j] (times f'_l, or, for
//             the input layer, E_scale times the normalization scale)
// v_mfma_f32_16x16x4f32 operands: lane L gives A[L % 16][L / 16] and
// B[L / 16][L % 16] of each K step of 4; accumulator r of lane L is
// D[4 (L / 16) + r][L % 16].
// One workgroup barrier ahead of the tile (the rows' inputs and species),
// one after each layer; every wavefront of the workgroup executes all of
// them, with or without an atom.  On return each row's N.act holds its
// activations and, with GRAD, N.coef holds dE/dG.
template <bool GRAD, class WPtr>
__device__ inline void network_rows(WPtr W, float *regions, const ChebLaunch &P, const int lane) {
  typedef __attribute__((address_space(3))) float lfloat;
  typedef const __attribute__((address_space(3))) SpeciesDev lspecies;
  lfloat *R = (lfloat *)regions;
  const int wave = threadIdx.x / WAVE;
  const int nw = P.waves;
  const int wl = wave_lds_words(P);
  const int hs = P.hs;
  const int nl = P.nl;
  // word offsets within a region (the layout of carve / carve_net)
  const int o_row = SPECIES_WORDS;
  const int o_gp = SPECIES_WORDS + 4 + 4 * AUXN + 10 * P.cap;
  const int o_coef = o_gp + P.nsf;
  const int o_d0 = o_coef + P.nsf;
  const int o_d1 = o_d0 + net_dmax(P.nsf, hs);
  const int o_act = o_d1 + net_dmax(P.nsf, hs);
  const int o_dact = o_act + nl * hs;
  const int rq = lane & 15;  // A row and B/D column of this lane
  const int kq = lane >> 4;  // A/B k of this lane within a step, D row group
  __syncthreads();
  const int myt = rq < nw ? reinterpret_cast<__attribute__((address_space(3))) int *>(R)[rq * wl + o_row] : 0;
  // rows that are the first of their species
  bool dup = false;
#pragma unroll
  for (int m = 0; m < 15; ++m) {
    const int tm = __shfl(myt, m, WAVE);
    if (m < rq && tm == myt) dup = true;
  }
  const unsigned long long lead = wave_ballot(lane < 16 && myt > 0 && !dup);
  const int ns = popc64(lead);
  // species of this lane's four D rows, gathered here with every lane
  // active (a cross-lane read from a lane that a branch has switched off
  // does not return that lane's value)
  int ot[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) ot[r] = __shfl(myt, 4 * kq + r, WAVE);
  auto nth_row = [&](int s) {
    unsigned long long m = lead;
    for (int c = 0; c < s; ++c) m &= m - 1;
    return (int)__builtin_ctzll(m);
  };
  auto species = [&](int q) -> lspecies & { return *(lspecies *)(R + q * wl); };
  const int arow_base = rq * wl;        // this lane's A row (read only when of the task's species)
  const int drow_base = 4 * kq * wl;    // first of this lane's four D rows

  for (int l = 0; l + 1 < nl; ++l) {
    if (ns > 0) {
      lspecies &sp0 = species(nth_row(0));
      const int n1 = sp0.nnodes[l], n2 = sp0.nnodes[l + 1];
      const int ntile = (n2 + 15) >> 4;
      const lfloat *X = R + arow_base + (l == 0 ? o_gp : o_act + l * hs);
      for (int task = wave; task < ns * ntile; task += nw) {
        const int s = task / ntile, jt = task - s * ntile;
        const int lr = nth_row(s);
        const int t = __shfl(myt, lr, WAVE);
        lspecies &st = species(lr);
        const WPtr Wl = W + st.w_off[l];
        const int col = jt * 16 + rq;
        const bool arow = myt == t;
        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = 0; k0 < n1; k0 += 4) {
          const int kk = k0 + kq;
          const float a = (arow && kk < n1) ? X[kk] : 0.0f;
          const float bw = (col < n2 && kk < n1) ? Wl[kk * n2 + col] : 0.0f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bw, acc, 0, 0, 0);
        }
        if (col < n2) {
          const float bias = Wl[n1 * n2 + col];
          const int fa = st.fa[l + 1];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (ot[r] == t) {
              lfloat *D = R + drow_base + r * wl + (l + 1) * hs + col;
              float y, dy;
              activate(fa, acc[r] + bias, y, dy);
              D[o_act] = y;
              D[o_dact] = dy;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if (!GRAD) return;
  for (int l = nl - 2; l >= 0; --l) {
    if (ns > 0) {
      lspecies &sp0 = species(nth_row(0));
      const int n1 = sp0.nnodes[l], n2 = sp0.nnodes[l + 1];
      const int ntile = (n1 + 15) >> 4;
      // D of layer l+1: the output layer's f', else the buffer of layer
      // l+1's parity; layer l's goes to the other one (the dE/dG words for
      // l = 0)
      const lfloat *src = R + arow_base + (l == nl - 2 ? o_dact + (nl - 1) * hs : ((l + 1) & 1 ? o_d1 : o_d0));
      const int o_dst = l == 0 ? o_coef : (l & 1 ? o_d1 : o_d0);
      for (int task = wave; task < ns * ntile; task += nw) {
        const int s = task / ntile, jt = task - s * ntile;
        const int lr = nth_row(s);
        const int t = __shfl(myt, lr, WAVE);
        lspecies &st = species(lr);
        const WPtr Wl = W + st.w_off[l];
        const int col = jt * 16 + rq;
        const bool arow = myt == t;
        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int j0 = 0; j0 < n2; j0 += 4) {
          const int jj = j0 + kq;
          const float a = (arow && jj < n2) ? src[jj] : 0.0f;
          const float bw = (col < n1 && jj < n2) ? Wl[col * n2 + jj] : 0.0f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bw, acc, 0, 0, 0);
        }
        if (col < n1) {
          // for l = 0 the destination holds the normalization scale
          const float m0 = l == 0 ? (float)st.E_scale : 1.0f;
          const int o_mul = l == 0 ? o_coef : o_dact + l * hs;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (ot[r] == t) {
              lfloat *D = R + drow_base + r * wl + col;
              D[o_dst] = m0 * acc[r] * D[o_mul];
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

// MODE 0: structures (compact lists, fixed-point atomic force scatter)
// MODE 1: explicit environments (per-neighbor forces out)
// TILE: the network of a round's atoms as one matrix tile (MODE 0)
template <int RN, int AN, int MODE, bool FORCES, bool TILE, bool WIDE>
__global__ __launch_bounds__(64 * CHEB_MAX_WAVES, 1) void k_atom_cheb(ModelDev md, BatchDev b, NbDev nbl, OutDev o, EnvDev env,
                                                          ChebLaunch P) {
  static_assert(2 * (RN + AN) <= 64, "features of one chunk");
  extern __shared__ float lds[];
  const int wave = threadIdx.x / WAVE;
  const int nw = P.waves;
  // weights and normalization of every species, once per workgroup
  const int wW = (P.wsize + 3) / 4 * 4;
  double *lnorm = reinterpret_cast<double *>(lds + wW);  // [2*nnorm]: averages, scales
  if (P.w_in_lds) {
    for (int k = threadIdx.x; k < P.wsize; k += 64 * nw) lds[k] = md.W[k];
    for (int k = threadIdx.x; k < P.nnorm; k += 64 * nw) {
      lnorm[k] = md.norm_avg[k];
      lnorm[P.nnorm + k] = md.norm_scale[k];
    }
    __syncthreads();
  }
  const int wsz = P.w_in_lds ? staged_words(P) : 0;
  WaveLds S = carve(lds + wsz + wave * wave_lds_words(P), P);

  const int count = P.active ? *P.nactive : P.natoms;
  // rounds of gridDim.x * nw entries: wavefront w of workgroup b takes
  // entry b * nw + w of each round.  With the tile network every wavefront
  // runs every round of its workgroup (the tile's barriers), with or
  // without an atom.
  constexpr bool tile = MODE == 0 && TILE;
  const int first = blockIdx.x * nw, per_round = gridDim.x * nw;
  const int nrounds = count > first ? (count - first + per_round - 1) / per_round : 0;
  // Forces of a workgroup that runs one round with its weights staged are
  // summed per target atom in an LDS table that takes the weights' place
  // after the round (workgroup-uniform); the wavefront's gradients wait in
  // its own region until then.
  const int ftab = MODE == 0 && FORCES && P.w_in_lds && nrounds == 1 ? P.ftab : 0;
  int agg_g = -1, agg_n = 0;  // this wavefront's atom and neighbor count for the table
  for (int rd = 0; rd < nrounds; ++rd) {
    // the lane index made opaque once per round: lane masks are computed
    // where they are used instead of being hoisted out of the loop and
    // held in scalar registers across it
    int lane = lane_id();
    asm volatile("" : "+v"(lane));
    const int k = first + rd * per_round + wave;
    const bool live = k < count;
    if (!tile && !live) break;
    const int g = live ? (P.active ? P.active[k] : k) : 0;
    wave_sync();
    // phases 1-3; false for an atom of an unknown species
    int ti = 0, n = 0, na = 0;
    int lane_f = -1;       // the species feature this lane holds after the reduce-scatter (-1: none)
    float lane_x = 0.0f;   // its normalized value
    const bool front_ok = live && [&]() -> bool {
      // ------------------------------------------------------------ 1
      int spw;
      if (MODE == 0) {
        // Two dependent memory rounds per atom: (type, count, list slots),
        // which depend on g alone, then (species record, neighbor types).
        // Every slot up to the capacity is read (all in bounds; those beyond
        // the count go unused), so the list loads do not wait for the count.
        ti = b.types[g];
        const int cnt = nbl.count[g];
        const float4 *R4 = reinterpret_cast<const float4 *>(nbl.R) + (size_t)g * nbl.cap;
        const int *J = nbl.j + (size_t)g * nbl.cap;
        auto load_slots = [&](int q0, float4 (&v)[2], int (&j)[2]) {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int q = q0 + u * WAVE + lane;
            v[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            j[u] = 0;
            if (q < P.cap) {
              v[u] = R4[q];
              j[u] = J[q];
            }
          }
        };
        auto gather_types = [&](int q0, const int (&j)[2], int (&tj)[2]) {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            // unconditional (slots past the count read the atom's own type),
            // so both gathers issue back to back
            const int jj = q0 + u * WAVE + lane < n ? j[u] : g;
            tj[u] = b.types[jj];
          }
        };
        auto put_slots = [&](int q0, const float4 (&v)[2], const int (&j)[2], const int (&tj)[2]) {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int q = q0 + u * WAVE + lane;
            if (q < n) {
              S.R[3 * q] = v[u].x;
              S.R[3 * q + 1] = v[u].y;
              S.R[3 * q + 2] = v[u].z;
              S.r[q] = v[u].w;
              S.j[q] = j[u];
              S.t[q] = tj[u];
            }
          }
        };
        float4 v[2];
        int j[2], tj[2];
        load_slots(0, v, j);
        n = min(cnt, P.cap);
        gather_types(0, j, tj);
        // issued after the gathers: loads complete in order, so the stores
        // below wait for the gathers without waiting for this one
        spw = species_word(md, min(max(ti, 1), md.ntypes) - 1);  // clamped: a bad type is refused below
        put_slots(0, v, j, tj);
        for (int q0 = 2 * WAVE; q0 < P.cap; q0 += 2 * WAVE) {
          load_slots(q0, v, j);
          gather_types(q0, j, tj);
          put_slots(q0, v, j, tj);
        }
      } else {
        ti = env.env_type[g];
        spw = species_word(md, min(max(ti, 1), md.ntypes) - 1);
        int e0 = env.env_start[g];
        n = env.env_start[g + 1] - e0;
        if (n > P.cap) {
          if (lane == 0) atomicOr(b.flags, AENET_AMD_FLAG_NB_OVERFLOW);
          n = P.cap;
        }
        const double ci[3] = {env.coo_i[3 * g], env.coo_i[3 * g + 1], env.coo_i[3 * g + 2]};
        for (int q = lane; q < n; q += WAVE) {
          double c[3];
          for (int k = 0; k < 3; ++k) c[k] = env.coo_j[3 * (size_t)(e0 + q) + k] - ci[k];
          S.R[3 * q] = (float)c[0];
          S.R[3 * q + 1] = (float)c[1];
          S.R[3 * q + 2] = (float)c[2];
          S.r[q] = (float)sqrt(dist2_exact(c));
          S.j[q] = e0 + q;
          S.t[q] = env.type_j[e0 + q];
        }
      }
      ti = uniform(ti);
      n = uniform(n);
      if (ti < 1 || ti > md.ntypes) {
        if (lane == 0) atomicOr(b.flags, AENET_AMD_FLAG_BAD_TYPE);
        return false;
      }
      wave_sync();
      const SpeciesDev &sp = store_species(spw, S.spc);
      const int rN = sp.r_N, aN = sp.a_N;
      const bool multi = sp.multi != 0;
      const float rRc = sp.r_Rc, aRc = sp.a_Rc;
      const float inv_rRc = 1.0f / rRc;
      const float inv_aRc = aRc > 0.0f ? 1.0f / aRc : 0.0f;

      // ------------------------------------------------------------ 2
      // accumulators in template layout: [radial RN | angular AN |
      // radial*spin RN | angular*spin AN | padding]
      float acc[64];
#pragma unroll
      for (int k = 0; k < 64; ++k) acc[k] = 0.0f;
      for (int q = lane; q < n; q += WAVE) {
        float r = S.r[q];
        if (r <= rRc && r > EPS_R) {
          float fc, dfc;
          cutoff(r, rRc, inv_rRc, fc, dfc);
          // kept for the radial force pass, in the words of the neighbor's
          // force, which that pass writes afterwards
          S.fs[3 * q] = fc;
          S.fs[3 * q + 1] = dfc;
          float x = 2.0f * r * inv_rRc - 1.0f;
          float sfc = sp.spin[S.t[q] - 1] * fc;
          float Tm = 1.0f, T = x;
          acc[0] += fc;
          acc[RN + AN] += sfc;
#pragma unroll
          for (int k = 1; k < RN; ++k) {
            acc[k] += fc * T;
            acc[RN + AN + k] += sfc * T;
            float Tn = 2.0f * x * T - Tm;
            Tm = T;
            T = Tn;
          }
        }
      }
      // angular neighbor list (order preserved)
      na = 0;
      for (int base = 0; base < n; base += WAVE) {
        int q = base + lane;
        bool in = false;
        if (q < n) {
          float r = S.r[q];
          in = r <= aRc && r >= EPS_R;
        }
        unsigned long long m = wave_ballot(in);
        if (in) S.ang[na + mask_prefix(m)] = q;
        na += popc64(m);
      }
      wave_sync();
      // per angular neighbor: 1/r, cutoff and its derivative, typespin --
      // tabulated once per atom for both triplet passes
      const bool use_aux = na <= AUXN;
      auto ang_terms = [&](int q) -> float4 {
        const float r = S.r[q];
        float fc, dfc;
        cutoff(r, aRc, inv_aRc, fc, dfc);
        return make_float4(1.0f / r, fc, dfc, multi ? sp.spin[S.t[q] - 1] : 0.0f);
      };
      if (use_aux) {
        if (lane < na) S.aux[lane] = ang_terms(S.ang[lane]);
        wave_sync();
      }
      auto ang_get = [&](int a, int q) -> float4 { return use_aux ? S.aux[a] : ang_terms(q); };
      const int ntrip = na * (na - 1) / 2;
      for (int p = lane; p < ntrip; p += WAVE) {
        int a, c;
        tri_decode_diag(p, na, a, c);
        int qj = S.ang[a], qk = S.ang[c];
        const float4 Aj = ang_get(a, qj), Ak = ang_get(c, qk);
        float cosv = (S.R[3 * qj] * S.R[3 * qk] + S.R[3 * qj + 1] * S.R[3 * qk + 1] + S.R[3 * qj + 2] * S.R[3 * qk + 2]) *
                     (Aj.x * Ak.x);
        float w = Aj.y * Ak.y;
        float sw = Aj.w * Ak.w * w;
        float Tm = 1.0f, T = cosv;
        acc[RN] += w;
        acc[2 * RN + AN] += sw;
#pragma unroll
        for (int k = 1; k < AN; ++k) {
          acc[RN + k] += w * T;
          acc[2 * RN + AN + k] += sw * T;
          float Tn = 2.0f * cosv * T - Tm;
          Tm = T;
          T = Tn;
        }
      }
      double G = reduce_scatter<64>(acc);
      // lane f holds template feature f; map it to the species' layout
      {
        int f = -1;
        if (lane < RN) {
          if (lane < rN) f = sp.r_i1 + lane;
        } else if (lane < RN + AN) {
          if (lane - RN < aN) f = sp.a_i1 + lane - RN;
        } else if (lane < 2 * RN + AN) {
          if (multi && lane - RN - AN < rN) f = sp.r_i2 + lane - RN - AN;
        } else if (lane < 2 * (RN + AN)) {
          if (multi && lane - 2 * RN - AN < aN) f = sp.a_i2 + lane - 2 * RN - AN;
        }
        if (f >= 0) {
          // ------------------------------------------------------ 3
          // (staged copies read through LDS-typed pointers: DS loads)
          typedef const __attribute__((address_space(3))) double *lds_f64;
          const double sc = P.w_in_lds ? ((lds_f64)lnorm)[P.nnorm + sp.norm_off + f] : md.norm_scale[sp.norm_off + f];
          const double av = P.w_in_lds ? ((lds_f64)lnorm)[sp.norm_off + f] : md.norm_avg[sp.norm_off + f];
          double Gs = sc * (G - av);
          S.net.gp[f] = (float)Gs;
          lane_f = f;
          lane_x = (float)Gs;
          S.net.coef[f] = (float)sc;  // for the network's dE/dG
          if (o.G) o.G[(size_t)g * o.G_stride + f] = o.G_scaled ? Gs : G;
        }
      }
      wave_sync();
      return true;
    }();
    const bool ok = uniform((int)front_ok) != 0;

    // ------------------------------------------------------------ 4
    float yout = 0.0f;
    if constexpr (tile) {
      {
        if (lane == 0) *S.row = ok ? ti : 0;
        if (P.w_in_lds)
          network_rows<FORCES>((lds_weights)lds, lds + wsz, P, lane);
        else
          network_rows<FORCES>((global_weights)md.W, lds + wsz, P, lane);
        if (!ok) continue;
        yout = S.net.act[(S.spc->nlayers - 1) * S.net.hs];
      }
    } else {
      if (!ok || o.desc_only) continue;
      const SpeciesDev &sp = *S.spc;
      if (MODE == 1 && o.coef_unit > 0) {
        // derivatives of one descriptor feature instead of the energy
        const int fu = g % o.coef_unit;
        for (int f = lane; f < sp.nsf; f += WAVE) S.net.coef[f] = f == fu ? 1.0f : 0.0f;
        wave_sync();
      } else {
        // the first layer from the lanes' features when it has at most 32 nodes
        const int first = sp.nnodes[1] <= 32 ? 1 : 0;
        if (first) {
          if (P.w_in_lds)
            first_layer_from_lanes(sp, (lds_weights)lds, lane_f, lane_x, S.net);
          else
            first_layer_from_lanes(sp, (global_weights)md.W, lane_f, lane_x, S.net);
        }
        yout = P.w_in_lds ? network<WIDE>(md, sp, (lds_weights)lds, S.net, FORCES, true, first)
                          : network<WIDE>(md, sp, (global_weights)md.W, S.net, FORCES, true, first);
      }
    }
    const SpeciesDev &sp = *S.spc;
    const int rN = sp.r_N, aN = sp.a_N;
    const bool multi = sp.multi != 0;
    const float rRc = sp.r_Rc, aRc = sp.a_Rc;
    const float inv_rRc = 1.0f / rRc;
    const float inv_aRc = aRc > 0.0f ? 1.0f / aRc : 0.0f;
    const bool use_aux = na <= AUXN;
    auto ang_terms = [&](int q) -> float4 {
      const float r = S.r[q];
      float fc, dfc;
      cutoff(r, aRc, inv_aRc, fc, dfc);
      return make_float4(1.0f / r, fc, dfc, multi ? sp.spin[S.t[q] - 1] : 0.0f);
    };
    auto ang_get = [&](int a, int q) -> float4 { return use_aux ? S.aux[a] : ang_terms(q); };
    const int ntrip = na * (na - 1) / 2;
    if (lane == 0) {
      if (MODE == 0)
        o.E_atom[g] = atom_energy(sp, yout);
      else
        env.E[g] = atom_energy(sp, yout);
    }
    if (!FORCES) continue;
    if (MODE == 0 && o.g_atom) {
      // forces of g_i * E_i: scale dE_i/dG
      const float gs = (float)o.g_atom[g];
      for (int f = lane; f < sp.nsf; f += WAVE) S.net.coef[f] *= gs;
      wave_sync();
    }

    // ------------------------------------------------------------ 5
    float *fs = S.fs;
    const float *cr1 = S.net.coef + sp.r_i1, *cr2 = S.net.coef + sp.r_i2;
    const float *ca1 = S.net.coef + sp.a_i1, *ca2 = S.net.coef + sp.a_i2;
    // dE/dG of the radial terms, plain (p) and typespin (s): in registers
    // up to 19 terms, read from LDS beyond (register budget)
    constexpr bool RREG = RN <= 19;
    constexpr int RR = RREG ? RN : 1;
    float crp_[RR], crs_[RR];
#pragma unroll
    for (int k = 0; k < RR; ++k) {
      crp_[k] = k < rN ? cr1[k] : 0.0f;
      crs_[k] = (multi && k < rN) ? cr2[k] : 0.0f;
    }
    auto rcoef = [&](int k, float sj) -> float {
      if constexpr (RREG)
        return crp_[k] + sj * crs_[k];
      else
        return k < rN ? cr1[k] + sj * (multi ? cr2[k] : 0.0f) : 0.0f;
    };
    for (int q = lane; q < n; q += WAVE) {
      float gx = 0.0f, gy = 0.0f, gz = 0.0f;
      float r = S.r[q];
      if (r <= rRc && r > EPS_R) {
        const float fc = fs[3 * q], dfc = fs[3 * q + 1];  // from the descriptor pass
        float x = 2.0f * r * inv_rRc - 1.0f;
        float sj = multi ? sp.spin[S.t[q] - 1] : 0.0f;
        float fcx = fc * 2.0f * inv_rRc;
        // sigma = dfc S + fc S' dx/dr with S = sum_n a_n T_n(x), both sums by
        // Clenshaw's recurrence: b_n = a_n + 2x b_n+1 - b_n+2 gives
        // S = a_0 + x b_1 - b_2, and its derivative d_n = 2 b_n+1 + 2x d_n+1 -
        // d_n+2 gives S' = b_1 + x d_1 - d_2 (zero coefficients beyond the
        // species' rN)
        const float x2 = 2.0f * x;
        float b1 = 0.0f, b2 = 0.0f, d1 = 0.0f, d2 = 0.0f;
#pragma unroll
        for (int k = RN - 1; k >= 1; --k) {
          const float a = rcoef(k, sj);
          const float b0 = fmaf(x2, b1, a - b2);
          const float d0 = fmaf(x2, d1, fmaf(2.0f, b1, -d2));
          b2 = b1;
          b1 = b0;
          d2 = d1;
          d1 = d0;
        }
        const float Sv = fmaf(x, b1, rcoef(0, sj) - b2);
        const float Sd = fmaf(x, d1, b1 - d2);
        const float sig = fmaf(dfc, Sv, fcx * Sd);
        float s = sig / r;
        gx = s * S.R[3 * q];
        gy = s * S.R[3 * q + 1];
        gz = s * S.R[3 * q + 2];
      }
      fs[3 * q] = gx;
      fs[3 * q + 1] = gy;
      fs[3 * q + 2] = gz;
    }
    wave_sync();
    // dE/dG of the angular terms in registers: plain (p) and typespin (s)
    float cap_[AN], cas_[AN];
#pragma unroll
    for (int k = 0; k < AN; ++k) {
      cap_[k] = k < aN ? ca1[k] : 0.0f;
      cas_[k] = (multi && k < aN) ? ca2[k] : 0.0f;
    }
    // gradients of triplet (angular neighbors a, c) with respect to R_j and R_k
    auto trip_grad = [&](int a, int c, float (&gj)[3], float (&gk)[3]) {
      const int qj = S.ang[a], qk = S.ang[c];
      const float4 Aj = ang_get(a, qj), Ak = ang_get(c, qk);
      const float ij = Aj.x, ik = Ak.x;
      float Rj[3] = {S.R[3 * qj], S.R[3 * qj + 1], S.R[3 * qj + 2]};
      float Rk[3] = {S.R[3 * qk], S.R[3 * qk + 1], S.R[3 * qk + 2]};
      float ijk = ij * ik;
      float cosv = (Rj[0] * Rk[0] + Rj[1] * Rk[1] + Rj[2] * Rk[2]) * ijk;
      const float fcj = Aj.y, dfcj = Aj.z, fck = Ak.y, dfck = Ak.z;
      float w = fcj * fck;
      float ss = Aj.w * Ak.w;
      // A = sum_n c_n T_n(cos), B = sum_n c_n T'_n(cos), c_n = p_n + ss s_n,
      // by Clenshaw's recurrence and its derivative (as in the radial pass)
      const float c2 = 2.0f * cosv;
      float b1 = 0.0f, b2 = 0.0f, d1 = 0.0f, d2 = 0.0f;
#pragma unroll
      for (int k = AN - 1; k >= 1; --k) {
        const float c = fmaf(ss, cas_[k], cap_[k]);
        const float b0 = fmaf(c2, b1, c - b2);
        const float d0 = fmaf(c2, d1, fmaf(2.0f, b1, -d2));
        b2 = b1;
        b1 = b0;
        d2 = d1;
        d1 = d0;
      }
      const float A = fmaf(cosv, b1, fmaf(ss, cas_[0], cap_[0]) - b2);
      const float Bd = fmaf(cosv, d1, b1 - d2);
      float wB = w * Bd;
      float aj = A * dfcj * fck * ij, ak = A * fcj * dfck * ik;
      float cj = cosv * ij * ij, ck = cosv * ik * ik;
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        gj[x] = aj * Rj[x] + wB * (Rk[x] * ijk - cj * Rj[x]);
        gk[x] = ak * Rk[x] + wB * (Rj[x] * ijk - ck * Rk[x]);
      }
    };
    // Schedule of the angular force pass.  With the per-neighbor table
    // (na <= AUXN), the lanes form G = 64 / na groups of na lanes and lane
    // (group h, a) owns angular neighbor a: in round d (1 .. na/2, the rounds
    // dealt round-robin to the groups) it takes the triplet (a, a + d mod na)
    // -- every unordered pair once, the half round d = na/2 only for a < d --
    // keeps its own gradient in registers and hands the partner's to the
    // partner's owner in its group with one cross-lane read per component.
    // The groups' sums are added per neighbor at the end: no atomics.  When
    // that takes more than twice the rounds of 64 triplets at a time, the
    // triplets go over the lanes as in the value pass and add to the
    // per-neighbor sums with LDS atomics.
    const int rounds = na / 2;
    const int groups = na >= 2 ? WAVE / na : 1;
    const int rot_iters = na >= 2 ? (rounds + groups - 1) / groups : 0;
    if (use_aux && na >= 2 && rot_iters <= 2 * ((ntrip + WAVE - 1) / WAVE)) {
      // lane / na from a float reciprocal: (lane + 1/2) / na stays at least
      // 1/128 from an integer (lane < 64, na <= 64), far beyond the rounding
      const int h = (int)(((float)lane + 0.5f) * __builtin_amdgcn_rcpf((float)na)), a = lane - h * na;
      float acc[3] = {0.0f, 0.0f, 0.0f};
      for (int it = 0; it < rot_iters; ++it) {
        const int d = it * groups + h + 1;
        const bool in_round = h < groups && d <= rounds;
        const bool act = in_round && !(2 * d == na && a >= d);
        float gk[3] = {0.0f, 0.0f, 0.0f};
        if (act) {
          int c = a + d;
          if (c >= na) c -= na;
          float gj[3];
          trip_grad(a, c, gj, gk);
#pragma unroll
          for (int x = 0; x < 3; ++x) acc[x] += gj[x];
        }
        // lane (h, a) takes the gradient of lane (h, a - d), whose partner is a
        int sa = a - d;
        if (sa < 0) sa += na;
        const int src = in_round ? h * na + sa : lane;
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[x] += __shfl(gk[x], src, WAVE);
      }
      float tot[3] = {acc[0], acc[1], acc[2]};
      for (int hh = 1; hh < groups; ++hh) {
        const int src = h == 0 ? hh * na + a : lane;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const float t = __shfl(acc[x], src, WAVE);
          if (h == 0) tot[x] += t;
        }
      }
      if (h == 0) {  // lanes a < na: distinct neighbor slots
        const int q = S.ang[a];
#pragma unroll
        for (int x = 0; x < 3; ++x) fs[3 * q + x] += tot[x];
      }
    } else {
      for (int p = lane; p < ntrip; p += WAVE) {
        int a, c;
        tri_decode_diag(p, na, a, c);
        float gj[3], gk[3];
        trip_grad(a, c, gj, gk);
        const int qj = S.ang[a], qk = S.ang[c];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          atomicAdd(&fs[3 * qj + x], gj[x]);
          atomicAdd(&fs[3 * qk + x], gk[x]);
        }
      }
    }
    wave_sync();
    // F_j -= dE_i/dR_j ; F_i += sum_j dE_i/dR_j
    if constexpr (MODE == 0) {
      if (ftab) {
        agg_g = g;
        agg_n = n;
      } else {
        scatter_forces(o.F, b.fmark, g, n, S.j, [&](int q, double (&c)[3]) {
          c[0] = fs[3 * q];
          c[1] = fs[3 * q + 1];
          c[2] = fs[3 * q + 2];
        });
      }
    } else {
      double own[3] = {0.0, 0.0, 0.0};
      for (int q = lane; q < n; q += WAVE) {
        double gx = fs[3 * q], gy = fs[3 * q + 1], gz = fs[3 * q + 2];
        own[0] += gx;
        own[1] += gy;
        own[2] += gz;
        size_t e = (size_t)S.j[q];
        env.dF_j[3 * e] = -gx;
        env.dF_j[3 * e + 1] = -gy;
        env.dF_j[3 * e + 2] = -gz;
      }
      for (int x = 0; x < 3; ++x) own[x] = wave_sum(own[x]);
      if (lane == 0) {
        env.dF_i[3 * g] = own[0];
        env.dF_i[3 * g + 1] = own[1];
        env.dF_i[3 * g + 2] = own[2];
      }
    }
    wave_sync();
  }
  if constexpr (MODE == 0 && FORCES) {
    if (ftab) {
      __syncthreads();  // every wavefront is done with the staged weights
      const ForceTable t = force_table_at(lds, ftab);
      for (int k = threadIdx.x; k < ftab; k += 64 * nw) t.key[k] = -1;
      for (int k = threadIdx.x; k < 3 * ftab; k += 64 * nw) t.val[k] = 0ull;
      __syncthreads();
      if (agg_g >= 0)
        scatter_forces_table(t, o.F, b.fmark, agg_g, agg_n, S.j, [&](int q, double (&c)[3]) {
          c[0] = S.fs[3 * q];
          c[1] = S.fs[3 * q + 1];
          c[2] = S.fs[3 * q + 2];
        });
      __syncthreads();
      // each target's sums into one of its slots (one 32-bit atomic for the
      // slot, plain stores) or, without a free slot, into F directly
      for (int k = threadIdx.x; k < ftab; k += 64 * nw) {
        const int a = t.key[k];
        if (a >= 0) {
          const long long v[3] = {(long long)t.val[3 * k], (long long)t.val[3 * k + 1], (long long)t.val[3 * k + 2]};
          const int slot = b.fslots ? atomicAdd(&b.fcnt[a], 1) : b.fslots;
          if (slot < b.fslots) {
            unsigned long long *d = b.fslot + 3 * ((size_t)a * b.fslots + slot);
            d[0] = (unsigned long long)v[0];
            d[1] = (unsigned long long)v[1];
            d[2] = (unsigned long long)v[2];
          } else {
            force_add_fixed(o.F, a, v);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------
// launch: template dispatch on the (radial, angular) term counts
// ---------------------------------------------------------------------

template <int RN, int AN, int MODE, bool F>
static void launch_one(const ModelDev &md, const BatchDev &b, const NbDev &nb, const OutDev &o, const EnvDev &e,
                       ChebLaunch P, hipStream_t stream) {
  PersistentShape ps;
  const size_t wbytes = P.wsize <= 16384 ? 4 * (size_t)staged_words(P) : 0;
  if (!persistent_shape(wbytes, 4 * (size_t)wave_lds_words(P), P.natoms, 4, ps, CHEB_MAX_WAVES, P.tune_waves,
                        P.tune_staged)) {
    // cannot happen within the capacity limits; refuse rather than overrun LDS
    raise_flag(b.flags, AENET_AMD_FLAG_NB_OVERFLOW, stream);
    return;
  }
  P.w_in_lds = ps.w_in_lds;
  P.waves = ps.waves;
  if (MODE == 0 && P.tile_net && !o.desc_only)
    hipLaunchKernelGGL((k_atom_cheb<RN, AN, MODE, F, MODE == 0, false>), dim3(ps.grid), dim3(64 * ps.waves), ps.lds,
                       stream, md, b, nb, o, e, P);
  else if (P.hs > 64)  // hidden layers wider than the wavefront
    hipLaunchKernelGGL((k_atom_cheb<RN, AN, MODE, F, false, true>), dim3(ps.grid), dim3(64 * ps.waves), ps.lds, stream,
                       md, b, nb, o, e, P);
  else
    hipLaunchKernelGGL((k_atom_cheb<RN, AN, MODE, F, false, false>), dim3(ps.grid), dim3(64 * ps.waves), ps.lds,
                       stream, md, b, nb, o, e, P);
}

template <int RN, int AN>
static void launch_cheb(int mode, bool forces, const ModelDev &md, const BatchDev &b, const NbDev &nb,
                        const OutDev &o, const EnvDev &e, ChebLaunch P, hipStream_t stream) {
  if (mode == 0) {
    if (forces)
      launch_one<RN, AN, 0, true>(md, b, nb, o, e, P, stream);
    else
      launch_one<RN, AN, 0, false>(md, b, nb, o, e, P, stream);
  } else {
    if (forces)
      launch_one<RN, AN, 1, true>(md, b, nb, o, e, P, stream);
    else
      launch_one<RN, AN, 1, false>(md, b, nb, o, e, P, stream);
  }
}

static ChebLaunch cheb_launch(int mode, const AtomSizes &z) {
  ChebLaunch P;
  P.natoms = z.natoms;
  P.cap = z.cap;
  P.nsf = z.nsf_max;
  P.nl = std::max(z.nl_max, 2);
  P.hs = net_stride(z.hmax);
  P.wsize = z.wsize;
  P.w_in_lds = 0;
  P.waves = 1;
  P.active = mode == 0 ? z.active : nullptr;
  P.nactive = mode == 0 ? z.nactive : nullptr;
  P.tune_waves = z.tune_waves;
  P.tune_staged = z.tune_staged;
  P.nnorm = z.nnorm;
  // the matrix-tile network: structures, one architecture, chosen by measurement
  P.tile_net = mode == 0 && z.net_uniform && z.tune_tile == 1;
  // the force table in the staged weights' words: the largest power of two
  // up to 512 entries that fits (7 words each)
  P.ftab = 0;
  for (int e = 512; e >= 64 && !P.ftab; e >>= 1)
    if (force_table_words(e) <= staged_words(P)) P.ftab = e;
  return P;
}

bool atom_cheb_shape_ok(const AtomSizes &z) {
  const ChebLaunch P = cheb_launch(0, z);
  PersistentShape ps;
  const size_t wbytes = P.wsize <= 16384 ? 4 * (size_t)staged_words(P) : 0;
  return persistent_shape(wbytes, 4 * (size_t)wave_lds_words(P), P.natoms, 4, ps, CHEB_MAX_WAVES, P.tune_waves,
                          P.tune_staged);
}

void launch_atom_cheb(int rN, int aN, int mode, bool forces, const AtomSizes &z, const ModelDev &md,
                      const BatchDev &b, const NbDev &nb, const OutDev &o, const EnvDev &e, hipStream_t stream) {
  if (z.natoms <= 0) return;
  const ChebLaunch P = cheb_launch(mode, z);
  if (rN <= 9 && aN <= 1)
    launch_cheb<9, 1>(mode, forces, md, b, nb, o, e, P, stream);
  else if (rN <= 17 && aN <= 5)
    launch_cheb<17, 5>(mode, forces, md, b, nb, o, e, P, stream);
  else if (rN <= 19 && aN <= 7)
    launch_cheb<19, 7>(mode, forces, md, b, nb, o, e, P, stream);
  else if (rN <= 24 && aN <= 8)
    launch_cheb<24, 8>(mode, forces, md, b, nb, o, e, P, stream);
  else  // rN, aN <= 16 (cheb_orders_ok)
    launch_cheb<16, 16>(mode, forces, md, b, nb, o, e, P, stream);
}

}  // namespace aenet_amd

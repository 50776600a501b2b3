// atom_common.h -- pieces shared by the per-atom kernels (atom_cheb.hip,
// atom_bp.hip): wavefront synchronisation, reduce-scatter of per-lane
// partial sums, the per-species network (forward and dE/dG backward), the
// per-wavefront LDS layout and the persistent-grid sizing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "aenet_amd_internal.h"
#include "kernels.h"
#include "kernels_common.h"

namespace aenet_amd {


__device__ __forceinline__ void activate(int f, float x, float &y, float &dy) {
  // ff_activate (src/ext/feedforward.f90:932-969)
  switch (f) {
    case 0: y = x; dy = 1.0f; break;
    case 1: y = tanhf(x); dy = 1.0f - y * y; break;
    case 2: y = 1.0f / (1.0f + expf(-x)); dy = y * (1.0f - y); break;
    case 3: { float t = tanhf(0.666666666666667f * x); y = 1.7159f * t; dy = 1.7159f * (1.0f - t * t) * 0.666666666666667f; } break;
    case 4: { float t = tanhf(0.666666666666667f * x); y = 1.7159f * t + 0.1f * x; dy = 1.7159f * (1.0f - t * t) * 0.666666666666667f + 0.1f; } break;
    default: y = 0.0f; dy = 0.0f;
  }
}

// cutoff function and its derivative: fc = (cos(pi r/Rc) + 1)/2 for
// r < Rc, else 0 (sfb_fc / sfb_fc_d1, src/ext/sfbasis.f90:719-752; sf_cut,
// src/ext/symmfunc.f90:823-838 is the same function)
__device__ __forceinline__ void cutoff(float r, float Rc, float inv_Rc, float &fc, float &dfc) {
  if (r >= Rc) {
    fc = 0.0f;
    dfc = 0.0f;
  } else {
    float sp, cp;
    sincospif(r * inv_Rc, &sp, &cp);
    fc = 0.5f * (cp + 1.0f);
    dfc = -0.5f * 3.14159265358979f * inv_Rc * sp;
  }
}

// triangular index p -> (a, b), 0 <= a < b < n, row-major
__device__ __forceinline__ void tri_decode(int p, int n, int &a, int &b) {
  float nn = (float)(2 * n - 1);
  // hardware square root (1 ulp): the estimate stays within one row of the
  // exact one (n <= 512), and the two steps below make the decode exact
  int r = (int)((nn - __builtin_amdgcn_sqrtf(nn * nn - 8.0f * (float)p)) * 0.5f);
  r = min(max(r, 0), n - 2);
  if (r > 0 && r * n - r * (r + 1) / 2 > p) --r;
  if (r < n - 2 && (r + 1) * n - (r + 1) * (r + 2) / 2 <= p) ++r;
  a = r;
  b = p - (r * n - r * (r + 1) / 2) + r + 1;
}

// triangular index p -> (a, b), 0 <= a < b < n, by diagonals: d = b - a
// outer (1 .. n-1), a inner.  Consecutive indexes of one diagonal have
// distinct a and distinct b, so the lanes of a wavefront that scatter into
// per-neighbor sums seldom hit the same address.
__device__ __forceinline__ void tri_decode_diag(int p, int n, int &a, int &b) {
  int r, c;
  tri_decode(p, n, r, c);  // row r of the row-major order = diagonal r + 1
  a = c - r - 1;
  b = c;
}

// Reduce-scatter of V per-lane values (V a power of two <= 64) over the
// wavefront: on return lane l holds the sum over all 64 lanes of
// v[l >> (6 - log2 V)] (64/V consecutive lanes hold the same feature).
// Level O (32, 16, .., 64/V) halves the values a lane holds: the lanes with
// bit O clear keep the lower half, the others the upper half, each adding
// its butterfly partner's copy (permlane swaps for O = 32, 16; DPP below);
// the remaining levels sum the copies of one feature.
template <int O, int H, class T>
__device__ __forceinline__ void rs_level(T (&v)[2 * H]) {
  if constexpr (O >= 16) {
#pragma unroll
    for (int k = 0; k < H; ++k) {
      T a = v[k], b = v[k + H];
      half_swap<O>(a, b);
      v[k] = a + b;
    }
  } else {
    const bool up = lane_id() & O;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const T lo = v[k], hi = v[k + H];
      const T send = up ? lo : hi;
      const T keep = up ? hi : lo;
      v[k] = keep + partner<O>(send);
    }
  }
}

template <int O, int V, class T>
__device__ __forceinline__ T rs_levels(T (&v)[V]) {
  if constexpr (V == 1) {
    return butterfly_sum<O>(v[0]);
  } else {
    rs_level<O, V / 2>(v);
    T (&w)[V / 2] = *reinterpret_cast<T (*)[V / 2]>(&v[0]);
    return rs_levels<O / 2, V / 2>(w);
  }
}

template <int V>
__device__ __forceinline__ double reduce_scatter(double (&v)[V]) {
  static_assert(V >= 1 && V <= 64 && (V & (V - 1)) == 0, "power of two up to 64");
  return rs_levels<32, V>(v);
}

// the same for FP32 partial sums: the first level in FP32, the rest FP64
template <int V>
__device__ __forceinline__ double reduce_scatter(float (&v)[V]) {
  static_assert(V >= 2 && V <= 64 && (V & (V - 1)) == 0, "power of two 2..64");
  rs_level<32, V / 2>(v);
  double d[V / 2];
#pragma unroll
  for (int k = 0; k < V / 2; ++k) d[k] = (double)v[k];
  return rs_levels<16, V / 2>(d);
}

// LDS of the network stage of one wavefront
struct NetLds {
  float *gp;    // [nsf] normalized descriptor
  float *coef;  // [nsf] E_scale * dy/dG' * norm scale = dE/dG (unnormalized G)
  float *d0;    // [dmax] back-propagation buffers, dmax = max(nsf, hs)
  float *d1;    // [dmax]
  float *act;   // [nl*hs] outputs of layer l at act[hs*l]
  float *dact;  // [nl*hs] activation derivatives
  int hs;       // layer stride: the widest non-input layer, rounded up to 32
};

__host__ __device__ inline int net_dmax(int nsf, int hs) { return nsf > hs ? nsf : hs; }
__host__ __device__ inline int net_lds_words(int nsf, int nl, int hs) {
  return 2 * nsf + 2 * net_dmax(nsf, hs) + 2 * nl * hs;
}
__device__ inline NetLds carve_net(float *p, int nsf, int nl, int hs) {
  NetLds n;
  n.hs = hs;
  n.gp = p;
  n.coef = n.gp + nsf;
  n.d0 = n.coef + nsf;
  n.d1 = n.d0 + net_dmax(nsf, hs);
  n.act = n.d1 + net_dmax(nsf, hs);
  n.dact = n.act + nl * hs;
  return n;
}

// layer stride for a model whose widest non-input layer has hmax nodes:
// 32, 64, then multiples of 32
__host__ __device__ inline int net_stride(int hmax) { return hmax <= 32 ? 32 : hmax <= 64 ? 64 : (hmax + 31) / 32 * 32; }

// A species record copied into a wavefront's LDS once per central atom:
// its fields (layer sizes, offsets, typespin table, ...) are then LDS reads.
// Read through the global pointer they are vector loads (the compiler does
// not take the table as read-only), each a memory round trip on the atom's
// critical path.
constexpr int SPECIES_WORDS = (int)((sizeof(SpeciesDev) + 15) / 16 * 4);

// The copy in two halves, one word of the record per lane: species_word's
// load can issue as soon as the type is known, beside the neighbor-list
// loads, and store_species writes the words to LDS once they are needed.
static_assert(sizeof(SpeciesDev) % 4 == 0 && sizeof(SpeciesDev) / 4 <= WAVE, "one word per lane");

__device__ __forceinline__ int species_word(const ModelDev &md, int t) {
  const int w = lane_id();
  return w < (int)(sizeof(SpeciesDev) / 4) ? reinterpret_cast<const int *>(md.species + t)[w] : 0;
}

__device__ __forceinline__ const SpeciesDev &store_species(int word, SpeciesDev *dst) {
  const int w = lane_id();
  if (w < (int)(sizeof(SpeciesDev) / 4)) reinterpret_cast<int *>(dst)[w] = word;
  wave_sync();
  return *dst;
}

__device__ inline const SpeciesDev &stage_species(const ModelDev &md, int t, SpeciesDev *dst) {
  return store_species(species_word(md, t), dst);
}

// Weight pointers with their address space stated: a plain pointer that
// may point to LDS or to global memory is read with FLAT loads, each waited
// for on both memory counters; these are read with DS or global loads.
typedef const __attribute__((address_space(3))) float *lds_weights;
typedef const __attribute__((address_space(1))) float *global_weights;

// First weight layer of a species' network (ff_eval,
// src/ext/feedforward.f90:586-655, layer 0 -> 1) straight from a
// descriptor spread over the lanes: lane l holds feature f (f < 0: none)
// with normalized value x.  Each lane forms its products W[f][j] x for the
// first 32 nodes, and a reduce-scatter over the wavefront sums them, leaving
// node j in lanes 2j and 2j+1; those lanes add the bias, activate and store
// the layer's outputs and derivatives in N.act / N.dact (layer 1).  Needs
// n2 = nnodes[1] <= 32.  network(.., first = 1) then goes on from layer 1.
template <class WPtr>
__device__ inline void first_layer_from_lanes(const SpeciesDev &sp, WPtr W, int f, float x, const NetLds &N) {
  const int lane = lane_id();
  const int n1 = sp.nnodes[0], n2 = sp.nnodes[1];
  const WPtr W0 = W + sp.w_off[0];
  // Branch-free, at immediate offsets: every lane reads 32 words from the
  // start of its row (row 0 for a lane without a feature, with x = 0).
  // Words j >= n2 lie past the row (the next rows, the bias column, the
  // next layer or species; the device copy of the weights is padded by 32
  // words, and the staged copy is followed by more of the workgroup's LDS);
  // their products only enter the sums of columns j >= n2, which belong to
  // no node and are never read.
  const WPtr row = W0 + (f >= 0 ? f : 0) * n2;
  const float xf = f >= 0 ? x : 0.0f;
  float v[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) v[j] = row[j] * xf;
  const float z = rs_levels<32, 32>(v);
  const int node = lane >> 1;
  if ((lane & 1) == 0 && node < n2) {
    float y, dy;
    activate(sp.fa[1], z + W0[n1 * n2 + node], y, dy);  // bias column
    N.act[N.hs + node] = y;
    N.dact[N.hs + node] = dy;
  }
  wave_sync();
}

// Feed-forward network of one species (hidden layers wider than 64 nodes
// need WIDE) on the normalized descriptor in
// N.gp: ff_eval (src/ext/feedforward.f90:586-655), lanes = nodes of a
// layer.  With first = 1 the first layer's outputs are already in N.act
// (first_layer_from_lanes) and the forward pass starts from them.  With grad, also ff_deriv (:661-757) and the chain through the
// normalization and the energy scale: N.coef[f] = dE_i/dG_f.  Returns the
// network output y (all lanes).  W is lds_weights or global_weights.  With
// scale_in_coef, N.coef[f] holds the normalization scale of feature f on
// entry (left there by the caller's normalization) instead of it being read
// from global memory again.
template <bool WIDE = false, class WPtr>
__device__ inline float network(const ModelDev &md, const SpeciesDev &sp, WPtr W, const NetLds &N, bool grad,
                                bool scale_in_coef = false, int first = 0) {
  const int lane = lane_id();
  const int nl = sp.nlayers;
  const float *in = first == 0 ? N.gp : N.act + first * N.hs;
  for (int l = first; l + 1 < nl; ++l) {
    const int n1 = sp.nnodes[l], n2 = sp.nnodes[l + 1];
    const WPtr Wl = W + sp.w_off[l];
    // lanes = (node, part): with NP the node count rounded up to a power of
    // two (at most 64), the 64 / NP parts share the inputs and a butterfly
    // over the part bits of the lane sums them; with WIDE, a layer wider
    // than the wavefront takes its nodes 64 at a time (one part, every
    // input).  Without WIDE (every layer at most 64 nodes) the node loop is
    // one pass, and the kernels that run common networks keep their
    // registers for the rest.
    // (shifts: NP is not known at compile time, and a division by it
    // would be a full integer division)
    int lg = 0;
    while ((1 << lg) < n2 && lg < 6) ++lg;
    const int NP = 1 << lg;
    const int part = lane >> lg;
    const int per = (n1 * NP + WAVE - 1) >> 6;
    const int c0 = min(n1, part * per), c1 = min(n1, c0 + per);
    for (int nb = 0; nb < (WIDE ? n2 : 1); nb += NP) {
      const int node = nb + (lane & (NP - 1));
      float z0 = 0.0f, z1 = 0.0f, z2 = 0.0f, z3 = 0.0f;
      if (node < n2) {
        int c = c0;
        for (; c + 3 < c1; c += 4) {
          z0 += Wl[c * n2 + node] * in[c];
          z1 += Wl[(c + 1) * n2 + node] * in[c + 1];
          z2 += Wl[(c + 2) * n2 + node] * in[c + 2];
          z3 += Wl[(c + 3) * n2 + node] * in[c + 3];
        }
        for (; c < c1; ++c) z0 += Wl[c * n2 + node] * in[c];
      }
      float z = (z0 + z1) + (z2 + z3);
      for (int o = WAVE / 2; o >= NP; o >>= 1) z += __shfl_xor(z, o, WAVE);
      if (part == 0 && node < n2) {
        float y, dy;
        activate(sp.fa[l + 1], z + Wl[n1 * n2 + node], y, dy);  // bias column
        N.act[(l + 1) * N.hs + node] = y;
        N.dact[(l + 1) * N.hs + node] = dy;
      }
    }
    wave_sync();
    in = N.act + (l + 1) * N.hs;
  }
  const float yout = N.act[(nl - 1) * N.hs];
  if (!grad) return yout;
  float *dcur = N.d0;
  float *dnext = N.d1;
  if (lane == 0) dcur[0] = N.dact[(nl - 1) * N.hs];
  wave_sync();
  for (int l = nl - 2; l >= 0; --l) {
    const int n1 = sp.nnodes[l], n2 = sp.nnodes[l + 1];
    const WPtr Wl = W + sp.w_off[l];
    if (n1 <= WAVE / 2) {
      // lanes = (input c, part) as in the forward pass, parts share the nodes
      int lg = 0;
      while ((1 << lg) < n1) ++lg;
      const int NP = 1 << lg;
      const int c = lane & (NP - 1), part = lane >> lg;
      const int per = (n2 * NP + WAVE - 1) >> 6;
      const int h0 = min(n2, part * per), h1 = min(n2, h0 + per);
      float s0 = 0.0f, s1 = 0.0f;
      if (c < n1) {
        const WPtr row = Wl + c * n2;
        int hh = h0;
        for (; hh + 1 < h1; hh += 2) {
          s0 += row[hh] * dcur[hh];
          s1 += row[hh + 1] * dcur[hh + 1];
        }
        if (hh < h1) s0 += row[hh] * dcur[hh];
      }
      float s = s0 + s1;
      for (int o = WAVE / 2; o >= NP; o >>= 1) s += __shfl_xor(s, o, WAVE);
      if (part == 0 && c < n1) dnext[c] = l > 0 ? s * N.dact[l * N.hs + c] : s;
    } else {
      for (int c = lane; c < n1; c += WAVE) {
        const WPtr row = Wl + c * n2;
        float s0 = 0.0f, s1 = 0.0f;
        int hh = 0;
        for (; hh + 1 < n2; hh += 2) {
          s0 += row[hh] * dcur[hh];
          s1 += row[hh + 1] * dcur[hh + 1];
        }
        if (hh < n2) s0 += row[hh] * dcur[hh];
        float s = s0 + s1;
        dnext[c] = l > 0 ? s * N.dact[l * N.hs + c] : s;
      }
    }
    wave_sync();
    float *t = dcur;
    dcur = dnext;
    dnext = t;
  }
  const float Es = (float)sp.E_scale;
  for (int f = lane; f < sp.nsf; f += WAVE)
    N.coef[f] = Es * dcur[f] * (scale_in_coef ? N.coef[f] : (float)md.norm_scale[sp.norm_off + f]);
  wave_sync();
  return yout;
}

// E_i = E_scale*y + E_shift + E_atom (src/aenet.f90:468-469), FP64
__device__ __forceinline__ double atom_energy(const SpeciesDev &sp, float y) {
  double E = sp.E_scale * (double)y + sp.E_shift;
  return E + sp.E_atom;
}

// Persistent launch shape for an atom kernel whose workgroups may hold
// wbytes of shared data (the network weights) plus wave_bytes per
// wavefront.  simd_waves is the kernel's register-limited wavefronts per
// SIMD.  The weights are staged in LDS only when that does not lower the
// number of wavefronts resident per CU (otherwise they are read through
// the caches); as many wavefronts per workgroup as LDS allows (up to
// ATOM_MAX_WAVES) and as many workgroups as fit on the device at once.
// Returns false when not even one wavefront fits.
struct PersistentShape {
  int waves, grid;
  size_t lds;
  bool w_in_lds;
};

inline int resident_waves(size_t wbytes, size_t wave_bytes, int simd_waves, int max_waves, int &waves_per_wg) {
  const DeviceLimits &dl = device_limits();
  if (wbytes + wave_bytes > dl.lds_per_block) return 0;
  waves_per_wg = (int)std::min<size_t>(max_waves, (dl.lds_per_block - wbytes) / wave_bytes);
  const size_t wg_bytes = wbytes + (size_t)waves_per_wg * wave_bytes;
  const int by_lds = (int)(dl.lds_per_cu / wg_bytes) * waves_per_wg;
  return std::min(by_lds, 4 * simd_waves);
}

inline bool persistent_shape(size_t wbytes, size_t wave_bytes, int natoms, int simd_waves, PersistentShape &ps,
                             int max_waves = ATOM_MAX_WAVES, int force_waves = 0, int force_staged = -1) {
  const DeviceLimits &dl = device_limits();
  if (force_waves > 0 || force_staged >= 0) {
    // a measured choice: waves per workgroup and weight staging as given,
    // as many workgroups per CU as LDS and registers allow
    const bool staged = force_staged >= 0 ? force_staged != 0 : wbytes > 0;
    if (staged && (wbytes == 0 || wbytes > 64 * 1024)) return false;
    const int waves = force_waves > 0 ? force_waves : ATOM_MAX_WAVES;
    if (waves > max_waves) return false;
    const size_t wg_bytes = (staged ? wbytes : 0) + (size_t)waves * wave_bytes;
    if (wg_bytes > dl.lds_per_block) return false;
    const int per_cu = std::min((int)(dl.lds_per_cu / wg_bytes), std::max(1, 4 * simd_waves / waves));
    if (per_cu < 1) return false;
    ps.w_in_lds = staged;
    ps.waves = waves;
    ps.lds = wg_bytes;
    ps.grid = std::max(1, std::min((natoms + waves - 1) / waves, dl.cus * per_cu));
    return true;
  }
  int wg_plain = 0, wg_staged = 0;
  // without staging, workgroups of at most ATOM_MAX_WAVES suffice
  const int r_plain = resident_waves(0, wave_bytes, simd_waves, std::min(max_waves, ATOM_MAX_WAVES), wg_plain);
  const int r_staged =
      (wbytes > 0 && wbytes <= 64 * 1024) ? resident_waves(wbytes, wave_bytes, simd_waves, max_waves, wg_staged) : 0;
  if (r_plain == 0 && r_staged == 0) return false;
  ps.w_in_lds = r_staged > 0 && r_staged >= r_plain;
  ps.waves = ps.w_in_lds ? wg_staged : wg_plain;
  ps.lds = (ps.w_in_lds ? wbytes : 0) + (size_t)ps.waves * wave_bytes;
  const int per_cu = std::max(1, (ps.w_in_lds ? r_staged : r_plain) / ps.waves);
  ps.grid = std::min((natoms + ps.waves - 1) / ps.waves, dl.cus * per_cu);
  ps.grid = std::max(ps.grid, 1);
  return true;
}

}  // namespace aenet_amd

// kernels_common.h -- wavefront helpers and the ordered FP64 neighbor
// search shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "aenet_amd_internal.h"

#define WAVE 64

namespace aenet_amd {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ unsigned long long wave_ballot(bool p) { return __ballot(p); }

// number of set bits of mask below this lane
__device__ __forceinline__ int mask_prefix(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ int popc64(unsigned long long m) { return __popcll(m); }

template <class T>
__device__ __forceinline__ T uniform(T v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// orders the LDS accesses of one wavefront (lanes exchange data through
// LDS between phases of one atom)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane exchanges without the LDS crossbar (gfx950).
// half_swap<O>(a, b), O = 32 or 16: v_permlane32_swap / v_permlane16_swap
// exchange the upper O lanes of every 2*O-lane group of a with the lower O
// lanes of b, so that afterwards a + b holds, in the lower half of each
// group, the sum of a over both halves and, in the upper half, the sum of b
// over both halves (one step of a reduce-scatter; with a = b = x the pair
// sum of x in every lane).
// partner<O>(x), O = 8, 4, 2, 1: x of the butterfly partner within the
// lane's row by DPP -- l ^ 8 (row_ror:8), the mirror l ^ 7 in the group of
// 8 (row_half_mirror), l ^ 2 and l ^ 1 (quad_perm).  The partner differs
// from l in bit O and agrees with it in the bits above O.
__device__ __forceinline__ void half_swap_u32(unsigned &a, unsigned &b, int o) {
  if (o == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = r[0];
    b = r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    a = r[0];
    b = r[1];
  }
}

template <int O>
__device__ __forceinline__ void half_swap(float &a, float &b) {
  static_assert(O == 32 || O == 16, "permlane swaps");
  unsigned x = __float_as_uint(a), y = __float_as_uint(b);
  half_swap_u32(x, y, O);
  a = __uint_as_float(x);
  b = __uint_as_float(y);
}

template <int O>
__device__ __forceinline__ void half_swap(double &a, double &b) {
  static_assert(O == 32 || O == 16, "permlane swaps");
  const unsigned long long x = (unsigned long long)__double_as_longlong(a);
  const unsigned long long y = (unsigned long long)__double_as_longlong(b);
  unsigned xl = (unsigned)x, xh = (unsigned)(x >> 32), yl = (unsigned)y, yh = (unsigned)(y >> 32);
  half_swap_u32(xl, yl, O);
  half_swap_u32(xh, yh, O);
  a = __longlong_as_double((long long)(((unsigned long long)xh << 32) | xl));
  b = __longlong_as_double((long long)(((unsigned long long)yh << 32) | yl));
}

template <int O>
__device__ __forceinline__ void half_swap(long long &a, long long &b) {
  double x = __longlong_as_double(a), y = __longlong_as_double(b);
  half_swap<O>(x, y);  // moves the bits only
  a = __double_as_longlong(x);
  b = __double_as_longlong(y);
}

template <int O>
__device__ __forceinline__ unsigned partner_u32(unsigned x) {
  static_assert(O == 8 || O == 4 || O == 2 || O == 1, "DPP within a row");
  constexpr int ctrl = O == 8 ? 0x128 : O == 4 ? 0x141 : O == 2 ? 0x4E : 0xB1;
  // every lane's source is inside its row for these patterns, so no lane
  // keeps the old value and none needs to be set first
  return (unsigned)__builtin_amdgcn_mov_dpp((int)x, ctrl, 0xf, 0xf, false);
}

template <int O>
__device__ __forceinline__ float partner(float x) {
  return __uint_as_float(partner_u32<O>(__float_as_uint(x)));
}

template <int O>
__device__ __forceinline__ double partner(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = partner_u32<O>((unsigned)u), hi = partner_u32<O>((unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int O>
__device__ __forceinline__ long long partner(long long x) {
  return __double_as_longlong(partner<O>(__longlong_as_double(x)));  // moves the bits only
}

// sum over the lanes of the butterfly levels O, O/2, .., 1 (every lane
// ends with the same value: each step adds two equal operand pairs)
template <int O, class T>
__device__ __forceinline__ T butterfly_sum(T v) {
  if constexpr (O >= 16) {
    T a = v, b = v;
    half_swap<O>(a, b);
    v = a + b;
  } else if constexpr (O >= 1) {
    v += partner<O>(v);
  }
  if constexpr (O > 1) return butterfly_sum<O / 2>(v);
  return v;
}

__device__ __forceinline__ float wave_sum(float v) { return butterfly_sum<32>(v); }

__device__ __forceinline__ double wave_sum(double v) { return butterfly_sum<32>(v); }

__device__ __forceinline__ long long wave_sum(long long v) { return butterfly_sum<32>(v); }

// The wavefront sums of three per-lane integers at once, returned in lane 0
// (other lanes: partial sums): two reduce-scatter levels leave the three
// sums' partials in lanes 0-15, 32-47 and 16-31, a butterfly over each
// 16-lane row finishes them, and lane 0 reads the other two rows' results.
// Integer addition, so the result is the same as three wave_sum calls.
__device__ __forceinline__ void wave_sum3_lane0(long long (&v)[3]) {
  long long a = v[0], b = v[1], c = v[2], d = 0;
  half_swap<32>(a, b);
  a += b;  // lanes 0-31: v[0] over lanes l, l+32; lanes 32-63: v[1]
  half_swap<32>(c, d);
  c += d;  // lanes 0-31: v[2]; lanes 32-63: zero
  half_swap<16>(a, c);
  a += c;  // rows: v[0], v[2], v[1], zero
  a = butterfly_sum<8>(a);
  const unsigned long long u = (unsigned long long)a;
  const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
  auto row = [&](int l) {
    return (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)hi, l) << 32) |
                       (unsigned)__builtin_amdgcn_readlane((int)lo, l));
  };
  v[0] = a;
  v[1] = row(32);
  v[2] = row(16);
}

// Forces of structures are summed as 64-bit fixed-point integers in units
// of 2^-32 eV/A: integer addition is associative, so the sum of the
// contributions of all atoms to F_j does not depend on the order in which
// the atomics arrive, and the forces are bitwise the same from run to run.
// The force array is reinterpreted during the sum (zero bits are zero in
// both forms) and turned into FP64 by k_struct_energy.  A contribution of
// 2^18 eV/A or more, or not finite, is not added (the sum over up to 512
// neighbors stays below 2^27 eV/A, inside the 2^31 eV/A of the format), and
// the atom is marked instead, which makes its force NaN.
constexpr double FORCE_FIX_SCALE = 4294967296.0;        // 2^32
constexpr double FORCE_FIX_UNIT = 1.0 / 4294967296.0;   // 2^-32
constexpr double FORCE_FIX_MAX = 262144.0;              // 2^18

// the fixed-point form of a contribution c; false when out of range.
// |c| 2^32 < 2^50: adding 1.5 2^52 rounds it to an integer (to nearest,
// ties to even) held in the low mantissa bits, and subtracting the bits of
// 1.5 2^52 leaves that integer in two's complement
__device__ __forceinline__ bool force_fix(const double (&c)[3], long long (&v)[3]) {
  if (!(fabs(c[0]) < FORCE_FIX_MAX && fabs(c[1]) < FORCE_FIX_MAX && fabs(c[2]) < FORCE_FIX_MAX)) return false;
  constexpr double magic = 6755399441055744.0;  // 1.5 2^52
#pragma unroll
  for (int x = 0; x < 3; ++x)
    v[x] = __double_as_longlong(c[x] * FORCE_FIX_SCALE + magic) - __double_as_longlong(magic);
  return true;
}

__device__ __forceinline__ void force_add_fixed(double *F, int atom, const long long (&v)[3]) {
  unsigned long long *Fi = reinterpret_cast<unsigned long long *>(F) + 3 * (size_t)atom;
#pragma unroll
  for (int x = 0; x < 3; ++x) atomicAdd(Fi + x, (unsigned long long)v[x]);
}

// The scatter of one central atom i over its n neighbors: per neighbor q
// the FP32 gradient dE_i/dR_j (grad(q) -> double[3]) goes to F_j with a
// minus sign, and F_i receives minus the integer sum of exactly those
// fixed-point values, so the forces of a structure sum to zero exactly in
// the fixed-point units (translation invariance).  A contribution out of
// range marks j and i.
template <class Grad>
__device__ inline void scatter_forces(double *F, int *fmark, int i, int n, const int *J, Grad grad) {
  long long own[3] = {0, 0, 0};
  bool bad = false;
  for (int q = lane_id(); q < n; q += WAVE) {
    double c[3];
    grad(q, c);
    const double m[3] = {-c[0], -c[1], -c[2]};
    long long v[3];
    const int j = J[q];
    if (force_fix(m, v)) {
      force_add_fixed(F, j, v);
      own[0] -= v[0];
      own[1] -= v[1];
      own[2] -= v[2];
    } else {
      fmark[j] = 1;
      bad = true;
    }
  }
  wave_sum3_lane0(own);
  bad = wave_ballot(bad) != 0;
  if (lane_id() == 0) {
    if (bad)
      fmark[i] = 1;
    else
      force_add_fixed(F, i, own);
  }
}

// Workgroup aggregation of the force scatter.  The atoms of a workgroup
// round are spatial neighbors and share most of their neighbors, so their
// contributions are first summed per target atom in an LDS table (open
// addressing, 2^k entries: key words, then three 64-bit sums per entry)
// and each target then takes one set of global atomics.  A contribution
// that finds no free entry goes to global memory directly.  The sums are
// integers, so the grouping changes no result.
struct ForceTable {
  unsigned long long *val;  // [3 * entries]
  int *key;                 // [entries], -1: free
  int entries;              // power of two
};

__host__ __device__ inline int force_table_words(int entries) { return 7 * entries; }

__device__ inline ForceTable force_table_at(float *p, int entries) {
  ForceTable t;
  t.val = reinterpret_cast<unsigned long long *>(p);  // 8-byte aligned base
  t.key = reinterpret_cast<int *>(t.val + 3 * (size_t)entries);
  t.entries = entries;
  return t;
}

__device__ inline void force_table_add(const ForceTable &t, double *F, int atom, const long long (&v)[3]) {
  unsigned s = ((unsigned)atom * 2654435761u) & (unsigned)(t.entries - 1);
  for (int probe = 0; probe < t.entries; ++probe) {
    const int prev = atomicCAS(&t.key[s], -1, atom);
    if (prev == -1 || prev == atom) {
#pragma unroll
      for (int x = 0; x < 3; ++x) atomicAdd(&t.val[3 * s + x], (unsigned long long)v[x]);
      return;
    }
    s = (s + 1) & (unsigned)(t.entries - 1);
  }
  force_add_fixed(F, atom, v);  // table full
}

// scatter_forces into the workgroup's table instead of global memory
template <class Grad>
__device__ inline void scatter_forces_table(const ForceTable &t, double *F, int *fmark, int i, int n, const int *J,
                                            Grad grad) {
  long long own[3] = {0, 0, 0};
  bool bad = false;
  for (int q = lane_id(); q < n; q += WAVE) {
    double c[3];
    grad(q, c);
    const double m[3] = {-c[0], -c[1], -c[2]};
    long long v[3];
    const int j = J[q];
    if (force_fix(m, v)) {
      force_table_add(t, F, j, v);
      own[0] -= v[0];
      own[1] -= v[1];
      own[2] -= v[2];
    } else {
      fmark[j] = 1;
      bad = true;
    }
  }
  wave_sum3_lane0(own);
  bad = wave_ballot(bad) != 0;
  if (lane_id() == 0) {
    if (bad)
      fmark[i] = 1;
    else
      force_table_add(t, F, i, own);
  }
}

// inclusive prefix sum over the wavefront
__device__ __forceinline__ int wave_scan(int x) {
  for (int o = 1; o < WAVE; o <<= 1) {
    int y = __shfl_up(x, o, WAVE);
    if (lane_id() >= o) x += y;
  }
  return x;
}

// L*v for a column-major 3x3 lattice, summed in the reference's order
// (matmul of src/ext/lclist.f90:621-651) without contraction.
__device__ __forceinline__ void matvec_exact(const double *L, const double *v, double *out) {
#pragma clang fp contract(off)
  for (int r = 0; r < 3; ++r) {
    double acc = L[r] * v[0] + L[3 + r] * v[1];
    out[r] = acc + L[6 + r] * v[2];
  }
}

__device__ __forceinline__ double dist2_exact(const double *c) {
#pragma clang fp contract(off)
  double d2 = c[0] * c[0] + c[1] * c[1];
  return d2 + c[2] * c[2];
}

__device__ __forceinline__ void add_T_exact(const double *s, const int *T, double *out) {
#pragma clang fp contract(off)
  out[0] = s[0] + (double)T[0];
  out[1] = s[1] + (double)T[1];
  out[2] = s[2] + (double)T[2];
}

// LDS work space of one search, sized by the batch's largest number of
// (cell, image) combinations of a home cell (kc)
struct SearchLds {
  int kc;
  int *kept;    // [2*kc] (cell, image rank) kept for this atom
  int *gk;      // [kc] group: first kept entry
  int *gn;      // [kc] group: number of images
  int *gpos;    // [kc] group: first atom position in cell order
  int *gitems;  // [kc+1] group: prefix of (atoms x images)
};

__host__ __device__ inline int search_lds_ints(int kc) { return 6 * kc + 1; }

__device__ inline SearchLds carve_search(int *p, int kc) {
  SearchLds W;
  W.kc = kc;
  W.kept = p;
  W.gk = p + 2 * kc;
  W.gn = W.gk + kc;
  W.gpos = W.gn + kc;
  W.gitems = W.gpos + kc;
  return W;
}

// Ordered neighbor search of atom g (global index) for one wavefront.  The
// entries come out in the order lcl_nbdist_cart produces them
// (src/ext/lclist.f90:520-677): periodic self images first, then the
// candidates in linked-cell order (home cell, then the stencil cells; atoms
// in descending index within a cell), each candidate with its home image
// followed by the lattice translations in the reference's order.  Only the
// (cell, image) combinations of the home cell whose bounding sphere reaches
// this atom's cutoff sphere are visited; the test that decides membership
// is the reference's FP64 expression, so the list is the reference's.
//
// Sink::put(slot, j, coo2, cart, dist) receives every entry and must check
// slot against its own capacity.  Returns the number of entries.
// The periodic images of the central atom g within the cutoff, first in
// the reference's list (src/ext/lclist.f90:599-624), through sink slots
// 0..nself-1.  Returns their number.
template <class Sink>
__device__ int self_images(const BatchDev &b, const StructDev *S, int g, const double *L, const double *si,
                           const double *coo1, Sink &sink) {
  const int lane = lane_id();
  const int *Tv = b.Tvec + 3 * S->T0;
  const int nself = S->nself;
  int count = 0;
  for (int base = 0; base < nself; base += WAVE) {
    int q = base + lane;
    if (q < nself) {
      int t = b.self_t[S->self0 + q];
      double c2[3], coo2[3], cart[3];
      add_T_exact(si, Tv + 3 * (t - 1), c2);
      matvec_exact(L, c2, coo2);
      cart[0] = coo2[0] - coo1[0];
      cart[1] = coo2[1] - coo1[1];
      cart[2] = coo2[2] - coo1[2];
      sink.put(count + q, g, coo2, cart, b.self_d[S->self0 + q]);
    }
    count += min(WAVE, nself - base);
  }
  return count;
}

template <class Sink>
__device__ int search_neighbors(const BatchDev &b, int g, Sink &sink, SearchLds &W) {
  const int lane = lane_id();
  const int sid = uniform(b.atom_struct[g]);
  const StructDev *S = b.st + sid;
  double L[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) L[k] = S->L[k];
  const double si[3] = {b.s[3 * (size_t)g], b.s[3 * (size_t)g + 1], b.s[3 * (size_t)g + 2]};
  double coo1[3];
  matvec_exact(L, si, coo1);
  const double Rc2 = S->Rc2;
  const int *Tv = b.Tvec + 3 * S->T0;
  // (2) periodic images of the central atom
  int count = self_images(b, S, g, L, si, coo1, sink);
  // (3) candidates: keep the home cell's (cell, image) combinations that
  // can reach this atom
  const int h = b.cellof[g];
  const int K = b.combo_count[h];
  const int *combos = b.combos + 2 * (size_t)h * b.combo_cap;
  const double lim = S->rc_list + S->cell_rad + 1.0e-6;
  const double lim2 = lim * lim;
  int nk = 0;
  for (int base = 0; base < K; base += WAVE) {
    int q = base + lane;
    bool keep = false;
    int c = 0, t = 0;
    if (q < K) {
      c = combos[2 * q];
      t = combos[2 * q + 1];
      double ctr[3] = {b.cell_ctr[3 * c], b.cell_ctr[3 * c + 1], b.cell_ctr[3 * c + 2]};
      if (t > 0) {
        const int *T = Tv + 3 * (t - 1);
        for (int r = 0; r < 3; ++r) ctr[r] += L[r] * T[0] + L[3 + r] * T[1] + L[6 + r] * T[2];
      }
      double dx = ctr[0] - coo1[0], dy = ctr[1] - coo1[1], dz = ctr[2] - coo1[2];
      keep = (dx * dx + dy * dy + dz * dz <= lim2) && (b.cell_start[c + 1] > b.cell_start[c]);
    }
    unsigned long long m = wave_ballot(keep);
    int pos = nk + mask_prefix(m);
    if (keep && pos < W.kc) {
      W.kept[2 * pos] = c;
      W.kept[2 * pos + 1] = t;
    }
    nk += popc64(m);
  }
  if (nk > W.kc) {
    if (lane == 0) atomicOr(b.flags, AENET_AMD_FLAG_KEPT_OVERFLOW);
    nk = W.kc;
  }
  __syncthreads();
  // groups = runs of equal cell; items = atoms x images of each group
  int ng = 0;
  for (int base = 0; base < nk; base += WAVE) {
    int q = base + lane;
    bool start = q < nk && (q == 0 || W.kept[2 * q] != W.kept[2 * (q - 1)]);
    unsigned long long m = wave_ballot(start);
    if (start) {
      int k = ng + mask_prefix(m);
      int c = W.kept[2 * q];
      W.gk[k] = q;
      W.gpos[k] = b.cell_start[c];
    }
    ng += popc64(m);
  }
  __syncthreads();
  int carry = 0;
  for (int base = 0; base < ng; base += WAVE) {
    int k = base + lane;
    int items = 0;
    if (k < ng) {
      int q0 = W.gk[k];
      int q1 = k + 1 < ng ? W.gk[k + 1] : nk;
      int c = W.kept[2 * q0];
      W.gn[k] = q1 - q0;
      items = (b.cell_start[c + 1] - b.cell_start[c]) * (q1 - q0);
    }
    int incl = wave_scan(items);
    if (k < ng) W.gitems[k] = carry + incl - items;
    carry += __shfl(incl, WAVE - 1, WAVE);
  }
  if (lane == 0) W.gitems[ng] = carry;
  __syncthreads();
  const int total = carry;
  for (int base = 0; base < total; base += WAVE) {
    int q = base + lane;
    bool pass = false;
    int j = 0;
    double coo2[3], cart[3], d2 = 0.0;
    if (q < total) {
      // group of item q: last k with gitems[k] <= q
      int lo = 0, hi = ng - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (W.gitems[mid] <= q)
          lo = mid;
        else
          hi = mid - 1;
      }
      int off = q - W.gitems[lo];
      int gn = W.gn[lo];
      int a = off / gn;
      int tt = off - a * gn;
      int pos = W.gpos[lo] + a;
      j = b.cell_atoms[pos];
      int t = W.kept[2 * (W.gk[lo] + tt) + 1];
      if (j != g) {
        const double sj[3] = {b.sorted_s[3 * (size_t)pos], b.sorted_s[3 * (size_t)pos + 1],
                              b.sorted_s[3 * (size_t)pos + 2]};
        if (t == 0) {
          matvec_exact(L, sj, coo2);
        } else {
          double c2[3];
          add_T_exact(sj, Tv + 3 * (t - 1), c2);
          matvec_exact(L, c2, coo2);
        }
        cart[0] = coo2[0] - coo1[0];
        cart[1] = coo2[1] - coo1[1];
        cart[2] = coo2[2] - coo1[2];
        d2 = dist2_exact(cart);
        pass = d2 <= Rc2;
      }
    }
    unsigned long long msk = wave_ballot(pass);
    if (pass) sink.put(count + mask_prefix(msk), j, coo2, cart, sqrt(d2));
    count += popc64(msk);
  }
  return count;
}

}  // namespace aenet_amd

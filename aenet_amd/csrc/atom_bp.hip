// atom_bp.hip -- per-atom kernel for Behler-Parrinello symmetry functions
// (G1-G5 of src/ext/symmfunc.f90, the "Behler2011" descriptor).
//
// Persistent workgroups; each wavefront evaluates one central atom at a
// time:
//   1. neighbors  : compact list of k_search (structures) or the caller's
//                   list (per-atom entry points); neighbors are bucketed by
//                   their type local to the central atom's descriptor, so
//                   that every block of functions -- one per neighbor type
//                   (radial) and per unordered neighbor-type pair
//                   (angular), sf_build_index (src/ext/symmfunc.f90:284-305)
//                   -- is uniform across the wavefront
//   2. descriptors: sf_fingerprint (src/ext/symmfunc.f90:311-431).  Radial
//                   functions: lanes = neighbors of the block's type.
//                   Angular functions: lanes = neighbor pairs (j, k) of the
//                   block's type pair; pairs that cannot contribute (r_jk
//                   beyond the block's cutoff, for blocks without G5) are
//                   dropped by a ballot-compacted queue first, so the
//                   function loop runs on full wavefronts.  Per-function
//                   partial sums in registers, reduce-scatter across lanes.
//                   G4 = 2 [(1+lambda cos)/2]^zeta exp(-eta S) fc fc fc is
//                   one exp2 per function: 2^(zeta log2(arg) - eta S log2 e)
//   3. normalize  : stp_normalize (src/sfsetup.f90:985-1035) in FP64
//   4. network    : shared with the Chebyshev kernel (atom_common.h)
//   5. forces     : the same traversal again with dE/dG: per pair / triplet
//                   the function sums collapse to dE/dr_ij, dE/dr_ik,
//                   dE/dr_jk and dE/dcos, which are turned into per-neighbor
//                   gradients in LDS, then fixed-point (64-bit integer) atomics
//                   into F
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "aenet_amd_internal.h"
#include "atom_common.h"
#include "kernels.h"
#include "kernels_common.h"

namespace aenet_amd {

namespace {

constexpr float LOG2E = 1.44269504088896341f;
constexpr int RCH = 16;   // radial functions per register chunk
constexpr int ACH = 32;   // angular functions per register chunk
constexpr int QCAP = 128; // triplet queue entries per wavefront
constexpr int BP_SIMD_WAVES = 4;  // wavefronts per SIMD the kernel's registers allow (<= 128 VGPRs)
constexpr int MT = AENET_AMD_MAX_TYPES;

}  // namespace

// The symmetry-function tables do not change during a launch: read through
// constant-address-space pointers they become scalar loads, and the
// per-function branches (cutoff, kind, lambda) wave-uniform.
typedef const __attribute__((address_space(4))) BPFunc *bp_funcs;
typedef const __attribute__((address_space(4))) BPDev *bp_tables;
typedef const __attribute__((address_space(4))) BPAngK *bp_angk;
typedef const __attribute__((address_space(4))) BPAngGrid *bp_grid;

// 2^x and log2(x) as the hardware instructions (1 ulp).  The library forms
// add scaling for subnormal inputs and results only: the log2 arguments here
// are clamped at 1e-30, and a term below 2^-126 vanishes beside the sums it
// joins.
__device__ __forceinline__ float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float log2_hw(float x) { return __builtin_amdgcn_logf(x); }

__device__ __forceinline__ BPFunc load_fn(bp_funcs fn, int k) {
  BPFunc f;
  f.kind = fn[k].kind;
  f.Rc = fn[k].Rc;
  f.p1 = fn[k].p1;
  f.p2 = fn[k].p2;
  f.p3 = fn[k].p3;
  return f;
}

struct BPWave {
  float2 *ct;   // [cap] cutoff fc and dfc/dr of each neighbor at the angular table radius
  float *R;     // [3*cap] R_j - R_i
  float *r;     // [cap]
  int *t;       // [cap] local type of the neighbor (0-based; -1: not in the descriptor); the
                // first words of fs, used only before the gradients are cleared
  int *bidx;    // [cap] neighbor slots grouped by local type
  float *fs;    // [3*cap] per-neighbor gradients dE_i/dR_j
  int *queue;   // [QCAP] packed (j slot, k slot) of pending triplets
  int *bstart;  // [MT+1] bucket starts in bidx
  NetLds net;
};

struct BPLaunch {
  int natoms;    // central atoms (structures) or environments
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
};

__host__ __device__ inline int bp_wave_words(const BPLaunch &P) {
  return (10 * P.cap + QCAP + 16 + net_lds_words(P.nsf, P.nl, P.hs) + 3) / 4 * 4;
}

__device__ inline BPWave bp_carve(float *base, const BPLaunch &P) {
  const int cap = P.cap;
  BPWave w;
  w.ct = reinterpret_cast<float2 *>(base);  // base is 16-byte aligned
  w.R = base + 2 * cap;
  w.r = w.R + 3 * cap;
  w.bidx = reinterpret_cast<int *>(w.r + cap);
  w.fs = reinterpret_cast<float *>(w.bidx + cap);
  w.t = reinterpret_cast<int *>(w.fs);
  w.queue = reinterpret_cast<int *>(w.fs + 3 * cap);
  w.bstart = w.queue + QCAP;
  w.net = carve_net(reinterpret_cast<float *>(w.bstart + 16), P.nsf, P.nl, P.hs);
  return w;
}

// value and radial derivative of one G1/G2/G3 function at distance r
// (sf_G1_ij, sf_G2_ij, sf_G3_ij, src/ext/symmfunc.f90:464-580)
__device__ __forceinline__ void radial_fn(const BPFunc &f, float r, float fc, float dfc, float &v, float &dv) {
  if (f.kind == 2) {
    const float a = r - f.p1;
    const float e = exp2_hw(-f.p2 * LOG2E * a * a);
    v = e * fc;
    dv = e * (dfc - 2.0f * f.p2 * a * fc);
  } else if (f.kind == 3) {
    float s, c;
    sincosf(f.p1 * r, &s, &c);
    v = c * fc;
    dv = c * dfc - f.p1 * s * fc;
  } else {
    v = fc;
    dv = dfc;
  }
}

// geometry of one triplet (i; j, k): unit vectors, distances, cos(theta)
// clamped to [-1, 1] (src/ext/symmfunc.f90:409-411)
struct Trip {
  float uj[3], uk[3], d[3];  // d = R_k - R_j
  float rj, rk, rjk, cosv;
  float ij, ik;  // 1 / rj, 1 / rk
};

__device__ __forceinline__ Trip make_trip(const float *R, const float *r, int ja, int kb) {
  Trip T;
  T.rj = r[ja];
  T.rk = r[kb];
  // hardware reciprocal and square root (1 ulp; FP32 geometry feeding FP32
  // function values)
  const float ij = __builtin_amdgcn_rcpf(T.rj), ik = __builtin_amdgcn_rcpf(T.rk);
  T.ij = ij;
  T.ik = ik;
  float dot = 0.0f, d2 = 0.0f;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const float a = R[3 * ja + x], b = R[3 * kb + x];
    T.uj[x] = a * ij;
    T.uk[x] = b * ik;
    T.d[x] = b - a;
    d2 += T.d[x] * T.d[x];
  }
  dot = T.uj[0] * T.uk[0] + T.uj[1] * T.uk[1] + T.uj[2] * T.uk[2];
  T.cosv = fminf(1.0f, fmaxf(-1.0f, dot));
  T.rjk = __builtin_amdgcn_sqrtf(d2);
  return T;
}

// Angular traversal of one block: every unordered neighbor pair (j, k) of
// local types (t2, t3) that can contribute goes through `op` with full
// wavefronts.  Candidates are decoded from the bucket pair, filtered by
// r_ij, r_ik < rc and (unless keep_all: blocks with G5) r_jk < rc, where
// rc is the block's largest cutoff, compacted into the LDS queue, and
// handed out 64 at a time.
template <class Op>
__device__ __forceinline__ void for_triplets(const BPWave &S, int b2, int m2, int b3, int m3, bool same, bool keep_all,
                                             float rc2, Op &op) {
  const int lane = lane_id();
  const int ncand = same ? m2 * (m2 - 1) / 2 : m2 * m3;
  const float inv_m2 = m2 > 0 ? 1.0f / (float)m2 : 0.0f;
  const float inv_m3 = m3 > 0 ? 1.0f / (float)m3 : 0.0f;
  // one call site of op (it is inlined once): groups of 64 while
  // candidates remain, then the rest
  int qn = 0;
  for (int base = 0;;) {
    const int p = base + lane;
    bool ok = false;
    int packed = 0;
    if (p < ncand) {
      // Diagonal orders: a run of consecutive candidates has distinct j and
      // distinct k, so the force pass's per-neighbor atomics seldom meet at
      // one address within a wavefront.
      int a, c;
      if (same) {
        tri_decode_diag(p, m2, a, c);
      } else {
        // p = d m2 + a, c = (a + d) mod m3: for each a, d over [0, m3)
        // takes c over [0, m3) once
        int d = (int)(((float)p + 0.5f) * inv_m2);
        a = p - d * m2;
        if (a < 0) {
          --d;
          a += m2;
        } else if (a >= m2) {
          ++d;
          a -= m2;
        }
        const int sum = a + d;
        int q = (int)(((float)sum + 0.5f) * inv_m3);
        c = sum - q * m3;
        if (c < 0) {
          c += m3;
        } else if (c >= m3) {
          c -= m3;
        }
      }
      const int ja = S.bidx[b2 + a], kb = S.bidx[b3 + c];
      const float rj = S.r[ja], rk = S.r[kb];
      if (!(rj * rj < rc2 && rk * rk < rc2)) {
        ok = false;  // fc(r_ij) fc(r_ik) = 0 for every function of the block
      } else if (keep_all) {
        ok = true;
      } else {
        const float dx = S.R[3 * kb] - S.R[3 * ja], dy = S.R[3 * kb + 1] - S.R[3 * ja + 1],
                    dz = S.R[3 * kb + 2] - S.R[3 * ja + 2];
        ok = dx * dx + dy * dy + dz * dz < rc2;
      }
      packed = ja | (kb << 16);
    }
    if (base < ncand) {
      const unsigned long long m = wave_ballot(ok);
      if (ok) S.queue[qn + mask_prefix(m)] = packed;
      qn += popc64(m);
      base += WAVE;
      wave_sync();
    }
    const bool more = base < ncand;
    if (qn >= WAVE || (!more && qn > 0)) {
      const int e = lane < qn ? S.queue[lane] : 0;
      op(e & 0xffff, e >> 16, lane < qn);
      wave_sync();
      if (lane < qn - WAVE) S.queue[lane] = S.queue[WAVE + lane];
      qn = max(qn - WAVE, 0);
      wave_sync();
    }
    if (!more && qn == 0) break;
  }
}

// Wraps an op of for_triplets and writes the triplets it is handed, in
// that order, to list[0 ..]: entries beyond cap are counted, not written.
// for_triplets hands an op a prefix of the lanes as valid, with every lane
// active.
template <class Op>
struct RecordTriplets {
  Op &op;
  int *list;
  int cap;
  int n;  // triplets handed over so far
  __device__ __forceinline__ void operator()(int ja, int kb, bool valid) {
    const int i = n + lane_id();
    if (valid && i < cap) list[i] = ja | (kb << 16);
    n += popc64(wave_ballot(valid));
    op(ja, kb, valid);
  }
};

// the n triplets of a list written by RecordTriplets through op, in the
// groups of 64 that for_triplets formed; the next group's load is issued
// before the current group's op
template <class Op>
__device__ __forceinline__ void replay_triplets(const int *list, int n, Op &op) {
  const int lane = lane_id();
  int e = lane < n ? list[lane] : 0;
  for (int base = 0; base < n; base += WAVE) {
    const int i = base + lane;
    const int next = i + WAVE < n ? list[i + WAVE] : 0;
    op(e & 0xffff, e >> 16, i < n);
    e = next;
  }
}

// cutoffs of the three legs of a triplet for one cutoff radius
struct TripCut {
  float rc;
  float cj, dcj, ck, dck, cjk, dcjk;
};

// per-neighbor cutoff table at one radius (the row's first angular cutoff)
struct CutTab {
  const float2 *ct;
  float rc, inv;
};

// the legs j and k come from the table when rc is its radius (the same
// cutoff() of the same r, computed once per neighbor instead of per triplet)
__device__ __forceinline__ void trip_cut(const Trip &T, float rc, TripCut &c, const CutTab &tab, int ja, int kb) {
  c.rc = rc;
  float inv;
  if (rc == tab.rc) {
    inv = tab.inv;
    const float2 a = tab.ct[ja], b2 = tab.ct[kb];
    c.cj = a.x;
    c.dcj = a.y;
    c.ck = b2.x;
    c.dck = b2.y;
  } else {
    inv = 1.0f / rc;
    cutoff(T.rj, rc, inv, c.cj, c.dcj);
    cutoff(T.rk, rc, inv, c.ck, c.dck);
  }
  cutoff(T.rjk, rc, inv, c.cjk, c.dcjk);
}

// energy pass of an angular chunk: G_f += 2 arg^zeta exp(-eta S) C; each
// lane sums its own triplets of a block (a few tens) in FP32, the
// reduce-scatter over the lanes continues in FP64
struct AngValueOp {
  const float *R, *r;
  CutTab tab;
  bp_funcs fn;
  int nf;  // functions in this chunk
  float acc[ACH];
  __device__ __forceinline__ void operator()(int ja, int kb, bool valid) {
    if (!valid) return;
    const Trip T = make_trip(R, r, ja, kb);
    const float argp = 0.5f * (1.0f + T.cosv), argm = 0.5f * (1.0f - T.cosv);
    const float Lp = log2_hw(fmaxf(argp, 1.0e-30f)), Lm = log2_hw(fmaxf(argm, 1.0e-30f));
    const float S2 = T.rj * T.rj + T.rk * T.rk, S3 = S2 + T.rjk * T.rjk;
    TripCut c;
    c.rc = -1.0f;
#pragma unroll
    for (int k = 0; k < ACH; ++k) {
      if (k < nf) {
        const BPFunc f = load_fn(fn, k);
        if (f.Rc != c.rc) trip_cut(T, f.Rc, c, tab, ja, kb);
        const float lam = f.p1, zeta = f.p2, eta = f.p3;
        float arg, L;
        if (lam == 1.0f) {
          arg = argp;
          L = Lp;
        } else if (lam == -1.0f) {
          arg = argm;
          L = Lm;
        } else {
          arg = 0.5f * (1.0f + lam * T.cosv);
          L = log2_hw(fmaxf(arg, 1.0e-30f));
        }
        const bool g4 = f.kind == 4;
        const float P = 2.0f * exp2_hw((zeta - 1.0f) * L - eta * LOG2E * (g4 ? S3 : S2));
        const float C = g4 ? c.cj * c.ck * c.cjk : c.cj * c.ck;
        acc[k] += P * arg * C;
      }
    }
  }
};

// force pass of an angular block: dE/d(r_ij, r_ik, r_jk, cos) from every
// function of the block, then the per-neighbor gradients
struct AngForceOp {
  const float *R, *r;
  CutTab tab;
  float *fs;
  bp_funcs fn;
  const float *coef;  // dE/dG of the block's functions
  int nf;
  bool one_seg;       // every function of the block has one cutoff and one kind
  __device__ __forceinline__ void operator()(int ja, int kb, bool valid) {
    if (!valid) return;
    const Trip T = make_trip(R, r, ja, kb);
    const float argp = 0.5f * (1.0f + T.cosv), argm = 0.5f * (1.0f - T.cosv);
    const float Lp = log2_hw(fmaxf(argp, 1.0e-30f)), Lm = log2_hw(fmaxf(argm, 1.0e-30f));
    const float S2 = T.rj * T.rj + T.rk * T.rk, S3 = S2 + T.rjk * T.rjk;
    float dRj = 0.0f, dRk = 0.0f, dRjk = 0.0f, dcos = 0.0f;
    // sums over a segment of functions with the same kind and cutoff
    float s_cos = 0.0f, s_E = 0.0f, s_eta = 0.0f;
    TripCut c;
    c.rc = -1.0f;
    bool seg_g4 = true;
    auto flush = [&]() {
      const float cjk = seg_g4 ? c.cjk : 1.0f;
      const float C = c.cj * c.ck * cjk;
      dcos += s_cos * C;
      dRj += s_E * (c.dcj * c.ck * cjk) + s_eta * T.rj * C;
      dRk += s_E * (c.cj * c.dck * cjk) + s_eta * T.rk * C;
      if (seg_g4) dRjk += s_E * (c.cj * c.ck * c.dcjk) + s_eta * T.rjk * C;
      s_cos = s_E = s_eta = 0.0f;
    };
    auto term = [&](const BPFunc &f, int k, bool g4) {
      const float lam = f.p1, zeta = f.p2, eta = f.p3;
      float arg, L;
      if (lam == 1.0f) {
        arg = argp;
        L = Lp;
      } else if (lam == -1.0f) {
        arg = argm;
        L = Lm;
      } else {
        arg = 0.5f * (1.0f + lam * T.cosv);
        L = log2_hw(fmaxf(arg, 1.0e-30f));
      }
      const float P = 2.0f * exp2_hw((zeta - 1.0f) * L - eta * LOG2E * (g4 ? S3 : S2));
      const float w = coef[k];
      const float wE = w * P * arg;
      s_cos += w * (0.5f * zeta * lam) * P;
      s_E += wE;
      s_eta += wE * (-2.0f * eta);
    };
    if (one_seg) {
      // one segment: cutoffs once, no per-function segment checks
      const BPFunc f0 = load_fn(fn, 0);
      seg_g4 = f0.kind == 4;
      if (nf > 0) trip_cut(T, f0.Rc, c, tab, ja, kb);
#pragma unroll 4
      for (int k = 0; k < nf; ++k) term(load_fn(fn, k), k, seg_g4);
    } else {
      for (int k = 0; k < nf; ++k) {
        const BPFunc f = load_fn(fn, k);
        const bool g4 = f.kind == 4;
        if (f.Rc != c.rc || g4 != seg_g4) {
          if (c.rc > 0.0f) flush();
          if (f.Rc != c.rc) trip_cut(T, f.Rc, c, tab, ja, kb);
          seg_g4 = g4;
        }
        term(f, k, g4);
      }
    }
    if (c.rc > 0.0f) flush();
    const float inv_rjk = T.rjk > 1.0e-8f ? __builtin_amdgcn_rcpf(T.rjk) : 1.0f;
    const float cj = dcos * T.ij, ck = dcos * T.ik;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const float e = dRjk * T.d[x] * inv_rjk;
      const float gj = dRj * T.uj[x] + cj * (T.uk[x] - T.cosv * T.uj[x]) - e;
      const float gk = dRk * T.uk[x] + ck * (T.uj[x] - T.cosv * T.uk[x]) + e;
      atomicAdd(&fs[3 * ja + x], gj);
      atomicAdd(&fs[3 * kb + x], gk);
    }
  }
};

// Fast path of an angular block with one cutoff, one kind and lambda = +-1
// (BPAngK constants, sign-sorted: the first np functions have lambda = +1).
// Per triplet log2 of both arguments once; per function
//   P = 2^(A L + B S)      (G = 2 arg P C)
// and no per-function branches on the parameters.
struct AngFastGeom {
  float Lp, Lm, argp, argm, S;
  TripCut c;
  float C;  // cutoff product of the segment's kind
};

__device__ __forceinline__ void fast_geom(const Trip &T, bool g4, float rc, const CutTab &tab, int ja, int kb,
                                          AngFastGeom &q) {
  q.argp = 0.5f * (1.0f + T.cosv);
  q.argm = 0.5f * (1.0f - T.cosv);
  q.Lp = log2_hw(fmaxf(q.argp, 1.0e-30f));
  q.Lm = log2_hw(fmaxf(q.argm, 1.0e-30f));
  const float S2 = T.rj * T.rj + T.rk * T.rk;
  q.S = g4 ? S2 + T.rjk * T.rjk : S2;
  trip_cut(T, rc, q.c, tab, ja, kb);
  q.C = g4 ? q.c.cj * q.c.ck * q.c.cjk : q.c.cj * q.c.ck;
}

// energy pass, fast path: acc[k] += P_k (2 arg C) for function k of the
// chunk (sign-sorted order)
struct AngValueFastOp {
  const float *R, *r;
  CutTab tab;
  bp_angk K;  // the chunk's constants
  int nf, np;  // functions in this chunk, of which lambda = +1 first
  float rc;
  bool g4;
  float acc[ACH];
  __device__ __forceinline__ void operator()(int ja, int kb, bool valid) {
    if (!valid) return;
    const Trip T = make_trip(R, r, ja, kb);
    AngFastGeom q;
    fast_geom(T, g4, rc, tab, ja, kb, q);
    const float aCp = 2.0f * q.argp * q.C, aCm = 2.0f * q.argm * q.C;
#pragma unroll
    for (int k = 0; k < ACH; ++k) {
      if (k < nf) {
        const bool pos = k < np;
        const float x = fmaf(K[k].A, pos ? q.Lp : q.Lm, K[k].B * q.S);
        acc[k] = fmaf(exp2_hw(x), pos ? aCp : aCm, acc[k]);
      }
    }
  }
};

// force pass, fast path: the sums of the generic force op from the
// constants, the block's lambda = +1 functions then its lambda = -1 ones
struct AngForceFastOp {
  const float *R, *r;
  CutTab tab;
  float *fs;
  bp_angk K;
  const float *coef;  // dE/dG of the block's functions (file order)
  int nf, np;
  float rc;
  bool g4;
  __device__ __forceinline__ void operator()(int ja, int kb, bool valid) {
    if (!valid) return;
    const Trip T = make_trip(R, r, ja, kb);
    AngFastGeom q;
    fast_geom(T, g4, rc, tab, ja, kb, q);
    // s_cos = sum w K1 P, s_E = sum w P arg, s_eta = sum w K2 P arg (P
    // without its factor 2, applied below)
    float sc = 0.0f, sep = 0.0f, sem = 0.0f, shp = 0.0f, shm = 0.0f;
#pragma unroll 4
    for (int k = 0; k < np; ++k) {
      const float t = coef[K[k].feat] * exp2_hw(fmaf(K[k].A, q.Lp, K[k].B * q.S));
      sc = fmaf(K[k].K1, t, sc);
      sep += t;
      shp = fmaf(K[k].K2, t, shp);
    }
#pragma unroll 4
    for (int k = np; k < nf; ++k) {
      const float t = coef[K[k].feat] * exp2_hw(fmaf(K[k].A, q.Lm, K[k].B * q.S));
      sc = fmaf(K[k].K1, t, sc);
      sem += t;
      shm = fmaf(K[k].K2, t, shm);
    }
    const float s_cos = 2.0f * sc;
    const float s_E = 2.0f * (q.argp * sep + q.argm * sem);
    const float s_eta = 2.0f * (q.argp * shp + q.argm * shm);
    const TripCut &c = q.c;
    const float cjk = g4 ? c.cjk : 1.0f;
    const float C = q.C;
    const float dcos = s_cos * C;
    const float dRj = s_E * (c.dcj * c.ck * cjk) + s_eta * T.rj * C;
    const float dRk = s_E * (c.cj * c.dck * cjk) + s_eta * T.rk * C;
    const float dRjk = g4 ? s_E * (c.cj * c.ck * c.dcjk) + s_eta * T.rjk * C : 0.0f;
    const float inv_rjk = T.rjk > 1.0e-8f ? __builtin_amdgcn_rcpf(T.rjk) : 1.0f;
    const float cj = dcos * T.ij, ck = dcos * T.ik;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const float e = dRjk * T.d[x] * inv_rjk;
      const float gj = dRj * T.uj[x] + cj * (T.uk[x] - T.cosv * T.uj[x]) - e;
      const float gk = dRk * T.uk[x] + ck * (T.uj[x] - T.cosv * T.uk[x]) + e;
      atomicAdd(&fs[3 * ja + x], gj);
      atomicAdd(&fs[3 * kb + x], gk);
    }
  }
};

// energy pass, grid path (BPAngGrid): per triplet E_ie = 2^(B_ie S) C for
// the ne etas and Z_iz = 2^(zeta_iz log2(arg) + 1) for the nz (lambda,
// zeta) pairs; slot iz BP_GRID_NE + ie adds Z_iz E_ie (slots past ne add
// zero)
static_assert(BP_GRID_NZ * BP_GRID_NE == ACH, "one register chunk of grid slots");
// The grid's constants, read once per block (not per triplet round)
struct GridConst {
  int nz, ne, nzp;
  float zeta[BP_GRID_NZ];
  float K1[BP_GRID_NZ];
  float B[BP_GRID_NE];
  float K2[BP_GRID_NE];
  __device__ __forceinline__ void load(bp_grid G) {
    nz = G->nz;
    ne = G->ne;
    nzp = G->nzp;
#pragma unroll
    for (int iz = 0; iz < BP_GRID_NZ; ++iz) {
      zeta[iz] = G->zeta[iz];
      K1[iz] = G->K1[iz];
    }
#pragma unroll
    for (int ie = 0; ie < BP_GRID_NE; ++ie) {
      B[ie] = G->B[ie];
      K2[ie] = G->K2[ie];
    }
  }
};

struct AngValueGridOp {
  const float *R, *r;
  CutTab tab;
  GridConst K;
  float rc;
  bool g4;
  float acc[ACH];
  __device__ __forceinline__ void operator()(int ja, int kb, bool valid) {
    if (!valid) return;
    const Trip T = make_trip(R, r, ja, kb);
    AngFastGeom q;
    fast_geom(T, g4, rc, tab, ja, kb, q);
    float E[BP_GRID_NE];
#pragma unroll
    for (int ie = 0; ie < BP_GRID_NE; ++ie) E[ie] = ie < K.ne ? exp2_hw(K.B[ie] * q.S) * q.C : 0.0f;
#pragma unroll
    for (int iz = 0; iz < BP_GRID_NZ; ++iz) {
      if (iz < K.nz) {
        const float Z = exp2_hw(fmaf(K.zeta[iz], iz < K.nzp ? q.Lp : q.Lm, 1.0f));
#pragma unroll
        for (int ie = 0; ie < BP_GRID_NE; ++ie) acc[iz * BP_GRID_NE + ie] = fmaf(Z, E[ie], acc[iz * BP_GRID_NE + ie]);
      }
    }
  }
};

// force pass, grid path: the sums of the fast path's force op by pairs,
//   u0_iz = sum_ie c_s E_ie,  u2_iz = sum_ie c_s K2_ie E_ie  (c_s = dE/dG of
//   slot s, E_ie = 2^(B_ie S) without the cutoffs),
// then with Zm_iz = 2^((zeta_iz - 1) log2(arg)): s_cos = 2 sum K1 Zm u0, and
// per sign of lambda sum Zm u0 and sum Zm u2
struct AngForceGridOp {
  const float *R, *r;
  CutTab tab;
  float *fs;
  GridConst K;
  float cs[ACH];  // dE/dG per slot (0: no function)
  float rc;
  bool g4;
  __device__ __forceinline__ void operator()(int ja, int kb, bool valid) {
    if (!valid) return;
    const Trip T = make_trip(R, r, ja, kb);
    AngFastGeom q;
    fast_geom(T, g4, rc, tab, ja, kb, q);
    float E[BP_GRID_NE], EK[BP_GRID_NE];
#pragma unroll
    for (int ie = 0; ie < BP_GRID_NE; ++ie) {
      E[ie] = ie < K.ne ? exp2_hw(K.B[ie] * q.S) : 0.0f;
      EK[ie] = K.K2[ie] * E[ie];
    }
    float sc = 0.0f, sep = 0.0f, sem = 0.0f, shp = 0.0f, shm = 0.0f;
#pragma unroll
    for (int iz = 0; iz < BP_GRID_NZ; ++iz) {
      if (iz < K.nz) {
        float u0 = 0.0f, u2 = 0.0f;
#pragma unroll
        for (int ie = 0; ie < BP_GRID_NE; ++ie) {
          u0 = fmaf(cs[iz * BP_GRID_NE + ie], E[ie], u0);
          u2 = fmaf(cs[iz * BP_GRID_NE + ie], EK[ie], u2);
        }
        const bool pos = iz < K.nzp;
        const float Zm = exp2_hw((K.zeta[iz] - 1.0f) * (pos ? q.Lp : q.Lm));
        const float t0 = Zm * u0, t2 = Zm * u2;
        sc = fmaf(K.K1[iz], t0, sc);
        if (pos) {
          sep += t0;
          shp += t2;
        } else {
          sem += t0;
          shm += t2;
        }
      }
    }
    const float s_cos = 2.0f * sc;
    const float s_E = 2.0f * (q.argp * sep + q.argm * sem);
    const float s_eta = 2.0f * (q.argp * shp + q.argm * shm);
    const TripCut &c = q.c;
    const float cjk = g4 ? c.cjk : 1.0f;
    const float C = q.C;
    const float dcos = s_cos * C;
    const float dRj = s_E * (c.dcj * c.ck * cjk) + s_eta * T.rj * C;
    const float dRk = s_E * (c.cj * c.dck * cjk) + s_eta * T.rk * C;
    const float dRjk = g4 ? s_E * (c.cj * c.ck * c.dcjk) + s_eta * T.rjk * C : 0.0f;
    const float inv_rjk = T.rjk > 1.0e-8f ? __builtin_amdgcn_rcpf(T.rjk) : 1.0f;
    const float cj = dcos * T.ij, ck = dcos * T.ik;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const float e = dRjk * T.d[x] * inv_rjk;
      const float gj = dRj * T.uj[x] + cj * (T.uk[x] - T.cosv * T.uj[x]) - e;
      const float gk = dRk * T.uk[x] + ck * (T.uj[x] - T.cosv * T.uk[x]) + e;
      atomicAdd(&fs[3 * ja + x], gj);
      atomicAdd(&fs[3 * kb + x], gk);
    }
  }
};

// MODE 0: structures (compact lists, fixed-point atomic force scatter)
// MODE 1: explicit environments (per-neighbor forces out)
// Registers are held to what BP_SIMD_WAVES wavefronts per SIMD allow.
template <int MODE, bool FORCES, bool WIDE>
__global__ __launch_bounds__(64 * ATOM_MAX_WAVES) __attribute__((amdgpu_waves_per_eu(BP_SIMD_WAVES))) void k_atom_bp(
    ModelDev md, BatchDev b, NbDev nbl, OutDev o, EnvDev env, BPLaunch P) {
  extern __shared__ float lds[];
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  const int nw = P.waves;
  if (P.w_in_lds) {
    for (int k = threadIdx.x; k < P.wsize; k += 64 * nw) lds[k] = md.W[k];
    __syncthreads();
  }
  const int wsz = P.w_in_lds ? (P.wsize + 3) / 4 * 4 : 0;
  BPWave S = bp_carve(lds + wsz + wave * bp_wave_words(P), P);
  const bp_tables bpt = (bp_tables)md.bp;
  const bp_funcs bpf = (bp_funcs)md.bpf;
  const int nt = bpt->ntypes;
  // this wavefront's triplet list (forces of structures only): per angular
  // block a count word (-1: the block's triplets are not listed), then its
  // triplets; blocks are listed while the count word fits
  const int tslot = blockIdx.x * nw + wave;
  int *const tlist = MODE == 0 && FORCES && b.trip && tslot < b.trip_slots ? b.trip + (size_t)tslot * b.trip_cap : nullptr;

  const int count = P.active ? *P.nactive : P.natoms;
  for (int k = blockIdx.x * nw + wave; k < count; k += gridDim.x * nw) {
    const int g = P.active ? P.active[k] : k;
    wave_sync();
    // ------------------------------------------------------------ 1
    int ti, n;
    if (MODE == 0) {
      ti = b.types[g];
      n = min(nbl.count[g], P.cap);
    } else {
      ti = env.env_type[g];
      n = env.env_start[g + 1] - env.env_start[g];
      if (n > P.cap) {
        if (lane == 0) atomicOr(b.flags, AENET_AMD_FLAG_NB_OVERFLOW);
        n = P.cap;
      }
    }
    ti = uniform(ti);
    n = uniform(n);
    if (ti < 1 || ti > md.ntypes) {
      if (lane == 0) atomicOr(b.flags, AENET_AMD_FLAG_BAD_TYPE);
      continue;
    }
    const SpeciesDev &sp = md.species[ti - 1];
    bool bad_nb = false;
    if (MODE == 0) {
      const float4 *R4 = reinterpret_cast<const float4 *>(nbl.R) + (size_t)g * nbl.cap;
      const int *J = nbl.j + (size_t)g * nbl.cap;
      for (int q = lane; q < n; q += WAVE) {
        float4 v = R4[q];
        int j = J[q];
        int tj = b.types[j];
        int lt = (tj >= 1 && tj <= md.ntypes) ? sp.bp_ltype[tj - 1] - 1 : -1;
        bad_nb |= lt < 0;
        S.R[3 * q] = v.x;
        S.R[3 * q + 1] = v.y;
        S.R[3 * q + 2] = v.z;
        S.r[q] = v.w;
        S.t[q] = lt;
      }
    } else {
      const int e0 = env.env_start[g];
      const double ci[3] = {env.coo_i[3 * g], env.coo_i[3 * g + 1], env.coo_i[3 * g + 2]};
      for (int q = lane; q < n; q += WAVE) {
        double c[3];
        for (int k = 0; k < 3; ++k) c[k] = env.coo_j[3 * (size_t)(e0 + q) + k] - ci[k];
        int tj = env.type_j[e0 + q];
        int lt = (tj >= 1 && tj <= md.ntypes) ? sp.bp_ltype[tj - 1] - 1 : -1;
        bad_nb |= lt < 0;
        S.R[3 * q] = (float)c[0];
        S.R[3 * q + 1] = (float)c[1];
        S.R[3 * q + 2] = (float)c[2];
        S.r[q] = (float)sqrt(dist2_exact(c));
        S.t[q] = lt;
      }
    }
    if (wave_ballot(bad_nb) != 0ull && lane == 0) atomicOr(b.flags, AENET_AMD_FLAG_BAD_TYPE);
    wave_sync();
    // buckets by local type
    {
      int cnt = 0;
      for (int t = 0; t < nt; ++t) {
        if (lane == 0) S.bstart[t] = cnt;
        for (int base = 0; base < n; base += WAVE) {
          const int q = base + lane;
          const bool in = q < n && S.t[q] == t;
          const unsigned long long m = wave_ballot(in);
          if (in) S.bidx[cnt + mask_prefix(m)] = q;
          cnt += popc64(m);
        }
      }
      if (lane == 0) S.bstart[nt] = cnt;
    }
    wave_sync();
    const int row = sp.bp_row;
    // the local types (the first words of fs) are done with: clear the
    // gradients, and tabulate each neighbor's cutoff at the radius of the
    // row's first angular function
    CutTab tab;
    tab.ct = S.ct;
    tab.rc = -1.0f;
    tab.inv = 0.0f;
    for (int t2 = 0; t2 < nt && tab.rc < 0.0f; ++t2)
      for (int t3 = t2; t3 < nt && tab.rc < 0.0f; ++t3) {
        const int key3 = (row * MT + t2) * MT + t3;
        if (bpt->ang_n[key3] > 0) tab.rc = bpf[bpt->ang_p0[key3]].Rc;
      }
    if (tab.rc > 0.0f) tab.inv = 1.0f / tab.rc;
    for (int q = lane; q < n; q += WAVE) {
      S.fs[3 * q] = 0.0f;
      S.fs[3 * q + 1] = 0.0f;
      S.fs[3 * q + 2] = 0.0f;
      float fc = 0.0f, dfc = 0.0f;
      if (tab.rc > 0.0f) cutoff(S.r[q], tab.rc, tab.inv, fc, dfc);
      S.ct[q] = make_float2(fc, dfc);
    }
    wave_sync();

    // ------------------------------------------------------------ 2 + 3
    // radial blocks
    for (int t2 = 0; t2 < nt; ++t2) {
      const int key2 = row * MT + t2;
      const int nf = bpt->rad_n[key2], off = bpt->rad_off[key2];
      const bp_funcs fn = bpf + bpt->rad_p0[key2];
      const int bs = S.bstart[t2], m = S.bstart[t2 + 1] - bs;
      for (int c0 = 0; c0 < nf; c0 += RCH) {
        float acc[RCH];
#pragma unroll
        for (int k = 0; k < RCH; ++k) acc[k] = 0.0f;
        for (int q = lane; q < m; q += WAVE) {
          const float r = S.r[S.bidx[bs + q]];
          float rc = -1.0f, fc = 0.0f, dfc = 0.0f;
#pragma unroll
          for (int k = 0; k < RCH; ++k) {
            if (c0 + k < nf) {
              const BPFunc f = load_fn(fn, c0 + k);
              if (f.Rc != rc) {
                rc = f.Rc;
                cutoff(r, rc, 1.0f / rc, fc, dfc);
              }
              float v, dv;
              radial_fn(f, r, fc, dfc, v, dv);
              acc[k] += v;
            }
          }
        }
        const double G = reduce_scatter<RCH>(acc);
        const int idx = lane >> 2;
        if ((lane & 3) == 0 && c0 + idx < nf) {
          const int ft = off + c0 + idx;
          const double Gs = md.norm_scale[sp.norm_off + ft] * (G - md.norm_avg[sp.norm_off + ft]);
          S.net.gp[ft] = (float)Gs;
          if (o.G) o.G[(size_t)g * o.G_stride + ft] = o.G_scaled ? Gs : G;
        }
      }
    }
    // angular blocks, unordered type pairs t2 <= t3; with a triplet list,
    // the first chunk of a block records the triplets it is handed and the
    // block's later chunks (and the force pass) replay them
    int toff = 0;  // words of the list in use
    for (int t2 = 0; t2 < nt; ++t2) {
      for (int t3 = t2; t3 < nt; ++t3) {
        const int key3 = (row * MT + t2) * MT + t3;
        const int nf = bpt->ang_n[key3], off = bpt->ang_off[key3];
        if (nf == 0) continue;
        const bp_funcs fn = bpf + bpt->ang_p0[key3];
        const int b2 = S.bstart[t2], m2 = S.bstart[t2 + 1] - b2;
        const int b3 = S.bstart[t3], m3 = S.bstart[t3 + 1] - b3;
        const float rc = bpt->ang_rcmax[key3];
        int tn = -1;  // triplets of this block in the list
        const bool listed = tlist && toff < b.trip_cap;
        int *const tl = listed ? tlist + toff + 1 : nullptr;
        const int tcap = listed ? b.trip_cap - toff - 1 : 0;
        // the chunk's triplets through op: from the list, or found (and
        // recorded by the first chunk while there is room)
        auto traverse = [&](auto &op, int c0) {
          if (tn >= 0) {
            replay_triplets(tl, tn, op);
          } else {
            RecordTriplets<std::remove_reference_t<decltype(op)>> rec{op, tl, c0 == 0 ? tcap : 0, 0};
            for_triplets(S, b2, m2, b3, m3, t2 == t3, bpt->ang_g5[key3] != 0, rc * rc, rec);
            if (listed && c0 == 0 && rec.n <= tcap) tn = rec.n;
          }
        };
        const bool fast = bpt->ang_fast[key3] != 0;
        const bp_angk K = (bp_angk)md.bpk + bpt->ang_p0[key3];
        const int gidx = bpt->ang_grid[key3];
        if (gidx > 0) {
          // one chunk: the grid's slots
          AngValueGridOp op;
          op.R = S.R;
          op.r = S.r;
          op.tab = tab;
          const bp_grid Gr = (bp_grid)md.bpg + (gidx - 1);
          op.K.load(Gr);
          op.rc = fn[0].Rc;
          op.g4 = fn[0].kind == 4;
#pragma unroll
          for (int k = 0; k < ACH; ++k) op.acc[k] = 0.0f;
          traverse(op, 0);
          const double G = reduce_scatter<ACH>(op.acc);
          const int slot = lane >> 1;
          const int fb = Gr->feat[slot];
          if ((lane & 1) == 0 && fb >= 0) {
            const int ft = off + fb;
            const double Gs = md.norm_scale[sp.norm_off + ft] * (G - md.norm_avg[sp.norm_off + ft]);
            S.net.gp[ft] = (float)Gs;
            if (o.G) o.G[(size_t)g * o.G_stride + ft] = o.G_scaled ? Gs : G;
          }
        }
        for (int c0 = 0; c0 < (gidx > 0 ? 0 : nf); c0 += ACH) {
          double G;
          if (fast) {
            AngValueFastOp op;
            op.R = S.R;
            op.r = S.r;
            op.tab = tab;
            op.K = K + c0;
            op.nf = min(ACH, nf - c0);
            op.np = min(op.nf, max(0, bpt->ang_np[key3] - c0));
            op.rc = fn[0].Rc;
            op.g4 = fn[0].kind == 4;
#pragma unroll
            for (int k = 0; k < ACH; ++k) op.acc[k] = 0.0f;
            traverse(op, c0);
            G = reduce_scatter<ACH>(op.acc);
          } else {
            AngValueOp op;
            op.R = S.R;
            op.r = S.r;
            op.tab = tab;
            op.fn = fn + c0;
            op.nf = min(ACH, nf - c0);
#pragma unroll
            for (int k = 0; k < ACH; ++k) op.acc[k] = 0.0f;
            traverse(op, c0);
            G = reduce_scatter<ACH>(op.acc);
          }
          const int idx = lane >> 1;
          if ((lane & 1) == 0 && c0 + idx < nf) {
            // (the fast path's functions are sign-sorted)
            const int ft = off + (fast ? K[c0 + idx].feat : c0 + idx);
            const double Gs = md.norm_scale[sp.norm_off + ft] * (G - md.norm_avg[sp.norm_off + ft]);
            S.net.gp[ft] = (float)Gs;
            if (o.G) o.G[(size_t)g * o.G_stride + ft] = o.G_scaled ? Gs : G;
          }
        }
        if (listed) {
          if (lane == 0) tlist[toff] = tn;
          toff += 1 + max(tn, 0);
        }
      }
    }
    wave_sync();
    if (o.desc_only) continue;

    // ------------------------------------------------------------ 4
    float yout;
    if (MODE == 1 && o.coef_unit > 0) {
      // derivatives of one descriptor feature instead of the energy
      const int fu = g % o.coef_unit;
      for (int f = lane; f < sp.nsf; f += WAVE) S.net.coef[f] = f == fu ? 1.0f : 0.0f;
      wave_sync();
      yout = 0.0f;
    } else {
      yout = P.w_in_lds ? network<WIDE>(md, sp, (lds_weights)lds, S.net, FORCES)
                        : network<WIDE>(md, sp, (global_weights)md.W, S.net, FORCES);
    }
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
    const float *coef = S.net.coef;
    for (int t2 = 0; t2 < nt; ++t2) {
      const int key2 = row * MT + t2;
      const int nf = bpt->rad_n[key2], off = bpt->rad_off[key2];
      if (nf == 0) continue;
      const bp_funcs fn = bpf + bpt->rad_p0[key2];
      const int bs = S.bstart[t2], m = S.bstart[t2 + 1] - bs;
      for (int q = lane; q < m; q += WAVE) {
        const int nq = S.bidx[bs + q];
        const float r = S.r[nq];
        float rc = -1.0f, fc = 0.0f, dfc = 0.0f, s = 0.0f;
        for (int k = 0; k < nf; ++k) {
          const BPFunc f = load_fn(fn, k);
          if (f.Rc != rc) {
            rc = f.Rc;
            cutoff(r, rc, 1.0f / rc, fc, dfc);
          }
          float v, dv;
          radial_fn(f, r, fc, dfc, v, dv);
          s += coef[off + k] * dv;
        }
        // each neighbor sits in exactly one bucket: one writer per slot
        const float sr = s / r;
        S.fs[3 * nq] += sr * S.R[3 * nq];
        S.fs[3 * nq + 1] += sr * S.R[3 * nq + 1];
        S.fs[3 * nq + 2] += sr * S.R[3 * nq + 2];
      }
    }
    wave_sync();
    // the value pass's list entries were written by other lanes of this
    // wavefront
    if (tlist) __threadfence_block();
    toff = 0;
    for (int t2 = 0; t2 < nt; ++t2) {
      for (int t3 = t2; t3 < nt; ++t3) {
        const int key3 = (row * MT + t2) * MT + t3;
        const int nf = bpt->ang_n[key3], off = bpt->ang_off[key3];
        if (nf == 0) continue;
        const int b2 = S.bstart[t2], m2 = S.bstart[t2 + 1] - b2;
        const int b3 = S.bstart[t3], m3 = S.bstart[t3 + 1] - b3;
        const float rc = bpt->ang_rcmax[key3];
        int tn = -1;
        if (tlist && toff < b.trip_cap) {
          tn = uniform(tlist[toff]);
          toff += 1;
        }
        const int *const tl = tn >= 0 ? tlist + toff : nullptr;
        if (tn > 0) toff += tn;
        auto traverse = [&](auto &op) {
          if (tn >= 0)
            replay_triplets(tl, tn, op);
          else
            for_triplets(S, b2, m2, b3, m3, t2 == t3, bpt->ang_g5[key3] != 0, rc * rc, op);
        };
        const bp_funcs fn = bpf + bpt->ang_p0[key3];
        if (bpt->ang_grid[key3] > 0) {
          AngForceGridOp op;
          op.R = S.R;
          op.r = S.r;
          op.tab = tab;
          op.fs = S.fs;
          const bp_grid Gr = (bp_grid)md.bpg + (bpt->ang_grid[key3] - 1);
          op.K.load(Gr);
#pragma unroll
          for (int s2 = 0; s2 < ACH; ++s2) {
            const int fb = Gr->feat[s2];
            op.cs[s2] = fb >= 0 ? coef[off + fb] : 0.0f;
          }
          op.rc = fn[0].Rc;
          op.g4 = fn[0].kind == 4;
          traverse(op);
        } else if (bpt->ang_fast[key3]) {
          AngForceFastOp op;
          op.R = S.R;
          op.r = S.r;
          op.tab = tab;
          op.fs = S.fs;
          op.K = (bp_angk)md.bpk + bpt->ang_p0[key3];
          op.coef = coef + off;
          op.nf = nf;
          op.np = bpt->ang_np[key3];
          op.rc = fn[0].Rc;
          op.g4 = fn[0].kind == 4;
          traverse(op);
        } else {
          AngForceOp op;
          op.R = S.R;
          op.r = S.r;
          op.tab = tab;
          op.fs = S.fs;
          op.fn = fn;
          op.coef = coef + off;
          op.nf = nf;
          op.one_seg = bpt->ang_one_seg[key3] != 0;
          traverse(op);
        }
      }
    }
    wave_sync();
    // F_j -= dE_i/dR_j ; F_i += sum_j dE_i/dR_j  (src/aenet.f90:471-477)
    if constexpr (MODE == 0) {
      const int *Jl = nbl.j + (size_t)g * nbl.cap;  // neighbor ids from the list
      scatter_forces(o.F, b.fmark, g, n, Jl, [&](int q, double (&c)[3]) {
        c[0] = S.fs[3 * q];
        c[1] = S.fs[3 * q + 1];
        c[2] = S.fs[3 * q + 2];
      });
    } else {
      double own[3] = {0.0, 0.0, 0.0};
      const int e0 = env.env_start[g];
      for (int q = lane; q < n; q += WAVE) {
        double gx = S.fs[3 * q], gy = S.fs[3 * q + 1], gz = S.fs[3 * q + 2];
        own[0] += gx;
        own[1] += gy;
        own[2] += gz;
        size_t e = (size_t)(e0 + q);
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
  }
}

template <int MODE, bool F>
static void launch_bp(const ModelDev &md, const BatchDev &b, const NbDev &nb, const OutDev &o, const EnvDev &e,
                      BPLaunch P, hipStream_t stream) {
  PersistentShape ps;
  const size_t wbytes = P.wsize <= 16384 ? 4 * (size_t)((P.wsize + 3) / 4 * 4) : 0;
  if (!persistent_shape(wbytes, 4 * (size_t)bp_wave_words(P), P.natoms, BP_SIMD_WAVES, ps, ATOM_MAX_WAVES, P.tune_waves,
                        P.tune_staged)) {
    raise_flag(b.flags, AENET_AMD_FLAG_NB_OVERFLOW, stream);
    return;
  }
  P.w_in_lds = ps.w_in_lds;
  P.waves = ps.waves;
  if (P.hs > 64)  // hidden layers wider than the wavefront
    hipLaunchKernelGGL((k_atom_bp<MODE, F, true>), dim3(ps.grid), dim3(64 * ps.waves), ps.lds, stream, md, b, nb, o, e,
                       P);
  else
    hipLaunchKernelGGL((k_atom_bp<MODE, F, false>), dim3(ps.grid), dim3(64 * ps.waves), ps.lds, stream, md, b, nb, o, e,
                       P);
}

static BPLaunch bp_launch(int mode, const AtomSizes &z) {
  BPLaunch P;
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
  return P;
}

bool atom_bp_shape_ok(const AtomSizes &z) {
  const BPLaunch P = bp_launch(0, z);
  PersistentShape ps;
  const size_t wbytes = P.wsize <= 16384 ? 4 * (size_t)((P.wsize + 3) / 4 * 4) : 0;
  return persistent_shape(wbytes, 4 * (size_t)bp_wave_words(P), P.natoms, BP_SIMD_WAVES, ps, ATOM_MAX_WAVES, P.tune_waves,
                          P.tune_staged);
}

void launch_atom_bp(int mode, bool forces, const AtomSizes &z, const ModelDev &md, const BatchDev &b,
                    const NbDev &nb, const OutDev &o, const EnvDev &e, hipStream_t stream) {
  if (z.natoms <= 0) return;
  const BPLaunch P = bp_launch(mode, z);
  if (mode == 0) {
    if (forces)
      launch_bp<0, true>(md, b, nb, o, e, P, stream);
    else
      launch_bp<0, false>(md, b, nb, o, e, P, stream);
  } else {
    if (forces)
      launch_bp<1, true>(md, b, nb, o, e, P, stream);
    else
      launch_bp<1, false>(md, b, nb, o, e, P, stream);
  }
}

}  // namespace aenet_amd

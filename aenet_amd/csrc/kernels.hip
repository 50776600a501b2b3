// kernels.hip -- gfx950 kernels of the energy/force path (everything but
// the descriptor/network/force kernels of atom_cheb.hip / atom_bp.hip).
//
// Per step:  k_frac -> k_bin_cells or k_bin_small (or k_frac_bin ->
//            k_scan_cells -> k_place -> k_cell_sort) -> k_search_cells
//            (compact lists) -> atom kernel -> k_struct_energy
// Subsets of atoms (Monte Carlo, autograd): k_search, a wavefront per atom.
// Once per batch: k_combos (cell/image combinations of each home cell).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "aenet_amd_internal.h"
#include "kernels.h"
#include "kernels_common.h"

namespace aenet_amd {

// ---------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------

__global__ void k_fill_i32(int *p, int v, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *p, double v, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_scale_f64(double *p, double a, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = a * p[i];
}

// The cell-major search (k_search_cells) runs one workgroup of CS_WAVES
// wavefronts per work item (c, j): atoms CS_WAVES*j .. CS_WAVES*j + CS_WAVES-1
// of home cell c, in cell order.  The binning kernels list the items: item
// (c, j) sits at slot start_c / CS_WAVES + c + j (start_c = first atom of c in
// cell order), which is below the next cell's first slot; the slots between
// are -1.  The slot count is at most natoms / CS_WAVES + ncells + 1.  A slot
// holds (c, start_c, atoms of c, structure of c), so the search starts from
// one load.  CS_WAVES is set in aenet_amd_internal.h (the host sizes the
// list from it).

__device__ inline void write_cell_work(const BatchDev &b, int c, int start, int next, int first, int stride) {
  const int sid = b.cell_struct[c];
  const int base = start / CS_WAVES + c;
  const int end = c + 1 < b.ncells ? next / CS_WAVES + c + 1 : b.work_cap;
  const int items = (next - start + CS_WAVES - 1) / CS_WAVES;
  for (int k = base + first; k < end; k += stride)
    reinterpret_cast<int4 *>(b.cell_work)[k] =
        k - base < items ? make_int4(c, start, next - start, sid) : make_int4(-1, 0, 0, 0);
}

// ---------------------------------------------------------------------
// binning: Cartesian -> lattice coordinates (geo_init_xsf,
// src/geometry.f90:716), wrap (wrap_cooLatt, src/ext/lclist.f90:953-987),
// cell (cell_get, src/ext/lclist.f90:833-845)
// ---------------------------------------------------------------------

// lattice coordinates and cell of atom g of structure S at Cartesian x
// (returns the global cell id)
__device__ __forceinline__ int frac_cell_at(const BatchDev &b, int g, const StructDev *S, double x[3], double *F) {
  if (F) {
    F[3 * (size_t)g] = 0.0;
    F[3 * (size_t)g + 1] = 0.0;
    F[3 * (size_t)g + 2] = 0.0;
  }
  double s[3];
  bool bad = false;
  {
#pragma clang fp contract(off)
    if (!S->pbc) {
      x[0] = x[0] - S->shift[0];
      x[1] = x[1] - S->shift[1];
      x[2] = x[2] - S->shift[2];
    }
    const double div = S->frac_div;
    for (int r = 0; r < 3; ++r) {
      double acc = S->B[r] * x[0] + S->B[3 + r] * x[1];
      acc = acc + S->B[6 + r] * x[2];
      s[r] = acc / div;
    }
    for (int r = 0; r < 3; ++r)
      if (!(fabs(s[r]) < 1.0e6)) {  // NaN, inf or absurdly far: refuse, do not loop
        bad = true;
        s[r] = 0.0;
      }
    if (S->pbc) {
      for (int r = 0; r < 3; ++r) {
        while (s[r] < 0.0) s[r] = s[r] + 1.0;
        while (s[r] >= 1.0) s[r] = s[r] - 1.0;
      }
    } else {
      // an isolated structure must stay inside the bounding box its plan
      // was made for
      for (int r = 0; r < 3; ++r)
        if (s[r] < -1.0e-9 || s[r] > 1.0 + 1.0e-9) bad = true;
    }
  }
  if (bad) atomicOr(b.flags, AENET_AMD_FLAG_BAD_COORD);
  int ic[3];
  for (int r = 0; r < 3; ++r) {
    double t;
    {
#pragma clang fp contract(off)
      t = s[r] * (double)S->nc[r] + 0.5;
    }
    int v = (int)round(t);
    v = v < S->nc[r] ? v : S->nc[r];
    ic[r] = v < 1 ? 1 : v;
  }
  int cell = S->cell0 + ((ic[0] - 1) * S->nc[1] + (ic[1] - 1)) * S->nc[2] + (ic[2] - 1);
  b.s[3 * (size_t)g] = s[0];
  b.s[3 * (size_t)g + 1] = s[1];
  b.s[3 * (size_t)g + 2] = s[2];
  b.cellof[g] = cell;
  return cell;
}

__device__ __forceinline__ int frac_cell(const BatchDev &b, int g, double *F) {
  double x[3] = {b.cart[3 * (size_t)g], b.cart[3 * (size_t)g + 1], b.cart[3 * (size_t)g + 2]};
  return frac_cell_at(b, g, b.st + b.atom_struct[g], x, F);
}

__global__ void k_frac_bin(BatchDev b, double *F) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= b.natoms) return;
  atomicAdd(&b.cell_count[frac_cell(b, g, F)], 1);
}

// lattice coordinates and cells only (small batches: k_bin_small counts)
__global__ __launch_bounds__(256) void k_frac(BatchDev b, double *F) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= b.natoms) return;
  frac_cell(b, g, F);
}

// the binning of a small batch after k_frac, in one workgroup: counts in
// LDS, exclusive scan, placement, and the order within each cell
// (descending atom index, as the reference's linked lists) with the lattice
// coordinates copied into that order.  Each thread takes the atoms
// tid + u*1024 (u < SMALL_PER_THREAD) and keeps their cells and lattice
// coordinates in registers.  In LDS each cell's atoms start at a multiple
// of 4 slots, padded with -1, so the rank count reads 4 slots at a time.
constexpr int SMALL_ATOMS = 8192, SMALL_CELLS = 3072, SMALL_PER_THREAD = SMALL_ATOMS / 1024;
constexpr int SMALL_SLOTS = SMALL_ATOMS + 3 * SMALL_CELLS;
__global__ __launch_bounds__(1024) void k_bin_small(BatchDev b) {
  __shared__ int cnt[SMALL_CELLS + 1];  // counts, then first atom of each cell in cell order
  __shared__ int pst[SMALL_CELLS + 1];  // first slot of each cell (multiple of 4)
  __shared__ int fill[SMALL_CELLS];
  __shared__ int4 slot4[SMALL_SLOTS / 4];
  __shared__ int wsum[16];
  int *slot = reinterpret_cast<int *>(slot4);
  const int tid = threadIdx.x;
  const int nat = b.natoms;
  for (int c = tid; c < b.ncells; c += blockDim.x) {
    cnt[c] = 0;
    fill[c] = 0;
  }
  // cells and lattice coordinates of this thread's atoms
  int cs[SMALL_PER_THREAD];
  double ss[SMALL_PER_THREAD][3];
#pragma unroll
  for (int u = 0; u < SMALL_PER_THREAD; ++u) {
    const int g = tid + u * 1024;
    cs[u] = 0;
    if (g < nat) {
      cs[u] = b.cellof[g];
      for (int k = 0; k < 3; ++k) ss[u][k] = b.s[3 * (size_t)g + k];
    }
  }
  __syncthreads();  // counters cleared
#pragma unroll
  for (int u = 0; u < SMALL_PER_THREAD; ++u)
    if (tid + u * 1024 < nat) atomicAdd(&cnt[cs[u]], 1);
  __syncthreads();
  // exclusive scan of (count, count rounded up to 4), packed in the low and
  // high 16 bits (totals below 8192 and 17408) -> cell_start and the slots
  int carry = 0;
  for (int base = 0; base < b.ncells; base += 1024) {
    const int i = base + tid;
    const int n = i < b.ncells ? cnt[i] : 0;
    const int v = n | (((n + 3) & ~3) << 16);
    const int x = wave_scan(v);
    const int w = tid / WAVE;
    if (lane_id() == WAVE - 1) wsum[w] = x;
    __syncthreads();
    if (tid < 16) {
      int t = wsum[tid];
      for (int o = 1; o < 16; o <<= 1) {
        const int y = __shfl_up(t, o, 16);
        if (tid >= o) t += y;
      }
      wsum[tid] = t;
    }
    __syncthreads();
    const int off = carry + (w > 0 ? wsum[w - 1] : 0) + x - v;
    const int tot = wsum[15];
    __syncthreads();
    if (i < b.ncells) {
      cnt[i] = off & 0xffff;
      pst[i] = off >> 16;
      b.cell_start[i] = off & 0xffff;
    }
    carry += tot;
  }
  if (tid == 0) {
    cnt[b.ncells] = carry & 0xffff;
    pst[b.ncells] = carry >> 16;
    b.cell_start[b.ncells] = carry & 0xffff;
  }
  __syncthreads();
  for (int c = tid; c < b.ncells; c += blockDim.x) {
    write_cell_work(b, c, cnt[c], cnt[c + 1], 0, 1);
    for (int p = pst[c] + cnt[c + 1] - cnt[c]; p < pst[c + 1]; ++p) slot[p] = -1;  // padding
  }
#pragma unroll
  for (int u = 0; u < SMALL_PER_THREAD; ++u) {
    const int g = tid + u * 1024;
    if (g < nat) slot[pst[cs[u]] + atomicAdd(&fill[cs[u]], 1)] = g;
  }
  __syncthreads();
  // rank of each of this thread's atoms within its cell, then the atom and
  // its lattice coordinates into that place
#pragma unroll
  for (int u = 0; u < SMALL_PER_THREAD; ++u) {
    const int g = tid + u * 1024;
    if (g < nat) {
      const int c = cs[u];
      const int p0 = pst[c] >> 2, p1 = pst[c + 1] >> 2;
      int r0 = 0, r1 = 0;
#pragma unroll 4
      for (int p = p0; p < p1; ++p) {
        const int4 e = slot4[p];
        r0 += (e.x > g) + (e.y > g);
        r1 += (e.z > g) + (e.w > g);
      }
      const int q = cnt[c] + r0 + r1;
      b.cell_atoms[q] = g;
      b.sorted_s[3 * (size_t)q] = ss[u][0];
      b.sorted_s[3 * (size_t)q + 1] = ss[u][1];
      b.sorted_s[3 * (size_t)q + 2] = ss[u][2];
    }
  }
}

// The binning of a small batch after k_frac with one workgroup per cell
// instead of one for the batch: each workgroup reads every atom's cell,
// counts the atoms of lower cells (its cell's start in cell order) and
// places its own atoms in descending index order (the reference's linked
// lists), with their lattice coordinates and the cell's search work items.
// Wavefront w takes the contiguous atom range [N - (w+1) P, N - w P) in
// chunks of 64 going down, lane l the atom top - 64 u - l of chunk u
// (coalesced loads, all issued first); ballots rank the cell's atoms within
// a chunk and running counts across chunks and wavefronts.  The results are
// those of k_bin_small.
constexpr int BIN_CELLS_THREADS = 256, BIN_CELLS_WAVES = BIN_CELLS_THREADS / WAVE;
constexpr int BIN_CELLS_CHUNKS = SMALL_ATOMS / BIN_CELLS_THREADS;  // chunks of 64 per wavefront at most
constexpr long BIN_CELLS_WORK = 4L << 20;                          // natoms x ncells
__global__ __launch_bounds__(BIN_CELLS_THREADS) void k_bin_cells(BatchDev b) {
  __shared__ int wb[BIN_CELLS_WAVES], wm[BIN_CELLS_WAVES];
  const int c = blockIdx.x;
  const int tid = threadIdx.x, w = tid / WAVE, lane = lane_id();
  const int N = b.natoms;
  const int P = (N + BIN_CELLS_THREADS - 1) / BIN_CELLS_THREADS * WAVE;  // atoms per wavefront, multiple of 64
  const int top = N - w * P, bottom = max(top - P, 0);
  int cc[BIN_CELLS_CHUNKS];
#pragma unroll
  for (int u = 0; u < BIN_CELLS_CHUNKS; ++u) {
    const int g = top - 1 - WAVE * u - lane;
    cc[u] = g >= bottom ? b.cellof[g] : -1;
  }
  int before = 0, mine = 0;
  unsigned long long mk[BIN_CELLS_CHUNKS];
#pragma unroll
  for (int u = 0; u < BIN_CELLS_CHUNKS; ++u) {
    before += cc[u] >= 0 && cc[u] < c;
    mk[u] = wave_ballot(cc[u] == c);
    mine += popc64(mk[u]);
  }
  before = (int)wave_sum((long long)before);
  if (lane == 0) {
    wb[w] = before;
    wm[w] = mine;
  }
  __syncthreads();
  int start = 0, total = 0, off = 0;
#pragma unroll
  for (int v = 0; v < BIN_CELLS_WAVES; ++v) {
    start += wb[v];
    total += wm[v];
    if (v < w) off += wm[v];
  }
  int q = start + off;
#pragma unroll
  for (int u = 0; u < BIN_CELLS_CHUNKS; ++u) {
    if (cc[u] == c) {
      const int g = top - 1 - WAVE * u - lane;
      const int p = q + mask_prefix(mk[u]);
      b.cell_atoms[p] = g;
      b.sorted_s[3 * (size_t)p] = b.s[3 * (size_t)g];
      b.sorted_s[3 * (size_t)p + 1] = b.s[3 * (size_t)g + 1];
      b.sorted_s[3 * (size_t)p + 2] = b.s[3 * (size_t)g + 2];
    }
    q += popc64(mk[u]);
  }
  if (tid == 0) {
    b.cell_start[c] = start;
    if (c == b.ncells - 1) b.cell_start[b.ncells] = start + total;
  }
  write_cell_work(b, c, start, start + total, tid, BIN_CELLS_THREADS);
}

// exclusive scan of the cell counts (one workgroup)
__global__ __launch_bounds__(1024) void k_scan_cells(BatchDev b) {
  __shared__ int wsum[16];
  __shared__ int carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < b.ncells; base += 1024) {
    int i = base + tid;
    int v = i < b.ncells ? b.cell_count[i] : 0;
    int x = wave_scan(v);
    int w = tid / WAVE;
    if (lane_id() == WAVE - 1) wsum[w] = x;
    __syncthreads();
    if (tid < 16) {
      int s = wsum[tid];
      for (int o = 1; o < 16; o <<= 1) {
        int y = __shfl_up(s, o, 16);
        if (tid >= o) s += y;
      }
      wsum[tid] = s;
    }
    __syncthreads();
    int off = carry + (w > 0 ? wsum[w - 1] : 0);
    if (i < b.ncells) {
      b.cell_start[i] = off + x - v;
      b.cell_fill[i] = 0;  // k_place counts into it next
    }
    __syncthreads();
    if (tid == 0) carry += wsum[15];
    __syncthreads();
  }
  if (tid == 0) b.cell_start[b.ncells] = carry;
}

__global__ void k_place(BatchDev b, int *tmp) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= b.natoms) return;
  int c = b.cellof[g];
  int pos = atomicAdd(&b.cell_fill[c], 1);
  tmp[b.cell_start[c] + pos] = g;
}

// order each cell like the reference's linked list (atoms are prepended, so
// descending atom index) and copy the lattice coordinates into that order;
// one wavefront per cell, every atom placed by its rank
__global__ __launch_bounds__(64) void k_cell_sort(BatchDev b, const int *tmp) {
  const int c = blockIdx.x;
  const int start = b.cell_start[c], n = b.cell_start[c + 1] - start;
  if (lane_id() == 0) b.cell_count[c] = 0;  // ready for the next step's k_frac_bin
  write_cell_work(b, c, start, start + n, lane_id(), WAVE);
  for (int a = lane_id(); a < n; a += WAVE) {
    int v = tmp[start + a];
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += tmp[start + k] > v;
    int p = start + rank;
    b.cell_atoms[p] = v;
    b.sorted_s[3 * (size_t)p] = b.s[3 * (size_t)v];
    b.sorted_s[3 * (size_t)p + 1] = b.s[3 * (size_t)v + 1];
    b.sorted_s[3 * (size_t)p + 2] = b.s[3 * (size_t)v + 2];
  }
}

// (cell, image) combinations of each home cell, in the reference's visiting
// order (stencil rank major, image rank minor), that can hold a neighbor of
// an atom of the home cell: cells are parallelepipeds of the structure's
// grid, bounded by spheres of the grid cell's circumradius and separated by
// whole plane spacings of the grid.  Computed once per batch; one
// workgroup of 256 per home cell.
__global__ __launch_bounds__(256) void k_combos(BatchDev b, int *combo_count, int *combos_out) {
  const int h = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ int wcount[4];
  __shared__ int total;
  if (tid == 0) total = 0;
  __syncthreads();
  const StructDev *S = b.st + b.cell_struct[h];
  const int nT1 = S->nT + 1;
  const int nitems = (S->nC + 1) * nT1;
  const double hc[3] = {b.cell_ctr[3 * h], b.cell_ctr[3 * h + 1], b.cell_ctr[3 * h + 2]};
  const int hijk[3] = {b.cell_ijk[3 * h], b.cell_ijk[3 * h + 1], b.cell_ijk[3 * h + 2]};
  const double lim = S->rc_list + 2.0 * S->cell_rad + 1.0e-6;
  int *out = combos_out + 2 * (size_t)h * b.combo_cap;
  for (int base = 0; base < nitems; base += 256) {
    int q = base + tid;
    bool keep = false;
    int c = 0, t = 0;
    if (q < nitems) {
      int r = q / nT1;
      t = q - r * nT1;
      int ic[3];
      bool inside = true;
      if (r == 0) {
        ic[0] = hijk[0];
        ic[1] = hijk[1];
        ic[2] = hijk[2];
      } else {
        const int *cv = b.Cvec + 3 * (S->C0 + r - 1);
        for (int k = 0; k < 3; ++k) {
          if (S->pbc) {
            int m = (hijk[k] + cv[k] - 1) % S->nc[k];
            ic[k] = (m < 0 ? m + S->nc[k] : m) + 1;
          } else {
            ic[k] = hijk[k] + cv[k];
            if (ic[k] < 1 || ic[k] > S->nc[k]) inside = false;
          }
        }
      }
      if (inside) {
        c = S->cell0 + ((ic[0] - 1) * S->nc[1] + (ic[1] - 1)) * S->nc[2] + (ic[2] - 1);
        double ctr[3] = {b.cell_ctr[3 * c], b.cell_ctr[3 * c + 1], b.cell_ctr[3 * c + 2]};
        // grid offset of the image cell; cells |d| apart along direction k
        // are at least (|d| - 1) plane spacings apart
        int d[3] = {ic[0] - hijk[0], ic[1] - hijk[1], ic[2] - hijk[2]};
        if (t > 0) {
          const int *T = b.Tvec + 3 * (S->T0 + t - 1);
          for (int k = 0; k < 3; ++k) ctr[k] += S->L[k] * T[0] + S->L[3 + k] * T[1] + S->L[6 + k] * T[2];
          for (int k = 0; k < 3; ++k) d[k] += S->nc[k] * T[k];
        }
        double gap = 0.0;
        for (int k = 0; k < 3; ++k) gap = fmax(gap, (double)max(abs(d[k]) - 1, 0) * S->cell_gap[k]);
        double dx = ctr[0] - hc[0], dy = ctr[1] - hc[1], dz = ctr[2] - hc[2];
        keep = dx * dx + dy * dy + dz * dz <= lim * lim && gap <= S->rc_list + 1.0e-5;
      }
    }
    unsigned long long m = wave_ballot(keep);
    int w = tid / WAVE;
    if (lane_id() == 0) wcount[w] = popc64(m);
    __syncthreads();
    int off = total;
    for (int k = 0; k < w; ++k) off += wcount[k];
    int pos = off + mask_prefix(m);
    if (keep && pos < b.combo_cap) {
      out[2 * pos] = c;
      out[2 * pos + 1] = t;
    }
    __syncthreads();
    if (tid == 0) total += wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
  }
  if (tid == 0) {
    if (total > b.combo_cap) {
      atomicOr(b.flags, AENET_AMD_FLAG_COMBO_OVERFLOW);
      total = b.combo_cap;
    }
    combo_count[h] = total;
  }
}

// The end of an evaluation, one launch:
// - workgroups 0 .. nsum-1: per-structure energy sums in FP64 (fixed
//   reduction order: thread t sums atoms t, t+256, ... in turn, then a tree
//   over the threads); the loads of SE_UNROLL atoms issue before their sums;
// - the workgroups after them, with F: the fixed-point force sums of the
//   atom kernels (scatter_forces) turned into FP64 times fscale, a thread per
//   atom; an atom marked out of range gets NaN and its mark is cleared.
constexpr int SE_UNROLL = 8;
__global__ __launch_bounds__(256) void k_struct_energy(BatchDev b, ModelDev md, const double *E_atom,
                                                       double *E_tot, double *E_coh, int nsum, double *F,
                                                       double fscale) {
  __shared__ double st[256], sc[256];
  __shared__ double e_free[AENET_AMD_MAX_TYPES];
  if ((int)blockIdx.x >= nsum) {
    const int cb = blockIdx.x - nsum;
    const double nan = __builtin_nan("");
    if (b.fslots) {
      // 32 lanes per atom, lane k adds slot k: one memory round trip, then
      // an integer butterfly over the half-wavefront (the order of integer
      // sums changes nothing)
      static_assert(FORCE_SLOTS == 32, "one lane per slot");
      const int g = cb * 8 + (int)(threadIdx.x >> 5), k = threadIdx.x & 31;
      long long u[3] = {0, 0, 0};
      if (g < b.natoms) {
        if (k < min(b.fcnt[g], b.fslots)) {
          const unsigned long long *sl = b.fslot + 3 * ((size_t)g * b.fslots + k);
          u[0] = (long long)sl[0];
          u[1] = (long long)sl[1];
          u[2] = (long long)sl[2];
        }
        if (k == 0) {  // the sums that went to F directly
          const long long *Fi = reinterpret_cast<const long long *>(F) + 3 * (size_t)g;
          u[0] += Fi[0];
          u[1] += Fi[1];
          u[2] += Fi[2];
        }
      }
#pragma unroll
      for (int x = 0; x < 3; ++x) u[x] = butterfly_sum<16>(u[x]);
      if (g < b.natoms && k == 0) {
        const bool bad = b.fmark[g] != 0;
        F[3 * (size_t)g] = bad ? nan : (double)u[0] * fscale;
        F[3 * (size_t)g + 1] = bad ? nan : (double)u[1] * fscale;
        F[3 * (size_t)g + 2] = bad ? nan : (double)u[2] * fscale;
        if (bad) b.fmark[g] = 0;
        b.fcnt[g] = 0;
      }
    } else {
      const int g = cb * 256 + threadIdx.x;
      if (g < b.natoms) {
        const long long *Fi = reinterpret_cast<const long long *>(F) + 3 * (size_t)g;
        const long long v0 = Fi[0], v1 = Fi[1], v2 = Fi[2];
        const bool bad = b.fmark[g] != 0;
        F[3 * (size_t)g] = bad ? nan : (double)v0 * fscale;
        F[3 * (size_t)g + 1] = bad ? nan : (double)v1 * fscale;
        F[3 * (size_t)g + 2] = bad ? nan : (double)v2 * fscale;
        if (bad) b.fmark[g] = 0;
      }
    }
    return;
  }
  const int s = blockIdx.x;
  const StructDev *S = b.st + s;
  const int n = S->natoms, a0 = S->atom0;
  if (threadIdx.x < AENET_AMD_MAX_TYPES)
    e_free[threadIdx.x] = (int)threadIdx.x < md.ntypes ? md.species[threadIdx.x].E_atom : 0.0;
  __syncthreads();
  double et = 0.0, ec = 0.0;
  for (int i0 = 0; i0 < n; i0 += 256 * SE_UNROLL) {
    double e[SE_UNROLL];
    int t[SE_UNROLL];
#pragma unroll
    for (int u = 0; u < SE_UNROLL; ++u) {
      const int i = i0 + u * 256 + (int)threadIdx.x;
      e[u] = 0.0;
      t[u] = 1;
      if (i < n) {
        e[u] = E_atom[a0 + i];
        t[u] = b.types[a0 + i];
      }
    }
#pragma unroll
    for (int u = 0; u < SE_UNROLL; ++u) {
      const int i = i0 + u * 256 + (int)threadIdx.x;
      if (i < n) {
        et += e[u];
        ec += e[u] - e_free[min(max(t[u], 1), AENET_AMD_MAX_TYPES) - 1];
      }
    }
  }
  st[threadIdx.x] = et;
  sc[threadIdx.x] = ec;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      st[threadIdx.x] += st[threadIdx.x + o];
      sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    E_tot[s] = st[0];
    E_coh[s] = sc[0];
  }
}

// ---------------------------------------------------------------------
// neighbor search kernels (one wavefront per atom)
// ---------------------------------------------------------------------

// compact lists for the atom kernels: R_ij and |R_ij| in FP32, j
struct CompactSink {
  NbDev nb;
  int g;
  int *overflow;
  __device__ void put(int slot, int j, const double *, const double *cart, double) {
    if (slot >= nb.cap) {
      *overflow = 1;
      return;
    }
    size_t e = (size_t)g * nb.cap + slot;
    float4 v = make_float4((float)cart[0], (float)cart[1], (float)cart[2], (float)sqrt(dist2_exact(cart)));
    reinterpret_cast<float4 *>(nb.R)[e] = v;
    nb.j[e] = j;
  }
};

__global__ __launch_bounds__(64) void k_search(BatchDev b, NbDev nb, const int *active, const int *nactive) {
  extern __shared__ int sdyn[];
  SearchLds W = carve_search(sdyn, b.kept_cap);
  __shared__ int ovf;
  int g = blockIdx.x;
  if (active) {
    if (g >= *nactive) return;  // the whole workgroup leaves together
    g = active[g];
  }
  if (threadIdx.x == 0) ovf = 0;
  __syncthreads();
  CompactSink sink{nb, g, &ovf};
  int n = search_neighbors(b, g, sink, W);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (ovf || n > nb.cap) {
      atomicOr(b.flags, AENET_AMD_FLAG_NB_OVERFLOW);
      n = min(n, nb.cap);
    }
    nb.count[g] = n;
  }
}

// bit-exact FP64 lists (aenet_nbl_* and the batch neighbor entry point)
struct ListSink {
  const BatchDev *b;
  OutDev o;
  int g;
  int *overflow;
  __device__ void put(int slot, int j, const double *coo2, const double *, double dist) {
    if (slot >= o.nb_cap) {
      *overflow = 1;
      return;
    }
    size_t base = (size_t)g * o.nb_cap + slot;
    o.nb_coo[3 * base] = coo2[0];
    o.nb_coo[3 * base + 1] = coo2[1];
    o.nb_coo[3 * base + 2] = coo2[2];
    o.nb_dist[base] = dist;
    const StructDev *S = b->st + b->atom_struct[j];
    o.nb_list[base] = j - S->atom0 + 1;
    o.nb_type[base] = b->types[j];
  }
};

// ---------------------------------------------------------------------
// cell-major search: one workgroup per (home cell, split)
// ---------------------------------------------------------------------
// Every atom of a home cell tests the same candidates: the atoms of the
// home cell's (cell, image) combinations in visiting order.  The workgroup
// computes their Cartesian coordinates L (s_j + T) with the reference's
// FP64 expression once, into LDS, and each wavefront scans them for one
// central atom at a time (distance test only).  The entries that pass come
// out in the order of search_neighbors: a combination that its per-atom
// sphere test skips holds no neighbor of that atom, and runs of equal cell
// are the stencil entries in both (stencil cells are distinct).  Passing
// candidates queue in LDS and reach the sink 64 at a time.
constexpr int CS_CHUNK = 1024; // candidates staged at once
constexpr int CS_QUEUE = 128;  // pending passing candidates per wavefront

struct CellSearchLds {
  double *cx, *cy, *cz;  // [CS_CHUNK] candidate coordinates (Cartesian, A)
  float4 *cf;            // [CS_CHUNK] FP32 coordinates relative to the home cell centre; w: atom (bits)
  int *queue;            // [CS_WAVES * CS_QUEUE]
  int *ovf;              // [CS_WAVES] sink overflow word (the count decides)
  int *hdr;              // [2] groups, candidates
  int *cc;               // [2*kc] the home cell's (cell, image rank) combinations
  int *gk, *gn;          // [kc] groups (runs of equal cell): first combination, images
  int *gpos;             // [2*kc] group: first atom position in cell order; [kc + k]: one past its last
  int *gitems;           // [kc+1] prefix of atoms x images
};

static inline size_t cell_search_lds_bytes(int kc) {
  return (size_t)CS_CHUNK * (3 * sizeof(double) + sizeof(float4)) +
         sizeof(int) * ((size_t)CS_WAVES * (CS_QUEUE + 1) + 2 + 7 * (size_t)kc + 1);
}

__device__ inline CellSearchLds carve_cell_search(double *p, int kc) {
  CellSearchLds W;
  W.cx = p;
  W.cy = p + CS_CHUNK;
  W.cz = p + 2 * CS_CHUNK;
  W.cf = reinterpret_cast<float4 *>(p + 3 * CS_CHUNK);
  int *q = reinterpret_cast<int *>(W.cf + CS_CHUNK);
  W.queue = q;
  W.ovf = W.queue + CS_WAVES * CS_QUEUE;
  W.hdr = W.ovf + CS_WAVES;
  W.cc = W.hdr + 2;
  W.gk = W.cc + 2 * kc;
  W.gn = W.gk + kc;
  W.gpos = W.gn + kc;
  W.gitems = W.gpos + 2 * kc;
  return W;
}

// Workgroup k takes work item (h, j) of slot k: wavefront w searches for
// atom a = CS_WAVES*j + w of home cell h.  Candidates are staged chunk by
// chunk; each wavefront scans a chunk before the next is staged.
template <bool LISTS>
__global__ __launch_bounds__(CS_WAVES *WAVE) void k_search_cells(BatchDev b, NbDev nb, OutDev o) {
  extern __shared__ double cs_dyn[];
  const int4 item = reinterpret_cast<const int4 *>(b.cell_work)[blockIdx.x];
  const int h = item.x;
  if (h < 0) return;  // empty slot (uniform)
  const int tid = threadIdx.x, lane = lane_id(), w = tid / WAVE;
  const int start = item.y, n = item.z;
  const int j = (int)blockIdx.x - (start / CS_WAVES + h);
  const int kc = b.combo_max;
  CellSearchLds W = carve_cell_search(cs_dyn, kc);
  const StructDev *S = b.st + item.w;
  // this wavefront's central atom: its loads issue first
  const int a = j * CS_WAVES + w;
  const bool have = a < n;  // uniform per wavefront
  const int g = have ? b.cell_atoms[start + a] : 0;
  double si[3] = {0.0, 0.0, 0.0};
  if (have) {
    si[0] = b.s[3 * (size_t)g];
    si[1] = b.s[3 * (size_t)g + 1];
    si[2] = b.s[3 * (size_t)g + 2];
  }
  double L[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) L[k] = S->L[k];
  const int *Tv = b.Tvec + 3 * S->T0;
  const double Rc2 = S->Rc2;
  const double ctr[3] = {b.cell_ctr[3 * h], b.cell_ctr[3 * h + 1], b.cell_ctr[3 * h + 2]};
  // FP32 prefilter: candidates and the central atom lie within M = rc +
  // 3 cell radii of the cell centre; rounding to FP32 and the FP32
  // distance then err by less than 84 M^2 2^-24 < 1e-5 M^2, so a distance
  // outside Rc2 -+ band decides like the FP64 test, and only candidates
  // inside the band take the FP64 test
  const double M = S->rc_list + 3.0 * S->cell_rad;
  const double band = 1.0e-5 * M * M + 1.0e-6 * Rc2;
  const float lo2 = (float)(Rc2 - band), hi2 = (float)(Rc2 + band);
  // the home cell's combinations into LDS
  const int K = min(b.combo_count[h], kc);
  {
    const int *combos = b.combos + 2 * (size_t)h * b.combo_cap;
    for (int q = tid; q < 2 * K; q += CS_WAVES * WAVE) W.cc[q] = combos[q];
  }
  __syncthreads();
  // groups of the combinations and their candidate counts
  if (w == 0) {
    int ng = 0;
    for (int base = 0; base < K; base += WAVE) {
      const int q = base + lane;
      const int c = q < K ? W.cc[2 * q] : -1;
      const bool st = q < K && (q == 0 || W.cc[2 * (q - 1)] != c);
      const unsigned long long m = wave_ballot(st);
      if (st) {
        const int k = ng + mask_prefix(m);
        W.gk[k] = q;
        W.gpos[k] = b.cell_start[c];
        W.gpos[kc + k] = b.cell_start[c + 1];
      }
      ng += popc64(m);
    }
    wave_sync();
    int carry = 0;
    for (int base = 0; base < ng; base += WAVE) {
      const int k = base + lane;
      int items = 0;
      if (k < ng) {
        const int q1 = k + 1 < ng ? W.gk[k + 1] : K;
        W.gn[k] = q1 - W.gk[k];
        items = (W.gpos[kc + k] - W.gpos[k]) * (q1 - W.gk[k]);
      }
      const int incl = wave_scan(items);
      if (k < ng) W.gitems[k] = carry + incl - items;
      carry += __shfl(incl, WAVE - 1, WAVE);
    }
    if (lane == 0) {
      W.gitems[ng] = carry;
      W.hdr[0] = ng;
      W.hdr[1] = carry;
    }
  }
  __syncthreads();
  const int ng = W.hdr[0], total = W.hdr[1];
  const int nchunks = max(1, (total + CS_CHUNK - 1) / CS_CHUNK);
  const int cap = LISTS ? o.nb_cap : nb.cap;
  int *count_out = LISTS ? o.nb_count : nb.count;
  int *queue = W.queue + w * CS_QUEUE;
  double coo1[3] = {0.0, 0.0, 0.0};
  if (have) matvec_exact(L, si, coo1);
  const float c1x = (float)(coo1[0] - ctr[0]), c1y = (float)(coo1[1] - ctr[1]), c1z = (float)(coo1[2] - ctr[2]);

  // entries of one central atom g among the cn staged candidates, appended
  // from slot count on; returns the new count
  auto scan = [&](auto &sink, int g, const double *coo1, int count, int cn) {
    auto emit = [&](int q, int slot) {
      const double coo2[3] = {W.cx[q], W.cy[q], W.cz[q]};
      double cart[3];
      cart[0] = coo2[0] - coo1[0];
      cart[1] = coo2[1] - coo1[1];
      cart[2] = coo2[2] - coo1[2];
      sink.put(slot, __float_as_int(W.cf[q].w), coo2, cart, sqrt(dist2_exact(cart)));
    };
    int qn = 0;
    for (int base = 0; base < cn; base += WAVE) {
      const int q = base + lane;
      bool pass = false, border = false;
      if (q < cn) {
        const float4 v = W.cf[q];
        const float dx = v.x - c1x, dy = v.y - c1y, dz = v.z - c1z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        const bool other = __float_as_int(v.w) != g;
        pass = other && d2 < lo2;
        border = other && d2 >= lo2 && d2 <= hi2;
      }
      if (wave_ballot(border)) {
        if (border) {
          double cart[3];
          cart[0] = W.cx[q] - coo1[0];
          cart[1] = W.cy[q] - coo1[1];
          cart[2] = W.cz[q] - coo1[2];
          pass = dist2_exact(cart) <= Rc2;
        }
      }
      const unsigned long long m = wave_ballot(pass);
      if (pass) queue[qn + mask_prefix(m)] = q;
      qn += popc64(m);
      wave_sync();
      if (qn >= WAVE) {
        emit(queue[lane], count + lane);
        count += WAVE;
        wave_sync();
        if (lane < qn - WAVE) queue[lane] = queue[WAVE + lane];
        qn -= WAVE;
        wave_sync();
      }
    }
    if (qn > 0) {
      if (lane < qn) emit(queue[lane], count + lane);
      count += qn;
      wave_sync();
    }
    return count;
  };

  auto run = [&](auto &sink) {
    int count = have ? self_images(b, S, g, L, si, coo1, sink) : 0;
    for (int ch = 0; ch < nchunks; ++ch) {
      const int cn = max(0, min(CS_CHUNK, total - ch * CS_CHUNK));
      __syncthreads();  // the previous chunk's scans are done
      for (int i = tid; i < cn; i += CS_WAVES * WAVE) {
        const int q = ch * CS_CHUNK + i;
        // group of candidate q: last k with gitems[k] <= q
        int lo = 0, hi = ng - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (W.gitems[mid] <= q)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int off = q - W.gitems[lo], gn = W.gn[lo];
        const int ai = gn == 1 ? off : off / gn;
        const int tt = off - ai * gn;
        const int pos = W.gpos[lo] + ai;
        const int t = W.cc[2 * (W.gk[lo] + tt) + 1];
        const double sj[3] = {b.sorted_s[3 * (size_t)pos], b.sorted_s[3 * (size_t)pos + 1],
                              b.sorted_s[3 * (size_t)pos + 2]};
        double coo2[3];
        if (t == 0) {
          matvec_exact(L, sj, coo2);
        } else {
          double c2[3];
          add_T_exact(sj, Tv + 3 * (t - 1), c2);
          matvec_exact(L, c2, coo2);
        }
        W.cx[i] = coo2[0];
        W.cy[i] = coo2[1];
        W.cz[i] = coo2[2];
        W.cf[i] = make_float4((float)(coo2[0] - ctr[0]), (float)(coo2[1] - ctr[1]), (float)(coo2[2] - ctr[2]),
                              __int_as_float(b.cell_atoms[pos]));
      }
      __syncthreads();
      if (have) count = scan(sink, g, coo1, count, cn);
    }
    if (have && lane == 0) {
      // a put beyond the capacity was dropped: the count says so
      if (count > cap) {
        atomicOr(b.flags, AENET_AMD_FLAG_NB_OVERFLOW);
        count = cap;
      }
      count_out[g] = count;
    }
  };
  if constexpr (LISTS) {
    ListSink sink{&b, o, g, W.ovf + w};
    run(sink);
  } else {
    CompactSink sink{nb, g, W.ovf + w};
    run(sink);
  }
}

__global__ __launch_bounds__(64) void k_neighbor_lists(BatchDev b, OutDev o) {
  extern __shared__ int sdyn[];
  SearchLds W = carve_search(sdyn, b.kept_cap);
  __shared__ int ovf;
  const int g = blockIdx.x;
  if (threadIdx.x == 0) ovf = 0;
  __syncthreads();
  ListSink sink{&b, o, g, &ovf};
  int n = search_neighbors(b, g, sink, W);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (ovf || n > o.nb_cap) {
      atomicOr(b.flags, AENET_AMD_FLAG_NB_OVERFLOW);
      n = min(n, o.nb_cap);
    }
    o.nb_count[g] = n;
  }
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------

static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

__global__ void k_or_flag(int *flags, int bits) {
  if (threadIdx.x == 0) atomicOr(flags, bits);
}

void raise_flag(int *flags, int bits, hipStream_t s) {
  hipLaunchKernelGGL(k_or_flag, dim3(1), dim3(64), 0, s, flags, bits);
}

const DeviceLimits &device_limits() {
  // one process drives one device; the first launch fixes the values
  static DeviceLimits dl = [] {
    DeviceLimits d{256, 64 * 1024, 160 * 1024};
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return d;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) d.cus = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      d.lds_per_block = (size_t)v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && v > 0)
      d.lds_per_cu = (size_t)v;
    if (d.lds_per_cu < d.lds_per_block) d.lds_per_cu = d.lds_per_block;
    return d;
  }();
  return dl;
}

void launch_fill_i32(int *p, int v, long n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_fill_f64(double *p, double v, long n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_scale_f64(double *p, double a, long n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_scale_f64, dim3(nblk(n, 256)), dim3(256), 0, s, p, a, n);
}

void launch_combos(const BatchDev &b, int *combo_count, int *combos, hipStream_t s) {
  if (b.ncells > 0) hipLaunchKernelGGL(k_combos, dim3(b.ncells), dim3(256), 0, s, b, combo_count, combos);
}

// Neighbor-count buckets.  Thread per atom; within a workgroup the atoms
// of a bucket are ranked in atom order (wavefront ballots), and one global
// atomic per bucket and workgroup reserves their places.  The order of the
// workgroups' runs in a list is the order of those atomics; the atom
// kernels' results do not depend on it.
__global__ __launch_bounds__(256) void k_bucket_atoms(const int *nb_count, int natoms, BucketCaps bc, int *lists,
                                                      int *counts) {
  __shared__ int wcnt[4][MAX_BUCKETS];
  __shared__ int base[MAX_BUCKETS];
  const int tid = threadIdx.x, w = tid / WAVE, lane = lane_id();
  const int g = blockIdx.x * 256 + tid;
  const bool in = g < natoms;
  const int n = in ? nb_count[g] : 0;
  int bk = bc.n - 1;
  for (int k = bc.n - 2; k >= 0; --k)
    if (n <= bc.cap[k]) bk = k;
  int rank = 0;
  for (int k = 0; k < bc.n; ++k) {
    const unsigned long long m = wave_ballot(in && bk == k);
    if (in && bk == k) rank = mask_prefix(m);
    if (lane == 0) wcnt[w][k] = popc64(m);
  }
  __syncthreads();
  if (tid < bc.n) {
    const int t = wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
    base[tid] = t ? atomicAdd(&counts[tid], t) : 0;
  }
  __syncthreads();
  if (in) {
    int off = base[bk] + rank;
    for (int v = 0; v < w; ++v) off += wcnt[v][bk];
    lists[(size_t)bk * natoms + off] = g;
  }
}

void launch_bucket_atoms(const int *nb_count, int natoms, const BucketCaps &bc, int *lists, int *counts,
                         hipStream_t s) {
  launch_fill_i32(counts, 0, bc.n, s);
  if (natoms > 0) hipLaunchKernelGGL(k_bucket_atoms, dim3(nblk(natoms, 256)), dim3(256), 0, s, nb_count, natoms, bc,
                                     lists, counts);
}

void launch_binning(const BatchDev &b, int *tmp, double *F, hipStream_t s) {
  if (b.natoms <= SMALL_ATOMS && b.ncells <= SMALL_CELLS) {
    hipLaunchKernelGGL(k_frac, dim3(nblk(b.natoms, 256)), dim3(256), 0, s, b, F);
    // a workgroup per cell while every workgroup's pass over all atoms stays
    // small; one workgroup for the batch otherwise
    if ((long)b.natoms * b.ncells <= BIN_CELLS_WORK)
      hipLaunchKernelGGL(k_bin_cells, dim3(b.ncells), dim3(BIN_CELLS_THREADS), 0, s, b);
    else
      hipLaunchKernelGGL(k_bin_small, dim3(1), dim3(1024), 0, s, b);
    return;
  }
  // cell_count is zero on entry (cleared at allocation and by k_cell_sort)
  hipLaunchKernelGGL(k_frac_bin, dim3(nblk(b.natoms, 256)), dim3(256), 0, s, b, F);
  hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, b);
  hipLaunchKernelGGL(k_place, dim3(nblk(b.natoms, 256)), dim3(256), 0, s, b, tmp);
  hipLaunchKernelGGL(k_cell_sort, dim3(b.ncells), dim3(64), 0, s, b, (const int *)tmp);
}

void launch_search(const BatchDev &b, const NbDev &nb, hipStream_t s, const int *active, const int *nactive,
                   int max_active) {
  if (active) {
    if (max_active > 0)
      hipLaunchKernelGGL(k_search, dim3(max_active), dim3(64), 4 * (size_t)search_lds_ints(b.kept_cap), s, b, nb,
                         active, nactive);
    return;
  }
  if (b.natoms > 0 && b.ncells > 0) {
    OutDev o;
    memset(&o, 0, sizeof(o));
    hipLaunchKernelGGL(k_search_cells<false>, dim3(b.work_cap), dim3(CS_WAVES * WAVE),
                       cell_search_lds_bytes(b.combo_max), s, b, nb, o);
  }
}

void launch_neighbor_lists(const BatchDev &b, const OutDev &o, hipStream_t s) {
  if (b.natoms > 0 && b.ncells > 0) {
    NbDev nb;
    memset(&nb, 0, sizeof(nb));
    hipLaunchKernelGGL(k_search_cells<true>, dim3(b.work_cap), dim3(CS_WAVES * WAVE),
                       cell_search_lds_bytes(b.combo_max), s, b, nb, o);
  }
}

void launch_struct_energy(const BatchDev &b, const ModelDev &md, const double *E_atom, double *E_tot,
                          double *E_coh, double *F, double fscale, hipStream_t s) {
  const int nsum = E_tot ? b.nstruct : 0;
  const int nconv = F ? (b.fslots ? nblk(b.natoms, 8) : nblk(b.natoms, 256)) : 0;
  if (nsum + nconv > 0)
    hipLaunchKernelGGL(k_struct_energy, dim3(nsum + nconv), dim3(256), 0, s, b, md, E_atom, E_tot, E_coh, nsum, F,
                       fscale * FORCE_FIX_UNIT);
}

}  // namespace aenet_amd

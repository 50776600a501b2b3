// kernels.hip -- gfx950 kernels of the energy/force path (everything but
// the descriptor/network/force kernels of atom_cheb.hip / atom_bp.hip).
//
// Per step:  k_frac_bin -> k_scan_cells -> k_place -> k_cell_sort ->
//            k_search (compact lists) -> atom kernel -> k_struct_energy
// Once per batch: k_combos (cell/image combinations of each home cell).
#include <hip/hip_runtime.h>

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

// the whole binning of a small batch in one workgroup: counts in LDS,
// exclusive scan, placement, and the order within each cell (descending
// atom index, as the reference's linked lists) with the lattice
// coordinates copied into that order.  Each thread takes the atoms
// tid + u*1024 (u < SMALL_PER_THREAD) and issues all their global loads
// before the arithmetic; the cell of every atom stays in LDS.
constexpr int SMALL_ATOMS = 8192, SMALL_CELLS = 3072, SMALL_PER_THREAD = SMALL_ATOMS / 1024;
__global__ __launch_bounds__(1024) void k_bin_small(BatchDev b, double *F) {
  __shared__ int cnt[SMALL_CELLS + 1];
  __shared__ int fill[SMALL_CELLS];
  __shared__ int slot[SMALL_ATOMS];
  __shared__ int cel[SMALL_ATOMS];
  __shared__ int wsum[16];
  const int tid = threadIdx.x;
  const int nat = b.natoms;
  for (int c = tid; c < b.ncells; c += blockDim.x) {
    cnt[c] = 0;
    fill[c] = 0;
  }
  {
    int as[SMALL_PER_THREAD];
    double xs[SMALL_PER_THREAD][3];
#pragma unroll
    for (int u = 0; u < SMALL_PER_THREAD; ++u) {
      const int g = tid + u * 1024;
      if (g < nat) {
        as[u] = b.atom_struct[g];
        xs[u][0] = b.cart[3 * (size_t)g];
        xs[u][1] = b.cart[3 * (size_t)g + 1];
        xs[u][2] = b.cart[3 * (size_t)g + 2];
      }
    }
    __syncthreads();  // counters cleared
#pragma unroll
    for (int u = 0; u < SMALL_PER_THREAD; ++u) {
      const int g = tid + u * 1024;
      if (g < nat) {
        const int c = frac_cell_at(b, g, b.st + as[u], xs[u], F);
        cel[g] = c;
        atomicAdd(&cnt[c], 1);
      }
    }
  }
  __syncthreads();
  // exclusive scan of cnt -> cell_start (global) and cnt (LDS, in place)
  int carry = 0;
  for (int base = 0; base < b.ncells; base += 1024) {
    const int i = base + tid;
    const int v = i < b.ncells ? cnt[i] : 0;
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
      cnt[i] = off;
      b.cell_start[i] = off;
    }
    carry += tot;
  }
  if (tid == 0) {
    cnt[b.ncells] = carry;
    b.cell_start[b.ncells] = carry;
  }
  __syncthreads();
  for (int g = tid; g < nat; g += blockDim.x) {
    const int c = cel[g];
    slot[cnt[c] + atomicAdd(&fill[c], 1)] = g;
  }
  __syncthreads();
  // rank within the cell, then the lattice coordinates into that order
  int vs[SMALL_PER_THREAD], qs[SMALL_PER_THREAD];
#pragma unroll
  for (int u = 0; u < SMALL_PER_THREAD; ++u) {
    const int p = tid + u * 1024;
    vs[u] = -1;
    if (p < nat) {
      const int v = slot[p];
      const int c = cel[v];
      const int start = cnt[c], n = cnt[c + 1] - start;
      // eight independent counts: the LDS reads of a cell issue together
      int r8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      int k = 0;
      for (; k + 8 <= n; k += 8) {
#pragma unroll
        for (int e = 0; e < 8; ++e) r8[e] += slot[start + k + e] > v;
      }
      for (; k < n; ++k) r8[0] += slot[start + k] > v;
      const int rank = ((r8[0] + r8[1]) + (r8[2] + r8[3])) + ((r8[4] + r8[5]) + (r8[6] + r8[7]));
      vs[u] = v;
      qs[u] = start + rank;
    }
  }
  double ss[SMALL_PER_THREAD][3];
#pragma unroll
  for (int u = 0; u < SMALL_PER_THREAD; ++u)
    if (vs[u] >= 0)
      for (int k = 0; k < 3; ++k) ss[u][k] = b.s[3 * (size_t)vs[u] + k];
#pragma unroll
  for (int u = 0; u < SMALL_PER_THREAD; ++u) {
    if (vs[u] >= 0) {
      const int q = qs[u];
      b.cell_atoms[q] = vs[u];
      b.sorted_s[3 * (size_t)q] = ss[u][0];
      b.sorted_s[3 * (size_t)q + 1] = ss[u][1];
      b.sorted_s[3 * (size_t)q + 2] = ss[u][2];
    }
  }
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
// grid, bounded by spheres of the grid cell's circumradius.  Computed once
// per batch; one workgroup of 256 per home cell.
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
        if (t > 0) {
          const int *T = b.Tvec + 3 * (S->T0 + t - 1);
          for (int k = 0; k < 3; ++k) ctr[k] += S->L[k] * T[0] + S->L[3 + k] * T[1] + S->L[6 + k] * T[2];
        }
        double dx = ctr[0] - hc[0], dy = ctr[1] - hc[1], dz = ctr[2] - hc[2];
        keep = dx * dx + dy * dy + dz * dz <= lim * lim;
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

// per-structure energy sums in FP64 (fixed reduction order)
__global__ __launch_bounds__(256) void k_struct_energy(BatchDev b, ModelDev md, const double *E_atom,
                                                       double *E_tot, double *E_coh) {
  __shared__ double st[256], sc[256];
  const int s = blockIdx.x;
  const StructDev *S = b.st + s;
  double et = 0.0, ec = 0.0;
  for (int i = threadIdx.x; i < S->natoms; i += 256) {
    int g = S->atom0 + i;
    double e = E_atom[g];
    et += e;
    ec += e - md.species[b.types[g] - 1].E_atom;
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

void launch_binning(const BatchDev &b, int *tmp, double *F, hipStream_t s) {
  if (b.natoms <= SMALL_ATOMS && b.ncells <= SMALL_CELLS) {
    hipLaunchKernelGGL(k_bin_small, dim3(1), dim3(1024), 0, s, b, F);
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
  const int n = active ? max_active : b.natoms;
  if (n > 0)
    hipLaunchKernelGGL(k_search, dim3(n), dim3(64), 4 * (size_t)search_lds_ints(b.kept_cap), s, b, nb, active, nactive);
}

void launch_neighbor_lists(const BatchDev &b, const OutDev &o, hipStream_t s) {
  if (b.natoms > 0)
    hipLaunchKernelGGL(k_neighbor_lists, dim3(b.natoms), dim3(64), 4 * (size_t)search_lds_ints(b.kept_cap), s, b, o);
}

void launch_struct_energy(const BatchDev &b, const ModelDev &md, const double *E_atom, double *E_tot,
                          double *E_coh, hipStream_t s) {
  if (b.nstruct > 0)
    hipLaunchKernelGGL(k_struct_energy, dim3(b.nstruct), dim3(256), 0, s, b, md, E_atom, E_tot, E_coh);
}

}  // namespace aenet_amd

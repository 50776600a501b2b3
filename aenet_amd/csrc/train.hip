// train.hip -- network training kernels: weight gradients of the atomic
// energies (ff_wderiv, src/ext/feedforward.f90:763-860, contracted with
// per-atom upstream gradients) and the Adam step (src/optimize.f90:1012-1046).
//
// k_net_grad: persistent workgroups, one wavefront per atom.  The atom's
// normalized descriptor G' goes through the species' network (forward,
// activations kept in LDS); the atomic energy E_i = E_scale*y + E_shift +
// E_atom is written out; the upstream gradient g_i = dL/dE_i is either
// given or the squared-error residual coef*(E_i - E_ref_i); then the
// backward pass forms, per layer, the outer product of the pre-activation
// deltas and the layer inputs (bias input 1), flattened over the 64 lanes
// in the weight array's column-major order, and adds g_i*E_scale times it
// into a workgroup copy of the gradient in LDS.  Each workgroup adds its
// copy to the global gradient once at the end.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "aenet_amd_internal.h"
#include "atom_common.h"
#include "kernels.h"
#include "kernels_common.h"

namespace aenet_amd {

struct NetGradLaunch {
  int natoms;
  int nsf;       // max features (LDS carve)
  int nl;        // max layers
  int wsize;
  int grad_in_lds;
  int waves;
  int hs;        // network layer stride
};

__host__ __device__ inline int net_grad_wave_words(const NetGradLaunch &P) { return net_lds_words(P.nsf, P.nl, P.hs); }

template <bool GRAD>
__global__ __launch_bounds__(64 * ATOM_MAX_WAVES) void k_net_grad(ModelDev md, NetGradArgs a, NetGradLaunch P) {
  extern __shared__ float lds[];
  const int lane = lane_id();
  const int wave = threadIdx.x / WAVE;
  const int nw = P.waves;
  float *gacc = nullptr;
  int gw = 0;
  if (GRAD && P.grad_in_lds) {
    gacc = lds;
    gw = (P.wsize + 3) / 4 * 4;
    for (int k = threadIdx.x; k < P.wsize; k += 64 * nw) gacc[k] = 0.0f;
    __syncthreads();
  }
  NetLds N = carve_net(lds + gw + wave * net_grad_wave_words(P), P.nsf, P.nl, P.hs);
  for (int i = blockIdx.x * nw + wave; i < P.natoms; i += gridDim.x * nw) {
    wave_sync();
    const int ti = uniform(a.types[i]);
    if (ti < 1 || ti > md.ntypes) {
      if (lane == 0 && a.flags) atomicOr(a.flags, AENET_AMD_FLAG_BAD_TYPE);
      continue;
    }
    const SpeciesDev &sp = md.species[ti - 1];
    for (int f = lane; f < sp.nsf; f += WAVE) N.gp[f] = (float)a.G[(size_t)i * a.stride + f];
    wave_sync();
    const float y = network<true>(md, sp, (global_weights)md.W, N, false);
    const double E = atom_energy(sp, y);
    double g;
    if (a.E_ref) {
      const double w = a.weight ? a.weight[i] : 1.0;
      const double r = E - a.E_ref[i];
      g = a.coef * w * r;
      if (lane == 0 && a.loss) unsafeAtomicAdd(a.loss, w * r * r);
    } else {
      g = a.upstream ? a.upstream[i] : 0.0;
    }
    if (lane == 0 && a.E_out) a.E_out[i] = E;
    if (!GRAD) continue;
    // backward: dcur = dy/dz of layer l+1 (pre-activation), starting at the output
    const int nl = sp.nlayers;
    float *dcur = N.d0;
    float *dnext = N.d1;
    if (lane == 0) dcur[0] = N.dact[(nl - 1) * N.hs];
    wave_sync();
    const float s = (float)(g * sp.E_scale);
    for (int l = nl - 2; l >= 0; --l) {
      const int n1 = sp.nnodes[l], n2 = sp.nnodes[l + 1];
      const float *in = l == 0 ? N.gp : N.act + l * N.hs;
      const int nblk = n2 * (n1 + 1);
      const int w0 = sp.w_off[l];
      // dE/dW_l[h + c*n2] = s * dcur[h] * in[c] (in[n1] = 1: bias)
      for (int idx = lane; idx < nblk; idx += WAVE) {
        const int c = idx / n2, h = idx - c * n2;
        const float v = s * dcur[h] * (c < n1 ? in[c] : 1.0f);
        if (gacc)
          atomicAdd(&gacc[w0 + idx], v);
        else
          atomicAdd(&a.grad[w0 + idx], v);
      }
      if (l > 0) {
        const float *Wl = md.W + w0;
        for (int c = lane; c < n1; c += WAVE) {
          const float *row = Wl + c * n2;
          float acc = 0.0f;
          for (int h = 0; h < n2; ++h) acc += row[h] * dcur[h];
          dnext[c] = acc * N.dact[l * N.hs + c];
        }
      }
      wave_sync();
      float *t = dcur;
      dcur = dnext;
      dnext = t;
    }
  }
  if (GRAD && gacc) {
    __syncthreads();
    for (int k = threadIdx.x; k < P.wsize; k += 64 * nw) {
      const float v = gacc[k];
      if (v != 0.0f) atomicAdd(&a.grad[k], v);
    }
  }
}

// Adam (src/optimize.f90:1012-1046): m = b1 m + (1-b1) g, v = b2 v +
// (1-b2) g^2, W -= lr * mhat / (sqrt(vhat) + eps), bias corrections with
// step t; FP32 weights, moments and gradient
__global__ void k_adam(float *W, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  const float mh = c1 * mi, vh = c2 * vi;
  W[i] = W[i] - lr / (sqrtf(vh) + eps) * mh;
}

void launch_net_grad(const ModelDev &md, const NetGradArgs &a, int natoms, int nsf_max, int nl_max, int wsize,
                     bool grad, hipStream_t stream, int hmax) {
  if (natoms <= 0) return;
  NetGradLaunch P;
  P.hs = net_stride(hmax);
  P.natoms = natoms;
  P.nsf = nsf_max;
  P.nl = std::max(nl_max, 2);
  P.wsize = wsize;
  PersistentShape ps;
  const size_t gbytes = grad ? 4 * (size_t)((wsize + 3) / 4 * 4) : 0;
  if (!persistent_shape(gbytes, 4 * (size_t)net_grad_wave_words(P), natoms, 4, ps)) {
    if (a.flags) raise_flag(a.flags, AENET_AMD_FLAG_NB_OVERFLOW, stream);
    return;
  }
  P.grad_in_lds = ps.w_in_lds;
  P.waves = ps.waves;
  if (grad)
    hipLaunchKernelGGL(k_net_grad<true>, dim3(ps.grid), dim3(64 * ps.waves), ps.lds, stream, md, a, P);
  else
    hipLaunchKernelGGL(k_net_grad<false>, dim3(ps.grid), dim3(64 * ps.waves), ps.lds, stream, md, a, P);
}

void launch_adam(float *W, const float *g, float *m, float *v, int n, int t, float lr, float b1, float b2,
                 float eps, hipStream_t stream) {
  if (n <= 0) return;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)t)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)t)));
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, stream, W, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

}  // namespace aenet_amd

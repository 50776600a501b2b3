// atom_bp.hip -- Behler-Parrinello atom kernel (not built yet: evaluations
// of Behler2011 potentials report an unsupported-type error).
#include <hip/hip_runtime.h>

#include "aenet_amd_internal.h"
#include "kernels.h"

namespace aenet_amd {

__global__ void k_bp_unsupported(int *flags) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicOr(flags, AENET_AMD_FLAG_BAD_TYPE);
}

void launch_atom_bp(int, bool, const AtomSizes &, const ModelDev &, const BatchDev &b, const NbDev &,
                    const OutDev &, const EnvDev &, hipStream_t stream) {
  hipLaunchKernelGGL(k_bp_unsupported, dim3(1), dim3(64), 0, stream, b.flags);
}

}  // namespace aenet_amd

// noalloc.hip -- aenet_structures_energy_and_forces without device
// allocation: every device buffer of the call is carved out of one
// file-scope device array of fixed size (include/aenet_amd.h).
//
// Host code never uses the array's name as an address: the device address
// of this GPU's copy comes from hipGetSymbolAddress, once per device, and
// only that pointer is handed to the kernels and copies.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "aenet_amd.h"
#include "aenet_capi_status.h"
#include "engine.h"

#define AENET_AMD_SCRATCH_BYTES (64ull << 20)
#define AENET_AMD_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[AENET_AMD_SCRATCH_BYTES];

namespace {

std::mutex s_mu;
void *s_ptr[AENET_AMD_MAX_DEVICES];  // device address of each GPU's copy

// this device's copy of the scratch array (looked up on the device's
// first call, with that device current)
int scratch_ptr(void **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= AENET_AMD_MAX_DEVICES)
    return aenet_amd::set_last_error(AENET_STATUS_ERR_DEVICE, "no usable HIP device");
  std::lock_guard<std::mutex> lk(s_mu);
  if (!s_ptr[dev]) {
    void *q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q)
      return aenet_amd::set_last_error(AENET_STATUS_ERR_DEVICE, "device scratch array not found on this device");
    s_ptr[dev] = q;
  }
  *out = s_ptr[dev];
  return AENET_STATUS_OK;
}

}  // namespace

extern "C" void aenet_structures_energy_and_forces_noalloc(int nstruct, const int natoms[], const double latt[],
                                                           const int pbc[], const int types[],
                                                           const double coords[], double E_tot[], double E_coh[],
                                                           double E_atoms[], double F[], int *stat) {
  void *p = nullptr;
  int st = scratch_ptr(&p);
  if (st) {
    *stat = st;
    return;
  }
  aenet_amd::ScratchArena arena;
  arena.base = static_cast<char *>(p);
  arena.cap = AENET_AMD_SCRATCH_BYTES;
  aenet_amd::structures_energy_and_forces(nstruct, natoms, latt, pbc, types, coords, E_tot, E_coh, E_atoms, F,
                                          &arena, "aenet_structures_energy_and_forces_noalloc", stat);
}

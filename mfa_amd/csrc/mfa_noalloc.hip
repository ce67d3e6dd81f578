// mfa_verlet_velocities_noalloc: verlet_velocities on arrays the caller owns,
// allocating nothing.  Its scratch (the kinetic-energy block partials and
// their sum) is the file-scope device array below.  It sits in this source
// file of its own, so it forms a code object of its own: the library's main
// code object carries no large zero-initialised segment, and this one is
// loaded on a device only when the entry point first runs there.
//
// Each GPU holds its own copy of `scratch`; its address on the current
// device is looked up once per device (with that device current) and kept.
// Host code passes only that address, never the array's name, to kernels and
// copies.  Calls on one GPU share the array: calls that overlap on one GPU
// race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdio.h>

#include "../../include/mfa_hip.h"

#define NOALLOC_BLOCK 256  // particles per block of the kernels (MFA_BLOCK)
#define NOALLOC_TERMS 12   // partial sums per block (MFA_NTERMS)
#define NOALLOC_BLOCKS (MFA_NOALLOC_MAX_N / NOALLOC_BLOCK)
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[(NOALLOC_BLOCKS + 1) * NOALLOC_TERMS];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES];

// in mfa_device.hip
int mfa_internal_vv_scratch(const mfa_params* p, double* force, double* v, double* a,
                            const int32_t* a_type, void* stream, double* t_kin, double* part,
                            double* red);
int mfa_internal_set_error(int code, const char* msg);

// The address of `scratch` on the current device, looked up on the first
// call on that device.
static int scratch_of_current_device(double** out) {
  char msg[256];
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    snprintf(msg, sizeof msg, "hipGetDevice: %s", hipGetErrorString(e));
    return mfa_internal_set_error(MFA_ERR_HIP, msg);
  }
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    snprintf(msg, sizeof msg, "device %d beyond the %d the library tracks", dev, NOALLOC_MAX_DEVICES);
    return mfa_internal_set_error(MFA_ERR_HIP, msg);
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      snprintf(msg, sizeof msg, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return mfa_internal_set_error(MFA_ERR_HIP, msg);
    }
    g_scratch_ptr[dev] = (double*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return MFA_OK;
}

extern "C" int32_t mfa_verlet_velocities_noalloc(const mfa_params* p, double* force, double* v,
                                                 double* a, const int32_t* a_type, void* stream,
                                                 double* t_kin) {
  char msg[256];
  if (!p || !force || !v || !a || !a_type)
    return mfa_internal_set_error(MFA_ERR_ARG, "mfa_verlet_velocities_noalloc: null argument");
  if (p->n_part < 1 || p->n_part > MFA_NOALLOC_MAX_N) {
    snprintf(msg, sizeof msg,
             "mfa_verlet_velocities_noalloc: n_part = %d outside 1..%d (the scratch limit)", p->n_part,
             MFA_NOALLOC_MAX_N);
    return mfa_internal_set_error(MFA_ERR_ARG, msg);
  }
  double* scr = nullptr;
  const int rc = scratch_of_current_device(&scr);
  if (rc) return rc;
  return mfa_internal_vv_scratch(p, force, v, a, a_type, stream, t_kin, scr,
                                 scr + (size_t)NOALLOC_BLOCKS * NOALLOC_TERMS);
}
